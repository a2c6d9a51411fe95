// Symmetric eigensolver for orders above 128 (the exact SVD of wide
// cross-covariances, e.g. 1024 features x 1000 classes, and the large SPD
// slice solves of PLSSVD), by two-sided block Jacobi.
//
// The order n is split into nb (even) blocks of at most 64 rows.  A sweep is
// the round-robin tournament over the blocks (circle method: nb - 1 rounds of
// nb / 2 disjoint pairs).  In a round every pair (I, J) is diagonalised
// exactly: the pair's principal submatrix A[I u J, I u J] (order <= 128) goes
// to the one-workgroup divide-and-conquer solver (k_eig.hip, one workgroup
// per pair), and its eigenvector matrix Z_p forms the block rotation J of the
// round, its columns ordered nearest to the identity (k_jac_nearid).  Then
// A <- J^T A J and V <- V J, computed as column updates
// X[:, I u J] <- X[:, I u J] Z_p: B = A J, A <- (B^T) J (A symmetric), V <- V J.
// After a round the pair blocks of A are diagonal; sweeps repeat until the
// off-diagonal Frobenius norm falls below 1e-14 of the total (block Jacobi
// with exact pair solves converges quadratically once the off-diagonal mass
// is small; Drmac & Veselic 2008 for the method, Golub & Van Loan 8.5 for the
// tournament ordering).  Eigenvalues are the final diagonal, sorted
// descending with their eigenvector columns.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "fpls_internal.h"

// Convergence hook (sweep, off-diagonal Frobenius^2, total Frobenius^2);
// compiled out unless a probe build defines it.
#ifndef FPLS_JAC_SWEEP
#define FPLS_JAC_SWEEP(sweep, off2, tot2)
#endif

namespace fpls {

namespace {

constexpr int JB_PAIR = 128;  // largest pair order (two blocks of <= 64)

// M_p = A[ix_p, ix_p] for every pair of the round (ld 128, stride 128^2).
__global__ void k_jac_gather(const double* __restrict__ A, int64_t lda,
                             const int* __restrict__ idx, const int* __restrict__ ns,
                             double* __restrict__ M) {
  const int p = blockIdx.y, s = ns[p];
  const int* ix = idx + p * JB_PAIR;
  double* Mp = M + (int64_t)p * JB_PAIR * JB_PAIR;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < s * s; e += gridDim.x * blockDim.x) {
    const int i = e % s, j = e / s;
    Mp[i + j * JB_PAIR] = A[ix[i] + (int64_t)ix[j] * lda];
  }
}

// D[:, ix_p] = S[:, ix_p] Z_p for every pair p of the round (Z_p: s_p x s_p,
// ld s_p, at Z + p 128^2).  Tile: 64 rows x 64 output columns of one pair,
// K in chunks of 32 through LDS, 4 x 4 outputs per thread, FP64 FMA.
__global__ __launch_bounds__(256) void k_jac_cols(int n, const double* __restrict__ S,
                                                  int64_t lds, double* __restrict__ D,
                                                  int64_t ldd, const int* __restrict__ idx,
                                                  const int* __restrict__ ns,
                                                  const double* __restrict__ Z) {
  __shared__ double As[32][64 + 1];  // [k][row]
  __shared__ double Zs[32][64 + 1];  // [k][col]
  const int p = blockIdx.z, s = ns[p];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  if (c0 >= s) return;
  const int* ix = idx + p * JB_PAIR;
  const double* Zp = Z + (int64_t)p * JB_PAIR * JB_PAIR;
  const int t = threadIdx.x, tr = t % 16, tc = t / 16;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < s; k0 += 32) {
    for (int e = t; e < 32 * 64; e += 256) {
      const int r = e % 64, k = e / 64;
      const int gr = r0 + r, gk = k0 + k;
      As[k][r] = (gr < n && gk < s) ? S[gr + (int64_t)ix[gk] * lds] : 0.0;
      const int gc = c0 + r;  // the same (k, 64-wide) shape for Z
      Zs[k][r] = (gc < s && gk < s) ? Zp[gk + (int64_t)gc * s] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; ++k) {
      double a[4], z[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][tr + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) z[j] = Zs[k][tc + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], z[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gr = r0 + tr + 16 * i, gc = c0 + tc + 16 * j;
      if (gr < n && gc < s) D[gr + (int64_t)ix[gc] * ldd] = acc[i][j];
    }
}

// The pair's eigenvector columns, reordered (and signed) to the permutation
// closest to the identity: columns in order of decreasing largest entry, each
// taking the row of its largest |entry| not yet taken, sign so that this
// entry is positive.  Block Jacobi converges only when the block rotations
// stay near the identity (sorting the pair's eigenvalues instead makes them
// permutation-like and the sweeps stall; checked in a float64 model).
//
// Fast path: when the largest-|entry| rows of the columns (ties -> lower row)
// are pairwise distinct -- the usual case once the sweeps approach
// convergence -- every column takes its own argmax row whatever the order,
// which is exactly the greedy result; the greedy loop runs only on a
// conflict.  Column maxima and argmaxes: one wave per column, 2 rows per lane.
__global__ __launch_bounds__(256) void k_jac_nearid(const int* __restrict__ ns,
                                                    const double* __restrict__ Z,
                                                    double* __restrict__ Zo) {
  __shared__ double Zs[JB_PAIR * JB_PAIR];  // the pair's eigenvectors (ld s)
  __shared__ double cmax[JB_PAIR];
  __shared__ int dest[JB_PAIR], amax[JB_PAIR], claim[JB_PAIR];
  __shared__ int conflict;
  const int p = blockIdx.x, s = ns[p], t = threadIdx.x, lane = t & 63, w = t >> 6;
  const double* Zp = Z + (int64_t)p * JB_PAIR * JB_PAIR;
  double* Zq = Zo + (int64_t)p * JB_PAIR * JB_PAIR;
  for (int e = t; e < s * s; e += 256) Zs[e] = Zp[e];
  if (t < JB_PAIR) claim[t] = 0;
  if (t == 0) conflict = 0;
  __syncthreads();
  for (int j = w; j < s; j += 4) {
    double v = -1.0;
    int ix = 0;
    for (int i = lane; i < s; i += 64) {
      const double a = fabs(Zs[i + j * s]);
      if (a > v) {  // rows visited in increasing order: ties keep the lower row
        v = a;
        ix = i;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_down(v, o, 64);
      const int oi = __shfl_down(ix, o, 64);
      if (ov > v || (ov == v && oi < ix)) {
        v = ov;
        ix = oi;
      }
    }
    if (lane == 0) {
      cmax[j] = v;
      amax[j] = ix;
    }
  }
  __syncthreads();
  if (t < s && atomicAdd(&claim[amax[t]], 1) != 0) conflict = 1;
  __syncthreads();
  // columns whose argmax row no other column wants keep it
  if (t < s) {
    const bool mine = claim[amax[t]] == 1;
    dest[t] = mine ? amax[t] : -1;
    if (!mine) conflict = 1;
  }
  __syncthreads();
  if (conflict && w == 0) {
    // the other columns, greedily in one wave: by decreasing maximum (ties ->
    // lower column), each taking its largest row not yet held; lane owns rows
    // lane and lane + 64, their held flags in registers
    const int r0 = lane, r1 = lane + 64;
    bool tk0 = r0 < s && claim[r0] == 1, tk1 = r1 < s && claim[r1] == 1;
    int todo = 0;
    for (int j = lane; j < s; j += 64) {
      if (dest[j] >= 0) cmax[j] = -3.0;  // placed
      else ++todo;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) todo += __shfl_xor(todo, o, 64);
    int done = 0;
    while (done < todo) {
      // next column in the order: the largest cmax among columns not yet placed
      double cv = -2.0;
      int cj = JB_PAIR;
      for (int j = lane; j < s; j += 64) {
        const double c = cmax[j];
        if (c > cv || (c == cv && j < cj)) {
          cv = c;
          cj = j;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(cv, o, 64);
        const int oj = __shfl_xor(cj, o, 64);
        if (ov > cv || (ov == cv && oj < cj)) {
          cv = ov;
          cj = oj;
        }
      }
      const int j = cj;
      // free rows rank by |entry| with NaN below every number (-0.5) and
      // held rows below that (-1): a total order on (value, row), so every
      // lane's butterfly ends on the same free row
      double v = -1.0;
      int ix = JB_PAIR;
      if (r0 < s && !tk0) {
        const double a = fabs(Zs[r0 + j * s]);
        v = a >= 0.0 ? a : -0.5;
        ix = r0;
      }
      if (r1 < s && !tk1) {
        double a = fabs(Zs[r1 + j * s]);
        a = a >= 0.0 ? a : -0.5;
        if (a > v) {
          v = a;
          ix = r1;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(ix, o, 64);
        if (ov > v || (ov == v && oi < ix)) {
          v = ov;
          ix = oi;
        }
      }
      if (ix == r0) tk0 = true;
      if (ix == r1) tk1 = true;
      if (lane == 0) {
        dest[j] = ix;
        cmax[j] = -3.0;  // placed
      }
      __builtin_amdgcn_wave_barrier();
      ++done;
    }
  }
  __syncthreads();
  for (int e = t; e < s * s; e += 256) {
    const int i = e % s, j = e / s, d = min(max(dest[j], 0), s - 1);
    const double sg = Zs[d + j * s] >= 0.0 ? 1.0 : -1.0;
    Zq[i + d * s] = sg * Zs[i + j * s];
  }
}

// A (n x n, ld n) = (S + S^T) / 2; V = I
__global__ void k_jac_init(int n, const double* __restrict__ S, int64_t lds,
                           double* __restrict__ A, double* __restrict__ V) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    A[e] = 0.5 * (S[i + (int64_t)j * lds] + S[j + (int64_t)i * lds]);
    V[e] = i == j ? 1.0 : 0.0;
  }
}

// out[0] = sum of squared off-diagonal entries, out[1] = sum of all squares
// (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void k_jac_offnorm(int n, const double* __restrict__ A,
                                                      double* __restrict__ out) {
  __shared__ double r0[16], r1[16];
  double off = 0.0, tot = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 1024) {
    const double v = A[e];
    const int i = (int)(e % n), j = (int)(e / n);
    tot = fma(v, v, tot);
    if (i != j) off = fma(v, v, off);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    off += __shfl_down(off, o, 64);
    tot += __shfl_down(tot, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = off;
    r1[threadIdx.x >> 6] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < 16; ++q) {
      a += r0[q];
      b += r1[q];
    }
    out[0] = a;
    out[1] = b;
  }
}

// w[j] = A[perm[j], perm[j]], V[:, j] = W[:, perm[j]]
__global__ void k_jac_finish(int n, const double* __restrict__ A, const double* __restrict__ W,
                             const int* __restrict__ perm, double* __restrict__ w,
                             double* __restrict__ V) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    const int pj = perm[j];
    V[e] = W[i + (int64_t)pj * n];
    if (i == 0) w[j] = A[pj + (int64_t)pj * n];
  }
}

unsigned jgrid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

}  // namespace

void sym_eig_large(Ctx* ctx, const double* S, int64_t lds, int n, double* w, double* V,
                   int kwant) {
  FPLS_REQUIRE(n >= 1, FPLS_ERR_ARG, "sym_eig_large: empty matrix");
  if (n <= kJacobiMax) {
    sym_eig(ctx, S, lds, n, w, V, kwant);
    return;
  }
  // the tridiagonal path on the whole chip; block Jacobi when it declines
  if (sym_eig_chip(ctx, S, lds, n, w, V, kwant)) return;
  const int nb = 2 * ((n + JB_PAIR - 1) / JB_PAIR);  // even, blocks of <= 64 rows
  const int q = nb / 2;
  std::vector<int> lo(nb + 1);
  for (int i = 0; i <= nb; ++i) lo[i] = (int)((int64_t)i * n / nb);
  // round-robin tournament: player 0 fixed, the others rotate
  const int rounds = nb - 1;
  std::vector<int> idx((size_t)rounds * q * JB_PAIR, 0), ns((size_t)rounds * q, 0);
  std::vector<int> ring(nb - 1);
  std::iota(ring.begin(), ring.end(), 1);
  for (int r = 0; r < rounds; ++r) {
    std::vector<int> L(nb);
    L[0] = 0;
    for (int i = 0; i < nb - 1; ++i) L[i + 1] = ring[(i + r) % (nb - 1)];
    for (int p = 0; p < q; ++p) {
      const int bi = L[p], bj = L[nb - 1 - p];
      int* ix = &idx[((size_t)r * q + p) * JB_PAIR];
      int s = 0;
      for (int v = lo[bi]; v < lo[bi + 1]; ++v) ix[s++] = v;
      for (int v = lo[bj]; v < lo[bj + 1]; ++v) ix[s++] = v;
      ns[(size_t)r * q + p] = s;
    }
  }
  DBuf<int> d_idx(ctx, idx.size()), d_ns(ctx, ns.size());
  h2d(ctx, d_idx.get(), idx.data(), idx.size() * sizeof(int));
  h2d(ctx, d_ns.get(), ns.data(), ns.size() * sizeof(int));
  const size_t nn = (size_t)n * n;
  DBuf<double> Aw(ctx, nn), Bw(ctx, nn), Bt(ctx, nn), Vw(ctx, nn), Vt(ctx, nn);
  DBuf<double> M(ctx, (size_t)q * JB_PAIR * JB_PAIR), Z(ctx, (size_t)q * JB_PAIR * JB_PAIR);
  DBuf<double> Zn(ctx, (size_t)q * JB_PAIR * JB_PAIR);
  DBuf<double> wb(ctx, (size_t)q * JB_PAIR), norms(ctx, 2);
  launch(ctx, "jacobi_init", k_jac_init, dim3(jgrid(nn)), dim3(256), 0, n, S, lds, Aw.get(),
         Vw.get());
  double* Vcur = Vw.get();
  double* Vnext = Vt.get();
  {
    // refuse a non-finite matrix before the first sweep
    launch(ctx, "jacobi_offnorm", k_jac_offnorm, dim3(1), dim3(1024), 0, n,
           (const double*)Aw.get(), norms.get());
    double h0[2];
    to_host(ctx, h0, norms.get(), 2);
    FPLS_REQUIRE(std::isfinite(h0[1]), FPLS_ERR_ARG,
                 "symmetric eigensolver: the matrix holds non-finite values (non-finite X or Y?)");
  }
  const dim3 cgrid((unsigned)((n + 63) / 64), (unsigned)(JB_PAIR / 64), (unsigned)q);
  const int max_sweeps = 40;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    for (int r = 0; r < rounds; ++r) {
      const int* ix = d_idx.get() + (size_t)r * q * JB_PAIR;
      const int* nsr = d_ns.get() + (size_t)r * q;
      launch(ctx, "jacobi_gather", k_jac_gather, dim3(16, (unsigned)q), dim3(256), 0,
             (const double*)Aw.get(), (int64_t)n, ix, nsr, M.get());
      sym_eig_batched(ctx, M.get(), JB_PAIR, (int64_t)JB_PAIR * JB_PAIR, nsr, JB_PAIR, q,
                      wb.get(), JB_PAIR, Z.get(), (int64_t)JB_PAIR * JB_PAIR, 0, 0, true);
      launch(ctx, "jacobi_nearid", k_jac_nearid, dim3((unsigned)q), dim3(256), 0, nsr,
             (const double*)Z.get(), Zn.get());
      // B = A J; A <- B^T J; V <- V J
      launch(ctx, "jacobi_cols", k_jac_cols, cgrid, dim3(256), 0, n, (const double*)Aw.get(),
             (int64_t)n, Bw.get(), (int64_t)n, ix, nsr, (const double*)Zn.get());
      transpose_f64(ctx, Bw.get(), n, n, n, Bt.get(), n);
      launch(ctx, "jacobi_cols", k_jac_cols, cgrid, dim3(256), 0, n, (const double*)Bt.get(),
             (int64_t)n, Aw.get(), (int64_t)n, ix, nsr, (const double*)Zn.get());
      launch(ctx, "jacobi_cols", k_jac_cols, cgrid, dim3(256), 0, n, (const double*)Vcur,
             (int64_t)n, Vnext, (int64_t)n, ix, nsr, (const double*)Zn.get());
      std::swap(Vcur, Vnext);
    }
    launch(ctx, "jacobi_offnorm", k_jac_offnorm, dim3(1), dim3(1024), 0, n,
           (const double*)Aw.get(), norms.get());
    double h[2];
    to_host(ctx, h, norms.get(), 2);
    FPLS_REQUIRE(std::isfinite(h[1]), FPLS_ERR_ARG, "sym_eig_large: non-finite input");
    FPLS_JAC_SWEEP(sweep, h[0], h[1]);
    if (h[0] <= 1e-28 * h[1]) break;
  }
  // eigenvalues descending (stable on ties by index)
  std::vector<double> diag(n);
  {
    std::vector<double> tmp(nn);
    to_host(ctx, tmp.data(), Aw.get(), nn);
    for (int i = 0; i < n; ++i) diag[i] = tmp[i + (size_t)i * n];
  }
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return diag[a] > diag[b]; });
  DBuf<int> d_perm(ctx, n);
  h2d(ctx, d_perm.get(), perm.data(), n * sizeof(int));
  launch(ctx, "jacobi_finish", k_jac_finish, dim3(jgrid(nn)), dim3(256), 0, n,
         (const double*)Aw.get(), (const double*)Vcur, (const int*)d_perm.get(), w, V);
}

// Slice solves of order above 128 (spd_solve_slices' modes) through the
// eigendecomposition A_k = V diag(lam) V^T: X = V diag(1 / (lam + shift)) V^T
// (times diag(d_k) in modes 0 and 2).  The shift rules are those of the
// one-workgroup sweep kernel.  Status mirrors the reference's
// solve(likely_sympd) -> general solve fallback: 0 when every lam + shift > 0
// (the Cholesky would succeed), 2 when the matrix is indefinite but
// nonsingular (the LU fallback's solution is the same X), 1 when singular.
__global__ void k_jac_diag(const double* __restrict__ A, int64_t lda, int k,
                           double* __restrict__ d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
    d[i] = A[i + (int64_t)i * lda];
}

__global__ void k_jac_inv_cols(double* __restrict__ Y, int k, const double* __restrict__ lam,
                               double shift) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)k * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / k);
    Y[e] /= (lam[c] + shift);
  }
}

void spd_solve_large(Ctx* ctx, const double* A, int64_t lda, const std::vector<int>& ks,
                     int n_slices, int mode, const double* dvals, double ridge, double* X,
                     int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status) {
  for (int sl = 0; sl < n_slices; ++sl) {
    const int k = ks[sl];
    FPLS_REQUIRE(k >= 1, FPLS_ERR_ARG, "spd_solve: order out of range");
    double shift = 0.0;
    if (mode == 1 || mode == 2) {
      DBuf<double> dg(ctx, k);
      launch(ctx, "jacobi_diag", k_jac_diag, dim3(jgrid(k)), dim3(256), 0, A, lda, k, dg.get());
      std::vector<double> hd(k);
      to_host(ctx, hd.data(), dg.get(), k);
      double tr = 0.0;
      for (double v : hd) tr += v;
      if (mode == 1) {
        double scale = tr / (double)(k > 1 ? k : 1);
        if (!std::isfinite(scale) || scale <= 0.0) scale = 1.0;
        double lam = ridge;
        if (!std::isfinite(lam) || lam < 0.0) lam = 1e-8;
        shift = lam * scale;
        if (lambda_out) to_device(ctx, lambda_out + sl, &shift, 1);
      } else {
        shift = 1e-10 * (tr / (double)k);
        if (!std::isfinite(shift) || shift <= 0.0) shift = 1e-10;
      }
    }
    DBuf<double> w(ctx, k), V(ctx, (size_t)k * k), Y(ctx, (size_t)k * k);
    sym_eig_large(ctx, A, lda, k, w.get(), V.get());
    std::vector<double> hw(k);
    to_host(ctx, hw.data(), w.get(), k);
    int32_t st = 0;
    for (double v : hw) {
      const double e = v + shift;
      if (!std::isfinite(e) || e == 0.0) {
        st = 1;
        break;
      }
      if (e < 0.0) st = 2;
    }
    if (status) h2d(ctx, status + sl, &st, sizeof(int32_t));
    if (st == 1) continue;
    copy_f64(ctx, Y.get(), V.get(), (int64_t)k * k);
    launch(ctx, "jacobi_invcols", k_jac_inv_cols, dim3(jgrid((int64_t)k * k)), dim3(256), 0,
           Y.get(), k, (const double*)w.get(), shift);
    double* Xs = X + (int64_t)sl * slice_stride;
    gemm_f64(ctx, false, true, k, k, k, 1.0, Y.get(), k, nullptr, V.get(), k, nullptr, 0.0, Xs,
             ldx);
    if (mode != 1) scale_cols(ctx, Xs, ldx, k, k, dvals, false);
  }
}

}  // namespace fpls
