// Symmetric eigenpairs of orders 129 .. 2048 on the whole chip (the Gram
// matrix of a wide class-sum matrix, latent Gram matrices of large ncomp):
//
//  1. Householder tridiagonalisation (Golub & Van Loan, Algorithm 8.3.1),
//     one launch per column j: every workgroup owns one 64 x 128 tile of A
//     for the whole reduction (the same workgroup index, so the same XCD and
//     L2), applies the pending rank-2 update of column j - 1 to it and
//     multiplies it by v_j; the last workgroup to finish (an atomic ticket)
//     adds the partial products, forms w_j = tau_j A v_j with
//     w_j += -tau_j / 2 (w_j^T v_j) v_j, updates column j + 1 and forms
//     reflector j + 1, so each launch moves the trailing block once.
//  2. The wanted eigenvalues of T (the kwant largest) by multisection on the
//     Sturm count and their vectors from the twisted factorisation, one
//     workgroup each, as for the small orders (k_eig_ii.hip).
//  3. Symmetric re-orthogonalisation Z <- Z - Z (Z^T Z - I) / 2, twice (two
//     GEMMs each): vectors of close eigenvalues, each accurate to
//     eps |T| / gap, come out orthonormal; a first-pass deviation above 0.05
//     (eigenvalues too close for independent vectors) flags the problem.
//  4. Back-transformation V = H_0 ... H_{n-3} Z by panels, last first:
//     Z <- Z - V_p (T_p (V_p^T Z)), with the compact-WY factors T_p of every
//     panel (dlarft, forward, columnwise) from one launch.
//
// A flagged problem (a residual above 1000 n eps |T|, a non-finite vector, a
// deviation above the bound) is solved again by the block Jacobi method.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "fpls_internal.h"
#include "fpls_wave.h"

namespace fpls {

namespace {

constexpr int LT_NB = 32;    // reflectors per panel
constexpr int LT_RB = 64;    // tile rows of the per-column product
constexpr int LT_CB = 128;   // tile columns
constexpr int LT_T = 256;    // threads per workgroup (start, pairs, T factors)
constexpr int LC_T = 256;    // threads of the per-column kernel
constexpr int LT_NMAX = 2048;
constexpr double LT_EPS = DBL_EPSILON * 0.5;

// sum over the workgroup of NT threads (every thread gets the total); red:
// NT / 64 doubles of LDS
template <int NT>
__device__ __forceinline__ double lt_block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum_d(v);
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) s += red[q];
  __syncthreads();
  return s;
}

// Reflector j from the corrected column xs[j .. n-1] (LDS, visible to the
// whole workgroup): d_j = xs[j], H_j maps xs[j+1 ..] to beta e_1.  Column j of
// Rf gets v_j (zero above j + 1, one at j + 1); tri gets d_j, e_j = beta, tau_j.
template <int NT>
__device__ void lt_reflector(const double* xs, int j, int n, double* __restrict__ Rf,
                             double* __restrict__ tri, double* red) {
  const int t = threadIdx.x;
  double s = 0.0;
  for (int r = j + 2 + t; r < n; r += NT) s = fma(xs[r], xs[r], s);
  const double xn2 = lt_block_sum<NT>(s, red);
  const double alpha = xs[j + 1];
  double beta = alpha, tau = 0.0, sv = 0.0;
  if (xn2 > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    sv = 1.0 / (alpha - beta);
  }
  for (int r = t; r < n; r += NT)
    Rf[r + (int64_t)j * n] = (r == j + 1) ? 1.0 : (r > j + 1 ? xs[r] * sv : 0.0);
  if (t == 0) {
    tri[j] = xs[j];
    tri[n + j] = beta;
    tri[2 * n + j] = tau;
  }
}

// A = (S + S^T) / 2 into the working buffer (ld n)
__global__ void k_lt_sym(const double* __restrict__ S, int64_t lds, int n, double* __restrict__ A) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    A[e] = 0.5 * (S[i + (int64_t)j * lds] + S[j + (int64_t)i * lds]);
  }
}

// Reflector 0 from column 0 of A.
__global__ __launch_bounds__(LT_T) void k_lt_start(const double* __restrict__ A, int n,
                                                   double* __restrict__ Rf,
                                                   double* __restrict__ tri) {
  extern __shared__ double lt_sm[];
  double* xs = lt_sm;  // n
  __shared__ double red[4];
  for (int r = threadIdx.x; r < n; r += LT_T) xs[r] = A[r];
  __syncthreads();
  lt_reflector<LT_T>(xs, 0, n, Rf, tri, red);
}

// Column j: grid (column tiles, row tiles) over the whole matrix, so a tile
// keeps its workgroup (and XCD, and L2) from launch to launch.  Every tile
// applies the pending update of column j - 1, A -= v w^T + w v^T (v = v_{j-1},
// w = w_{j-1}, both zero above row j), to its entries at rows and columns
// >= j, then multiplies them by v_j (wave wv: columns wv, wv + 4, ..., lane =
// tile row).  Partial products and the updated column j + 1 travel through
// agent-scope atomic stores and loads (coherent across XCDs without an L2
// write-back), ordered before the ticket by waiting for their completion.
// The last workgroup adds the partials, forms w_j = tau_j A v_j,
// w_j += -tau_j / 2 (w_j^T v_j) v_j, updates column j + 1 by (v_j, w_j) and
// forms reflector j + 1.
__global__ __launch_bounds__(LC_T) void k_lt_col(double* __restrict__ A, int n, int j,
                                                 double* __restrict__ Rf,
                                                 double* __restrict__ Wc,
                                                 double* __restrict__ tri,
                                                 double* __restrict__ Yp,
                                                 double* __restrict__ xcol,
                                                 unsigned* __restrict__ counter) {
  extern __shared__ double lt_sm[];
  constexpr int NW = LC_T / 64;
  __shared__ double part[NW][LT_RB];
  __shared__ double red[NW];
  __shared__ unsigned ticket;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int cc = blockIdx.x, rb = blockIdx.y;
  const int r0 = rb * LT_RB, c0 = cc * LT_CB;
  const double* __restrict__ v = Rf + (int64_t)j * n;
  const bool pend = j > 0;
  const double* __restrict__ vp = Rf + (int64_t)(pend ? j - 1 : 0) * n;
  const double* __restrict__ wp = Wc + (int64_t)((j - 1) & 1) * n;
  {
    // wave wv takes columns c0 + wv + NW i; every load of the thread's
    // LT_CB / NW columns is issued before the first is used, so a tile costs
    // about one memory latency rather than one per pair of columns
    // (the tile's entries of v_j, v_{j-1} and w_{j-1} go through LDS)
    constexpr int CPT = LT_CB / NW;
    __shared__ double sv[LT_CB], svp[LT_CB], swp[LT_CB];
    const int r = r0 + lane;
    const bool row_ok = r < n && r >= j;
    const double vr = (pend && row_ok) ? vp[r] : 0.0, wr = (pend && row_ok) ? wp[r] : 0.0;
    double a[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int c = c0 + wv + NW * i;
      a[i] = (row_ok && c >= j && c < n) ? A[r + (int64_t)c * n] : 0.0;
    }
    if (t < LT_CB) {
      const int c = c0 + t;
      sv[t] = c < n ? v[c] : 0.0;
      svp[t] = (pend && c < n) ? vp[c] : 0.0;
      swp[t] = (pend && c < n) ? wp[c] : 0.0;
    }
    __syncthreads();
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int c = c0 + wv + NW * i, cl = wv + NW * i;
      if (row_ok && c >= j && c < n) {
        double x = a[i];
        if (pend) {
          x = fma(-vr, swp[cl], fma(-wr, svp[cl], x));
          A[r + (int64_t)c * n] = x;
        }
        if (i & 1) acc1 = fma(x, sv[cl], acc1);
        else acc0 = fma(x, sv[cl], acc0);
        if (c == j + 1 && r > j)
          __hip_atomic_store(&xcol[r], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    part[wv][lane] = acc0 + acc1;
  }
  __syncthreads();
  if (t < LT_RB && r0 + t < n) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) s += part[q][t];
    __hip_atomic_store(&Yp[(int64_t)cc * n + r0 + t], s, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
  __syncthreads();
  if (t == 0)
    ticket = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (ticket != gridDim.x * gridDim.y - 1) return;

  // ---- the last workgroup
  if (t == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int ncc = gridDim.x;
  double* xs = lt_sm + n;   // n: column j + 1
  __shared__ double wj1s;
  const double tau = tri[2 * n + j];
  const bool more = j + 1 <= n - 3;
  // this thread's rows r = j + 1 + t + LC_T q: the partials of (A v_j)(r) and
  // column j + 1 are loaded together (one round of memory latency)
  constexpr int NQ = LT_NMAX / LC_T;
  static_assert(LT_NMAX % LC_T == 0, "rows per thread");
  double ysr[NQ], xr[NQ], vq[NQ];
  double dv = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = j + 1 + t + LC_T * q;
    ysr[q] = 0.0;
    xr[q] = 0.0;
    vq[q] = 0.0;
    if (r < n) {
      double s = 0.0;
      for (int c = 0; c < ncc; ++c)
        s += __hip_atomic_load(&Yp[(int64_t)c * n + r], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      if (more)
        xr[q] = __hip_atomic_load(&xcol[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      vq[q] = v[r];
      s *= tau;
      ysr[q] = s;
      dv = fma(s, vq[q], dv);
    }
  }
  const double alpha = -0.5 * tau * lt_block_sum<LC_T>(dv, red);
  double* __restrict__ wj = Wc + (int64_t)(j & 1) * n;
  for (int r = t; r <= j && r < n; r += LC_T) wj[r] = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = j + 1 + t + LC_T * q;
    if (r < n) {
      ysr[q] = fma(alpha, vq[q], ysr[q]);  // w_j(r)
      wj[r] = ysr[q];
      if (r == j + 1) wj1s = ysr[q];
    }
  }
  if (!more) return;
  __syncthreads();
  // column j + 1 after the update by (v_j, w_j)
  const double wj1 = wj1s, vj1 = v[j + 1];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = j + 1 + t + LC_T * q;
    if (r < n) xs[r] = fma(-vq[q], wj1, fma(-ysr[q], vj1, xr[q]));
  }
  __syncthreads();
  lt_reflector<LC_T>(xs, j + 1, n, Rf, tri, red);
}

// d[n-2], e[n-2], d[n-1] after the last update (v_{n-3}, w_{n-3}) of the
// trailing 2 x 2 block; the last two reflector columns and taus are zero.
__global__ void k_lt_end(const double* __restrict__ A, int n, double* __restrict__ Rf,
                         const double* __restrict__ Wc, double* __restrict__ tri) {
  const int t = threadIdx.x;
  if (t == 0) {
    const double* v = Rf + (int64_t)(n - 3) * n;
    const double* w = Wc + (int64_t)((n - 3) & 1) * n;
    const int a = n - 2, b = n - 1;
    tri[a] = A[a + (int64_t)a * n] - 2.0 * v[a] * w[a];
    tri[n + a] = A[b + (int64_t)a * n] - v[b] * w[a] - w[b] * v[a];
    tri[b] = A[b + (int64_t)b * n] - 2.0 * v[b] * w[b];
    tri[n + b] = 0.0;
    tri[2 * n + a] = 0.0;
    tri[2 * n + b] = 0.0;
  }
  __syncthreads();
  for (int r = t; r < 2 * n; r += blockDim.x) {
    const int jj = n - 2 + r / n;
    Rf[(r % n) + (int64_t)jj * n] = 0.0;
  }
}

// ---------------------------------------------------------------- stage 2
__device__ __forceinline__ double lt_rcp1(double d) {
  const double y = __builtin_amdgcn_rcp(d);
  return fma(y, fma(-d, y, 1.0), y);
}

// Gershgorin interval, |T| bound and pivmin over n entries (wave-uniform).
struct LtBounds {
  double gl, gu, tnorm, pivmin;
};
__device__ LtBounds lt_bounds(const double* d, const double* e, const double* ee, int n) {
  const int lane = threadIdx.x & 63;
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i + 1 < n) emax = fmax(emax, ee[i]);
  }
  LtBounds b;
  b.gl = -wave_max_d(-lo);
  b.gu = wave_max_d(hi);
  b.tnorm = fmax(fabs(b.gl), fabs(b.gu));
  b.pivmin = DBL_MIN * fmax(1.0, wave_max_d(emax));
  const double fudge = 2.1 * (b.tnorm * LT_EPS * n + 2.0 * b.pivmin);
  b.gl -= fudge;
  b.gu += fudge;
  return b;
}

constexpr int LP_PTS = 2 * LT_T;

// Workgroup oc: eigenvalue oc from the top (ascending index n - 1 - oc) and
// its eigenvector of T into Z[:, oc] (ld n); w[oc] = eigenvalue.
__global__ __launch_bounds__(LT_T) void k_lt_pairs(int n, const double* __restrict__ tri,
                                                   double* __restrict__ w,
                                                   double* __restrict__ Z,
                                                   int* __restrict__ flag) {
  extern __shared__ double lt_sm[];
  double* d = lt_sm;
  double* e = lt_sm + n;
  double* ee = lt_sm + 2 * n;
  double* f0 = lt_sm + 3 * n;
  double* f1 = lt_sm + 4 * n;
  double* zv = lt_sm + 5 * n;
  __shared__ unsigned long long above_w[LT_T / 64][2];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int oc = blockIdx.x, k = n - 1 - oc;
  for (int q = t; q < n; q += LT_T) {
    d[q] = tri[q];
    e[q] = tri[n + q];
    ee[q] = tri[n + q] * tri[n + q];
  }
  __syncthreads();
  const LtBounds b = lt_bounds(d, e, ee, n);
  double lo = b.gl, hi = b.gu;
  for (int it = 0; it < 12; ++it) {
    const double width = hi - lo;
    const double tol = fmax(LT_EPS * b.tnorm,
                            fmax(b.pivmin, 2.0 * LT_EPS * fmax(fabs(lo), fabs(hi))));
    if (!(width > tol)) break;  // uniform
    const double step = width / (double)(LP_PTS + 1);
    const double x0 = lo + step * (double)(t + 1), x1 = lo + step * (double)(t + 1 + LT_T);
    double q0 = d[0] - x0, q1 = d[0] - x1;
    if (fabs(q0) < b.pivmin) q0 = -b.pivmin;
    if (fabs(q1) < b.pivmin) q1 = -b.pivmin;
    int c0 = q0 < 0.0 ? 1 : 0, c1 = q1 < 0.0 ? 1 : 0;
    for (int q = 1; q < n; ++q) {
      const double dq = d[q], eq = ee[q - 1];
      q0 = (dq - x0) - eq * lt_rcp1(q0);
      q1 = (dq - x1) - eq * lt_rcp1(q1);
      if (fabs(q0) < b.pivmin) q0 = -b.pivmin;
      if (fabs(q1) < b.pivmin) q1 = -b.pivmin;
      c0 += q0 < 0.0 ? 1 : 0;
      c1 += q1 < 0.0 ? 1 : 0;
    }
    const unsigned long long m0 = __ballot(c0 > k), m1 = __ballot(c1 > k);
    if (lane == 0) {
      above_w[wv][0] = m0;
      above_w[wv][1] = m1;
    }
    __syncthreads();
    int f = LP_PTS;
    for (int h = 0; h < 2 && f == LP_PTS; ++h)
      for (int q = 0; q < LT_T / 64; ++q) {
        const unsigned long long m = above_w[q][h];
        if (m) {
          f = h * LT_T + 64 * q + __ffsll((long long)m) - 1;
          break;
        }
      }
    __syncthreads();
    const double xf = f < LP_PTS ? lo + step * (double)(f + 1) : hi;
    const double xb = f > 0 ? lo + step * (double)f : lo;
    lo = xb;
    hi = xf;
  }
  const double lambda = 0.5 * (lo + hi);
  if (wv != 0) return;
  // twisted factorisation: top-down pivots on lane 0, bottom-up on lane 1
  if (lane < 2) {
    const bool up = lane == 1;
    double* fo = up ? f1 : f0;
    int ii = up ? n - 1 : 0;
    double q = d[ii] - lambda;
    if (fabs(q) < b.pivmin) q = -b.pivmin;
    fo[ii] = q;
    for (int s = 1; s < n; ++s) {
      ii = up ? n - 1 - s : s;
      const double e2 = up ? ee[ii] : ee[ii - 1];
      q = (d[ii] - lambda) - e2 * lt_rcp1(q);
      if (fabs(q) < b.pivmin) q = -b.pivmin;
      fo[ii] = q;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double gbest = DBL_MAX;
  int rbest = n;
  for (int r = lane; r < n; r += 64) {
    const double g0 = f0[r], g1 = f1[r];
    const double g = fabs(g0 + g1 - (d[r] - lambda));
    if (g < gbest) {
      gbest = g;
      rbest = r;
    }
    f0[r] = r + 1 < n ? e[r] * lt_rcp1(g0) : 0.0;
    f1[r] = r >= 1 ? e[r - 1] * lt_rcp1(g1) : 0.0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double og = __shfl_xor(gbest, o, 64);
    const int orr = __shfl_xor(rbest, o, 64);
    if (og < gbest || (og == gbest && orr < rbest)) {
      gbest = og;
      rbest = orr;
    }
  }
  const int rt = rbest < n ? rbest : 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < 2) {
    const bool down = lane == 1;
    const double* rat = down ? f1 : f0;
    if (lane == 0) zv[rt] = 1.0;
    const int steps = down ? n - 1 - rt : rt;
    double zp = 1.0, zpp = 0.0;
    for (int s = 1; s <= steps; ++s) {
      const int q = down ? rt + s : rt - s;
      double z = -rat[q] * zp;
      if (zp == 0.0 && s >= 2) {
        const double eq = down ? e[q - 1] : e[q], efar = down ? e[q - 2] : e[q + 1];
        z = eq != 0.0 ? -(efar / eq) * zpp : 0.0;
      }
      zv[q] = z;
      zpp = zp;
      zp = z;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // unit length, largest entry positive (lowest index on ties)
  double amax = -1.0, av = 0.0, s2 = 0.0;
  int ai = n;
  for (int q = lane; q < n; q += 64) {
    const double z = zv[q];
    s2 = fma(z, z, s2);
    if (fabs(z) > amax) {
      amax = fabs(z);
      av = z;
      ai = q;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ou = __shfl_xor(amax, o, 64), ov = __shfl_xor(av, o, 64);
    const int oi = __shfl_xor(ai, o, 64);
    if (ou > amax || (ou == amax && oi < ai)) {
      amax = ou;
      av = ov;
      ai = oi;
    }
  }
  const double nrm2 = wave_sum_d(s2);
  bool bad = !(nrm2 > 0.0) || !isfinite(nrm2);
  const double inv = (av < 0.0 ? -1.0 : 1.0) / sqrt(nrm2);
  for (int q = lane; q < n; q += 64) zv[q] *= inv;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double res = 0.0;
  for (int q = lane; q < n; q += 64) {
    double ri = (d[q] - lambda) * zv[q];
    if (q > 0) ri = fma(e[q - 1], zv[q - 1], ri);
    if (q + 1 < n) ri = fma(e[q], zv[q + 1], ri);
    res = fmax(res, fabs(ri));
    Z[q + (int64_t)oc * n] = zv[q];
  }
  res = wave_max_d(res);
  if (!(res <= 1000.0 * (double)n * LT_EPS * fmax(b.tnorm, DBL_MIN))) bad = true;
  if (lane == 0) {
    w[oc] = lambda;
    if (bad || !isfinite(lambda)) *flag = 1;
  }
}

// ---------------------------------------------------------------- stage 3
// E -= I (K x K), and the largest |entry| as ordered bits into *dev
__global__ void k_lt_eminus(double* __restrict__ E, int K, unsigned long long* __restrict__ dev) {
  double mx = 0.0;
  const int64_t kk = (int64_t)K * K;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < kk;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % K), c = (int)(q / K);
    double x = E[q];
    if (r == c) x -= 1.0;
    E[q] = x;
    mx = fmax(mx, fabs(x));
  }
  mx = wave_max_d(mx);
  if ((threadIdx.x & 63) == 0) {
    if (!isfinite(mx)) mx = DBL_MAX;
    atomicMax(dev, (unsigned long long)__double_as_longlong(mx));
  }
}

__global__ void k_lt_devflag(const unsigned long long* __restrict__ dev, double bound,
                             int* __restrict__ flag) {
  if (threadIdx.x == 0 && __longlong_as_double((long long)*dev) > bound) *flag = 1;
}

// ---------------------------------------------------------------- stage 4
// T factor of every panel (dlarft, forward, columnwise):
// T[:a, a] = -tau_a T[:a, :a] (V[:, :a]^T v_a), T[a, a] = tau_a; one
// workgroup per panel, T (b x b, ld LT_NB) at Tall + p LT_NB^2.
__global__ __launch_bounds__(LT_T) void k_lt_tfac(int n, const double* __restrict__ Rf,
                                                  const double* __restrict__ tri, int m,
                                                  double* __restrict__ Tall) {
  __shared__ double G[LT_NB][LT_NB + 1];
  __shared__ double T[LT_NB][LT_NB + 1];
  const int p = blockIdx.x, j0 = p * LT_NB, bsz = min(LT_NB, m - j0);
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  // G = V^T V over rows j0 + 1 .. n - 1 (one wave per entry, strided)
  for (int q = wv; q < bsz * bsz; q += LT_T / 64) {
    const int a = q % bsz, c = q / bsz;
    double s = 0.0;
    if (a < c) {
      const double* va = Rf + (int64_t)(j0 + a) * n;
      const double* vc = Rf + (int64_t)(j0 + c) * n;
      for (int r = j0 + 1 + lane; r < n; r += 64) s = fma(va[r], vc[r], s);
      s = wave_sum_d(s);
    }
    if (lane == 0) G[a][c] = s;
  }
  __syncthreads();
  if (t == 0) {
    for (int c = 0; c < bsz; ++c) {
      const double tc = tri[2 * n + j0 + c];
      for (int a = 0; a < c; ++a) {
        double s = 0.0;
        for (int q = a; q < c; ++q) s = fma(T[a][q], G[q][c], s);
        T[a][c] = -tc * s;
      }
      T[c][c] = tc;
      for (int a = c + 1; a < bsz; ++a) T[a][c] = 0.0;
    }
  }
  __syncthreads();
  double* To = Tall + (int64_t)p * LT_NB * LT_NB;
  for (int q = t; q < LT_NB * LT_NB; q += LT_T) {
    const int a = q % LT_NB, c = q / LT_NB;
    To[q] = (a < bsz && c < bsz) ? T[a][c] : 0.0;
  }
}

unsigned lt_grid(int64_t count) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 4096));
}

}  // namespace

bool sym_eig_chip(Ctx* ctx, const double* S, int64_t lds, int n, double* w, double* V,
                  int kwant) {
  if (n < 3 || n > LT_NMAX) return false;
  const int K = (kwant > 0 && kwant < n) ? kwant : n;
  const int m = n - 2;  // reflectors
  const int nrb = (n + LT_RB - 1) / LT_RB, ncc = (n + LT_CB - 1) / LT_CB;
  const int npan = (m + LT_NB - 1) / LT_NB;  // back-transformation panels
  DBuf<double> A(ctx, (size_t)n * n), Rf(ctx, (size_t)n * n), Wc(ctx, (size_t)2 * n);
  DBuf<double> tri(ctx, (size_t)3 * n), Yp(ctx, (size_t)ncc * n), xcol(ctx, (size_t)n);
  DBuf<unsigned> counter(ctx, 1);
  DBuf<int> flag(ctx, 1);
  fill_i32(ctx, reinterpret_cast<int32_t*>(counter.get()), 0, 1);
  fill_i32(ctx, flag.get(), 0, 1);
  launch(ctx, "eig_chip_sym", k_lt_sym, dim3(lt_grid((int64_t)n * n)), dim3(256), 0, S, lds, n,
         A.get());
  // ---- stage 1: reflector 0, then one launch per column
  launch(ctx, "eig_chip_start", k_lt_start, dim3(1), dim3(LT_T), (size_t)n * sizeof(double),
         (const double*)A.get(), n, Rf.get(), tri.get());
  for (int j = 0; j < m; ++j)
    launch(ctx, "eig_chip_col", k_lt_col, dim3((unsigned)ncc, (unsigned)nrb), dim3(LC_T),
           (size_t)2 * n * sizeof(double), A.get(), n, j, Rf.get(), Wc.get(), tri.get(),
           Yp.get(), xcol.get(), counter.get());
  launch(ctx, "eig_chip_end", k_lt_end, dim3(1), dim3(256), 0, (const double*)A.get(), n,
         Rf.get(), (const double*)Wc.get(), tri.get());
  // ---- stage 2: the K largest eigenpairs of T
  DBuf<double> Z(ctx, (size_t)n * K), Z2(ctx, (size_t)n * K), E(ctx, (size_t)K * K);
  launch(ctx, "eig_chip_pairs", k_lt_pairs, dim3((unsigned)K), dim3(LT_T),
         (size_t)6 * n * sizeof(double), n, (const double*)tri.get(), w, Z.get(), flag.get());
  // ---- stage 3: Z <- Z - Z (Z^T Z - I) / 2, twice
  DBuf<unsigned long long> dev(ctx, 1);
  double* Zc = Z.get();
  double* Zn = Z2.get();
  for (int pass = 0; pass < 2; ++pass) {
    fill_i32(ctx, reinterpret_cast<int32_t*>(dev.get()), 0, 2);
    gemm_f64(ctx, true, false, K, K, n, 1.0, Zc, n, nullptr, Zc, n, nullptr, 0.0, E.get(), K);
    launch(ctx, "eig_chip_orth", k_lt_eminus, dim3(lt_grid((int64_t)K * K)), dim3(256), 0,
           E.get(), K, dev.get());
    if (pass == 0)
      launch(ctx, "eig_chip_orth", k_lt_devflag, dim3(1), dim3(64), 0,
             (const unsigned long long*)dev.get(), 0.05, flag.get());
    copy_f64(ctx, Zn, Zc, (int64_t)n * K);
    gemm_f64(ctx, false, false, n, K, K, -0.5, Zc, n, nullptr, E.get(), K, nullptr, 1.0, Zn, n);
    std::swap(Zc, Zn);
  }
  // ---- stage 4: V = H_0 ... H_{n-3} Z, panels from the last
  DBuf<double> Tall(ctx, (size_t)npan * LT_NB * LT_NB), M1(ctx, (size_t)LT_NB * K),
      M2(ctx, (size_t)LT_NB * K);
  launch(ctx, "eig_chip_tfac", k_lt_tfac, dim3((unsigned)npan), dim3(LT_T), 0, n,
         (const double*)Rf.get(), (const double*)tri.get(), m, Tall.get());
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * LT_NB, b = std::min(LT_NB, m - j0);
    const int64_t rows = n - j0 - 1;
    const double* Vp = Rf.get() + (j0 + 1) + (int64_t)j0 * n;
    gemm_f64(ctx, true, false, b, K, rows, 1.0, Vp, n, nullptr, Zc + j0 + 1, n, nullptr, 0.0,
             M1.get(), b);
    gemm_f64(ctx, false, false, b, K, b, 1.0, Tall.get() + (int64_t)p * LT_NB * LT_NB, LT_NB,
             nullptr, M1.get(), b, nullptr, 0.0, M2.get(), b);
    gemm_f64(ctx, false, false, rows, K, b, -1.0, Vp, n, nullptr, M2.get(), b, nullptr, 1.0,
             Zc + j0 + 1, n);
  }
  copy_f64(ctx, V, Zc, (int64_t)n * K);
  int32_t hflag = 0;
  to_host_i32(ctx, &hflag, flag.get(), 1);
  return hflag == 0;
}

}  // namespace fpls
