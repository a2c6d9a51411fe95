// Small FP64 factorisations kept on the GPU (one workgroup each, matrix in LDS;
// the symmetric eigensolver lives in k_eig.hip):
//
//  * spd_solve_slices -- per ncomp slice: Cholesky of the leading block (with
//                        the reference's diagonal shift rules) and solves
//                        against a diagonal / identity right-hand side; LU with
//                        partial pivoting when the block is not SPD.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "fpls_internal.h"

namespace fpls {

constexpr int JMAX = kJacobiMax;  // 128

// orders of the ncomp slices of one launch, passed by value
constexpr int kSliceMax = 64;
struct SliceOrders {
  int k[kSliceMax];
};
constexpr int JLD = JMAX + 1;

// --------------------------------------------------------------- slice solves
// mode 0: (A_k) X = diag(d_k)                     -- PLSSVD latent coefficients
// mode 1: (A_k + lambda I) X = I, lambda from the  -- LDA inverse covariance
//         reference's ridge rule; lambda_out[s] = lambda
// mode 2: (A_k + r I) X = diag(d_k), r = 1e-10 mean(diag A_k)  -- label-aware PLSSVD
// Cholesky and the inverse of the factor each take one barrier per step: the
// trailing update of step j uses column j unscaled (A -= c c^T / d), and the
// scaling of column j by 1/sqrt(d) is deferred to step j+1, when no thread
// reads it any more.  The solution is then X = Z^T Z RHS with Z = L^{-1}.
template <int KMAX, bool Z_IN_LDS>
__global__ __launch_bounds__(1024) void k_spd_solve(const double* __restrict__ Ain, int64_t lda,
                                                    SliceOrders ks, int mode,
                                                    const double* __restrict__ dvals,
                                                    double ridge, double* __restrict__ X,
                                                    int64_t ldx, int64_t slice_stride,
                                                    double* __restrict__ lambda_out,
                                                    int32_t* __restrict__ status,
                                                    double* __restrict__ Zg) {
  __shared__ double L[KMAX * (KMAX + 1)];
  __shared__ double sZ[Z_IN_LDS ? KMAX * (KMAX + 1) : 1];
  __shared__ int piv[KMAX];
  __shared__ double shift;
  __shared__ int chol_ok;
  constexpr int LD = KMAX + 1;
  const int s = blockIdx.x;
  const int k = ks.k[s];
  const int t = threadIdx.x, nt = blockDim.x;
  double* Xs = X + (int64_t)s * slice_stride;
  double* Z = Z_IN_LDS ? sZ : Zg + (int64_t)s * KMAX * LD;
  if (t == 0) {
    double sh = 0.0;
    if (mode == 1) {
      double tr = 0.0;
      for (int i = 0; i < k; ++i) tr += Ain[i + (int64_t)i * lda];
      double scale = tr / (double)(k > 1 ? k : 1);
      if (!isfinite(scale) || scale <= 0.0) scale = 1.0;
      double lam = ridge;
      if (!isfinite(lam) || lam < 0.0) lam = 1e-8;
      sh = lam * scale;
      if (lambda_out) lambda_out[s] = sh;
    } else if (mode == 2) {
      double tr = 0.0;
      for (int i = 0; i < k; ++i) tr += Ain[i + (int64_t)i * lda];
      double r = 1e-10 * (tr / (double)k);
      if (!isfinite(r) || r <= 0.0) r = 1e-10;
      sh = r;
    }
    shift = sh;
    chol_ok = 1;
  }
  __syncthreads();
  // 2-D thread map (no integer division in the loops): tx runs down a column
  const int tx = t & 31, ty = t >> 5;
  for (int j = ty; j < k; j += 32) {
    for (int i = tx; i < k; i += 32) {
      double v = Ain[i + (int64_t)j * lda];
      if (i == j) v += shift;
      L[i + j * LD] = v;
      Z[i + j * LD] = (i == j) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  // ---- Cholesky, lower triangle, one barrier per column
  double prev_d = 1.0;  // pivot of the previous column (same in every thread)
  for (int j = 0; j < k; ++j) {
    const double d = L[j + j * LD];
    if (!(d > 0.0) || !isfinite(d)) {
      if (t == 0) chol_ok = 0;
      break;
    }
    if (j > 0) {  // deferred scaling of column j-1
      const double r = 1.0 / sqrt(prev_d);
      for (int i = j - 1 + t; i < k; i += nt) L[i + (j - 1) * LD] *= r;
    }
    prev_d = d;
    const double inv_d = 1.0 / d;
    for (int jj = j + 1 + ty; jj < k; jj += 32) {
      const double cj = L[jj + j * LD] * inv_d;
      for (int ii = jj + tx; ii < k; ii += 32) L[ii + jj * LD] -= L[ii + j * LD] * cj;
    }
    __syncthreads();
  }
  __syncthreads();
  if (chol_ok) {
    if (t == 0) L[(k - 1) + (k - 1) * LD] = sqrt(L[(k - 1) + (k - 1) * LD]);
    __syncthreads();
    // ---- Z = L^{-1} by a column sweep, one barrier per step
    for (int q = 0; q < k; ++q) {
      const double lqq = L[q + q * LD];
      if (q > 0) {  // deferred scaling of row q-1
        const double r = 1.0 / L[(q - 1) + (q - 1) * LD];
        for (int c = t; c < q; c += nt) Z[(q - 1) + c * LD] *= r;
      }
      const double ilqq = 1.0 / lqq;
      for (int c = ty; c <= q; c += 32) {
        const double zq = Z[q + c * LD] * ilqq;
        for (int i = q + 1 + tx; i < k; i += 32) Z[i + c * LD] -= L[i + q * LD] * zq;
      }
      __syncthreads();
    }
    {
      const double r = 1.0 / L[(k - 1) + (k - 1) * LD];
      for (int c = t; c < k; c += nt) Z[(k - 1) + c * LD] *= r;
    }
    __syncthreads();
    // ---- X = Z^T Z diag(rhs), Z lower triangular.  Each wave owns a 32 x 32
    // output block as an 8 x 8 grid of 4 x 4 register tiles, so a step over r
    // reads 64 distinct Z values per wave (the rest are LDS broadcasts).
    {
      const int wv = t >> 6, ln = t & 63;
      const int i0 = (wv & 3) * 32 + (ln & 7) * 4, j0 = (wv >> 2) * 32 + (ln >> 3) * 4;
      const int wi0 = (wv & 3) * 32, wj0 = (wv >> 2) * 32;
      if (wi0 < k && wj0 < k) {
        double acc[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
        // rows below max(i, j) vanish in both factors
        for (int r = wi0 > wj0 ? wi0 : wj0; r < k; ++r) {
          double zi[4], zj[4];
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            zi[x] = i0 + x < k ? Z[r + (i0 + x) * LD] : 0.0;
            zj[x] = j0 + x < k ? Z[r + (j0 + x) * LD] : 0.0;
          }
#pragma unroll
          for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = fma(zi[x], zj[y], acc[x][y]);
        }
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          const int j = j0 + y;
          if (j >= k) continue;
          const double sc = (mode == 1) ? 1.0 : dvals[j];
#pragma unroll
          for (int x = 0; x < 4; ++x)
            if (i0 + x < k) Xs[(i0 + x) + (int64_t)j * ldx] = acc[x][y] * sc;
        }
      }
    }
    if (t == 0 && status) status[s] = 0;
    return;
  }
  // LU with partial pivoting (Doolittle, in place)
  __syncthreads();
  for (int idx = t; idx < k * k; idx += nt) {
    const int i = idx % k, j = idx / k;
    double v = Ain[i + (int64_t)j * lda];
    if (i == j) v += shift;
    L[i + j * LD] = v;
  }
  __shared__ int singular;
  if (t == 0) singular = 0;
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (t == 0) {
      int pr = j;
      double best = fabs(L[j + j * LD]);
      for (int i = j + 1; i < k; ++i) {
        const double v = fabs(L[i + j * LD]);
        if (v > best) { best = v; pr = i; }
      }
      piv[j] = pr;
      if (best == 0.0 || !isfinite(best)) singular = 1;
    }
    __syncthreads();
    if (singular) break;
    const int pr = piv[j];
    if (pr != j) {
      for (int c = t; c < k; c += nt) {
        const double a = L[j + c * LD];
        L[j + c * LD] = L[pr + c * LD];
        L[pr + c * LD] = a;
      }
    }
    __syncthreads();
    const double d = L[j + j * LD];
    for (int i = j + 1 + t; i < k; i += nt) L[i + j * LD] /= d;
    __syncthreads();
    const int m = k - j - 1;
    for (int idx = t; idx < m * m; idx += nt) {
      const int ii = j + 1 + idx % m, jj = j + 1 + idx / m;
      L[ii + jj * LD] -= L[ii + j * LD] * L[j + jj * LD];
    }
    __syncthreads();
  }
  if (singular) {
    if (t == 0 && status) status[s] = 1;
    return;
  }
  for (int c = t; c < k; c += nt) {
    double* x = Xs + (int64_t)c * ldx;
    for (int i = 0; i < k; ++i) x[i] = 0.0;
    x[c] = (mode == 1) ? 1.0 : dvals[c];
    for (int j = 0; j < k; ++j) {  // apply row swaps
      const int pr = piv[j];
      if (pr != j) { const double a = x[j]; x[j] = x[pr]; x[pr] = a; }
    }
    for (int i = 0; i < k; ++i) {
      double v = x[i];
      for (int q = 0; q < i; ++q) v -= L[i + q * LD] * x[q];
      x[i] = v;
    }
    for (int i = k - 1; i >= 0; --i) {
      double v = x[i];
      for (int q = i + 1; q < k; ++q) v -= L[i + q * LD] * x[q];
      x[i] = v / L[i + i * LD];
    }
  }
  if (t == 0 && status) status[s] = 2;  // solved through the LU fallback
}

void spd_solve_slices(Ctx* ctx, const double* A, int64_t lda, int kmax, const std::vector<int>& ks,
                      int n_slices, int mode, const double* dvals, double ridge, double* X,
                      int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status) {
  FPLS_REQUIRE(kmax >= 1 && kmax <= kJacobiMax, FPLS_ERR_ARG,
               "spd_solve_slices: order must be in 1..128");
  FPLS_REQUIRE((int)ks.size() >= n_slices, FPLS_ERR_ARG, "spd_solve_slices: missing orders");
  // slice orders travel in the kernel arguments, kSliceMax per launch
  for (int c0 = 0; c0 < n_slices; c0 += kSliceMax) {
    const int cnt = std::min(kSliceMax, n_slices - c0);
    SliceOrders so{};
    for (int i = 0; i < cnt; ++i) {
      FPLS_REQUIRE(ks[c0 + i] >= 1 && ks[c0 + i] <= kmax, FPLS_ERR_ARG,
                   "spd_solve_slices: order out of range");
      so.k[i] = ks[c0 + i];
    }
    double* Xc = X + (int64_t)c0 * slice_stride;
    double* lc = lambda_out ? lambda_out + c0 : nullptr;
    int32_t* sc = status ? status + c0 : nullptr;
    if (kmax <= 100) {
      launch(ctx, "spd_solve", k_spd_solve<100, true>, dim3(cnt), dim3(1024), 0, A, lda, so,
             mode, dvals, ridge, Xc, ldx, slice_stride, lc, sc, (double*)nullptr);
    } else {
      DBuf<double> Zg(ctx, (size_t)cnt * 128 * 129);
      launch(ctx, "spd_solve", k_spd_solve<128, false>, dim3(cnt), dim3(1024), 0, A, lda, so,
             mode, dvals, ridge, Xc, ldx, slice_stride, lc, sc, Zg.get());
    }
  }
}

// --------------------------------------------------------------- column scaling
// X[:, j] *= f[j] (or /= f[j] when divide; zero divisors leave the column).
__global__ void k_scale_cols(double* X, int64_t ld, int64_t rows, int64_t cols,
                             const double* f, int divide) {
  const int64_t n = rows * cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    const double g = f[j];
    if (divide) {
      if (g != 0.0) X[i + j * ld] /= g;
    } else {
      X[i + j * ld] *= g;
    }
  }
}

void scale_cols(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols, const double* f,
                bool divide) {
  if (rows <= 0 || cols <= 0) return;
  const int64_t n = rows * cols;
  launch(ctx, "scale_cols", k_scale_cols, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
         dim3(256), 0, X, ld, rows, cols, f, divide ? 1 : 0);
}

// X[i, :] /= f[i] (row scaling; zero / non-finite f treated as 1)
__global__ void k_scale_rows(double* X, int64_t ld, int64_t rows, int64_t cols, const double* f) {
  const int64_t n = rows * cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    double g = f[i];
    if (!isfinite(g) || g == 0.0) g = 1.0;
    X[i + j * ld] /= g;
  }
}

void scale_rows_div(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols,
                    const double* f) {
  if (rows <= 0 || cols <= 0) return;
  const int64_t n = rows * cols;
  launch(ctx, "scale_rows", k_scale_rows, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
         dim3(256), 0, X, ld, rows, cols, f);
}

}  // namespace fpls

namespace fpls {

// --------------------------------------------------------------- CholeskyQR step
// Linv = L^{-1} where L L^T = G + shift I (G: l x l SPD Gram).  When the
// factorisation breaks down the shift is raised (x100, up to 6 times) starting
// from the shifted-CholeskyQR bound 11 (rows*l + l(l+1)) u tr(G) (Fukaya et al.,
// SIAM J. Sci. Comput. 2020), so a numerically rank-deficient sample still
// yields a usable orthogonalisation; status[0] records the retries.
__global__ __launch_bounds__(1024) void k_chol_inv_lower(const double* __restrict__ G, int64_t ldg,
                                                         int l, int64_t rows, double shift0,
                                                         double* __restrict__ Linv,
                                                         int32_t* __restrict__ status) {
  __shared__ double L[JMAX * JLD];
  __shared__ int ok;
  __shared__ double shift;
  const int t = threadIdx.x, nt = blockDim.x;
  double tr = 0.0;
  for (int i = 0; i < l; ++i) tr += G[i + (int64_t)i * ldg];
  const double u = DBL_EPSILON * 0.5;
  double boost = 11.0 * ((double)rows * l + (double)l * (l + 1)) * u * (tr > 0.0 ? tr : 1.0);
  if (shift0 < 0.0) shift0 = boost;  // shifted first pass of CholeskyQR3
  int attempt = 0;
  for (;;) {
    if (t == 0) {
      shift = attempt == 0 ? shift0 : boost;
      ok = 1;
    }
    __syncthreads();
    for (int idx = t; idx < l * l; idx += nt) {
      const int i = idx % l, j = idx / l;
      double v = G[i + (int64_t)j * ldg];
      if (i == j) v += shift;
      L[i + j * JLD] = v;
    }
    __syncthreads();
    for (int j = 0; j < l; ++j) {
      if (t == 0) {
        const double d = L[j + j * JLD];
        if (!(d > 0.0) || !isfinite(d)) ok = 0;
        else L[j + j * JLD] = sqrt(d);
      }
      __syncthreads();
      if (!ok) break;
      const double djj = L[j + j * JLD];
      for (int i = j + 1 + t; i < l; i += nt) L[i + j * JLD] /= djj;
      __syncthreads();
      const int m = l - j - 1;
      for (int idx = t; idx < m * m; idx += nt) {
        const int ii = j + 1 + idx % m, jj = j + 1 + idx / m;
        if (ii >= jj) L[ii + jj * JLD] -= L[ii + j * JLD] * L[jj + j * JLD];
      }
      __syncthreads();
    }
    if (ok || attempt >= 6) break;
    if (attempt > 0) boost *= 100.0;
    ++attempt;
    __syncthreads();
  }
  if (t == 0 && status) status[0] = ok ? attempt : -1;
  // forward substitution, one column of the identity per thread
  for (int c = t; c < l; c += nt) {
    double* x = Linv + (int64_t)c * l;
    for (int i = 0; i < c; ++i) x[i] = 0.0;
    for (int i = c; i < l; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int q = c; q < i; ++q) v -= L[i + q * JLD] * x[q];
      x[i] = v / L[i + i * JLD];
    }
  }
}

void chol_inv_lower(Ctx* ctx, const double* G, int64_t ldg, int l, int64_t rows, double shift,
                    double* Linv, int32_t* status) {
  FPLS_REQUIRE(l >= 1 && l <= kJacobiMax, FPLS_ERR_ARG, "chol_inv_lower: order must be 1..128");
  launch(ctx, "chol_inv", k_chol_inv_lower, dim3(1), dim3(1024), 0, G, ldg, l, rows, shift, Linv,
         status);
}

}  // namespace fpls
