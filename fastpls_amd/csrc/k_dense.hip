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
// The SPD inverse is a symmetric sweep held in registers (one barrier per
// pivot); X = A^{-1} RHS.
// GS: side of the thread grid of the sweep (GS x GS threads, each holding
// (KMAX / GS)^2 entries in registers).
template <int KMAX, int GS>
__global__ __launch_bounds__(GS * GS) void k_spd_solve(const double* __restrict__ Ain, int64_t lda,
                                                    SliceOrders ks, int mode,
                                                    const double* __restrict__ dvals,
                                                    double ridge, double* __restrict__ X,
                                                    int64_t ldx, int64_t slice_stride,
                                                    double* __restrict__ lambda_out,
                                                    int32_t* __restrict__ status) {
  __shared__ double L[KMAX * (KMAX + 1)];
  __shared__ int piv[KMAX];
  __shared__ double shift;
  __shared__ int chol_ok;
  constexpr int LD = KMAX + 1;
  const int s = blockIdx.x;
  const int k = ks.k[s];
  const int t = threadIdx.x, nt = blockDim.x;
  double* Xs = X + (int64_t)s * slice_stride;
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
  // ---- symmetric sweep (Gauss-Jordan without pivoting; Goodnight 1979).
  // Thread (tr, tc) of a 32 x 32 grid holds the lower-triangle entries among
  // (tr + 32a, tc + 32b), a, b < 4, in registers.  Sweeping pivot j:
  //   d = A(j,j); A(j,j) = -1/d; A(i,j) = A(i,j)/d; A(i,c) -= A(i,j) A(c,j)/d,
  // and after all k pivots A = -A_in^{-1}.  Column j of the symmetric matrix
  // is read from a double-buffered LDS snapshot written by its owners (column
  // j below the diagonal, row j left of it) during the previous pivot, so each
  // pivot costs one barrier.  A non-positive pivot is the Cholesky failure (a
  // leading minor is not positive definite) and sends the slice to the LU
  // fallback, as the reference's solve(likely_sympd) does.
  {
    constexpr int G = GS, R = KMAX / GS;  // grid side, register blocks per side
    static_assert(KMAX == 128 && G * R == KMAX && G * G >= 256, "k_spd_solve: grid shape");
    __shared__ double snap[2][128];
    const int tr = t / G, tc = t % G;
    const bool active = t < G * G;
    auto lower = [&](int a, int b) { return a > b || (a == b && tr >= tc); };
    double v[R][R];
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
      for (int b = 0; b < R; ++b) {
        const int i = tr + G * a, c = tc + G * b;
        double x = 0.0;
        if (active && i < k && c < k && lower(a, b)) {
          x = Ain[i + (int64_t)c * lda];
          if (i == c) x += shift;
        }
        v[a][b] = x;
      }
    // snapshot of column jn: owners of (i, jn), i >= jn, and of (jn, c), c < jn
    auto publish = [&](int jn, double* dst) {
      if (!active) return;
      if (tc == jn % G) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) {
            const int i = tr + G * a;
            if (b == jn / G && lower(a, b) && i < k) dst[i] = v[a][b];
          }
      }
      if (tr == jn % G) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) {
            const int c = tc + G * b;
            if (a == jn / G && lower(a, b) && c < jn) dst[c] = v[a][b];
          }
      }
    };
    if (t < 256) (&snap[0][0])[t] = 0.0;  // 2 x 128 entries
    __syncthreads();
    publish(0, snap[0]);
    __syncthreads();
    for (int j = 0; j < k; ++j) {
      const double* cur = snap[j & 1];
      // all snapshot reads issued together (entries past k are never used
      // for outputs, so no bounds test is needed on the 128-entry buffer)
      const double d = cur[j];
      double ns[R], sc[R];
#pragma unroll
      for (int a = 0; a < R; ++a) ns[a] = -cur[(tr + G * a) & 127];
#pragma unroll
      for (int b = 0; b < R; ++b) sc[b] = cur[(tc + G * b) & 127];
      if (!(d > 0.0) || !isfinite(d)) {  // uniform: every thread reads the same pivot
        if (t == 0) chol_ok = 0;
        break;
      }
      const double rd = 1.0 / d;
#pragma unroll
      for (int b = 0; b < R; ++b) sc[b] *= rd;
#pragma unroll
      for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b)
          if (lower(a, b)) v[a][b] = fma(ns[a], sc[b], v[a][b]);
      // row j and column j of the swept matrix
      if (tr == j % G) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) {
            const int c = tc + G * b;
            if (a == j / G && lower(a, b) && c <= j) v[a][b] = (c == j) ? -rd : sc[b];
          }
      }
      if (tc == j % G) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) {
            const int i = tr + G * a;
            if (b == j / G && lower(a, b) && i > j) v[a][b] = -ns[a] * rd;
          }
      }
      if (j + 1 < k) publish(j + 1, snap[(j + 1) & 1]);
      __syncthreads();
    }
    __syncthreads();
    if (chol_ok) {
      // X = A^{-1} diag(rhs) = -A_swept diag(rhs), both triangles from the lower
      if (active) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) {
            const int i = tr + G * a, c = tc + G * b;
            if (!lower(a, b) || i >= k || c >= k) continue;
            const double x = -v[a][b];
            Xs[i + (int64_t)c * ldx] = (mode == 1) ? x : x * dvals[c];
            if (i != c) Xs[c + (int64_t)i * ldx] = (mode == 1) ? x : x * dvals[i];
          }
      }
      if (t == 0 && status) status[s] = 0;
      return;
    }
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
  FPLS_REQUIRE(kmax >= 1, FPLS_ERR_ARG, "spd_solve_slices: order must be positive");
  FPLS_REQUIRE((int)ks.size() >= n_slices, FPLS_ERR_ARG, "spd_solve_slices: missing orders");
  if (kmax > kJacobiMax) {
    // orders above the one-workgroup sweep go through the eigendecomposition
    spd_solve_large(ctx, A, lda, ks, n_slices, mode, dvals, ridge, X, ldx, slice_stride,
                    lambda_out, status);
    return;
  }
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
    launch(ctx, "spd_solve", k_spd_solve<128, 32>, dim3(cnt), dim3(1024), 0, A, lda, so, mode,
           dvals, ridge, Xc, ldx, slice_stride, lc, sc);
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
