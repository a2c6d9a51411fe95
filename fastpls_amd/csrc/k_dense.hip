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
#include "fpls_wave.h"

// Timing hook for stage-level measurements of the slice solve; compiled out
// unless a probe build defines it.
#ifndef FPLS_SPD_STAMP
#define FPLS_SPD_STAMP(slot)
#endif

namespace fpls {

constexpr int JMAX = kJacobiMax;  // 128


constexpr int JLD = JMAX + 1;

// --------------------------------------------------------------- slice solves
// mode 0: (A_k) X = diag(d_k)                     -- PLSSVD latent coefficients
// mode 1: (A_k + lambda I) X = I, lambda from the  -- LDA inverse covariance
//         reference's ridge rule; lambda_out[s] = lambda
// mode 2: (A_k + r I) X = diag(d_k), r = 1e-10 mean(diag A_k)  -- label-aware PLSSVD
// The SPD inverse is a symmetric sweep held in registers (one barrier per
// pivot); X = A^{-1} RHS.
// Blocked symmetric sweep (Gauss-Jordan without pivoting on an SPD matrix;
// Goodnight 1979) over pivot blocks of SPD_B = 16: with D = A_JJ,
//   A_JJ <- -D^{-1},  A_iJ <- A_iJ D^{-1},  A_ij <- A_ij - A_iJ D^{-1} A_Jj,
// and after every block A = -A_in^{-1}.  256 threads hold the (padded)
// 128 x 128 matrix as a 16 x 16 grid of 8 x 8 FP64 register blocks.  Per
// block: the owners publish column panel J to LDS; one wave sweeps the
// 16 x 16 diagonal block in registers with cross-lane reads (no barrier); all
// threads form W = P D^{-1} (128 x 16) and apply the rank-16 update with FP64
// FMAs -- three barriers per block instead of one per pivot.  Orders that are
// not multiples of 16 are padded with identity rows.  A non-positive pivot
// inside a diagonal block is the Cholesky failure (a leading minor is not
// positive definite) and sends the slice to the LU fallback below, as the
// reference's solve(likely_sympd) does.
constexpr int SPD_B = 16, SPD_T = 256, SPD_R = 8, SPD_PLD = SPD_B + 2;  // 16-byte rows

// 1/d from the hardware reciprocal and two Newton steps (a few ulp; d is a
// positive pivot here)
__device__ __forceinline__ double spd_rcp(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// a double of lane `src` (a compile-time lane) in every lane
__device__ __forceinline__ double spd_bcast(double v, int src) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffu), src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// One workgroup per job of the launch (jobs passed by value).
template <int KMAX>
__global__ __launch_bounds__(SPD_T) void k_spd_solve(SpdJobs jobs) {
  static_assert(KMAX == 128, "k_spd_solve: 16 x 16 grid of 8 x 8 blocks");
  const SpdJob& J = jobs.j[blockIdx.x];
  const double* __restrict__ Ain = J.A;
  const int64_t lda = J.lda;
  const int mode = J.mode;
  const double* __restrict__ dvals = J.d;
  const double ridge = J.ridge;
  const int64_t ldx = J.ldx;
  double* __restrict__ lambda_out = J.lambda_out;
  int32_t* __restrict__ status = J.status;
  __shared__ __align__(16) double L[KMAX * (KMAX + 1)];  // LU fallback; panels of the sweep
  __shared__ int piv[KMAX];
  __shared__ double shift;
  __shared__ int chol_ok;
  constexpr int LD = KMAX + 1;
  const int k = J.k;
  const int t = threadIdx.x, nt = blockDim.x;
  double* Xs = J.X;
  // trace of A_k for the shift rules: one diagonal entry per thread, summed in
  // a fixed order (a serial loop of dependent global loads in one thread
  // costs about a microsecond per entry)
  __shared__ double trpart[SPD_T / 64];
  {
    double dv = 0.0;
    for (int i = t; i < k; i += nt) dv += Ain[i + (int64_t)i * lda];
    dv = wave_sum_d(dv);
    if ((t & 63) == 0) trpart[t >> 6] = dv;
  }
  __syncthreads();
  if (t == 0) {
    double sh = 0.0;
    double trace = 0.0;
    for (int w = 0; w < SPD_T / 64; ++w) trace += trpart[w];
    if (mode == 1) {
      const double tr = trace;
      double scale = tr / (double)(k > 1 ? k : 1);
      if (!isfinite(scale) || scale <= 0.0) scale = 1.0;
      double lam = ridge;
      if (!isfinite(lam) || lam < 0.0) lam = 1e-8;
      sh = lam * scale;
      if (lambda_out) *lambda_out = sh;
    } else if (mode == 2) {
      const double tr = trace;
      double r = 1e-10 * (tr / (double)k);
      if (!isfinite(r) || r <= 0.0) r = 1e-10;
      sh = r;
    }
    shift = sh;
    chol_ok = 1;
  }
  __syncthreads();
  FPLS_SPD_STAMP(0);
  {
    const int tr = t >> 4, tc = t & 15;
    const int kp = (k + SPD_B - 1) / SPD_B * SPD_B;  // padded order
    const bool live = tr * SPD_R < kp && tc * SPD_R < kp;
    double a[SPD_R][SPD_R];
#pragma unroll
    for (int r = 0; r < SPD_R; ++r)
#pragma unroll
      for (int c = 0; c < SPD_R; ++c) {
        const int i = tr * SPD_R + r, j = tc * SPD_R + c;
        double x = (i == j) ? 1.0 : 0.0;
        if (i < k && j < k) {
          x = Ain[i + (int64_t)j * lda];
          if (i == j) x += shift;
        }
        a[r][c] = x;
      }
    double* P = L;                         // column panel, 128 x SPD_PLD
    double* W = L + KMAX * SPD_PLD;        // P D^{-1}, 128 x SPD_PLD
    double* Dm = L + 2 * KMAX * SPD_PLD;   // swept diagonal block, 16 x SPD_PLD
    FPLS_SPD_STAMP(1);
    for (int J = 0; J < kp / SPD_B; ++J) {
      const int j0 = J * SPD_B;
      if ((tc >> 1) == J) {
        const int q0 = (tc & 1) * SPD_R;
#pragma unroll
        for (int r = 0; r < SPD_R; ++r)
#pragma unroll
          for (int c = 0; c < SPD_R; ++c) P[(tr * SPD_R + r) * SPD_PLD + q0 + c] = a[r][c];
      }
      __syncthreads();
      if (t < 64) {
        // lane (i = t % 16, cg = t / 16) holds D[i][4 cg .. 4 cg + 3]; per pivot
        // the pivot comes by readlane, the pivot row's entries from lane p of
        // the same 16-lane segment and D[i][p] from lane i of segment p / 4
        const int i = t & 15, cg = t >> 4;
        double d[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) d[e] = P[(j0 + i) * SPD_PLD + 4 * cg + e];
        bool ok = true;
#pragma unroll
        for (int p = 0; p < SPD_B; ++p) {
          const double pv = spd_bcast(d[p & 3], p + 16 * (p >> 2));
          double rowp[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) rowp[e] = __shfl(d[e], p, 16);
          const double colp = __shfl(d[p & 3], i + 16 * (p >> 2), 64);
          if (!(pv > 0.0) || !isfinite(pv)) {  // uniform
            ok = false;
            break;
          }
          const double rd = spd_rcp(pv);
          const double f = -colp * rd;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int j = 4 * cg + e;
            const double upd = fma(f, rowp[e], d[e]);
            const double rowv = (j == p) ? -rd : rowp[e] * rd;
            const double colv = colp * rd;
            d[e] = (i == p) ? rowv : ((j == p) ? colv : upd);
          }
        }
        if (!ok && t == 0) chol_ok = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) Dm[i * SPD_PLD + 4 * cg + e] = d[e];
      }
      __syncthreads();
      FPLS_SPD_STAMP(2);
      if (!chol_ok) break;
      {
        // W = P D^{-1} = -P Dm: thread -> row t / 2, 8 columns at 8 (t % 2)
        const int i = t >> 1, q0 = (t & 1) * SPD_R;
        double w[SPD_R];
#pragma unroll
        for (int c = 0; c < SPD_R; ++c) w[c] = 0.0;
#pragma unroll
        for (int q = 0; q < SPD_B; ++q) {
          const double pv = P[i * SPD_PLD + q];
#pragma unroll
          for (int c = 0; c < SPD_R; ++c) w[c] = fma(pv, Dm[q * SPD_PLD + q0 + c], w[c]);
        }
#pragma unroll
        for (int c = 0; c < SPD_R; ++c) W[i * SPD_PLD + q0 + c] = -w[c];
      }
      __syncthreads();
      FPLS_SPD_STAMP(3);
      if (live) {
        const bool rowJ = (tr >> 1) == J, colJ = (tc >> 1) == J;
        if (rowJ && colJ) {
#pragma unroll
          for (int r = 0; r < SPD_R; ++r)
#pragma unroll
            for (int c = 0; c < SPD_R; ++c)
              a[r][c] = Dm[((tr & 1) * SPD_R + r) * SPD_PLD + (tc & 1) * SPD_R + c];
        } else if (colJ) {
#pragma unroll
          for (int r = 0; r < SPD_R; ++r)
#pragma unroll
            for (int c = 0; c < SPD_R; ++c)
              a[r][c] = W[(tr * SPD_R + r) * SPD_PLD + (tc & 1) * SPD_R + c];
        } else if (rowJ) {
#pragma unroll
          for (int r = 0; r < SPD_R; ++r)
#pragma unroll
            for (int c = 0; c < SPD_R; ++c)
              a[r][c] = W[(tc * SPD_R + c) * SPD_PLD + (tr & 1) * SPD_R + r];
        } else {
#pragma unroll 2
          for (int q = 0; q < SPD_B; q += 2) {
            double2 wv[SPD_R], pv[SPD_R];
#pragma unroll
            for (int r = 0; r < SPD_R; ++r)
              wv[r] = *reinterpret_cast<const double2*>(W + (tr * SPD_R + r) * SPD_PLD + q);
#pragma unroll
            for (int c = 0; c < SPD_R; ++c)
              pv[c] = *reinterpret_cast<const double2*>(P + (tc * SPD_R + c) * SPD_PLD + q);
#pragma unroll
            for (int r = 0; r < SPD_R; ++r)
#pragma unroll
              for (int c = 0; c < SPD_R; ++c)
                a[r][c] = fma(-wv[r].y, pv[c].y, fma(-wv[r].x, pv[c].x, a[r][c]));
          }
        }
      }
      __syncthreads();  // the next panel overwrites P, W and Dm
      FPLS_SPD_STAMP(4);
    }
    if (chol_ok) {
      // X = A^{-1} diag(rhs) = -A_swept diag(rhs): both triangles from the
      // lower one, staged in LDS so that the global writes are coalesced
      if (live && tr >= tc) {
#pragma unroll
        for (int r = 0; r < SPD_R; ++r)
#pragma unroll
          for (int c = 0; c < SPD_R; ++c) {
            const int i = tr * SPD_R + r, j = tc * SPD_R + c;
            if (i < j || i >= k || j >= k) continue;
            L[i + j * LD] = -a[r][c];
            L[j + i * LD] = -a[r][c];
          }
      }
      __syncthreads();
      for (int idx = t; idx < k * k; idx += nt) {
        const int i = idx % k, j = idx / k;
        const double x = L[i + j * LD];
        Xs[i + (int64_t)j * ldx] = (mode == 1) ? x : x * dvals[j];
      }
      FPLS_SPD_STAMP(5);
      if (t == 0 && status) *status = 0;
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
    if (t == 0 && status) *status = 1;
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
  if (t == 0 && status) *status = 2;  // solved through the LU fallback
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
  std::vector<SpdJob> jobs;
  spd_slice_jobs(A, lda, kmax, ks, n_slices, mode, dvals, ridge, X, ldx, slice_stride,
                 lambda_out, status, jobs);
  spd_solve_jobs(ctx, jobs);
}

void spd_slice_jobs(const double* A, int64_t lda, int kmax, const std::vector<int>& ks,
                    int n_slices, int mode, const double* dvals, double ridge, double* X,
                    int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status,
                    std::vector<SpdJob>& jobs) {
  FPLS_REQUIRE(kmax >= 1 && kmax <= kJacobiMax, FPLS_ERR_ARG,
               "spd_slice_jobs: order must be in 1..128");
  FPLS_REQUIRE((int)ks.size() >= n_slices, FPLS_ERR_ARG, "spd_slice_jobs: missing orders");
  for (int s = 0; s < n_slices; ++s) {
    FPLS_REQUIRE(ks[s] >= 1 && ks[s] <= kmax, FPLS_ERR_ARG,
                 "spd_solve_slices: order out of range");
    jobs.push_back(SpdJob{A, lda, ks[s], mode, dvals, ridge, X + (int64_t)s * slice_stride, ldx,
                          lambda_out ? lambda_out + s : nullptr, status ? status + s : nullptr});
  }
}

void spd_solve_jobs(Ctx* ctx, const std::vector<SpdJob>& jobs) {
  // jobs travel in the kernel arguments, kSpdJobMax per launch
  for (size_t c0 = 0; c0 < jobs.size(); c0 += kSpdJobMax) {
    const int cnt = (int)std::min<size_t>(kSpdJobMax, jobs.size() - c0);
    SpdJobs js{};
    for (int i = 0; i < cnt; ++i) {
      FPLS_REQUIRE(jobs[c0 + i].k >= 1 && jobs[c0 + i].k <= JMAX, FPLS_ERR_ARG,
                   "spd_solve_jobs: order out of range");
      js.j[i] = jobs[c0 + i];
    }
    launch(ctx, "spd_solve", k_spd_solve<128>, dim3(cnt), dim3(SPD_T), 0, js);
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
  __shared__ double trpart[1024 / 64];
  const int t = threadIdx.x, nt = blockDim.x;
  // trace: one diagonal entry per thread, fixed-order sum of the wave sums
  {
    double dv = 0.0;
    for (int i = t; i < l; i += nt) dv += G[i + (int64_t)i * ldg];
    dv = wave_sum_d(dv);
    if ((t & 63) == 0) trpart[t >> 6] = dv;
  }
  __syncthreads();
  double tr = 0.0;
  for (int w = 0; w < (nt >> 6); ++w) tr += trpart[w];
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
  // L^{-1} in place, columns from the right: with columns > j already
  // inverted, Linv(i, j) = -(sum_{j < q <= i} Linv(i, q) L(q, j)) / L(j, j)
  // for i > j (one row per thread) and Linv(j, j) = 1 / L(j, j)
  __syncthreads();
  for (int j = l - 1; j >= 0; --j) {
    const double rjj = 1.0 / L[j + j * JLD];
    double v = 0.0;
    const int i = j + 1 + t;
    if (i < l) {
      for (int q = j + 1; q <= i; ++q) v = fma(L[i + q * JLD], L[q + j * JLD], v);
    }
    __syncthreads();  // every read of column j is done
    if (i < l) L[i + j * JLD] = -v * rjj;
    if (t == 0) L[j + j * JLD] = rjj;
    __syncthreads();
  }
  for (int idx = t; idx < l * l; idx += nt) {
    const int i = idx % l, c = idx / l;
    Linv[i + (int64_t)c * l] = i >= c ? L[i + c * JLD] : 0.0;
  }
}

void chol_inv_lower(Ctx* ctx, const double* G, int64_t ldg, int l, int64_t rows, double shift,
                    double* Linv, int32_t* status) {
  FPLS_REQUIRE(l >= 1 && l <= kJacobiMax, FPLS_ERR_ARG, "chol_inv_lower: order must be 1..128");
  launch(ctx, "chol_inv", k_chol_inv_lower, dim3(1), dim3(1024), 0, G, ldg, l, rows, shift, Linv,
         status);
}

}  // namespace fpls
