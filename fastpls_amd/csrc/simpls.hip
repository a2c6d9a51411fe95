// SIMPLS-fast fit on the device: pls_model2_fast (src/fastPLS.cpp:2454-2778)
// under the options .with_fastpls_fast_options sets (R/main.R:55-80): refresh
// block 1, inc_iters power iterations of S S^T on the warm start, no
// t-centring, single Gram-Schmidt pass against V, deflation S -= v (v^T S).
//
// Per component (all FP64 vectors, X streamed from HBM twice):
//   refresh   y <- S (S^T y) x inc_iters, rr = y / |y|          (p x m, tiny)
//   append    t = Xs rr, |t|, p = Xs^T t, q = Yc^T t             (2 passes over X)
//             v = p - V (V^T p), |v|, S -= v (v^T S), B += rr q^T
// Only component 0 draws from the Gaussian stream (seed): later components
// start from the previous rr (src/fastPLS.cpp:1124-1131).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "fpls_wave.h"
#include "model.h"

namespace fpls {

void fit_preprocess(Ctx* ctx, Model* M, const DMat& X, const DMat* Y, const int32_t* y, int32_t C,
                    int scaling, bool guard_sd, bool means_later = false);
void cross_cov(Ctx* ctx, const Model* M, const DMat& X, const DMat* Y, const int32_t* y, int64_t m,
               double* S, int precision, bool means_later = false, bool guard_sd = false);
bool store_b_policy(int store_b, int64_t p, int64_t m, int n_slices);
DBuf<double> centered_response(Ctx* ctx, const Model* M, const DMat* Y, const int32_t* y,
                               int64_t n, int64_t m);
void r2_into(Ctx* ctx, const double* Yc, int64_t n, int64_t m, const double* Yhat, double* out);
void add_row(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols, const double* v);

// ---------------------------------------------------------------- device-side steps
// The component loop runs without host round trips: the stop tests of
// append_component (|t| or |v| zero or not finite, src/fastPLS.cpp:2580-2620)
// set flags[0] on the device, every later write of the fit (R, Q, V, B,
// deflation, Yfit) is gated on it, and flags[1] counts the components kept.
// Norms are scaled (max |x| first), so tiny or huge vectors neither
// underflow nor overflow where the reference's QR would not.
constexpr int SP_NT = 1024;

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < SP_NT / 64; ++w) t += red[w];
  return t;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < SP_NT / 64; ++w) t = fmax(t, red[w]);
  return t;
}

// ||x|| scaled by max |x|; NaN when x holds a non-finite value.  x may be
// global or LDS memory; every thread returns the same value.
__device__ double block_norm(const double* x, int64_t n, double* red) {
  double amax = 0.0, bad = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += SP_NT) {
    const double a = fabs(x[i]);
    if (!(a <= DBL_MAX)) bad = 1.0;
    else if (a > amax) amax = a;
  }
  bad = block_max(bad, red);
  amax = block_max(amax, red);
  if (bad != 0.0) return __builtin_nan("");
  if (amax == 0.0) return 0.0;
  const double inv = 1.0 / amax;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += SP_NT) {
    const double u = x[i] * inv;
    s = fma(u, u, s);
  }
  return amax * sqrt(block_sum(s, red));
}

// ---- chip-wide normalisation for the large path: the scaled norm of
// x = src - sub combined from per-chunk partials (max |x|, sum (x / max)^2)
// as in a two-level dnrm2, then dst = x / |x| (and also /= |x|).
constexpr int SP_NB = 256, SP_NE = 8;
constexpr int64_t SP_NCH = (int64_t)SP_NB * SP_NE;  // elements per partial

template <int NT>
__device__ __forceinline__ double blk_reduce(double v, double* red, bool is_max) {
  v = is_max ? wave_max_d(v) : wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) t = is_max ? fmax(t, red[w]) : t + red[w];
  return t;
}

__global__ __launch_bounds__(SP_NB) void k_sp_norm_part(const double* __restrict__ src,
                                                        const double* __restrict__ sub, int64_t n,
                                                        double2* __restrict__ part,
                                                        const int32_t* __restrict__ flags) {
  __shared__ double red[SP_NB / 64];
  if (flags[0]) return;
  const int64_t b = (int64_t)blockIdx.x * SP_NCH + threadIdx.x;
  double v[SP_NE], amax = 0.0, bad = 0.0;
#pragma unroll
  for (int e = 0; e < SP_NE; ++e) {
    const int64_t i = b + (int64_t)e * SP_NB;
    v[e] = i < n ? src[i] - (sub ? sub[i] : 0.0) : 0.0;
    const double a = fabs(v[e]);
    if (!(a <= DBL_MAX)) bad = 1.0;
    else amax = fmax(amax, a);
  }
  bad = blk_reduce<SP_NB>(bad, red, true);
  amax = blk_reduce<SP_NB>(amax, red, true);
  double ssq = 0.0;
  if (bad == 0.0 && amax > 0.0) {
    const double inv = 1.0 / amax;
#pragma unroll
    for (int e = 0; e < SP_NE; ++e) {
      const double u = v[e] * inv;
      ssq = fma(u, u, ssq);
    }
  }
  ssq = blk_reduce<SP_NB>(ssq, red, false);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(bad != 0.0 ? __builtin_nan("") : amax, ssq);
}

__global__ __launch_bounds__(SP_NB) void k_sp_norm_apply(const double* src, const double* sub,
                                                         double* dst, int64_t n, double* also,
                                                         int64_t n2,
                                                         const double2* __restrict__ part,
                                                         int nparts, int32_t* flags) {
  __shared__ double red[SP_NB / 64];
  if (flags[0]) return;
  double am = 0.0, bad = 0.0;
  for (int q = threadIdx.x; q < nparts; q += SP_NB) {
    const double a = part[q].x;
    if (!(a <= DBL_MAX)) bad = 1.0;
    else am = fmax(am, a);
  }
  bad = blk_reduce<SP_NB>(bad, red, true);
  am = blk_reduce<SP_NB>(am, red, true);
  double ss = 0.0;
  if (bad == 0.0 && am > 0.0)
    for (int q = threadIdx.x; q < nparts; q += SP_NB) {
      const double r = part[q].x / am;
      ss = fma(r * r, part[q].y, ss);
    }
  ss = blk_reduce<SP_NB>(ss, red, false);
  const double nr = bad != 0.0 ? __builtin_nan("") : am * sqrt(ss);
  if (!(nr > 0.0 && nr <= DBL_MAX)) {  // every block computes the same nr
    if (blockIdx.x == 0 && threadIdx.x == 0) flags[0] = 1;
    return;
  }
  const int64_t t0 = (int64_t)blockIdx.x * SP_NB + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * SP_NB;
  for (int64_t i = t0; i < n; i += stride) dst[i] = (src[i] - (sub ? sub[i] : 0.0)) / nr;
  if (also)
    for (int64_t i = t0; i < n2; i += stride) also[i] /= nr;
}

// One-workgroup form of k_sp_norm_part + k_sp_norm_apply for n <= SP_NCH
// (one partial): the same arithmetic, so the same result, in one launch.
__global__ __launch_bounds__(SP_NB) void k_sp_norm_one(const double* src, const double* sub,
                                                       double* dst, int64_t n, double* also,
                                                       int64_t n2, int32_t* flags) {
  __shared__ double red[SP_NB / 64];
  if (flags[0]) return;
  double v[SP_NE], amax = 0.0, bad = 0.0;
#pragma unroll
  for (int e = 0; e < SP_NE; ++e) {
    const int64_t i = threadIdx.x + (int64_t)e * SP_NB;
    v[e] = i < n ? src[i] - (sub ? sub[i] : 0.0) : 0.0;
    const double a = fabs(v[e]);
    if (!(a <= DBL_MAX)) bad = 1.0;
    else amax = fmax(amax, a);
  }
  bad = blk_reduce<SP_NB>(bad, red, true);
  amax = blk_reduce<SP_NB>(amax, red, true);
  double ssq = 0.0;
  if (bad == 0.0 && amax > 0.0) {
    const double inv = 1.0 / amax;
#pragma unroll
    for (int e = 0; e < SP_NE; ++e) {
      const double u = v[e] * inv;
      ssq = fma(u, u, ssq);
    }
  }
  ssq = blk_reduce<SP_NB>(ssq, red, false);
  // k_sp_norm_apply with one partial: r = amax / amax = 1, ss = fma(1, ssq, 0)
  const double nr = bad != 0.0 ? __builtin_nan("") : amax * sqrt(ssq);
  if (!(nr > 0.0 && nr <= DBL_MAX)) {
    if (threadIdx.x == 0) flags[0] = 1;
    return;
  }
#pragma unroll
  for (int e = 0; e < SP_NE; ++e) {
    const int64_t i = threadIdx.x + (int64_t)e * SP_NB;
    if (i < n) dst[i] = v[e] / nr;
  }
  if (also)
    for (int64_t i = threadIdx.x; i < n2; i += SP_NB) also[i] /= nr;
}

// dst = (src - sub) / |src - sub| (sub may be NULL; dst may be src), also /=
// the same norm; a zero or non-finite norm stops the fit.  part holds
// ceil(n / SP_NCH) partials.
static void sp_normalize(Ctx* c, const double* src, const double* sub, double* dst, int64_t n,
                         double* also, int64_t n2, double2* part, int32_t* flags) {
  const int nparts = (int)((n + SP_NCH - 1) / SP_NCH);
  if (nparts == 1 && (!also || n2 <= 16 * SP_NCH)) {
    launch(c, "simpls_norm", k_sp_norm_one, dim3(1), dim3(SP_NB), 0, src, sub, dst, n, also, n2,
           flags);
    return;
  }
  launch(c, "simpls_norm", k_sp_norm_part, dim3(nparts), dim3(SP_NB), 0, src, sub, n, part,
         (const int32_t*)flags);
  const int64_t work = std::max(n, also ? n2 : 0);
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((work + 1023) / 1024, 1024));
  launch(c, "simpls_norm", k_sp_norm_apply, dim3(g), dim3(SP_NB), 0, src, sub, dst, n, also, n2,
         (const double2*)part, nparts, flags);
}

// X^T X branch: tn = sqrt(r . pp) from the device scalar, r /= tn, pp /= tn.
__global__ __launch_bounds__(SP_NT) void k_sp_xtx_scale(const double* tn2p, double* r, double* pp,
                                                        int64_t p, int32_t* flags) {
  if (flags[0]) return;
  const double tn2 = *tn2p;
  if (!(tn2 > 0.0 && tn2 <= DBL_MAX)) {
    if (threadIdx.x == 0) flags[0] = 1;
    return;
  }
  const double tn = sqrt(tn2);
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) {
    r[i] /= tn;
    pp[i] /= tn;
  }
}

// S += alpha u w^T (S: rows x cols, ld), unless stopped.
__global__ void k_sp_rank1(double* S, int64_t ld, int64_t rows, int64_t cols, const double* u,
                           const double* w, double alpha, const int32_t* flags) {
  if (flags[0]) return;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < rows * cols;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    S[i + j * ld] += alpha * u[i] * w[j];
  }
}

// Component a kept: R[:, a] = r, Q[:, a] = q, V[:, a] = v, and Bcur += r q^T
// when Bcur is set.
__global__ void k_sp_commit(int a, const double* r, const double* q, const double* v, int64_t p,
                            int64_t m, double* R, double* Q, double* V, double* Bcur,
                            int32_t* flags) {
  if (flags[0]) return;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = t0; i < p; i += stride) {
    R[i + a * p] = r[i];
    V[i + a * p] = v[i];
  }
  for (int64_t j = t0; j < m; j += stride) Q[j + a * m] = q[j];
  if (Bcur)
    for (int64_t e = t0; e < p * m; e += stride) {
      const int64_t i = e % p, j = e / p;
      Bcur[e] += r[i] * q[j];
    }
  if (t0 == 0) flags[1] = a + 1;
}

// dst = src (+ row vector add over an n x m block when given), unless stopped.
__global__ void k_sp_copy(double* dst, const double* src, int64_t n, int64_t m,
                          const double* add_row, const int32_t* flags) {
  if (flags[0]) return;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * m;
       e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[e] + (add_row ? add_row[e / n] : 0.0);
}

// ---- fused one-workgroup steps for small S (p <= SP_PMAX, m, K <= SP_MMAX)
constexpr int SP_PMAX = 8192, SP_MMAX = 256;

// Refresh: y = warm (previous r, or the Gaussian start), inc_iters times
// y <- S (S^T y), then r = y / |y| (src/fastPLS.cpp:1117-1189, k_block = 1).
__global__ __launch_bounds__(SP_NT) void k_sp_refresh_small(const double* S, int64_t p, int64_t m,
                                                            int iters, const double* warm,
                                                            double* r, int32_t* flags) {
  __shared__ double y[SP_PMAX];
  __shared__ double z[SP_MMAX];
  __shared__ double red[SP_NT / 64];
  if (flags[0]) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) y[i] = warm[i];
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int64_t j = w; j < m; j += SP_NT / 64) {
      double acc = 0.0;
      for (int64_t i = lane; i < p; i += 64) acc = fma(S[i + j * p], y[i], acc);
      acc = wave_sum_d(acc);
      if (lane == 0) z[j] = acc;
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < p; i += SP_NT) {
      double acc = 0.0;
      for (int64_t j = 0; j < m; ++j) acc = fma(S[i + j * p], z[j], acc);
      y[i] = acc;
    }
    __syncthreads();
  }
  const double nr = block_norm(y, p, red);
  if (!(nr > 0.0 && nr <= DBL_MAX)) {
    if (threadIdx.x == 0) flags[0] = 1;
    return;
  }
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) r[i] = y[i] / nr;
}

// v = pp - V_a (V_a^T pp), |v| (stop test), v /= |v|, S -= v (v^T S), and the
// commit of component a (R, Q, V columns, Bcur += r q^T, flags[1] = a + 1).
__global__ __launch_bounds__(SP_NT) void k_sp_orth_small(int a, const double* pp, double* S,
                                                         int64_t p, int64_t m, const double* r,
                                                         const double* q, double* R, double* Q,
                                                         double* V, double* Bcur,
                                                         int32_t* flags) {
  __shared__ double v[SP_PMAX];
  __shared__ double c[SP_MMAX];
  __shared__ double red[SP_NT / 64];
  if (flags[0]) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = w; j < a; j += SP_NT / 64) {
    double acc = 0.0;
    for (int64_t i = lane; i < p; i += 64) acc = fma(V[i + (int64_t)j * p], pp[i], acc);
    acc = wave_sum_d(acc);
    if (lane == 0) c[j] = acc;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) {
    double acc = 0.0;
    for (int j = 0; j < a; ++j) acc = fma(V[i + (int64_t)j * p], c[j], acc);
    v[i] = pp[i] - acc;
  }
  __syncthreads();
  const double nr = block_norm(v, p, red);
  if (!(nr > 0.0 && nr <= DBL_MAX)) {
    if (threadIdx.x == 0) flags[0] = 1;
    return;
  }
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) v[i] /= nr;
  __syncthreads();
  for (int64_t j = w; j < m; j += SP_NT / 64) {
    double acc = 0.0;
    for (int64_t i = lane; i < p; i += 64) acc = fma(v[i], S[i + j * p], acc);
    acc = wave_sum_d(acc);
    if (lane == 0) c[j] = acc;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < p; i += SP_NT) {
    const double vi = v[i], ri = r[i];
    for (int64_t j = 0; j < m; ++j) {
      S[i + j * p] -= vi * c[j];
      Bcur[i + j * p] += ri * q[j];
    }
    R[i + (int64_t)a * p] = ri;
    V[i + (int64_t)a * p] = vi;
  }
  for (int64_t j = threadIdx.x; j < m; j += SP_NT) Q[j + (int64_t)a * m] = q[j];
  if (threadIdx.x == 0) flags[1] = a + 1;
}

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

__global__ void k_div_vec(const double* __restrict__ x, const double* __restrict__ d, int64_t n,
                          double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = x[i] / d[i];
}

}  // namespace fpls

using namespace fpls;

namespace fpls {

// pls_model2_fast on device-resident inputs (dense Y or labels with m classes).
void simpls_fit_dev(Ctx* c, Model* M, const DMat& X, const DMat& Y, const int32_t* y, int64_t m,
                    const int32_t* ncomp, int32_t n_ncomp, const fpls_opts& o) {
  M->ctx = c;
  M->method = FPLS_METHOD_SIMPLS;
  M->precision = o.precision;
  M->scaling = o.scaling;
  if (y) M->n_classes = (int)m;
  const int64_t n = X.rows, p = X.cols;
  FPLS_REQUIRE(n >= 1 && p >= 1 && m >= 1, FPLS_ERR_ARG, "empty input");
  M->n = n;
  M->p = p;
  M->m = m;
  M->in_dtype = X.dtype;
  FPLS_REQUIRE(ncomp && n_ncomp >= 1, FPLS_ERR_ARG, "ncomp must contain at least one value");
  std::vector<int> nc(ncomp, ncomp + n_ncomp);
  for (int& v : nc) v = std::max(v, 1);
  const int K = *std::max_element(nc.begin(), nc.end());
  const int L = (int)nc.size();
  M->ncomp = nc;
  M->kmax = K;
  const bool can_incremental = (p > 5) && (m > 1);
  // component 0's Gaussian start, drawn before any device work is queued
  std::vector<double> omega0;
  if (can_incremental) {
    omega0.resize(p);
    gaussian_host(omega0.data(), p, o.seed);
  }
  fit_preprocess(c, M, X, y ? nullptr : &Y, y, (int32_t)m, o.scaling, false);
  DBuf<double> S(c, (size_t)p * m), S0;
  cross_cov(c, M, X, y ? nullptr : &Y, y, m, S.get(), o.precision);
  M->R = DBuf<double>(c, (size_t)p * K);
  M->Q = DBuf<double>(c, (size_t)m * K);
  M->V = DBuf<double>(c, (size_t)p * K);
  fill_f64(c, M->R.get(), 0.0, p * K);
  fill_f64(c, M->Q.get(), 0.0, m * K);
  fill_f64(c, M->V.get(), 0.0, p * K);
  M->b_stored = store_b_policy(o.store_b, p, m, L);
  DBuf<double> Bcur(c, (size_t)p * m);
  fill_f64(c, Bcur.get(), 0.0, p * m);
  if (M->b_stored) {
    M->B = DBuf<double>(c, (size_t)p * m * L);
    fill_f64(c, M->B.get(), 0.0, p * m * (int64_t)L);
  }
  DBuf<double> Yc, Yfit_cur, r2tmp;
  if (o.fit) {
    M->R2Y = DBuf<double>(c, L);
    fill_f64(c, M->R2Y.get(), 0.0, L);
    M->Yfit = DBuf<double>(c, (size_t)n * m * L);
    fill_f64(c, M->Yfit.get(), 0.0, n * m * (int64_t)L);
    Yc = centered_response(c, M, y ? nullptr : &Y, y, n, m);
    Yfit_cur = DBuf<double>(c, (size_t)n * m);
    fill_f64(c, Yfit_cur.get(), 0.0, n * m);
    r2tmp = DBuf<double>(c, 1);
  }
  // X^T X shortcut of the reference (src/fastPLS.cpp:2556-2568)
  const bool use_xtx = (K >= 20) && (p <= n) && (n >= 8 * p) && (p <= 512);
  DBuf<double> XtX;
  if (use_xtx) {
    XtX = DBuf<double>(c, (size_t)p * p);
    if (X.dtype == FPLS_F32)
      gemm_tn_f32(c, o.precision, p, p, n, X.f(), X.ld, M->mX.get(), X.f(), X.ld, nullptr,
                  M->mX.get(), XtX.get(), p);
    else
      gemm_f64(c, true, false, p, p, n, 1.0, X.d(), X.ld, M->mX.get(), X.d(), X.ld, M->mX.get(),
               0.0, XtX.get(), p);
    if (o.scaling == 2) {
      scale_rows_div(c, XtX.get(), p, p, p, M->vX.get());
      DBuf<double> tmp(c, (size_t)p * p);
      transpose_f64(c, XtX.get(), p, p, p, tmp.get(), p);
      scale_rows_div(c, tmp.get(), p, p, p, M->vX.get());
      copy_f64(c, XtX.get(), tmp.get(), p * p);
    }
    S0 = DBuf<double>(c, (size_t)p * m);
    copy_f64(c, S0.get(), S.get(), p * m);
  }
  DBuf<double> yv_(c, p), zv(c, m), rr(c, p), rs(c, p), tt(c, n), pp(c, p), qq(c, m), vv(c, p),
      cv(c, K), vS(c, m), tn2(c, 1), om(c, std::max<int64_t>(p, 1)), vc(c, p);
  DBuf<double2> npart(c, (size_t)((std::max(n, p) + SP_NCH - 1) / SP_NCH));
  const DMat Sv = dview(S.get(), p, m, p);
  DBuf<int32_t> flags(c, 2);
  fill_i32(c, flags.get(), 0, 2);
  if (can_incremental)
    FPLS_HIP(hipMemcpyAsync(om.get(), omega0.data(), p * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
  int32_t* fl = flags.get();
  const bool small = can_incremental && p <= SP_PMAX && m <= SP_MMAX && K <= SP_MMAX;
  const bool unit_scale = M->scaling != 2;  // vX = 1: centring only, or none
  int i_out = 0;
  for (int a = 0; a < K; ++a) {
    // an occasional look at the stop flag spares a long tail of dead
    // components after an early stop (one round trip per 32 components)
    if (a > 0 && a % 32 == 0) {
      int32_t hs = 0;
      to_host_i32(c, &hs, fl, 1);
      if (hs) break;
    }
    // ---- refresh: leading left singular direction of the deflated S
    const double* warm = a == 0 ? om.get() : rr.get();
    if (can_incremental && small) {
      launch(c, "simpls_refresh", k_sp_refresh_small, dim3(1), dim3(SP_NT), 0,
             (const double*)S.get(), p, m, std::max(o.inc_iters, 0), warm, rr.get(), fl);
    } else if (can_incremental) {
      // power steps y <- S (S^T y) from the warm start, then normalised
      const double* yin = warm;
      for (int it = 0; it < std::max(o.inc_iters, 0); ++it) {
        gemv_t(c, Sv, nullptr, yin, zv.get());
        gemv_n(c, Sv, nullptr, zv.get(), yv_.get());
        yin = yv_.get();
      }
      sp_normalize(c, yin, nullptr, rr.get(), p, nullptr, 0, npart.get(), fl);
    } else {
      SvdOut sv;
      svd_exact_dev(c, S.get(), p, p, m, 1, sv, true);
      copy_f64(c, rr.get(), sv.U.get(), p);
    }
    // ---- append component (src/fastPLS.cpp:2570-2663)
    if (use_xtx) {
      // pp = XtX rr; tn = sqrt(rr . pp); rr /= tn; pp /= tn; qq = S0^T rr
      gemm_f64(c, false, false, p, 1, p, 1.0, XtX.get(), p, nullptr, rr.get(), p, nullptr, 0.0,
               pp.get(), p);
      gemm_f64(c, true, false, 1, 1, p, 1.0, rr.get(), p, nullptr, pp.get(), p, nullptr, 0.0,
               tn2.get(), 1);
      launch(c, "simpls_xtx_scale", k_sp_xtx_scale, dim3(1), dim3(SP_NT), 0,
             (const double*)tn2.get(), rr.get(), pp.get(), p, fl);
      gemm_f64(c, true, false, m, 1, p, 1.0, S0.get(), p, nullptr, rr.get(), p, nullptr, 0.0,
               qq.get(), m);
      if (o.fit) {
        launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, rr.get(), M->vX.get(), p,
               rs.get());
        gemv_n(c, X, M->mX.get(), rs.get(), tt.get());
      }
    } else {
      // tt = Xs rr = Xc (rr / vX); |t| (stop test); tt, rr /= |t|
      // (without autoscaling vX is all ones and the divisions are identities)
      if (!unit_scale)
        launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, rr.get(), M->vX.get(), p,
               rs.get());
      const double* xin = unit_scale ? rr.get() : rs.get();
      if (!gemv_n_normalize(c, X, M->mX.get(), xin, tt.get(), rr.get(), p, fl)) {
        gemv_n(c, X, M->mX.get(), xin, tt.get());
        sp_normalize(c, tt.get(), nullptr, tt.get(), n, rr.get(), p, npart.get(), fl);
      }
      // pp = Xs^T tt = (Xc^T tt) / vX
      gemv_t(c, X, M->mX.get(), tt.get(), pp.get());
      if (!unit_scale)
        launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, pp.get(), M->vX.get(), p,
               pp.get());
      // qq = Yc^T tt
      if (y)
        gemm_f64(c, true, false, 1, m, n, 1.0, tt.get(), n, nullptr, nullptr, n, M->mY.get(), 0.0,
                 qq.get(), 1, y);
      else
        gemv_t(c, Y, M->mY.get(), tt.get(), qq.get());
    }
    if (small) {
      launch(c, "simpls_orth", k_sp_orth_small, dim3(1), dim3(SP_NT), 0, a,
             (const double*)pp.get(), S.get(), p, m, (const double*)rr.get(),
             (const double*)qq.get(), M->R.get(), M->Q.get(), M->V.get(), Bcur.get(), fl);
    } else {
      // vv = pp - V_a (V_a^T pp); |v| (stop test); vv /= |v|
      if (a > 0) {
        const DMat Va = dview(M->V.get(), p, a, p);
        gemv_t(c, Va, nullptr, pp.get(), cv.get());
        gemv_n(c, Va, nullptr, cv.get(), vc.get());
      }
      sp_normalize(c, pp.get(), a > 0 ? vc.get() : nullptr, vv.get(), p, nullptr, 0, npart.get(),
                   fl);
      // S -= vv (vv^T S)
      gemv_t(c, Sv, nullptr, vv.get(), vS.get());
      launch(c, "simpls_deflate", k_sp_rank1, dim3(g1(p * m)), dim3(256), 0, S.get(), p, p, m,
             (const double*)vv.get(), (const double*)vS.get(), -1.0, (const int32_t*)fl);
      // the large path forms B at the requested orders from R and Q instead
      // of a running sum (no per-component pass over a p x m matrix)
      launch(c, "simpls_commit", k_sp_commit, dim3(g1(std::max(p, m))), dim3(256), 0, a,
             (const double*)rr.get(), (const double*)qq.get(), (const double*)vv.get(), p, m,
             M->R.get(), M->Q.get(), M->V.get(), (double*)nullptr, fl);
    }
    while (i_out < L && a == nc[i_out] - 1) {
      if (M->b_stored) {
        // B_a = R[:, :a+1] Q[:, :a+1]^T, stored unless the fit has stopped
        if (!small)
          gemm_f64(c, false, true, p, m, a + 1, 1.0, M->R.get(), p, nullptr, M->Q.get(), m,
                   nullptr, 0.0, Bcur.get(), p);
        launch(c, "simpls_store", k_sp_copy, dim3(g1(p * m)), dim3(256), 0,
               M->B.get() + (size_t)i_out * p * m, (const double*)Bcur.get(), p * m, (int64_t)1,
               (const double*)nullptr, (const int32_t*)fl);
      }
      if (o.fit) {
        // reference quirk kept: Yfit_cur only accumulates at requested ncomp
        launch(c, "simpls_yfit", k_sp_rank1, dim3(g1(n * m)), dim3(256), 0, Yfit_cur.get(), n, n,
               m, (const double*)tt.get(), (const double*)qq.get(), 1.0, (const int32_t*)fl);
        r2_into(c, Yc.get(), n, m, Yfit_cur.get(), r2tmp.get());
        launch(c, "simpls_store", k_sp_copy, dim3(1), dim3(64), 0, M->R2Y.get() + i_out,
               (const double*)r2tmp.get(), (int64_t)1, (int64_t)1, (const double*)nullptr,
               (const int32_t*)fl);
        launch(c, "simpls_store", k_sp_copy, dim3(g1(n * m)), dim3(256), 0,
               M->Yfit.get() + (size_t)i_out * n * m, (const double*)Yfit_cur.get(), n, m,
               (const double*)M->mY.get(), (const int32_t*)fl);
      }
      ++i_out;
    }
  }
  int32_t done[2] = {0, 0};
  to_host_i32(c, done, fl, 2);  // the fit's one round trip (waits for the whole chain)
  M->n_components = done[1];
  build_projection_cache(c, M, M->scaling != 2);
}

}  // namespace fpls

extern "C" int fpls_simpls_fit(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                               const fpls_ivec* yv, int32_t n_classes, const int32_t* ncomp,
                               int32_t n_ncomp, const fpls_opts* opts, fpls_model** out) {
  if (out) *out = nullptr;
  std::unique_ptr<fpls_model> mdl(new fpls_model());
  int rc = fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(out != nullptr, FPLS_ERR_ARG, "out is NULL");
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, hy;
    DBuf<int32_t> hl;
    DMat X = stage(c, Xm, hx, "simpls: X");
    DMat Y;
    const int32_t* y = nullptr;
    int64_t m = 0;
    if (yv && yv->ptr) {
      FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "n_classes must be >= 1 with labels");
      y = stage_labels(c, yv, X.rows, hl, "simpls: y");
      m = n_classes;
    } else {
      Y = stage(c, Ym, hy, "simpls: Y");
      FPLS_REQUIRE(Y.rows == X.rows, FPLS_ERR_ARG, "X and Y must have the same rows");
      m = Y.cols;
    }
    simpls_fit_dev(c, &mdl->m, X, Y, y, m, ncomp, n_ncomp, o);
  });
  if (rc == FPLS_OK) *out = mdl.release();
  return rc;
}
