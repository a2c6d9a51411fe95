// PLSSVD fit (pls_model1, src/fastPLS.cpp:4780-4986, and the label-aware
// stream model, R/main.R:513-655), model access, latent projection and
// prediction (pls_predict, src/fastPLS.cpp:3211-3445).
//
// Data flow for an FP32 X (n x p) resident in HBM:
//   col_sums          -> mX                               (1 pass over X)
//   gemm_tn (labels)  -> S = Xc^T (onehot(y) - 1 mY^T)    (1 pass over X, MFMA,
//                        one-hot generated from int32 labels, never stored)
//   Gram + Jacobi     -> U, s, V of S                     (FP64, one workgroup)
//   gemm_nn           -> T = Xc U                         (1 pass over X, MFMA)
//   gemm_tn           -> G = T^T T                        (FP64 reduce)
//   slice solves      -> C_k = G_k^{-1} diag(s_k), W_k = C_k V_k^T, B_k = U_k W_k
// Centring is applied on load (X is never rewritten); T stays resident for the
// classifier heads.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <string>
#include <cmath>
#include <set>
#include <vector>

#include "model.h"
#include "fpls_wave.h"

namespace fpls {

__global__ void k_counts_to_prop(const unsigned long long* counts, double* mY, double* cnt_d,
                                 int C, int64_t n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) {
    const double v = (double)counts[c];
    mY[c] = v / (double)n;
    if (cnt_d) cnt_d[c] = v;
  }
}

static unsigned g1(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

// The projection cache in one pass, block j per column of the padded FP32
// operand (Kp x Np, ld Kp; j >= k and rows >= p are zero):
//   R_predict(:, j) = R(:, j) / vX (or R when unit; zero or non-finite vX
//   count as 1), its FP32 copy (ld p) and
//   padded copy, offset[j] = mX . R_predict(:, j) (fixed-order block sum), and
//   from block 0 the FP32 mX.
__global__ __launch_bounds__(256) void k_proj_cache(const double* __restrict__ R, int64_t p,
                                                    int64_t k, const double* __restrict__ vX,
                                                    const double* __restrict__ mX, int unit,
                                                    double* __restrict__ Rp,
                                                    float* __restrict__ Rf,
                                                    float* __restrict__ Rpad, int64_t Kp,
                                                    double* __restrict__ offset,
                                                    float* __restrict__ mXf) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int64_t i = t; i < Kp; i += 256) {
    float f = 0.f;
    if (j < k && i < p) {
      double g = unit ? 1.0 : vX[i];
      if (!isfinite(g) || g == 0.0) g = 1.0;  // the rule of scale_rows_div
      const double x = unit ? R[i + j * p] : R[i + j * p] / g;
      Rp[i + j * p] = x;
      f = (float)x;
      Rf[i + j * p] = f;
      acc = fma(mX[i], x, acc);
    }
    Rpad[i + j * Kp] = f;
    if (j == 0 && i < p) mXf[i] = (float)mX[i];
  }
  if (j >= k) return;  // uniform
  acc = wave_sum_d(acc);
  if ((t & 63) == 0) red[t >> 6] = acc;
  __syncthreads();
  if (t == 0) offset[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

void build_projection_cache(Ctx* ctx, Model* M, bool unit_scale) {
  const int64_t p = M->p;
  const int k = M->kmax;
  const int64_t Kp = nn_pad_k(p), Np = nn_pad_n(k);
  M->R_predict = DBuf<double>(ctx, (size_t)p * k);
  M->R_offset = DBuf<double>(ctx, k);
  M->R_predict_f32 = DBuf<float>(ctx, (size_t)p * k);
  M->R_pad_f32 = DBuf<float>(ctx, (size_t)Kp * Np);
  M->mX_f32 = DBuf<float>(ctx, (size_t)p);
  // vX == 1 (no autoscaling): the division would return R itself
  launch(ctx, "proj_cache", k_proj_cache, dim3((unsigned)Np), dim3(256), 0,
         (const double*)M->R.get(), p, (int64_t)k, (const double*)M->vX.get(),
         (const double*)M->mX.get(), unit_scale ? 1 : 0, M->R_predict.get(),
         M->R_predict_f32.get(), M->R_pad_f32.get(), Kp, M->R_offset.get(), M->mX_f32.get());
}

void project_scores(Ctx* ctx, const Model* M, const DMat& X, int k, float* Tf, double* Td,
                    int64_t ldt) {
  FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG, "projection: X has the wrong number of columns");
  FPLS_REQUIRE(k >= 1 && k <= M->kmax, FPLS_ERR_ARG, "projection: k out of range");
  // (X - 1 mX^T) R_predict == X R_predict - 1 offset^T (R/main.R:1003-1016);
  // centring on load keeps FP32 accuracy for uncentred data.
  if (X.dtype == FPLS_F32) {
    if (k == M->kmax && M->R_pad_f32.get() && M->mX_f32.get())
      gemm_nn_f32_padded(ctx, M->precision, X.rows, k, X.cols, X.f(), X.ld, M->mX.get(),
                         M->mX_f32.get(), M->R_pad_f32.get(), nullptr, Tf, Td, ldt);
    else
      gemm_nn_f32(ctx, M->precision, X.rows, k, X.cols, X.f(), X.ld, M->mX.get(),
                  M->R_predict_f32.get(), M->p, nullptr, Tf, Td, ldt);
  } else {
    FPLS_REQUIRE(Td != nullptr, FPLS_ERR_ARG, "projection: FP64 input needs an FP64 output");
    gemm_f64(ctx, false, false, X.rows, k, X.cols, 1.0, X.d(), X.ld, M->mX.get(),
             M->R_predict.get(), M->p, nullptr, 0.0, Td, ldt);
  }
}

// mX, vX, mY (+ counts) for a fit; shared by PLSSVD and SIMPLS.
void fit_preprocess(Ctx* ctx, Model* M, const DMat& X, const DMat* Y, const int32_t* y,
                    int32_t C, int scaling, bool guard_sd, bool means_later) {
  const int64_t n = X.rows, p = X.cols;
  M->mX = DBuf<double>(ctx, p);
  M->vX = DBuf<double>(ctx, p);
  // means (and unit scales) in the column pass itself; sd = sqrt(ssd / (n-1));
  // means_later: the one-pass class sums of cross_cov take mX (and vX = 1 or
  // the sd pass after it)
  if (means_later) {
  } else if (scaling < 3) {
    col_stat(ctx, X, nullptr, M->mX.get(), (double)n, 0, scaling == 2 ? nullptr : M->vX.get());
  } else {
    fill_f64(ctx, M->mX.get(), 0.0, p);
  }
  if (means_later) {
  } else if (scaling == 2) {
    col_stat(ctx, X, M->mX.get(), M->vX.get(), (double)(n > 1 ? n - 1 : 1), guard_sd ? 2 : 1);
  } else if (scaling >= 3) {
    fill_f64(ctx, M->vX.get(), 1.0, p);
  }
  if (y) {
    M->mY = DBuf<double>(ctx, C);
    M->counts = DBuf<double>(ctx, C);
    // rows with codes outside 1..C are counted in no class by every kernel of
    // the fit; the count is checked when the entry point returns
    int32_t* bad =
        ctx->defer_check(1, FPLS_ERR_LABEL_RANGE, "labels must be encoded as 1..n_classes");
    if (!label_counts_prop(ctx, y, n, C, M->mY.get(), M->counts.get(), bad)) {
      DBuf<int64_t> cnt(ctx, C);
      label_counts(ctx, y, n, C, cnt.get(), bad, true);
      launch(ctx, "counts", k_counts_to_prop, dim3((C + 255) / 256), dim3(256), 0,
             reinterpret_cast<const unsigned long long*>(cnt.get()), M->mY.get(),
             M->counts.get(), C, n);
    }
  } else {
    M->mY = DBuf<double>(ctx, Y->cols);
    col_stat(ctx, *Y, nullptr, M->mY.get(), (double)n, 0);
  }
}

// Whether the PLSSVD fit of labels takes the column means from the one-pass
// class sums (FP32 X, centring or autoscaling, at most 1663 classes).
bool fit_means_from_class_sums(const DMat& X, const int32_t* y, int64_t C, int scaling) {
  return y != nullptr && X.dtype == FPLS_F32 && scaling < 3 && C >= 1 && C <= 1663;
}

// S = ((X - mX)/vX)^T (Y - mY)  (p x m, FP64); Y dense or labels.  With
// means_later (fit_means_from_class_sums) the class-sum pass also writes mX,
// and vX (1, or the sd pass that follows it).
void cross_cov(Ctx* ctx, const Model* M, const DMat& X, const DMat* Y, const int32_t* y, int64_t m,
               double* S, int precision, bool means_later, bool guard_sd) {
  const int64_t n = X.rows, p = X.cols;
  if (means_later) {
    double* ones = M->scaling == 2 ? nullptr : M->vX.get();
    FPLS_REQUIRE(class_cross_cov_f32(ctx, precision, p, m, n, X.f(), X.ld, y, M->mY.get(), S, p,
                                     M->mX.get(), ones) ||
                     class_sums_f32(ctx, n, p, X.f(), X.ld, y, (int)m, M->mY.get(), S, p,
                                    M->mX.get(), ones),
                 FPLS_ERR_ARG, "cross_cov: too many classes for the one-pass class sums");
    if (M->scaling == 2) {
      col_stat(ctx, X, M->mX.get(), M->vX.get(), (double)(n > 1 ? n - 1 : 1), guard_sd ? 2 : 1);
      scale_rows_div(ctx, S, p, p, m, M->vX.get());
    }
    return;
  }
  if (X.dtype == FPLS_F32) {
    const float* Bp = nullptr;
    int64_t ldb = 0;
    DBuf<float> Yf;
    if (!y) {
      if (Y->dtype == FPLS_F32) {
        Bp = Y->f();
        ldb = Y->ld;
      } else {
        Yf = DBuf<float>(ctx, (size_t)n * m);
        for (int64_t j = 0; j < m; ++j) f64_to_f32(ctx, Yf.get() + j * n, Y->d() + j * Y->ld, n);
        Bp = Yf.get();
        ldb = n;
      }
    }
    gemm_tn_f32(ctx, precision, p, m, n, X.f(), X.ld, M->mX.get(), Bp, ldb, y, M->mY.get(), S, p);
  } else {
    DBuf<double> Yd;
    const double* Bp = nullptr;
    int64_t ldb = n;
    if (!y) {
      if (Y->dtype == FPLS_F64) {
        Bp = Y->d();
        ldb = Y->ld;
      } else {
        Yd = DBuf<double>(ctx, (size_t)n * m);
        for (int64_t j = 0; j < m; ++j) f32_to_f64(ctx, Yd.get() + j * n, Y->f() + j * Y->ld, n);
        Bp = Yd.get();
      }
    }
    gemm_f64(ctx, true, false, p, m, n, 1.0, X.d(), X.ld, M->mX.get(), Bp, ldb, M->mY.get(), 0.0,
             S, p, y);
  }
  if (M->scaling == 2) scale_rows_div(ctx, S, p, p, m, M->vX.get());
}

// Whether a fit keeps the dense coefficient cube B (p x m x slices) beside the
// compact factors, should_store_coefficients (src/fastPLS.cpp:63-85): the
// caller's choice, else FASTPLS_STORE_B (always / 1 / true / yes, never / 0 /
// false / no), else B fits in FASTPLS_STORE_B_MAX_MB (default 256, clamped to
// 0..1048576; an unparsable value keeps the default).
bool store_b_policy(int store_b, int64_t p, int64_t m, int n_slices) {
  if (store_b == 0) return false;
  if (store_b == 1) return true;
  if (const char* mode = std::getenv("FASTPLS_STORE_B")) {
    std::string v(mode);
    for (char& ch : v) ch = (char)std::tolower((unsigned char)ch);
    if (v == "always" || v == "1" || v == "true" || v == "yes") return true;
    if (v == "never" || v == "0" || v == "false" || v == "no") return false;
  }
  long max_mb = 256;
  if (const char* raw = std::getenv("FASTPLS_STORE_B_MAX_MB")) {
    char* end = nullptr;
    const long v = std::strtol(raw, &end, 10);
    if (end != raw) max_mb = std::min<long>(1048576, std::max<long>(0, v));
  }
  const double mb = (double)p * (double)m * (double)std::max(n_slices, 1) * 8.0 / (1024.0 * 1024.0);
  return mb <= (double)max_mb;
}

// Compute T = Xs R (kmax columns) into the model, and G = T^T T.
void fit_scores_and_gram(Ctx* ctx, Model* M, const DMat& X) {
  const int64_t n = X.rows;
  const int k = M->kmax;
  if (X.dtype == FPLS_F32) {
    M->T_f32 = DBuf<float>(ctx, (size_t)n * k);
    project_scores(ctx, M, X, k, M->T_f32.get(), nullptr, n);
  } else {
    M->T_f64 = DBuf<double>(ctx, (size_t)n * k);
    project_scores(ctx, M, X, k, nullptr, M->T_f64.get(), n);
  }
  M->G = DBuf<double>(ctx, (size_t)k * k);
  DMat T = M->T();
  if (T.dtype == FPLS_F32)
    gemm_tn_f32(ctx, M->precision, k, k, n, T.f(), n, nullptr, T.f(), n, nullptr, nullptr,
                M->G.get(), k);
  else
    gemm_f64(ctx, true, false, k, k, n, 1.0, T.d(), n, nullptr, T.d(), n, nullptr, 0.0,
             M->G.get(), k);
}

// R2 = 1 - PRESS/TSS over all columns (RQ, src/fastPLS.cpp:1483-1500):
// acc[0] += (y - yhat)^2, acc[1] += (y - mean_col(y))^2.  Y centred already.
__global__ void k_press_tss(const double* __restrict__ Yc, int64_t ldy, const double* __restrict__ Yh,
                            int64_t n, int64_t m, const double* __restrict__ ymean,
                            double* __restrict__ acc) {
  __shared__ double r0[4], r1[4];
  double a = 0.0, b = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * m;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, j = t / n;
    const double y = Yc[i + j * ldy], d = y - Yh[i + j * n], e = y - ymean[j];
    a += d * d;
    b += e * e;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = a;
    r1[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], r0[0] + r0[1] + r0[2] + r0[3]);
    atomicAdd(&acc[1], r1[0] + r1[1] + r1[2] + r1[3]);
  }
}

__global__ void k_r2(const double* acc, double* out) { *out = 1.0 - acc[0] / acc[1]; }

__global__ void k_add_row(double* X, int64_t ld, int64_t rows, int64_t cols, const double* v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < rows * cols;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    X[i + j * ld] += v[j];
  }
}

void add_row(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols, const double* v) {
  if (rows > 0 && cols > 0)
    launch(ctx, "add_row", k_add_row, dim3(g1(rows * cols)), dim3(256), 0, X, ld, rows, cols, v);
}

// Centered response Yc (n x m, FP64) materialised for the fit statistics.
DBuf<double> centered_response(Ctx* ctx, const Model* M, const DMat* Y, const int32_t* y,
                               int64_t n, int64_t m);

void r2_into(Ctx* ctx, const double* Yc, int64_t n, int64_t m, const double* Yhat, double* out) {
  DBuf<double> acc(ctx, 2), ymean(ctx, m), ones(ctx, n);
  fill_f64(ctx, acc.get(), 0.0, 2);
  fill_f64(ctx, ones.get(), 1.0 / (double)n, n);
  gemm_f64(ctx, true, false, m, 1, n, 1.0, Yc, n, nullptr, ones.get(), n, nullptr, 0.0,
           ymean.get(), m);
  launch(ctx, "r2", k_press_tss, dim3(g1(n * m)), dim3(256), 0, Yc, n, Yhat, n, m, ymean.get(),
         acc.get());
  launch(ctx, "r2", k_r2, dim3(1), dim3(1), 0, acc.get(), out);
}

__global__ void k_onehot_centered(const int32_t* y, int64_t n, int64_t C, const double* mY,
                                  double* out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * C;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, c = t / n;
    out[t] = (y[i] == c + 1 ? 1.0 : 0.0) - mY[c];
  }
}

__global__ void k_center_cols(const double* Y, int64_t ldy, int64_t n, int64_t m, const double* mY,
                              double* out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * m;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, j = t / n;
    out[t] = Y[i + j * ldy] - mY[j];
  }
}
__global__ void k_center_cols_f32(const float* Y, int64_t ldy, int64_t n, int64_t m,
                                  const double* mY, double* out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * m;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, j = t / n;
    out[t] = (double)Y[i + j * ldy] - mY[j];
  }
}

DBuf<double> centered_response(Ctx* ctx, const Model* M, const DMat* Y, const int32_t* y,
                               int64_t n, int64_t m) {
  DBuf<double> out(ctx, (size_t)n * m);
  if (y)
    launch(ctx, "onehot", k_onehot_centered, dim3(g1(n * m)), dim3(256), 0, y, n, m, M->mY.get(),
           out.get());
  else if (Y->dtype == FPLS_F64)
    launch(ctx, "center", k_center_cols, dim3(g1(n * m)), dim3(256), 0, Y->d(), Y->ld, n, m,
           M->mY.get(), out.get());
  else
    launch(ctx, "center", k_center_cols_f32, dim3(g1(n * m)), dim3(256), 0, Y->f(), Y->ld, n, m,
           M->mY.get(), out.get());
  return out;
}

}  // namespace fpls

using namespace fpls;

static std::vector<int> read_ncomp(const int32_t* ncomp, int32_t n_ncomp) {
  FPLS_REQUIRE(ncomp != nullptr && n_ncomp >= 1, FPLS_ERR_ARG,
               "ncomp must contain at least one value");
  return std::vector<int>(ncomp, ncomp + n_ncomp);
}

namespace fpls {

// pls_model1 on device-resident inputs (dense Y or labels with m classes).
void plssvd_fit_dev(Ctx* c, Model* M, const DMat& X, const DMat& Y, const int32_t* y, int64_t m,
                    const int32_t* ncomp, int32_t n_ncomp, const fpls_opts& o,
                    const LdaFuse* lda) {
  {
    M->ctx = c;
    M->method = FPLS_METHOD_PLSSVD;
    M->precision = o.precision;
    M->scaling = o.scaling;
    const int64_t n = X.rows, p = X.cols;
    FPLS_REQUIRE(n >= 1 && p >= 1 && m >= 1, FPLS_ERR_ARG, "empty input");
    M->n = n;
    M->p = p;
    M->m = m;
    M->in_dtype = X.dtype;
    M->label_aware = (y != nullptr) && o.label_aware;
    // ncomp capping (src/fastPLS.cpp:4801-4814; R/main.R:530-532)
    const int max_rank = (int)std::min<int64_t>(n, std::min(p, m));
    FPLS_REQUIRE(max_rank >= 1, FPLS_ERR_RANK, "plssvd effective rank is < 1");
    std::vector<int> nc = read_ncomp(ncomp, n_ncomp);
    for (int& v : nc) v = std::min(std::max(v, 1), max_rank);
    int kmax = *std::max_element(nc.begin(), nc.end());
    const int L = (int)nc.size();
    const bool one_pass = fit_means_from_class_sums(X, y, m, o.scaling);
    fit_preprocess(c, M, X, y ? nullptr : &Y, y, (int32_t)m, o.scaling, M->label_aware,
                   one_pass);
    if (y) M->n_classes = (int)m;
    // truncated SVD of S (src/fastPLS.cpp:4843-4856, plssvd_use_small_exact_svd :318-324)
    SvdOut sv;
    sv.want_v = true;
    const bool full = M->label_aware || o.svd_method == FPLS_SVD_EXACT || max_rank <= 32;
    const int64_t xrank = std::min(p, m);
    const int l_sk = (int)std::min<int64_t>(xrank, (int64_t)kmax + std::max(o.oversample, 0));
    // matrix-free route (truncated_rsvd_crossprod_double): dense Y, the range
    // finder, a sketch narrower than S (the reference forms S otherwise)
    const bool implicit = o.xprod && !y && !full && l_sk < xrank && l_sk <= kJacobiMax;
    M->xprod_mode = M->label_aware ? 2 : (implicit ? 1 : 0);
    if (implicit) {
      DBuf<double> Yd(c, (size_t)n * m);
      if (Y.dtype == FPLS_F64) copy2d_f64(c, Yd.get(), n, Y.d(), Y.ld, n, m);
      else
        for (int64_t j = 0; j < m; ++j) f32_to_f64(c, Yd.get() + j * n, Y.f() + j * Y.ld, n);
      std::vector<double> om((size_t)m * l_sk);
      gaussian_host(om.data(), (int64_t)om.size(), o.seed);
      rsvd_xprod_core(c, X, M->mX.get(), o.scaling == 2 ? M->vX.get() : nullptr, Yd.get(), m,
                      M->mY.get(), kmax, l_sk, o.power, om.data(), o.precision, true, sv);
    } else {
      DBuf<double> S(c, (size_t)p * m);
      cross_cov(c, M, X, y ? nullptr : &Y, y, m, S.get(), o.precision, one_pass,
                M->label_aware);
      rsvd_dev(c, S.get(), p, p, m, kmax, o.oversample, o.power, o.seed, nullptr, false, full,
               sv);
    }
    kmax = std::min(kmax, sv.k);
    FPLS_REQUIRE(kmax >= 1, FPLS_ERR_RANK, "plssvd effective rank is < 1 after SVD");
    M->kmax = kmax;
    // the factors move into the model when the SVD kept exactly kmax of them
    if (sv.k == kmax) {
      M->R = std::move(sv.U);
      M->s = std::move(sv.s);
    } else {
      M->R = DBuf<double>(c, (size_t)p * kmax);
      copy_f64(c, M->R.get(), sv.U.get(), p * kmax);
      M->s = DBuf<double>(c, kmax);
      copy_f64(c, M->s.get(), sv.s.get(), kmax);
    }
    if (sv.V.get()) {
      M->Q = std::move(sv.V);  // V (m x >= kmax, ld m)
    } else {
      M->Q = DBuf<double>(c, (size_t)m * kmax);  // V = Vt^T (m x kmax)
      transpose_f64(c, sv.Vt.get(), sv.k, kmax, m, M->Q.get(), m);
    }
    build_projection_cache(c, M, o.scaling != 2);
    fit_scores_and_gram(c, M, X);
    // per-slice latent coefficients
    for (int& v : nc) v = std::min(v, kmax);
    M->ncomp = nc;
    M->C_latent = DBuf<double>(c, (size_t)kmax * kmax * L);
    M->W_latent = DBuf<double>(c, (size_t)kmax * m * L);
    // slices below kmax leave the trailing rows and columns of their blocks
    // unwritten; at kmax the solves and products write every entry
    if (std::any_of(nc.begin(), nc.end(), [&](int v) { return v < kmax; })) {
      fill_f64(c, M->C_latent.get(), 0.0, (int64_t)kmax * kmax * L);
      fill_f64(c, M->W_latent.get(), 0.0, (int64_t)kmax * m * L);
    }
    int32_t* st = c->defer_check(L, FPLS_ERR_SOLVE, "plssvd latent solve failed", 1);
    const bool fuse = lda != nullptr && y != nullptr && kmax <= kJacobiMax;
    LdaPending lpend;
    if (fuse) {
      // the LDA head on the training scores (reusing G = T^T T); its inverse
      // covariances join the latent solves in one launch
      std::set<int> kset;
      for (int i = 0; i < lda->n_ncomp; ++i)
        kset.insert(std::min(std::max<int>(lda->ncomp[i], 1), kmax));
      FPLS_REQUIRE(!kset.empty(), FPLS_ERR_ARG, "LDA requires at least one component count");
      const std::vector<int> ks(kset.begin(), kset.end());
      std::vector<SpdJob> jobs;
      lda_prepare(c, M->T(), M->G.get(), kmax, y, (int32_t)m, ks, lda->ridge, M->precision,
                  M->counts, M->lda, lpend, jobs, true);
      spd_slice_jobs(M->G.get(), kmax, kmax, nc, L, M->label_aware ? 2 : 0, M->s.get(), 0.0,
                     M->C_latent.get(), kmax, (int64_t)kmax * kmax, nullptr, st, jobs);
      spd_solve_jobs(c, jobs);
    } else {
      spd_solve_slices(c, M->G.get(), kmax, kmax, nc, L, M->label_aware ? 2 : 0, M->s.get(), 0.0,
                       M->C_latent.get(), kmax, (int64_t)kmax * kmax, nullptr, st);
    }
    M->b_stored = !M->label_aware && store_b_policy(o.store_b, p, m, L);
    if (M->b_stored) M->B = DBuf<double>(c, (size_t)p * m * L);
    DBuf<double> Yc;
    DBuf<double> Td;
    if (o.fit) {
      M->R2Y = DBuf<double>(c, L);
      M->Yfit = DBuf<double>(c, (size_t)n * m * L);
      Yc = centered_response(c, M, y ? nullptr : &Y, y, n, m);
      if (M->T_f32.get()) {
        Td = DBuf<double>(c, (size_t)n * kmax);
        f32_to_f64(c, Td.get(), M->T_f32.get(), n * (int64_t)kmax);
      }
    }
    for (int a = 0; a < L; ++a) {
      const int mc = nc[a];
      double* Ca = M->C_latent.get() + (size_t)a * kmax * kmax;
      double* Wa = M->W_latent.get() + (size_t)a * kmax * m;
      // W_a = C_a V_mc^T  (mc x m, ld kmax)
      gemm_f64(c, false, true, mc, m, mc, 1.0, Ca, kmax, nullptr, M->Q.get(), m, nullptr, 0.0, Wa,
               kmax);
      if (M->b_stored)
        gemm_f64(c, false, false, p, m, mc, 1.0, M->R.get(), p, nullptr, Wa, kmax, nullptr, 0.0,
                 M->B.get() + (size_t)a * p * m, p);
      if (o.fit) {
        double* Yf = M->Yfit.get() + (size_t)a * n * m;
        const double* Tp = M->T_f32.get() ? Td.get() : M->T_f64.get();
        gemm_f64(c, false, false, n, m, mc, 1.0, Tp, n, nullptr, Wa, kmax, nullptr, 0.0, Yf, n);
        r2_into(c, Yc.get(), n, m, Yf, M->R2Y.get() + a);
        add_row(c, Yf, n, n, m, M->mY.get());
      }
    }
    M->n_components = kmax;
    if (fuse) {
      lda_finish(c, lpend, M->lda);
    } else if (lda != nullptr) {
      // orders above the one-workgroup solves: the LDA head on its own
      std::set<int> kset;
      for (int i = 0; i < lda->n_ncomp; ++i)
        kset.insert(std::min(std::max<int>(lda->ncomp[i], 1), kmax));
      lda_train_core(c, M->T(), M->G.get(), kmax, y, (int32_t)m,
                     std::vector<int>(kset.begin(), kset.end()), lda->ridge, M->precision,
                     M->counts, M->lda);
    }
  }
}

}  // namespace fpls

extern "C" int fpls_plssvd_fit(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
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
    DMat X = stage(c, Xm, hx, "plssvd: X");
    DMat Y;
    const int32_t* y = nullptr;
    int64_t m = 0;
    if (yv && yv->ptr) {
      FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "n_classes must be >= 1 with labels");
      y = stage_labels(c, yv, X.rows, hl, "plssvd: y");
      m = n_classes;
    } else {
      Y = stage(c, Ym, hy, "plssvd: Y");
      FPLS_REQUIRE(Y.rows == X.rows, FPLS_ERR_ARG, "X and Y must have the same rows");
      m = Y.cols;
    }
    plssvd_fit_dev(c, &mdl->m, X, Y, y, m, ncomp, n_ncomp, o);
  });
  if (rc == FPLS_OK) *out = mdl.release();
  return rc;
}

extern "C" int fpls_plssvd_lda_fit(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_ivec* yv,
                                   int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                                   const fpls_opts* opts, double lda_ridge, fpls_model** out) {
  if (out) *out = nullptr;
  std::unique_ptr<fpls_model> mdl(new fpls_model());
  int rc = fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(out != nullptr, FPLS_ERR_ARG, "out is NULL");
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx;
    DBuf<int32_t> hl;
    DMat X = stage(c, Xm, hx, "plssvd_lda: X");
    FPLS_REQUIRE(yv && yv->ptr, FPLS_ERR_ARG, "plssvd_lda: class labels are required");
    FPLS_REQUIRE(n_classes >= 2, FPLS_ERR_ARG, "LDA requires at least two classes");
    const int32_t* y = stage_labels(c, yv, X.rows, hl, "plssvd_lda: y");
    LdaFuse lf;
    lf.ridge = lda_ridge;
    lf.ncomp = ncomp;
    lf.n_ncomp = n_ncomp;
    DMat Y;
    plssvd_fit_dev(c, &mdl->m, X, Y, y, n_classes, ncomp, n_ncomp, o, &lf);
  });
  if (rc == FPLS_OK) *out = mdl.release();
  return rc;
}

// ---------------------------------------------------------------- access
extern "C" int fpls_model_info_get(const fpls_model* model, fpls_model_info* info) {
  if (!model || !info) return FPLS_ERR_ARG;
  const Model& M = model->m;
  std::memset(info, 0, sizeof(*info));
  info->method = M.method;
  info->n = M.n;
  info->p = M.p;
  info->m = M.m;
  info->kmax = M.kmax;
  info->n_ncomp = (int32_t)M.ncomp.size();
  info->n_components = M.n_components;
  info->b_stored = M.b_stored ? 1 : 0;
  info->has_lda = M.lda.empty() ? 0 : 1;
  info->has_cknn = M.cknn.empty() ? 0 : 1;
  info->n_classes = M.n_classes;
  for (size_t i = 0; i < M.ncomp.size() && i < 64; ++i) info->ncomp[i] = M.ncomp[i];
  info->xprod_mode = M.xprod_mode;
  return FPLS_OK;
}

extern "C" int fpls_model_read(fpls_ctx* ctx, const fpls_model* model, int32_t field,
                               double* host_out, int64_t n_elem) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model& M = model->m;
    const int64_t L = (int64_t)M.ncomp.size();
    const double* src = nullptr;
    int64_t need = 0;
    DBuf<double> tmp;
    switch (field) {
      case FPLS_FIELD_R: src = M.R.get(); need = M.p * M.kmax; break;
      case FPLS_FIELD_Q: src = M.Q.get(); need = M.m * M.kmax; break;
      case FPLS_FIELD_MX: src = M.mX.get(); need = M.p; break;
      case FPLS_FIELD_VX: src = M.vX.get(); need = M.p; break;
      case FPLS_FIELD_MY: src = M.mY.get(); need = M.m; break;
      case FPLS_FIELD_B: src = M.B.get(); need = M.p * M.m * L; break;
      case FPLS_FIELD_C_LATENT: src = M.C_latent.get(); need = (int64_t)M.kmax * M.kmax * L; break;
      case FPLS_FIELD_W_LATENT: src = M.W_latent.get(); need = (int64_t)M.kmax * M.m * L; break;
      case FPLS_FIELD_S: src = M.s.get(); need = M.kmax; break;
      case FPLS_FIELD_R2Y: src = M.R2Y.get(); need = L; break;
      case FPLS_FIELD_YFIT: src = M.Yfit.get(); need = M.n * M.m * L; break;
      case FPLS_FIELD_V: src = M.V.get(); need = M.p * M.kmax; break;
      case FPLS_FIELD_R_PREDICT: src = M.R_predict.get(); need = M.p * M.kmax; break;
      case FPLS_FIELD_R_OFFSET: src = M.R_offset.get(); need = M.kmax; break;
      case FPLS_FIELD_TTRAIN:
        need = M.n * M.kmax;
        if (M.T_f64.get()) src = M.T_f64.get();
        else if (M.T_f32.get()) {
          tmp = DBuf<double>(c, need);
          f32_to_f64(c, tmp.get(), M.T_f32.get(), need);
          src = tmp.get();
        }
        break;
      default: fail(FPLS_ERR_ARG, "unknown model field");
    }
    FPLS_REQUIRE(src != nullptr, FPLS_ERR_STATE, "model does not hold this field");
    FPLS_REQUIRE(host_out != nullptr && n_elem >= need, FPLS_ERR_ARG,
                 "output buffer too small for the model field");
    to_host(c, host_out, src, need);
  });
}

extern "C" void fpls_model_free(fpls_model* model) { delete model; }

// ---------------------------------------------------------------- projection
extern "C" int fpls_project(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm, int32_t k,
                            double* T_out) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "project: X");
    const int kk = std::min<int>(std::max<int>(k, 1), M->kmax);
    DBuf<double> T(c, (size_t)X.rows * kk);
    if (X.dtype == FPLS_F32) gemm_nn_f32(c, M->precision, X.rows, kk, X.cols, X.f(), X.ld,
                                         M->mX.get(), M->R_predict_f32.get(), M->p, nullptr,
                                         nullptr, T.get(), X.rows);
    else project_scores(c, M, X, kk, nullptr, T.get(), X.rows);
    to_host(c, T_out, T.get(), (size_t)X.rows * kk);
  });
}

// ---------------------------------------------------------------- prediction
extern "C" int fpls_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm,
                            double* ypred) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "predict: X");
    FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG, "predict: Xtest has the wrong number of columns");
    const int64_t w = X.rows, m = M->m;
    const int L = (int)M->ncomp.size();
    DBuf<double> out(c, (size_t)w * m * L);
    // scores T = Xs R (kmax) once; every slice is a prefix
    const int kmax = M->kmax;
    DBuf<double> T(c, (size_t)w * kmax);
    if (X.dtype == FPLS_F32)
      gemm_nn_f32(c, M->precision, w, kmax, X.cols, X.f(), X.ld, M->mX.get(),
                  M->R_predict_f32.get(), M->p, nullptr, nullptr, T.get(), w);
    else
      project_scores(c, M, X, kmax, nullptr, T.get(), w);
    for (int a = 0; a < L; ++a) {
      const int mc = M->ncomp[a];
      double* Ya = out.get() + (size_t)a * w * m;
      if (M->method == FPLS_METHOD_SIMPLS) {
        // Xs R_mc Q_mc^T
        gemm_f64(c, false, true, w, m, mc, 1.0, T.get(), w, nullptr, M->Q.get(), m, nullptr, 0.0,
                 Ya, w);
      } else {
        gemm_f64(c, false, false, w, m, mc, 1.0, T.get(), w, nullptr,
                 M->W_latent.get() + (size_t)a * kmax * m, kmax, nullptr, 0.0, Ya, w);
      }
      add_row(c, Ya, w, w, m, M->mY.get());
    }
    to_host(c, ypred, out.get(), (size_t)w * m * L);
  });
}
