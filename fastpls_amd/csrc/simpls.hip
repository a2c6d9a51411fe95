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
#include <cmath>
#include <vector>

#include "model.h"

namespace fpls {

void fit_preprocess(Ctx* ctx, Model* M, const DMat& X, const DMat* Y, const int32_t* y, int32_t C,
                    int scaling, bool guard_sd);
void cross_cov(Ctx* ctx, const Model* M, const DMat& X, const DMat* Y, const int32_t* y, int64_t m,
               double* S, int precision);
bool store_b_policy(int store_b, int64_t p, int64_t m, int n_slices);
DBuf<double> centered_response(Ctx* ctx, const Model* M, const DMat* Y, const int32_t* y,
                               int64_t n, int64_t m);
void r2_into(Ctx* ctx, const double* Yc, int64_t n, int64_t m, const double* Yhat, double* out);
void add_row(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols, const double* v);

// ||x||_2 with a fixed-order block reduction (one workgroup).
__global__ __launch_bounds__(1024) void k_norm2(const double* __restrict__ x, int64_t n,
                                                double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s = fma(x[i], x[i], s);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    *out = sqrt(t);
  }
}

// x[i] /= *d (and the same for up to one more vector y of length ny)
__global__ void k_div_by_scalar(double* __restrict__ x, int64_t n, double* __restrict__ y,
                                int64_t ny, const double* __restrict__ d) {
  const double v = *d;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] /= v;
  if (y)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ny;
         i += (int64_t)gridDim.x * blockDim.x)
      y[i] /= v;
}

__global__ void k_div_vec(const double* __restrict__ x, const double* __restrict__ d, int64_t n,
                          double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = x[i] / d[i];
}

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

static void norm2(Ctx* ctx, const double* x, int64_t n, double* out) {
  launch(ctx, "norm2", k_norm2, dim3(1), dim3(1024), 0, x, n, out);
}

}  // namespace fpls

using namespace fpls;

namespace fpls {

// pls_model2_fast on device-resident inputs (dense Y or labels with m classes).
void simpls_fit_dev(Ctx* c, Model* M, const DMat& X, const DMat& Y, const int32_t* y, int64_t m,
                    const int32_t* ncomp, int32_t n_ncomp, const fpls_opts& o) {
  {
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
    DBuf<double> Yc, Yfit_cur;
    if (o.fit) {
      M->R2Y = DBuf<double>(c, L);
      fill_f64(c, M->R2Y.get(), 0.0, L);
      M->Yfit = DBuf<double>(c, (size_t)n * m * L);
      fill_f64(c, M->Yfit.get(), 0.0, n * m * (int64_t)L);
      Yc = centered_response(c, M, y ? nullptr : &Y, y, n, m);
      Yfit_cur = DBuf<double>(c, (size_t)n * m);
      fill_f64(c, Yfit_cur.get(), 0.0, n * m);
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
    const bool can_incremental = (p > 5) && (m > 1);
    DBuf<double> yv_(c, p), zv(c, m), rr(c, p), rs(c, p), tt(c, n), pp(c, p), qq(c, m), vv(c, p),
        cv(c, K), vS(c, m), norms(c, 2);
    int a = 0, i_out = 0;
    std::vector<double> omega0;
    for (a = 0; a < K; ++a) {
      // ---- refresh: leading left singular direction of the deflated S
      if (can_incremental) {
        if (a == 0) {
          omega0.resize(p);
          gaussian_host(omega0.data(), p, o.seed + (uint32_t)a);
          to_device(c, yv_.get(), omega0.data(), p);
        } else {
          copy_f64(c, yv_.get(), rr.get(), p);  // warm start = previous rr
        }
        for (int it = 0; it < std::max(o.inc_iters, 0); ++it) {
          gemm_f64(c, true, false, m, 1, p, 1.0, S.get(), p, nullptr, yv_.get(), p, nullptr, 0.0,
                   zv.get(), m);
          gemm_f64(c, false, false, p, 1, m, 1.0, S.get(), p, nullptr, zv.get(), m, nullptr, 0.0,
                   yv_.get(), p);
        }
        norm2(c, yv_.get(), p, norms.get());
        copy_f64(c, rr.get(), yv_.get(), p);
        launch(c, "div", k_div_by_scalar, dim3(g1(p)), dim3(256), 0, rr.get(), p,
               (double*)nullptr, (int64_t)0, (const double*)norms.get());
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
                 norms.get(), 1);
        double tn2 = 0.0;
        to_host(c, &tn2, norms.get(), 1);
        if (!std::isfinite(tn2) || tn2 <= 0.0) break;
        const double tn = std::sqrt(tn2);
        to_device(c, norms.get(), &tn, 1);
        launch(c, "div", k_div_by_scalar, dim3(g1(p)), dim3(256), 0, rr.get(), p, pp.get(), p,
               (const double*)norms.get());
        gemm_f64(c, true, false, m, 1, p, 1.0, S0.get(), p, nullptr, rr.get(), p, nullptr, 0.0,
                 qq.get(), m);
        if (o.fit) {
          launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, rr.get(), M->vX.get(), p,
                 rs.get());
          gemv_n(c, X, M->mX.get(), rs.get(), tt.get());
        }
      } else {
        // tt = Xs rr = Xc (rr / vX)
        launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, rr.get(), M->vX.get(), p,
               rs.get());
        gemv_n(c, X, M->mX.get(), rs.get(), tt.get());
        norm2(c, tt.get(), n, norms.get());
        double tn = 0.0;
        to_host(c, &tn, norms.get(), 1);
        if (!std::isfinite(tn) || tn <= 0.0) break;
        launch(c, "div", k_div_by_scalar, dim3(g1(std::max(n, p))), dim3(256), 0, tt.get(), n,
               rr.get(), p, (const double*)norms.get());
        // pp = Xs^T tt = (Xc^T tt) / vX
        gemv_t(c, X, M->mX.get(), tt.get(), pp.get());
        launch(c, "div_vec", k_div_vec, dim3(g1(p)), dim3(256), 0, pp.get(), M->vX.get(), p,
               pp.get());
        // qq = Yc^T tt
        if (y) {
          gemm_f64(c, true, false, 1, m, n, 1.0, tt.get(), n, nullptr, nullptr, n, M->mY.get(),
                   0.0, qq.get(), 1, y);
        } else if (Y.dtype == FPLS_F64) {
          gemv_t(c, Y, M->mY.get(), tt.get(), qq.get());
        } else {
          gemv_t(c, Y, M->mY.get(), tt.get(), qq.get());
        }
      }
      // vv = pp - V_a (V_a^T pp)
      copy_f64(c, vv.get(), pp.get(), p);
      if (a > 0) {
        gemm_f64(c, true, false, a, 1, p, 1.0, M->V.get(), p, nullptr, pp.get(), p, nullptr, 0.0,
                 cv.get(), a);
        gemm_f64(c, false, false, p, 1, a, -1.0, M->V.get(), p, nullptr, cv.get(), a, nullptr, 1.0,
                 vv.get(), p);
      }
      norm2(c, vv.get(), p, norms.get() + 1);
      double vn = 0.0;
      to_host(c, &vn, norms.get() + 1, 1);
      if (!std::isfinite(vn) || vn <= 0.0) break;
      launch(c, "div", k_div_by_scalar, dim3(g1(p)), dim3(256), 0, vv.get(), p, (double*)nullptr,
             (int64_t)0, (const double*)(norms.get() + 1));
      // S -= vv (vv^T S)
      gemm_f64(c, true, false, 1, m, p, 1.0, vv.get(), p, nullptr, S.get(), p, nullptr, 0.0,
               vS.get(), 1);
      gemm_f64(c, false, false, p, m, 1, -1.0, vv.get(), p, nullptr, vS.get(), 1, nullptr, 1.0,
               S.get(), p);
      copy_f64(c, M->R.get() + (size_t)a * p, rr.get(), p);
      copy_f64(c, M->Q.get() + (size_t)a * m, qq.get(), m);
      copy_f64(c, M->V.get() + (size_t)a * p, vv.get(), p);
      gemm_f64(c, false, false, p, m, 1, 1.0, rr.get(), p, nullptr, qq.get(), 1, nullptr, 1.0,
               Bcur.get(), p);
      while (i_out < L && a == nc[i_out] - 1) {
        if (M->b_stored) copy_f64(c, M->B.get() + (size_t)i_out * p * m, Bcur.get(), p * m);
        if (o.fit) {
          // reference quirk kept: Yfit_cur only accumulates at requested ncomp
          gemm_f64(c, false, false, n, m, 1, 1.0, tt.get(), n, nullptr, qq.get(), 1, nullptr, 1.0,
                   Yfit_cur.get(), n);
          r2_into(c, Yc.get(), n, m, Yfit_cur.get(), M->R2Y.get() + i_out);
          double* Yf = M->Yfit.get() + (size_t)i_out * n * m;
          copy_f64(c, Yf, Yfit_cur.get(), n * m);
          add_row(c, Yf, n, n, m, M->mY.get());
        }
        ++i_out;
      }
    }
    M->n_components = a;
    build_projection_cache(c, M);
    sync(c);
  }
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
