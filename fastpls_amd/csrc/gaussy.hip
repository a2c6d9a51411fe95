// Gaussian response compression (the reference's gaussian_y option) and its
// decoders, R/main.R:1945-2142 and src/fastPLS.cpp:4203-4215:
//
//   regression      y_center = colMeans(Y), Yc = Y - y_center,
//                   P (m x dim) = N(0, 1) / sqrt(dim) (column-major fill),
//                   Z = Yc P, decoder = (Z^T Z + ridge I)^{-1} Z^T Yc with
//                   ridge = 1e-8 mean(diag Z^T Z) (1e-8 when that is not a
//                   positive finite number); decode: Yhat = Zhat decoder + y_center.
//   classification  codes (C x dim) = N(0, 1) / sqrt(dim), columns centred and
//                   scaled to unit norm, rows scaled to unit norm (zero or
//                   non-finite norms -> 1); Z = codes[y, ]; decode: scores =
//                   2 Zhat codes^T - ||code_c||^2, label = first maximum.
//
// The reference draws P and the codes from R's RNG (set.seed + rnorm); this
// library draws them from the std::mt19937 + normal_distribution<double>
// stream of the seed that its rSVD sketch uses.  Products and the decoder
// solve run in FP64 on the device (gemm_f64; spd_solve_slices with the ridge
// rule above and its LU fallback for an indefinite Gram).
#include <algorithm>
#include <cmath>
#include <vector>

#include "model.h"

namespace fpls {

namespace {

unsigned gyg(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

// out (rows x cols, ld) <- src (FP64, ld rows), as FP32 or FP64
__global__ void k_gy_emit(const double* __restrict__ src, int64_t rows, int64_t cols,
                          void* __restrict__ out, int64_t ld, int f32) {
  const int64_t tot = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    if (f32) static_cast<float*>(out)[i + j * ld] = static_cast<float>(src[e]);
    else static_cast<double*>(out)[i + j * ld] = src[e];
  }
}

// Z (n x dim, FP64) = codes[y_i - 1, :]; codes C x dim (ld C)
__global__ void k_gy_gather(const double* __restrict__ codes, int64_t C, int64_t dim,
                            const int32_t* __restrict__ y, int64_t n, double* __restrict__ Z,
                            int32_t* __restrict__ bad) {
  const int64_t tot = n * dim;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, j = e / n;
    const int32_t c = y[i];
    if (c < 1 || c > C) {
      if (j == 0) atomicAdd(bad, 1);
      Z[e] = 0.0;
    } else {
      Z[e] = codes[(c - 1) + j * C];
    }
  }
}

// Y (n x m) += 1 center^T
__global__ void k_gy_add_center(double* __restrict__ Y, int64_t n, int64_t m,
                                const double* __restrict__ center) {
  const int64_t tot = n * m;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x)
    Y[e] += center[e / n];
}

}  // namespace

// caller output matrix (host or device, FP32 or FP64) <- device FP64 (ld rows)
static void emit(Ctx* c, const fpls_mat* out, const double* src, int64_t rows, int64_t cols,
                 const char* what) {
  FPLS_REQUIRE(out && out->ptr, FPLS_ERR_ARG, std::string(what) + ": output matrix is NULL");
  FPLS_REQUIRE(out->rows == rows && out->cols == cols && out->ld >= std::max<int64_t>(rows, 1),
               FPLS_ERR_ARG, std::string(what) + ": output matrix has the wrong shape");
  FPLS_REQUIRE(out->dtype == FPLS_F32 || out->dtype == FPLS_F64, FPLS_ERR_ARG,
               std::string(what) + ": output dtype must be FPLS_F32 or FPLS_F64");
  if (rows == 0 || cols == 0) return;
  const int f32 = out->dtype == FPLS_F32 ? 1 : 0;
  void* dst = const_cast<void*>(out->ptr);
  if (out->loc == FPLS_DEVICE) {
    launch(c, "gy_emit", k_gy_emit, dim3(gyg(rows * cols)), dim3(256), 0, src, rows, cols, dst,
           out->ld, f32);
    return;
  }
  FPLS_REQUIRE(out->loc == FPLS_HOST, FPLS_ERR_ARG, std::string(what) + ": bad location");
  const size_t es = f32 ? 4 : 8;
  DBuf<unsigned char> tmp(c, (size_t)rows * cols * es);
  launch(c, "gy_emit", k_gy_emit, dim3(gyg(rows * cols)), dim3(256), 0, src, rows, cols,
         (void*)tmp.get(), rows, f32);
  FPLS_HIP(hipMemcpy2DAsync(dst, out->ld * es, tmp.get(), rows * es, rows * es, cols,
                            hipMemcpyDeviceToHost, c->stream));
  FPLS_HIP(hipStreamSynchronize(c->stream));
}

// a staged matrix as a packed FP64 device copy
static DBuf<double> as_f64(Ctx* c, const DMat& A) {
  DBuf<double> out(c, std::max<size_t>(1, (size_t)A.rows * A.cols));
  if (A.dtype == FPLS_F64) {
    copy2d_f64(c, out.get(), A.rows, A.d(), A.ld, A.rows, A.cols);
  } else if (A.ld == A.rows) {
    f32_to_f64(c, out.get(), A.f(), A.rows * A.cols);
  } else {
    for (int64_t j = 0; j < A.cols; ++j) f32_to_f64(c, out.get() + j * A.rows, A.f() + j * A.ld, A.rows);
  }
  return out;
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_gaussian_y_encode(fpls_ctx* ctx, const fpls_mat* Ym, int32_t dim,
                                      uint32_t seed, const fpls_mat* Z_out, double* center,
                                      double* decoder, double* projection) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> hy;
    DMat Y = stage(c, Ym, hy, "gaussian_y_encode: Y");
    const int64_t n = Y.rows, m = Y.cols;
    FPLS_REQUIRE(n >= 1 && m >= 1, FPLS_ERR_ARG, "gaussian_y_encode: Y is empty");
    FPLS_REQUIRE(dim >= 1, FPLS_ERR_ARG, "gaussian_y_dim must be a positive integer");
    FPLS_REQUIRE(center && decoder, FPLS_ERR_ARG, "gaussian_y_encode: center / decoder missing");
    // FP64 Y is read in place; FP32 Y is widened once
    DBuf<double> Yd;
    const double* yd = Y.d();
    int64_t ldy = Y.ld;
    if (Y.dtype != FPLS_F64) {
      Yd = as_f64(c, Y);
      yd = Yd.get();
      ldy = n;
    }
    // y_center and Yc
    DBuf<double> cs(c, m);
    DMat Yv;
    Yv.p = yd;
    Yv.dtype = FPLS_F64;
    Yv.rows = n;
    Yv.cols = m;
    Yv.ld = ldy;
    col_sums(c, Yv, cs.get());
    std::vector<double> hc(m);
    to_host(c, hc.data(), cs.get(), m);
    for (double& v : hc) v /= (double)n;
    to_device(c, cs.get(), hc.data(), m);
    std::copy(hc.begin(), hc.end(), center);
    // projection P = N(0, 1) / sqrt(dim), m x dim column-major
    std::vector<double> P((size_t)m * dim);
    gaussian_host(P.data(), (int64_t)P.size(), seed);
    const double sc = 1.0 / std::sqrt((double)dim);
    for (double& v : P) v *= sc;
    if (projection) std::copy(P.begin(), P.end(), projection);
    DBuf<double> Pd(c, P.size()), Z(c, (size_t)n * dim), G(c, (size_t)dim * dim),
        R(c, (size_t)dim * m), Gi(c, (size_t)dim * dim), D(c, (size_t)dim * m);
    to_device(c, Pd.get(), P.data(), P.size());
    // Z = (Y - 1 center^T) P
    gemm_f64(c, false, false, n, dim, m, 1.0, yd, ldy, cs.get(), Pd.get(), m, nullptr, 0.0,
             Z.get(), n, nullptr, "gy_encode");
    // decoder = (Z^T Z + ridge I)^{-1} Z^T Yc
    gemm_f64(c, true, false, dim, dim, n, 1.0, Z.get(), n, nullptr, Z.get(), n, nullptr, 0.0,
             G.get(), dim);
    {
      // R = Z^T Yc as (Yc^T Z)^T: the long side (m) on the rows of the product
      DBuf<double> Rt(c, (size_t)m * dim);
      gemm_f64(c, true, false, m, dim, n, 1.0, yd, ldy, cs.get(), Z.get(), n, nullptr, 0.0,
               Rt.get(), m, nullptr, "gy_cross");
      transpose_f64(c, Rt.get(), m, m, dim, R.get(), dim);
      sync(c);
    }
    DBuf<int32_t> st(c, 1);
    spd_solve_slices(c, G.get(), dim, dim, std::vector<int>{dim}, 1, 1, nullptr, 1e-8, Gi.get(),
                     dim, (int64_t)dim * dim, nullptr, st.get());
    int32_t hs = 0;
    to_host_i32(c, &hs, st.get(), 1);
    FPLS_REQUIRE(hs != 1, FPLS_ERR_SOLVE, "gaussian_y decoder: singular system");
    gemm_f64(c, false, false, dim, m, dim, 1.0, Gi.get(), dim, nullptr, R.get(), dim, nullptr, 0.0,
             D.get(), dim);
    to_host(c, decoder, D.get(), (size_t)dim * m);
    if (Z_out) emit(c, Z_out, Z.get(), n, dim, "gaussian_y_encode: Z");
  });
}

extern "C" int fpls_gaussian_y_codes(fpls_ctx* ctx, const fpls_ivec* y, int32_t n_classes,
                                     int32_t dim, uint32_t seed, const fpls_mat* Z_out,
                                     double* codes) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "gaussian_y_codes: no classes");
    FPLS_REQUIRE(dim >= 1, FPLS_ERR_ARG, "gaussian_y_dim must be a positive integer");
    FPLS_REQUIRE(codes, FPLS_ERR_ARG, "gaussian_y_codes: codes output missing");
    const int64_t C = n_classes;
    // codebook on the host (C x dim): R/main.R:1982-1995
    std::vector<double> K((size_t)C * dim);
    gaussian_host(K.data(), (int64_t)K.size(), seed);
    const double sc = 1.0 / std::sqrt((double)dim);
    for (double& v : K) v *= sc;
    for (int64_t j = 0; j < dim; ++j) {
      double* col = K.data() + j * C;
      double mu = 0.0;
      for (int64_t i = 0; i < C; ++i) mu += col[i];
      mu /= (double)C;
      double ss = 0.0;
      for (int64_t i = 0; i < C; ++i) {
        col[i] -= mu;
        ss += col[i] * col[i];
      }
      double sd = std::sqrt(ss);
      if (!std::isfinite(sd) || sd == 0.0) sd = 1.0;
      for (int64_t i = 0; i < C; ++i) col[i] /= sd;
    }
    for (int64_t i = 0; i < C; ++i) {
      double ss = 0.0;
      for (int64_t j = 0; j < dim; ++j) ss += K[i + j * C] * K[i + j * C];
      double nr = std::sqrt(ss);
      if (!std::isfinite(nr) || nr == 0.0) nr = 1.0;
      for (int64_t j = 0; j < dim; ++j) K[i + j * C] /= nr;
    }
    std::copy(K.begin(), K.end(), codes);
    if (!Z_out) return;
    FPLS_REQUIRE(y != nullptr, FPLS_ERR_ARG, "gaussian_y_codes: labels missing");
    const int64_t n = y->n;
    DBuf<int32_t> hy;
    const int32_t* yd = stage_labels(c, y, n, hy, "gaussian_y_codes: y");
    DBuf<double> Kd(c, K.size()), Z(c, std::max<size_t>(1, (size_t)n * dim));
    to_device(c, Kd.get(), K.data(), K.size());
    int32_t* bad = c->defer_check(1, FPLS_ERR_LABEL_RANGE, "gaussian_y labels must be 1..n_classes");
    launch(c, "gy_gather", k_gy_gather, dim3(gyg(n * dim)), dim3(256), 0, (const double*)Kd.get(),
           C, (int64_t)dim, yd, n, Z.get(), bad);
    emit(c, Z_out, Z.get(), n, dim, "gaussian_y_codes: Z");
  });
}

extern "C" int fpls_gaussian_y_decode(fpls_ctx* ctx, const fpls_mat* Zm, const double* decoder,
                                      int32_t m, const double* center, const fpls_mat* Y_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> hz;
    DMat Zs = stage(c, Zm, hz, "gaussian_y_decode: Z");
    const int64_t n = Zs.rows, dim = Zs.cols;
    FPLS_REQUIRE(m >= 1 && decoder && center, FPLS_ERR_ARG, "gaussian_y_decode: decoder missing");
    DBuf<double> Zd = as_f64(c, Zs), Dd(c, (size_t)dim * m), cd(c, m),
        Y(c, std::max<size_t>(1, (size_t)n * m));
    to_device(c, Dd.get(), decoder, (size_t)dim * m);
    to_device(c, cd.get(), center, m);
    if (n > 0) {
      gemm_f64(c, false, false, n, m, dim, 1.0, Zd.get(), n, nullptr, Dd.get(), dim, nullptr, 0.0,
               Y.get(), n, nullptr, "gy_decode");
      launch(c, "gy_center", k_gy_add_center, dim3(gyg(n * m)), dim3(256), 0, Y.get(), n,
             (int64_t)m, (const double*)cd.get());
    }
    emit(c, Y_out, Y.get(), n, m, "gaussian_y_decode: Y");
  });
}

extern "C" int fpls_gaussian_y_classes(fpls_ctx* ctx, const fpls_mat* Zm, const double* codes,
                                       int32_t n_classes, int32_t* labels, double* scores) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> hz;
    DMat Zs = stage(c, Zm, hz, "gaussian_y_classes: Z");
    const int64_t n = Zs.rows, dim = Zs.cols, C = n_classes;
    FPLS_REQUIRE(C >= 1 && codes && labels, FPLS_ERR_ARG, "gaussian_y_classes: codes missing");
    if (n == 0) return;
    std::vector<double> nrm(C, 0.0);
    for (int64_t i = 0; i < C; ++i)
      for (int64_t j = 0; j < dim; ++j) nrm[i] -= codes[i + j * C] * codes[i + j * C];
    DBuf<double> Zd = as_f64(c, Zs), Kd(c, (size_t)C * dim), nd(c, C), S(c, (size_t)n * C);
    DBuf<int32_t> lab(c, n);
    to_device(c, Kd.get(), codes, (size_t)C * dim);
    to_device(c, nd.get(), nrm.data(), C);
    // scores = 2 Z codes^T - ||code||^2 (nearest_code_classes, src/fastPLS.cpp:4203-4215)
    gemm_f64(c, false, true, n, C, dim, 2.0, Zd.get(), n, nullptr, Kd.get(), C, nullptr, 0.0,
             S.get(), n);
    argmax_rows(c, S.get(), n, n, C, nd.get(), lab.get());
    to_host_i32(c, labels, lab.get(), n);
    if (scores) {
      launch(c, "gy_center", k_gy_add_center, dim3(gyg(n * C)), dim3(256), 0, S.get(), n, C,
             (const double*)nd.get());
      to_host(c, scores, S.get(), (size_t)n * C);
    }
  });
}
