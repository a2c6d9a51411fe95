// Orthogonal-signal filter (OPLS) and kernel matrices for kernel PLS, the
// pre-stages that feed the SIMPLS core (src/fastPLS.cpp:4314-4483):
//
//   kernel_matrix_cpp       K = X1 X2^T (linear), (gamma X1 X2^T + coef0)^degree
//                           (poly), exp(-gamma max0(|x1|^2 + |x2|^2 - 2 X1 X2^T))
//                           (rbf; only values in (-1e-10, 0) are clamped to 0)
//   center_kernel_train_cpp Kc = K - 1 colmeans^T - rowmeans 1^T + grand mean
//   center_kernel_test_cpp  the same with the training column means / grand mean
//   opls_filter_cpp         centre / scale X, then per orthogonal component:
//                           w = first left singular vector of X^T Yc, t = X w,
//                           p = X^T t / t't, w_o = p - w (w'p / w'w) normalised,
//                           t_o = X w_o, p_o = X^T t_o / t_o't_o, X -= t_o p_o^T;
//                           stops early when a norm is zero or not finite
//   opls_apply_filter_cpp   (X - mX) / vX, then X -= (X w_o) p_o^T per component
//
// Everything runs in FP64 on the device, as the reference computes in double:
// the products through gemm_f64 / gemv_n / gemv_t, the first singular vector
// through svd_exact_dev (the eigenpairs of the small Gram), the scalar
// bookkeeping of a component in one workgroup whose checks set a device flag
// (the deflation kernel skips its update when the flag is set), read back
// once per component.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "model.h"

namespace fpls {

namespace {

constexpr int OP_T = 1024;

unsigned opg(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384));
}

__device__ double op_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  __syncthreads();
  if (threadIdx.x == 0) red[0] = s;
  __syncthreads();
  return red[0];
}

// out (rows x cols, FP64, ld ldo) = (X - mX) / vX; mX / vX may be NULL (0 / 1)
template <typename T>
__global__ void k_op_prep(const T* __restrict__ X, int64_t ldx, int64_t rows, int64_t cols,
                          const double* __restrict__ mX, const double* __restrict__ vX,
                          double* __restrict__ out, int64_t ldo) {
  const int64_t tot = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    double v = (double)X[i + j * ldx];
    if (mX) v = v - mX[j];
    if (vX) v = v / vX[j];
    out[i + j * ldo] = v;
  }
}

// mX = sums / n; vX = sqrt(ssd / (n - 1)) with 0 or non-finite -> 1
__global__ void k_op_moments(const double* __restrict__ sums, const double* __restrict__ ssd,
                             int64_t n, int64_t p, int center, int scale, double* __restrict__ mX,
                             double* __restrict__ vX) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= p) return;
  if (center) mX[j] = sums[j] / (double)n;
  if (scale) {
    double s = n > 1 ? sqrt(ssd[j] / (double)(n - 1)) : 0.0;
    if (!isfinite(s) || s == 0.0) s = 1.0;
    vX[j] = s;
  }
}

// v /= |v| (flag on a zero or non-finite norm)
__global__ __launch_bounds__(OP_T) void k_op_unit(double* __restrict__ v, int64_t n,
                                                 int32_t* __restrict__ flag) {
  __shared__ double red[OP_T / 64];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += OP_T) s += v[i] * v[i];
  const double nrm = sqrt(op_block_sum(s, red));
  if (!isfinite(nrm) || nrm <= 0.0) {
    if (threadIdx.x == 0) flag[0] = 1;
    return;
  }
  for (int64_t i = threadIdx.x; i < n; i += OP_T) v[i] = v[i] / nrm;
}

// v /= *ss (flag unless *ss is finite and positive)
__global__ void k_op_div(double* __restrict__ v, int64_t n, const double* __restrict__ ss,
                         int32_t* __restrict__ flag) {
  const double d = ss[0];
  const bool ok = isfinite(d) && d > 0.0;
  if (!ok) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 1;
    return;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = v[i] / d;
}

// wo = p - w (w'p / w'w), then wo /= |wo|
__global__ __launch_bounds__(OP_T) void k_op_orth(const double* __restrict__ w,
                                                 const double* __restrict__ pv, int64_t n,
                                                 double* __restrict__ wo,
                                                 int32_t* __restrict__ flag) {
  __shared__ double red[OP_T / 64];
  double sww = 0.0, swp = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += OP_T) {
    sww += w[i] * w[i];
    swp += w[i] * pv[i];
  }
  const double ww = op_block_sum(sww, red);
  const double wp = op_block_sum(swp, red);
  const double c = wp / ww;
  double so = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += OP_T) {
    const double v = pv[i] - __dmul_rn(w[i], c);
    wo[i] = v;
    so += v * v;
  }
  const double nrm = sqrt(op_block_sum(so, red));
  if (!isfinite(nrm) || nrm <= 0.0) {
    if (threadIdx.x == 0) flag[0] = 1;
    return;
  }
  for (int64_t i = threadIdx.x; i < n; i += OP_T) wo[i] = wo[i] / nrm;
}

// X -= t p^T (skipped when the flag is set)
__global__ void k_op_deflate(double* __restrict__ X, int64_t ld, int64_t rows, int64_t cols,
                             const double* __restrict__ t, const double* __restrict__ p,
                             const int32_t* __restrict__ flag) {
  if (flag && flag[0]) return;
  const int64_t tot = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    X[i + j * ld] = X[i + j * ld] - __dmul_rn(t[i], p[j]);
  }
}

// r[i] = sum_j f(X[i, j]) / div, f = square (SQ) or identity.  A
// workgroup takes 64 rows (one per lane, coalesced down each column) and
// splits the columns over its 16 waves; the 16 partial sums of a row are
// added in wave order through LDS.
constexpr int OP_RR_SEG = 16;
template <bool SQ>
__global__ __launch_bounds__(64 * OP_RR_SEG) void k_op_row_reduce(const double* __restrict__ X,
                                                                  int64_t ld, int64_t rows,
                                                                  int64_t cols, double div,
                                                                  double* __restrict__ r) {
  __shared__ double part[OP_RR_SEG][64];
  const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const int64_t per = (cols + OP_RR_SEG - 1) / OP_RR_SEG;
  const int64_t j0 = seg * per, j1 = std::min<int64_t>(cols, j0 + per);
  double s0 = 0.0, s1 = 0.0;
  if (i < rows) {
    const double* x = X + i;
    int64_t j = j0;
    for (; j + 1 < j1; j += 2) {
      const double a = x[j * ld], b = x[(j + 1) * ld];
      s0 += SQ ? a * a : a;
      s1 += SQ ? b * b : b;
    }
    if (j < j1) {
      const double a = x[j * ld];
      s0 += SQ ? a * a : a;
    }
  }
  part[seg][lane] = s0 + s1;
  __syncthreads();
  if (seg == 0 && i < rows) {
    double s = 0.0;
    for (int q = 0; q < OP_RR_SEG; ++q) s += part[q][lane];
    r[i] = s / div;
  }
}

void row_reduce(Ctx* c, bool sq, const double* X, int64_t ld, int64_t rows, int64_t cols,
                double div, double* r) {
  if (rows <= 0) return;
  const dim3 g((unsigned)((rows + 63) / 64)), b(64 * OP_RR_SEG);
  if (sq) launch(c, "row_sumsq", k_op_row_reduce<true>, g, b, 0, X, ld, rows, cols, div, r);
  else launch(c, "row_means", k_op_row_reduce<false>, g, b, 0, X, ld, rows, cols, div, r);
}

// kernel epilogue on D = X1 X2^T in place: kind 2 rbf, 3 poly
__global__ void k_op_kernel(double* __restrict__ D, int64_t ld, int64_t rows, int64_t cols,
                            int kind, double gamma, double degree, double coef0,
                            const double* __restrict__ n1, const double* __restrict__ n2) {
  const int64_t tot = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    const double d = D[i + j * ld];
    double v;
    if (kind == 3) {
      v = pow(__dadd_rn(__dmul_rn(gamma, d), coef0), degree);
    } else {
      double d2 = __dsub_rn(__dadd_rn(n1[i], n2[j]), __dmul_rn(2.0, d));
      if (d2 < 0.0 && d2 > -1e-10) d2 = 0.0;
      v = exp(__dmul_rn(-gamma, d2));
    }
    D[i + j * ld] = v;
  }
}

// out = ((K - 1 cm^T) - rm 1^T) + gm
__global__ void k_op_center(const double* __restrict__ K, int64_t ldk, int64_t rows, int64_t cols,
                            const double* __restrict__ cm, const double* __restrict__ rm,
                            const double* __restrict__ gm, double* __restrict__ out,
                            int64_t ldo) {
  const double g = gm[0];
  const int64_t tot = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    out[i + j * ldo] = __dadd_rn(__dsub_rn(__dsub_rn(K[i + j * ldk], cm[j]), rm[i]), g);
  }
}

// cm[j] = colsums[j] / rows; gm = sum(cm) / cols (one workgroup)
__global__ __launch_bounds__(OP_T) void k_op_col_means(double* __restrict__ cm, int64_t rows,
                                                      int64_t cols, double* __restrict__ gm) {
  __shared__ double red[OP_T / 64];
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < cols; j += OP_T) {
    const double v = cm[j] / (double)rows;
    cm[j] = v;
    s += v;
  }
  s = op_block_sum(s, red);
  if (threadIdx.x == 0) gm[0] = s / (double)cols;
}

// A caller FP64 output matrix: device with ld == rows is written in place,
// anything else goes through a device buffer and a copy.
struct OutMat {
  double* dev = nullptr;
  int64_t ld = 0;
  DBuf<double> hold;
  const fpls_mat* m = nullptr;
};

OutMat out_mat(Ctx* c, const fpls_mat* o, int64_t rows, int64_t cols, const char* what) {
  FPLS_REQUIRE(o && o->ptr, FPLS_ERR_ARG, std::string(what) + ": output matrix is NULL");
  FPLS_REQUIRE(o->rows == rows && o->cols == cols && o->ld >= std::max<int64_t>(rows, 1),
               FPLS_ERR_ARG, std::string(what) + ": output matrix has the wrong shape");
  FPLS_REQUIRE(o->dtype == FPLS_F64, FPLS_ERR_ARG, std::string(what) + ": output must be FP64");
  OutMat r;
  r.m = o;
  if (o->loc == FPLS_DEVICE) {
    r.dev = const_cast<double*>(static_cast<const double*>(o->ptr));
    r.ld = o->ld;
  } else {
    r.hold = DBuf<double>(c, (size_t)std::max<int64_t>(rows * cols, 1));
    r.dev = r.hold.get();
    r.ld = std::max<int64_t>(rows, 1);
  }
  return r;
}

void finish_out(Ctx* c, OutMat& r, int64_t rows, int64_t cols) {
  if (r.m->loc == FPLS_DEVICE || rows * cols == 0) return;
  double* h = const_cast<double*>(static_cast<const double*>(r.m->ptr));
  if (r.m->ld == rows) {
    to_host(c, h, r.dev, (size_t)rows * cols);
  } else {
    std::vector<double> tmp((size_t)rows * cols);
    to_host(c, tmp.data(), r.dev, tmp.size());
    for (int64_t j = 0; j < cols; ++j)
      std::copy(tmp.begin() + j * rows, tmp.begin() + (j + 1) * rows, h + j * r.m->ld);
  }
}

// (X - mX) / vX into an FP64 device matrix (mX / vX device, may be NULL)
void prep_into(Ctx* c, const DMat& X, const double* mX, const double* vX, double* out,
               int64_t ldo) {
  if (X.rows * X.cols == 0) return;
  if (X.dtype == FPLS_F32)
    launch(c, "opls_prep", k_op_prep<float>, dim3(opg(X.rows * X.cols)), dim3(256), 0, X.f(),
           X.ld, X.rows, X.cols, mX, vX, out, ldo);
  else
    launch(c, "opls_prep", k_op_prep<double>, dim3(opg(X.rows * X.cols)), dim3(256), 0, X.d(),
           X.ld, X.rows, X.cols, mX, vX, out, ldo);
}

}  // namespace

// The filter on a device matrix: X (n x p, FP32 / FP64) preprocessed and
// deflated into Xout (FP64, ld ldo); the response is Y (dense) or the one-hot
// of the labels lab (m classes).  Outputs mX, vX (p), Wo, Po (p x north,
// columns 0..used-1 valid) on the device.
void opls_filter_dev(Ctx* c, const DMat& X, const DMat* Y, const int32_t* lab, int64_t m,
                     int north, int scaling, double* Xout, int64_t ldo, DBuf<double>& mX,
                     DBuf<double>& vX, DBuf<double>& Wo, DBuf<double>& Po, int& used) {
  const int64_t n = X.rows, p = X.cols;
  // preprocessing (scaling 1 centre, 2 centre + sd, 3 none)
  mX = DBuf<double>(c, std::max<int64_t>(p, 1));
  vX = DBuf<double>(c, std::max<int64_t>(p, 1));
  fill_f64(c, mX.get(), 0.0, p);
  fill_f64(c, vX.get(), 1.0, p);
  if (scaling < 3 && n > 0) {
    DBuf<double> sums(c, std::max<int64_t>(p, 1)), ssd(c, std::max<int64_t>(p, 1));
    col_sums(c, X, sums.get());
    launch(c, "opls_moments", k_op_moments, dim3((unsigned)((p + 255) / 256)), dim3(256), 0,
           sums.get(), (const double*)nullptr, n, p, 1, 0, mX.get(), vX.get());
    if (scaling == 2) {
      col_ssd(c, X, mX.get(), ssd.get());
      launch(c, "opls_moments", k_op_moments, dim3((unsigned)((p + 255) / 256)), dim3(256), 0,
             sums.get(), ssd.get(), n, p, 0, 1, mX.get(), vX.get());
    }
    sync(c);
  }
  prep_into(c, X, scaling < 3 ? mX.get() : nullptr, scaling == 2 ? vX.get() : nullptr, Xout, ldo);
  // centred response: dense Y (FP64 copy) or the one-hot of the labels
  DBuf<double> Yd, mY(c, std::max<int64_t>(m, 1)), ysum(c, std::max<int64_t>(m, 1));
  fill_f64(c, mY.get(), 0.0, m);
  if (n > 0) {
    if (Y) {
      Yd = DBuf<double>(c, (size_t)std::max<int64_t>(n * m, 1));
      prep_into(c, *Y, nullptr, nullptr, Yd.get(), n);
      col_sums(c, dview(Yd.get(), n, m, n), ysum.get());
    } else {
      // the one-hot column sums are the class counts
      DBuf<double> ones(c, n);
      fill_f64(c, ones.get(), 1.0, n);
      gemm_f64(c, true, false, 1, m, n, 1.0, ones.get(), n, nullptr, nullptr, 0, nullptr, 0.0,
               ysum.get(), 1, lab);
      sync(c);
    }
    // mY = column sums / n
    launch(c, "opls_moments", k_op_moments, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
           (const double*)ysum.get(), (const double*)nullptr, n, m, 1, 0, mY.get(),
           (double*)nullptr);
  }
  DBuf<double> S(c, (size_t)std::max<int64_t>(p * m, 1));
  DBuf<double> w(c, std::max<int64_t>(p, 1)), pv(c, std::max<int64_t>(p, 1)),
      t(c, std::max<int64_t>(n, 1)), ss(c, 1);
  Wo = DBuf<double>(c, (size_t)std::max<int64_t>(p * north, 1));
  Po = DBuf<double>(c, (size_t)std::max<int64_t>(p * north, 1));
  DBuf<int32_t> flag(c, 1);
  const DMat Xw = dview(Xout, n, p, ldo);
  used = 0;
  for (int a = 0; a < north && n > 0 && p > 0 && m > 0; ++a) {
    fill_i32(c, flag.get(), 0, 1);
    // S = X^T (Y - 1 mY^T)
    if (Y)
      gemm_f64(c, true, false, p, m, n, 1.0, Xout, ldo, nullptr, Yd.get(), n, mY.get(), 0.0,
               S.get(), p);
    else
      gemm_f64(c, true, false, p, m, n, 1.0, Xout, ldo, nullptr, nullptr, 0, mY.get(), 0.0,
               S.get(), p, lab);
    SvdOut sv;
    svd_exact_dev(c, S.get(), p, p, m, 1, sv, true);
    copy_f64(c, w.get(), sv.U.get(), p);
    launch(c, "opls_unit", k_op_unit, dim3(1), dim3(OP_T), 0, w.get(), p, flag.get());
    // t = X w, p = X^T t / t't
    gemv_n(c, Xw, nullptr, w.get(), t.get());
    gemv_t(c, dview(t.get(), n, 1, n), nullptr, t.get(), ss.get());
    gemv_t(c, Xw, nullptr, t.get(), pv.get());
    launch(c, "opls_div", k_op_div, dim3(opg(p)), dim3(256), 0, pv.get(), p,
           (const double*)ss.get(), flag.get());
    double* wo = Wo.get() + (size_t)used * p;
    double* po = Po.get() + (size_t)used * p;
    launch(c, "opls_orth", k_op_orth, dim3(1), dim3(OP_T), 0, (const double*)w.get(),
           (const double*)pv.get(), p, wo, flag.get());
    // t_o = X w_o, p_o = X^T t_o / t_o't_o, X -= t_o p_o^T
    gemv_n(c, Xw, nullptr, wo, t.get());
    gemv_t(c, dview(t.get(), n, 1, n), nullptr, t.get(), ss.get());
    gemv_t(c, Xw, nullptr, t.get(), po);
    launch(c, "opls_div", k_op_div, dim3(opg(p)), dim3(256), 0, po, p, (const double*)ss.get(),
           flag.get());
    launch(c, "opls_deflate", k_op_deflate, dim3(opg(n * p)), dim3(256), 0, Xout, ldo, n, p,
           (const double*)t.get(), (const double*)po, (const int32_t*)flag.get());
    int32_t stop = 0;
    to_host_i32(c, &stop, flag.get(), 1);
    if (stop) break;
    ++used;
  }
}

// Xout (FP64, ld ldo) = the stored filter (device mX, vX, Wo, Po) on X.
void opls_apply_dev(Ctx* c, const DMat& X, const double* mX, const double* vX, const double* Wo,
                    const double* Po, int north, double* Xout, int64_t ldo) {
  const int64_t n = X.rows, p = X.cols;
  prep_into(c, X, mX, vX, Xout, ldo);
  if (north > 0 && n > 0) {
    DBuf<double> t(c, n);
    const DMat Xw = dview(Xout, n, p, ldo);
    for (int a = 0; a < north; ++a) {
      gemv_n(c, Xw, nullptr, Wo + (size_t)a * p, t.get());
      launch(c, "opls_deflate", k_op_deflate, dim3(opg(n * p)), dim3(256), 0, Xout, ldo, n, p,
             (const double*)t.get(), Po + (size_t)a * p, (const int32_t*)nullptr);
    }
    sync(c);
  }
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_opls_filter(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                                const fpls_ivec* y, int32_t n_classes, int32_t north,
                                int32_t scaling, const fpls_mat* X_out, double* mX_out,
                                double* vX_out, double* W_orth, double* P_orth,
                                int32_t* north_used) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(north >= 0, FPLS_ERR_ARG, "north must be >= 0");
    FPLS_REQUIRE(scaling >= 1 && scaling <= 3, FPLS_ERR_ARG, "scaling must be 1, 2 or 3");
    FPLS_REQUIRE((Ym != nullptr) != (y != nullptr), FPLS_ERR_ARG,
                 "opls_filter: give exactly one of Y and y");
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "opls_filter: X");
    const int64_t n = X.rows, p = X.cols;
    DMat Y;
    DBuf<int32_t> hl;
    const int32_t* lab = nullptr;
    int64_t m = 0;
    if (Ym) {
      Y = stage(c, Ym, hy, "opls_filter: Y");
      FPLS_REQUIRE(Y.rows == n, FPLS_ERR_ARG, "X and Y must have the same number of rows");
      m = Y.cols;
    } else {
      FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "opls_filter: n_classes must be positive");
      lab = stage_labels(c, y, n, hl, "opls_filter: y");
      m = n_classes;
    }
    OutMat W = out_mat(c, X_out, n, p, "opls_filter: X_out");
    DBuf<double> mX, vX, Wo, Po;
    int used = 0;
    opls_filter_dev(c, X, Ym ? &Y : nullptr, lab, m, north, scaling, W.dev, W.ld, mX, vX, Wo, Po,
                    used);
    if (north_used) *north_used = used;
    if (mX_out) to_host(c, mX_out, mX.get(), p);
    if (vX_out) to_host(c, vX_out, vX.get(), p);
    if (W_orth && used) to_host(c, W_orth, Wo.get(), (size_t)p * used);
    if (P_orth && used) to_host(c, P_orth, Po.get(), (size_t)p * used);
    finish_out(c, W, n, p);
  });
}

extern "C" int fpls_opls_apply(fpls_ctx* ctx, const fpls_mat* Xm, const double* mX,
                               const double* vX, const double* W_orth, const double* P_orth,
                               int32_t north, const fpls_mat* X_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(mX && vX, FPLS_ERR_ARG, "opls_apply: mX and vX are required");
    FPLS_REQUIRE(north >= 0 && (north == 0 || (W_orth && P_orth)), FPLS_ERR_ARG,
                 "Invalid OPLS orthogonal filter dimensions");
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "opls_apply: X");
    const int64_t n = X.rows, p = X.cols;
    OutMat W = out_mat(c, X_out, n, p, "opls_apply: X_out");
    DBuf<double> dm(c, std::max<int64_t>(p, 1)), dv(c, std::max<int64_t>(p, 1));
    h2d(c, dm.get(), mX, sizeof(double) * p);
    h2d(c, dv.get(), vX, sizeof(double) * p);
    DBuf<double> Wo, Po;
    if (north > 0) {
      Wo = DBuf<double>(c, (size_t)p * north);
      Po = DBuf<double>(c, (size_t)p * north);
      h2d(c, Wo.get(), W_orth, sizeof(double) * p * north);
      h2d(c, Po.get(), P_orth, sizeof(double) * p * north);
    }
    opls_apply_dev(c, X, dm.get(), dv.get(), Wo.get(), Po.get(), north, W.dev, W.ld);
    finish_out(c, W, n, p);
  });
}

extern "C" int fpls_kernel_matrix(fpls_ctx* ctx, const fpls_mat* X1m, const fpls_mat* X2m,
                                  int32_t kernel, double gamma, int32_t degree, double coef0,
                                  const fpls_mat* K_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(kernel >= 1 && kernel <= 3, FPLS_ERR_ARG, "Unknown kernel id");
    DBuf<unsigned char> h1, h2;
    DMat X1 = stage(c, X1m, h1, "kernel_matrix: X1");
    DMat X2 = stage(c, X2m, h2, "kernel_matrix: X2");
    FPLS_REQUIRE(X1.cols == X2.cols, FPLS_ERR_ARG,
                 "X1 and X2 must have the same number of columns");
    const int64_t n1 = X1.rows, n2 = X2.rows, p = X1.cols;
    OutMat K = out_mat(c, K_out, n1, n2, "kernel_matrix: K");
    // FP64 operands (FP32 inputs are widened once)
    DBuf<double> A1, A2;
    const double* a1 = X1.d();
    int64_t l1 = X1.ld;
    const double* a2 = X2.d();
    int64_t l2 = X2.ld;
    if (X1.dtype == FPLS_F32) {
      A1 = DBuf<double>(c, (size_t)std::max<int64_t>(n1 * p, 1));
      prep_into(c, X1, nullptr, nullptr, A1.get(), n1);
      a1 = A1.get();
      l1 = n1;
    }
    if (X2.dtype == FPLS_F32) {
      A2 = DBuf<double>(c, (size_t)std::max<int64_t>(n2 * p, 1));
      prep_into(c, X2, nullptr, nullptr, A2.get(), n2);
      a2 = A2.get();
      l2 = n2;
    }
    if (n1 > 0 && n2 > 0) {
      gemm_f64(c, false, true, n1, n2, p, 1.0, a1, std::max<int64_t>(l1, 1), nullptr, a2,
               std::max<int64_t>(l2, 1), nullptr, 0.0, K.dev, K.ld, nullptr, "kernel_gemm");
      if (kernel != 1) {
        DBuf<double> q1(c, n1), q2(c, n2);
        if (kernel == 2) {
          row_reduce(c, true, a1, l1, n1, p, 1.0, q1.get());
          row_reduce(c, true, a2, l2, n2, p, 1.0, q2.get());
        }
        launch(c, "kernel_epilogue", k_op_kernel, dim3(opg(n1 * n2)), dim3(256), 0, K.dev, K.ld,
               n1, n2, (int)kernel, gamma, (double)degree, coef0, (const double*)q1.get(),
               (const double*)q2.get());
        sync(c);
      }
    }
    finish_out(c, K, n1, n2);
  });
}

static void center_kernel(Ctx* c, const fpls_mat* Km, const fpls_mat* K_out, const double* cm_in,
                          double gm_in, double* cm_out, double* gm_out, const char* what) {
  DBuf<unsigned char> hk;
  DMat K = stage(c, Km, hk, what);
  FPLS_REQUIRE(K.dtype == FPLS_F64, FPLS_ERR_ARG, std::string(what) + ": K must be FP64");
  const int64_t r = K.rows, q = K.cols;
  OutMat O = out_mat(c, K_out, r, q, what);
  DBuf<double> cm(c, std::max<int64_t>(q, 1)), rm(c, std::max<int64_t>(r, 1)), gm(c, 1);
  if (cm_in) {
    h2d(c, cm.get(), cm_in, sizeof(double) * q);
    h2d(c, gm.get(), &gm_in, sizeof(double));
  } else if (r > 0 && q > 0) {
    col_sums(c, K, cm.get());
    launch(c, "kernel_colmeans", k_op_col_means, dim3(1), dim3(OP_T), 0, cm.get(), r, q,
           gm.get());
  }
  if (r > 0 && q > 0) {
    row_reduce(c, false, K.d(), K.ld, r, q, (double)q, rm.get());
    launch(c, "kernel_center", k_op_center, dim3(opg(r * q)), dim3(256), 0, K.d(), K.ld, r, q,
           (const double*)cm.get(), (const double*)rm.get(), (const double*)gm.get(), O.dev,
           O.ld);
  }
  if (cm_out) to_host(c, cm_out, cm.get(), q);
  if (gm_out) to_host(c, gm_out, gm.get(), 1);
  finish_out(c, O, r, q);
  sync(c);
}

extern "C" int fpls_center_kernel_train(fpls_ctx* ctx, const fpls_mat* K, const fpls_mat* K_out,
                                        double* col_means, double* grand_mean) {
  return fpls_guard(ctx, [&] {
    center_kernel(&ctx->c, K, K_out, nullptr, 0.0, col_means, grand_mean, "center_kernel_train");
  });
}

extern "C" int fpls_center_kernel_test(fpls_ctx* ctx, const fpls_mat* K, const double* col_means,
                                       double grand_mean, const fpls_mat* K_out) {
  if (!col_means) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    center_kernel(&ctx->c, K, K_out, col_means, grand_mean, nullptr, nullptr,
                  "center_kernel_test");
  });
}
