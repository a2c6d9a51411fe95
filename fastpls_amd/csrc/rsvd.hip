// Truncated SVD on the device, FP64:
//
//  * svd_exact_dev -- thin SVD through the Gram matrix of the narrow side and
//                     the symmetric eigensolver (one workgroup up to order
//                     128, block Jacobi above);
//                     stands in for svd_econ (src/svd_cpu_exact.cpp:7-38).
//  * rsvd_dev      -- randomized range finder of src/svd_cpu_rsvd.cpp:95-124
//                     with the finalisation of :26-93.  The sketch is the
//                     reference's own mt19937 Gaussian stream (or a caller
//                     sketch).  Householder QR is replaced by shifted
//                     CholeskyQR3 (same column span; QR signs do not reach the
//                     returned factors because Q enters only through Q Uhat).
//  * fpls_rsvd / fpls_rsvd_xprod -- the C ABI (fastsvd and the matrix-free
//                     crossprod rSVD of src/fastPLS.cpp:550-623).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "model.h"

namespace fpls {

__global__ void k_sqrt_clamp(const double* __restrict__ w, double* __restrict__ s, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = sqrt(fmax(w[i], 0.0));
}

// Left factor of the exact SVD from the eigenpairs of A^T A:
//   s_j = sqrt(max(w_j, 0)),  U(:, j) = A E(:, j) / s_j  (s_j = 0 leaves A E),
// one output per thread (four accumulators over the cols-long product), the
// columns of E broadcast per block; block (0, j) writes s_j.
__global__ __launch_bounds__(256) void k_svd_left(const double* __restrict__ A, int64_t lda,
                                                  int64_t rows, int64_t cols,
                                                  const double* __restrict__ E, int64_t lde,
                                                  const double* __restrict__ w,
                                                  double* __restrict__ U,
                                                  double* __restrict__ s) {
  const int64_t j = blockIdx.y, i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const double sj = sqrt(fmax(w[j], 0.0));
  if (blockIdx.x == 0 && threadIdx.x == 0) s[j] = sj;
  if (i >= rows) return;
  const double* e = E + j * lde;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t c = 0;
  for (; c + 4 <= cols; c += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = fma(A[i + (c + u) * lda], e[c + u], a[u]);
  }
  for (; c < cols; ++c) a[0] = fma(A[i + c * lda], e[c], a[0]);
  const double v = (a[0] + a[1]) + (a[2] + a[3]);
  U[i + j * rows] = sj != 0.0 ? v / sj : v;
}

void orth_cholqr(Ctx* ctx, double* Y, int64_t rows, int l) {
  DBuf<double> G(ctx, (size_t)l * l), Linv(ctx, (size_t)l * l), tmp(ctx, (size_t)rows * l);
  // a Cholesky that still breaks down after its shifted retries (status -1,
  // e.g. a non-finite Gram) fails the entry point when it returns
  int32_t* st = ctx->defer_check(3, FPLS_ERR_SOLVE,
                                 "Cholesky-QR of the range finder failed (non-finite sketch?)", -1);
  for (int pass = 0; pass < 3; ++pass) {
    gemm_f64(ctx, true, false, l, l, rows, 1.0, Y, rows, nullptr, Y, rows, nullptr, 0.0, G.get(),
             l);
    chol_inv_lower(ctx, G.get(), l, l, rows, pass == 0 ? -1.0 : 0.0, Linv.get(), st + pass);
    // Y <- Y L^{-T}
    gemm_f64(ctx, false, true, rows, l, l, 1.0, Y, rows, nullptr, Linv.get(), l, nullptr, 0.0,
             tmp.get(), rows);
    copy_f64(ctx, Y, tmp.get(), rows * (int64_t)l);
  }
}

void svd_exact_dev(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols, int k,
                   SvdOut& out, bool left_only) {
  const int64_t mn = std::min(rows, cols);
  FPLS_REQUIRE(mn >= 1, FPLS_ERR_RANK, "svd: empty matrix");
  // orders above 128 take the block-Jacobi eigensolver (k_jacobi.hip)
  const int kk = (int)std::min<int64_t>(mn, std::max(k, 1));
  const int g = (int)mn;
  DBuf<double> Gm(ctx, (size_t)g * g), w(ctx, g), E(ctx, (size_t)g * g);
  out.s = DBuf<double>(ctx, kk);
  out.U = DBuf<double>(ctx, (size_t)rows * kk);
  if (!left_only && !(out.want_v && cols <= rows)) out.Vt = DBuf<double>(ctx, (size_t)kk * cols);
  out.k = kk;
  if (cols <= rows) {
    gemm_f64(ctx, true, false, cols, cols, rows, 1.0, A, lda, nullptr, A, lda, nullptr, 0.0,
             Gm.get(), g);
    sym_eig_large(ctx, Gm.get(), g, g, w.get(), E.get(), kk);
    // s = sqrt(w), U = A V / s in one pass (the product is rows x cols x kk)
    launch(ctx, "svd_left", k_svd_left, dim3((unsigned)((rows + 255) / 256), (unsigned)kk),
           dim3(256), 0, A, lda, rows, cols, (const double*)E.get(), (int64_t)g,
           (const double*)w.get(), out.U.get(), out.s.get());
    if (!left_only) {
      if (out.want_v) {
        out.V = std::move(E);
        out.Vt = DBuf<double>();
      } else {
        transpose_f64(ctx, E.get(), g, cols, kk, out.Vt.get(), kk);
      }
    }
  } else {
    gemm_f64(ctx, false, true, rows, rows, cols, 1.0, A, lda, nullptr, A, lda, nullptr, 0.0,
             Gm.get(), g);
    sym_eig_large(ctx, Gm.get(), g, g, w.get(), E.get(), kk);
    launch(ctx, "sqrt", k_sqrt_clamp, dim3((kk + 255) / 256), dim3(256), 0, w.get(), out.s.get(),
           kk);
    copy2d_f64(ctx, out.U.get(), rows, E.get(), g, rows, kk);
    if (!left_only) {
      // Vt = U^T A / s (row-wise)
      gemm_f64(ctx, true, false, kk, cols, rows, 1.0, out.U.get(), rows, nullptr, A, lda,
               nullptr, 0.0, out.Vt.get(), kk);
      scale_rows_div(ctx, out.Vt.get(), kk, kk, cols, out.s.get());
    }
  }
}

void rsvd_dev(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols, int k,
              int oversample, int power, uint32_t seed, const double* omega_host,
              bool left_only, bool use_full, SvdOut& out) {
  FPLS_REQUIRE(k >= 1, FPLS_ERR_ARG, "truncated_svd: k must be >= 1");
  const int64_t max_rank = std::min(rows, cols);
  const int target = (int)std::min<int64_t>(max_rank, std::max(k, 1));
  const int l = (int)std::min<int64_t>(max_rank, (int64_t)target + std::max(oversample, 0));
  // dispatch rules of src/svd_iface.cpp:64-70 and src/svd_cpu_rsvd.cpp:100-105;
  // a sketch wider than the one-workgroup factorisations (l > 128) takes the
  // exact SVD, whose leading factors the range finder approximates
  if (use_full || max_rank < 6 || l >= max_rank || l > kJacobiMax) {
    svd_exact_dev(ctx, A, lda, rows, cols, target, out, left_only);
    return;
  }
  std::vector<double> om((size_t)cols * l);
  if (omega_host) std::copy(omega_host, omega_host + om.size(), om.begin());
  else gaussian_host(om.data(), (int64_t)om.size(), seed);
  DBuf<double> Om(ctx, om.size()), Y(ctx, (size_t)rows * l), Z(ctx, (size_t)cols * l);
  to_device(ctx, Om.get(), om.data(), om.size());
  gemm_f64(ctx, false, false, rows, l, cols, 1.0, A, lda, nullptr, Om.get(), cols, nullptr, 0.0,
           Y.get(), rows);
  const int q = std::max(power, 0);
  if (q == 1) {
    gemm_f64(ctx, true, false, cols, l, rows, 1.0, A, lda, nullptr, Y.get(), rows, nullptr, 0.0,
             Z.get(), cols);
    gemm_f64(ctx, false, false, rows, l, cols, 1.0, A, lda, nullptr, Z.get(), cols, nullptr, 0.0,
             Y.get(), rows);
  } else {
    for (int it = 0; it < q; ++it) {
      gemm_f64(ctx, true, false, cols, l, rows, 1.0, A, lda, nullptr, Y.get(), rows, nullptr,
               0.0, Z.get(), cols);
      orth_cholqr(ctx, Z.get(), cols, l);
      gemm_f64(ctx, false, false, rows, l, cols, 1.0, A, lda, nullptr, Z.get(), cols, nullptr,
               0.0, Y.get(), rows);
    }
  }
  orth_cholqr(ctx, Y.get(), rows, l);  // Y now holds Q
  DBuf<double> Bm(ctx, (size_t)l * cols);
  gemm_f64(ctx, true, false, l, cols, rows, 1.0, Y.get(), rows, nullptr, A, lda, nullptr, 0.0,
           Bm.get(), l);
  if (l <= cols) {
    // eig(B B^T), descending; usable count as in src/svd_cpu_rsvd.cpp:45-56
    DBuf<double> BB(ctx, (size_t)l * l), w(ctx, l), E(ctx, (size_t)l * l);
    gemm_f64(ctx, false, true, l, l, cols, 1.0, Bm.get(), l, nullptr, Bm.get(), l, nullptr, 0.0,
             BB.get(), l);
    sym_eig(ctx, BB.get(), l, l, w.get(), E.get(), target);
    std::vector<double> wh(l);
    to_host(ctx, wh.data(), w.get(), l);
    const double tol = DBL_EPSILON * (double)std::max<int64_t>(l, cols) * std::max(wh[0], 1.0);
    int usable = 0;
    while (usable < l && wh[usable] > tol) ++usable;
    usable = std::min(usable, target);
    if (usable > 0) {
      out.k = usable;
      out.s = DBuf<double>(ctx, usable);
      out.U = DBuf<double>(ctx, (size_t)rows * usable);
      launch(ctx, "sqrt", k_sqrt_clamp, dim3((usable + 255) / 256), dim3(256), 0, w.get(),
             out.s.get(), usable);
      gemm_f64(ctx, false, false, rows, usable, l, 1.0, Y.get(), rows, nullptr, E.get(), l,
               nullptr, 0.0, out.U.get(), rows);
      if (!left_only) {
        out.Vt = DBuf<double>(ctx, (size_t)usable * cols);
        gemm_f64(ctx, true, false, usable, cols, l, 1.0, E.get(), l, nullptr, Bm.get(), l,
                 nullptr, 0.0, out.Vt.get(), usable);
        scale_rows_div(ctx, out.Vt.get(), usable, usable, cols, out.s.get());
      }
      return;
    }
  }
  // svd_econ(B) fallback, then U = Q Uhat
  SvdOut sb;
  svd_exact_dev(ctx, Bm.get(), l, l, cols, target, sb, left_only);
  out.k = sb.k;
  out.s = std::move(sb.s);
  out.U = DBuf<double>(ctx, (size_t)rows * sb.k);
  gemm_f64(ctx, false, false, rows, sb.k, l, 1.0, Y.get(), rows, nullptr, sb.U.get(), l, nullptr,
           0.0, out.U.get(), rows);
  if (!left_only) out.Vt = std::move(sb.Vt);
}

}  // namespace fpls

using namespace fpls;

static DBuf<double> as_f64(Ctx* ctx, const DMat& A) {
  DBuf<double> out(ctx, std::max<int64_t>(1, A.rows * A.cols));
  if (A.dtype == FPLS_F64) {
    copy2d_f64(ctx, out.get(), A.rows, A.d(), A.ld, A.rows, A.cols);
  } else {
    // convert column by column (ld may exceed rows)
    for (int64_t j = 0; j < A.cols; ++j) f32_to_f64(ctx, out.get() + j * A.rows, A.f() + j * A.ld, A.rows);
  }
  return out;
}

extern "C" int fpls_rsvd(fpls_ctx* ctx, const fpls_mat* Am, int32_t k, const fpls_opts* opts,
                         const fpls_mat* omega, double* U, double* s, double* Vt,
                         int32_t* k_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hold;
    DMat A = stage(c, Am, hold, "fpls_rsvd: A");
    FPLS_REQUIRE(k >= 1, FPLS_ERR_ARG, "truncated_svd: k must be >= 1");
    DBuf<double> A64 = as_f64(c, A);
    std::vector<double> om;
    const double* omp = nullptr;
    if (omega) {
      FPLS_REQUIRE(omega->loc == FPLS_HOST && omega->dtype == FPLS_F64, FPLS_ERR_ARG,
                   "fpls_rsvd: omega must be a host FP64 matrix");
      om.resize((size_t)omega->rows * omega->cols);
      const double* src = static_cast<const double*>(omega->ptr);
      for (int64_t j = 0; j < omega->cols; ++j)
        for (int64_t i = 0; i < omega->rows; ++i) om[i + j * omega->rows] = src[i + j * omega->ld];
      const int64_t max_rank = std::min(A.rows, A.cols);
      const int64_t tgt = std::min<int64_t>(max_rank, k);
      const int64_t l = std::min<int64_t>(max_rank, tgt + std::max(o.oversample, 0));
      FPLS_REQUIRE(omega->rows == A.cols && omega->cols == l, FPLS_ERR_ARG,
                   "fpls_rsvd: omega must be cols(A) x (k + oversample)");
      omp = om.data();
    }
    SvdOut r;
    const bool full = o.svd_method == FPLS_SVD_EXACT;
    rsvd_dev(c, A64.get(), A.rows, A.rows, A.cols, k, o.oversample, o.power, o.seed, omp,
             Vt == nullptr, full, r);
    if (k_out) *k_out = r.k;
    if (U) to_host(c, U, r.U.get(), (size_t)A.rows * r.k);
    if (s) to_host(c, s, r.s.get(), r.k);
    if (Vt && r.Vt.get()) to_host(c, Vt, r.Vt.get(), (size_t)r.k * A.cols);
  });
}

namespace fpls {

// out (rows x l, FP64) = (A - 1 shift^T) M, A the data matrix (FP32 on MFMA or
// FP64), M cols x l; shift (per column of A, FP64) may be null
static void data_times(Ctx* ctx, const DMat& A, const double* M, int l, double* out,
                       int precision, const double* shift = nullptr) {
  if (A.dtype == FPLS_F32) {
    DBuf<float> Mf(ctx, (size_t)A.cols * l);
    f64_to_f32(ctx, Mf.get(), M, A.cols * (int64_t)l);
    gemm_nn_f32(ctx, precision, A.rows, l, A.cols, A.f(), A.ld, shift, Mf.get(), A.cols,
                nullptr, nullptr, out, A.rows);
  } else {
    gemm_f64(ctx, false, false, A.rows, l, A.cols, 1.0, A.d(), A.ld, shift, M, A.cols, nullptr,
             0.0, out, A.rows);
  }
}

// out (cols x l, FP64) = (A - 1 shift^T)^T N, N rows x l
static void data_t_times(Ctx* ctx, const DMat& A, const double* N, int l, double* out,
                         int precision, const double* shift = nullptr) {
  if (A.dtype == FPLS_F32) {
    DBuf<float> Nf(ctx, (size_t)A.rows * l);
    f64_to_f32(ctx, Nf.get(), N, A.rows * (int64_t)l);
    gemm_tn_f32(ctx, precision, A.cols, l, A.rows, A.f(), A.ld, shift, Nf.get(), A.rows,
                nullptr, nullptr, out, A.cols);
  } else {
    gemm_f64(ctx, true, false, A.cols, l, A.rows, 1.0, A.d(), A.ld, shift, N, A.rows, nullptr,
             0.0, out, A.cols);
  }
}

// truncated_rsvd_crossprod_double (src/fastPLS.cpp:550-623) and
// finalize_rsvd_from_q_b_double (:434-475) for S = Xs^T Yc without forming
// S: Xs = (X - 1 mX^T) diag(1 / vX), Yc = Yd - 1 mY^T, both applied on load
// (vX null: no scaling), a_times(M) = Xs^T (Yc M), at_times(M) = Yc^T (Xs M);
// Yd FP64 (n x m, ld n), om_host the m x l sketch (column-major).
void rsvd_xprod_core(Ctx* c, const DMat& X, const double* mX, const double* vX,
                     const double* Yd, int64_t m, const double* mY, int target, int l,
                     int power, const double* om_host, int precision, bool want_v,
                     SvdOut& r) {
  const int64_t n = X.rows, p = X.cols;
  DBuf<double> Om(c, (size_t)m * l), Ys(c, (size_t)p * l), Z(c, (size_t)m * l),
      Nn(c, (size_t)n * l), Ms;
  if (vX) Ms = DBuf<double>(c, (size_t)p * l);
  to_device(c, Om.get(), om_host, (size_t)m * l);
  auto xs_times = [&](const double* M, double* out) {  // Xs M (n x l)
    const double* Mx = M;
    if (vX) {
      copy_f64(c, Ms.get(), M, p * (int64_t)l);
      scale_rows_div(c, Ms.get(), p, p, l, vX);
      Mx = Ms.get();
    }
    data_times(c, X, Mx, l, out, precision, mX);
  };
  auto a_times = [&](const double* M, double* out) {  // Xs^T (Yc M)
    gemm_f64(c, false, false, n, l, m, 1.0, Yd, n, mY, M, m, nullptr, 0.0, Nn.get(), n);
    data_t_times(c, X, Nn.get(), l, out, precision, mX);
    if (vX) scale_rows_div(c, out, p, p, l, vX);
  };
  auto at_times = [&](const double* M, double* out) {  // Yc^T (Xs M)
    xs_times(M, Nn.get());
    gemm_f64(c, true, false, m, l, n, 1.0, Yd, n, mY, Nn.get(), n, nullptr, 0.0, out, m);
  };
  a_times(Om.get(), Ys.get());
  const int q = std::max(power, 0);
  if (q == 1) {
    at_times(Ys.get(), Z.get());
    a_times(Z.get(), Ys.get());
  } else {
    for (int it = 0; it < q; ++it) {
      at_times(Ys.get(), Z.get());
      orth_cholqr(c, Z.get(), m, l);
      a_times(Z.get(), Ys.get());
    }
  }
  orth_cholqr(c, Ys.get(), p, l);  // Q
  // B = (Xs Q)^T Yc  (l x m)
  xs_times(Ys.get(), Nn.get());
  DBuf<double> Bm(c, (size_t)l * m);
  gemm_f64(c, true, false, l, m, n, 1.0, Nn.get(), n, nullptr, Yd, n, mY, 0.0, Bm.get(), l);
  SvdOut sb;
  sb.want_v = want_v;
  svd_exact_dev(c, Bm.get(), l, l, m, target, sb, false);
  r.k = sb.k;
  r.s = std::move(sb.s);
  r.U = DBuf<double>(c, (size_t)p * sb.k);
  gemm_f64(c, false, false, p, sb.k, l, 1.0, Ys.get(), p, nullptr, sb.U.get(), l, nullptr, 0.0,
           r.U.get(), p);
  r.V = std::move(sb.V);
  r.Vt = std::move(sb.Vt);
}

// S = X^T Y materialised in FP64 (p x m)
static void cross(Ctx* ctx, const DMat& X, const DMat& Y, double* S, int precision) {
  const int64_t p = X.cols, m = Y.cols;
  if (X.dtype == FPLS_F32 && Y.dtype == FPLS_F32) {  // both FP32: Y is the MFMA operand as is
    gemm_tn_f32(ctx, precision, p, m, X.rows, X.f(), X.ld, nullptr, Y.f(), Y.ld, nullptr, nullptr,
                S, p);
    return;
  }
  DBuf<double> Yd(ctx, (size_t)Y.rows * m);
  if (Y.dtype == FPLS_F64) copy2d_f64(ctx, Yd.get(), Y.rows, Y.d(), Y.ld, Y.rows, m);
  else
    for (int64_t j = 0; j < m; ++j) f32_to_f64(ctx, Yd.get() + j * Y.rows, Y.f() + j * Y.ld, Y.rows);
  data_t_times(ctx, X, Yd.get(), (int)m, S, precision);
  (void)p;
}

}  // namespace fpls

// truncated_rsvd_crossprod_double (src/fastPLS.cpp:550-623): rSVD of
// S = X^T Y through a_times(M) = X^T (Y M) and at_times(M) = Y^T (X M).
extern "C" int fpls_rsvd_xprod(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym, int32_t k,
                               const fpls_opts* opts, const fpls_mat* omega, double* U, double* s,
                               double* Vt, int32_t* k_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "rsvd_xprod: X");
    DMat Y = stage(c, Ym, hy, "rsvd_xprod: Y");
    FPLS_REQUIRE(X.rows == Y.rows, FPLS_ERR_ARG, "rsvd_xprod: X and Y need the same rows");
    const int64_t p = X.cols, m = Y.cols;
    const int64_t max_rank = std::min(p, m);
    FPLS_REQUIRE(max_rank >= 1, FPLS_ERR_RANK, "rsvd_xprod: empty cross-product");
    const int target = (int)std::min<int64_t>(max_rank, std::max<int>(k, 1));
    const int l = (int)std::min<int64_t>(max_rank, (int64_t)target + std::max(o.oversample, 0));
    SvdOut r;
    std::vector<double> om((size_t)m * l);
    if (omega) {  // a caller sketch (the learned-sketch extension), m x l host FP64
      FPLS_REQUIRE(omega->loc == FPLS_HOST && omega->dtype == FPLS_F64 && omega->rows == m &&
                       omega->cols == l,
                   FPLS_ERR_ARG, "rsvd_xprod: omega must be a host FP64 cols(Y) x (k + oversample)");
      const double* src = static_cast<const double*>(omega->ptr);
      for (int64_t j = 0; j < l; ++j)
        for (int64_t i = 0; i < m; ++i) om[i + j * m] = src[i + j * omega->ld];
    }
    // wider sketches than the one-workgroup factorisations hold take the
    // exact SVD of the formed S (as rsvd_dev does)
    const bool exact = o.svd_method == FPLS_SVD_EXACT || l >= max_rank || l > kJacobiMax;
    // the exact route never reads the sketch: draw it only for the range finder
    if (!omega && !exact) gaussian_host(om.data(), (int64_t)om.size(), o.seed);
    if (exact || o.xprod_route == 1) {
      // S formed once (one pass over X), then the exact SVD or the range
      // finder on S with the same sketch and power schedule
      DBuf<double> S(c, (size_t)p * m);
      cross(c, X, Y, S.get(), o.precision);
      rsvd_dev(c, S.get(), p, p, m, target, o.oversample, o.power, o.seed, om.data(),
               Vt == nullptr, exact, r);
    } else {
      const int64_t n = X.rows;
      // Y and X operands in FP64 scratch for the small products
      DBuf<double> Yd(c, (size_t)n * m);
      if (Y.dtype == FPLS_F64) copy2d_f64(c, Yd.get(), n, Y.d(), Y.ld, n, m);
      else if (Y.ld == n) f32_to_f64(c, Yd.get(), Y.f(), n * m);
      else
        for (int64_t j = 0; j < m; ++j) f32_to_f64(c, Yd.get() + j * n, Y.f() + j * Y.ld, n);
      DMat Yv;
      Yv.p = Yd.get();
      Yv.dtype = FPLS_F64;
      Yv.rows = n;
      Yv.cols = m;
      Yv.ld = n;
      DBuf<double> Om(c, om.size()), Ys(c, (size_t)p * l), Z(c, (size_t)m * l), Nn(c, (size_t)n * l);
      to_device(c, Om.get(), om.data(), om.size());
      auto a_times = [&](const double* M, double* out) {  // X^T (Y M)
        data_times(c, Yv, M, l, Nn.get(), o.precision);
        data_t_times(c, X, Nn.get(), l, out, o.precision);
      };
      auto at_times = [&](const double* M, double* out) {  // Y^T (X M)
        data_times(c, X, M, l, Nn.get(), o.precision);
        data_t_times(c, Yv, Nn.get(), l, out, o.precision);
      };
      a_times(Om.get(), Ys.get());
      const int q = std::max(o.power, 0);
      if (q == 1) {
        at_times(Ys.get(), Z.get());
        a_times(Z.get(), Ys.get());
      } else {
        for (int it = 0; it < q; ++it) {
          at_times(Ys.get(), Z.get());
          orth_cholqr(c, Z.get(), m, l);
          a_times(Z.get(), Ys.get());
        }
      }
      orth_cholqr(c, Ys.get(), p, l);  // Q
      // B = (X Q)^T Y  (l x m)
      data_times(c, X, Ys.get(), l, Nn.get(), o.precision);
      DBuf<double> Bm(c, (size_t)l * m);
      gemm_f64(c, true, false, l, m, n, 1.0, Nn.get(), n, nullptr, Yd.get(), n, nullptr, 0.0,
               Bm.get(), l);
      // finalize_rsvd_from_q_b_double: svd_econ(B), U = Q Uhat
      SvdOut sb;
      svd_exact_dev(c, Bm.get(), l, l, m, target, sb, Vt == nullptr);
      r.k = sb.k;
      r.s = std::move(sb.s);
      r.U = DBuf<double>(c, (size_t)p * sb.k);
      gemm_f64(c, false, false, p, sb.k, l, 1.0, Ys.get(), p, nullptr, sb.U.get(), l, nullptr,
               0.0, r.U.get(), p);
      if (Vt) r.Vt = std::move(sb.Vt);
    }
    if (k_out) *k_out = r.k;
    if (U) to_host(c, U, r.U.get(), (size_t)p * r.k);
    if (s) to_host(c, s, r.s.get(), r.k);
    if (Vt && r.Vt.get()) to_host(c, Vt, r.Vt.get(), (size_t)r.k * m);
  });
}
