// Learned sketch for the rSVD range finder of S = Xc^T Yc (p x m).
//
// The range finder (src/svd_cpu_rsvd.cpp:95-124, matrix free at
// src/fastPLS.cpp:550-623) multiplies S by a Gaussian Omega (m x l) and, for an
// accurate top-k subspace when the spectrum decays slowly, by q power
// iterations (S S^T)^q: every iteration costs two passes over X in the matrix-
// free form.  Here Omega is learned instead, on cross-covariances the method
// generates itself from row-fold draws of (X, Y):
//
//   draws      S_d = Xc_d^T Yc_d for the rows of mask d, all draws of a chunk
//              from one TN product against the masked, shifted responses
//              (per-draw centring by the rank-1 correction of cvb.hip);
//   targets    G_d = S_d^T S_d = V_d diag(sig_d^2) V_d^T (batched one-workgroup
//              eigensolver, m <= 128);
//   loss       L(Omega) = sum_d || U_dk - P(S_d Omega) U_dk ||_F^2
//                       = sum_d [ k - tr(M_d^-1 N_d) ],
//              with W = V_d^T Omega, M_d = W^T Sig^2 W, N_d = W^T Sig_k^2 W (the
//              squared Frobenius sines between range(S_d Omega) and the top-k
//              left singular subspace U_dk; exact, no sampling);
//   gradient   dL/dOmega = V_d 2 (Sig^2 W P - Sig_k^2 W M^-1), P = M^-1 N M^-1
//              (checked against finite differences on the CPU);
//   training   Adam on the summed gradient, Omega re-orthonormalised after each
//              step (the loss depends only on range(Omega)).
//
// Forward / backward run one workgroup per draw (k_sk_fwdbwd) with the small
// matrices in a per-draw global scratch; the gradients are summed in a fixed
// order (deterministic) and one workgroup applies the Adam step and the
// Cholesky-QR re-orthonormalisation.  fpls_sketch_eval measures, per draw,
// sin(theta) between the range finder's subspace (learned or Gaussian Omega,
// q power iterations with the reference's schedule) and the exact top-k left
// singular subspace.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fpls_wave.h"
#include "model.h"

namespace fpls {

namespace {

constexpr int SK_NT = 256;
constexpr int SK_MAXL = 128;

unsigned skg(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

// Z (n x (D m + D), FP32): mask ? Y - gy : 0 per draw and column, then the
// D masks themselves (for the per-draw row sums of Xc).
__global__ void k_sk_zbuild(const int8_t* __restrict__ masks, int64_t n, int D, int m,
                            const float* __restrict__ Yf, const double* __restrict__ Yd,
                            int64_t ldy, const double* __restrict__ gy, float* __restrict__ Z) {
  const int64_t ncols = (int64_t)D * m + D;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * ncols;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, col = e / n;
    float z = 0.f;
    if (col < (int64_t)D * m) {
      const int d = (int)(col / m), c = (int)(col % m);
      if (masks[i + (int64_t)d * n]) {
        const double yv = Yd ? Yd[i + (int64_t)c * ldy] : (double)Yf[i + (int64_t)c * ldy];
        z = (float)(yv - gy[c]);
      }
    } else {
      z = masks[i + (col - (int64_t)D * m) * n] ? 1.f : 0.f;
    }
    Z[e] = z;
  }
}

// S_d[:, c] = C0[:, d m + c] - C0[:, D m + d] (zs[d m + c] / n_d), n_d = zs[D m + d]
__global__ void k_sk_sfinish(const double* __restrict__ C0, const double* __restrict__ zs,
                             int64_t p, int D, int m, double* __restrict__ S) {
  const int64_t tot = p * D * m;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % p, col = e / p;
    const int d = (int)(col / m);
    const double nd = zs[(int64_t)D * m + d];
    const double mu = nd > 0.0 ? zs[col] / nd : 0.0;
    S[e] = C0[j + col * p] - C0[j + ((int64_t)D * m + d) * p] * mu;
  }
}

// One workgroup per draw: loss_d and dL_d/dOmega (m x l) from the draw's
// eigenpairs (V ld m, columns by descending lam) and Omega (m x l).  Scratch
// per draw: Ws (m x l), M, N, Minv, T (l x l).
__global__ __launch_bounds__(SK_NT) void k_sk_fwdbwd(int m, int l, int k,
                                                     const double* __restrict__ Vall,
                                                     const double* __restrict__ lamall,
                                                     const double* __restrict__ Om,
                                                     double* __restrict__ scratch,
                                                     double* __restrict__ loss,
                                                     double* __restrict__ grad) {
  __shared__ double red[SK_NT / 64];
  __shared__ double sig[SK_MAXL];
  const int d = blockIdx.x, t = threadIdx.x;
  const double* V = Vall + (int64_t)d * m * m;
  const double* lam = lamall + (int64_t)d * m;
  double* Ws = scratch + (int64_t)d * (2 * m * l + 4 * l * l);
  double* M = Ws + (int64_t)m * l;
  double* N = M + (int64_t)l * l;
  double* Mi = N + (int64_t)l * l;
  double* T = Mi + (int64_t)l * l;
  double* Gw = T + (int64_t)l * l;  // m x l
  double* G = grad + (int64_t)d * m * l;
  for (int i = t; i < m; i += SK_NT) sig[i] = sqrt(fmax(lam[i], 0.0));
  __syncthreads();
  // Ws = Sig V^T Omega
  for (int e = t; e < m * l; e += SK_NT) {
    const int i = e % m, j = e / m;
    double s = 0.0;
    for (int r = 0; r < m; ++r) s = fma(V[r + i * m], Om[r + j * m], s);
    Ws[e] = sig[i] * s;
  }
  __syncthreads();
  // M = Ws^T Ws, N = Ws_k^T Ws_k
  for (int e = t; e < l * l; e += SK_NT) {
    const int a = e % l, b = e / l;
    double sm = 0.0, sn = 0.0;
    for (int i = 0; i < m; ++i) {
      const double v = Ws[i + a * m] * Ws[i + b * m];
      sm += v;
      if (i < k) sn += v;
    }
    M[e] = sm;
    N[e] = sn;
  }
  __syncthreads();
  // a tiny relative ridge keeps the Cholesky defined for a rank-deficient sketch
  if (t == 0) {
    double tr = 0.0;
    for (int a = 0; a < l; ++a) tr += M[a + a * l];
    const double r = 1e-14 * tr / l + 1e-300;
    for (int a = 0; a < l; ++a) M[a + a * l] += r;
  }
  __syncthreads();
  // Cholesky M = L L^T in place (lower)
  for (int j = 0; j < l; ++j) {
    if (t == 0) M[j + j * l] = sqrt(fmax(M[j + j * l], 1e-300));
    __syncthreads();
    const double djj = M[j + j * l];
    for (int i = j + 1 + t; i < l; i += SK_NT) M[i + j * l] /= djj;
    __syncthreads();
    for (int e = t; e < (l - j - 1) * (l - j - 1); e += SK_NT) {
      const int i = j + 1 + e % (l - j - 1), c = j + 1 + e / (l - j - 1);
      if (c <= i) M[i + c * l] -= M[i + j * l] * M[c + j * l];
    }
    __syncthreads();
  }
  // T = L^-1 (column c by forward substitution), Mi = T^T T
  for (int c = t; c < l; c += SK_NT) {
    for (int i = 0; i < l; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int q = c; q < i; ++q) v -= M[i + q * l] * T[q + c * l];
      T[i + c * l] = (i < c) ? 0.0 : v / M[i + i * l];
    }
  }
  __syncthreads();
  for (int e = t; e < l * l; e += SK_NT) {
    const int a = e % l, b = e / l;
    double s = 0.0;
    for (int i = 0; i < l; ++i) s = fma(T[i + a * l], T[i + b * l], s);
    Mi[e] = s;
  }
  __syncthreads();
  // loss = k - tr(Mi N); T <- Mi N
  double tr = 0.0;
  for (int e = t; e < l * l; e += SK_NT) {
    const int a = e % l, b = e / l;
    double s = 0.0;
    for (int i = 0; i < l; ++i) s = fma(Mi[a + i * l], N[i + b * l], s);
    T[e] = s;
    if (a == b) tr += s;
  }
  tr = wave_sum_d(tr);
  if ((t & 63) == 0) red[t >> 6] = tr;
  __syncthreads();
  if (t == 0) loss[d] = (double)k - ((red[0] + red[1]) + (red[2] + red[3]));
  // N <- P = (Mi N) Mi
  for (int e = t; e < l * l; e += SK_NT) {
    const int a = e % l, b = e / l;
    double s = 0.0;
    for (int i = 0; i < l; ++i) s = fma(T[a + i * l], Mi[i + b * l], s);
    N[e] = s;
  }
  __syncthreads();
  // Gw = 2 sig (Ws P - [i < k] Ws Mi)
  for (int e = t; e < m * l; e += SK_NT) {
    const int i = e % m, j = e / m;
    double sp = 0.0, sq = 0.0;
    for (int a = 0; a < l; ++a) {
      const double w = Ws[i + a * m];
      sp = fma(w, N[a + j * l], sp);
      sq = fma(w, Mi[a + j * l], sq);
    }
    Gw[e] = 2.0 * sig[i] * (sp - (i < k ? sq : 0.0));
  }
  __syncthreads();
  // G = V Gw
  for (int e = t; e < m * l; e += SK_NT) {
    const int r = e % m, j = e / m;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s = fma(V[r + i * m], Gw[i + j * m], s);
    G[e] = s;
  }
}

// gsum = sum_d grad_d, lsum = sum_d loss_d (fixed order)
__global__ void k_sk_reduce(int D, int64_t ml, const double* __restrict__ grad,
                            const double* __restrict__ loss, double* __restrict__ gsum,
                            double* __restrict__ lsum) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ml;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) s += grad[(int64_t)d * ml + e];
    gsum[e] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) s += loss[d];
    *lsum = s;
  }
}

// Adam step (bias-corrected) on Omega, then Omega <- Omega chol(Omega^T Omega)^-T
// (one workgroup; scratch: l x l Gram / factor).
__global__ __launch_bounds__(SK_NT) void k_sk_adam_orth(int m, int l, int step, double lr,
                                                        double b1, double b2, double eps,
                                                        const double* __restrict__ g,
                                                        double* __restrict__ mom,
                                                        double* __restrict__ vel,
                                                        double* __restrict__ Om,
                                                        double* __restrict__ Gm,
                                                        double* __restrict__ tmp) {
  const int t = threadIdx.x;
  const double c1 = 1.0 - pow(b1, (double)step), c2 = 1.0 - pow(b2, (double)step);
  if (lr > 0.0) {
    for (int e = t; e < m * l; e += SK_NT) {
      const double ge = g[e];
      const double mv = b1 * mom[e] + (1.0 - b1) * ge;
      const double vv = b2 * vel[e] + (1.0 - b2) * ge * ge;
      mom[e] = mv;
      vel[e] = vv;
      Om[e] -= lr * (mv / c1) / (sqrt(vv / c2) + eps);
    }
  }
  __syncthreads();
  for (int e = t; e < l * l; e += SK_NT) {
    const int a = e % l, b = e / l;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s = fma(Om[i + a * m], Om[i + b * m], s);
    Gm[e] = s;
  }
  __syncthreads();
  for (int j = 0; j < l; ++j) {
    if (t == 0) Gm[j + j * l] = sqrt(fmax(Gm[j + j * l], 1e-300));
    __syncthreads();
    const double djj = Gm[j + j * l];
    for (int i = j + 1 + t; i < l; i += SK_NT) Gm[i + j * l] /= djj;
    __syncthreads();
    for (int e = t; e < (l - j - 1) * (l - j - 1); e += SK_NT) {
      const int i = j + 1 + e % (l - j - 1), c = j + 1 + e / (l - j - 1);
      if (c <= i) Gm[i + c * l] -= Gm[i + j * l] * Gm[c + j * l];
    }
    __syncthreads();
  }
  // rows of Omega: x L^T = omega_row, solved per row (forward substitution)
  for (int r = t; r < m; r += SK_NT) {
    for (int j = 0; j < l; ++j) {
      double v = Om[r + j * m];
      for (int q = 0; q < j; ++q) v -= tmp[r + q * m] * Gm[j + q * l];
      tmp[r + j * m] = v / Gm[j + j * l];
    }
  }
  __syncthreads();
  for (int e = t; e < m * l; e += SK_NT) Om[e] = tmp[e];
}

// sin(theta) = sqrt(lam_max(R^T R)) for the residual R = U_k - Q Q^T U_k;
// lam_max is w[0] of the descending eigenvalues.
__global__ void k_sk_sintheta(const double* __restrict__ w, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = sqrt(fmax(0.0, w[0]));
}

}  // namespace

// Cross-covariances S (p x m x D, FP64, ld p) of the D draws in masks_dev
// (n x D int8), centred per draw.
static void sketch_draws(Ctx* c, const DMat& X, const DMat& Y, const double* g, const double* gy,
                         const int8_t* masks_dev, int D, int precision, double* S) {
  const int64_t n = X.rows, p = X.cols, m = Y.cols;
  const int64_t ncols = (int64_t)D * m + D;
  DBuf<float> Z(c, (size_t)n * ncols);
  launch(c, "sketch_zbuild", k_sk_zbuild, dim3(skg(n * ncols)), dim3(256), 0, masks_dev, n, D,
         (int)m, Y.dtype == FPLS_F32 ? Y.f() : (const float*)nullptr,
         Y.dtype == FPLS_F64 ? Y.d() : (const double*)nullptr, Y.ld, gy, Z.get());
  DBuf<double> zs(c, ncols), C0(c, (size_t)p * ncols);
  DMat Zm;
  Zm.p = Z.get();
  Zm.dtype = FPLS_F32;
  Zm.rows = n;
  Zm.cols = ncols;
  Zm.ld = n;
  col_sums(c, Zm, zs.get());
  gemm_tn_f32(c, precision, p, ncols, n, X.f(), X.ld, g, Z.get(), n, nullptr, nullptr, C0.get(),
              p);
  launch(c, "sketch_sfinish", k_sk_sfinish, dim3(skg(p * D * m)), dim3(256), 0,
         (const double*)C0.get(), (const double*)zs.get(), p, D, (int)m, S);
}

// column means of X (FP64 and the shift) and of Y
static void sketch_means(Ctx* c, const DMat& X, const DMat& Y, DBuf<double>& g,
                         DBuf<double>& gy) {
  g = DBuf<double>(c, X.cols);
  gy = DBuf<double>(c, Y.cols);
  col_sums(c, X, g.get());
  col_sums(c, Y, gy.get());
  std::vector<double> hg(X.cols), hy(Y.cols);
  to_host(c, hg.data(), g.get(), X.cols);
  to_host(c, hy.data(), gy.get(), Y.cols);
  for (double& v : hg) v /= (double)X.rows;
  for (double& v : hy) v /= (double)Y.rows;
  to_device(c, g.get(), hg.data(), X.cols);
  to_device(c, gy.get(), hy.data(), Y.cols);
}

static int sketch_chunk(int64_t p, int64_t m, int64_t n, int D) {
  // keep C0 (p x chunk (m + 1) FP64) and Z (n x chunk (m + 1) FP32) near 2 GB
  const int64_t per = (p * 8 + n * 4) * (m + 1) * 2;
  return (int)std::max<int64_t>(1, std::min<int64_t>(D, (int64_t(2) << 30) / std::max<int64_t>(per, 1)));
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_sketch_train(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                                 const int8_t* masks, int32_t n_draws, int32_t k, int32_t l,
                                 int32_t iters, double lr, const fpls_opts* opts, double* omega,
                                 double* loss_trace) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "sketch_train: X");
    DMat Y = stage(c, Ym, hy, "sketch_train: Y");
    FPLS_REQUIRE(X.dtype == FPLS_F32, FPLS_ERR_ARG, "sketch_train: X must be FP32");
    FPLS_REQUIRE(X.rows == Y.rows, FPLS_ERR_ARG, "sketch_train: X and Y need the same rows");
    const int64_t n = X.rows, p = X.cols, m = Y.cols;
    const int D = n_draws;
    FPLS_REQUIRE(D >= 1 && masks && omega, FPLS_ERR_ARG, "sketch_train: masks / omega missing");
    FPLS_REQUIRE(m >= 2 && m <= kJacobiMax, FPLS_ERR_ARG, "sketch_train: Y needs 2..128 columns");
    FPLS_REQUIRE(l >= 1 && l <= std::min<int64_t>(m, SK_MAXL) && k >= 1 && k <= l, FPLS_ERR_ARG,
                 "sketch_train: need 1 <= k <= l <= min(m, 128)");
    FPLS_REQUIRE(iters >= 0, FPLS_ERR_ARG, "sketch_train: iters must be >= 0");
    DBuf<double> g, gy;
    sketch_means(c, X, Y, g, gy);
    DBuf<int8_t> mk(c, (size_t)n * D);
    h2d(c, mk.get(), masks, (size_t)n * D);
    // targets: eigenpairs of every draw's Gram S_d^T S_d
    DBuf<double> Vall(c, (size_t)m * m * D), lam(c, (size_t)m * D);
    const int chunk = sketch_chunk(p, m, n, D);
    for (int d0 = 0; d0 < D; d0 += chunk) {
      const int cnt = std::min(chunk, D - d0);
      DBuf<double> S(c, (size_t)p * m * cnt), Gr(c, (size_t)m * m * cnt);
      sketch_draws(c, X, Y, g.get(), gy.get(), mk.get() + (size_t)n * d0, cnt, o.precision,
                   S.get());
      for (int d = 0; d < cnt; ++d)
        gemm_f64(c, true, false, m, m, p, 1.0, S.get() + (size_t)d * p * m, p, nullptr,
                 S.get() + (size_t)d * p * m, p, nullptr, 0.0, Gr.get() + (size_t)d * m * m, m);
      sym_eig_batched(c, Gr.get(), m, (int64_t)m * m, nullptr, (int)m, cnt,
                      lam.get() + (size_t)m * d0, m, Vall.get() + (size_t)m * m * d0,
                      (int64_t)m * m);
    }
    // Omega: the Gaussian stream of opts->seed (the reference's sketch),
    // orthonormalised
    std::vector<double> om((size_t)m * l);
    gaussian_host(om.data(), (int64_t)om.size(), o.seed);
    DBuf<double> Om(c, om.size()), mom(c, om.size()), vel(c, om.size()), gsum(c, om.size());
    DBuf<double> Gm(c, (size_t)l * l), tmp(c, om.size()), grads(c, om.size() * D), loss(c, D);
    DBuf<double> scratch(c, (size_t)D * (2 * m * l + 4 * l * l)), ltr(c, (size_t)iters + 1);
    to_device(c, Om.get(), om.data(), om.size());
    fill_f64(c, mom.get(), 0.0, (int64_t)om.size());
    fill_f64(c, vel.get(), 0.0, (int64_t)om.size());
    launch(c, "sketch_adam", k_sk_adam_orth, dim3(1), dim3(SK_NT), 0, (int)m, (int)l, 1, 0.0,
           0.9, 0.999, 1e-12, (const double*)gsum.get(), mom.get(), vel.get(), Om.get(),
           Gm.get(), tmp.get());
    for (int it = 0; it <= iters; ++it) {
      launch(c, "sketch_fwdbwd", k_sk_fwdbwd, dim3((unsigned)D), dim3(SK_NT), 0, (int)m, (int)l,
             (int)k, (const double*)Vall.get(), (const double*)lam.get(),
             (const double*)Om.get(), scratch.get(), loss.get(), grads.get());
      launch(c, "sketch_reduce", k_sk_reduce, dim3(skg((int64_t)m * l)), dim3(256), 0, D,
             (int64_t)m * l, (const double*)grads.get(), (const double*)loss.get(), gsum.get(),
             ltr.get() + it);
      if (it < iters)
        launch(c, "sketch_adam", k_sk_adam_orth, dim3(1), dim3(SK_NT), 0, (int)m, (int)l, it + 1,
               lr, 0.9, 0.999, 1e-12, (const double*)gsum.get(), mom.get(), vel.get(),
               Om.get(), Gm.get(), tmp.get());
    }
    to_host(c, omega, Om.get(), om.size());
    if (loss_trace) {
      std::vector<double> h((size_t)iters + 1);
      to_host(c, h.data(), ltr.get(), h.size());
      for (double& v : h) v /= (double)D;
      std::copy(h.begin(), h.end(), loss_trace);
    }
  });
}

extern "C" int fpls_sketch_eval(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                                const int8_t* masks, int32_t n_draws, int32_t k,
                                const double* omega, int32_t l, int32_t power,
                                const fpls_opts* opts, double* sin_theta) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "sketch_eval: X");
    DMat Y = stage(c, Ym, hy, "sketch_eval: Y");
    FPLS_REQUIRE(X.dtype == FPLS_F32, FPLS_ERR_ARG, "sketch_eval: X must be FP32");
    FPLS_REQUIRE(X.rows == Y.rows, FPLS_ERR_ARG, "sketch_eval: X and Y need the same rows");
    const int64_t n = X.rows, p = X.cols, m = Y.cols;
    const int D = n_draws;
    FPLS_REQUIRE(D >= 1 && masks && sin_theta, FPLS_ERR_ARG, "sketch_eval: masks missing");
    FPLS_REQUIRE(m >= 2 && m <= kJacobiMax && l >= k && l <= m && k >= 1, FPLS_ERR_ARG,
                 "sketch_eval: need 1 <= k <= l <= m <= 128");
    DBuf<double> g, gy;
    sketch_means(c, X, Y, g, gy);
    DBuf<int8_t> mk(c, (size_t)n * D);
    h2d(c, mk.get(), masks, (size_t)n * D);
    std::vector<double> om((size_t)m * l);
    if (omega) std::copy(omega, omega + om.size(), om.begin());
    else gaussian_host(om.data(), (int64_t)om.size(), o.seed);
    DBuf<double> Om(c, om.size());
    to_device(c, Om.get(), om.data(), om.size());
    DBuf<double> Gr(c, (size_t)m * m), w(c, m), E(c, (size_t)m * m), Uk(c, (size_t)p * k);
    DBuf<double> Yq(c, (size_t)p * l), Zq(c, (size_t)m * l), QU(c, (size_t)l * k);
    DBuf<double> EE(c, (size_t)k * k), we(c, k), Ve(c, (size_t)k * k), st(c, 1);
    const int chunk = sketch_chunk(p, m, n, D);
    for (int d0 = 0; d0 < D; d0 += chunk) {
      const int cnt = std::min(chunk, D - d0);
      DBuf<double> S(c, (size_t)p * m * cnt);
      sketch_draws(c, X, Y, g.get(), gy.get(), mk.get() + (size_t)n * d0, cnt, o.precision,
                   S.get());
      for (int d = 0; d < cnt; ++d) {
        const double* Sd = S.get() + (size_t)d * p * m;
        // exact top-k left singular subspace: U_k = S V_k / sig_k
        gemm_f64(c, true, false, m, m, p, 1.0, Sd, p, nullptr, Sd, p, nullptr, 0.0, Gr.get(), m);
        sym_eig(c, Gr.get(), (int)m, (int)m, w.get(), E.get());
        DBuf<double> sk(c, k);
        std::vector<double> hw(k);
        to_host(c, hw.data(), w.get(), k);
        for (double& v : hw) v = std::sqrt(std::max(v, 0.0));
        to_device(c, sk.get(), hw.data(), k);
        gemm_f64(c, false, false, p, k, m, 1.0, Sd, p, nullptr, E.get(), m, nullptr, 0.0,
                 Uk.get(), p);
        scale_cols(c, Uk.get(), p, p, k, sk.get(), true);
        // range finder with the reference's power schedule (src/svd_cpu_rsvd.cpp:110-121)
        gemm_f64(c, false, false, p, l, m, 1.0, Sd, p, nullptr, Om.get(), m, nullptr, 0.0,
                 Yq.get(), p);
        const int q = std::max(power, 0);
        if (q == 1) {
          gemm_f64(c, true, false, m, l, p, 1.0, Sd, p, nullptr, Yq.get(), p, nullptr, 0.0,
                   Zq.get(), m);
          gemm_f64(c, false, false, p, l, m, 1.0, Sd, p, nullptr, Zq.get(), m, nullptr, 0.0,
                   Yq.get(), p);
        } else {
          for (int it = 0; it < q; ++it) {
            gemm_f64(c, true, false, m, l, p, 1.0, Sd, p, nullptr, Yq.get(), p, nullptr, 0.0,
                     Zq.get(), m);
            orth_cholqr(c, Zq.get(), m, l);
            gemm_f64(c, false, false, p, l, m, 1.0, Sd, p, nullptr, Zq.get(), m, nullptr, 0.0,
                     Yq.get(), p);
          }
        }
        orth_cholqr(c, Yq.get(), p, l);
        // sin theta = || (I - Q Q^T) U_k ||_2: the residual form stays accurate
        // for small angles, where sqrt(1 - sigma_min(Q^T U_k)^2) cancels
        gemm_f64(c, true, false, l, k, p, 1.0, Yq.get(), p, nullptr, Uk.get(), p, nullptr, 0.0,
                 QU.get(), l);
        gemm_f64(c, false, false, p, k, l, -1.0, Yq.get(), p, nullptr, QU.get(), l, nullptr, 1.0,
                 Uk.get(), p);
        gemm_f64(c, true, false, k, k, p, 1.0, Uk.get(), p, nullptr, Uk.get(), p, nullptr, 0.0,
                 EE.get(), k);
        sym_eig_large(c, EE.get(), k, k, we.get(), Ve.get());
        launch(c, "sketch_sintheta", k_sk_sintheta, dim3(1), dim3(64), 0,
               (const double*)we.get(), st.get());
        to_host(c, sin_theta + d0 + d, st.get(), 1);
      }
    }
  });
}
