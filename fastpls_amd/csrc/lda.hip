// LDA classifier head on the latent scores:
//   training   lda_train_prefix_cpp      src/fastPLS.cpp:1817-1927
//              lda_project_train_prefix_cpp                :1930-1957
//   prediction lda_project_predict_labels_cpp              :2100-2147
//              lda_labels_from_scores (strict '>', lowest class wins ties) :1718-1734
//
// Class sums of T come from the MFMA cross-product with the one-hot operand
// generated from int32 labels; T^T T is the PLSSVD Gram already resident on
// the model when the training rows are the fit's own rows.  Inverses,
// linear terms and constants are FP64 on the device.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <memory>
#include <set>
#include <vector>

#include "model.h"
#include "fpls_wave.h"

namespace fpls {

void fit_scores_and_gram(Ctx* ctx, Model* M, const DMat& X);

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// labels[i] = 1 + argmax_c (scores[i,c] + cst[c]); first maximum wins.  With
// MARGIN, margin[i] = the winning value minus the runner-up (0 at an exact
// tie, +inf with one class): the row's entry in the tie log.
template <bool MARGIN>
__global__ void k_argmax_rows(const double* __restrict__ scores, int64_t lds, int64_t n,
                              int64_t C, const double* __restrict__ cst,
                              int32_t* __restrict__ labels, double* __restrict__ margin) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int best = 0;
  double bv = scores[i] + cst[0];
  double sv = -INFINITY;  // runner-up value
  auto take = [&](double v, int c) {
    if (v > bv) {
      if (MARGIN) sv = bv;
      bv = v;
      best = c;
    } else if (MARGIN && v > sv) {
      sv = v;
    }
  };
  int64_t c = 1;
  // eight loads in flight, then the comparisons in column order (first max)
  for (; c + 8 <= C; c += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = scores[i + (c + u) * lds] + cst[c + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) take(v[u], (int)(c + u));
  }
  for (; c < C; ++c) take(scores[i + c * lds] + cst[c], (int)c);
  labels[i] = best + 1;
  if (MARGIN) margin[i] = bv - sv;
}

__global__ void k_add_consts(double* S, int64_t lds, int64_t n, int64_t C, const double* cst) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * C;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, c = t / n;
    S[i + c * lds] += cst[c];
  }
}

void argmax_rows(Ctx* ctx, const double* scores, int64_t lds, int64_t n, int64_t C,
                 const double* cst, int32_t* labels, double* margin) {
  if (n <= 0) return;
  if (margin)
    launch(ctx, "lda_argmax", k_argmax_rows<true>, dim3((unsigned)((n + 255) / 256)), dim3(256),
           0, scores, lds, n, C, cst, labels, margin);
  else
    launch(ctx, "lda_argmax", k_argmax_rows<false>, dim3((unsigned)((n + 255) / 256)), dim3(256),
           0, scores, lds, n, C, cst, labels, (double*)nullptr);
}

// FP64 class counts; empty[0] = number of empty classes
__global__ void k_counts_checked(const unsigned long long* __restrict__ cnt, int C,
                                 double* __restrict__ out, int32_t* __restrict__ empty) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) {
    out[c] = (double)cnt[c];
    if (cnt[c] == 0ull) atomicAdd(empty, 1);
  }
}

// One pass over the class sums (k x C, ld kmax) for the training state:
//   means (C x kmax, ld C)  mu[c, a] = sums[a, c] / n_c, and the leading k_s
//                           columns again into each slice's means,
//   pooled (kmax x kmax)    (symmatu(G) - sum_c mu[c, a] (mu[c, b] n_c)) / df,
//   empty[0]                the number of empty classes (when set).
// 16 x 16 tiles of pooled, the class terms staged through LDS 64 at a time.
struct LdaMeanCopies {
  double* p[kSpdJobMax];
  int k[kSpdJobMax];
  int count;
};
constexpr int LP_CH = 64;
__global__ __launch_bounds__(256) void k_lda_pooled(const double* __restrict__ sums,
                                                    int64_t kmax, int64_t C,
                                                    const double* __restrict__ counts,
                                                    const double* __restrict__ G, int64_t ldg,
                                                    double inv_df, double* __restrict__ pooled,
                                                    double* __restrict__ means,
                                                    LdaMeanCopies copies,
                                                    int32_t* __restrict__ empty) {
  __shared__ double mua[16][LP_CH + 1], msb[16][LP_CH + 1];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int64_t a0 = 16 * (int64_t)blockIdx.x, b0 = 16 * (int64_t)blockIdx.y;
  double acc = 0.0;
  for (int64_t c0 = 0; c0 < C; c0 += LP_CH) {
    __syncthreads();
    for (int idx = t; idx < 16 * LP_CH; idx += 256) {
      const int i = idx & 15, cc = idx >> 4;
      const int64_t c = c0 + cc;
      double va = 0.0, vb = 0.0;
      if (c < C) {
        const double nc = counts[c];
        if (a0 + i < kmax) va = sums[(a0 + i) + c * kmax] / nc;
        if (b0 + i < kmax) vb = (sums[(b0 + i) + c * kmax] / nc) * nc;
      }
      mua[i][cc] = va;
      msb[i][cc] = vb;
    }
    __syncthreads();
    const int cn = C - c0 < LP_CH ? (int)(C - c0) : LP_CH;
    // zero-padded chunk: four independent accumulators over all LP_CH terms
    double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int cc = 0; cc < LP_CH; cc += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = fma(mua[tx][cc + u], msb[ty][cc + u], q[u]);
    }
    (void)cn;
    acc += (q[0] + q[1]) + (q[2] + q[3]);
  }
  const int64_t a = a0 + tx, b = b0 + ty;
  if (a < kmax && b < kmax) {
    const double g = (a <= b) ? G[a + b * ldg] : G[b + a * ldg];
    pooled[a + b * kmax] = (g - acc) * inv_df;
  }
  if (blockIdx.y == 0) {
    for (int64_t idx = t; idx < 16 * C; idx += 256) {
      const int64_t i = idx & 15, c = idx >> 4, aa = a0 + i;
      if (aa >= kmax) continue;
      const double mu = sums[aa + c * kmax] / counts[c];
      means[c + aa * C] = mu;
      for (int s = 0; s < copies.count; ++s)
        if (aa < copies.k[s]) copies.p[s][c + aa * C] = mu;
    }
  }
  if (empty && blockIdx.x == 0 && blockIdx.y == 0)
    for (int64_t c = t; c < C; c += 256)
      if (counts[c] == 0.0) atomicAdd(empty, 1);
}

// Per slice and class c: linear[c, :] = means[c, :] inv_cov (k x k, staged
// in LDS), constants[c] = -1/2 means_c . linear_c + log(max(n_c / n, DBL_MIN))
// (the dot product in the lane order of one wave, fixed).  Block: one class,
// 256 threads; thread pair (t, t + 128) splits the k-long product of output
// column t into two interleaved halves with four accumulators each.
struct LdaLinJob {
  const double* means;  // C x k
  const double* inv;    // k x k
  double* linear;       // C x k
  double* constants;    // C
  int k;
};
struct LdaLinJobs {
  LdaLinJob j[kSpdJobMax];
  int count;
};
__global__ __launch_bounds__(256) void k_lda_linear(LdaLinJobs jobs, int64_t C,
                                                    const double* __restrict__ counts, double n) {
  __shared__ double inv[kJacobiMax * (kJacobiMax + 1)];
  __shared__ double mu[kJacobiMax], half[kJacobiMax], lin[kJacobiMax];
  const LdaLinJob J = jobs.j[blockIdx.y];
  const int k = J.k, t = threadIdx.x, lane = t & 63, col = t & 127, hb = t >> 7;
  const int64_t c = blockIdx.x;
  const int ld = k + 1;  // odd: column reads of lanes spread over the banks
  for (int idx = t; idx < k * k; idx += 256) inv[(idx % k) + (idx / k) * ld] = J.inv[idx];
  if (t < k) mu[t] = J.means[c + t * C];
  __syncthreads();
  double s = 0.0;
  if (col < k) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const double* ic = inv + col * ld;
    int b = hb;
    for (; b + 6 < k; b += 8) {
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = fma(mu[b + 2 * u], ic[b + 2 * u], a[u]);
    }
    for (; b < k; b += 2) a[0] = fma(mu[b], ic[b], a[0]);
    s = (a[0] + a[1]) + (a[2] + a[3]);
  }
  if (hb == 1 && col < k) half[col] = s;
  __syncthreads();
  if (hb == 0 && col < k) {
    const double v = s + half[col];
    J.linear[c + col * C] = v;
    lin[col] = v;
  }
  __syncthreads();
  if (t < 64) {
    double qv = 0.0;
    for (int j = lane; j < k; j += 64) qv = fma(mu[j], lin[j], qv);
    qv = wave_sum_d(qv);
    if (lane == 0) {
      const double prior = fmax(counts[c] / n, DBL_MIN);
      J.constants[c] = -0.5 * qv + log(prior);
    }
  }
}

// Prefix-LDA training (src/fastPLS.cpp:1817-1927) in two halves around the
// inverse-covariance solves, so that a caller can batch those solves with
// its own (lda_train_core runs all three steps).
//
// lda_prepare: counts, class means and the pooled covariance on scores T
// (n x >= max ks; G = T^T T optional), the per-k slices with their buffers,
// and one SpdJob per slice appended to jobs.
void lda_prepare(Ctx* ctx, const DMat& T, const double* G, int64_t ldg, const int32_t* y,
                 int32_t C, const std::vector<int>& ks, double ridge, int precision,
                 DBuf<double>& counts_out, std::vector<LdaSlice>& slices, LdaPending& pend,
                 std::vector<SpdJob>& jobs, bool counts_ready) {
  const int64_t n = T.rows;
  FPLS_REQUIRE(n > 0 && T.cols > 0, FPLS_ERR_ARG, "LDA requires a non-empty score matrix");
  FPLS_REQUIRE(C >= 2, FPLS_ERR_ARG, "LDA requires at least two classes");
  const int kmax = *std::max_element(ks.begin(), ks.end());
  FPLS_REQUIRE(kmax >= 1 && kmax <= T.cols, FPLS_ERR_ARG,
               "LDA component counts must be in 1..ncol(Ttrain)");
  FPLS_REQUIRE(kmax <= kJacobiMax, FPLS_ERR_ARG, "LDA on the device supports up to 128 components");
  FPLS_REQUIRE((int)ks.size() <= kSpdJobMax, FPLS_ERR_ARG,
               "LDA supports up to 16 component counts per fit");
  // counts + label validation (src/fastPLS.cpp:1853-1867): codes outside 1..C
  // and empty classes are counted on the device and checked when the entry
  // point returns (kernels below count such rows in no class)
  int32_t* empty = ctx->defer_check(1, FPLS_ERR_EMPTY_CLASS, "LDA received an empty class");
  if (!counts_ready) {
    DBuf<int64_t> cnt(ctx, C);
    int32_t* bad =
        ctx->defer_check(1, FPLS_ERR_LABEL_RANGE, "LDA labels must be encoded as 1..n_classes");
    label_counts(ctx, y, n, C, cnt.get(), bad, true);
    counts_out = DBuf<double>(ctx, C);
    launch(ctx, "lda_counts", k_counts_checked, dim3((C + 255) / 256), dim3(256), 0,
           reinterpret_cast<const unsigned long long*>(cnt.get()), C, counts_out.get(), empty);
    empty = nullptr;  // counted already
  }
  // class sums of the leading kmax score columns: k x C
  DBuf<double> sums(ctx, (size_t)kmax * C);
  pend.means = DBuf<double>(ctx, (size_t)C * kmax);
  if (T.dtype == FPLS_F32)
    gemm_tn_f32(ctx, precision, kmax, C, n, T.f(), T.ld, nullptr, nullptr, 0, y, nullptr,
                sums.get(), kmax);
  else
    gemm_f64(ctx, true, false, kmax, C, n, 1.0, T.d(), T.ld, nullptr, nullptr, n, nullptr, 0.0,
             sums.get(), kmax, y);
  DBuf<double> Gl;
  pend.pooled = DBuf<double>(ctx, (size_t)kmax * kmax);
  if (!G) {
    Gl = DBuf<double>(ctx, (size_t)kmax * kmax);
    if (T.dtype == FPLS_F32)
      gemm_tn_f32(ctx, precision, kmax, kmax, n, T.f(), T.ld, nullptr, T.f(), T.ld, nullptr,
                  nullptr, Gl.get(), kmax);
    else
      gemm_f64(ctx, true, false, kmax, kmax, n, 1.0, T.d(), T.ld, nullptr, T.d(), T.ld, nullptr,
               0.0, Gl.get(), kmax);
    G = Gl.get();
    ldg = kmax;
  }
  // per k: the means prefix and the buffers of the ridge rule and the inverse
  slices.clear();
  LdaMeanCopies copies{};
  for (int k : ks) {
    LdaSlice sl;
    sl.k = k;
    sl.means = DBuf<double>(ctx, (size_t)C * k);
    sl.inv_cov = DBuf<double>(ctx, (size_t)k * k);
    sl.ridge_d = DBuf<double>(ctx, 1);
    copies.p[copies.count] = sl.means.get();
    copies.k[copies.count] = k;
    ++copies.count;
    slices.push_back(std::move(sl));
  }
  // means and pooled = (symmatu(T^T T) - sum_c n_c mu_c mu_c^T) / max(1, n - C)
  const double df = std::max(1.0, (double)n - (double)C);
  const unsigned tiles = (unsigned)((kmax + 15) / 16);
  launch(ctx, "lda_pooled", k_lda_pooled, dim3(tiles, tiles), dim3(256), 0,
         (const double*)sums.get(), (int64_t)kmax, (int64_t)C, (const double*)counts_out.get(),
         G, ldg, 1.0 / df, pend.pooled.get(), pend.means.get(), copies, empty);
  pend.n = n;
  pend.C = C;
  pend.kmax = kmax;
  pend.counts = counts_out.get();
  pend.f32_scores = T.dtype == FPLS_F32;
  // the ridge rule and the inverse (mode 1 of the slice solves); the solve
  // writes its status (1: singular) straight into the deferred check
  for (auto& sl : slices) {
    int32_t* sing = ctx->defer_check(1, FPLS_ERR_SOLVE, "LDA pooled covariance is singular", 1);
    jobs.push_back(SpdJob{pend.pooled.get(), (int64_t)kmax, sl.k, 1, nullptr, ridge,
                          sl.inv_cov.get(), (int64_t)sl.k, sl.ridge_d.get(), sing});
  }
}

// lda_finish: after the solves, the linear terms means inv_cov and the
// constants -1/2 mu linear + log(max(prior, DBL_MIN)) of every slice in one
// launch, and (FP32 scores) linear^T as the padded operand of the latent scores.
void lda_finish(Ctx* ctx, const LdaPending& pend, std::vector<LdaSlice>& slices) {
  const int64_t C = pend.C;
  LdaLinJobs lj{};
  int kbig = 1;
  for (auto& sl : slices) {
    sl.linear = DBuf<double>(ctx, (size_t)C * sl.k);
    sl.constants = DBuf<double>(ctx, C);
    lj.j[lj.count++] = LdaLinJob{sl.means.get(), sl.inv_cov.get(), sl.linear.get(),
                                 sl.constants.get(), sl.k};
    kbig = std::max(kbig, sl.k);
  }
  FPLS_REQUIRE(kbig <= kJacobiMax, FPLS_ERR_ARG, "LDA supports up to 128 components");
  launch(ctx, "lda_linear", k_lda_linear, dim3((unsigned)C, (unsigned)lj.count), dim3(256), 0,
         lj, C, pend.counts, (double)pend.n);
  if (pend.f32_scores)
    for (auto& sl : slices) sl.linT_pad = nn_pad_operand(ctx, sl.linear.get(), C, true, sl.k, C);
}

void lda_train_core(Ctx* ctx, const DMat& T, const double* G, int64_t ldg, const int32_t* y,
                    int32_t C, const std::vector<int>& ks, double ridge, int precision,
                    DBuf<double>& counts_out, std::vector<LdaSlice>& slices) {
  LdaPending pend;
  std::vector<SpdJob> jobs;
  lda_prepare(ctx, T, G, ldg, y, C, ks, ridge, precision, counts_out, slices, pend, jobs);
  spd_solve_jobs(ctx, jobs);
  lda_finish(ctx, pend, slices);
}

static std::vector<int> unique_sorted(const int32_t* ncomp, int32_t L, int kcap) {
  FPLS_REQUIRE(ncomp && L >= 1, FPLS_ERR_ARG, "LDA requires at least one component count");
  std::set<int> s;
  for (int i = 0; i < L; ++i) s.insert(std::min(std::max<int>(ncomp[i], 1), kcap));
  return std::vector<int>(s.begin(), s.end());
}

// Scores (w x C, FP64) for one LDA slice of kk components; T_pre holds the
// projected test rows (w x >= kk) when the latent form is used.
void lda_scores(Ctx* ctx, const Model* M, const LdaSlice& sl, const DMat& X, const DMat& Tpre,
                double* scores) {
  const int64_t w = X.rows, p = M->p, C = M->n_classes;
  const int kk = sl.k;
  const double latent_ops = (double)w * kk * ((double)p + (double)C);
  const double direct_ops = (double)w * p * (double)C;
  if (direct_ops < 0.5 * latent_ops) {
    // direct form: (X - 1 mX^T) (R_predict lin^T)  (src/fastPLS.cpp:2128-2136)
    DBuf<double> W(ctx, (size_t)p * C);
    gemm_f64(ctx, false, true, p, C, kk, 1.0, M->R_predict.get(), p, nullptr, sl.linear.get(), C,
             nullptr, 0.0, W.get(), p);
    if (X.dtype == FPLS_F32) {
      DBuf<float> Wf(ctx, (size_t)p * C);
      f64_to_f32(ctx, Wf.get(), W.get(), p * C);
      gemm_nn_f32(ctx, M->precision, w, C, p, X.f(), X.ld, M->mX.get(), Wf.get(), p, nullptr,
                  nullptr, scores, w);
    } else {
      gemm_f64(ctx, false, false, w, C, p, 1.0, X.d(), X.ld, M->mX.get(), W.get(), p, nullptr,
               0.0, scores, w);
    }
  } else {
    // latent form: T lin^T with T = Xs R_predict (src/fastPLS.cpp:2139-2146)
    if (Tpre.dtype == FPLS_F32 && sl.linT_pad.get()) {
      gemm_nn_f32_padded(ctx, M->precision, w, C, kk, Tpre.f(), Tpre.ld, nullptr, nullptr,
                         sl.linT_pad.get(), nullptr, nullptr, scores, w);
    } else if (Tpre.dtype == FPLS_F32) {
      DBuf<double> linT(ctx, (size_t)kk * C);
      transpose_f64(ctx, sl.linear.get(), C, C, kk, linT.get(), kk);
      DBuf<float> linTf(ctx, (size_t)kk * C);
      f64_to_f32(ctx, linTf.get(), linT.get(), (int64_t)kk * C);
      gemm_nn_f32(ctx, M->precision, w, C, kk, Tpre.f(), Tpre.ld, nullptr, linTf.get(), kk,
                  nullptr, nullptr, scores, w);
    } else {
      gemm_f64(ctx, false, true, w, C, kk, 1.0, Tpre.d(), Tpre.ld, nullptr, sl.linear.get(), C,
               nullptr, 0.0, scores, w);
    }
  }
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_lda_train_prefix(fpls_ctx* ctx, fpls_model* model, const fpls_mat* Xm,
                                     const fpls_ivec* yv, int32_t n_classes,
                                     const int32_t* ncomp, int32_t n_ncomp, double ridge) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    Model* M = &model->m;
    std::vector<int> ks = unique_sorted(ncomp, n_ncomp, M->kmax);
    DBuf<int32_t> hl;
    const int32_t* y = stage_labels(c, yv, M->n, hl, "lda: y");
    DBuf<unsigned char> hx;
    DMat T;
    const double* G = nullptr;
    int64_t ldg = 0;
    DBuf<float> Tf;
    DBuf<double> Td;
    if (Xm == nullptr) {
      FPLS_REQUIRE(M->has_T(), FPLS_ERR_STATE, "model holds no training scores; pass Xtrain");
      T = M->T();
      if (M->G.get()) {
        G = M->G.get();
        ldg = M->kmax;
      }
    } else {
      DMat X = stage(c, Xm, hx, "lda: Xtrain");
      FPLS_REQUIRE(X.rows == M->n, FPLS_ERR_ARG, "LDA requires one class label per training row");
      const int kmax = ks.back();
      T.rows = X.rows;
      T.cols = kmax;
      T.ld = X.rows;
      if (X.dtype == FPLS_F32) {
        Tf = DBuf<float>(c, (size_t)X.rows * kmax);
        project_scores(c, M, X, kmax, Tf.get(), nullptr, X.rows);
        T.p = Tf.get();
        T.dtype = FPLS_F32;
      } else {
        Td = DBuf<double>(c, (size_t)X.rows * kmax);
        project_scores(c, M, X, kmax, nullptr, Td.get(), X.rows);
        T.p = Td.get();
        T.dtype = FPLS_F64;
      }
    }
    M->n_classes = n_classes;
    M->priors.reset();
    lda_train_core(c, T, G, ldg, y, n_classes, ks, ridge, M->precision, M->counts, M->lda);
  });
}

extern "C" int fpls_lda_train_scores(fpls_ctx* ctx, const fpls_mat* Tm, const fpls_ivec* yv,
                                     int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                                     double ridge, double* means, double* inv_cov,
                                     double* linear, double* constants, double* priors,
                                     double* ridge_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> ht;
    DMat T = stage(c, Tm, ht, "lda_train_scores: T");
    FPLS_REQUIRE(ncomp && n_ncomp >= 1, FPLS_ERR_ARG, "LDA requires at least one component count");
    std::vector<int> ks(ncomp, ncomp + n_ncomp);
    for (int k : ks)
      FPLS_REQUIRE(k >= 1 && k <= T.cols, FPLS_ERR_ARG,
                   "LDA component counts must be in 1..ncol(Ttrain)");
    DBuf<int32_t> hl;
    const int32_t* y = stage_labels(c, yv, T.rows, hl, "lda_train_scores: y");
    DBuf<double> counts;
    std::vector<LdaSlice> sl;
    lda_train_core(c, T, nullptr, 0, y, n_classes, ks, ridge, FPLS_PREC_FP32, counts, sl);
    size_t om = 0, oi = 0, oc = 0;
    for (size_t s = 0; s < sl.size(); ++s) {
      const int k = sl[s].k;
      if (means) to_host(c, means + om, sl[s].means.get(), (size_t)n_classes * k);
      if (linear) to_host(c, linear + om, sl[s].linear.get(), (size_t)n_classes * k);
      if (inv_cov) to_host(c, inv_cov + oi, sl[s].inv_cov.get(), (size_t)k * k);
      if (constants) to_host(c, constants + oc, sl[s].constants.get(), n_classes);
      if (ridge_out) to_host(c, ridge_out + s, sl[s].ridge_d.get(), 1);
      om += (size_t)n_classes * k;
      oi += (size_t)k * k;
      oc += n_classes;
    }
    if (priors) {
      std::vector<double> h(n_classes);
      to_host(c, h.data(), counts.get(), n_classes);
      for (int i = 0; i < n_classes; ++i) priors[i] = h[i] / (double)T.rows;
    }
  });
}

extern "C" int fpls_lda_read(fpls_ctx* ctx, const fpls_model* model, int32_t slice, int32_t field,
                             double* host_out, int64_t n_elem) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model& M = model->m;
    FPLS_REQUIRE(slice >= 0 && slice < (int)M.lda.size(), FPLS_ERR_STATE, "no such LDA slice");
    const LdaSlice& s = M.lda[slice];
    const int64_t C = M.n_classes, k = s.k;
    const double* src = nullptr;
    int64_t need = 0;
    std::vector<double> pri;
    switch (field) {
      case FPLS_LDA_MEANS: src = s.means.get(); need = C * k; break;
      case FPLS_LDA_INV_COV: src = s.inv_cov.get(); need = k * k; break;
      case FPLS_LDA_LINEAR: src = s.linear.get(); need = C * k; break;
      case FPLS_LDA_CONSTANTS: src = s.constants.get(); need = C; break;
      case FPLS_LDA_PRIORS: {
        FPLS_REQUIRE(host_out && n_elem >= C, FPLS_ERR_ARG, "output buffer too small");
        if (M.priors.get()) {
          to_host(c, host_out, M.priors.get(), C);
          return;
        }
        FPLS_REQUIRE(M.counts.get() && M.n > 0, FPLS_ERR_STATE,
                     "this LDA head holds no priors");
        to_host(c, host_out, M.counts.get(), C);
        for (int64_t i = 0; i < C; ++i) host_out[i] /= (double)M.n;
        return;
      }
      case FPLS_LDA_RIDGE: src = s.ridge_d.get(); need = 1; break;
      default: fail(FPLS_ERR_ARG, "unknown LDA field");
    }
    FPLS_REQUIRE(src != nullptr, FPLS_ERR_STATE, "this LDA head holds no such field");
    FPLS_REQUIRE(host_out && n_elem >= need, FPLS_ERR_ARG, "output buffer too small");
    to_host(c, host_out, src, need);
  });
}

namespace fpls {

// Labels (w x L, device, 1-based) of every ncomp slice for the test rows X;
// `on_scores(a, sc)` receives slice a's discriminant scores plus constants
// (w x C, device) when set.
void lda_predict_core(Ctx* c, const Model* M, const DMat& X, int32_t* lab,
                      const std::function<void(int, const double*)>& on_scores,
                      double* margins) {
  FPLS_REQUIRE(!M->lda.empty(), FPLS_ERR_STATE,
               "this model does not contain fitted LDA classifier parameters");
  FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG, "lda_predict: X has the wrong number of columns");
  const int64_t w = X.rows, C = M->n_classes;
  const int L = (int)M->ncomp.size();
  const int lda_kmax = M->lda.back().k;
  // projected test scores for the largest slice; prefixes serve the rest
  DMat T;
  T.rows = w;
  T.cols = lda_kmax;
  T.ld = w;
  DBuf<float> Tf;
  DBuf<double> Td;
  if (X.dtype == FPLS_F32) {
    Tf = DBuf<float>(c, (size_t)w * lda_kmax);
    project_scores(c, M, X, lda_kmax, Tf.get(), nullptr, w);
    T.p = Tf.get();
    T.dtype = FPLS_F32;
  } else {
    Td = DBuf<double>(c, (size_t)w * lda_kmax);
    project_scores(c, M, X, lda_kmax, nullptr, Td.get(), w);
    T.p = Td.get();
    T.dtype = FPLS_F64;
  }
  DBuf<double> sc(c, (size_t)w * C);
  for (int a = 0; a < L; ++a) {
    const int kk = std::min(std::max(M->ncomp[a], 1), lda_kmax);
    const LdaSlice* sl = nullptr;
    for (const auto& s : M->lda)
      if (s.k == kk) sl = &s;
    FPLS_REQUIRE(sl != nullptr, FPLS_ERR_STATE, "no fitted LDA classifier for this ncomp");
    lda_scores(c, M, *sl, X, T, sc.get());
    argmax_rows(c, sc.get(), w, w, C, sl->constants.get(), lab + (size_t)a * w,
                margins ? margins + (size_t)a * w : nullptr);
    if (on_scores) {
      launch(c, "lda_consts", k_add_consts, dim3(g1(w * C)), dim3(256), 0, sc.get(), w, w, C,
             sl->constants.get());
      on_scores(a, sc.get());
    }
  }
}

}  // namespace fpls

extern "C" int fpls_plssvd_lda_fit_predict(fpls_ctx* ctx, const fpls_mat* Xm,
                                           const fpls_ivec* yv, int32_t n_classes,
                                           const int32_t* ncomp, int32_t n_ncomp,
                                           const fpls_opts* opts, double lda_ridge,
                                           const fpls_mat* Xtest, int32_t* labels,
                                           double* margins, fpls_model** out) {
  if (out) *out = nullptr;
  std::unique_ptr<fpls_model> mdl(new fpls_model());
  int rc = fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(out != nullptr && labels != nullptr, FPLS_ERR_ARG, "out or labels is NULL");
    FPLS_REQUIRE(Xtest != nullptr, FPLS_ERR_ARG, "plssvd_lda_fit_predict: Xtest is NULL");
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, ht;
    DBuf<int32_t> hl;
    DMat X = stage(c, Xm, hx, "plssvd_lda: X");
    FPLS_REQUIRE(yv && yv->ptr, FPLS_ERR_ARG, "plssvd_lda: class labels are required");
    FPLS_REQUIRE(n_classes >= 2, FPLS_ERR_ARG, "LDA requires at least two classes");
    const int32_t* y = stage_labels(c, yv, X.rows, hl, "plssvd_lda: y");
    DMat Xt = stage(c, Xtest, ht, "plssvd_lda: Xtest");
    FPLS_REQUIRE(Xt.cols == X.cols, FPLS_ERR_ARG,
                 "plssvd_lda: Xtest has the wrong number of columns");
    LdaFuse lf;
    lf.ridge = lda_ridge;
    lf.ncomp = ncomp;
    lf.n_ncomp = n_ncomp;
    DMat Y;
    Model* M = &mdl->m;
    plssvd_fit_dev(c, M, X, Y, y, n_classes, ncomp, n_ncomp, o, &lf);
    FPLS_REQUIRE(!M->lda.empty(), FPLS_ERR_STATE, "plssvd_lda: the LDA head was not fitted");
    const int L = (int)M->ncomp.size();
    DBuf<int32_t> lab(c, (size_t)Xt.rows * L);
    DBuf<double> mg;
    if (margins) mg = DBuf<double>(c, (size_t)Xt.rows * L);
    lda_predict_core(c, M, Xt, lab.get(), nullptr, mg.get());
    // completed by the entry point's one synchronisation (the checks' read-back)
    c->to_host_later(labels, lab.get(), (size_t)Xt.rows * L * sizeof(int32_t));
    if (margins) c->to_host_later(margins, mg.get(), (size_t)Xt.rows * L * sizeof(double));
  });
  if (rc == FPLS_OK) *out = mdl.release();
  return rc;
}

// The reference's model$lda list (R/main.R:1244-1250) as the device LDA head:
// each model's linear / constants (what lda_project_predict_labels_cpp reads,
// src/fastPLS.cpp:2100-2147) and the optional means, inv_cov, priors and
// applied ridge, plus the padded FP32 linear^T of the latent score product.
extern "C" int fpls_model_import_lda(fpls_ctx* ctx, fpls_model* model,
                                     const fpls_lda_fields* f) {
  if (!model || !f) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    Model* M = &model->m;
    const int C = f->n_classes, L = f->n_models;
    FPLS_REQUIRE(C >= 1, FPLS_ERR_ARG, "LDA import requires at least one class");
    FPLS_REQUIRE(L >= 1 && f->ncomp, FPLS_ERR_ARG, "LDA import requires at least one model");
    FPLS_REQUIRE(f->linear && f->constants, FPLS_ERR_ARG,
                 "LDA import requires linear and constants");
    FPLS_REQUIRE(M->cknn.empty() || M->n_classes == C, FPLS_ERR_ARG,
                 "the LDA and candidate-kNN heads disagree on the number of classes");
    for (int a = 0; a < L; ++a) {
      FPLS_REQUIRE(f->ncomp[a] >= 1 && f->ncomp[a] <= M->kmax, FPLS_ERR_ARG,
                   "LDA score dimension does not match the model");
      FPLS_REQUIRE(a == 0 || f->ncomp[a] > f->ncomp[a - 1], FPLS_ERR_ARG,
                   "LDA ncomp must be strictly ascending");
    }
    std::vector<LdaSlice> slices(L);
    size_t om = 0, oi = 0;
    for (int a = 0; a < L; ++a) {
      LdaSlice& sl = slices[a];
      const int k = f->ncomp[a];
      sl.k = k;
      auto up = [&](DBuf<double>& dst, const double* src, size_t count) {
        dst = DBuf<double>(c, count);
        to_device(c, dst.get(), src, count);
      };
      up(sl.linear, f->linear + om, (size_t)C * k);
      up(sl.constants, f->constants + (size_t)a * C, (size_t)C);
      if (f->means) up(sl.means, f->means + om, (size_t)C * k);
      if (f->inv_cov) up(sl.inv_cov, f->inv_cov + oi, (size_t)k * k);
      if (f->ridge) up(sl.ridge_d, f->ridge + a, 1);
      sl.linT_pad = nn_pad_operand(c, sl.linear.get(), C, true, k, C);
      om += (size_t)C * k;
      oi += (size_t)k * k;
    }
    DBuf<double> pri;
    if (f->priors) {
      pri = DBuf<double>(c, C);
      to_device(c, pri.get(), f->priors, C);
    }
    sync(c);
    M->lda = std::move(slices);
    M->priors = std::move(pri);
    M->counts.reset();
    M->n_classes = C;
  });
}

static int lda_predict_entry(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm,
                             int32_t* labels, double* scores_out, double* margins_out) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    FPLS_REQUIRE(!M->lda.empty(), FPLS_ERR_STATE,
                 "this model does not contain fitted LDA classifier parameters");
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "lda_predict: X");
    const int64_t w = X.rows, C = M->n_classes;
    const int L = (int)M->ncomp.size();
    DBuf<int32_t> lab(c, (size_t)w * L);
    std::function<void(int, const double*)> cb;
    if (scores_out)
      cb = [&](int a, const double* sc) {
        to_host(c, scores_out + (size_t)a * w * C, sc, (size_t)w * C);
      };
    DBuf<double> mg;
    if (margins_out) mg = DBuf<double>(c, (size_t)w * L);
    lda_predict_core(c, M, X, lab.get(), cb, mg.get());
    to_host_i32(c, labels, lab.get(), (size_t)w * L);
    if (margins_out) to_host(c, margins_out, mg.get(), (size_t)w * L);
  });
}

extern "C" int fpls_lda_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm,
                                int32_t* labels, double* scores_out) {
  return lda_predict_entry(ctx, model, Xm, labels, scores_out, nullptr);
}

extern "C" int fpls_lda_predict_margins(fpls_ctx* ctx, const fpls_model* model,
                                        const fpls_mat* Xm, int32_t* labels, double* scores_out,
                                        double* margins) {
  if (!margins) return FPLS_ERR_ARG;
  return lda_predict_entry(ctx, model, Xm, labels, scores_out, margins);
}
