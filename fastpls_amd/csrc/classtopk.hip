// Argmax-head class ranking with per-class offsets, src/fastPLS.cpp:3838-4113
// (class_bias_offsets, fill_topk_from_yblock, class_bias_topk_from_cube,
// pls_class_predict_topk_cpp) and R/main.R:377-492 (.class_topk_from_score_cube,
// .class_bias_predict):
//
//   value(i, j) = score(i, j) + offset(j),  offset = mY + class_bias[, a]
//   (the model path) or class_bias[, a] alone (a given score cube);
//   per row the top_k classes in the order of a stable descending sort:
//   larger value first, equal values by smaller class index, values that are
//   not greater than -inf (NaN, -inf) never ranked; empty slots hold class 0
//   with score -inf.  top_k is clamped to [1, m].
//
// The reference walks each row inserting with a strict '>' into a sorted
// list.  Here one thread owns one (row, class segment): the segment's values
// pass through a register list by compare-exchange on (value desc, index asc),
// which keeps the same order whatever the insertion sequence, so segments of
// one row merge exactly.  Requests above 32 ranks use a rank-by-count kernel
// (rank(j) = #{j' ranked ahead of j}), exact for any top_k.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "model.h"

namespace fpls {

namespace {

constexpr int CTK_T = 256;          // threads per workgroup
constexpr int CTK_MAX_REG = 32;     // largest register list
constexpr int CTK_U = 8;            // score loads in flight per thread

__device__ __forceinline__ bool ctk_ahead(double v, int iv, double w, int iw) {
  return v > w || (v == w && iv < iw);
}

// Insert (v, iv) into the sorted list by a compare-exchange pass; the
// displaced tail element falls off the end.
template <int KM>
__device__ __forceinline__ void ctk_push(double (&bv)[KM], int (&bi)[KM], double v, int iv) {
  if (!ctk_ahead(v, iv, bv[KM - 1], bi[KM - 1])) return;
#pragma unroll
  for (int q = 0; q < KM; ++q) {
    const bool sw = ctk_ahead(v, iv, bv[q], bi[q]);
    const double tv = bv[q];
    const int ti = bi[q];
    bv[q] = sw ? v : tv;
    bi[q] = sw ? iv : ti;
    v = sw ? tv : v;
    iv = sw ? ti : iv;
  }
}

// off[j] = base[j] + bias[j]  (base, bias may be NULL: 0)
__global__ void k_ctk_offsets(const double* __restrict__ base, const double* __restrict__ bias,
                              int m, double* __restrict__ off) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const double b = base ? base[j] : 0.0;
  off[j] = b + (bias ? bias[j] : 0.0);
}

// Register lists.  Thread t of a workgroup owns row r = t % RB of the block
// and class segment p = t / RB; RB = CTK_T / P.  S: rows x m (ld ldS), FP64
// or FP32.  Output (row i, rank q): oi[i + ldo q], os[i + ldo q].
template <int KM, typename TS>
__global__ __launch_bounds__(CTK_T) void k_ctk_reg(const TS* __restrict__ S, int64_t ldS,
                                                  int64_t rows, int m, int P,
                                                  const double* __restrict__ off, int k,
                                                  int64_t ldo, int32_t* __restrict__ oi,
                                                  double* __restrict__ os) {
  extern __shared__ unsigned char ctk_smem[];
  const int RB = CTK_T / P;
  const int t = threadIdx.x;
  const int r = t % RB, p = t / RB;
  const int64_t i = (int64_t)blockIdx.x * RB + r;
  const bool live = i < rows;
  const int seg = (m + P - 1) / P;
  const int j0 = p * seg, j1 = min(m, j0 + seg);

  double bv[KM];
  int bi[KM];
#pragma unroll
  for (int q = 0; q < KM; ++q) { bv[q] = -INFINITY; bi[q] = 0; }

  if (live) {
    // two register sets of CTK_U scores: the loads of the next set are in
    // flight while the current one is ranked
    const TS* s = S + (int64_t)j0 * ldS + i;
    const int n = j1 > j0 ? j1 - j0 : 0;
    const int nfull = n / CTK_U * CTK_U;
    TS va[CTK_U], vb[CTK_U];
    if (nfull > 0) {
#pragma unroll
      for (int u = 0; u < CTK_U; ++u) va[u] = s[(int64_t)u * ldS];
    }
    for (int c = 0; c < nfull; c += CTK_U) {
      const bool more = c + CTK_U < nfull;
      if (more) {
#pragma unroll
        for (int u = 0; u < CTK_U; ++u) vb[u] = s[(int64_t)(c + CTK_U + u) * ldS];
      }
#pragma unroll
      for (int u = 0; u < CTK_U; ++u)
        ctk_push<KM>(bv, bi, (double)va[u] + off[j0 + c + u], j0 + c + u + 1);
#pragma unroll
      for (int u = 0; u < CTK_U; ++u) va[u] = vb[u];
    }
    for (int c = nfull; c < n; ++c)
      ctk_push<KM>(bv, bi, (double)s[(int64_t)c * ldS] + off[j0 + c], j0 + c + 1);
  }

  if (P > 1) {
    // segments 1..P-1 hand their lists to segment 0 through LDS
    double* lv = reinterpret_cast<double*>(ctk_smem);
    int* li = reinterpret_cast<int*>(lv + (size_t)(P - 1) * RB * KM);
    if (p > 0) {
#pragma unroll
      for (int q = 0; q < KM; ++q) {
        const size_t e = ((size_t)(p - 1) * RB + r) * KM + q;
        lv[e] = bv[q];
        li[e] = bi[q];
      }
    }
    __syncthreads();
    if (p == 0) {
      for (int pp = 1; pp < P; ++pp)
        for (int q = 0; q < KM; ++q) {
          const size_t e = ((size_t)(pp - 1) * RB + r) * KM + q;
          const double v = lv[e];
          if (!(v > -INFINITY)) break;  // the rest of that list is empty
          ctk_push<KM>(bv, bi, v, li[e]);
        }
    }
  }
  if (live && p == 0) {
#pragma unroll
    for (int q = 0; q < KM; ++q)
      if (q < k) {
        oi[i + ldo * q] = bi[q];
        if (os) os[i + ldo * q] = bv[q];
      }
  }
}

// Rank by count, for top_k above the register lists.  grid (ceil(m/256),
// rows); the row's values are staged in LDS chunks of CTK_RC.  Slots no class
// reaches keep the (0, -inf) the caller filled in.
constexpr int CTK_RC = 2048;
template <typename TS>
__global__ __launch_bounds__(CTK_T) void k_ctk_rank(const TS* __restrict__ S, int64_t ldS, int m,
                                                   const double* __restrict__ off, int k,
                                                   int64_t ldo, int32_t* __restrict__ oi,
                                                   double* __restrict__ os) {
  __shared__ double cv[CTK_RC];
  const int64_t i = blockIdx.y;
  const int j = blockIdx.x * CTK_T + threadIdx.x;
  const TS* s = S + i;
  const double v = j < m ? (double)s[(int64_t)j * ldS] + off[j] : -INFINITY;
  const bool ranked = j < m && v > -INFINITY;
  int rank = 0;
  for (int c0 = 0; c0 < m; c0 += CTK_RC) {
    const int cn = min(CTK_RC, m - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < cn; e += CTK_T)
      cv[e] = (double)s[(int64_t)(c0 + e) * ldS] + off[c0 + e];
    __syncthreads();
    if (ranked)
      for (int e = 0; e < cn; ++e) rank += ctk_ahead(cv[e], c0 + e, v, j) ? 1 : 0;
  }
  if (ranked && rank < k) {
    oi[i + ldo * rank] = j + 1;
    if (os) os[i + ldo * rank] = v;
  }
}

int ctk_kmax(int k) {
  return k <= 1 ? 1 : k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= CTK_MAX_REG ? CTK_MAX_REG : 0;
}

template <int KM, typename TS>
void ctk_reg_launch(Ctx* c, const TS* S, int64_t ldS, int64_t rows, int m, int P,
                    const double* off, int k, int64_t ldo, int32_t* oi, double* os) {
  const int RB = CTK_T / P;
  const size_t sm = P > 1 ? (size_t)(P - 1) * RB * KM * (sizeof(double) + sizeof(int)) : 0;
  const unsigned blocks = (unsigned)((rows + RB - 1) / RB);
  launch(c, "class_topk", k_ctk_reg<KM, TS>, dim3(blocks), dim3(CTK_T), sm, S, ldS, rows, m, P,
         off, k, ldo, oi, os);
}

// Segments per row: enough workgroups to cover the chip, each segment at
// least 16 classes, at least 64 rows per workgroup (a wave reads one
// segment, so its offsets are uniform loads), and the merge staging under
// 64 KiB of LDS.
int ctk_segments(Ctx* c, int64_t rows, int m, int KM) {
  int P = 1;
  const int64_t want = 4LL * std::max(1, c->num_cus);
  while (P < CTK_T / 64) {
    const int P2 = P * 2;
    const int RB = CTK_T / P2;
    const size_t sm = (size_t)(P2 - 1) * RB * KM * (sizeof(double) + sizeof(int));
    if (m / P2 < 16 || sm > 64 * 1024) break;
    if ((rows + (CTK_T / P) - 1) / (CTK_T / P) >= want) break;
    P = P2;
  }
  return P;
}

template <typename TS>
void ctk_dispatch(Ctx* c, const TS* S, int64_t ldS, int64_t rows, int m, const double* off,
                  int k, int64_t ldo, int32_t* oi, double* os) {
  const int KM = ctk_kmax(k);
  if (KM == 0) {
    for (int q = 0; q < k; ++q) {
      fill_i32(c, oi + ldo * q, 0, rows);
      if (os) fill_f64(c, os + ldo * q, -INFINITY, rows);
    }
    for (int64_t r0 = 0; r0 < rows; r0 += 65535) {
      const int64_t nr = std::min<int64_t>(65535, rows - r0);
      launch(c, "class_topk_rank", k_ctk_rank<TS>,
             dim3((unsigned)((m + CTK_T - 1) / CTK_T), (unsigned)nr), dim3(CTK_T), 0, S + r0,
             ldS, m, off, k, ldo, oi + r0, os ? os + r0 : nullptr);
    }
    return;
  }
  const int P = ctk_segments(c, rows, m, KM);
  switch (KM) {
    case 1: ctk_reg_launch<1>(c, S, ldS, rows, m, P, off, k, ldo, oi, os); break;
    case 4: ctk_reg_launch<4>(c, S, ldS, rows, m, P, off, k, ldo, oi, os); break;
    case 8: ctk_reg_launch<8>(c, S, ldS, rows, m, P, off, k, ldo, oi, os); break;
    case 16: ctk_reg_launch<16>(c, S, ldS, rows, m, P, off, k, ldo, oi, os); break;
    default: ctk_reg_launch<CTK_MAX_REG>(c, S, ldS, rows, m, P, off, k, ldo, oi, os); break;
  }
}

// The m x n_slices offset matrix: base (m, device, may be NULL) plus the
// caller's class_bias (m x bias_cols host FP64, column-major; bias_cols 0, 1
// or n_slices), as class_bias_offsets (src/fastPLS.cpp:3838-3861).
DBuf<double> ctk_offsets(Ctx* c, const double* base, int m, int L, const double* class_bias,
                         int bias_cols) {
  FPLS_REQUIRE(bias_cols == 0 || class_bias, FPLS_ERR_ARG,
               "class_bias is NULL but bias_cols is not 0");
  FPLS_REQUIRE(bias_cols == 0 || bias_cols == 1 || bias_cols == L, FPLS_ERR_ARG,
               "class_bias must have one column or one column per ncomp");
  DBuf<double> off(c, (size_t)m * L);
  DBuf<double> bias;
  if (bias_cols > 0) {
    bias = DBuf<double>(c, (size_t)m * bias_cols);
    h2d(c, bias.get(), class_bias, sizeof(double) * (size_t)m * bias_cols);
  }
  for (int a = 0; a < L; ++a) {
    const double* b = bias_cols == 0 ? nullptr : bias.get() + (size_t)(bias_cols == 1 ? 0 : a) * m;
    launch(c, "class_topk_offsets", k_ctk_offsets, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
           base, b, m, off.get() + (size_t)a * m);
  }
  sync(c);  // bias is released on return
  return off;
}

int clamp_topk(int top_k, int64_t m) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(top_k, m));
}

}  // namespace

// Device ranking over one slice's scores: S (rows x m, ld ldS, FP64) with the
// offsets off (m); writes ranks 0..k-1 of row i at oi/os[i + ldo q].
void class_topk_dev(Ctx* c, const double* S, int64_t ldS, int64_t rows, int m, const double* off,
                    int k, int64_t ldo, int32_t* oi, double* os) {
  if (rows <= 0) return;
  ctk_dispatch<double>(c, S, ldS, rows, m, off, k, ldo, oi, os);
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_class_predict_topk(fpls_ctx* ctx, const fpls_model* model,
                                       const fpls_mat* Xm, const double* class_bias,
                                       int32_t bias_cols, int32_t top_k, int32_t* top_index,
                                       double* top_score) {
  if (!model || !top_index) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "class_predict_topk: X");
    FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG,
                 "class_predict_topk: Xtest has the wrong number of columns");
    FPLS_REQUIRE(top_k >= 1, FPLS_ERR_ARG, "top must be a positive integer");
    const int64_t w = X.rows;
    const int m = (int)M->m;
    const int L = (int)M->ncomp.size();
    const int k = clamp_topk(top_k, m);
    DBuf<double> off = ctk_offsets(c, M->mY.get(), m, L, class_bias, bias_cols);
    if (w == 0) return;
    DBuf<int32_t> oi(c, (size_t)w * k * L);
    DBuf<double> os(c, top_score ? (size_t)w * k * L : 0);
    const int kmax = M->kmax;
    // compact prediction: T = Xs R once; slice a ranks T[:, :mc] W_a (+ offsets)
    DBuf<double> T(c, (size_t)w * kmax);
    if (X.dtype == FPLS_F32)
      gemm_nn_f32(c, M->precision, w, kmax, X.cols, X.f(), X.ld, M->mX.get(),
                  M->R_predict_f32.get(), M->p, nullptr, nullptr, T.get(), w);
    else
      project_scores(c, M, X, kmax, nullptr, T.get(), w);
    // rows per block: the score block stays under 2 GiB
    const int64_t cap = std::max<int64_t>(4096, (int64_t(1) << 28) / std::max(1, m));
    const int64_t rb = std::min<int64_t>(w, cap);
    DBuf<double> Sb(c, (size_t)rb * m);
    for (int a = 0; a < L; ++a) {
      const int mc = M->ncomp[a];
      for (int64_t r0 = 0; r0 < w; r0 += rb) {
        const int64_t nr = std::min<int64_t>(rb, w - r0);
        if (M->method == FPLS_METHOD_SIMPLS)
          gemm_f64(c, false, true, nr, m, mc, 1.0, T.get() + r0, w, nullptr, M->Q.get(), m,
                   nullptr, 0.0, Sb.get(), nr);
        else
          gemm_f64(c, false, false, nr, m, mc, 1.0, T.get() + r0, w, nullptr,
                   M->W_latent.get() + (size_t)a * kmax * m, kmax, nullptr, 0.0, Sb.get(), nr);
        const size_t base = (size_t)a * w * k + r0;
        class_topk_dev(c, Sb.get(), nr, nr, m, off.get() + (size_t)a * m, k, w, oi.get() + base,
                       top_score ? os.get() + base : nullptr);
      }
    }
    to_host_i32(c, top_index, oi.get(), (size_t)w * k * L);
    if (top_score) to_host(c, top_score, os.get(), (size_t)w * k * L);
  });
}

extern "C" int fpls_class_topk_scores(fpls_ctx* ctx, const fpls_mat* scores, int32_t m,
                                      int32_t n_slices, const double* class_bias,
                                      int32_t bias_cols, int32_t top_k, int32_t* top_index,
                                      double* top_score) {
  if (!top_index) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(m >= 1 && n_slices >= 1, FPLS_ERR_ARG,
                 "class_topk_scores: m and n_slices must be positive");
    FPLS_REQUIRE(top_k >= 1, FPLS_ERR_ARG, "top must be a positive integer");
    DBuf<unsigned char> hs;
    DMat S = stage(c, scores, hs, "class_topk_scores: scores");
    FPLS_REQUIRE(S.cols == (int64_t)m * n_slices, FPLS_ERR_ARG,
                 "class_topk_scores: scores must have m * n_slices columns");
    const int64_t w = S.rows;
    const int L = n_slices;
    const int k = clamp_topk(top_k, m);
    DBuf<double> off = ctk_offsets(c, nullptr, m, L, class_bias, bias_cols);
    if (w == 0) return;
    DBuf<int32_t> oi(c, (size_t)w * k * L);
    DBuf<double> os(c, top_score ? (size_t)w * k * L : 0);
    for (int a = 0; a < L; ++a) {
      const size_t base = (size_t)a * w * k;
      const int64_t col0 = (int64_t)a * m * S.ld;
      if (S.dtype == FPLS_F32)
        ctk_dispatch<float>(c, S.f() + col0, S.ld, w, m, off.get() + (size_t)a * m, k, w,
                            oi.get() + base, top_score ? os.get() + base : nullptr);
      else
        ctk_dispatch<double>(c, S.d() + col0, S.ld, w, m, off.get() + (size_t)a * m, k, w,
                             oi.get() + base, top_score ? os.get() + base : nullptr);
    }
    to_host_i32(c, top_index, oi.get(), (size_t)w * k * L);
    if (top_score) to_host(c, top_score, os.get(), (size_t)w * k * L);
  });
}

// Nearest-code classes of a Gaussian-Y classification model, ranked on the
// device from the compact factors: with z = T W_a + mY the code score
// 2 z c - |c|^2 (nearest_code_classes, src/fastPLS.cpp:4203-4215, applied to
// the compact prediction of pls_predict_code_classes_compact_cpu, :4217-4270)
// is T (2 W_a codes^T) + (2 codes mY - |c|^2), so the per-slice weights are
// 2 W_a codes^T and the offsets 2 codes mY - |c|^2; the ranking is that of
// .class_topk_from_score_cube on the decoded scores (R/main.R:2744-2750).
extern "C" int fpls_gaussian_y_predict_classes(fpls_ctx* ctx, const fpls_model* model,
                                               const fpls_mat* Xm, const double* codes,
                                               int32_t n_classes, int32_t top_k,
                                               int32_t* top_index, double* top_score) {
  if (!model || !top_index || !codes) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "gaussian_y_predict_classes: X");
    FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG,
                 "gaussian_y_predict_classes: Xtest has the wrong number of columns");
    FPLS_REQUIRE(top_k >= 1, FPLS_ERR_ARG, "top must be a positive integer");
    FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "gaussian_y_predict_classes: no classes");
    const int64_t w = X.rows;
    const int dim = (int)M->m, C = n_classes;
    const int L = (int)M->ncomp.size();
    const int kmax = M->kmax;
    const int k = clamp_topk(top_k, C);
    // offsets 2 codes mY - |c|^2 (host, FP64)
    std::vector<double> hmy(dim), hoff(C);
    to_host(c, hmy.data(), M->mY.get(), dim);
    for (int q = 0; q < C; ++q) {
      double dot = 0.0, nn = 0.0;
      for (int j = 0; j < dim; ++j) {
        const double v = codes[q + (size_t)j * C];
        dot += v * hmy[j];
        nn += v * v;
      }
      hoff[q] = 2.0 * dot - nn;
    }
    DBuf<double> off(c, C), dcodes(c, (size_t)C * dim);
    h2d(c, off.get(), hoff.data(), sizeof(double) * C);
    h2d(c, dcodes.get(), codes, sizeof(double) * (size_t)C * dim);
    // per-slice weights 2 W_a codes^T (first mc rows of a kmax x C block)
    DBuf<double> Wc(c, (size_t)kmax * C * L);
    for (int a = 0; a < L; ++a) {
      const int mc = M->ncomp[a];
      double* Wa = Wc.get() + (size_t)a * kmax * C;
      if (M->method == FPLS_METHOD_SIMPLS)  // W_a = Q[:, :mc]^T
        gemm_f64(c, true, true, mc, C, dim, 2.0, M->Q.get(), dim, nullptr, dcodes.get(), C,
                 nullptr, 0.0, Wa, kmax);
      else
        gemm_f64(c, false, true, mc, C, dim, 2.0, M->W_latent.get() + (size_t)a * kmax * dim,
                 kmax, nullptr, dcodes.get(), C, nullptr, 0.0, Wa, kmax);
    }
    if (w == 0) return;
    DBuf<int32_t> oi(c, (size_t)w * k * L);
    DBuf<double> os(c, top_score ? (size_t)w * k * L : 0);
    DBuf<double> T(c, (size_t)w * kmax);
    if (X.dtype == FPLS_F32)
      gemm_nn_f32(c, M->precision, w, kmax, X.cols, X.f(), X.ld, M->mX.get(),
                  M->R_predict_f32.get(), M->p, nullptr, nullptr, T.get(), w);
    else
      project_scores(c, M, X, kmax, nullptr, T.get(), w);
    const int64_t cap = std::max<int64_t>(4096, (int64_t(1) << 28) / std::max(1, C));
    const int64_t rb = std::min<int64_t>(w, cap);
    DBuf<double> Sb(c, (size_t)rb * C);
    for (int a = 0; a < L; ++a) {
      const int mc = M->ncomp[a];
      for (int64_t r0 = 0; r0 < w; r0 += rb) {
        const int64_t nr = std::min<int64_t>(rb, w - r0);
        gemm_f64(c, false, false, nr, C, mc, 1.0, T.get() + r0, w, nullptr,
                 Wc.get() + (size_t)a * kmax * C, kmax, nullptr, 0.0, Sb.get(), nr);
        const size_t base = (size_t)a * w * k + r0;
        class_topk_dev(c, Sb.get(), nr, nr, C, off.get(), k, w, oi.get() + base,
                       top_score ? os.get() + base : nullptr);
      }
    }
    to_host_i32(c, top_index, oi.get(), (size_t)w * k * L);
    if (top_score) to_host(c, top_score, os.get(), (size_t)w * k * L);
  });
}
