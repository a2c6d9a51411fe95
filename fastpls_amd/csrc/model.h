// Device-resident fitted model shared by the fit, classifier and prediction
// entry points.  Field names follow the reference's model list
// (src/fastPLS.cpp:2759-2777, 4966-4980; R/main.R:1019-1041, 1236-1250, 858-871).
#pragma once

#include <functional>

#include "fpls_internal.h"

namespace fpls {

struct LdaSlice {
  int k = 0;                // ncomp of this LDA model
  DBuf<double> means;       // C x k
  DBuf<double> inv_cov;     // k x k
  DBuf<double> linear;      // C x k
  DBuf<double> constants;   // C
  DBuf<double> ridge_d;     // applied ridge (1), read back on request
  DBuf<int32_t> status_d;   // 0 Cholesky, 1 singular, 2 LU fallback (1)
  DBuf<float> linT_pad;     // linear^T (k x C) as the padded FP32 operand of the
                            // latent scores (optional; built by the fit)
};

struct CknnSlice {
  int k = 0;
  DBuf<float> Tn;           // n x k row-major, row-normalised, rows sorted by class
  DBuf<double> centroids;   // C x k row-major (row-normalised class means)
  DBuf<float> centroids_f;  // FP32 copy for the MFMA base scores
};

struct Model {
  Ctx* ctx = nullptr;
  int method = FPLS_METHOD_PLSSVD;
  int64_t n = 0, p = 0, m = 0;
  int kmax = 0;                 // columns of R / Q
  int n_components = 0;         // SIMPLS: components actually extracted
  int in_dtype = FPLS_F64;      // dtype of the training X
  int precision = FPLS_PREC_FP32;
  int scaling = 1;
  std::vector<int> ncomp;       // effective ncomp per slice
  bool b_stored = false;
  bool label_aware = false;
  int xprod_mode = 0;           // 0 S formed, 1 matrix-free rSVD of X^T Y, 2 label-aware

  DBuf<double> R, Q, V, s, mX, vX, mY;
  DBuf<double> C_latent, W_latent, B;
  DBuf<double> R2Y, Yfit;
  // projection cache (R/main.R:1019-1041): R_predict = R / vX, offset = mX R_predict
  DBuf<double> R_predict, R_offset;
  DBuf<float> R_predict_f32;    // for the MFMA projection of FP32 inputs
  DBuf<float> R_pad_f32;        // R_predict (kmax columns) as the padded operand
  DBuf<float> mX_f32;           // mX in FP32, the shift of that projection
  // training scores T = Xc R (n x kmax): FP32 for FP32 inputs, FP64 otherwise
  DBuf<float> T_f32;
  DBuf<double> T_f64;
  DBuf<double> G;               // T^T T (kmax x kmax), reused by LDA

  // LDA head
  int n_classes = 0;
  DBuf<double> counts;          // C (as double, like the reference)
  DBuf<double> priors;          // C, held only by an imported head (else counts / n)
  std::vector<LdaSlice> lda;
  // candidate-kNN head
  std::vector<CknnSlice> cknn;
  DBuf<int32_t> y_codes;        // n training labels (device)
  DBuf<int32_t> class_offsets;  // C + 1 CSR offsets over rows sorted by class
  DBuf<int32_t> class_rows;     // n row ids grouped by class (ascending inside a class)
  std::vector<int> class_offsets_host;
  int knn_k = 10, top_m = 20;
  double tau = 0.2, alpha = 0.75;

  bool has_T() const { return T_f32.get() != nullptr || T_f64.get() != nullptr; }
  DMat T() const {
    DMat d;
    d.rows = n;
    d.cols = kmax;
    d.ld = n;
    if (T_f32.get()) { d.p = T_f32.get(); d.dtype = FPLS_F32; }
    else { d.p = T_f64.get(); d.dtype = FPLS_F64; }
    return d;
  }
};

// R_predict / R_offset / FP32 copy from R, vX, mX (k columns).
void build_projection_cache(Ctx* ctx, Model* M, bool unit_scale = false);
// T = (X - 1 mX^T) R_predict[:, :k]  (FP32 MFMA for FP32 X, FP64 otherwise),
// written into an n x k buffer of the input's dtype (Tf or Td).
void project_scores(Ctx* ctx, const Model* M, const DMat& X, int k, float* Tf, double* Td,
                    int64_t ldt);

// Exact / randomized truncated SVD of a device FP64 matrix (rsvd.hip).
struct SvdOut {
  DBuf<double> U, s, Vt;        // U rows x k, s k, Vt k x cols
  // want_v: the exact path of a tall A (cols <= rows) returns V (cols x >= k,
  // ld cols; the eigenvector matrix itself) instead of Vt
  DBuf<double> V;
  bool want_v = false;
  int k = 0;
};
void svd_exact_dev(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols, int k,
                   SvdOut& out, bool left_only);
void rsvd_dev(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols, int k,
              int oversample, int power, uint32_t seed, const double* omega_host,
              bool left_only, bool use_full, SvdOut& out);
// The matrix-free range finder of S = Xs^T Yc (truncated_rsvd_crossprod_double):
// Xs = (X - 1 mX^T) diag(1 / vX) (vX null: unscaled), Yc = Yd - 1 mY^T on
// load; Yd FP64 n x m (ld n); om_host the m x l sketch; target <= l factors.
void rsvd_xprod_core(Ctx* c, const DMat& X, const double* mX, const double* vX,
                     const double* Yd, int64_t m, const double* mY, int target, int l,
                     int power, const double* om_host, int precision, bool want_v,
                     SvdOut& r);
// Orthonormal basis of the columns of Y (rows x l) by shifted CholeskyQR3, in place.
void orth_cholqr(Ctx* ctx, double* Y, int64_t rows, int l);

// OPLS filter on a device matrix (opls.hip): Xout (FP64, ld ldo) = X
// preprocessed and deflated; mX, vX, Wo, Po (p x north, `used` columns valid)
// on the device.  The response is *Y (dense) or the one-hot of lab (m classes).
void opls_filter_dev(Ctx* c, const DMat& X, const DMat* Y, const int32_t* lab, int64_t m,
                     int north, int scaling, double* Xout, int64_t ldo, DBuf<double>& mX,
                     DBuf<double>& vX, DBuf<double>& Wo, DBuf<double>& Po, int& used);
// The stored filter (device arrays) applied to X into Xout.
void opls_apply_dev(Ctx* c, const DMat& X, const double* mX, const double* vX, const double* Wo,
                    const double* Po, int north, double* Xout, int64_t ldo);

// Fits on device-resident inputs (plssvd.hip, simpls.hip).
// With lda (labels y required), the LDA head is trained on the training
// scores inside the fit, its inverse covariances solved in the same launch as
// the latent coefficients (pls(..., classifier = "lda")).
struct LdaFuse {
  double ridge = 1e-8;
  const int32_t* ncomp = nullptr;
  int32_t n_ncomp = 0;
};
void plssvd_fit_dev(Ctx* c, Model* M, const DMat& X, const DMat& Y, const int32_t* y, int64_t m,
                    const int32_t* ncomp, int32_t n_ncomp, const fpls_opts& o,
                    const LdaFuse* lda = nullptr);
void simpls_fit_dev(Ctx* c, Model* M, const DMat& X, const DMat& Y, const int32_t* y, int64_t m,
                    const int32_t* ncomp, int32_t n_ncomp, const fpls_opts& o);

// LDA training state between lda_prepare and lda_finish (lda.hip).
struct LdaPending {
  int64_t n = 0;
  int C = 0, kmax = 0;
  bool f32_scores = false;        // T was FP32: keep the padded FP32 linear^T
  DBuf<double> means, pooled;     // C x kmax, kmax x kmax
  const double* counts = nullptr; // C class counts (the model's)
};
// counts_ready: counts_out already holds the FP64 class counts of y (its
// producer checked the label range), so they are not counted again.
void lda_prepare(Ctx* ctx, const DMat& T, const double* G, int64_t ldg, const int32_t* y,
                 int32_t C, const std::vector<int>& ks, double ridge, int precision,
                 DBuf<double>& counts_out, std::vector<LdaSlice>& slices, LdaPending& pend,
                 std::vector<SpdJob>& jobs, bool counts_ready = false);
void lda_finish(Ctx* ctx, const LdaPending& pend, std::vector<LdaSlice>& slices);
// LDA (lda.hip): prefix models for every k in ks on scores T; G = T^T T optional.
void lda_train_core(Ctx* ctx, const DMat& T, const double* G, int64_t ldg, const int32_t* y,
                    int32_t C, const std::vector<int>& ks, double ridge, int precision,
                    DBuf<double>& counts_out, std::vector<LdaSlice>& slices);
// LDA labels (w x n_ncomp, device, 1-based) of the test rows X; on_scores(a,
// scores) receives slice a's discriminant scores (w x C, device) when set;
// margins (w x n_ncomp, device, optional) the rows' winning margins.
void lda_predict_core(Ctx* ctx, const Model* M, const DMat& X, int32_t* labels,
                      const std::function<void(int, const double*)>& on_scores,
                      double* margins = nullptr);
// labels[i] = 1 + argmax_c(scores[i,c] + cst[c]) (first maximum wins);
// margin[i] (optional) = winning value - runner-up value
void argmax_rows(Ctx* ctx, const double* scores, int64_t lds, int64_t n, int64_t C,
                 const double* cst, int32_t* labels, double* margin = nullptr);

// Candidate kNN building blocks (cknn.hip).
struct CkClassIndex {
  std::vector<int> off;  // C + 1 offsets into the class-sorted rows
  DBuf<int> d_off;
  DBuf<int> d_rows;      // sorted position -> original row
};
void stable_class_sort(Ctx* ctx, const int32_t* y, int64_t n, int C, CkClassIndex& ci);
void class_centroids(Ctx* ctx, const float* Tn, int k, const CkClassIndex& ci, int C,
                     DBuf<double>& cen, DBuf<float>& cen_f);
void score_space(Ctx* ctx, const Model* M, const DMat& T, int k, float* out);
void rownorm_rows(Ctx* ctx, const float* S, int64_t n, int k, const int* rows, float* Tn);
void cknn_predict_core(Ctx* ctx, int precision, const float* Ttn, int64_t w, int k,
                       const float* Tn, const std::vector<int>& off, const float* cen_f, int C,
                       const double* bias, int top, int top_m, int knn_k, double tau,
                       double alpha, int32_t* top_index, double* top_score, int64_t ld_out,
                       double* margins = nullptr);  // w x 2: rerank gap, candidate boundary

}  // namespace fpls

struct fpls_model {
  fpls::Model m;
};

// Runs `body` and converts library errors into status codes + last error.
template <typename F>
int fpls_guard(fpls_ctx* ctx, F&& body) {
  if (!ctx) return FPLS_ERR_ARG;
  try {
    FPLS_HIP(hipSetDevice(ctx->c.device));
    body();
    ctx->c.flush_checks();
    return FPLS_OK;
  } catch (const fpls::Error& e) {
    ctx->c.drop_checks();
    ctx->c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.drop_checks();
    ctx->c.err = e.what();
    return FPLS_ERR_ARG;
  }
}
