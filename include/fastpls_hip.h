/*
 * fastpls_hip.h -- C ABI of the MI355X (gfx950) fastPLS hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status (FPLS_OK == 0, negative on error; the message is
 * available from fpls_last_error()).  No C++ types or exceptions cross it.
 *
 * Matrices are column-major (R / Armadillo layout) and described by
 * fpls_mat: a host or device pointer, an element type (FP32 / FP64), the
 * shape and the leading dimension.  Device pointers must live on the
 * context's device; host buffers are staged through the context's stream.
 * Results are written to caller-allocated host buffers (FP64 / int32), the
 * way an Rcpp wrapper hands R-allocated vectors to C.
 *
 * A fitted model (fpls_model) stays resident in HBM between fit, classifier
 * training and prediction, so the latent scores of the training rows are
 * computed once and reused (the reference recomputes them for LDA and
 * candidate kNN: R/main.R:1216-1251, 826-836).
 *
 * Reference interfaces replaced (paths relative to the fastPLS 0.3 tree):
 *   fpls_available            has_cuda()                 src/fastPLS.cpp:1607-1609
 *   fpls_plssvd_fit           pls_model1                 src/fastPLS.cpp:4780-4986
 *                             .plssvd_label_aware_stream_model (labels) R/main.R:513-655
 *   fpls_plssvd_lda_fit       pls_model1 + .attach_lda_classifier R/main.R:1216-1251
 *   fpls_plssvd_lda_fit_predict  pls(..., Xtest) fit-then-predict tail R/main.R:5640-5646
 *   fpls_simpls_fit           pls_model2_fast            src/fastPLS.cpp:2454-2778
 *   fpls_model_read           the Rcpp::List fields returned by the fits
 *   fpls_model_import         those fields back into a device model (predict()
 *                             on a model fitted by the reference; the B_stored
 *                             policy src/fastPLS.cpp:63-90)
 *   fpls_model_import_lda     the model$lda list back onto a device model
 *                             (R/main.R:1244-1250; per model src/fastPLS.cpp:1914-1921)
 *   fpls_model_import_cknn    the model$candidate_knn list likewise (R/main.R:861-872)
 *   fpls_predict              pls_predict                src/fastPLS.cpp:3211-3445
 *   fpls_project              .fastpls_latent_scores     R/main.R:965-1017
 *   fpls_lda_train_prefix     lda_project_train_prefix_cpp src/fastPLS.cpp:1930-1957
 *   fpls_lda_train_scores     lda_train_prefix_cpp       src/fastPLS.cpp:1817-1927
 *   fpls_lda_read             the per-ncomp LDA list fields (means, linear, ...)
 *   fpls_lda_predict          lda_project_predict_labels_cpp src/fastPLS.cpp:2100-2147
 *   fpls_cknn_fit             .fit_candidate_knn         R/main.R:795-873
 *   fpls_cknn_predict         .candidate_knn_predictions R/main.R:875-950
 *   fpls_cknn_predict_scores  candidate_knn_predict_cpp  src/fastPLS.cpp:1627-1648
 *   fpls_rsvd                 truncated_svd_cpu_rsvd / fastsvd  src/svd_cpu_rsvd.cpp:95-124
 *   fpls_rsvd_xprod           truncated_rsvd_crossprod_double   src/fastPLS.cpp:550-623
 *   fpls_class_sums           rowsum() class sums of the label-aware stream R/main.R:563-588
 *   fpls_cv_batched           the fold loops of pls_cv_predict_compiled src/fastPLS.cpp:6777-7142
 *                             and pls.double.cv R/main.R:6676-6764
 *   fpls_gaussian_streams     gaussian_matrix_local      src/fastPLS.cpp:99-115
 *   fpls_sketch_train / fpls_sketch_eval  the learned-sketch extension of the
 *                             Gaussian range finder (no reference counterpart)
 *   fpls_gaussian_y_encode    .prepare_gaussian_y (regression) R/main.R:1997-2052,
 *                             .gaussian_y_ridge_decoder        R/main.R:1967-1980
 *   fpls_gaussian_y_codes     .gaussian_y_class_codes          R/main.R:1982-1995
 *   fpls_gaussian_y_decode    .decode_gaussian_y_matrix (regr.) R/main.R:2077-2096
 *   fpls_gaussian_y_classes   nearest_code_classes       src/fastPLS.cpp:4203-4215
 *   fpls_class_predict_topk   pls_class_predict_topk_cpp src/fastPLS.cpp:3954-4113
 *                             (compact_prediction_weights :3737, class_bias_offsets
 *                             :3838, fill_topk_from_yblock :3863)
 *   fpls_class_topk_scores    class_bias_topk_from_cube  src/fastPLS.cpp:3908-3952,
 *                             .class_topk_from_score_cube R/main.R:453-474
 *   fpls_gaussian_y_predict_classes  pls_predict_code_classes_compact_cpu
 *                             src/fastPLS.cpp:4217-4270 with the ranking of
 *                             .class_topk_from_score_cube R/main.R:2744-2750
 *   fpls_opls_filter          opls_filter_cpp            src/fastPLS.cpp:4377-4461
 *   fpls_opls_apply           opls_apply_filter_cpp      src/fastPLS.cpp:4463-4483
 *   fpls_kernel_matrix        kernel_matrix_cpp          src/fastPLS.cpp:4314-4341
 *   fpls_center_kernel_train  center_kernel_train_cpp    src/fastPLS.cpp:4344-4357
 *   fpls_center_kernel_test   center_kernel_test_cpp     src/fastPLS.cpp:4360-4374
 */
#ifndef FASTPLS_HIP_H
#define FASTPLS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FPLS_OK                 0
#define FPLS_ERR_ARG           -1   /* bad argument / shape                     */
#define FPLS_ERR_HIP           -2   /* HIP runtime error                        */
#define FPLS_ERR_EMPTY_CLASS   -3   /* LDA: a class has no rows (src/fastPLS.cpp:1766-1768) */
#define FPLS_ERR_LABEL_RANGE   -4   /* labels outside 1..C (src/fastPLS.cpp:1757-1760) */
#define FPLS_ERR_RANK          -5   /* effective rank < 1 (src/fastPLS.cpp:4812-4814) */
#define FPLS_ERR_SOLVE         -6   /* singular system, no fallback left       */
#define FPLS_ERR_NO_DEVICE     -7   /* no usable gfx950 device                  */
#define FPLS_ERR_STATE         -8   /* model lacks the requested part          */

/* ---- element types and locations --------------------------------------- */
#define FPLS_F32     0
#define FPLS_F64     1
#define FPLS_HOST    0
#define FPLS_DEVICE  1

typedef struct fpls_mat {
  const void* ptr;   /* first element, column-major                      */
  int32_t dtype;     /* FPLS_F32 | FPLS_F64                              */
  int32_t loc;       /* FPLS_HOST | FPLS_DEVICE                          */
  int64_t rows;
  int64_t cols;
  int64_t ld;        /* leading dimension >= rows                       */
} fpls_mat;

typedef struct fpls_ivec {
  const int32_t* ptr;  /* 1-based class codes                            */
  int32_t loc;         /* FPLS_HOST | FPLS_DEVICE                         */
  int64_t n;
} fpls_ivec;

/* ---- options ------------------------------------------------------------ */
#define FPLS_SVD_EXACT   3   /* reference SVD method ids (src/svd_iface.h:39-44) */
#define FPLS_SVD_RSVD    4
#define FPLS_PREC_FP32   0   /* plain FP32 MFMA contractions                    */
#define FPLS_PREC_BF16X3 1   /* split-BF16 x3 MFMA contractions                 */

typedef struct fpls_opts {
  int32_t scaling;     /* 1 centering, 2 autoscaling, 3 none (R/main.R:5890)  */
  int32_t svd_method;  /* FPLS_SVD_EXACT | FPLS_SVD_RSVD                      */
  int32_t oversample;  /* rSVD oversampling (default 10)                     */
  int32_t power;       /* rSVD power iterations (default 1)                  */
  uint32_t seed;       /* sketch seed (default 1)                            */
  int32_t inc_iters;   /* SIMPLS refresh power iterations (default 2)        */
  int32_t precision;   /* FPLS_PREC_*                                        */
  int32_t fit;         /* compute Yfit / R2Y                                 */
  int32_t store_b;     /* -1 auto (<= 256 MB), 0 never, 1 always             */
  int32_t label_aware; /* labels + 1: .plssvd_label_aware_stream_model rules
                          (ridge 1e-10 mean(diag G), guarded sd, no B)
                          R/main.R:513-655                                    */
  int32_t xprod_route; /* fpls_rsvd_xprod: 0 matrix free (a_times/at_times,
                          src/fastPLS.cpp:550-623), 1 form S = X^T Y once
                          and sketch S (one pass over X)                      */
  int32_t opls_north;  /* fpls_cv_predict with FPLS_METHOD_OPLS: orthogonal
                          components removed per training fold (default 1)   */
  int32_t xprod;       /* fpls_plssvd_fit with a dense Y and the rSVD: 1 takes
                          the matrix-free FP64 range finder of S = Xs^T Yc
                          (pls.model1.rsvd.xprod.precision "implicit64",
                          R/main.R:6080-6108, src/fastPLS.cpp:5518-5590,
                          550-623); S is never formed.  0 forms S (pls_model1).
                          pls() sets it by .should_use_xprod_default
                          (R/main.R:1883-1892)                                */
  int32_t reserved[3];
} fpls_opts;

void fpls_opts_default(fpls_opts* opts);

/* ---- context ------------------------------------------------------------ */
typedef struct fpls_ctx fpls_ctx;
typedef struct fpls_model fpls_model;

int fpls_available(void);
int fpls_ctx_create(int device, fpls_ctx** out);
void fpls_ctx_destroy(fpls_ctx* ctx);
const char* fpls_last_error(const fpls_ctx* ctx);
int fpls_ctx_synchronize(fpls_ctx* ctx);
/* Kernel timing: when enabled, every launch of a named kernel is bracketed
   by HIP events on the context stream; fpls_kernel_time returns the count
   and the summed milliseconds since the last reset. */
int fpls_timing_enable(fpls_ctx* ctx, int enable);
/* Extension (no reference counterpart; the reference's CUDA kernels use fixed
   launch shapes): the autotuner of the skinny projection tiles (T = X R and
   the other thin products of the data matrix).  mode 1 (the default, also
   FPLS_AUTOTUNE=1 in the environment) times the candidate tile heights (64,
   128, 256 rows per workgroup) and split-K counts on a shape's first call
   and keeps the fastest per context; mode 0 uses the built-in rule.
   force_mb (1, 2, 4; 0 = none) and force_splits (>= 1; 0 = none) pin a
   configuration.  Clears the table. */
int fpls_autotune_set(fpls_ctx* ctx, int32_t mode, int32_t force_mb, int32_t force_splits);
/* The tuning table as JSON text ({"key": {"mb", "rows", "splits", "us",
   "tuned"}, ...}); returns FPLS_OK, or the buffer size it needs when cap is
   too small. */
int fpls_autotune_report(fpls_ctx* ctx, char* buf, int64_t cap);
int fpls_kernel_time(fpls_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms);

/* ---- fits ---------------------------------------------------------------- */
/* PLSSVD.  Give either a dense response Y (n x m) or 1-based labels y with
   n_classes (the one-hot response is never formed: class sums are streamed
   from the int32 labels). */
int fpls_plssvd_fit(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const fpls_ivec* y,
                    int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                    const fpls_opts* opts, fpls_model** out);

/* PLSSVD with the LDA head: pls(method = "plssvd", classifier = "lda")'s fit
   (pls_model1 src/fastPLS.cpp:4780-4986 on the one-hot of y, then
   .attach_lda_classifier on the training scores R/main.R:1216-1251 with
   lda_train_prefix_cpp src/fastPLS.cpp:1817-1927) in one call: the same
   model as fpls_plssvd_fit followed by fpls_lda_train_prefix(model, NULL, y,
   n_classes, ncomp, n_ncomp, lda_ridge), with the latent solves and the LDA
   inverse covariances run in one launch. */
int fpls_plssvd_lda_fit(fpls_ctx* ctx, const fpls_mat* X, const fpls_ivec* y, int32_t n_classes,
                        const int32_t* ncomp, int32_t n_ncomp, const fpls_opts* opts,
                        double lda_ridge, fpls_model** out);

/* pls(Xtrain, Ytrain, Xtest, method = "plssvd", classifier = "lda"): the
   fit of fpls_plssvd_lda_fit and predict(model, Xtest) of the LDA head
   (R/main.R:5640-5646, the fit-then-predict tail of pls(); prediction
   lda_project_predict_labels_cpp src/fastPLS.cpp:2100-2147) in one call with
   one host round trip at its end.  labels: Xtest rows x n_ncomp int32 (host,
   1-based), as fpls_lda_predict writes them; margins (same shape, FP64, host,
   may be NULL): the tie log of fpls_lda_predict_margins; *out receives the
   model. */
int fpls_plssvd_lda_fit_predict(fpls_ctx* ctx, const fpls_mat* X, const fpls_ivec* y,
                                int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                                const fpls_opts* opts, double lda_ridge, const fpls_mat* Xtest,
                                int32_t* labels, double* margins, fpls_model** out);

/* SIMPLS ("fast" incremental core).  Dense Y or labels (one-hot response). */
int fpls_simpls_fit(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const fpls_ivec* y,
                    int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                    const fpls_opts* opts, fpls_model** out);

/* ---- model access -------------------------------------------------------- */
#define FPLS_METHOD_PLSSVD 1
#define FPLS_METHOD_SIMPLS 3
/* fpls_cv_predict only (the reference's method ids 4 and 5,
   src/fastPLS.cpp:6664-6670): OPLS filter per training fold, then SIMPLS with
   scaling "none"; kernel PLS, whose compiled CV fits the direct SIMPLS core. */
#define FPLS_METHOD_OPLS 4
#define FPLS_METHOD_KERNELPLS 5

typedef struct fpls_model_info {
  int32_t method;      /* FPLS_METHOD_*                                      */
  int64_t n, p, m;
  int32_t kmax;        /* columns of R / Q                                   */
  int32_t n_ncomp;     /* slices                                             */
  int32_t n_components;/* components actually extracted (SIMPLS early stop)  */
  int32_t b_stored;
  int32_t has_lda;
  int32_t has_cknn;
  int32_t n_classes;
  int32_t ncomp[64];   /* effective ncomp per slice (first 64)             */
  int32_t xprod_mode;  /* how S was treated: 0 formed (dense or class sums),
                          1 matrix-free ("implicit"), 2 label-aware stream  */
} fpls_model_info;

/* Field ids for fpls_model_read; sizes in FP64 elements. */
#define FPLS_FIELD_R        1  /* p x kmax                                  */
#define FPLS_FIELD_Q        2  /* m x kmax                                  */
#define FPLS_FIELD_MX       3  /* p                                         */
#define FPLS_FIELD_VX       4  /* p                                         */
#define FPLS_FIELD_MY       5  /* m                                         */
#define FPLS_FIELD_B        6  /* p x m x n_ncomp (when b_stored)           */
#define FPLS_FIELD_C_LATENT 7  /* kmax x kmax x n_ncomp (PLSSVD)            */
#define FPLS_FIELD_W_LATENT 8  /* kmax x m x n_ncomp (PLSSVD)               */
#define FPLS_FIELD_S        9  /* kmax singular values (PLSSVD)             */
#define FPLS_FIELD_TTRAIN  10  /* n x kmax training scores                   */
#define FPLS_FIELD_R2Y     11  /* n_ncomp (fit)                              */
#define FPLS_FIELD_YFIT    12  /* n x m x n_ncomp (fit)                      */
#define FPLS_FIELD_V       13  /* p x kmax SIMPLS V basis                    */
#define FPLS_FIELD_R_PREDICT 14 /* p x kmax  R / vX                          */
#define FPLS_FIELD_R_OFFSET  15 /* kmax      mX R_predict                    */

int fpls_model_info_get(const fpls_model* model, fpls_model_info* info);

/* A fitted model in the reference's list format (host FP64, column-major),
   for fpls_model_import. */
typedef struct fpls_model_fields {
  int32_t method;          /* FPLS_METHOD_SIMPLS | FPLS_METHOD_PLSSVD          */
  int32_t kmax;            /* columns of R                                     */
  int64_t p, m;
  int32_t n_ncomp;
  int32_t reserved;
  const int32_t* ncomp;    /* n_ncomp slice orders, 1..kmax                    */
  const double* R;         /* p x kmax                                         */
  const double* Q;         /* m x kmax (SIMPLS required; PLSSVD with C_latent) */
  const double* C_latent;  /* kmax x kmax x n_ncomp (PLSSVD) or NULL          */
  const double* W_latent;  /* kmax x m x n_ncomp (PLSSVD) or NULL             */
  const double* B;         /* p x m x n_ncomp or NULL                          */
  const double* mX;        /* p                                                */
  const double* vX;        /* p                                                */
  const double* mY;        /* m                                                */
} fpls_model_fields;
/* Device model from those fields: fpls_predict, fpls_project and
   fpls_class_predict_topk then work on it (the LDA / cKNN heads are not part
   of the fields).  Errors follow compact_prediction_weights: SIMPLS needs Q,
   PLSSVD W_latent or C_latent + Q, and every ncomp within 1..kmax. */
int fpls_model_import(fpls_ctx* ctx, const fpls_model_fields* fields, fpls_model** out);
int fpls_model_read(fpls_ctx* ctx, const fpls_model* model, int32_t field, double* host_out,
                    int64_t n_elem);

/* The reference's model$lda list (R/main.R:1244-1250): ncomp, and per entry of
   models the fields of lda_train_prefix_cpp (src/fastPLS.cpp:1914-1921).  Host
   FP64, column-major, packed one model after another in the order of ncomp --
   the layout fpls_lda_train_scores writes. */
typedef struct fpls_lda_fields {
  int32_t n_classes;        /* C = length(lev)                                   */
  int32_t n_models;         /* length(lda$ncomp)                                 */
  const int32_t* ncomp;     /* lda$ncomp: strictly ascending, each in 1..kmax    */
  const double* means;      /* C x k per model, or NULL                          */
  const double* inv_cov;    /* k x k per model, or NULL                          */
  const double* linear;     /* C x k per model (required)                        */
  const double* constants;  /* C per model (required)                            */
  const double* priors;     /* C, or NULL                                        */
  const double* ridge;      /* n_models applied ridges, or NULL                  */
} fpls_lda_fields;
/* Attach that LDA head to a device model (one from fpls_model_import or a
   fit), replacing any LDA head it holds: fpls_lda_predict then scores with
   the given linear / constants exactly as lda_project_predict_labels_cpp
   (src/fastPLS.cpp:2100-2147).  fpls_lda_read returns the imported fields;
   one that was passed as NULL reads as FPLS_ERR_STATE. */
int fpls_model_import_lda(fpls_ctx* ctx, fpls_model* model, const fpls_lda_fields* fields);

/* The reference's model$candidate_knn list (R/main.R:861-872). */
typedef struct fpls_cknn_fields {
  int32_t n_classes;        /* C = length(lev)                                   */
  int32_t n_models;         /* length(candidate_knn$ncomp)                       */
  const int32_t* ncomp;     /* candidate_knn$ncomp: strictly ascending, 1..kmax  */
  const double* centroids;  /* C x k per model (models[[k]]$centroids), packed   */
  int64_t n_train;          /* nrow(Ttrain)                                      */
  int32_t ttrain_cols;      /* ncol(Ttrain) >= max(ncomp)                        */
  int32_t reserved;
  const double* Ttrain;     /* n_train x ttrain_cols training scores             */
  const int32_t* y_codes;   /* n_train training labels, 1..C                     */
  int32_t knn_k, top_m;     /* parameters$knn_k, parameters$top_m                */
  double tau, alpha;        /* parameters$tau, parameters$alpha                  */
} fpls_cknn_fields;
/* Attach that candidate-kNN head to a device model, replacing any it holds:
   the training scores are mapped into each model's score space with the
   device model's C_latent and row-normalised as at prediction time
   (R/main.R:903-918), the centroids are used as given, and
   fpls_cknn_predict then ranks as .candidate_knn_predictions.  Errors as
   .fit_candidate_knn (R/main.R:813-815, 840-843); labels outside 1..C are
   rejected with FPLS_ERR_LABEL_RANGE. */
int fpls_model_import_cknn(fpls_ctx* ctx, fpls_model* model, const fpls_cknn_fields* fields);
void fpls_model_free(fpls_model* model);

/* ---- prediction ------------------------------------------------------------ */
/* Ypred (w x m x n_ncomp) = ((Xtest - mX)/vX) B + mY, via B or the latent
   factors. */
int fpls_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest, double* ypred);
/* Latent scores T = X R_predict[:, :k] - offset[:k] (n x k). */
int fpls_project(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* X, int32_t k,
                 double* T_out);

/* ---- LDA head -------------------------------------------------------------- */
/* Train the prefix LDA on the model's projected training rows (reusing the
   resident training scores when X is the fit's own matrix: pass X = NULL). */
int fpls_lda_train_prefix(fpls_ctx* ctx, fpls_model* model, const fpls_mat* Xtrain,
                          const fpls_ivec* y, int32_t n_classes, const int32_t* ncomp,
                          int32_t n_ncomp, double ridge);
#define FPLS_LDA_MEANS     1  /* C x k                                        */
#define FPLS_LDA_INV_COV   2  /* k x k                                        */
#define FPLS_LDA_LINEAR    3  /* C x k                                        */
#define FPLS_LDA_CONSTANTS 4  /* C                                            */
#define FPLS_LDA_PRIORS    5  /* C                                            */
#define FPLS_LDA_RIDGE     6  /* 1                                            */
int fpls_lda_read(fpls_ctx* ctx, const fpls_model* model, int32_t slice, int32_t field,
                  double* host_out, int64_t n_elem);
/* Labels (w x n_ncomp, 1-based) and optionally the discriminant scores
   (w x C x n_ncomp, may be NULL). */
int fpls_lda_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                     int32_t* labels, double* scores);
/* fpls_lda_predict with the tie log (the per-row diagnostic the reference
   does not emit; its labels come from lda_labels_from_scores,
   src/fastPLS.cpp:1718-1734): margins (w x n_ncomp FP64, host, required) =
   the winning discriminant value minus the runner-up (score + constant; 0 at
   an exact tie, +inf with one class).  A label is a near-tie when its margin
   is below the error bound of the precision mode. */
int fpls_lda_predict_margins(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                             int32_t* labels, double* scores, double* margins);
/* No counterpart in the reference.  fpls_lda_predict for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc): Xtest (device), labels (w x n_ncomp
   int32, device) and scores (w x C x n_ncomp FP64, device, may be NULL) as
   fpls_lda_predict, with every temporary carved from one static 64 MiB
   scratch array per GPU.  Rows beyond what that array holds for the model
   (about 64 MiB / (4 k + 12 C) bytes per row for FP32 input) are rejected
   with FPLS_ERR_ARG before any work starts.  Returns after the device work
   has finished.  Calls that overlap on one GPU (several threads or contexts)
   share the array and race. */
int fpls_lda_predict_noalloc(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                             int32_t* labels, double* scores);
/* Array-level trainer on given scores T (n x kmax). Outputs per slice s with
   k_s = ncomp[s]: means/linear (C x k_s), inv_cov (k_s x k_s), constants (C),
   packed one slice after another; priors (C) and ridge (n_ncomp). */
int fpls_lda_train_scores(fpls_ctx* ctx, const fpls_mat* T, const fpls_ivec* y,
                          int32_t n_classes, const int32_t* ncomp, int32_t n_ncomp,
                          double ridge, double* means, double* inv_cov, double* linear,
                          double* constants, double* priors, double* ridge_out);

/* ---- candidate kNN head ------------------------------------------------------- */
int fpls_cknn_fit(fpls_ctx* ctx, fpls_model* model, const fpls_mat* Xtrain, const fpls_ivec* y,
                  int32_t n_classes, int32_t knn_k, double tau, double alpha, int32_t top_m);
/* top_index (w x top x n_ncomp, 1-based class), top_score (same, may be NULL). */
int fpls_cknn_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                      int32_t top, int32_t* top_index, double* top_score);
/* fpls_cknn_predict with the tie log (no counterpart in the reference, whose
   ranking is candidate_insert_top, src/fastPLS.cpp:1249-1267, 1389-1457):
   margins (w x 2 x n_ncomp FP64, host, required): [., 0, a] the rank-1 minus
   the rank-2 reranked score (+inf with one ranked class), [., 1, a] the
   top_m-th candidate's base score minus the best base score left out of the
   candidate set (+inf when every class is a candidate). */
int fpls_cknn_predict_margins(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                              int32_t top, int32_t* top_index, double* top_score,
                              double* margins);
/* Array-level scorer on given normalised scores (candidate_knn_predict_cpp). */
int fpls_cknn_predict_scores(fpls_ctx* ctx, const fpls_mat* Ttest_norm,
                             const fpls_mat* Ttrain_norm, const fpls_ivec* y_codes,
                             const fpls_mat* centroids, const double* candidate_bias,
                             int32_t top_k, int32_t top_m, int32_t knn_k, double tau,
                             double alpha, int32_t* top_index, double* top_score);
/* The same with the tie log of fpls_cknn_predict_margins (margins: ntest x 2). */
int fpls_cknn_predict_scores_margins(fpls_ctx* ctx, const fpls_mat* Ttest_norm,
                                     const fpls_mat* Ttrain_norm, const fpls_ivec* y_codes,
                                     const fpls_mat* centroids, const double* candidate_bias,
                                     int32_t top_k, int32_t top_m, int32_t knn_k, double tau,
                                     double alpha, int32_t* top_index, double* top_score,
                                     double* margins);

/* ---- argmax-head class ranking ------------------------------------------------ */
/* Top-k classes of the argmax head with per-class offsets: value(i, j) =
   (Xs R[:, :mc] W_a)(i, j) + mY[j] + class_bias[j, a], ranked per row in the
   order of a stable descending sort (larger value first, ties to the smaller
   class index; NaN and -inf never ranked, empty slots hold class 0 and score
   -inf).  k = max(1, min(top_k, m)); top_index (w x k x n_ncomp, 1-based
   class) and top_score (same, may be NULL) are host buffers, column-major.
   class_bias: host FP64 m x bias_cols (column-major), bias_cols 0 (NULL,
   zero offsets), 1 (broadcast over the slices) or n_ncomp. */
int fpls_class_predict_topk(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                            const double* class_bias, int32_t bias_cols, int32_t top_k,
                            int32_t* top_index, double* top_score);
/* The same ranking over a given score cube: scores is w x (m * n_slices)
   (element (i, j, a) at column j + m a), host or device, FP32 or FP64;
   offsets are class_bias alone. */
int fpls_class_topk_scores(fpls_ctx* ctx, const fpls_mat* scores, int32_t m, int32_t n_slices,
                           const double* class_bias, int32_t bias_cols, int32_t top_k,
                           int32_t* top_index, double* top_score);

/* Nearest-code classes of a Gaussian-Y classification model (a fit on the
   code matrix Z = codes[y]): the ranking above of the code scores
   2 z c - |c|^2 of the compact prediction z, straight from the factors.
   codes: host FP64 n_classes x m (column-major, m = the model's response
   dimension); outputs as fpls_class_predict_topk with k = max(1, min(top_k,
   n_classes)). */
int fpls_gaussian_y_predict_classes(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xtest,
                                    const double* codes, int32_t n_classes, int32_t top_k,
                                    int32_t* top_index, double* top_score);

/* ---- OPLS filter and kernel matrices (pre-stages of the SIMPLS fit) ---------- */
/* Orthogonal-signal filter: X is centred (scaling 1), centred and divided by
   its column sd (2, zero or non-finite sd -> 1) or left (3); then up to north
   orthogonal components are removed (early stop when a norm is zero or not
   finite).  The response is Y (n x m, dense) or the one-hot of y (1-based
   labels, n_classes columns): give exactly one.  Outputs: X_out (n x p FP64,
   host or device; the filtered X), mX, vX (p, host), W_orth, P_orth (p x
   north_used, host, column-major, may be NULL), north_used.  north = 0 is the
   plain preprocessing. */
int fpls_opls_filter(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const fpls_ivec* y,
                     int32_t n_classes, int32_t north, int32_t scaling, const fpls_mat* X_out,
                     double* mX, double* vX, double* W_orth, double* P_orth,
                     int32_t* north_used);
/* X_out = the stored filter applied to new rows: (X - mX) / vX, then
   X -= (X w_a) p_a^T for a = 1..north (mX, vX, W_orth, P_orth host). */
int fpls_opls_apply(fpls_ctx* ctx, const fpls_mat* X, const double* mX, const double* vX,
                    const double* W_orth, const double* P_orth, int32_t north,
                    const fpls_mat* X_out);
#define FPLS_KERNEL_LINEAR 1
#define FPLS_KERNEL_RBF    2
#define FPLS_KERNEL_POLY   3
/* K (n1 x n2 FP64, host or device) = X1 X2^T (linear),
   exp(-gamma (|x1|^2 + |x2|^2 - 2 x1'x2)) with (-1e-10, 0) clamped to 0 (rbf),
   or (gamma x1'x2 + coef0)^degree (poly). */
int fpls_kernel_matrix(fpls_ctx* ctx, const fpls_mat* X1, const fpls_mat* X2, int32_t kernel,
                       double gamma, int32_t degree, double coef0, const fpls_mat* K_out);
/* Kc = K - 1 colmeans^T - rowmeans 1^T + grand_mean; K_out may alias K.
   col_means (n2) and grand_mean are host outputs (may be NULL). */
int fpls_center_kernel_train(fpls_ctx* ctx, const fpls_mat* K, const fpls_mat* K_out,
                             double* col_means, double* grand_mean);
/* The test-side centring with the training col_means / grand_mean (host). */
int fpls_center_kernel_test(fpls_ctx* ctx, const fpls_mat* K, const double* col_means,
                            double grand_mean, const fpls_mat* K_out);

/* ---- truncated SVD ---------------------------------------------------------- */
/* rSVD (or exact, per opts->svd_method / the reference's dispatch rules) of A
   (rows x cols).  U (rows x k), s (k), Vt (k x cols); *k_out = columns
   produced.  omega (cols x l, FP64) replaces the Gaussian sketch when given
   (the learned-sketch extension); NULL reproduces the reference's stream. */
int fpls_rsvd(fpls_ctx* ctx, const fpls_mat* A, int32_t k, const fpls_opts* opts,
              const fpls_mat* omega, double* U, double* s, double* Vt, int32_t* k_out);
/* Matrix-free rSVD of S = X^T Y without forming S.  omega (cols(Y) x l,
   host FP64) replaces the Gaussian sketch when given (a learned sketch from
   fpls_sketch_train); NULL reproduces the reference's stream. */
int fpls_rsvd_xprod(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, int32_t k,
                    const fpls_opts* opts, const fpls_mat* omega, double* U, double* s,
                    double* Vt, int32_t* k_out);

/* ---- learned sketch ------------------------------------------------------------ */
/* Trains a dense sketch Omega (cols(Y) x l) for the rSVD range finder of
   S = Xc^T Yc on the cross-covariances of row-fold draws of (X, Y): masks is
   n x n_draws (host int8, nonzero = row in the draw), each draw centred on its
   own rows.  Loss: the squared Frobenius sine between range(S_d Omega) and the
   exact top-k left singular subspace of S_d, averaged over draws (exact, from
   the eigenpairs of S_d^T S_d); iters Adam steps at learning rate lr from the
   orthonormalised Gaussian stream of opts->seed, Omega re-orthonormalised
   after every step.  omega (cols(Y) x l, host) receives the result,
   loss_trace (iters + 1, optional) the loss before each step and at the end.
   X FP32; 1 <= k <= l <= cols(Y) <= 128.  No counterpart in the reference: the
   learned-sketch extension of the Gaussian range finder
   (src/svd_cpu_rsvd.cpp:95-124, src/fastPLS.cpp:550-623). */
int fpls_sketch_train(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const int8_t* masks,
                      int32_t n_draws, int32_t k, int32_t l, int32_t iters, double lr,
                      const fpls_opts* opts, double* omega, double* loss_trace);
/* Per draw of masks: sin(theta) (spectral) between the range finder's
   l-dimensional subspace of S_d (omega, or the Gaussian stream of opts->seed
   when NULL; power iterations with the reference's schedule) and S_d's exact
   top-k left singular subspace. */
int fpls_sketch_eval(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const int8_t* masks,
                     int32_t n_draws, int32_t k, const double* omega, int32_t l, int32_t power,
                     const fpls_opts* opts, double* sin_theta);

/* ---- Gaussian response compression (gaussian_y) --------------------------- */
/* Output matrices (Z_out, Y_out) are caller buffers described by fpls_mat
   (host or device, FPLS_F32 or FPLS_F64); the library writes through ptr.
   The random projection and codebook come from the std::mt19937 +
   normal_distribution<double> stream of `seed` (the reference draws them
   from R's RNG), scaled by 1/sqrt(dim), column-major.  */

/* Regression (.prepare_gaussian_y, R/main.R:2031-2052): center (m) = column
   means of Y, Yc = Y - center, P (m x dim), Z_out (n x dim) = Yc P,
   decoder (dim x m) = (Z^T Z + ridge I)^{-1} Z^T Yc with ridge = 1e-8
   mean(diag Z^T Z) (.gaussian_y_ridge_decoder, R/main.R:1967-1980);
   projection (m x dim) is optional.  center, decoder, projection: host FP64. */
int fpls_gaussian_y_encode(fpls_ctx* ctx, const fpls_mat* Y, int32_t dim, uint32_t seed,
                           const fpls_mat* Z_out, double* center, double* decoder,
                           double* projection);
/* Classification codebook (.gaussian_y_class_codes, R/main.R:1982-1995):
   codes (C x dim, host FP64) with centred unit-norm columns and unit-norm rows;
   when Z_out is given, Z_out (n x dim) = codes[y, ] (R/main.R:2019-2021). */
int fpls_gaussian_y_codes(fpls_ctx* ctx, const fpls_ivec* y, int32_t n_classes, int32_t dim,
                          uint32_t seed, const fpls_mat* Z_out, double* codes);
/* Regression decode (.decode_gaussian_y_matrix, R/main.R:2091-2096):
   Y_out (n x m) = Z decoder + 1 center^T. */
int fpls_gaussian_y_decode(fpls_ctx* ctx, const fpls_mat* Z, const double* decoder, int32_t m,
                           const double* center, const fpls_mat* Y_out);
/* Class decode (nearest_code_classes, src/fastPLS.cpp:4203-4215;
   R/main.R:2081-2087): scores = 2 Z codes^T - ||code_c||^2, labels (n, host,
   1-based) = first maximum; scores (n x C, host) optional. */
int fpls_gaussian_y_classes(fpls_ctx* ctx, const fpls_mat* Z, const double* codes,
                            int32_t n_classes, int32_t* labels, double* scores);

/* ---- cross-validation ----------------------------------------------------------- */
/* The row -> fold map of pls_cv_predict_compiled (src/fastPLS.cpp:6666-6775):
   the distinct values of constrain (n, host) are the groups; kfold < 0 or
   kfold >= number of groups leaves one group out (fold j = group j in
   ascending order); otherwise kfold (at least 2) folds take a permutation of
   the groups (samplewithoutreplace, src/fastPLS.cpp:4498-4509), per class of
   each group's first row when labels (1-based, host, may be NULL) and
   n_classes > 1 are given, drawn from R's Mersenne-Twister uniforms as
   set.seed(seed) leaves them (R/main.R:3883).  fold_out (n, 1-based), and the
   fold count.  Host only: needs no device. */
int fpls_cv_folds(const int32_t* constrain, int64_t n, int32_t kfold, const int32_t* labels,
                  int32_t n_classes, uint32_t seed, int32_t* fold_out, int32_t* n_folds_out);
/* pls_cv_predict_compiled (src/fastPLS.cpp:6615-7151) with the folds of
   fpls_cv_folds (kfold, opts->seed).  method: FPLS_METHOD_PLSSVD or
   FPLS_METHOD_SIMPLS; classifier: 0 argmax, 1 LDA, 2 candidate kNN (labels
   required).  Each fold fits with seed + fold.  Outputs: fold_out (n, 1-based),
   status_out (>= number of folds: 1 ok, 2 empty, 3 no training rows,
   4 degenerate classification fold), *n_folds_out, class_pred (n x n_ncomp,
   classification) and ypred (n x m x n_ncomp, regression or argmax scores;
   may be NULL for classification). */
int fpls_cv_predict(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const fpls_ivec* y,
                    int32_t n_classes, const int32_t* constrain, const int32_t* ncomp,
                    int32_t n_ncomp, int32_t kfold, int32_t method, const fpls_opts* opts,
                    int32_t classifier, double lda_ridge, int32_t knn_k, double tau,
                    double alpha, int32_t top_m, int32_t* fold_out, int32_t* status_out,
                    int32_t status_cap, int32_t* n_folds_out, int32_t* class_pred,
                    double* ypred);

/* Batched cross-validation: n_fits SIMPLS-fast fits (pls_model2_fast,
   src/fastPLS.cpp:2454-2778) over one resident FP32 X, each defined by a row
   mask, advanced together and replayed as one captured HIP graph.  Replaces
   the per-fold loop of pls_cv_predict_compiled (src/fastPLS.cpp:6777-7142,
   argmax classifier or regression) and, with the fits of every outer fold in
   one batch, the inner loops of pls.double.cv (R/main.R:6676-6764).
     masks   n x n_fits (column-major, host): 0 unused, 1 train, 2 test row
     seeds   n_fits seeds of the Gaussian starts (NULL: opts->seed + fit)
     y / Y   labels 1..n_classes (classification) or responses (regression)
   Outputs: status_out (n_fits: 1 ok, 2 no test rows, 3 no training rows,
   4 fewer than two training classes), and for the test rows of every fit
   class_pred (n x n_ncomp x n_fits, 0 elsewhere), n_correct (n_ncomp x
   n_fits, the accuracy numerators of optim.pls.cv, R/main.R:3744-3763) and
   ypred (n x m x n_ncomp x n_fits: predictions, or argmax scores for
   classification); each output may be NULL.  Limits: p > 5, n_fits <= 128,
   m <= 512, ncomp <= 512, scaling 1 or 3. */
int fpls_cv_batched(fpls_ctx* ctx, const fpls_mat* X, const fpls_mat* Y, const fpls_ivec* y,
                    int32_t n_classes, const int8_t* masks, int32_t n_fits,
                    const uint32_t* seeds, const int32_t* ncomp, int32_t n_ncomp,
                    const fpls_opts* opts, int32_t* status_out, int32_t* class_pred,
                    int32_t* n_correct, double* ypred);

/* ---- Gaussian sketch streams ---------------------------------------------------- */
/* gaussian_matrix_local (src/fastPLS.cpp:99-115) drawn on the device: column s
   of out (n_streams x len doubles, column-major, host) is the
   std::mt19937(seeds[s]) + std::normal_distribution<double>(0, 1) stream.
   Exact integer and IEEE-rounded arithmetic throughout; log is correctly
   rounded, so draws may differ from glibc's by an ulp or two where glibc's
   log is not (about 0.1% of pairs). */
int fpls_gaussian_streams(fpls_ctx* ctx, const uint32_t* seeds, int32_t n_streams, int64_t len,
                          double* out);

/* ---- streaming primitive ------------------------------------------------------ */
/* Class sums (p x C, FP64) and column sums (p) of X in one pass from int32
   labels; counts (C). */
int fpls_class_sums(fpls_ctx* ctx, const fpls_mat* X, const fpls_ivec* y, int32_t n_classes,
                    double* class_sums, double* col_sums, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif /* FASTPLS_HIP_H */
