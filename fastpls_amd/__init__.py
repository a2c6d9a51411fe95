"""fastPLS hot path on MI355X (gfx950): HIP kernels behind a C ABI.

Public API (mirrors the fastPLS R package with backend = "hip"):
``pls``, ``predict``, ``fastsvd``, ``pls_single_cv``, ``optim_pls_cv``,
``pls_double_cv``; ``cv_batched`` runs many cross-validation fits as one
batched launch.
"""
from ._lib import Context, FplsError, available, context, device_colmajor, load  # noqa: F401
from .api import (FastPLS, candidate_knn_predict, class_sums, cv_batched, fastsvd, pls,  # noqa: F401
                  predict, truncated_rsvd_crossprod, optim_pls_cv, pls_double_cv, pls_single_cv,
                  pls_cv_predict, sketch_eval, sketch_train, prepare_gaussian_y,
                  decode_gaussian_y_matrix, decode_gaussian_y_cube, gaussian_y_default_dim,
                  class_predict_topk, class_topk_from_score_cube, class_bias_matrix,
                  resolve_top_k, opls_filter, opls_apply_filter, kernel_matrix,
                  center_kernel_train, center_kernel_test, model_from_fields, compiled_kfold,
                  cv_folds)

__all__ = ["pls", "predict", "fastsvd", "optim_pls_cv", "pls_single_cv", "pls_cv_predict",
           "pls_double_cv", "cv_batched", "sketch_train", "sketch_eval", "prepare_gaussian_y",
           "decode_gaussian_y_matrix", "decode_gaussian_y_cube", "gaussian_y_default_dim",
           "class_predict_topk", "class_topk_from_score_cube", "class_bias_matrix",
           "resolve_top_k", "opls_filter", "opls_apply_filter", "kernel_matrix",
           "center_kernel_train", "center_kernel_test", "model_from_fields", "compiled_kfold",
           "cv_folds", "FastPLS", "Context", "FplsError", "available",
           "context", "device_colmajor", "load"]
