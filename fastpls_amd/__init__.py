"""fastPLS hot path on MI355X (gfx950): HIP kernels behind a C ABI.

Public API (mirrors the fastPLS R package with backend = "hip"):
``pls``, ``predict``, ``fastsvd``, ``pls_single_cv``, ``optim_pls_cv``.
"""
from ._lib import Context, FplsError, available, context, device_colmajor, load  # noqa: F401
from .api import FastPLS, candidate_knn_predict, fastsvd, pls, predict  # noqa: F401

__all__ = ["pls", "predict", "fastsvd", "FastPLS", "Context", "FplsError", "available",
           "context", "device_colmajor", "load"]
