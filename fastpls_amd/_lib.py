"""ctypes binding of the C ABI in include/fastpls_hip.h.

Loads the in-tree ``fastpls_amd/libfastpls_hip.so``.  There is no CPU
fallback: if the library is missing or no gfx950 device is present the calls
raise.
"""
from __future__ import annotations

import ctypes
import os
import sys
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastpls_hip.so")

FPLS_OK = 0
ERRORS = {
    -1: "FPLS_ERR_ARG", -2: "FPLS_ERR_HIP", -3: "FPLS_ERR_EMPTY_CLASS",
    -4: "FPLS_ERR_LABEL_RANGE", -5: "FPLS_ERR_RANK", -6: "FPLS_ERR_SOLVE",
    -7: "FPLS_ERR_NO_DEVICE", -8: "FPLS_ERR_STATE",
}
F32, F64 = 0, 1
HOST, DEVICE = 0, 1
SVD_EXACT, SVD_RSVD = 3, 4
PREC_FP32, PREC_BF16X3 = 0, 1
METHOD_PLSSVD, METHOD_SIMPLS, METHOD_OPLS, METHOD_KERNELPLS = 1, 3, 4, 5

FIELDS = {
    "R": 1, "Q": 2, "mX": 3, "vX": 4, "mY": 5, "B": 6, "C_latent": 7, "W_latent": 8,
    "s": 9, "Ttrain": 10, "R2Y": 11, "Yfit": 12, "V": 13, "R_predict": 14, "R_offset": 15,
}
LDA_FIELDS = {"means": 1, "inv_cov": 2, "linear": 3, "constants": 4, "priors": 5, "ridge": 6}


class FplsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Mat(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("loc", ctypes.c_int32),
                ("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


class IVec(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("loc", ctypes.c_int32), ("n", ctypes.c_int64)]


class Opts(ctypes.Structure):
    _fields_ = [("scaling", ctypes.c_int32), ("svd_method", ctypes.c_int32),
                ("oversample", ctypes.c_int32), ("power", ctypes.c_int32),
                ("seed", ctypes.c_uint32), ("inc_iters", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("fit", ctypes.c_int32),
                ("store_b", ctypes.c_int32), ("label_aware", ctypes.c_int32),
                ("xprod_route", ctypes.c_int32), ("opls_north", ctypes.c_int32),
                ("xprod", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class ModelInfo(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("n", ctypes.c_int64), ("p", ctypes.c_int64),
                ("m", ctypes.c_int64), ("kmax", ctypes.c_int32), ("n_ncomp", ctypes.c_int32),
                ("n_components", ctypes.c_int32), ("b_stored", ctypes.c_int32),
                ("has_lda", ctypes.c_int32), ("has_cknn", ctypes.c_int32),
                ("n_classes", ctypes.c_int32), ("ncomp", ctypes.c_int32 * 64),
                ("xprod_mode", ctypes.c_int32)]


class ModelFields(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("kmax", ctypes.c_int32), ("p", ctypes.c_int64),
                ("m", ctypes.c_int64), ("n_ncomp", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("ncomp", ctypes.c_void_p), ("R", ctypes.c_void_p), ("Q", ctypes.c_void_p),
                ("C_latent", ctypes.c_void_p), ("W_latent", ctypes.c_void_p),
                ("B", ctypes.c_void_p), ("mX", ctypes.c_void_p), ("vX", ctypes.c_void_p),
                ("mY", ctypes.c_void_p)]


class LdaFields(ctypes.Structure):
    _fields_ = [("n_classes", ctypes.c_int32), ("n_models", ctypes.c_int32),
                ("ncomp", ctypes.c_void_p), ("means", ctypes.c_void_p),
                ("inv_cov", ctypes.c_void_p), ("linear", ctypes.c_void_p),
                ("constants", ctypes.c_void_p), ("priors", ctypes.c_void_p),
                ("ridge", ctypes.c_void_p)]


class CknnFields(ctypes.Structure):
    _fields_ = [("n_classes", ctypes.c_int32), ("n_models", ctypes.c_int32),
                ("ncomp", ctypes.c_void_p), ("centroids", ctypes.c_void_p),
                ("n_train", ctypes.c_int64), ("ttrain_cols", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("Ttrain", ctypes.c_void_p),
                ("y_codes", ctypes.c_void_p), ("knn_k", ctypes.c_int32),
                ("top_m", ctypes.c_int32), ("tau", ctypes.c_double),
                ("alpha", ctypes.c_double)]


_P = ctypes.POINTER
_vp = ctypes.c_void_p
SIGNATURES = {
    "fpls_opts_default": (None, [_P(Opts)]),
    "fpls_available": (ctypes.c_int, []),
    "fpls_ctx_create": (ctypes.c_int, [ctypes.c_int, _P(_vp)]),
    "fpls_ctx_destroy": (None, [_vp]),
    "fpls_last_error": (ctypes.c_char_p, [_vp]),
    "fpls_ctx_synchronize": (ctypes.c_int, [_vp]),
    "fpls_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fpls_kernel_time": (ctypes.c_int, [_vp, ctypes.c_char_p, _P(ctypes.c_int64), _P(ctypes.c_double)]),
    "fpls_autotune_set": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fpls_autotune_report": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int64]),
    "fpls_plssvd_fit": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                       ctypes.c_int32, _P(Opts), _P(_vp)]),
    "fpls_simpls_fit": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                       ctypes.c_int32, _P(Opts), _P(_vp)]),
    "fpls_plssvd_lda_fit": (ctypes.c_int, [_vp, _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                           ctypes.c_int32, _P(Opts), ctypes.c_double, _P(_vp)]),
    "fpls_plssvd_lda_fit_predict": (ctypes.c_int, [_vp, _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                                   ctypes.c_int32, _P(Opts), ctypes.c_double,
                                                   _P(Mat), _vp, _vp, _P(_vp)]),
    "fpls_model_info_get": (ctypes.c_int, [_vp, _P(ModelInfo)]),
    "fpls_model_import": (ctypes.c_int, [_vp, _P(ModelFields), _P(_vp)]),
    "fpls_model_import_lda": (ctypes.c_int, [_vp, _vp, _P(LdaFields)]),
    "fpls_model_import_cknn": (ctypes.c_int, [_vp, _vp, _P(CknnFields)]),
    "fpls_model_read": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp, ctypes.c_int64]),
    "fpls_model_free": (None, [_vp]),
    "fpls_predict": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp]),
    "fpls_project": (ctypes.c_int, [_vp, _vp, _P(Mat), ctypes.c_int32, _vp]),
    "fpls_lda_train_prefix": (ctypes.c_int, [_vp, _vp, _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                             ctypes.c_int32, ctypes.c_double]),
    "fpls_lda_read": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int64]),
    "fpls_lda_predict": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp, _vp]),
    "fpls_lda_predict_margins": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp, _vp, _vp]),
    "fpls_lda_predict_noalloc": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp, _vp]),
    "fpls_lda_train_scores": (ctypes.c_int, [_vp, _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                             ctypes.c_int32, ctypes.c_double, _vp, _vp, _vp, _vp,
                                             _vp, _vp]),
    "fpls_cknn_fit": (ctypes.c_int, [_vp, _vp, _P(Mat), _P(IVec), ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int32]),
    "fpls_cknn_predict": (ctypes.c_int, [_vp, _vp, _P(Mat), ctypes.c_int32, _vp, _vp]),
    "fpls_cknn_predict_margins": (ctypes.c_int, [_vp, _vp, _P(Mat), ctypes.c_int32, _vp, _vp,
                                                 _vp]),
    "fpls_cknn_predict_scores": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), _P(Mat), _vp,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "fpls_cknn_predict_scores_margins": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), _P(Mat),
                                                        _vp, ctypes.c_int32, ctypes.c_int32,
                                                        ctypes.c_int32, ctypes.c_double,
                                                        ctypes.c_double, _vp, _vp, _vp]),
    "fpls_rsvd": (ctypes.c_int, [_vp, _P(Mat), ctypes.c_int32, _P(Opts), _P(Mat), _vp, _vp, _vp,
                                 _P(ctypes.c_int32)]),
    "fpls_rsvd_xprod": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), ctypes.c_int32, _P(Opts), _P(Mat),
                                       _vp, _vp, _vp, _P(ctypes.c_int32)]),
    "fpls_sketch_train": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _vp, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _P(Opts),
                                         _vp, _vp]),
    "fpls_sketch_eval": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _vp, ctypes.c_int32, ctypes.c_int32,
                                        _vp, ctypes.c_int32, ctypes.c_int32, _P(Opts), _vp]),
    "fpls_class_sums": (ctypes.c_int, [_vp, _P(Mat), _P(IVec), ctypes.c_int32, _vp, _vp, _vp]),
    "fpls_gaussian_y_encode": (ctypes.c_int, [_vp, _P(Mat), ctypes.c_int32, ctypes.c_uint32,
                                              _P(Mat), _vp, _vp, _vp]),
    "fpls_gaussian_y_codes": (ctypes.c_int, [_vp, _P(IVec), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_uint32, _P(Mat), _vp]),
    "fpls_gaussian_y_decode": (ctypes.c_int, [_vp, _P(Mat), _vp, ctypes.c_int32, _vp, _P(Mat)]),
    "fpls_gaussian_y_classes": (ctypes.c_int, [_vp, _P(Mat), _vp, ctypes.c_int32, _vp, _vp]),
    "fpls_class_predict_topk": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp, ctypes.c_int32,
                                               ctypes.c_int32, _vp, _vp]),
    "fpls_class_topk_scores": (ctypes.c_int, [_vp, _P(Mat), ctypes.c_int32, ctypes.c_int32, _vp,
                                              ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    "fpls_gaussian_y_predict_classes": (ctypes.c_int, [_vp, _vp, _P(Mat), _vp, ctypes.c_int32,
                                                       ctypes.c_int32, _vp, _vp]),
    "fpls_opls_filter": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_int32, _P(Mat), _vp, _vp, _vp,
                                        _vp, _vp]),
    "fpls_opls_apply": (ctypes.c_int, [_vp, _P(Mat), _vp, _vp, _vp, _vp, ctypes.c_int32,
                                       _P(Mat)]),
    "fpls_kernel_matrix": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int32, ctypes.c_double, _P(Mat)]),
    "fpls_center_kernel_train": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _vp, _vp]),
    "fpls_center_kernel_test": (ctypes.c_int, [_vp, _P(Mat), _vp, ctypes.c_double, _P(Mat)]),
    "fpls_cv_predict": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), ctypes.c_int32, _vp, _vp,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P(Opts),
                                       ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int32, _vp, _vp,
                                       ctypes.c_int32, _P(ctypes.c_int32), _vp, _vp]),
    "fpls_cv_batched": (ctypes.c_int, [_vp, _P(Mat), _P(Mat), _P(IVec), ctypes.c_int32, _vp,
                                       ctypes.c_int32, _vp, _vp, ctypes.c_int32, _P(Opts), _vp,
                                       _vp, _vp, _vp]),
    "fpls_cv_folds": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp, ctypes.c_int32,
                                     ctypes.c_uint32, _vp, _P(ctypes.c_int32)]),
    "fpls_gaussian_streams": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int64, _vp]),
}

_lib = None


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_opts(**kw) -> Opts:
    o = Opts()
    load().fpls_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ---------------------------------------------------------------- matrices

def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def as_mat(x, keep: list) -> Mat:
    """Describe a column-major host (numpy) or device (torch) matrix.

    numpy arrays are converted to Fortran order float32/float64 (a copy is
    kept alive in ``keep``).  torch tensors must already be on the GPU in
    column-major layout (stride (1, ld)) -- see ``device_colmajor``."""
    if _is_torch(x):
        import torch
        if x.dim() == 1:
            x = x.reshape(-1, 1)
        assert x.is_cuda, "torch input must live on the GPU"
        if x.dtype == torch.float32:
            dt = F32
        elif x.dtype == torch.float64:
            dt = F64
        else:
            raise TypeError("torch input must be float32 or float64")
        rows, cols = x.shape
        s0, s1 = x.stride()
        if not (s0 == 1 and (s1 >= rows or cols == 1)):
            raise ValueError("torch matrix must be column-major (use device_colmajor)")
        ld = s1 if cols > 1 else max(rows, 1)
        keep.append(x)
        return Mat(x.data_ptr(), dt, DEVICE, rows, cols, max(ld, 1))
    a = np.asarray(x)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    a = np.asfortranarray(a)
    keep.append(a)
    return Mat(a.ctypes.data, F32 if a.dtype == np.float32 else F64, HOST, a.shape[0],
               a.shape[1], max(a.shape[0], 1))


def as_ivec(y, keep: list) -> IVec:
    if _is_torch(y):
        import torch
        assert y.dtype == torch.int32 and y.is_cuda and y.is_contiguous()
        keep.append(y)
        return IVec(y.data_ptr(), DEVICE, y.numel())
    a = np.ascontiguousarray(np.asarray(y, dtype=np.int32).ravel())
    keep.append(a)
    return IVec(a.ctypes.data, HOST, a.size)


def device_colmajor(a: np.ndarray, device: str = "cuda:0"):
    """Copy a 2-D numpy array to the GPU as a column-major torch tensor."""
    import torch
    a = np.asarray(a)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to(device)  # (cols, rows) row-major
    return t.t()  # (rows, cols) with stride (1, rows)


# ---------------------------------------------------------------- context

class Context:
    """One HIP stream + HBM pool on one device (fpls_ctx)."""

    def __init__(self, device: int = 0):
        # torch bundles its own HIP runtime, which cannot initialise after this
        # library's has: when torch is in use, let it go first
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        lib = load()
        h = ctypes.c_void_p()
        rc = lib.fpls_ctx_create(device, ctypes.byref(h))
        if rc != FPLS_OK:
            raise FplsError(rc, f"cannot create a context on device {device} (gfx950 required)")
        self.h = h
        self.lib = lib
        self.device = device

    def check(self, rc: int):
        if rc != FPLS_OK:
            raise FplsError(rc, self.lib.fpls_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.fpls_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self, enable: bool = True):
        self.check(self.lib.fpls_timing_enable(self.h, int(enable)))

    def kernel_time(self, name: str):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self.check(self.lib.fpls_kernel_time(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def autotune(self, mode: int = 1, force_mb: int = 0, force_splits: int = 0):
        """The projection-tile autotuner: mode 1 times candidates on a shape's
        first call, 0 keeps the built-in rule; force_* pin a configuration."""
        self.check(self.lib.fpls_autotune_set(self.h, int(mode), int(force_mb),
                                              int(force_splits)))

    def autotune_report(self) -> dict:
        """The tuning table: {shape key: {mb, rows, splits, us, tuned}}."""
        import json
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self.lib.fpls_autotune_report(self.h, buf, cap)
            if rc == 0:
                return json.loads(buf.value.decode())
            if rc <= cap:
                raise RuntimeError("fpls_autotune_report failed")
            cap = rc

    def synchronize(self):
        self.check(self.lib.fpls_ctx_synchronize(self.h))


_default_ctx: Optional[Context] = None


def context(device: int = 0) -> Context:
    global _default_ctx
    if _default_ctx is None or _default_ctx.h is None or _default_ctx.device != device:
        _default_ctx = Context(device)
    return _default_ctx


def available() -> bool:
    try:
        return bool(load().fpls_available())
    except Exception:
        return False


def int_array(vals: Sequence[int]):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.int32).ravel())
    return a, a.ctypes.data
