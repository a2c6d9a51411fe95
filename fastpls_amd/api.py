"""Host-side mirror of the fastPLS R API for ``backend = "hip"``.

``pls()``, ``predict()``, ``fastsvd()``, ``pls_single_cv()`` and
``optim_pls_cv()`` keep the reference's argument names and meanings
(R/main.R:5860 ``pls``, :2577 ``predict.fastPLS``, :4818 ``fastsvd``,
:4316 ``pls.single.cv``, :6399 ``optim.pls.cv``); the numerical work runs in
the gfx950 library through the C ABI (``include/fastpls_hip.h``).  Returned
models are dicts with the reference's field names plus a ``handle`` to the
HBM-resident model.  Only ``backend="hip"`` exists here: anything else raises,
and there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import as_ivec, as_mat, int_array

SCALINGS = {"centering": 1, "autoscaling": 2, "none": 3}


class FastPLS(dict):
    """A fitted model: reference field names + the device-resident handle."""

    def __init__(self, ctx: _lib.Context, handle: ctypes.c_void_p):
        super().__init__()
        self.ctx = ctx
        self.handle = handle

    def free(self):
        if getattr(self, "handle", None) and self.ctx.h:
            self.ctx.lib.fpls_model_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self) -> _lib.ModelInfo:
        inf = _lib.ModelInfo()
        self.ctx.check(self.ctx.lib.fpls_model_info_get(self.handle, ctypes.byref(inf)))
        return inf

    def read(self, field: str, shape) -> np.ndarray:
        n = int(np.prod(shape))
        out = np.empty(n, dtype=np.float64)
        self.ctx.check(self.ctx.lib.fpls_model_read(self.ctx.h, self.handle, _lib.FIELDS[field],
                                                    out.ctypes.data, n))
        return out.reshape(shape, order="F")

    def lda_read(self, slice_idx: int, field: str, shape) -> np.ndarray:
        n = int(np.prod(shape))
        out = np.empty(n, dtype=np.float64)
        self.ctx.check(self.ctx.lib.fpls_lda_read(self.ctx.h, self.handle, slice_idx,
                                                  _lib.LDA_FIELDS[field], out.ctypes.data, n))
        return out.reshape(shape, order="F")


def should_use_xprod_default(p, q, ncomp) -> bool:
    """.should_use_xprod_default (R/main.R:1883-1892): the matrix-free route
    when S = X^T Y would exceed 32 MiB, or for q >= 100 responses and at most
    10 components."""
    try:
        p, q = float(p), float(q)
        nc = max(int(v) for v in np.atleast_1d(ncomp))
    except (TypeError, ValueError):
        return False
    if not (np.isfinite(p) and np.isfinite(q)):
        return False
    s_mb = p * q * 8 / 1024.0 ** 2
    return bool(s_mb > 32 or (q >= 100 and nc <= 10))


def _n_levels(y) -> int:
    """Level count of host labels (0 when it cannot be read on the host)."""
    try:
        return int(len(np.unique(np.asarray(y))))
    except (TypeError, ValueError, RuntimeError):
        return 0


def _is_factor(y) -> bool:
    if type(y).__module__.startswith("torch"):
        import torch
        return y.dtype in (torch.int32, torch.int64)
    a = np.asarray(y)
    return a.dtype.kind in "iubOUS" and (a.ndim == 1 or (a.ndim == 2 and a.shape[1] == 1))


def _encode_factor(y, n_classes=None):
    """Levels (sorted unique) and 1-based codes, as R's factor().  Device
    (torch) labels are taken as codes 1..C; C is ``n_classes`` when given (a
    factor's level count), else a device reduction."""
    if type(y).__module__.startswith("torch"):
        import torch
        codes = y if y.dtype == torch.int32 else y.to(torch.int32)
        C = int(n_classes) if n_classes is not None else int(codes.max().item())
        return list(range(1, C + 1)), codes, C
    a = np.asarray(y).ravel()
    lev, codes = np.unique(a, return_inverse=True)
    return list(lev), (codes + 1).astype(np.int32), len(lev)


def _labels_from_codes(lev, codes):
    """R's factor labels of 1-based class codes.  Levels that are the integers
    1..C (device labels, or host labels coded that way) are the codes
    themselves, returned without a per-row lookup."""
    C = len(lev)
    if C and isinstance(lev[0], (int, np.integer)) and lev[0] == 1 and lev[-1] == C:
        return codes  # sorted unique integers from 1 to C: exactly 1..C
    levs = np.asarray(lev)
    if levs.dtype.kind not in "iufUS":
        levs = np.asarray(lev, dtype=object)
    return levs[codes - 1]


def _check_backend(backend: str):
    if backend != "hip":
        raise ValueError("this package implements backend='hip' only")


def _fetch_fields(model: FastPLS, fit: bool):
    inf = model.info()
    p, m, k, L = inf.p, inf.m, inf.kmax, inf.n_ncomp
    model["p"], model["m"] = int(p), int(m)
    model["ncomp"] = np.array(inf.ncomp[:L], dtype=np.int64)
    model["R"] = model.read("R", (p, k))
    model["Q"] = model.read("Q", (m, k))
    model["mX"] = model.read("mX", (p,))
    model["vX"] = model.read("vX", (p,))
    model["mY"] = model.read("mY", (m,))
    if inf.b_stored:
        model["B"] = model.read("B", (p, m, L))
    model["B_stored"] = bool(inf.b_stored)
    model["compact_prediction"] = not bool(inf.b_stored)
    if inf.method == _lib.METHOD_PLSSVD:
        model["C_latent"] = model.read("C_latent", (k, k, L))
        model["W_latent"] = model.read("W_latent", (k, m, L))
        model["s"] = model.read("s", (k,))
        model["pls_method"] = "plssvd"
    else:
        model["pls_method"] = "simpls"
        model["n_components"] = int(inf.n_components)
    if fit:
        model["R2Y"] = model.read("R2Y", (L,))
        model["Yfit"] = model.read("Yfit", (inf.n, m, L))
    model["R_predict"] = model.read("R_predict", (p, k))
    model["R_offset"] = model.read("R_offset", (k,))


def pls(Xtrain, Ytrain, Xtest=None, Ytest=None, ncomp=2, scaling="centering",
        method="simpls", svd_method="rsvd", classifier="argmax", lda_ridge=1e-8,
        candidate_knn_k=10, candidate_tau=0.2, candidate_alpha=0.75, candidate_top_m=20,
        fit=False, backend="hip", seed=1, rsvd_oversample=10, rsvd_power=1,
        precision="fp32", label_aware=False, fetch=True, device=0, top=1,
        n_classes=None, gaussian_y=False, gaussian_y_dim=None, north=1, kernel="linear",
        gamma=None, degree=3, coef0=1.0, tie_log=False, xprod=None) -> FastPLS:
    """pls() (R/main.R:5860-6314) with backend = "hip".  ``n_classes``: level
    count of device-resident integer labels (optional).  ``gaussian_y``: fit a
    Gaussian compression of the response and decode the predictions
    (.prepare_gaussian_y, R/main.R:6057-6068; argmax classifier).
    ``method="opls"`` (``north`` orthogonal components) and
    ``method="kernelpls"`` (``kernel``, ``gamma``, ``degree``, ``coef0``) fit
    SIMPLS on the filtered X / the centred training kernel (R/main.R:5956-6006).
    ``tie_log``: with Xtest and an LDA / candidate-kNN head, also return the
    device's per-row winning margins as ``tie_margin`` (see predict()).
    ``xprod``: None applies .should_use_xprod_default with the response width
    (R/main.R:6080-6085, the rSVD routes of PLSSVD and SIMPLS); True / False
    force the route.  A dense-response PLSSVD fit on the route takes the
    matrix-free range finder of S = X^T Y (pls.model1.rsvd.xprod.precision,
    R/main.R:6091-6108); label responses keep the one-pass class sums and
    SIMPLS keeps S resident, so their ``xprod_mode`` stays "materialized"."""
    _check_backend(backend)
    if method in ("opls", "kernelpls"):
        inner_kw = dict(ncomp=ncomp, classifier=classifier, lda_ridge=lda_ridge,
                        candidate_knn_k=candidate_knn_k, candidate_tau=candidate_tau,
                        candidate_alpha=candidate_alpha, candidate_top_m=candidate_top_m,
                        fit=fit, seed=seed, rsvd_oversample=rsvd_oversample,
                        rsvd_power=rsvd_power, precision=precision, fetch=fetch, device=device,
                        n_classes=n_classes, gaussian_y=gaussian_y,
                        gaussian_y_dim=gaussian_y_dim)
        if method == "opls":
            model = _opls_fit(Xtrain, Ytrain, int(north), scaling, inner_kw)
        else:
            model = _kernel_pls_fit(Xtrain, Ytrain, scaling, kernel, gamma, degree, coef0,
                                    inner_kw)
        if Xtest is not None:
            model.update(predict(model, Xtest, Ytest, top=top, tie_log=tie_log))
        return model
    if gaussian_y:
        if classifier not in ("argmax", "lda", "cknn"):
            raise ValueError("classifier must be 'argmax', 'lda' or 'cknn'")
        yprep = prepare_gaussian_y(Ytrain, Xtrain, True, gaussian_y_dim, seed, device)
        model = pls(Xtrain, yprep["Ytrain"], None, None, ncomp, scaling, method, svd_method,
                    "argmax", lda_ridge, candidate_knn_k, candidate_tau, candidate_alpha,
                    candidate_top_m, fit, backend, seed, rsvd_oversample, rsvd_power,
                    precision, False, fetch, device, top)
        _attach_gaussian_y(model, yprep)
        if yprep["classification"] and classifier != "argmax":
            # .attach_lda_classifier with the original labels (R/main.R:6197-6203):
            # the head is trained on the latent scores of the model of the codes
            keep_h: list = []
            _, codes, C = _encode_factor(Ytrain, n_classes)
            yv = as_ivec(codes, keep_h)
            Xm = as_mat(Xtrain, keep_h)
            _attach_head(model, classifier, yv, C, lda_ridge, candidate_knn_k, candidate_tau,
                         candidate_alpha, candidate_top_m, Xm)
        if fit and fetch and not yprep["classification"] and "Yfit" in model:
            Yf = decode_gaussian_y_cube(model["gaussian_y"], model["Yfit"], device)
            model["Yfit"] = Yf
            Yo = np.asarray(Ytrain, dtype=np.float64).reshape(Yf.shape[0], -1)
            tss = float(((Yo - Yo.mean(axis=0)) ** 2).sum())
            model["R2Y"] = np.array([1.0 - float(((Yo - Yf[:, :, a]) ** 2).sum()) / tss
                                     for a in range(Yf.shape[2])])
        if Xtest is not None:
            model.update(predict(model, Xtest, Ytest, top=top, tie_log=tie_log))
        return model
    ctx = _lib.context(device)
    keep: list = []
    X = as_mat(Xtrain, keep)
    classification = _is_factor(Ytrain)
    if xprod is None:
        if classification:
            q_resp = int(n_classes) if n_classes else _n_levels(Ytrain)
        else:
            q_resp = 1 if len(getattr(Ytrain, "shape", np.shape(Ytrain))) == 1 \
                else int(Ytrain.shape[1] if hasattr(Ytrain, "shape") else np.shape(Ytrain)[1])
        xprod = svd_method != "exact" and method in ("plssvd", "simpls") and \
            should_use_xprod_default(X.cols, q_resp, ncomp)
    opts = _lib.default_opts(scaling=SCALINGS[scaling], seed=int(seed) & 0xFFFFFFFF,
                             oversample=int(rsvd_oversample), power=int(rsvd_power),
                             svd_method=_lib.SVD_EXACT if svd_method == "exact" else _lib.SVD_RSVD,
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32,
                             fit=int(bool(fit)), label_aware=int(bool(label_aware)),
                             xprod=int(bool(xprod)))
    nc_arr, nc_ptr = int_array(np.atleast_1d(ncomp))
    h = ctypes.c_void_p()
    yv = None
    Ym = None
    lev = None
    C = 0
    if classification:
        lev, codes, C = _encode_factor(Ytrain, n_classes)
        yv = as_ivec(codes, keep)
    else:
        Ym = as_mat(Ytrain, keep)
    fn = ctx.lib.fpls_plssvd_fit if method == "plssvd" else ctx.lib.fpls_simpls_fit
    if method not in ("plssvd", "simpls"):
        raise ValueError("method must be 'plssvd' or 'simpls'")
    fused_lda = method == "plssvd" and classification and classifier == "lda"
    # with test rows and top-1 labels, the prediction joins the same call (one
    # host round trip at its end)
    fused_pred = None
    if fused_lda and Xtest is not None and resolve_top_k(top) == 1:
        Xt = as_mat(Xtest, keep)
        fused_pred = np.empty((Xt.rows, nc_arr.size), dtype=np.int32, order="F")
        fused_margin = (np.empty((Xt.rows, nc_arr.size), dtype=np.float64, order="F")
                        if tie_log else None)
        ctx.check(ctx.lib.fpls_plssvd_lda_fit_predict(
            ctx.h, ctypes.byref(X), ctypes.byref(yv), C, nc_ptr, nc_arr.size,
            ctypes.byref(opts), float(lda_ridge), ctypes.byref(Xt), fused_pred.ctypes.data,
            fused_margin.ctypes.data if tie_log else None, ctypes.byref(h)))
    elif fused_lda:
        # the fit and the LDA head in one call (one launch for their solves)
        ctx.check(ctx.lib.fpls_plssvd_lda_fit(ctx.h, ctypes.byref(X), ctypes.byref(yv), C,
                                              nc_ptr, nc_arr.size, ctypes.byref(opts),
                                              float(lda_ridge), ctypes.byref(h)))
    else:
        ctx.check(fn(ctx.h, ctypes.byref(X), ctypes.byref(Ym) if Ym is not None else None,
                     ctypes.byref(yv) if yv is not None else None, C, nc_ptr, nc_arr.size,
                     ctypes.byref(opts), ctypes.byref(h)))
    model = FastPLS(ctx, h)
    model["classification"] = classification
    model["lev"] = lev
    model["predict_backend"] = "hip"
    inf = model.info()
    model["xprod_mode"] = {1: "implicit", 2: "label_aware_stream"}.get(int(inf.xprod_mode),
                                                                      "materialized")
    model["xprod_default"] = bool(xprod)
    model["ncomp"] = np.array(inf.ncomp[:inf.n_ncomp], dtype=np.int64)
    if fetch:
        _fetch_fields(model, bool(fit))
    if fused_lda:
        model["classification_rule"] = "lda_hip"
        model["lda"] = {"ncomp": sorted(set(int(min(max(c, 1), inf.kmax))
                                            for c in model["ncomp"])),
                        "ridge": float(lda_ridge)}
    elif classification and classifier in ("lda", "cknn"):
        _attach_head(model, classifier, yv, C, lda_ridge, candidate_knn_k, candidate_tau,
                     candidate_alpha, candidate_top_m)
    else:
        model["classification_rule"] = "argmax"
    if fused_pred is not None:
        res = {"Ypred_index": fused_pred,
               "Ypred": _labels_from_codes(lev, fused_pred)}
        if tie_log:
            res["tie_margin"] = fused_margin
        if Ytest is not None:
            L = fused_pred.shape[1]
            res["Q2Y"] = _q2_classes(Ytest, lev, np.stack(
                [(fused_pred[:, a][:, None] == np.arange(1, C + 1)[None, :]).astype(np.float64)
                 for a in range(L)], axis=2))
        model.update(res)
    elif Xtest is not None:
        model.update(predict(model, Xtest, Ytest, top=top, tie_log=tie_log))
    return model


def resolve_top_k(top=1, top5: bool = False) -> int:
    """.resolve_top_k (R/main.R:368-375)."""
    t = np.atleast_1d(top)[0] if top is not None else None
    if t is None or not np.isfinite(float(t)) or int(t) < 1:
        raise ValueError("top must be a positive integer")
    t = int(t)
    return max(t, 5) if top5 else t


def class_bias_matrix(class_bias, lev, n_slices: int) -> np.ndarray:
    """.class_bias_matrix (R/main.R:377-415): an m x n_slices FP64 offset
    matrix from None (zeros), a per-class vector (a dict is matched to the
    levels by name) or a matrix with one column or one column per slice
    (transposed when it is n_slices x m)."""
    C = len(lev)
    if class_bias is None:
        return np.zeros((C, n_slices), order="F")
    if isinstance(class_bias, dict):
        try:
            vals = [float(class_bias[l]) for l in lev]
        except KeyError:
            raise ValueError("class_bias must have one numeric value per class") from None
        class_bias = np.asarray(vals)
    b = np.asarray(class_bias, dtype=np.float64)
    if b.ndim <= 1:
        b = b.reshape(-1)
        if b.size != C or np.isnan(b).any():
            raise ValueError("class_bias must have one numeric value per class")
        return np.asfortranarray(np.repeat(b[:, None], n_slices, axis=1))
    if b.shape[0] != C and b.shape[1] == C:
        b = b.T
    if b.shape[0] != C or np.isnan(b).any():
        raise ValueError("class_bias must have one row per class")
    if b.shape[1] == 1 and n_slices > 1:
        b = np.repeat(b, n_slices, axis=1)
    if b.shape[1] != n_slices:
        raise ValueError("class_bias must have one column or one column per ncomp")
    return np.asfortranarray(b)


def _topk_result(ti: np.ndarray, ts: np.ndarray, lev) -> Dict:
    """.class_topk_to_labels (R/main.R:417-451): labels of rank 1 per slice,
    their indices, and for top > 1 the ranked index and score cubes."""
    labs = np.asarray(lev, dtype=object)
    res = {"Ypred_index": np.ascontiguousarray(ti[:, 0, :]),
           "Ypred": labs[ti[:, 0, :] - 1]}
    if ti.shape[1] > 1:
        res["Ypred_top"] = ti
        res["Ypred_top_score"] = ts
        res["Ypred_top_label"] = np.where(ti > 0, labs[np.maximum(ti, 1) - 1], None)
    return res


def class_predict_topk(model: FastPLS, newdata, class_bias=None, top: int = 1) -> Dict:
    """.class_bias_predict (R/main.R:476-492) -> pls_class_predict_topk
    (src/fastPLS.cpp:3954-4113): the argmax head's ranked classes with
    per-class offsets, computed on the device from the compact factors."""
    ctx = model.ctx
    keep: list = []
    X = as_mat(newdata, keep)
    inf = model.info()
    L, m = inf.n_ncomp, inf.m
    lev = model.get("lev")
    if lev is None:
        lev = [str(i + 1) for i in range(m)]
    bias = class_bias_matrix(class_bias, lev, L)
    k = max(1, min(int(top), m))
    ti = np.empty((X.rows, k, L), dtype=np.int32, order="F")
    ts = np.empty((X.rows, k, L), dtype=np.float64, order="F")
    ctx.check(ctx.lib.fpls_class_predict_topk(ctx.h, model.handle, ctypes.byref(X),
                                              bias.ctypes.data, L, int(top), ti.ctypes.data,
                                              ts.ctypes.data))
    return _topk_result(ti, ts, lev)


def class_topk_from_score_cube(score_cube, lev, class_bias=None, top: int = 1,
                               device: int = 0) -> Dict:
    """.class_topk_from_score_cube (R/main.R:453-474) on the device
    (fpls_class_topk_scores): score_cube is w x C x L (host array, or a device
    FP32/FP64 tensor of that shape)."""
    ctx = _lib.context(device)
    keep: list = []
    if hasattr(score_cube, "is_cuda"):
        w, C = int(score_cube.shape[0]), int(score_cube.shape[1])
        L = int(score_cube.shape[2]) if score_cube.dim() == 3 else 1
        flat = score_cube.reshape(w, C, L).permute(2, 1, 0).contiguous().reshape(L * C, w).t()
        S = as_mat(flat, keep)
    else:
        cube = np.asarray(score_cube, dtype=np.float64)
        if cube.ndim == 2:
            cube = cube[:, :, None]
        w, C, L = cube.shape
        S = as_mat(np.asfortranarray(cube).reshape(w, C * L, order="F"), keep)
    bias = class_bias_matrix(class_bias, lev, L)
    k = max(1, min(int(top), C))
    ti = np.empty((w, k, L), dtype=np.int32, order="F")
    ts = np.empty((w, k, L), dtype=np.float64, order="F")
    ctx.check(ctx.lib.fpls_class_topk_scores(ctx.h, ctypes.byref(S), C, L, bias.ctypes.data, L,
                                             int(top), ti.ctypes.data, ts.ctypes.data))
    return _topk_result(ti, ts, lev)


def _q2_classes(Ytest, lev, cube: np.ndarray) -> np.ndarray:
    """Q2Y of a classification prediction (R/main.R:2694-2719): the one-hot
    Ytest against each slice of the score cube, RQ (src/fastPLS.cpp:1483)."""
    yt = np.asarray(Ytest).astype(str)
    onehot = (yt[:, None] == np.asarray(lev, dtype=str)[None, :]).astype(np.float64)
    tss = float(((onehot - onehot.mean(axis=0)) ** 2).sum())
    return np.array([1.0 - float(((onehot - cube[:, :, a]) ** 2).sum()) / tss
                     for a in range(cube.shape[2])])


def _attach_head(model: FastPLS, classifier: str, yv, C: int, lda_ridge, candidate_knn_k,
                 candidate_tau, candidate_alpha, candidate_top_m, Xm=None) -> None:
    """.attach_lda_classifier (R/main.R:1098-1290): the LDA or candidate-kNN
    head trained on the model's resident training scores, or on the
    projection of ``Xm`` (the training rows) when given."""
    ctx = model.ctx
    inf = model.info()
    xp = ctypes.byref(Xm) if Xm is not None else None
    if classifier == "lda":
        ncomp_eff = sorted(set(int(min(max(c, 1), inf.kmax)) for c in model["ncomp"]))
        ks, kp = int_array(ncomp_eff)
        ctx.check(ctx.lib.fpls_lda_train_prefix(ctx.h, model.handle, xp, ctypes.byref(yv), C,
                                                kp, ks.size, float(lda_ridge)))
        model["classification_rule"] = "lda_hip"
        model["lda"] = {"ncomp": ncomp_eff, "ridge": float(lda_ridge)}
    else:
        ctx.check(ctx.lib.fpls_cknn_fit(ctx.h, model.handle, xp, ctypes.byref(yv), C,
                                        int(candidate_knn_k), float(candidate_tau),
                                        float(candidate_alpha), int(candidate_top_m)))
        model["classification_rule"] = "candidate_knn_hip"


def predict(model: FastPLS, newdata, Ytest=None, proj: bool = False, top: int = 1,
            top5: bool = False, raw_scores: bool = False, tie_log: bool = False) -> Dict:
    """predict.fastPLS (R/main.R:2577-2762) for a HIP model.

    Classification results follow .class_topk_to_labels: ``Ypred`` (labels,
    w x L), ``Ypred_index`` (1-based), and for top > 1 ``Ypred_top`` (ranked
    class indices, w x top x L), ``Ypred_top_score`` and ``Ypred_top_label``.
    The argmax head without Ytest / raw_scores ranks on the device straight
    from the compact factors (fpls_class_predict_topk); otherwise the score
    cube is formed and ranked (fpls_class_topk_scores).

    ``tie_log`` (LDA and candidate-kNN heads): ``tie_margin`` holds the
    device's per-row margins -- LDA w x L (winning minus runner-up
    discriminant value), candidate kNN w x 2 x L (rank-1 minus rank-2
    reranked score; the candidate-set boundary gap of the base scores) --
    by which a label is judged a near-tie."""
    ctx = model.ctx
    top = resolve_top_k(top, top5)
    if model.get("opls") is not None:
        newdata = opls_apply_filter(newdata, **model["opls"], device=ctx.device,
                                    out="device")
    elif model.get("kernel_pls") is not None:
        newdata = _kernel_test_matrix(model["kernel_pls"], newdata, ctx.device)
    keep: list = []
    X = as_mat(newdata, keep)
    w = X.rows
    inf = model.info()
    L = inf.n_ncomp
    lev = model.get("lev")
    rule = model.get("classification_rule", "argmax")
    res: Dict = {}
    if proj:
        T = np.empty((w, inf.kmax), dtype=np.float64, order="F")
        ctx.check(ctx.lib.fpls_project(ctx.h, model.handle, ctypes.byref(X), inf.kmax,
                                       T.ctypes.data))
        res["Ttest"] = T
    if model.get("classification") and rule == "lda_hip":
        C = len(lev)
        labels = np.empty((w, L), dtype=np.int32, order="F")
        want = raw_scores or top > 1
        sc = np.empty((w, C, L), dtype=np.float64, order="F") if want else None
        if tie_log:
            mg = np.empty((w, L), dtype=np.float64, order="F")
            ctx.check(ctx.lib.fpls_lda_predict_margins(ctx.h, model.handle, ctypes.byref(X),
                                                       labels.ctypes.data,
                                                       sc.ctypes.data if want else None,
                                                       mg.ctypes.data))
            res["tie_margin"] = mg
        else:
            ctx.check(ctx.lib.fpls_lda_predict(ctx.h, model.handle, ctypes.byref(X),
                                               labels.ctypes.data,
                                               sc.ctypes.data if want else None))
        res["Ypred_index"] = labels
        res["Ypred"] = _labels_from_codes(lev, labels)
        if raw_scores:
            res["LDA_scores"] = sc
        if top > 1:
            res.update(class_topk_from_score_cube(sc, lev, None, top, ctx.device))
        if Ytest is not None:
            res["Q2Y"] = _q2_classes(Ytest, lev, np.stack(
                [(res["Ypred_index"][:, a][:, None] == np.arange(1, C + 1)[None, :])
                 .astype(np.float64) for a in range(L)], axis=2))
        return res
    if model.get("classification") and rule == "candidate_knn_hip":
        k = max(1, min(int(top), len(lev)))
        ti = np.empty((w, k, L), dtype=np.int32, order="F")
        ts = np.empty((w, k, L), dtype=np.float64, order="F")
        if tie_log:
            mg = np.empty((w, 2, L), dtype=np.float64, order="F")
            ctx.check(ctx.lib.fpls_cknn_predict_margins(ctx.h, model.handle, ctypes.byref(X), k,
                                                        ti.ctypes.data, ts.ctypes.data,
                                                        mg.ctypes.data))
        else:
            ctx.check(ctx.lib.fpls_cknn_predict(ctx.h, model.handle, ctypes.byref(X), k,
                                                ti.ctypes.data, ts.ctypes.data))
        res.update(_topk_result(ti, ts, lev))
        if tie_log:
            res["tie_margin"] = mg
        if Ytest is not None:
            C = len(lev)
            res["Q2Y"] = _q2_classes(Ytest, lev, np.stack(
                [(res["Ypred_index"][:, a][:, None] == np.arange(1, C + 1)[None, :])
                 .astype(np.float64) for a in range(L)], axis=2))
        return res
    if (model.get("classification") and Ytest is None and not raw_scores
            and not model.get("gaussian_y")):
        res.update(class_predict_topk(model, newdata, None, top))
        return res
    gspec = model.get("gaussian_y")
    if (model.get("classification") and Ytest is None and not raw_scores and gspec
            and gspec.get("task") == "classification"):
        # nearest-code classes ranked on the device from the compact factors
        codes = np.asfortranarray(gspec["codes"], dtype=np.float64)
        C = codes.shape[0]
        k = max(1, min(int(top), C))
        ti = np.empty((w, k, L), dtype=np.int32, order="F")
        ts = np.empty((w, k, L), dtype=np.float64, order="F")
        ctx.check(ctx.lib.fpls_gaussian_y_predict_classes(ctx.h, model.handle, ctypes.byref(X),
                                                          codes.ctypes.data, C, int(top),
                                                          ti.ctypes.data, ts.ctypes.data))
        res.update(_topk_result(ti, ts, lev))
        return res
    m = inf.m
    yp = _host_f64(w, m, L)
    ctx.check(ctx.lib.fpls_predict(ctx.h, model.handle, ctypes.byref(X), yp.ctypes.data))
    if model.get("gaussian_y"):
        # .decode_gaussian_y_cube (R/main.R:2099-2118) after the prediction
        yp = decode_gaussian_y_cube(model["gaussian_y"], yp, ctx.device)
    if model.get("classification"):
        if Ytest is not None:
            res["Q2Y"] = _q2_classes(Ytest, lev, yp)
        if raw_scores:
            res["Yscore"] = yp
        res.update(class_topk_from_score_cube(yp, lev, None, top, ctx.device))
    else:
        res["Ypred"] = yp
        if Ytest is not None:
            yt = np.asarray(Ytest, dtype=np.float64).reshape(w, -1)
            q2 = []
            for a in range(L):
                press = float(((yt - yp[:, :, a]) ** 2).sum())
                tss = float(((yt - yt.mean(axis=0)) ** 2).sum())
                q2.append(1.0 - press / tss)
            res["Q2Y"] = np.array(q2)
    return res


def _host_f64(*shape: int) -> np.ndarray:
    """A Fortran-order FP64 host array for a large result copied from HBM:
    pinned (page-locked) when torch is loaded, so the copy runs at PCIe speed
    instead of through the runtime's pageable staging."""
    if int(np.prod(shape)) * 8 >= (4 << 20):
        try:
            import torch
            if torch.cuda.is_available():
                t = torch.empty(tuple(reversed(shape)), dtype=torch.float64, pin_memory=True)
                return t.numpy().T
        except (ImportError, RuntimeError):
            pass
    return np.empty(shape, dtype=np.float64, order="F")


def _device_f64(rows: int, cols: int, device: int):
    import torch
    return torch.empty((cols, rows), dtype=torch.float64, device=f"cuda:{device}").t()


def _f64_out(rows: int, cols: int, out: str, device: int):
    if out == "device":
        return _device_f64(rows, cols, device)
    return np.empty((rows, cols), dtype=np.float64, order="F")


def _mat_out(a, keep: list) -> _lib.Mat:
    if hasattr(a, "is_cuda"):
        return as_mat(a, keep)
    keep.append(a)
    return _lib.Mat(a.ctypes.data, _lib.F64, _lib.HOST, a.shape[0], a.shape[1],
                    max(a.shape[0], 1))


def _shape(x):
    return (int(x.shape[0]), int(x.shape[1]) if len(x.shape) > 1 else 1)


def opls_filter(X, Y, north: int = 1, scaling="centering", n_classes=None, device: int = 0,
                out: str = "host") -> Dict:
    """opls_filter_cpp (src/fastPLS.cpp:4377-4461) on the device.  Y: a dense
    response, or class labels (their one-hot, .supervised_response_matrix).
    ``out="device"`` leaves the filtered X in HBM (a column-major FP64 tensor)."""
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    n, p = Xm.rows, Xm.cols
    sc = SCALINGS[scaling] if isinstance(scaling, str) else int(scaling)
    Ym = yv = None
    C = 0
    if _is_factor(Y):
        _, codes, C = _encode_factor(Y, n_classes)
        yv = as_ivec(codes, keep)
    else:
        Ym = as_mat(Y, keep)
    Xo = _f64_out(n, p, out, ctx.device)
    Om = _mat_out(Xo, keep)
    mX = np.empty(p)
    vX = np.empty(p)
    Wo = np.zeros((p, max(int(north), 0)), order="F")
    Po = np.zeros((p, max(int(north), 0)), order="F")
    used = ctypes.c_int32(0)
    ctx.check(ctx.lib.fpls_opls_filter(ctx.h, ctypes.byref(Xm),
                                       ctypes.byref(Ym) if Ym is not None else None,
                                       ctypes.byref(yv) if yv is not None else None, C,
                                       int(north), sc, ctypes.byref(Om), mX.ctypes.data,
                                       vX.ctypes.data, Wo.ctypes.data if north > 0 else None,
                                       Po.ctypes.data if north > 0 else None,
                                       ctypes.byref(used)))
    k = used.value
    return {"X": Xo, "mX": mX, "vX": vX, "W_orth": np.asfortranarray(Wo[:, :k]),
            "P_orth": np.asfortranarray(Po[:, :k]), "north": k}


def opls_apply_filter(X, mX, vX, W_orth, P_orth, north=None, device: int = 0,
                      out: str = "host"):
    """opls_apply_filter_cpp (src/fastPLS.cpp:4463-4483) on the device."""
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    n, p = Xm.rows, Xm.cols
    mX = np.ascontiguousarray(mX, dtype=np.float64).reshape(-1)
    vX = np.ascontiguousarray(vX, dtype=np.float64).reshape(-1)
    Wo = np.asfortranarray(W_orth, dtype=np.float64).reshape(p, -1, order="F")
    Po = np.asfortranarray(P_orth, dtype=np.float64).reshape(p, -1, order="F")
    if mX.size != p or vX.size != p:
        raise ValueError("X columns must match stored OPLS preprocessing")
    if Wo.shape != Po.shape:
        raise ValueError("Invalid OPLS orthogonal filter dimensions")
    k = Wo.shape[1]
    Xo = _f64_out(n, p, out, ctx.device)
    Om = _mat_out(Xo, keep)
    ctx.check(ctx.lib.fpls_opls_apply(ctx.h, ctypes.byref(Xm), mX.ctypes.data, vX.ctypes.data,
                                      Wo.ctypes.data if k else None,
                                      Po.ctypes.data if k else None, k, ctypes.byref(Om)))
    return Xo


KERNELS = {"linear": 1, "rbf": 2, "poly": 3}


def kernel_matrix(X1, X2, kernel="rbf", gamma=None, degree: int = 3, coef0: float = 1.0,
                  device: int = 0, out: str = "host"):
    """kernel_matrix_cpp (src/fastPLS.cpp:4314-4341) on the device (FP64)."""
    ctx = _lib.context(device)
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    keep: list = []
    A = as_mat(X1, keep)
    B = as_mat(X2, keep)
    if gamma is None:
        gamma = 1.0 / max(1, A.cols)
    K = _f64_out(A.rows, B.rows, out, ctx.device)
    Km = _mat_out(K, keep)
    ctx.check(ctx.lib.fpls_kernel_matrix(ctx.h, ctypes.byref(A), ctypes.byref(B), kid,
                                         float(gamma), int(degree), float(coef0),
                                         ctypes.byref(Km)))
    return K


def center_kernel_train(K, device: int = 0, out: str = "host", in_place: bool = False) -> Dict:
    """center_kernel_train_cpp (src/fastPLS.cpp:4344-4357) on the device."""
    ctx = _lib.context(device)
    keep: list = []
    Km = as_mat(K, keep)
    Ko = K if in_place else _f64_out(Km.rows, Km.cols, out, ctx.device)
    Om = _mat_out(Ko, keep) if not in_place else Km
    cm = np.empty(Km.cols)
    gm = ctypes.c_double(0.0)
    ctx.check(ctx.lib.fpls_center_kernel_train(ctx.h, ctypes.byref(Km), ctypes.byref(Om),
                                               cm.ctypes.data, ctypes.byref(gm)))
    return {"K": Ko, "col_means": cm, "grand_mean": gm.value}


def center_kernel_test(Ktest, col_means, grand_mean: float, device: int = 0,
                       out: str = "host"):
    """center_kernel_test_cpp (src/fastPLS.cpp:4360-4374) on the device."""
    ctx = _lib.context(device)
    keep: list = []
    Km = as_mat(Ktest, keep)
    cm = np.ascontiguousarray(col_means, dtype=np.float64).reshape(-1)
    if cm.size != Km.cols:
        raise ValueError("Ktest columns must match the training kernel size")
    Ko = _f64_out(Km.rows, Km.cols, out, ctx.device)
    Om = _mat_out(Ko, keep)
    ctx.check(ctx.lib.fpls_center_kernel_test(ctx.h, ctypes.byref(Km), cm.ctypes.data,
                                              float(grand_mean), ctypes.byref(Om)))
    return Ko


def _opls_fit(Xtrain, Ytrain, north: int, scaling, inner_kw: Dict) -> FastPLS:
    """.opls_fit (R/main.R:3047-3091): the filter on the device, SIMPLS on the
    filtered X (HBM-resident) with scaling "none"."""
    filt = opls_filter(Xtrain, Ytrain, north, scaling, inner_kw.get("n_classes"),
                       inner_kw["device"], out="device")
    model = pls(filt["X"], Ytrain, None, None, method="simpls", scaling="none", **inner_kw)
    model["opls"] = {k: filt[k] for k in ("mX", "vX", "W_orth", "P_orth")}
    model["north"] = filt["north"]
    model["pls_method"] = "simpls"
    model["opls_engine"] = "hip"
    model["class"] = ["fastPLSOpls", "fastPLS"]
    return model


def _kernel_pls_fit(Xtrain, Ytrain, scaling, kernel, gamma, degree, coef0,
                    inner_kw: Dict) -> FastPLS:
    """.kernel_pls_fit (R/main.R:2826-2905): the linear kernel fits SIMPLS on X
    directly; rbf / poly preprocess X, build and centre the training kernel in
    HBM and fit SIMPLS on it with scaling "none"."""
    if kernel not in KERNELS:
        raise ValueError("kernel must be 'linear', 'rbf' or 'poly'")
    dev = inner_kw["device"]
    if kernel == "linear":
        model = pls(Xtrain, Ytrain, None, None, method="simpls", scaling=scaling, **inner_kw)
        model["kernel"] = "linear"
        model["kernel_engine"] = "hip_direct"
        model["kernel_linear_direct"] = True
        model["pls_method"] = "simpls"
        return model
    prep = opls_filter(Xtrain, np.zeros((_shape(Xtrain)[0], 1)), 0, scaling, device=dev,
                       out="device")
    p = prep["X"].shape[1]
    if gamma is None:
        gamma = 1.0 / max(1, p)
    gamma = float(np.atleast_1d(gamma)[0])
    if not np.isfinite(gamma) or gamma <= 0:
        raise ValueError("gamma must be a finite positive number")
    K = kernel_matrix(prep["X"], prep["X"], kernel, gamma, degree, coef0, dev, out="device")
    kc = center_kernel_train(K, dev, in_place=True)
    model = pls(kc["K"], Ytrain, None, None, method="simpls", scaling="none", **inner_kw)
    model["kernel_pls"] = {"Xref": prep["X"], "mX": prep["mX"], "vX": prep["vX"],
                           "kernel": kernel, "kernel_id": KERNELS[kernel], "gamma": gamma,
                           "degree": int(degree), "coef0": float(coef0),
                           "col_means": kc["col_means"], "grand_mean": kc["grand_mean"]}
    model["kernel"] = kernel
    model["kernel_engine"] = "hip"
    model["pls_method"] = "simpls"
    model["class"] = ["fastPLSKernel", "fastPLS"]
    return model


def _kernel_test_matrix(spec: Dict, newdata, device: int):
    """predict.fastPLSKernel (R/main.R:3031-3045): the centred test kernel."""
    p = spec["Xref"].shape[1]
    Xn = opls_apply_filter(newdata, spec["mX"], spec["vX"], np.zeros((p, 0)), np.zeros((p, 0)),
                           device=device, out="device")
    Kt = kernel_matrix(Xn, spec["Xref"], spec["kernel"], spec["gamma"], spec["degree"],
                       spec["coef0"], device, out="device")
    return center_kernel_test(Kt, spec["col_means"], spec["grand_mean"], device, out="device")


def model_from_fields(fields: Dict, device: int = 0) -> FastPLS:
    """A device model from the reference's model-list fields (R, Q, C_latent,
    W_latent, B, mX, vX, mY, ncomp, pls_method; optional classification /
    lev), e.g. a model fitted by the reference's CPU path: predict() then runs
    on the device (fpls_model_import; compact_prediction_weights,
    src/fastPLS.cpp:3737-3836)."""
    method = fields.get("pls_method", "simpls")
    if method not in ("simpls", "simpls_fast", "plssvd"):
        raise ValueError("pls_method must be 'simpls' or 'plssvd'")
    for name in ("R", "ncomp", "mX", "vX", "mY"):
        if fields.get(name) is None:
            raise ValueError(f"model field {name} is required")
    keep: list = []

    def arr(name, shape=None, required=False):
        v = fields.get(name)
        if v is None:
            if required:
                raise ValueError(f"model field {name} is required")
            return None
        a = np.asfortranarray(np.asarray(v, dtype=np.float64))
        if shape is not None:
            a = np.asfortranarray(a.reshape(shape, order="F"))
        keep.append(a)
        return a
    R = arr("R", required=True)
    if R.ndim != 2:
        raise ValueError("Model `R` must be a p x kmax matrix")
    p, kmax = R.shape
    nc = np.ascontiguousarray(np.atleast_1d(fields["ncomp"]), dtype=np.int32)
    keep.append(nc)
    L = nc.size
    mY = arr("mY", (-1,), True)
    m = mY.size
    f = _lib.ModelFields()
    f.method = _lib.METHOD_PLSSVD if method == "plssvd" else _lib.METHOD_SIMPLS
    f.kmax, f.p, f.m, f.n_ncomp = kmax, p, m, L
    f.ncomp = nc.ctypes.data
    f.R = R.ctypes.data
    for name, shape in (("Q", (m, kmax)), ("C_latent", (kmax, kmax, L)),
                        ("W_latent", (kmax, m, L)), ("B", (p, m, L))):
        v = fields.get(name)
        if v is not None and np.size(v):
            a = np.asarray(v, dtype=np.float64)
            if name in ("W_latent", "C_latent") and a.ndim == 3 and a.shape[2] > L:
                a = a[:, :, :L]          # dim[3] >= length(ncomp) in the reference
            setattr(f, name, arr_from(a, shape, keep).ctypes.data)
    mX = arr("mX", (-1,), True)
    vX = arr("vX", (-1,), True)
    if mX.size != p or vX.size != p:
        raise ValueError("Model `mX` and `vX` must have one value per predictor (nrow(R))")
    f.mX = mX.ctypes.data
    f.vX = vX.ctypes.data
    f.mY = mY.ctypes.data
    ctx = _lib.context(device)
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.fpls_model_import(ctx.h, ctypes.byref(f), ctypes.byref(h)))
    model = FastPLS(ctx, h)
    model["classification"] = bool(fields.get("classification", False))
    model["lev"] = fields.get("lev")
    if model["classification"] and model["lev"] is None:
        model["lev"] = list(range(1, m + 1))
    model["classification_rule"] = "argmax"
    model["pls_method"] = "plssvd" if method == "plssvd" else "simpls"
    model["predict_backend"] = "hip"
    model["ncomp"] = nc.astype(np.int64)
    model["m"], model["p"] = m, p
    inf = model.info()
    model["B_stored"] = bool(inf.b_stored)
    model["compact_prediction"] = not bool(inf.b_stored)
    _import_heads(model, fields)
    return model


def _head_models(sub: Dict, ks) -> list:
    """The per-k entries of a classifier sub-list's `models`, named by k (the
    reference's named list) or given in the order of `ncomp`."""
    models = sub["models"]
    if isinstance(models, dict):
        out = []
        for k in ks:
            e = models.get(k, models.get(str(k)))
            if e is None:
                raise ValueError(f"No fitted classifier for ncomp={k}")
            out.append(e)
        return out
    models = list(models)
    if len(models) != len(ks):
        raise ValueError("classifier models must have one entry per ncomp")
    return models


def _import_heads(model: FastPLS, fields: Dict) -> None:
    """The `lda` (R/main.R:1244-1250) and `candidate_knn` (R/main.R:861-872)
    sub-lists of a reference model onto the device model
    (fpls_model_import_lda / fpls_model_import_cknn); the classification rule
    follows the model's own, else the head that is present (LDA first)."""
    ctx = model.ctx
    lev = fields.get("lev")
    keep: list = []

    def f64(v, shape=None):
        a = np.asfortranarray(np.asarray(v, dtype=np.float64))
        if shape is not None:
            a = np.asfortranarray(a.reshape(shape, order="F"))
        keep.append(a)
        return a

    def packed(arrs):
        a = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64)
                                                 .ravel(order="F") for x in arrs]))
        keep.append(a)
        return a.ctypes.data
    rule = fields.get("classification_rule")
    heads: list = []
    lda = fields.get("lda")
    if lda is not None and lda.get("models") is not None:
        ks = [int(k) for k in np.atleast_1d(lda["ncomp"])]
        ms = _head_models(lda, ks)
        lin = [f64(e["linear"]) for e in ms]
        C = len(lev) if lev is not None else lin[0].shape[0]
        for e, k, li in zip(ms, ks, lin):
            if li.shape != (C, k) or np.size(e["constants"]) != C:
                raise ValueError(f"LDA model for ncomp={k} has inconsistent linear / constants")
        lf = _lib.LdaFields()
        lf.n_classes, lf.n_models = C, len(ks)
        kv, lf.ncomp = int_array(ks)
        keep.append(kv)
        lf.linear = packed(lin)
        lf.constants = packed([e["constants"] for e in ms])
        if all(e.get("means") is not None for e in ms):
            lf.means = packed([f64(e["means"], (C, k)) for e, k in zip(ms, ks)])
        if all(e.get("inv_cov") is not None for e in ms):
            lf.inv_cov = packed([f64(e["inv_cov"], (k, k)) for e, k in zip(ms, ks)])
        if ms[0].get("priors") is not None:
            lf.priors = packed([ms[0]["priors"]])
        if all(e.get("ridge") is not None for e in ms):
            lf.ridge = packed([[float(np.atleast_1d(e["ridge"])[0])] for e in ms])
        ctx.check(ctx.lib.fpls_model_import_lda(ctx.h, model.handle, ctypes.byref(lf)))
        model["lda"] = {"ncomp": ks, "ridge": lda.get("ridge")}
        heads.append("lda_hip")
    ck = fields.get("candidate_knn")
    if ck is not None and ck.get("models") is not None:
        ks = [int(k) for k in np.atleast_1d(ck["ncomp"])]
        ms = _head_models(ck, ks)
        cen = [f64(e["centroids"]) for e in ms]
        C = len(lev) if lev is not None else cen[0].shape[0]
        Tt = f64(ck["Ttrain"])
        if Tt.ndim != 2:
            raise ValueError("candidate_knn Ttrain must be a matrix")
        yc = np.ascontiguousarray(np.asarray(ck["y_codes"]).ravel(), dtype=np.int32)
        keep.append(yc)
        if yc.size != Tt.shape[0]:
            raise ValueError("candidate_knn y_codes must hold one label per Ttrain row")
        par = ck.get("parameters", {})
        cf = _lib.CknnFields()
        cf.n_classes, cf.n_models = C, len(ks)
        kv, cf.ncomp = int_array(ks)
        keep.append(kv)
        for c_, k in zip(cen, ks):
            if c_.shape != (C, k):
                raise ValueError(f"candidate-kNN centroids for ncomp={k} must be C x {k}")
        cf.centroids = packed(cen)
        cf.n_train, cf.ttrain_cols = Tt.shape
        cf.Ttrain = Tt.ctypes.data
        cf.y_codes = yc.ctypes.data
        cf.knn_k = int(np.atleast_1d(par.get("knn_k", 3))[0])
        cf.top_m = int(np.atleast_1d(par.get("top_m", 20))[0])
        cf.tau = float(np.atleast_1d(par.get("tau", 0.2))[0])
        cf.alpha = float(np.atleast_1d(par.get("alpha", 0.5))[0])
        ctx.check(ctx.lib.fpls_model_import_cknn(ctx.h, model.handle, ctypes.byref(cf)))
        heads.append("candidate_knn_hip")
    if rule is not None and str(rule).startswith("lda") and "lda_hip" in heads:
        rule = "lda_hip"
    elif rule is not None and "candidate_knn" in str(rule) and "candidate_knn_hip" in heads:
        rule = "candidate_knn_hip"
    elif rule is None and heads:
        rule = heads[0]
    if rule in ("lda_hip", "candidate_knn_hip"):
        model["classification"] = True
        model["classification_rule"] = rule
        if model.get("lev") is None:
            C = model.info().n_classes
            model["lev"] = list(range(1, C + 1))


def arr_from(a: np.ndarray, shape, keep: list) -> np.ndarray:
    a = np.asfortranarray(a.reshape(shape, order="F"))
    keep.append(a)
    return a


def gaussian_y_default_dim(Xtrain, gaussian_y_dim=None) -> int:
    """.gaussian_y_default_dim (R/main.R:1945-1954): min(ncol(X), 100)."""
    if gaussian_y_dim is None:
        p = Xtrain.shape[1] if hasattr(Xtrain, "shape") and len(Xtrain.shape) > 1 else 1
        return max(1, min(int(p), 100))
    d = int(np.atleast_1d(gaussian_y_dim)[0])
    if d < 1:
        raise ValueError("gaussian_y_dim must be a positive integer")
    return d


def prepare_gaussian_y(Ytrain, Xtrain, gaussian_y: bool = False, gaussian_y_dim=None,
                       seed: int = 1, device: int = 0) -> Dict:
    """.prepare_gaussian_y (R/main.R:1997-2052): the compressed response Z
    (host FP64, or a device FP64 column-major tensor for a device Y) and the
    decoding spec.  The projection / codebook come from this library's
    seeded Gaussian stream (fpls_gaussian_y_encode / fpls_gaussian_y_codes)."""
    classification = _is_factor(Ytrain)
    if not gaussian_y:
        return {"Ytrain": Ytrain, "classification": classification, "lev": None,
                "gaussian": None}
    dim = gaussian_y_default_dim(Xtrain, gaussian_y_dim)
    ctx = _lib.context(device)
    keep: list = []
    seed32 = int(seed) & 0xFFFFFFFF
    if classification:
        lev, codes_i, C = _encode_factor(Ytrain, None)
        yv = as_ivec(codes_i, keep)
        n = yv.n
        codes = np.empty((C, dim), order="F")
        Z, Zm = _out_matrix(n, dim, Ytrain, keep)
        ctx.check(ctx.lib.fpls_gaussian_y_codes(ctx.h, ctypes.byref(yv), C, dim, seed32,
                                                ctypes.byref(Zm), codes.ctypes.data))
        return {"Ytrain": Z, "classification": True, "lev": lev,
                "gaussian": {"enabled": True, "task": "classification", "dim": dim,
                             "original_dim": C, "codes": codes}}
    Ym = as_mat(Ytrain, keep)
    n, m = Ym.rows, Ym.cols
    center = np.empty(m)
    decoder = np.empty((dim, m), order="F")
    Z, Zm = _out_matrix(n, dim, Ytrain, keep)
    ctx.check(ctx.lib.fpls_gaussian_y_encode(ctx.h, ctypes.byref(Ym), dim, seed32,
                                             ctypes.byref(Zm), center.ctypes.data,
                                             decoder.ctypes.data, None))
    return {"Ytrain": Z, "classification": False, "lev": None,
            "gaussian": {"enabled": True, "task": "regression", "dim": dim, "original_dim": m,
                         "y_center": center, "decoder": decoder}}


def _out_matrix(rows: int, cols: int, like, keep: list):
    """An FP64 output matrix on the side of ``like``: a device column-major
    tensor for a torch input, else a host Fortran array."""
    if _lib._is_torch(like):
        import torch
        t = torch.empty((cols, rows), dtype=torch.float64, device=like.device).t()
        keep.append(t)
        return t, _lib.Mat(t.data_ptr(), _lib.F64, _lib.DEVICE, rows, cols, max(rows, 1))
    a = np.empty((rows, cols), order="F")
    keep.append(a)
    return a, _lib.Mat(a.ctypes.data, _lib.F64, _lib.HOST, rows, cols, max(rows, 1))


def _attach_gaussian_y(model: "FastPLS", yprep: Dict) -> None:
    """.attach_gaussian_y (R/main.R:2054-2075)."""
    spec = yprep["gaussian"]
    model["gaussian_y"] = spec
    model["gaussian_y_task"] = spec["task"]
    model["gaussian_y_dim"] = spec["dim"]
    model["gaussian_y_original_dim"] = spec["original_dim"]
    model["classification"] = yprep["classification"]
    model["lev"] = yprep["lev"]
    model["classification_rule"] = "argmax"


def decode_gaussian_y_matrix(spec: Dict, Z, device: int = 0, out=None) -> np.ndarray:
    """.decode_gaussian_y_matrix (R/main.R:2077-2096): regression predictions
    (n x m) or class scores 2 Z codes^T - ||code||^2 (n x C).  ``out``: an
    n x m column-major FP64 host array to decode into (regression)."""
    ctx = _lib.context(device)
    keep: list = []
    Zm = as_mat(np.asarray(Z, dtype=np.float64) if not _lib._is_torch(Z) else Z, keep)
    n = Zm.rows
    if spec["task"] == "classification":
        codes = np.asfortranarray(spec["codes"], dtype=np.float64)
        C = codes.shape[0]
        labels = np.empty(n, dtype=np.int32)
        scores = np.empty((n, C), order="F")
        ctx.check(ctx.lib.fpls_gaussian_y_classes(ctx.h, ctypes.byref(Zm), codes.ctypes.data, C,
                                                  labels.ctypes.data, scores.ctypes.data))
        return scores
    dec = np.asfortranarray(spec["decoder"], dtype=np.float64)
    cen = np.ascontiguousarray(spec["y_center"], dtype=np.float64)
    m = dec.shape[1]
    Y = _host_f64(n, m) if out is None else out
    if Y.shape != (n, m) or not Y.flags.f_contiguous:
        raise ValueError("out must be a column-major n x m array")
    Ym = _lib.Mat(Y.ctypes.data, _lib.F64, _lib.HOST, n, m, max(n, 1))
    ctx.check(ctx.lib.fpls_gaussian_y_decode(ctx.h, ctypes.byref(Zm), dec.ctypes.data, m,
                                             cen.ctypes.data, ctypes.byref(Ym)))
    return Y


def decode_gaussian_y_cube(spec: Dict, Zcube, device: int = 0) -> np.ndarray:
    """.decode_gaussian_y_cube (R/main.R:2099-2118): slice by slice."""
    Zc = np.asarray(Zcube, dtype=np.float64)
    if Zc.ndim == 2:
        return decode_gaussian_y_matrix(spec, Zc, device)
    if spec["task"] == "classification":
        out = [decode_gaussian_y_matrix(spec, np.asfortranarray(Zc[:, :, i]), device)
               for i in range(Zc.shape[2])]
        return np.stack(out, axis=2)
    n, L = Zc.shape[0], Zc.shape[2]
    cube = _host_f64(n, np.asarray(spec["decoder"]).shape[1], L)
    for i in range(L):
        decode_gaussian_y_matrix(spec, np.asfortranarray(Zc[:, :, i]), device, out=cube[:, :, i])
    return cube


def fastsvd(A, ncomp: int, method: str = "rsvd", rsvd_oversample: int = 10,
            rsvd_power: int = 1, seed: int = 1, omega=None, backend: str = "hip",
            device: int = 0) -> Dict:
    """fastsvd (R/main.R:4818-4886): truncated SVD on the device."""
    _check_backend(backend)
    ctx = _lib.context(device)
    keep: list = []
    Am = as_mat(A, keep)
    rows, cols = Am.rows, Am.cols
    k = int(min(max(int(ncomp), 1), min(rows, cols)))
    opts = _lib.default_opts(oversample=int(rsvd_oversample), power=int(rsvd_power),
                             seed=int(seed) & 0xFFFFFFFF,
                             svd_method=_lib.SVD_EXACT if method == "exact" else _lib.SVD_RSVD)
    U = np.empty(rows * k)
    s = np.empty(k)
    Vt = np.empty(k * cols)
    kout = ctypes.c_int32()
    om = as_mat(np.asarray(omega, dtype=np.float64), keep) if omega is not None else None
    ctx.check(ctx.lib.fpls_rsvd(ctx.h, ctypes.byref(Am), k, ctypes.byref(opts),
                                ctypes.byref(om) if om is not None else None, U.ctypes.data,
                                s.ctypes.data, Vt.ctypes.data, ctypes.byref(kout)))
    kk = kout.value
    U = U[: rows * kk].reshape((rows, kk), order="F")
    Vt = Vt[: kk * cols].reshape((kk, cols), order="F")
    return {"d": s[:kk], "u": U, "v": Vt.T.copy()}


CLASSIFIERS = {"argmax": 0, "lda": 1, "cknn": 2}


def _cv_metric(Ytrue: np.ndarray, Ypred: np.ndarray):
    """.cv_metric_from_matrix (R/main.R:3699-3709): q2 for one response, rmsd otherwise."""
    if Ytrue.shape[1] == 1:
        press = float(((Ypred[:, 0] - Ytrue[:, 0]) ** 2).sum())
        tss = float(((Ytrue[:, 0] - Ytrue[:, 0].mean()) ** 2).sum())
        return "q2", (1.0 - press / tss) if tss > 0 else float("nan")
    return "rmsd", float(np.sqrt(np.mean((Ypred - Ytrue) ** 2)))


def _cv_best_index(names, values) -> int:
    """.cv_best_index (R/main.R:6316-6330): max, or min for loss metrics."""
    v = np.asarray(values, dtype=float)
    fin = np.isfinite(v)
    if not fin.any():
        return 0
    if any(nm in ("rmsd", "rmse", "mae", "mse") for nm, f in zip(names, fin) if f):
        return int(np.argmin(np.where(fin, v, np.inf)))
    return int(np.argmax(np.where(fin, v, -np.inf)))


def cv_batched(Xdata, Ydata, masks, ncomp=2, seeds=None, scaling="centering", seed=1,
               precision="fp32", n_classes=None, return_class_pred=True, return_scores=False,
               device=0) -> Dict:
    """All fits of a cross-validation (or of every inner loop of a double
    cross-validation) as one batched SIMPLS-fast call (``fpls_cv_batched``).

    ``masks`` is n x F (0 unused, 1 train, 2 test row of fit f); ``seeds`` the
    F seeds of the fits (default ``seed + f``, the CV driver's rule,
    src/fastPLS.cpp:6906).  Returns ``status`` (F), and for the test rows of
    each fit ``class_pred`` (n x L x F), ``n_correct`` (L x F) and, when asked
    or for regression, ``Ypred`` (n x m x L x F)."""
    ctx = _lib.context(device)
    keep: list = []
    X = as_mat(Xdata, keep)
    n = X.rows
    M = np.asfortranarray(np.asarray(masks, dtype=np.int8))
    if M.ndim != 2 or M.shape[0] != n:
        raise ValueError("masks must be n x n_fits")
    keep.append(M)
    F = M.shape[1]
    classification = _is_factor(Ydata)
    yv = Ym = None
    C = 0
    lev = None
    if classification:
        lev, codes, C = _encode_factor(Ydata, n_classes)
        yv = as_ivec(codes, keep)
        m = C
    else:
        Ym = as_mat(Ydata, keep)
        m = Ym.cols
    nc_arr, nc_ptr = int_array(np.atleast_1d(ncomp))
    L = nc_arr.size
    opts = _lib.default_opts(scaling=SCALINGS[scaling], seed=int(seed) & 0xFFFFFFFF,
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32)
    sd = None
    if seeds is not None:
        sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64).ravel() & 0xFFFFFFFF,
                                  dtype=np.uint32)
        if sd.size != F:
            raise ValueError("seeds must have one value per fit")
        keep.append(sd)
    status = np.zeros(F, dtype=np.int32)
    cls = np.zeros((n, L, F), dtype=np.int32, order="F") if classification and return_class_pred else None
    ncor = np.zeros((L, F), dtype=np.int32, order="F") if classification else None
    want_scores = (not classification) or return_scores
    yp = np.zeros((n, m, L, F), dtype=np.float64, order="F") if want_scores else None
    ctx.check(ctx.lib.fpls_cv_batched(
        ctx.h, ctypes.byref(X), ctypes.byref(Ym) if Ym is not None else None,
        ctypes.byref(yv) if yv is not None else None, C, M.ctypes.data, F,
        sd.ctypes.data if sd is not None else None, nc_ptr, L, ctypes.byref(opts),
        status.ctypes.data, cls.ctypes.data if cls is not None else None,
        ncor.ctypes.data if ncor is not None else None,
        yp.ctypes.data if yp is not None else None))
    return {"status": status, "class_pred": cls, "n_correct": ncor, "Ypred": yp,
            "ncomp": nc_arr.astype(np.int64), "levels": lev, "classification": classification,
            "n_test": (M == 2).sum(axis=0)}


def _batched_eligible(Xdata, method, classifier, scaling, ngroups) -> bool:
    if method != "simpls" or classifier != "argmax" or scaling == "autoscaling":
        return False
    if not 1 <= ngroups <= 128:
        return False
    if type(Xdata).__module__.startswith("torch"):
        import torch
        ok = Xdata.dtype == torch.float32
        shape = Xdata.shape
    else:
        ok = np.asarray(Xdata).dtype == np.float32
        shape = np.asarray(Xdata).shape
    return ok and len(shape) == 2 and shape[1] > 5


_LOOCV_KEYS = ("loocv", "loo", "leave-one-out", "leave_one_out", "leave one out")


def compiled_kfold(kfold, constrain) -> int:
    """.compiled_cv_kfold_arg (R/main.R:4003-4009, .cv_kfold_int 3982-3996):
    -1 (leave one group out) for "loocv" or kfold >= number of groups, else
    max(2, kfold)."""
    n_groups = int(np.unique(np.asarray(constrain).ravel()).size)
    if isinstance(kfold, str):
        if kfold.strip().lower() in _LOOCV_KEYS:
            return -1
        raise ValueError("cross-validation: kfold must be a single integer or 'loocv'.")
    if np.ndim(kfold) != 0 or kfold is None or not np.isfinite(float(kfold)):
        raise ValueError("cross-validation: kfold must be a finite integer or 'loocv'.")
    k = int(kfold)
    if k < 0:
        return -1
    if k >= n_groups:
        return -1
    return max(2, k)


def cv_folds(constrain, kfold=-1, labels=None, n_classes: int = 0, seed: int = 1):
    """Row -> fold map (1-based) and fold count of pls_cv_predict_compiled
    (src/fastPLS.cpp:6666-6775) through ``fpls_cv_folds``: leave one group out
    for kfold < 0 or kfold >= #groups, otherwise a permutation of the groups
    (stratified by the class of each group's first row when ``labels`` are
    given) from R's uniforms after set.seed(seed)."""
    lib = _lib.load()
    cons = np.ascontiguousarray(np.asarray(constrain).ravel().astype(np.int32))
    n = cons.size
    fold = np.zeros(n, dtype=np.int32)
    nf = ctypes.c_int32()
    lab = None
    if labels is not None:
        lab = np.ascontiguousarray(np.asarray(labels).ravel().astype(np.int32))
        if lab.size != n:
            raise ValueError("labels must have one value per sample")
    rc = lib.fpls_cv_folds(cons.ctypes.data, n, int(kfold), lab.ctypes.data if lab is not None
                           else None, int(n_classes), int(seed) & 0xFFFFFFFF, fold.ctypes.data,
                           ctypes.byref(nf))
    if rc != _lib.FPLS_OK:
        raise _lib.FplsError(rc, "fold assignment failed (labels must be 1..n_classes)")
    return fold, nf.value


def pls_cv_predict(Xdata, Ydata, constrain, ncomp=2, scaling="centering", method="simpls",
                   classifier="argmax", lda_ridge=1e-8, candidate_knn_k=10, candidate_tau=0.2,
                   candidate_alpha=0.75, candidate_top_m=20, seed=1, rsvd_oversample=10,
                   rsvd_power=1, precision="fp32", return_scores=False, device=0,
                   engine="auto", north=1, kfold=-1) -> Dict:
    """pls_cv_predict_compiled (src/fastPLS.cpp:6615-7151) over the folds of
    ``cv_folds``: ``kfold`` = -1 (or >= the number of groups) leaves one
    distinct value of ``constrain`` out at a time, otherwise ``kfold`` folds
    of a seeded permutation of the groups.  ``method``: "plssvd",
    "simpls", "opls" (the filter with ``north`` components per training fold,
    then SIMPLS with scaling "none") or "kernelpls" (the direct SIMPLS core,
    as the compiled CV fits it).

    ``engine``: "folds" fits the folds one after another (every method and
    classifier); "batched" runs all folds as one batched call (SIMPLS, argmax
    classifier, FP32 X, scaling centering / none); "auto" takes "batched"
    when it applies.  Both run on the GPU and give the same predictions up to
    rounding."""
    cons0 = np.asarray(constrain).ravel()
    ngroups = int(np.unique(cons0).size)
    kfold = compiled_kfold(kfold, cons0) if isinstance(kfold, str) else int(kfold)
    if engine not in ("auto", "batched", "folds"):
        raise ValueError("engine must be 'auto', 'batched' or 'folds'")
    nfold_est = ngroups if (kfold < 0 or kfold >= ngroups) else max(2, kfold)
    use_batched = engine == "batched" or (
        engine == "auto" and _batched_eligible(Xdata, method, classifier, scaling, nfold_est))
    if use_batched:
        if _is_factor(Ydata):
            _, codes0, C0 = _encode_factor(Ydata)
            fold1, nfold = cv_folds(cons0, kfold, codes0, C0, seed)
        else:
            fold1, nfold = cv_folds(cons0, kfold, None, 0, seed)
        return _pls_cv_predict_batched(Xdata, Ydata, fold1 - 1, nfold, ncomp, scaling, seed,
                                       precision, return_scores, device, method, classifier)
    ctx = _lib.context(device)
    keep: list = []
    X = as_mat(Xdata, keep)
    n = X.rows
    classification = _is_factor(Ydata)
    lev = None
    yv = Ym = None
    C = 0
    if classification:
        lev, codes, C = _encode_factor(Ydata)
        yv = as_ivec(codes, keep)
        m = C
    else:
        Ym = as_mat(Ydata, keep)
        m = Ym.cols
    cons = np.ascontiguousarray(np.asarray(constrain).ravel().astype(np.int32))
    if cons.size != n:
        raise ValueError("constrain must have one value for each sample")
    keep.append(cons)
    nc_arr, nc_ptr = int_array(np.atleast_1d(ncomp))
    L = nc_arr.size
    opts = _lib.default_opts(scaling=SCALINGS[scaling], seed=int(seed) & 0xFFFFFFFF,
                             oversample=int(rsvd_oversample), power=int(rsvd_power),
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32)
    opts.opls_north = int(north)
    ngroups = int(np.unique(cons).size)
    fold = np.empty(n, dtype=np.int32)
    status = np.zeros(max(ngroups, kfold, 2), dtype=np.int32)
    nf = ctypes.c_int32()
    cls = np.zeros((n, L), dtype=np.int32, order="F") if classification else None
    want_scores = (not classification) or return_scores
    yp = np.zeros((n, m, L), dtype=np.float64, order="F") if want_scores else None
    methods = {"plssvd": _lib.METHOD_PLSSVD, "simpls": _lib.METHOD_SIMPLS,
               "simpls_fast": _lib.METHOD_SIMPLS, "opls": _lib.METHOD_OPLS,
               "kernelpls": _lib.METHOD_KERNELPLS}
    if method not in methods:
        raise ValueError("method must be 'plssvd', 'simpls', 'opls' or 'kernelpls'")
    meth = methods[method]
    ctx.check(ctx.lib.fpls_cv_predict(
        ctx.h, ctypes.byref(X), ctypes.byref(Ym) if Ym is not None else None,
        ctypes.byref(yv) if yv is not None else None, C, cons.ctypes.data, nc_ptr, L, kfold, meth,
        ctypes.byref(opts), CLASSIFIERS[classifier], float(lda_ridge), int(candidate_knn_k),
        float(candidate_tau), float(candidate_alpha), int(candidate_top_m), fold.ctypes.data,
        status.ctypes.data, status.size, ctypes.byref(nf),
        cls.ctypes.data if cls is not None else None,
        yp.ctypes.data if yp is not None else None))
    out = {"fold": fold, "status": status[: nf.value], "ncomp": nc_arr.astype(np.int64),
           "method": method, "backend": "hip", "classifier": classifier,
           "classification": classification, "levels": lev}
    names, values, preds = [], [], []
    if classification:
        out["class_pred"] = cls
        truth = np.asarray(Ydata).ravel()
        levs = np.asarray(lev, dtype=object)
        for i in range(L):
            ok = (cls[:, i] >= 1) & (cls[:, i] <= len(lev))
            pred = np.full(n, None, dtype=object)
            pred[ok] = levs[cls[ok, i] - 1]
            preds.append(pred)
            names.append("accuracy")
            values.append(float(np.mean(pred == truth)))
        if yp is not None:
            out["Ypred"] = yp
    else:
        out["Ypred"] = yp
        yt = np.asarray(Ydata, dtype=np.float64).reshape(n, -1)
        for i in range(L):
            preds.append(yp[:, :, i])
            nm, v = _cv_metric(yt, yp[:, :, i])
            names.append(nm)
            values.append(v)
    out["pred"] = preds
    out["metrics"] = {"ncomp_index": list(range(1, L + 1)), "metric_name": names,
                      "metric_value": values}
    return out


def _cv_summary(out, Ydata, cls, yp, classification, lev, n, L):
    names, values, preds = [], [], []
    if classification:
        out["class_pred"] = cls
        truth = np.asarray(Ydata).ravel()
        levs = np.asarray(lev, dtype=object)
        for i in range(L):
            ok = (cls[:, i] >= 1) & (cls[:, i] <= len(lev))
            pred = np.full(n, None, dtype=object)
            pred[ok] = levs[cls[ok, i] - 1]
            preds.append(pred)
            names.append("accuracy")
            values.append(float(np.mean(pred == truth)))
        if yp is not None:
            out["Ypred"] = yp
    else:
        out["Ypred"] = yp
        yt = np.asarray(Ydata, dtype=np.float64).reshape(n, -1)
        for i in range(L):
            preds.append(yp[:, :, i])
            nm, v = _cv_metric(yt, yp[:, :, i])
            names.append(nm)
            values.append(v)
    out["pred"] = preds
    out["metrics"] = {"ncomp_index": list(range(1, L + 1)), "metric_name": names,
                      "metric_value": values}
    return out


def _pls_cv_predict_batched(Xdata, Ydata, fold, F, ncomp, scaling, seed, precision,
                            return_scores, device, method, classifier):
    """The folds (0-based row -> fold, F folds) as one batched call; fold f
    fits with seed + f like the fold-by-fold driver."""
    if method != "simpls" or classifier != "argmax":
        raise ValueError("the batched engine runs SIMPLS with the argmax classifier")
    fold = np.asarray(fold, dtype=np.int32)
    n = fold.size
    masks = np.ones((n, F), dtype=np.int8, order="F")
    masks[np.arange(n), fold] = 2
    res = cv_batched(Xdata, Ydata, masks, ncomp, None, scaling, seed, precision,
                     return_class_pred=True, return_scores=return_scores, device=device)
    classification = res["classification"]
    L = res["ncomp"].size
    cls = yp = None
    if classification:
        cls = np.asfortranarray(res["class_pred"].sum(axis=2, dtype=np.int64).astype(np.int32))
    if res["Ypred"] is not None:
        yp = np.asfortranarray(res["Ypred"].sum(axis=3))
    out = {"fold": fold + 1, "status": res["status"], "ncomp": res["ncomp"],
           "method": method, "backend": "hip", "classifier": classifier,
           "classification": classification, "levels": res["levels"], "engine": "batched"}
    return _cv_summary(out, Ydata, cls, yp, classification, res["levels"], n, L)


def pls_double_cv(Xdata, Ydata, ncomp=2, outer_fold=None, inner_fold=None, scaling="centering",
                  seed=1, run=1, precision="fp32", device=0) -> Dict:
    """One run of pls.double.cv (R/main.R:6586-6894) for SIMPLS with the argmax
    classifier, on given folds (the reference draws them from R's RNG).

    ``outer_fold`` (n) assigns rows to outer folds; ``inner_fold`` (n x
    n_outer) assigns, for outer fold j, the rows outside it to inner folds
    (entries of rows inside outer fold j are ignored).  Every inner fit of
    every outer fold runs in one batched call, seeded
    seed + 1000 run + f + inner fold (R/main.R:6708 and src/fastPLS.cpp:6906;
    f counts outer folds from 1); the optimal ncomp of each outer fold
    maximises inner accuracy (first maximum, R/main.R:6527-6536); the outer
    fits (seed + 2000 run + f, R/main.R:6738) run as a second batched call.
    Missing values in Xdata stop it, as pls.double.cv does (R/main.R:6610-6612)."""
    _check_finite(Xdata)
    keep_shape = np.asarray(Xdata).shape if not type(Xdata).__module__.startswith("torch") \
        else tuple(Xdata.shape)
    n = keep_shape[0]
    if not _is_factor(Ydata):
        raise ValueError("pls_double_cv here covers classification (factor Ydata)")
    lev, codes, C = _encode_factor(Ydata)
    ofold = np.asarray(outer_fold).ravel()
    outer_vals = np.unique(ofold)
    J = outer_vals.size
    inner = np.asarray(inner_fold).reshape(n, J)
    nc = np.atleast_1d(np.asarray(ncomp, dtype=np.int64))
    cols, seeds, owner = [], [], []
    for j, ov in enumerate(outer_vals):
        tr = ofold != ov
        ivals = np.unique(inner[tr, j])
        for fi, iv in enumerate(ivals):
            mk = np.zeros(n, dtype=np.int8)
            mk[tr] = 1
            mk[tr & (inner[:, j] == iv)] = 2
            cols.append(mk)
            seeds.append(int(seed) + 1000 * int(run) + (j + 1) + fi)
            owner.append(j)
    masks = np.stack(cols, axis=1)
    owner = np.asarray(owner)
    inner_res = cv_batched(Xdata, np.asarray(Ydata), masks, nc, seeds, scaling, seed, precision,
                           return_class_pred=False, device=device)
    ncor = inner_res["n_correct"].astype(np.int64)  # L x F_inner
    best = np.zeros(J, dtype=np.int64)
    acc_inner = np.zeros((J, nc.size))
    for j in range(J):
        sel = owner == j
        ntr = (ofold != outer_vals[j]).sum()
        acc = ncor[:, sel].sum(axis=1) / max(int(ntr), 1)
        acc_inner[j] = acc
        best[j] = nc[int(np.argmax(acc))]
    # outer fits: train on the outer-train rows, predict the outer fold, one
    # slice per distinct optimal ncomp
    omask = np.ones((n, J), dtype=np.int8, order="F")
    for j, ov in enumerate(outer_vals):
        omask[ofold == ov, j] = 2
    grid = np.unique(best)
    oseeds = [int(seed) + 2000 * int(run) + (j + 1) for j in range(J)]
    outer_res = cv_batched(Xdata, np.asarray(Ydata), omask, grid, oseeds, scaling, seed,
                           precision, return_class_pred=True, device=device)
    pred_codes = np.zeros(n, dtype=np.int32)
    for j in range(J):
        s = int(np.searchsorted(grid, best[j]))
        rows = omask[:, j] == 2
        pred_codes[rows] = outer_res["class_pred"][rows, s, j]
    levs = np.asarray(lev, dtype=object)
    pred = np.full(n, None, dtype=object)
    ok = pred_codes >= 1
    pred[ok] = levs[pred_codes[ok] - 1]
    truth = np.asarray(Ydata).ravel()
    return {"Ypred": pred, "optim_comp": best, "inner_accuracy": acc_inner,
            "inner_status": inner_res["status"], "outer_status": outer_res["status"],
            "metric_name": "accuracy", "metric_value": float(np.mean(pred == truth)),
            "fold": ofold, "backend": "hip", "method": "simpls"}


def _check_finite(Xdata):
    """The reference's CV entry points stop on missing values
    (R/main.R:6426-6428); non-finite X is refused on the host."""
    if type(Xdata).__module__.startswith("torch"):
        import torch
        ok = bool(torch.isfinite(Xdata).all().item())
    else:
        ok = bool(np.isfinite(np.asarray(Xdata, dtype=np.float64)).all())
    if not ok:
        raise ValueError("Missing values are present")


def optim_pls_cv(Xdata, Ydata, ncomp=2, constrain=None, scaling="centering", method="simpls",
                 backend="hip", seed=1, kfold=10, classifier="argmax", **kw) -> Dict:
    """optim.pls.cv (R/main.R:6399-6537): ``kfold`` folds (default 10; "loocv"
    or kfold >= #groups leaves one group out) over the groups of
    ``constrain`` (default one group per sample), stratified by class for a
    factor response, drawn after set.seed(seed) as the compiled CV does."""
    _check_backend(backend)
    _check_finite(Xdata)
    n = np.asarray(Xdata).shape[0] if not type(Xdata).__module__.startswith("torch") else Xdata.shape[0]
    if constrain is None:
        constrain = np.arange(1, n + 1)
    res = pls_cv_predict(Xdata, Ydata, constrain, ncomp, scaling, method, classifier, seed=seed,
                         kfold=compiled_kfold(kfold, constrain), **kw)
    vals = res["metrics"]["metric_value"]
    bi = _cv_best_index(res["metrics"]["metric_name"], vals)
    res["optim_comp"] = int(res["ncomp"][bi])
    res["best_index"] = bi + 1
    res["best_metric_name"] = res["metrics"]["metric_name"][bi]
    res["best_metric_value"] = vals[bi]
    res["Q2Y"] = vals
    res["R2Y"] = vals
    res["Ypred_optim"] = res["pred"][bi]
    return res


def pls_single_cv(Xdata, Ydata, constrain=None, ncomp=2, kfold=10, scaling="centering",
                  method="simpls", backend="hip", seed=1, classifier="argmax", **kw) -> Dict:
    """pls.single.cv (R/main.R:4316-4435): cross-validated predictions for the
    given ncomp values over ``kfold`` folds (default 10) of the groups of
    ``constrain``."""
    return optim_pls_cv(Xdata, Ydata, ncomp, constrain, scaling, method, backend, seed, kfold,
                        classifier, **kw)


def truncated_rsvd_crossprod(X, Y, k: int, rsvd_oversample: int = 10, rsvd_power: int = 1,
                             seed: int = 1, exact: bool = False, omega=None,
                             fetch: bool = True, precision: str = "fp32",
                             route: str = "matrix_free", device: int = 0) -> Dict:
    """truncated_rsvd_crossprod_double (src/fastPLS.cpp:550-623): rSVD of X^T Y
    without forming it.  ``omega`` (cols(Y) x (k + oversample)) replaces the
    Gaussian sketch (a learned sketch from ``sketch_train``).  ``route =
    "materialise"`` forms S = X^T Y once (one pass over X) and runs the same
    range finder on S; ``precision`` selects the FP32 or split-BF16x3 passes
    over an FP32 X."""
    if route not in ("matrix_free", "materialise"):
        raise ValueError("route must be 'matrix_free' or 'materialise'")
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    Ym = as_mat(Y, keep)
    p, m = Xm.cols, Ym.cols
    kk = int(min(max(int(k), 1), min(p, m)))
    opts = _lib.default_opts(oversample=int(rsvd_oversample), power=int(rsvd_power),
                             seed=int(seed) & 0xFFFFFFFF,
                             svd_method=_lib.SVD_EXACT if exact else _lib.SVD_RSVD,
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32,
                             xprod_route=1 if route == "materialise" else 0)
    Om = as_mat(np.asarray(omega, dtype=np.float64), keep) if omega is not None else None
    U = np.empty(p * kk) if fetch else None
    s = np.empty(kk)
    Vt = np.empty(kk * m) if fetch else None
    kout = ctypes.c_int32()
    ctx.check(ctx.lib.fpls_rsvd_xprod(ctx.h, ctypes.byref(Xm), ctypes.byref(Ym), kk,
                                      ctypes.byref(opts),
                                      ctypes.byref(Om) if Om is not None else None,
                                      U.ctypes.data if U is not None else None, s.ctypes.data,
                                      Vt.ctypes.data if Vt is not None else None,
                                      ctypes.byref(kout)))
    if not fetch:
        return {"s": s[: kout.value]}
    r = kout.value
    return {"U": U[: p * r].reshape((p, r), order="F"), "s": s[:r],
            "Vt": Vt[: r * m].reshape((r, m), order="F")}


def sketch_train(X, Y, masks, k: int, l: int, iters: int = 200, lr: float = 0.02, seed: int = 1,
                 precision: str = "fp32", device: int = 0) -> Dict:
    """Learn a sketch Omega (cols(Y) x l) for the rSVD range finder of Xc^T Yc
    from the cross-covariances of the row-fold draws in ``masks`` (n x D,
    nonzero = row in the draw); ``fpls_sketch_train``.  Returns ``omega`` and
    the per-step mean loss (squared Frobenius sine of the top-k subspace)."""
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    Ym = as_mat(Y, keep)
    M = np.asfortranarray(np.asarray(masks, dtype=np.int8))
    keep.append(M)
    m = Ym.cols
    opts = _lib.default_opts(seed=int(seed) & 0xFFFFFFFF,
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32)
    om = np.zeros((m, l), order="F")
    trace = np.zeros(int(iters) + 1)
    ctx.check(ctx.lib.fpls_sketch_train(ctx.h, ctypes.byref(Xm), ctypes.byref(Ym), M.ctypes.data,
                                        M.shape[1], int(k), int(l), int(iters), float(lr),
                                        ctypes.byref(opts), om.ctypes.data, trace.ctypes.data))
    return {"omega": om, "loss": trace}


def sketch_eval(X, Y, masks, k: int, omega=None, l: int = None, power: int = 0, seed: int = 1,
                precision: str = "fp32", device: int = 0) -> np.ndarray:
    """sin(theta) per draw between the range finder's subspace of S_d (``omega``
    or the Gaussian stream of ``seed``, ``power`` iterations) and the exact
    top-k left singular subspace of S_d; ``fpls_sketch_eval``."""
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    Ym = as_mat(Y, keep)
    M = np.asfortranarray(np.asarray(masks, dtype=np.int8))
    keep.append(M)
    om = None
    if omega is not None:
        om = np.asfortranarray(np.asarray(omega, dtype=np.float64))
        keep.append(om)
        l = om.shape[1]
    opts = _lib.default_opts(seed=int(seed) & 0xFFFFFFFF,
                             precision=_lib.PREC_BF16X3 if precision == "bf16x3" else _lib.PREC_FP32)
    out = np.zeros(M.shape[1])
    ctx.check(ctx.lib.fpls_sketch_eval(ctx.h, ctypes.byref(Xm), ctypes.byref(Ym), M.ctypes.data,
                                       M.shape[1], int(k), om.ctypes.data if om is not None else None,
                                       int(l), int(power), ctypes.byref(opts), out.ctypes.data))
    return out


def class_sums(X, y, n_classes: int, device: int = 0) -> Dict:
    """Per-class column sums (p x C), column sums (p) and class counts from
    1-based int32 labels, streamed from HBM (R/main.R:563-588)."""
    ctx = _lib.context(device)
    keep: list = []
    Xm = as_mat(X, keep)
    yv = as_ivec(y, keep)
    p = Xm.cols
    S = np.empty((p, n_classes), order="F")
    cs = np.empty(p)
    cnt = np.empty(n_classes, dtype=np.int64)
    ctx.check(ctx.lib.fpls_class_sums(ctx.h, ctypes.byref(Xm), ctypes.byref(yv), int(n_classes),
                                      S.ctypes.data, cs.ctypes.data, cnt.ctypes.data))
    return {"class_sums": S, "col_sums": cs, "counts": cnt}


def candidate_knn_predict(Ttest_norm, Ttrain_norm, y_codes, centroids, candidate_bias=None,
                          top_k=1, top_m=20, knn_k=10, tau=0.2, alpha=0.75, device=0) -> Dict:
    """candidate_knn_predict_cpp (src/fastPLS.cpp:1627-1648) on the device:
    array-level candidate kNN on normalised scores."""
    ctx = _lib.context(device)
    keep: list = []
    Te = as_mat(Ttest_norm, keep)
    Tr = as_mat(Ttrain_norm, keep)
    Ce = as_mat(centroids, keep)
    yv = as_ivec(y_codes, keep)
    C = Ce.rows
    top = max(1, min(int(top_k), C))
    w = Te.rows
    ti = np.empty((w, top), dtype=np.int32, order="F")
    ts = np.empty((w, top), dtype=np.float64, order="F")
    mg = np.empty((w, 2), dtype=np.float64, order="F")
    bias = None
    if candidate_bias is not None and len(candidate_bias):
        bias = np.ascontiguousarray(np.asarray(candidate_bias, dtype=np.float64))
        keep.append(bias)
    ctx.check(ctx.lib.fpls_cknn_predict_scores_margins(
        ctx.h, ctypes.byref(Te), ctypes.byref(Tr), ctypes.byref(yv), ctypes.byref(Ce),
        bias.ctypes.data if bias is not None else None, top, int(top_m), int(knn_k), float(tau),
        float(alpha), ti.ctypes.data, ts.ctypes.data, mg.ctypes.data))
    return {"top_index": ti, "top_score": ts, "tie_margin": mg}
