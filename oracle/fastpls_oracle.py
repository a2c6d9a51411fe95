"""FP64 CPU restatement of the fastPLS hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the HIP library.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only to compute expected results (or to time the CPU baseline).  The
product path (``fastpls_amd``) never imports it.

Every function cites the reference file:line it restates (paths relative to
the fastPLS 0.3 source tree).  The reference is an R/Rcpp/Armadillo package;
Armadillo forwards its dense factorisations to LAPACK, and so does numpy:

* ``arma::svd_econ``  -> dgesdd   == ``numpy.linalg.svd(full_matrices=False)``
* ``arma::qr_econ``   -> dgeqrf + dorgqr == ``numpy.linalg.qr(mode="reduced")``
* ``arma::eig_sym``   -> dsyevd   == ``numpy.linalg.eigh``
* ``arma::solve(.., likely_sympd)`` -> dposv, falling back to dgesv
* ``arma::inv_sympd`` -> dpotrf + dpotri

The Gaussian sketch uses the libstdc++ ``std::mt19937`` +
``std::normal_distribution<double>`` stream through the small C++ helper in
``oracle/rng.cpp`` (built by ``oracle/Makefile``), so sketches are bit-equal to
the reference's on the same seed.

Pinned by: tests/test_oracle*.py (closed forms, brute force, library
routines, the reference's own tolerance tests restated, hand-worked cases) and
the golden fixtures under tests/golden/ (textbook SIMPLS / PLSSVD written
independently of this module, and the values R prints for its uniform
generator).  Functions without such a pin are marked "parity unpinned" in
their docstring.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np
import scipy.linalg as sla

_HERE = os.path.dirname(os.path.abspath(__file__))
_RNG_SO = os.path.join(_HERE, "_build", "liboracle_rng.so")
_rng_lib = None

DBL_MIN = np.finfo(np.float64).tiny
DBL_EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------------------
# RNG (src/svd_cpu_rsvd.cpp:10-22, src/fastPLS.cpp:99-115)
# ---------------------------------------------------------------------------

def build_rng(force: bool = False) -> str:
    """Compile oracle/rng.cpp into oracle/_build/liboracle_rng.so with g++."""
    if force or not os.path.exists(_RNG_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _RNG_SO


def _rng():
    global _rng_lib
    if _rng_lib is None:
        build_rng()
        lib = ctypes.CDLL(_RNG_SO)
        lib.fpls_oracle_gaussian.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint]
        lib.fpls_oracle_gaussian.restype = None
        lib.fpls_oracle_mt19937_raw.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint]
        lib.fpls_oracle_mt19937_raw.restype = None
        _rng_lib = lib
    return _rng_lib


def gaussian_matrix(n_rows: int, n_cols: int, seed: int) -> np.ndarray:
    """std::mt19937(seed) + normal_distribution<double>, column-major fill.

    src/svd_cpu_rsvd.cpp:10-22 (identical: src/fastPLS.cpp:99-115)."""
    out = np.empty((n_rows, n_cols), dtype=np.float64, order="F")
    if out.size:
        _rng().fpls_oracle_gaussian(out.ctypes.data, out.size, ctypes.c_uint(seed & 0xFFFFFFFF))
    return out


def mt19937_raw(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    _rng().fpls_oracle_mt19937_raw(out.ctypes.data, n, ctypes.c_uint(seed & 0xFFFFFFFF))
    return out


# ---------------------------------------------------------------------------
# small dense helpers (Armadillo -> LAPACK)
# ---------------------------------------------------------------------------

def svd_econ(A: np.ndarray):
    """arma::svd_econ (dgesdd).  Returns U, s, V (V, not V^T)."""
    if A.size == 0:
        return (np.zeros((A.shape[0], 0)), np.zeros(0), np.zeros((A.shape[1], 0)))
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    return U, s, Vt.T


def qr_econ(A: np.ndarray):
    """arma::qr_econ (dgeqrf + dorgqr)."""
    Q, R = np.linalg.qr(A, mode="reduced")
    return Q, R


def eig_sym_desc(A: np.ndarray):
    """arma::eig_sym (dsyevd) followed by sort_index(evals, "descend")."""
    w, V = np.linalg.eigh(A)
    order = np.argsort(-w, kind="stable")
    return w[order], V[:, order]


def solve_likely_sympd(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """arma::solve(X, A, B, solve_opts::likely_sympd): Cholesky, else LU."""
    try:
        c = sla.cho_factor(A, lower=False, check_finite=True)
        return sla.cho_solve(c, B)
    except (np.linalg.LinAlgError, sla.LinAlgError, ValueError):
        return np.linalg.solve(A, B)


def inv_sympd_or_fallback(A: np.ndarray) -> np.ndarray:
    """inv_sympd -> inv -> pinv chain of src/fastPLS.cpp:1789-1797."""
    try:
        c = sla.cho_factor(A, lower=False)
        return sla.cho_solve(c, np.eye(A.shape[0]))
    except (np.linalg.LinAlgError, sla.LinAlgError, ValueError):
        pass
    try:
        return np.linalg.inv(A)
    except np.linalg.LinAlgError:
        return np.linalg.pinv(A)


def variance(X: np.ndarray) -> np.ndarray:
    """Column sample sd, n-1 denominator, no zero guard (src/fastPLS.cpp:1570-1589).

    The reference accumulates with Welford's recurrence; the two-pass form here
    agrees to rounding."""
    n = X.shape[0]
    m = X.mean(axis=0)
    m2 = ((X - m) ** 2).sum(axis=0)
    return np.sqrt(m2 / max(n - 1, 1))


def transformy(y: Sequence[int]) -> np.ndarray:
    """Dense one-hot of 1-based labels (src/fastPLS.cpp:1593-1604)."""
    y = np.asarray(y, dtype=np.int64)
    nc = int(y.max())
    out = np.zeros((y.size, nc))
    out[np.arange(y.size), y - 1] = 1.0
    return out


def RQ(ydata: np.ndarray, ypred: np.ndarray) -> float:
    """1 - PRESS/TSS over all columns (src/fastPLS.cpp:1483-1500)."""
    my = ydata.mean(axis=0)
    press = float(((ydata - ypred) ** 2).sum())
    tss = float(((ydata - my) ** 2).sum())
    return 1.0 - press / tss


# ---------------------------------------------------------------------------
# truncated SVD (src/svd_iface.cpp:58-108, src/svd_cpu_exact.cpp, src/svd_cpu_rsvd.cpp)
# ---------------------------------------------------------------------------

def truncated_svd_exact(A: np.ndarray, k: int, left_only: bool = False):
    """src/svd_cpu_exact.cpp:7-38."""
    max_rank = min(A.shape)
    rank = min(max_rank, max(k, 1))
    U, s, V = svd_econ(A)
    return {"U": U[:, :rank], "s": s[:rank], "Vt": None if left_only else V[:, :rank].T}


def finalize_rsvd_from_sample(A: np.ndarray, Y: np.ndarray, k: int, left_only: bool = False):
    """src/svd_cpu_rsvd.cpp:26-93."""
    max_rank = min(A.shape)
    rank = min(max_rank, max(k, 1))
    Q, _ = qr_econ(Y)
    B = Q.T @ A
    if B.shape[0] <= B.shape[1]:
        try:
            evals, evecs = eig_sym_desc(B @ B.T)
            ok = True
        except np.linalg.LinAlgError:
            ok = False
        if ok:
            tol = DBL_EPS * max(B.shape) * (max(evals[0], 1.0) if evals.size else 1.0)
            usable = 0
            while usable < evals.size and evals[usable] > tol:
                usable += 1
            usable = min(usable, rank)
            if usable > 0:
                s = np.sqrt(np.clip(evals[:usable], 0.0, np.inf))
                Uhat = evecs[:, :usable]
                out = {"U": Q @ Uhat, "s": s, "Vt": None}
                if not left_only:
                    out["Vt"] = (Uhat.T @ B) / s[:, None]
                return out
    Uhat, s, V = svd_econ(B)
    U = Q @ Uhat
    return {"U": U[:, :rank], "s": s[:rank], "Vt": None if left_only else V[:, :rank].T}


def truncated_svd_rsvd(A: np.ndarray, k: int, oversample: int = 10, power: int = 1,
                       seed: int = 1, left_only: bool = False, omega: Optional[np.ndarray] = None):
    """src/svd_cpu_rsvd.cpp:95-124.  ``omega`` overrides the Gaussian sketch
    (the learned-sketch extension); None reproduces the reference."""
    max_rank = min(A.shape)
    target = min(max_rank, max(k, 1))
    l = min(max_rank, target + max(oversample, 0))
    if l >= max_rank:
        return truncated_svd_exact(A, target, left_only)
    Omega = gaussian_matrix(A.shape[1], l, seed) if omega is None else omega
    Y = A @ Omega
    q = max(power, 0)
    if q == 1:
        Y = A @ (A.T @ Y)
    else:
        for _ in range(q):
            Z = A.T @ Y
            Qz, _ = qr_econ(Z)
            Y = A @ Qz
    return finalize_rsvd_from_sample(A, Y, target, left_only)


def truncated_svd(A: np.ndarray, k: int, method: str = "rsvd", oversample: int = 10,
                  power: int = 1, seed: int = 1, left_only: bool = False,
                  use_full_svd: bool = False):
    """Dispatch of src/svd_iface.cpp:58-108 (CPU backend; irlba out of scope)."""
    if k < 1:
        raise ValueError("truncated_svd: k must be >= 1")
    if use_full_svd or min(A.shape) < 6 or method == "exact":
        return truncated_svd_exact(A, k, left_only)
    return truncated_svd_rsvd(A, k, oversample, power, seed, left_only)


def finalize_rsvd_from_q_b(Q: np.ndarray, B: np.ndarray, k: int, left_only: bool = False):
    """src/fastPLS.cpp:434-473 (svd_econ of B, not eig)."""
    max_rank = min(B.shape)
    rank = min(max_rank, max(k, 1))
    Uhat, s, V = svd_econ(B)
    actual = min(rank, Uhat.shape[1], s.size)
    return {"U": Q @ Uhat[:, :actual], "s": s[:actual],
            "Vt": None if left_only else V[:, :actual].T}


def truncated_rsvd_crossprod(X: np.ndarray, Y: np.ndarray, k: int, oversample: int = 10,
                             power: int = 1, seed: int = 1, left_only: bool = False,
                             use_full_svd: bool = False, omega: np.ndarray = None):
    """Matrix-free rSVD of S = X^T Y (src/fastPLS.cpp:550-623).  ``omega``
    (m x l) replaces the Gaussian sketch of ``seed`` (a learned sketch)."""
    p, m = X.shape[1], Y.shape[1]
    max_rank = min(p, m)
    target = min(max_rank, max(k, 1))
    l = min(max_rank, target + max(oversample, 0))
    if use_full_svd or l >= max_rank:
        return truncated_svd(X.T @ Y, target, "rsvd", oversample, power, seed, left_only, True)

    def a_times(M):
        return X.T @ (Y @ M)

    def at_times(M):
        return Y.T @ (X @ M)

    Omega = gaussian_matrix(m, l, seed) if omega is None else np.asarray(omega, np.float64)
    Ys = a_times(Omega)
    q = max(power, 0)
    if q == 1:
        Ys = a_times(at_times(Ys))
    else:
        for _ in range(q):
            Qz, _ = qr_econ(at_times(Ys))
            Ys = a_times(Qz)
    Q, _ = qr_econ(Ys)
    B = (X @ Q).T @ Y
    return finalize_rsvd_from_q_b(Q, B, target, left_only)


# ---------------------------------------------------------------------------
# preprocessing (src/fastPLS.cpp:2482-2495, 4819-4834)
# ---------------------------------------------------------------------------

def _preprocess(X: np.ndarray, Y: np.ndarray, scaling: int):
    X = np.array(X, dtype=np.float64, order="F", copy=True)
    Y = np.array(Y, dtype=np.float64, order="F", copy=True)
    p = X.shape[1]
    mX = np.zeros(p)
    if scaling < 3:
        mX = X.mean(axis=0)
        X -= mX
    vX = np.ones(p)
    if scaling == 2:
        vX = variance(X)
        X /= vX
    mY = Y.mean(axis=0)
    Y -= mY
    return X, Y, mX, vX, mY


def _store_B(p: int, m: int, n_slices: int, max_mb: float = 256.0) -> bool:
    """should_store_coefficients default policy (src/fastPLS.cpp:63-85)."""
    return p * m * max(n_slices, 1) * 8 / (1024.0 * 1024.0) <= max_mb


# ---------------------------------------------------------------------------
# PLSSVD (src/fastPLS.cpp:4780-4986)
# ---------------------------------------------------------------------------

def pls_model1(Xtrain, Ytrain, ncomp, scaling: int = 1, fit: bool = False,
               svd_method: str = "rsvd", oversample: int = 10, power: int = 1,
               seed: int = 1, xprod: bool = False) -> Dict:
    """pls_model1 (src/fastPLS.cpp:4780-4986), optimized branch (default).
    ``xprod``: pls_model1_rsvd_xprod_precision with precision 3 ("implicit64",
    src/fastPLS.cpp:5518-5690): the same preprocessing and latent
    coefficients, the SVD of S from the matrix-free range finder
    truncated_rsvd_crossprod_double (:550-623) without forming S."""
    Xtrain = np.asarray(Xtrain, dtype=np.float64)
    Ytrain = np.asarray(Ytrain, dtype=np.float64)
    if Ytrain.ndim == 1:
        Ytrain = Ytrain[:, None]
    n, p = Xtrain.shape
    m = Ytrain.shape[1]
    max_rank = min(n, p, m)
    ncomp = np.array([min(max(int(c), 1), max_rank) for c in np.atleast_1d(ncomp)], dtype=np.int64)
    max_ncomp_eff = min(int(ncomp.max()), max_rank)
    X, Y, mX, vX, mY = _preprocess(Xtrain, Ytrain, scaling)
    use_full = max_rank <= 32  # plssvd_use_small_exact_svd (src/fastPLS.cpp:318-324)
    if xprod and svd_method != "exact":
        svd = truncated_rsvd_crossprod(X, Y, max_ncomp_eff, oversample, power, seed, False,
                                       use_full)
    else:
        S = X.T @ Y
        svd = truncated_svd(S, max_ncomp_eff, "exact" if svd_method == "exact" else "rsvd",
                            oversample, power, seed, False, use_full)
    U, s, V = svd["U"], svd["s"], svd["Vt"].T
    L = ncomp.size
    C_latent = np.zeros((max_ncomp_eff, max_ncomp_eff, L))
    W_latent = np.zeros((max_ncomp_eff, m, L))
    max_ncomp_eff = min(max_ncomp_eff, U.shape[1], V.shape[1])
    U = U[:, :max_ncomp_eff]
    V = V[:, :max_ncomp_eff]
    T = X @ U
    G = T.T @ T
    store_B = _store_B(p, m, L)
    B = np.zeros((p, m, L)) if store_B else None
    Yfit = np.zeros((n, m, L)) if fit else None
    R2Y = np.zeros(L)
    for a in range(L):
        mc = min(int(ncomp[a]), max_ncomp_eff)
        D = np.diag(s[:mc])
        coeff = solve_likely_sympd(G[:mc, :mc], D)
        C_latent[:mc, :mc, a] = coeff
        W = coeff @ V[:, :mc].T
        W_latent[:mc, :, a] = W
        if store_B:
            B[:, :, a] = U[:, :mc] @ W
        if fit:
            t1 = T[:, :mc] @ W
            R2Y[a] = RQ(Y, t1)
            Yfit[:, :, a] = t1 + mY
    out = {"C_latent": C_latent, "W_latent": W_latent, "Q": V, "Ttrain": T, "R": U,
           "mX": mX, "vX": vX, "mY": mY, "p": p, "m": m, "ncomp": ncomp, "Yfit": Yfit,
           "R2Y": R2Y, "s": s[:max_ncomp_eff], "pls_method": "plssvd", "B_stored": store_B}
    if xprod and svd_method != "exact":
        out["xprod_precision"], out["xprod_mode"] = 3, "implicit"
    if store_B:
        out["B"] = B
    return out


# ---------------------------------------------------------------------------
# SIMPLS-fast (src/fastPLS.cpp:2454-2778, refresh 1068-1190, finalize 264-291)
# ---------------------------------------------------------------------------

def _refresh_left(S: np.ndarray, warm: Optional[np.ndarray], power_iters: int, seed: int):
    """SimplsFastRefreshWorkspace::refresh with k_block = 1 (src/fastPLS.cpp:1117-1189)."""
    Omega = gaussian_matrix(S.shape[0], 1, seed)
    if warm is not None:
        Omega[:, 0] = warm
    Y = Omega
    for _ in range(power_iters):
        Z = S.T @ Y
        Y = S @ Z
    Q, _ = qr_econ(Y)
    Bsmall = Q.T @ S
    # finalize_left_block_from_bsmall: 1 row < threshold 4 -> svd_econ "left"
    Uhat, _, _ = svd_econ(Bsmall)
    return (Q @ Uhat)[:, 0]


def pls_model2_fast(Xtrain, Ytrain, ncomp, scaling: int = 1, fit: bool = False,
                    svd_method: str = "rsvd", oversample: int = 10, power: int = 1,
                    seed: int = 1, inc_iters: int = 2, return_ttrain: bool = False) -> Dict:
    """pls_model2_fast (src/fastPLS.cpp:2454-2778) under the options that
    .with_fastpls_fast_options sets (R/main.R:55-80): center_t=0, reorth_v=0,
    inc_iters=2, deflation cache on, refresh block 1, adaptive policy off.

    Quirk kept: Yfit_cur only accumulates t q^T at the requested ncomp values
    (src/fastPLS.cpp:2651-2661)."""
    Xtrain = np.asarray(Xtrain, dtype=np.float64)
    Ytrain = np.asarray(Ytrain, dtype=np.float64)
    if Ytrain.ndim == 1:
        Ytrain = Ytrain[:, None]
    n, p = Xtrain.shape
    m = Ytrain.shape[1]
    ncomp = np.array([max(int(c), 1) for c in np.atleast_1d(ncomp)], dtype=np.int64)
    K = int(ncomp.max())
    L = ncomp.size
    X, Y, mX, vX, mY = _preprocess(Xtrain, Ytrain, scaling)
    S = X.T @ Y
    RR = np.zeros((p, K))
    QQ = np.zeros((m, K))
    VV = np.zeros((p, K))
    TT = np.zeros((n, K)) if return_ttrain else None
    store_B = _store_B(p, m, L)
    B = np.zeros((p, m, L)) if store_B else None
    Bcur = np.zeros((p, m))
    Yfit = np.zeros((n, m, L)) if fit else None
    Yfit_cur = np.zeros((n, m))
    R2Y = np.zeros(L)
    use_xtx = (K >= 20) and (p <= n) and (n >= 8 * p) and (p <= 512)
    XtX = X.T @ X if use_xtx else None
    state = {"i_out": 0, "rr_prev": None}

    def append(rr: np.ndarray, a: int) -> bool:
        rr = rr.copy()
        tt = None
        if use_xtx:
            pp = XtX @ rr
            tn2 = float(rr @ pp)
            if not math.isfinite(tn2) or tn2 <= 0.0:
                return False
            tn = math.sqrt(tn2)
            rr /= tn
            pp /= tn
            qq = S_orig.T @ rr
            if fit or return_ttrain:
                tt = X @ rr
        else:
            tt = X @ rr
            tn = float(np.linalg.norm(tt))
            if not math.isfinite(tn) or tn <= 0.0:
                return False
            tt = tt / tn
            rr /= tn
            pp = X.T @ tt
            qq = Y.T @ tt
        state["rr_prev"] = rr
        vv = pp.copy()
        if a > 0:
            Vp = VV[:, :a]
            vv -= Vp @ (Vp.T @ pp)
        vn = float(np.linalg.norm(vv))
        if not math.isfinite(vn) or vn <= 0.0:
            return False
        vv /= vn
        vS = vv @ S
        S[:] -= np.outer(vv, vS)
        RR[:, a] = rr
        QQ[:, a] = qq
        VV[:, a] = vv
        if return_ttrain and tt is not None:
            TT[:, a] = tt
        Bcur[:] += np.outer(rr, qq)
        while state["i_out"] < L and a == ncomp[state["i_out"]] - 1:
            io = state["i_out"]
            if store_B:
                B[:, :, io] = Bcur
            if fit:
                Yfit_cur[:] += np.outer(tt, qq)
                R2Y[io] = RQ(Y, Yfit_cur)
                Yfit[:, :, io] = Yfit_cur + mY
            state["i_out"] += 1
        return True

    S_orig = S.copy()  # Sxy_cache (only read on the X^T X branch)
    can_incremental = (p > 5) and (m > 1)
    a = 0
    while a < K:
        if can_incremental:
            rr = _refresh_left(S, state["rr_prev"], inc_iters, seed + a)
        else:
            sv = truncated_svd(S, 1, "rsvd", oversample, power, seed + a, True, False)
            if sv["U"].shape[1] < 1:
                break
            rr = sv["U"][:, 0]
        if not append(rr, a):
            break
        a += 1
    out = {"P": None, "Q": QQ, "Ttrain": TT, "R": RR, "V": VV, "mX": mX, "vX": vX, "mY": mY,
           "p": p, "m": m, "ncomp": ncomp, "Yfit": Yfit, "R2Y": R2Y, "n_components": a,
           "pls_method": "simpls", "B_stored": store_B}
    if store_B:
        out["B"] = B
    return out


# ---------------------------------------------------------------------------
# prediction (src/fastPLS.cpp:3211-3445, R/main.R:965-1041)
# ---------------------------------------------------------------------------

def pls_predict(model: Dict, Xtest) -> np.ndarray:
    """Ypred cube (w x m x |ncomp|): (Xtest - mX)/vX times B (or its latent
    factors), plus mY (src/fastPLS.cpp:3211-3445)."""
    X = (np.asarray(Xtest, dtype=np.float64) - model["mX"]) / model["vX"]
    ncomp = model["ncomp"]
    m = model["m"]
    out = np.zeros((X.shape[0], m, ncomp.size))
    for a in range(ncomp.size):
        if model.get("B") is not None:
            out[:, :, a] = X @ model["B"][:, :, a]
        elif model["pls_method"] == "simpls":
            mc = int(ncomp[a])
            out[:, :, a] = (X @ model["R"][:, :mc]) @ model["Q"][:, :mc].T
        else:
            mc = min(int(ncomp[a]), model["R"].shape[1])
            out[:, :, a] = (X @ model["R"][:, :mc]) @ model["W_latent"][:mc, :, a]
        out[:, :, a] += model["mY"]
    return out


def attach_latent_projection_cache(model: Dict, ncomp: Optional[int] = None) -> Dict:
    """R_predict = R / vX (row-wise), R_offset = mX R_predict (R/main.R:1019-1041)."""
    R = model["R"]
    k = min(int(ncomp if ncomp is not None else max(model["ncomp"])), R.shape[1])
    R = R[:, :k].copy()
    scale = np.asarray(model["vX"], dtype=np.float64).copy()
    scale[~np.isfinite(scale) | (scale == 0)] = 1.0
    R = R / scale[:, None]
    model["R_predict"] = R
    model["R_offset"] = np.asarray(model["mX"]) @ R
    return model


def latent_scores(model: Dict, X, k: int) -> np.ndarray:
    """T = X R_predict - offset (R/main.R:965-1017)."""
    if "R_predict" not in model:
        attach_latent_projection_cache(model)
    Rp = model["R_predict"][:, :k]
    return np.asarray(X, dtype=np.float64) @ Rp - model["R_offset"][:k]


# ---------------------------------------------------------------------------
# LDA (src/fastPLS.cpp:1718-1957, 2100-2147)
# ---------------------------------------------------------------------------

def lda_labels_from_scores(scores: np.ndarray, constants: np.ndarray) -> np.ndarray:
    """argmax of scores + constants, strict '>' so the lowest class wins ties,
    1-based (src/fastPLS.cpp:1718-1734)."""
    v = scores + constants
    return (np.argmax(v, axis=1) + 1).astype(np.int32)  # argmax returns the first max


def lda_train_prefix(T: np.ndarray, y, n_classes: int, ncomp: Sequence[int],
                     ridge: float = 1e-8) -> List[Dict]:
    """lda_train_prefix_cpp (src/fastPLS.cpp:1817-1927)."""
    T = np.asarray(T, dtype=np.float64)
    y = np.asarray(y, dtype=np.int64)
    n = T.shape[0]
    if n == 0 or T.shape[1] == 0:
        raise ValueError("lda_train_prefix_cpp requires a non-empty score matrix")
    if y.size != n:
        raise ValueError("lda_train_prefix_cpp requires one class label per training row")
    if n_classes < 2:
        raise ValueError("lda_train_prefix_cpp requires at least two classes")
    kmax = int(max(ncomp))
    if kmax < 1 or kmax > T.shape[1]:
        raise ValueError("lda_train_prefix_cpp component counts must be in 1..ncol(Ttrain)")
    Tk = T[:, :kmax]
    if np.any((y < 1) | (y > n_classes)):
        raise ValueError("lda_train_prefix_cpp labels must be encoded as 1..n_classes")
    counts = np.bincount(y - 1, minlength=n_classes).astype(np.float64)
    if np.any(counts <= 0):
        raise ValueError("lda_train_prefix_cpp received an empty class")
    means = np.zeros((n_classes, kmax))
    np.add.at(means, y - 1, Tk)
    means /= counts[:, None]
    pooled_full = Tk.T @ Tk
    pooled_full = np.triu(pooled_full) + np.triu(pooled_full, 1).T  # symmatu
    pooled_full -= (means.T * counts) @ means
    df = max(1.0, float(n) - float(n_classes))
    pooled_full /= df
    models = []
    for kk in ncomp:
        kk = int(kk)
        if kk < 1 or kk > kmax:
            raise ValueError("lda_train_prefix_cpp component counts must be in 1..max(ncomp)")
        pooled = pooled_full[:kk, :kk].copy()
        means_k = means[:, :kk]
        ridge_scale = np.trace(pooled) / max(1, kk)
        if not math.isfinite(ridge_scale) or ridge_scale <= 0.0:
            ridge_scale = 1.0
        lam = ridge
        if not math.isfinite(lam) or lam < 0.0:
            lam = 1e-8
        lam *= ridge_scale
        pooled[np.diag_indices(kk)] += lam
        inv_cov = inv_sympd_or_fallback(pooled)
        linear = means_k @ inv_cov
        priors = np.maximum(counts / n, DBL_MIN)
        constants = -0.5 * np.einsum("ck,ck->c", means_k, linear) + np.log(priors)
        models.append({"ncomp": kk, "means": means_k, "inv_cov": inv_cov, "linear": linear,
                       "constants": constants, "priors": counts / n, "ridge": lam,
                       "counts": counts})
    return models


def lda_project_train_prefix(X, R, offset, y, n_classes, ncomp, ridge=1e-8):
    """lda_project_train_prefix_cpp (src/fastPLS.cpp:1930-1957)."""
    T = np.asarray(X, dtype=np.float64) @ R
    if offset is not None and len(offset) >= R.shape[1]:
        T = T - np.asarray(offset)[: R.shape[1]]
    return lda_train_prefix(T, y, n_classes, ncomp, ridge)


def lda_project_predict_labels(X, R, offset, lda: Dict, return_scores: bool = False):
    """lda_project_predict_labels_cpp (src/fastPLS.cpp:2100-2147): direct form
    X (R lin^T) when n p C < 0.5 n k (p + C), else latent form."""
    X = np.asarray(X, dtype=np.float64)
    linear, constants = lda["linear"], lda["constants"]
    n, p = X.shape
    k = R.shape[1]
    C = linear.shape[0]
    latent_ops = n * k * (p + C)
    direct_ops = n * p * C
    if direct_ops < 0.5 * latent_ops:
        W = R @ linear.T
        cadj = constants.copy()
        if offset is not None and len(offset) >= k:
            cadj = cadj - np.asarray(offset)[:k] @ linear.T
        scores = X @ W
        labels = lda_labels_from_scores(scores, cadj)
        return (labels, scores + cadj) if return_scores else labels
    T = X @ R
    if offset is not None and len(offset) >= k:
        T = T - np.asarray(offset)[:k]
    scores = T @ linear.T
    labels = lda_labels_from_scores(scores, constants)
    return (labels, scores + constants) if return_scores else labels


# ---------------------------------------------------------------------------
# label-aware PLSSVD (R/main.R:506-655)
# ---------------------------------------------------------------------------

def class_sums(X: np.ndarray, y, n_classes: int, center: Optional[np.ndarray] = None):
    """rowsum(Xb, y) accumulated over blocks, transposed to p x C (R/main.R:563-586)."""
    y = np.asarray(y, dtype=np.int64)
    Xc = np.asarray(X, dtype=np.float64)
    if center is not None:
        Xc = Xc - center
    out = np.zeros((n_classes, X.shape[1]))
    np.add.at(out, y - 1, Xc)
    return out.T


def plssvd_label_aware(Xtrain, y, ncomp, scaling: int = 1) -> Dict:
    """.plssvd_label_aware_stream_model (R/main.R:513-655).  The block loop is
    a summation order only; the restatement sums whole columns."""
    X = np.asarray(Xtrain, dtype=np.float64)
    y = np.asarray(y, dtype=np.int64)
    n, p = X.shape
    m = int(y.max())
    if n < 1 or p < 1 or m < 2:
        raise ValueError("label-aware PLSSVD requires non-empty X and at least two classes")
    max_rank_cap = min(n, p, m)
    ncomp = np.array([min(max(int(c), 1), max_rank_cap) for c in np.atleast_1d(ncomp)])
    K = int(ncomp.max())
    sums = X.sum(axis=0)
    mX = sums / n if scaling < 3 else np.zeros(p)
    vX = np.ones(p)
    if scaling == 2:
        sums_sq = (X * X).sum(axis=0)
        vX = np.sqrt(np.maximum(sums_sq - n * mX * mX, 0) / max(1, n - 1))
        vX[~np.isfinite(vX) | (vX == 0)] = 1.0
    counts = np.bincount(y - 1, minlength=m).astype(np.float64)
    mY = counts / n
    Xc = X - mX if scaling < 3 else X.copy()
    if scaling == 2:
        Xc = Xc / vX
    cs = class_sums(Xc, y, m)
    total = cs.sum(axis=1)
    S = cs - np.outer(total, mY)
    U, d, V = svd_econ(S)
    R = U[:, :K]
    Q = V[:, :K]
    d = d[:K]
    T = Xc @ R
    G = T.T @ T
    L = ncomp.size
    C_latent = np.zeros((K, K, L))
    W_latent = np.zeros((K, m, L))
    for a, k in enumerate(ncomp):
        k = int(k)
        Gk = G[:k, :k]
        ridge = 1e-10 * float(np.mean(np.diag(Gk)))
        if not math.isfinite(ridge) or ridge <= 0:
            ridge = 1e-10
        A = Gk + ridge * np.eye(k)
        try:
            coeff = np.linalg.solve(A, np.diag(d[:k]))
        except np.linalg.LinAlgError:
            coeff = np.linalg.lstsq(A, np.diag(d[:k]), rcond=None)[0]
        C_latent[:k, :k, a] = coeff
        W_latent[:k, :, a] = coeff @ Q[:, :k].T
    return {"C_latent": C_latent, "W_latent": W_latent, "Q": Q, "R": R, "mX": mX, "vX": vX,
            "mY": mY, "p": p, "m": m, "ncomp": ncomp, "S": S, "s": d, "G": G,
            "pls_method": "plssvd", "B_stored": False, "counts": counts}


# ---------------------------------------------------------------------------
# candidate kNN (R/main.R:657-950, src/fastPLS.cpp:1249-1467)
# ---------------------------------------------------------------------------

def candidate_row_l2(X: np.ndarray) -> np.ndarray:
    """R/main.R:657-662: zero / non-finite norms map to 1."""
    nr = np.sqrt((X * X).sum(axis=1))
    nr[~np.isfinite(nr) | (nr == 0)] = 1.0
    return X / nr[:, None]


def candidate_centroids(Tn: np.ndarray, y_codes, n_classes: int) -> np.ndarray:
    """R/main.R:686-698."""
    y = np.asarray(y_codes, dtype=np.int64)
    sums = np.zeros((n_classes, Tn.shape[1]))
    np.add.at(sums, y - 1, Tn)
    counts = np.bincount(y - 1, minlength=n_classes).astype(np.float64)
    counts[counts == 0] = 1.0
    return candidate_row_l2(sums / counts[:, None])


def candidate_score_space(model: Dict, T: np.ndarray, k: int) -> np.ndarray:
    """R/main.R:700-730: PLSSVD scores are mapped through C_latent."""
    kk = min(int(k), T.shape[1])
    T = T[:, :kk]
    if model.get("pls_method") != "plssvd" or model.get("C_latent") is None:
        return T
    Cl = model["C_latent"]
    if Cl.ndim != 3 or Cl.shape[0] < kk or Cl.shape[1] < kk or Cl.shape[2] < 1:
        return T
    mn = [int(c) for c in model["ncomp"]]
    if kk in mn:
        sl = mn.index(kk)
    else:
        ge = [i for i, c in enumerate(mn) if c >= kk]
        sl = ge[0] if ge else len(mn) - 1
    sl = min(max(0, sl), Cl.shape[2] - 1)
    Ck = Cl[:kk, :kk, sl]
    if not np.all(np.isfinite(Ck)):
        return T
    return T @ Ck


def fit_candidate_knn(model: Dict, Xtrain, y_codes, knn_k=10, tau=0.2, alpha=0.75, top_m=20):
    """.fit_candidate_knn (R/main.R:795-873)."""
    attach_latent_projection_cache(model)
    K = int(max(model["ncomp"]))
    Ttrain = latent_scores(model, Xtrain, K)
    y = np.asarray(y_codes, dtype=np.int64)
    ncomp_eff = sorted(set(max(1, min(int(c), Ttrain.shape[1])) for c in model["ncomp"]))
    C = len(model["lev"]) if "lev" in model else int(y.max())
    models = {}
    for kk in ncomp_eff:
        Tn = candidate_row_l2(candidate_score_space(model, Ttrain, kk))
        models[kk] = {"ncomp": kk, "centroids": candidate_centroids(Tn, y, C)}
    model["candidate_knn"] = {"ncomp": ncomp_eff, "models": models,
                              "Ttrain": Ttrain[:, :max(ncomp_eff)], "y_codes": y,
                              "parameters": {"knn_k": max(1, int(knn_k)), "tau": float(tau),
                                             "alpha": float(alpha), "top_m": max(1, int(top_m))}}
    return model


def _stable_desc_order(v: np.ndarray) -> np.ndarray:
    """Order that repeated strict-'>' insertion (candidate_insert_top,
    src/fastPLS.cpp:1249-1267) produces: descending, ties keep visit order."""
    return np.argsort(-v, kind="stable")


def candidate_knn_predict(Ttest_norm, Ttrain_norm, y_codes, centroids, candidate_bias=None,
                          top_k=1, top_m=20, knn_k=10, tau=0.2, alpha=0.75,
                          return_candidates: bool = False):
    """candidate_knn_predict_native_impl (src/fastPLS.cpp:1322-1467)."""
    Ttest_norm = np.asarray(Ttest_norm, dtype=np.float64)
    Ttrain_norm = np.asarray(Ttrain_norm, dtype=np.float64)
    y = np.asarray(y_codes, dtype=np.int64)
    C = centroids.shape[0]
    ntest = Ttest_norm.shape[0]
    top_k = max(1, min(int(top_k), C))
    top_m = max(top_k, min(max(1, int(top_m)), C))
    knn_k = max(1, int(knn_k))
    tau = tau if (math.isfinite(tau) and tau > 0.0) else 0.2
    alpha = alpha if math.isfinite(alpha) else 0.5
    bias = np.zeros(C) if candidate_bias is None or len(candidate_bias) == 0 else np.asarray(candidate_bias, dtype=np.float64)
    base = Ttest_norm @ centroids.T
    order = np.argsort(-base, axis=1, kind="stable")[:, :top_m]
    cand = (order + 1).astype(np.int32)
    cand_base = np.take_along_axis(base, order, axis=1)
    class_rows = [np.nonzero(y == c + 1)[0] for c in range(C)]
    local = np.full((ntest, top_m), -np.inf)
    for c in range(C):
        rows = class_rows[c]
        ii, jj = np.nonzero(cand == c + 1)
        if ii.size == 0 or rows.size == 0:
            continue
        use_k = max(1, min(knn_k, rows.size))
        sims = Ttest_norm[ii] @ Ttrain_norm[rows].T
        if use_k < rows.size:
            part = -np.partition(-sims, use_k - 1, axis=1)[:, :use_k]
        else:
            part = sims
        top = -np.sort(-part, axis=1)
        mx = top[:, 0]
        acc = np.exp((top - mx[:, None]) / tau).sum(axis=1)
        local[ii, jj] = mx + tau * np.log(acc / use_k)
    score = local + alpha * cand_base + bias[cand - 1]
    top_index = np.zeros((ntest, top_k), dtype=np.int32)
    top_score = np.zeros((ntest, top_k))
    for i in range(ntest):
        o = _stable_desc_order(score[i])[:top_k]
        top_index[i] = cand[i, o]
        top_score[i] = score[i, o]
    out = {"top_index": top_index, "top_score": top_score}
    if return_candidates:
        out.update({"candidates": cand, "candidate_base": cand_base, "candidate_score": score})
    return out


def candidate_knn_predictions(model: Dict, Xtest, top: int = 1):
    """.candidate_knn_predictions (R/main.R:875-950): one result per ncomp."""
    ck = model["candidate_knn"]
    par = ck["parameters"]
    kmax = max(ck["ncomp"])
    Ttest = latent_scores(model, Xtest, kmax)
    C = ck["models"][ck["ncomp"][0]]["centroids"].shape[0]
    top = min(max(1, int(top)), C)
    res = []
    for kk_req in model["ncomp"]:
        kk = max(1, min(int(kk_req), kmax))
        cm = ck["models"][kk]
        r = candidate_knn_predict(candidate_row_l2(candidate_score_space(model, Ttest, kk)),
                                  candidate_row_l2(candidate_score_space(model, ck["Ttrain"], kk)),
                                  ck["y_codes"], cm["centroids"], None, top, max(top, par["top_m"]),
                                  par["knn_k"], par["tau"], par["alpha"])
        res.append(r)
    return res


# ---------------------------------------------------------------------------
# Gaussian response compression (R/main.R:1945-2142, src/fastPLS.cpp:4203-4215)
# The reference draws the projection and the codebook from R's RNG
# (set.seed + rnorm); here they come from gaussian_matrix(seed) -- that draw is
# "parity unpinned" (no R here); everything after the draw follows the
# reference and is pinned in tests/test_oracle_gaussian_y.py.
# ---------------------------------------------------------------------------

def gaussian_y_default_dim(p: int, gaussian_y_dim=None) -> int:
    """.gaussian_y_default_dim (R/main.R:1945-1954)."""
    if gaussian_y_dim is None:
        return max(1, min(int(p), 100))
    d = int(gaussian_y_dim)
    if d < 1:
        raise ValueError("gaussian_y_dim must be a positive integer")
    return d


def gaussian_y_ridge_decoder(Z: np.ndarray, Yc: np.ndarray) -> np.ndarray:
    """.gaussian_y_ridge_decoder (R/main.R:1967-1980): solve(Z'Z + ridge I, Z'Yc)."""
    ztz = Z.T @ Z
    ridge = 1e-8 * float(np.mean(np.diag(ztz)))
    if not math.isfinite(ridge) or ridge <= 0:
        ridge = 1e-8
    return np.linalg.solve(ztz + ridge * np.eye(ztz.shape[0]), Z.T @ Yc)


def gaussian_y_class_codes(n_classes: int, dim: int, seed: int) -> np.ndarray:
    """.gaussian_y_class_codes (R/main.R:1982-1995); draw from gaussian_matrix."""
    codes = gaussian_matrix(n_classes, dim, seed) / math.sqrt(dim)
    codes = codes - codes.mean(axis=0)
    sds = np.sqrt((codes * codes).sum(axis=0))
    sds[~np.isfinite(sds) | (sds == 0)] = 1.0
    codes = codes / sds
    rn = np.sqrt((codes * codes).sum(axis=1))
    rn[~np.isfinite(rn) | (rn == 0)] = 1.0
    return codes / rn[:, None]


def prepare_gaussian_y_regression(Y: np.ndarray, dim: int, seed: int) -> Dict:
    """.prepare_gaussian_y, regression branch (R/main.R:2031-2052)."""
    Y = np.asarray(Y, dtype=np.float64)
    center = Y.mean(axis=0)
    Yc = Y - center
    P = gaussian_matrix(Y.shape[1], dim, seed) / math.sqrt(dim)
    Z = Yc @ P
    return {"Z": Z, "y_center": center, "decoder": gaussian_y_ridge_decoder(Z, Yc),
            "projection": P}


def prepare_gaussian_y_classes(y_codes, n_classes: int, dim: int, seed: int) -> Dict:
    """.prepare_gaussian_y, classification branch (R/main.R:2015-2030)."""
    codes = gaussian_y_class_codes(n_classes, dim, seed)
    y = np.asarray(y_codes, dtype=np.int64)
    return {"Z": codes[y - 1], "codes": codes}


def decode_gaussian_y_regression(Z, decoder, center) -> np.ndarray:
    """.decode_gaussian_y_matrix, regression (R/main.R:2089-2096)."""
    return np.asarray(Z, dtype=np.float64) @ decoder + np.asarray(center)[None, :]


def gaussian_y_class_scores(Z, codes) -> np.ndarray:
    """.decode_gaussian_y_matrix, classification (R/main.R:2081-2087)."""
    return 2.0 * (np.asarray(Z, dtype=np.float64) @ codes.T) - (codes * codes).sum(axis=1)[None, :]


def nearest_code_classes(Z, codes) -> np.ndarray:
    """nearest_code_classes (src/fastPLS.cpp:4203-4215): 1-based, first maximum."""
    return (np.argmax(gaussian_y_class_scores(Z, codes), axis=1) + 1).astype(np.int32)


# ---------------------------------------------------------------------------
# end-to-end pls() flows used by the tests and the CPU baseline (R/main.R:5860-6314)
# ---------------------------------------------------------------------------

class RUniform:
    """R's default uniform generator after ``set.seed(seed)`` (R's RNG.c:
    RNG_Init scrambles the seed by 50 steps of s <- 69069 s + 1 and fills the
    Mersenne-Twister position slot and 624 state words with the next 625 steps;
    MT_genrand returns the tempered MT19937 word times 2^-32; fixup keeps the
    value inside (0, 1)).  The reference draws its CV permutations from it
    (unif_rand in samplewithoutreplace, src/fastPLS.cpp:4498-4509, after
    set.seed(seed) in R/main.R:3883).  Pinned by R's known runif() outputs."""

    _M32 = 0xFFFFFFFF

    def __init__(self, seed: int):
        s = int(seed) & self._M32
        for _ in range(50):
            s = (69069 * s + 1) & self._M32
        words = []
        for _ in range(625):
            s = (69069 * s + 1) & self._M32
            words.append(s)
        self.mt = words[1:]
        self.mti = 624

    def _twist(self):
        mt = self.mt
        for k in range(624):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
            mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self.mti = 0

    def unif_rand(self) -> float:
        if self.mti >= 624:
            self._twist()
        y = self.mt[self.mti]
        self.mti += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        x = y * 2.3283064365386963e-10
        i2_32m1 = 2.328306437080797e-10
        if x <= 0.0:
            return 0.5 * i2_32m1
        if 1.0 - x <= 0.0:
            return 1.0 - 0.5 * i2_32m1
        return x


def samplewithoutreplace(yy, size: int, rng: RUniform):
    """samplewithoutreplace (src/fastPLS.cpp:4498-4509)."""
    yy = list(yy)
    out = []
    for _ in range(size):
        it = int(rng.unif_rand() * len(yy))
        out.append(yy.pop(it))
    return out


def cv_folds(constrain, kfold: int, labels=None, n_classes: int = 0, seed: int = 1):
    """Row -> fold map (0-based) and fold count of pls_cv_predict_compiled
    (src/fastPLS.cpp:6666-6775): kfold < 0 or kfold >= #groups leaves one
    group out; otherwise each class's groups (class of the group's first row)
    are permuted and dealt round-robin, or all groups when unstratified."""
    cons = np.asarray(constrain).ravel()
    groups = np.unique(cons)
    g_of = np.searchsorted(groups, cons)
    ng = groups.size
    kfold = int(kfold)
    loo = kfold < 0
    if not loo and kfold < 2:
        kfold = 2
    group_fold = np.zeros(ng, dtype=np.int64)
    if loo or kfold >= ng:
        return g_of.astype(np.int64), max(ng, 1)
    rng = RUniform(seed)
    if labels is not None and n_classes > 1:
        lab = np.asarray(labels, dtype=np.int64).ravel()
        first = np.array([np.nonzero(g_of == j)[0][0] for j in range(ng)])
        cls = np.clip(lab[first], 1, n_classes)
        for c in range(1, n_classes + 1):
            members = [j for j in range(ng) if cls[j] == c]
            if not members:
                continue
            perm = samplewithoutreplace(range(1, len(members) + 1), len(members), rng)
            for pos, pv in enumerate(perm):
                group_fold[members[pv - 1]] = pos % kfold
    else:
        perm = samplewithoutreplace(range(1, ng + 1), ng, rng)
        for j in range(ng):
            group_fold[j] = (perm[j] - 1) % kfold
    return group_fold[g_of], kfold


def pls_cv_predict(X, Y, constrain, ncomp, method="simpls", classification=False,
                   n_classes=None, classifier="argmax", scaling=1, seed=1, lda_ridge=1e-8,
                   knn_k=10, tau=0.2, alpha=0.75, top_m=20, north=1, kfold=-1):
    """pls_cv_predict_compiled (src/fastPLS.cpp:6615-7151) with the folds of
    ``cv_folds`` (kfold < 0: leave one group out, group_fold(j) = j over the
    sorted distinct constrain values, :6720-6736).
    Y holds 1-based labels when ``classification``.  Returns fold (1-based),
    status, class_pred (n x L) and Ypred (n x m x L, argmax / regression).
    method "opls" filters each training fold (opls_filter_cpp with ``north``
    components) and applies the filter to the held-out rows before a SIMPLS
    fit with scaling "none"; "kernelpls" is the SIMPLS fit (its linear kernel,
    :6874-6891)."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    cons = np.asarray(constrain).ravel()
    L = len(np.atleast_1d(ncomp))
    nc = np.array(np.atleast_1d(ncomp), dtype=np.int64)
    if classification:
        y = np.asarray(Y, dtype=np.int64).ravel()
        C = int(n_classes or y.max())
        m = C
        Yd = np.zeros((n, C))
        Yd[np.arange(n), y - 1] = 1.0
    else:
        Yd = np.asarray(Y, dtype=np.float64).reshape(n, -1)
        m = Yd.shape[1]
    fold, nfold = cv_folds(cons, kfold, y if classification else None,
                           C if classification else 0, seed)
    if method == "plssvd":
        nc = np.minimum(np.maximum(nc, 1), min(n, X.shape[1], m))
    class_pred = np.zeros((n, L), dtype=np.int32)
    Ypred = np.zeros((n, m, L))
    # decision margin per row and slice (inf where no decision is made): the
    # gap between the best and the second class score, for the tie log
    margin = np.full((n, L), np.inf)

    def gap(scores):
        srt = np.sort(scores, axis=1)
        return srt[:, -1] - srt[:, -2] if scores.shape[1] > 1 else np.full(scores.shape[0], np.inf)

    status = np.zeros(nfold, dtype=np.int32)
    for f in range(nfold):
        te = np.nonzero(fold == f)[0]
        tr = np.nonzero(fold != f)[0]
        if te.size == 0:
            status[f] = 2
            continue
        if tr.size == 0:
            status[f] = 3
            Ypred[te] = Yd[te][:, :, None]
            if classification:
                class_pred[te] = y[te][:, None]
            continue
        if classification:
            active = np.unique(y[tr])
            if active.size <= 1:
                status[f] = 4
                row = np.zeros(m)
                if active.size == 1:
                    row[active[0] - 1] = 1.0
                Ypred[te] = row[None, :, None]
                class_pred[te] = active[0] if active.size == 1 else 1
                continue
        Xtr, Xte, fit_scaling = X[tr], X[te], scaling
        if method == "opls":
            filt = opls_filter(Xtr, Yd[tr], max(int(north), 0), scaling)
            Xtr = filt["X"]
            Xte = opls_apply_filter(Xte, filt["mX"], filt["vX"], filt["W_orth"], filt["P_orth"])
            fit_scaling = 3
        fit = pls_model1 if method == "plssvd" else pls_model2_fast
        mod = fit(Xtr, Yd[tr], nc, fit_scaling, False, "rsvd", 10, 1, seed + f)
        if classification and classifier in ("lda", "cknn"):
            attach_latent_projection_cache(mod)
            kmax = min(int(nc.max()), mod["R"].shape[1])
            Ttr = latent_scores(mod, Xtr, kmax)
            Tte = latent_scores(mod, Xte, kmax)
            if classifier == "lda":
                ks = [min(int(k), kmax) for k in nc]
                lda = lda_train_prefix(Ttr, y[tr], C, ks, lda_ridge)
                for s, k in enumerate(ks):
                    sc = Tte[:, :k] @ lda[s]["linear"].T
                    class_pred[te, s] = lda_labels_from_scores(sc, lda[s]["constants"])
                    margin[te, s] = gap(sc + lda[s]["constants"])
            else:
                for s, k in enumerate(nc):
                    k = min(int(k), kmax)
                    a = candidate_row_l2(candidate_score_space(mod, Ttr, k))
                    b = candidate_row_l2(candidate_score_space(mod, Tte, k))
                    cen = candidate_centroids(a, y[tr], C)
                    r = candidate_knn_predict(b, a, y[tr], cen, None, 2, top_m, knn_k, tau, alpha)
                    class_pred[te, s] = r["top_index"][:, 0]
                    mg = r["top_score"][:, 0] - r["top_score"][:, 1]
                    tm = max(2, min(int(top_m), C))
                    if C > tm:  # a class at the candidate boundary may swap in
                        bs = -np.sort(-(b @ cen.T), axis=1)
                        mg = np.minimum(mg, bs[:, tm - 1] - bs[:, tm])
                    margin[te, s] = mg
        else:
            yp = pls_predict(mod, Xte)
            Ypred[te] = yp
            if classification:
                class_pred[te] = (np.argmax(yp, axis=1) + 1).astype(np.int32)
                for s in range(L):
                    margin[te, s] = gap(yp[:, :, s])
        status[f] = 1
    return {"fold": fold + 1, "status": status, "class_pred": class_pred, "Ypred": Ypred,
            "margin": margin}


def pls_plssvd_lda(Xtrain, ytrain, Xtest, ncomp, lda_ridge=1e-8, scaling=1, seed=1,
                   oversample=10, power=1):
    """pls(method="plssvd", svd.method="rsvd", classifier="lda") then predict:
    transformy -> pls_model1 -> .attach_lda_classifier (projected training)
    -> lda_project_predict_labels_cpp (R/main.R:5860-6314, 1098-1290, 1544-1600)."""
    y = np.asarray(ytrain, dtype=np.int64)
    C = int(y.max())
    model = pls_model1(Xtrain, transformy(y), ncomp, scaling, False, "rsvd", oversample, power, seed)
    attach_latent_projection_cache(model)
    Rp = model["R_predict"]
    ncomp_eff = sorted(set(max(1, min(int(c), Rp.shape[1])) for c in model["ncomp"]))
    lda = lda_project_train_prefix(Xtrain, Rp[:, :max(ncomp_eff)], model["R_offset"], y, C,
                                   ncomp_eff, lda_ridge)
    model["lda"] = dict(zip(ncomp_eff, lda))
    preds = []
    for kk_req in model["ncomp"]:
        kk = max(1, min(int(kk_req), max(ncomp_eff)))
        preds.append(lda_project_predict_labels(Xtest, Rp[:, :kk], model["R_offset"][:kk],
                                                model["lda"][kk]))
    model["Ypred_labels"] = np.stack(preds, axis=1)
    return model


# ---------------------------------------------------------------------------
# Argmax-head class ranking with per-class offsets
# (src/fastPLS.cpp:3737-4113; R/main.R:377-492)
# ---------------------------------------------------------------------------

def compact_prediction_weights(model: Dict) -> np.ndarray:
    """compact_prediction_weights (src/fastPLS.cpp:3737-3836): Wflash
    (kmax x m x L); slice a holds Q[:, :mc]^T (SIMPLS) or W_latent[:, :, a]
    (PLSSVD) in its first mc rows."""
    R = model["R"]
    kmax = R.shape[1]
    m = int(model["m"])
    ncomp = np.asarray(model["ncomp"])
    W = np.zeros((kmax, m, ncomp.size))
    for a in range(ncomp.size):
        mc = int(ncomp[a])
        if mc < 1 or mc > kmax:
            raise ValueError("ncomp exceeds latent rank for compact class prediction")
        if model["pls_method"] == "simpls":
            W[:mc, :, a] = model["Q"][:, :mc].T
        else:
            W[:, :, a] = model["W_latent"][:, :, a]
    return W


def class_bias_offsets(mY, class_bias, L: int) -> np.ndarray:
    """class_bias_offsets (src/fastPLS.cpp:3838-3861): m x L, mY in every
    column plus class_bias (empty, one column broadcast, or one per slice)."""
    mY = np.asarray(mY, dtype=np.float64).reshape(-1)
    m = mY.size
    off = np.repeat(mY[:, None], L, axis=1)
    if class_bias is None or np.size(class_bias) == 0:
        return off
    b = np.asarray(class_bias, dtype=np.float64)
    if b.ndim == 1:
        b = b[:, None]
    if b.shape[0] != m:
        raise ValueError("class_bias must have one row per class")
    if b.shape[1] == 1:
        return off + b[:, :1]
    if b.shape[1] == L:
        return off + b
    raise ValueError("class_bias must have one column or one column per ncomp")


def fill_topk_from_yblock(yblock, offset, top_k: int):
    """fill_topk_from_yblock (src/fastPLS.cpp:3863-3906), loop for loop: per
    row, value = yblock[i, j] + offset[j] inserted with a strict '>' into a
    list initialised to (-inf, class 0).  Returns (index (rows x k, 1-based),
    score (rows x k)) with k = max(1, min(top_k, m))."""
    yblock = np.asarray(yblock, dtype=np.float64)
    offset = np.asarray(offset, dtype=np.float64).reshape(-1)
    n, m = yblock.shape
    k = max(1, min(int(top_k), m))
    idx = np.zeros((n, k), dtype=np.int32)
    sc = np.full((n, k), -np.inf)
    for i in range(n):
        bs = [-math.inf] * k
        bi = [0] * k
        row = yblock[i]
        for j in range(m):
            v = float(row[j]) + float(offset[j])
            for r in range(k):
                if v > bs[r]:
                    bs[r + 1:] = bs[r:k - 1]
                    bi[r + 1:] = bi[r:k - 1]
                    bs[r] = v
                    bi[r] = j + 1
                    break
        idx[i] = bi
        sc[i] = bs
    return idx, sc


def class_bias_topk_from_cube(Ypred, class_bias, top_k: int):
    """class_bias_topk_from_cube (src/fastPLS.cpp:3908-3952): ranks of a given
    score cube (w x m x L) with class_bias alone as the offsets.  Returns
    top_index / top_score cubes (w x k x L)."""
    Y = np.asarray(Ypred, dtype=np.float64)
    if Y.ndim == 2:
        Y = Y[:, :, None]
    w, m, L = Y.shape
    off = class_bias_offsets(np.zeros(m), class_bias, L)
    k = max(1, min(int(top_k), m))
    ti = np.zeros((w, k, L), dtype=np.int32)
    ts = np.zeros((w, k, L))
    for a in range(L):
        ti[:, :, a], ts[:, :, a] = fill_topk_from_yblock(Y[:, :, a], off[:, a], k)
    return ti, ts


def pls_class_predict_topk(model: Dict, Xtest, class_bias=None, top_k: int = 1,
                           block_size: int = 4096):
    """pls_class_predict_topk_cpp (src/fastPLS.cpp:3954-4113), compact path:
    scores = ((Xtest - mX)/vX) R per row block, yblock = scores[:, :mc] Wflash_a
    and fill_topk_from_yblock with offsets mY + class_bias."""
    X = (np.asarray(Xtest, dtype=np.float64) - model["mX"]) / model["vX"]
    ncomp = np.asarray(model["ncomp"])
    L = ncomp.size
    m = int(model["m"])
    k = max(1, min(int(top_k), m))
    off = class_bias_offsets(model["mY"], class_bias, L)
    W = compact_prediction_weights(model)
    R = model["R"]
    w = X.shape[0]
    ti = np.zeros((w, k, L), dtype=np.int32)
    ts = np.zeros((w, k, L))
    bs = max(1, min(int(block_size), max(w, 1)))
    for s in range(0, w, bs):
        sc = X[s:s + bs] @ R
        for a in range(L):
            mc = int(ncomp[a])
            yb = sc[:, :mc] @ W[:mc, :, a]
            ti[s:s + bs, :, a], ts[s:s + bs, :, a] = fill_topk_from_yblock(yb, off[:, a], k)
    return ti, ts


def class_topk_from_score_cube(score_cube, class_bias=None, top: int = 1):
    """.class_topk_from_score_cube (R/main.R:453-474): score + bias per slice;
    top == 1 takes max.col(ties.method = "first"), top > 1 the first `top` of
    order(decreasing = TRUE), R's stable order (ties by class index).
    Returns (top_index, top_score) cubes (w x top x L)."""
    Y = np.asarray(score_cube, dtype=np.float64)
    if Y.ndim == 2:
        Y = Y[:, :, None]
    w, m, L = Y.shape
    top = min(int(top), m)
    off = class_bias_offsets(np.zeros(m), class_bias, L)
    ti = np.zeros((w, top, L), dtype=np.int32)
    ts = np.zeros((w, top, L))
    for a in range(L):
        S = Y[:, :, a] + off[:, a]
        order = np.argsort(-S, axis=1, kind="stable")[:, :top]
        ti[:, :, a] = order + 1
        ts[:, :, a] = np.take_along_axis(S, order, axis=1)
    return ti, ts


# ---------------------------------------------------------------------------
# OPLS filter and kernel matrices (src/fastPLS.cpp:4314-4483; R/main.R:2764-3249)
# ---------------------------------------------------------------------------

def kernel_matrix(X1, X2, kernel: int, gamma: float, degree: int = 3, coef0: float = 1.0):
    """kernel_matrix_cpp (src/fastPLS.cpp:4314-4341): 1 linear, 2 rbf, 3 poly."""
    X1 = np.asarray(X1, dtype=np.float64)
    X2 = np.asarray(X2, dtype=np.float64)
    if X1.shape[1] != X2.shape[1]:
        raise ValueError("X1 and X2 must have the same number of columns")
    dots = X1 @ X2.T
    if kernel == 1:
        return dots
    if kernel == 3:
        return np.power(gamma * dots + coef0, float(degree))
    if kernel != 2:
        raise ValueError("Unknown kernel id")
    n1 = np.sum(X1 * X1, axis=1)
    n2 = np.sum(X2 * X2, axis=1)
    d2 = (n1[:, None] + n2[None, :]) - 2.0 * dots
    d2 = np.where((d2 < 0.0) & (d2 > -1e-10), 0.0, d2)
    return np.exp(-gamma * d2)


def center_kernel_train(K):
    """center_kernel_train_cpp (src/fastPLS.cpp:4344-4357)."""
    K = np.asarray(K, dtype=np.float64)
    cm = K.mean(axis=0)
    rm = K.mean(axis=1)
    gm = float(cm.mean())
    return {"K": ((K - cm[None, :]) - rm[:, None]) + gm, "col_means": cm, "grand_mean": gm}


def center_kernel_test(Ktest, col_means, grand_mean: float):
    """center_kernel_test_cpp (src/fastPLS.cpp:4360-4374)."""
    K = np.asarray(Ktest, dtype=np.float64)
    cm = np.asarray(col_means, dtype=np.float64).reshape(-1)
    if K.shape[1] != cm.size:
        raise ValueError("Ktest columns must match the training kernel size")
    rm = K.mean(axis=1)
    return ((K - cm[None, :]) - rm[:, None]) + grand_mean


def preprocess_train(X, scaling: int):
    """.fastpls_preprocess_train (R/main.R:2764-2779) == opls_filter_cpp's
    preprocessing: centre for scaling 1/2, divide by the column sd (n - 1)
    for 2 (zero / non-finite sd -> 1)."""
    X = np.asarray(X, dtype=np.float64)
    mX = np.zeros(X.shape[1])
    vX = np.ones(X.shape[1])
    if scaling < 3:
        mX = X.mean(axis=0)
        X = X - mX
    if scaling == 2:
        vX = X.std(axis=0, ddof=1) if X.shape[0] > 1 else np.zeros(X.shape[1])
        vX = np.where(np.isfinite(vX) & (vX != 0.0), vX, 1.0)
        X = X / vX
    return X, mX, vX


def opls_filter(X, Y, north: int, scaling: int = 1) -> Dict:
    """opls_filter_cpp (src/fastPLS.cpp:4377-4461)."""
    if north < 0:
        raise ValueError("north must be >= 0")
    X, mX, vX = preprocess_train(X, scaling)
    Y = np.asarray(Y, dtype=np.float64)
    if Y.ndim == 1:
        Y = Y[:, None]
    if X.shape[0] != Y.shape[0]:
        raise ValueError("X and Y must have the same number of rows")
    Y = Y - Y.mean(axis=0)
    Wo, Po = [], []
    for _ in range(north):
        S = X.T @ Y
        U = np.linalg.svd(S, full_matrices=False)[0]
        if U.shape[1] == 0:
            break
        w = U[:, 0]
        wn = np.linalg.norm(w)
        if not np.isfinite(wn) or wn <= 0.0:
            break
        w = w / wn
        t = X @ w
        tss = float(t @ t)
        if not np.isfinite(tss) or tss <= 0.0:
            break
        p = X.T @ t / tss
        wo = p - w * (float(w @ p) / float(w @ w))
        won = np.linalg.norm(wo)
        if not np.isfinite(won) or won <= 0.0:
            break
        wo = wo / won
        to = X @ wo
        toss = float(to @ to)
        if not np.isfinite(toss) or toss <= 0.0:
            break
        po = X.T @ to / toss
        X = X - np.outer(to, po)
        Wo.append(wo)
        Po.append(po)
    p_ = X.shape[1]
    return {"X": X, "mX": mX, "vX": vX,
            "W_orth": np.array(Wo).T if Wo else np.zeros((p_, 0)),
            "P_orth": np.array(Po).T if Po else np.zeros((p_, 0)), "north": len(Wo)}


def opls_apply_filter(X, mX, vX, W_orth, P_orth):
    """opls_apply_filter_cpp (src/fastPLS.cpp:4463-4483)."""
    X = (np.asarray(X, dtype=np.float64) - mX) / vX
    for a in range(W_orth.shape[1]):
        t = X @ W_orth[:, a]
        X = X - np.outer(t, P_orth[:, a])
    return X


def opls_pls(Xtrain, Ytrain, Xtest, ncomp, north: int = 1, scaling: int = 1, classes=None,
             oversample=10, power=1, seed=1):
    """.opls_fit (R/main.R:3047-3091) with the SIMPLS inner fit on the filtered
    X (scaling "none") and predict.fastPLSOpls (R/main.R:3230-3242).  classes:
    the number of classes when Ytrain holds 1-based labels (the filter sees
    their one-hot, .supervised_response_matrix)."""
    Yf = transformy(Ytrain) if classes else np.asarray(Ytrain, dtype=np.float64)
    filt = opls_filter(Xtrain, Yf, north, scaling)
    mod = pls_model2_fast(filt["X"], Yf, ncomp, 3, False, "rsvd", oversample, power, seed)
    Xt = opls_apply_filter(Xtest, filt["mX"], filt["vX"], filt["W_orth"], filt["P_orth"])
    return {"filter": filt, "inner": mod, "Ypred": pls_predict(mod, Xt)}


def kernel_pls(Xtrain, Ytrain, Xtest, ncomp, kernel: int = 2, gamma=None, degree: int = 3,
               coef0: float = 1.0, scaling: int = 1, classes=None, oversample=10, power=1,
               seed=1):
    """.kernel_pls_fit (R/main.R:2826-2905) for the rbf / poly kernels and
    predict.fastPLSKernel (R/main.R:3031-3045): SIMPLS on the centred training
    kernel (scaling "none"), prediction from the centred test kernel."""
    Xp, mX, vX = preprocess_train(Xtrain, scaling)
    if gamma is None:
        gamma = 1.0 / max(1, Xp.shape[1])
    K = kernel_matrix(Xp, Xp, kernel, gamma, degree, coef0)
    kc = center_kernel_train(K)
    Yf = transformy(Ytrain) if classes else np.asarray(Ytrain, dtype=np.float64)
    mod = pls_model2_fast(kc["K"], Yf, ncomp, 3, False, "rsvd", oversample, power, seed)
    Xn = (np.asarray(Xtest, dtype=np.float64) - mX) / vX
    Kt = center_kernel_test(kernel_matrix(Xn, Xp, kernel, gamma, degree, coef0),
                            kc["col_means"], kc["grand_mean"])
    return {"Kc": kc, "inner": mod, "gamma": gamma, "Ypred": pls_predict(mod, Kt), "Ktest": Kt}
