"""GPU parity: truncated SVD, PLSSVD and the LDA head against the FP64 oracle.

Tolerances: FP64 inputs run FP64 kernels end to end (rel 1e-9).  FP32 inputs
run FP32 MFMA contractions with FP64 factorisations (rel 1e-4 on
coefficients); predicted labels must match exactly except rows whose FP64
oracle top-2 margin is below 1e-3 of the score scale (documented near-ties).
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import align_signs, label_mismatches_excused, rel_err, sin_theta

pytestmark = pytest.mark.gpu


def _lowrank(rng, n, p, r, noise):
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    V, _ = np.linalg.qr(rng.standard_normal((p, r)))
    s = 10.0 * 0.8 ** np.arange(r)
    return (U * s) @ V.T + noise * rng.standard_normal((n, p))


@pytest.mark.parametrize("rows,cols,k,over,power", [
    (300, 40, 10, 10, 1),   # rSVD, q == 1 (no re-orthonormalisation)
    (300, 60, 8, 6, 2),     # rSVD, q == 2 (QR between passes)
    (500, 30, 25, 10, 1),   # l >= max_rank -> exact
    (50, 200, 5, 10, 0),    # wide, q == 0
])
def test_fastsvd_matches_oracle(ctx, rows, cols, k, over, power):
    import fastpls_amd as F
    rng = np.random.default_rng(7)
    A = _lowrank(rng, rows, cols, min(12, min(rows, cols)), 1e-3)
    got = F.fastsvd(A, k, rsvd_oversample=over, rsvd_power=power, seed=3)
    ref = O.truncated_svd(A, k, "rsvd", over, power, 3)
    kk = ref["U"].shape[1]
    assert got["u"].shape[1] == kk
    np.testing.assert_allclose(got["d"], ref["s"], rtol=1e-9, atol=1e-12 * ref["s"][0])
    # leading singular vectors with a clear gap agree column-wise after sign alignment
    kgap = min(kk, 8)
    U = align_signs(got["u"][:, :kgap], ref["U"][:, :kgap])
    assert rel_err(U, ref["U"][:, :kgap]) < 1e-7
    assert sin_theta(got["u"], ref["U"]) < 1e-7


def test_plssvd_regression_fp64(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(11)
    X = _lowrank(rng, 150, 60, 15, 0.1)
    Y = X @ rng.standard_normal((60, 8)) * 0.1 + 0.05 * rng.standard_normal((150, 8))
    ncomp = [2, 5, 7]
    m = F.pls(X, Y, method="plssvd", ncomp=ncomp, fit=True)
    ref = O.pls_model1(X, Y, ncomp, 1, True, "rsvd", 10, 1, 1)
    assert rel_err(m["mX"], ref["mX"]) < 1e-13
    assert rel_err(m["mY"], ref["mY"]) < 1e-13
    R = align_signs(m["R"], ref["R"])
    assert rel_err(R, ref["R"]) < 1e-9
    # B, Yfit and R2Y are sign invariant
    assert rel_err(m["B"], ref["B"]) < 1e-9
    assert rel_err(m["Yfit"], ref["Yfit"]) < 1e-9
    np.testing.assert_allclose(m["R2Y"], ref["R2Y"], rtol=1e-9)
    Xt = rng.standard_normal((20, 60))
    yp = F.predict(m, Xt)["Ypred"]
    assert rel_err(yp, O.pls_predict(ref, Xt)) < 1e-9


def test_plssvd_autoscaling_fp64(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(12)
    X = _lowrank(rng, 120, 40, 10, 0.2) * rng.uniform(0.5, 3.0, 40) + 2.0
    Y = rng.standard_normal((120, 6))
    m = F.pls(X, Y, method="plssvd", ncomp=[3, 6], scaling="autoscaling")
    ref = O.pls_model1(X, Y, [3, 6], 2, False, "rsvd", 10, 1, 1)
    assert rel_err(m["vX"], ref["vX"]) < 1e-12
    assert rel_err(m["B"], ref["B"]) < 1e-9


def test_plssvd_lda_fp32_labels(ctx):
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=12, per_class_train=150,
                                                 per_class_test=40, p=96, seed=5)
    ncomp = 11
    m = F.pls(Xtr, ytr, Xte, method="plssvd", ncomp=ncomp, classifier="lda", tie_log=True)
    ref = O.pls_plssvd_lda(Xtr.astype(np.float64), ytr, Xte.astype(np.float64), [ncomp])
    # coefficients within the FP32 budget
    assert rel_err(align_signs(m["R"], ref["R"]), ref["R"]) < 1e-4
    assert rel_err(m["B"], ref["B"]) < 1e-4
    lda_ref = ref["lda"][ncomp]
    C = 12
    lin = m.lda_read(0, "linear", (C, ncomp))
    cst = m.lda_read(0, "constants", (C,))
    # linear terms flip with the latent signs; constants are sign invariant
    np.testing.assert_allclose(cst, lda_ref["constants"], rtol=1e-4, atol=1e-4)
    assert rel_err(np.abs(lin), np.abs(lda_ref["linear"])) < 1e-3
    labels = m["Ypred_index"][:, 0]
    _, sc = O.lda_project_predict_labels(Xte.astype(np.float64), ref["R_predict"][:, :ncomp],
                                         ref["R_offset"][:ncomp], lda_ref, return_scores=True)
    scale = float(np.abs(sc).max())
    bad, unexcused = label_mismatches_excused(labels, ref["Ypred_labels"][:, 0], sc, 1e-3 * scale,
                                              m["tie_margin"][:, 0])
    assert not unexcused, f"label mismatches beyond ties: {unexcused}"


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_precision_modes_subspace_error(ctx, precision):
    """Both contraction modes keep the latent subspace within sin(theta) <= 1e-4
    of the FP64 oracle and labels equal up to near-ties."""
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=20, per_class_train=200,
                                                 per_class_test=30, p=256, seed=17)
    k = 19
    m = F.pls(Xtr, ytr, Xte, method="plssvd", ncomp=k, classifier="lda", precision=precision,
              tie_log=True)
    ref = O.pls_plssvd_lda(Xtr.astype(np.float64), ytr, Xte.astype(np.float64), [k])
    assert sin_theta(m["R"], ref["R"]) < 1e-4
    assert rel_err(m["B"], ref["B"]) < (1e-4 if precision == "fp32" else 5e-4)
    lda_ref = ref["lda"][k]
    _, sc = O.lda_project_predict_labels(Xte.astype(np.float64), ref["R_predict"][:, :k],
                                         ref["R_offset"][:k], lda_ref, return_scores=True)
    bad, unexcused = label_mismatches_excused(m["Ypred_index"][:, 0], ref["Ypred_labels"][:, 0],
                                              sc, 1e-3 * float(np.abs(sc).max()),
                                              m["tie_margin"][:, 0])
    assert not unexcused


def test_bf16x3_gemm_accuracy(ctx):
    """Projection X R through the split-BF16x3 MFMA path: relative error well
    below 1e-4 (three products of 8-bit-mantissa halves, FP32 accumulation)."""
    import fastpls_amd as F
    rng = np.random.default_rng(18)
    X = rng.standard_normal((3000, 300)).astype(np.float32)
    y = rng.standard_normal((3000, 5)).astype(np.float32)
    m32 = F.pls(X, y, method="plssvd", ncomp=5, precision="fp32")
    m16 = F.pls(X, y, method="plssvd", ncomp=5, precision="bf16x3")
    ref = O.pls_model1(X.astype(np.float64), y.astype(np.float64), [5])
    e32 = rel_err(m32["B"], ref["B"])
    e16 = rel_err(m16["B"], ref["B"])
    assert e32 < 1e-5 and e16 < 1e-4, (e32, e16)


def test_plssvd_label_aware_equals_dense_onehot(ctx):
    """Label-aware S == Xc^T onehot(y) (dense); one-hot never formed."""
    import fastpls_amd as F
    rng = np.random.default_rng(3)
    n, p, C = 400, 30, 6
    X = rng.standard_normal((n, p)) + 1.5
    y = rng.integers(1, C + 1, n).astype(np.int32)
    m_lab = F.pls(X, y, method="plssvd", ncomp=[4])
    m_dense = F.pls(X, O.transformy(y), method="plssvd", ncomp=[4])
    assert rel_err(m_lab["B"], m_dense["B"]) < 1e-11
    ref = O.pls_model1(X, O.transformy(y), [4], 1, False, "rsvd", 10, 1, 1)
    assert rel_err(m_lab["B"], ref["B"]) < 1e-9


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("power", [0, 1, 2])
def test_rsvd_crossprod_matches_oracle(ctx, dtype, power):
    import fastpls_amd as F
    rng = np.random.default_rng(40 + power)
    n, p, m = 400, 150, 60
    X = _lowrank(rng, n, p, 20, 0.05)
    Y = X @ rng.standard_normal((p, m)) * 0.05 + 0.01 * rng.standard_normal((n, m))
    Xd = X.astype(dtype)
    Yd = Y.astype(dtype)
    got = F.truncated_rsvd_crossprod(Xd, Yd, 8, rsvd_oversample=6, rsvd_power=power, seed=4)
    ref = O.truncated_rsvd_crossprod(Xd.astype(np.float64), Yd.astype(np.float64), 8, 6, power, 4)
    tol = 1e-9 if dtype == np.float64 else 2e-4
    np.testing.assert_allclose(got["s"], ref["s"], rtol=tol)
    assert sin_theta(got["U"][:, :6], ref["U"][:, :6]) < (1e-7 if dtype == np.float64 else 1e-3)


def test_class_sums_exact_counts(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(50)
    n, p, C = 20000, 33, 17
    X = rng.standard_normal((n, p)).astype(np.float32)
    y = rng.integers(1, C + 1, n).astype(np.int32)
    r = F.class_sums(X, y, C)
    np.testing.assert_array_equal(r["counts"], np.bincount(y - 1, minlength=C))
    ref = O.class_sums(X.astype(np.float64), y, C)
    np.testing.assert_allclose(r["class_sums"], ref, rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(r["col_sums"], X.astype(np.float64).sum(0), rtol=1e-9, atol=1e-6)
    with pytest.raises(F.FplsError, match="LABEL_RANGE"):
        F.class_sums(X, np.where(y == 3, C + 1, y).astype(np.int32), C)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_fused_plssvd_lda_equals_separate_calls(ctx, precision):
    """fpls_plssvd_lda_fit is fpls_plssvd_fit + fpls_lda_train_prefix with the
    solves batched: every model and LDA field equal to the last bit."""
    import ctypes
    import fastpls_amd as F
    from fastpls_amd import _lib, synth
    Xtr, ytr, Xte, _ = synth.classes_on_sphere(n_classes=9, per_class_train=120,
                                               per_class_test=30, p=64, seed=12)
    keep = []
    Xm = _lib.as_mat(Xtr, keep)
    yv = _lib.as_ivec(ytr, keep)
    nc, ncp = _lib.int_array([3, 8])
    opts = _lib.default_opts(precision=_lib.PREC_BF16X3 if precision == "bf16x3" else 0)
    h1, h2 = ctypes.c_void_p(), ctypes.c_void_p()
    ctx.check(ctx.lib.fpls_plssvd_lda_fit(ctx.h, ctypes.byref(Xm), ctypes.byref(yv), 9, ncp, 2,
                                          ctypes.byref(opts), 1e-8, ctypes.byref(h1)))
    ctx.check(ctx.lib.fpls_plssvd_fit(ctx.h, ctypes.byref(Xm), None, ctypes.byref(yv), 9, ncp, 2,
                                      ctypes.byref(opts), ctypes.byref(h2)))
    ctx.check(ctx.lib.fpls_lda_train_prefix(ctx.h, h2, None, ctypes.byref(yv), 9, ncp, 2, 1e-8))
    m1, m2 = F.FastPLS(ctx, h1), F.FastPLS(ctx, h2)
    p, k = 64, 8
    for name, shape in (("R", (p, k)), ("C_latent", (k, k, 2)), ("W_latent", (k, 9, 2)),
                        ("B", (p, 9, 2)), ("mX", (p,)), ("mY", (9,))):
        np.testing.assert_array_equal(m1.read(name, shape), m2.read(name, shape), err_msg=name)
    for s, kk in enumerate([3, 8]):
        for field, shape in (("means", (9, kk)), ("inv_cov", (kk, kk)), ("linear", (9, kk)),
                             ("constants", (9,)), ("priors", (9,)), ("ridge", (1,))):
            np.testing.assert_array_equal(m1.lda_read(s, field, shape),
                                          m2.lda_read(s, field, shape), err_msg=field)
    Xt = _lib.as_mat(Xte, keep)
    l1 = np.empty((Xte.shape[0], 2), dtype=np.int32, order="F")
    l2 = np.empty_like(l1)
    ctx.check(ctx.lib.fpls_lda_predict(ctx.h, h1, ctypes.byref(Xt), l1.ctypes.data, None))
    ctx.check(ctx.lib.fpls_lda_predict(ctx.h, h2, ctypes.byref(Xt), l2.ctypes.data, None))
    np.testing.assert_array_equal(l1, l2)
    # pls(..., Xtest) with the head: fit and prediction in one call
    h3 = ctypes.c_void_p()
    l3 = np.empty_like(l1)
    ctx.check(ctx.lib.fpls_plssvd_lda_fit_predict(ctx.h, ctypes.byref(Xm), ctypes.byref(yv), 9,
                                                  ncp, 2, ctypes.byref(opts), 1e-8,
                                                  ctypes.byref(Xt), l3.ctypes.data, None,
                                                  ctypes.byref(h3)))
    np.testing.assert_array_equal(l3, l1)
    m3 = F.FastPLS(ctx, h3)
    np.testing.assert_array_equal(m3.read("R", (p, k)), m1.read("R", (p, k)))
    m1.free()
    m2.free()
    m3.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("scaling", ["centering", "autoscaling"])
@pytest.mark.parametrize("C", [9, 150])
def test_label_fit_offset_columns(ctx, dtype, scaling, C):
    """Columns centred near 1000 with unit-scale spread: the one-pass
    labelled cross-product shifts X by its first row before the products
    (C <= 127: the MFMA class cross-product; C > 127: the segmented class
    sums).  mX, vX and the coefficients match the FP64 oracle on the same
    (rounded) data: FP64 inputs to 1e-9, FP32 inputs mX / vX to 1e-7 and B to
    1e-4 (FP32 products of exactly shifted values)."""
    import fastpls_amd as F
    rng = np.random.default_rng(70 + C + (scaling == "autoscaling"))
    n, p, k = 3000, 60, 5
    mu = rng.standard_normal((C, p))
    y = rng.integers(1, C + 1, n).astype(np.int32)
    y[:C] = np.arange(1, C + 1)
    X = (1000.0 + rng.uniform(-50, 50, p)) + mu[y - 1] + rng.standard_normal((n, p))
    X = X.astype(dtype)
    m = F.pls(X, y, method="plssvd", ncomp=[k], scaling=scaling, seed=5)
    sc = 1 if scaling == "centering" else 2
    ref = O.pls_model1(X.astype(np.float64), O.transformy(y), [k], sc, False, "rsvd", 10, 1, 5)
    tight = dtype == np.float64
    assert rel_err(m["mX"], ref["mX"]) < (1e-12 if tight else 1e-7)
    assert rel_err(m["vX"], ref["vX"]) < (1e-9 if tight else 1e-7)
    assert rel_err(m["B"], ref["B"]) < (1e-9 if tight else 1e-4)
    assert rel_err(align_signs(m["R"], ref["R"]), ref["R"]) < (1e-9 if tight else 1e-4)


@pytest.mark.parametrize("dtype,scaling,tol", [(np.float64, "centering", 1e-9),
                                                (np.float64, "autoscaling", 1e-9),
                                                (np.float32, "centering", 1e-4)])
def test_plssvd_xprod_route_matches_oracle(ctx, dtype, scaling, tol):
    """pls() takes the matrix-free route of .should_use_xprod_default for a
    dense response with q >= 100 and ncomp <= 10 (R/main.R:6080-6108): S is
    never formed, xprod_mode is "implicit", and the fit equals the oracle's
    pls_model1_rsvd_xprod_precision (implicit64) and the device's own
    formed-S fit."""
    import fastpls_amd as F
    rng = np.random.default_rng(23)
    n, p, q = 400, 250, 120
    X = _lowrank(rng, n, p, 20, 0.05) + 1.5
    Y = X[:, :q] @ rng.standard_normal((q, q)) * 0.1 + 0.1 * rng.standard_normal((n, q))
    ncomp = [3, 6]
    sc = {"centering": 1, "autoscaling": 2}[scaling]
    m = F.pls(X.astype(dtype), Y.astype(dtype), method="plssvd", ncomp=ncomp, fit=True,
              scaling=scaling, seed=4)
    assert m["xprod_mode"] == "implicit" and m["xprod_default"]
    ref = O.pls_model1(X.astype(dtype), Y.astype(dtype), ncomp, sc, True, "rsvd", 10, 1, 4,
                       xprod=True)
    assert sin_theta(m["R"], ref["R"]) < tol
    assert rel_err(m["B"], ref["B"]) < tol
    assert rel_err(m["Yfit"], ref["Yfit"]) < tol
    formed = F.pls(X.astype(dtype), Y.astype(dtype), method="plssvd", ncomp=ncomp, fit=True,
                   scaling=scaling, seed=4, xprod=False)
    assert formed["xprod_mode"] == "materialized"
    assert rel_err(formed["B"], m["B"]) < tol


@pytest.mark.parametrize("n", [3000, 40000])
def test_plssvd_fp32_wide_gram_fits(ctx, n):
    """FP32 dense-response PLSSVD with ncomp = 150: the training-score Gram
    T^T T (150 x 150) takes the tiled TN product that computes only the tiles
    on and above the diagonal and mirrors the rest in its reduction; at
    n = 40,000 rows the scores T = Xc R (150 columns) run as 128-column
    passes of the streaming projection.  Fitted values and coefficients
    against the FP64 oracle (exact SVD on both sides) to the FP32 input
    rounding (rel 1e-3; a wrong mirror or slice gives O(1))."""
    import fastpls_amd as F
    rng = np.random.default_rng(23)
    p, m, k = 300, 200, 150
    X = (rng.standard_normal((n, p)) * np.linspace(2.0, 0.5, p)).astype(np.float32)
    W = rng.standard_normal((p, m)) / np.sqrt(p)
    Y = (X.astype(np.float64) @ W + 0.1 * rng.standard_normal((n, m))).astype(np.float32)
    mod = F.pls(X, Y, method="plssvd", ncomp=k, fit=True, svd_method="exact")
    ref = O.pls_model1(X.astype(np.float64), Y.astype(np.float64), [k], 1, True, "exact")
    assert rel_err(mod["Yfit"], ref["Yfit"]) < 1e-3
    assert rel_err(mod["B"], ref["B"]) < 1e-3
