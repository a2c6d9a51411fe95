"""GPU parity: argmax-head class ranking (fpls_class_predict_topk,
fpls_class_topk_scores) against the oracle's restatement of
fill_topk_from_yblock / pls_class_predict_topk (src/fastPLS.cpp:3737-4113),
and predict()'s routing mirroring tests/testthat/test-class-bias-top5.R.

The ranking of a given score cube is integer/index work on FP64 values that
both sides form the same way (score + offset in FP64), so indices and scores
are compared bit-exactly.  Through a fitted model the scores come out of the
device GEMMs: scores within 1e-10 of the oracle's ranking of the same
device prediction cube, and labels equal except at near-ties."""
import numpy as np
import pytest

from oracle import fastpls_oracle as O

pytestmark = pytest.mark.gpu


def _cube(rng, w, m, L, ties=True, specials=False):
    if ties:
        c = rng.integers(-4, 5, size=(w, m, L)).astype(np.float64) * 0.5
    else:
        c = rng.standard_normal((w, m, L))
    if specials:
        mask = rng.random(c.shape)
        c[mask < 0.05] = np.nan
        c[(mask >= 0.05) & (mask < 0.08)] = -np.inf
        c[(mask >= 0.08) & (mask < 0.09)] = np.inf
    return np.asfortranarray(c)


def _rank(ctx, cube, bias, bias_cols, top):
    import ctypes
    import fastpls_amd as F
    from fastpls_amd._lib import as_mat
    w, m, L = cube.shape
    keep = []
    S = as_mat(np.asfortranarray(cube).reshape(w, m * L, order="F"), keep)
    k = max(1, min(top, m))
    ti = np.empty((w, k, L), dtype=np.int32, order="F")
    ts = np.empty((w, k, L), dtype=np.float64, order="F")
    b = None if bias is None else np.asfortranarray(bias, dtype=np.float64)
    ctx.check(ctx.lib.fpls_class_topk_scores(ctx.h, ctypes.byref(S), m, L,
                                             None if b is None else b.ctypes.data, bias_cols,
                                             top, ti.ctypes.data, ts.ctypes.data))
    return ti, ts


@pytest.mark.parametrize("top", [1, 3, 4, 5, 16, 17, 32, 33, 60])
def test_score_cube_ranking_bit_exact(ctx, top):
    """Every register-list width (1, 4, 16, 32) and the rank-by-count path
    (top > 32), with ties, NaN, +-inf and per-slice offsets."""
    rng = np.random.default_rng(100 + top)
    cube = _cube(rng, 700, 60, 2, ties=True, specials=True)
    bias = rng.integers(-2, 3, size=(60, 2)).astype(np.float64) * 0.25
    ti, ts = _rank(ctx, cube, bias, 2, top)
    ri, rs = O.class_bias_topk_from_cube(cube, bias, top)
    np.testing.assert_array_equal(ti, ri)
    np.testing.assert_array_equal(ts, rs)


@pytest.mark.parametrize("w,m,top", [(37, 3000, 5), (5, 2500, 40), (3, 70, 70), (2, 1, 3),
                                     (300, 17, 16)])
def test_score_cube_shapes(ctx, w, m, top):
    """Many classes (segmented rows merged through LDS; rank chunks past
    2048 classes), top >= m (clamped), a single class, ragged widths."""
    rng = np.random.default_rng(w * 7 + m)
    cube = _cube(rng, w, m, 1, ties=True)
    ti, ts = _rank(ctx, cube, None, 0, top)
    ri, rs = O.class_bias_topk_from_cube(cube, None, top)
    np.testing.assert_array_equal(ti, ri)
    np.testing.assert_array_equal(ts, rs)


def test_score_cube_broadcast_bias_fp32_and_device_input(ctx):
    import torch
    import fastpls_amd as F
    rng = np.random.default_rng(8)
    cube = _cube(rng, 500, 9, 3, ties=False).astype(np.float32)
    b1 = rng.standard_normal(9)
    # one bias column broadcast over the slices
    ti, ts = _rank(ctx, cube, b1[:, None], 1, 4)
    ri, rs = O.class_bias_topk_from_cube(cube.astype(np.float64), b1, 4)
    np.testing.assert_array_equal(ti, ri)
    np.testing.assert_array_equal(ts, rs)
    # the Python mirror with a device tensor cube (row-major torch layout)
    dev = torch.from_numpy(np.ascontiguousarray(cube)).cuda()
    got = F.class_topk_from_score_cube(dev, [str(c) for c in range(9)], b1, top=4)
    np.testing.assert_array_equal(got["Ypred_top"], ri)
    np.testing.assert_array_equal(got["Ypred_top_score"], rs)
    np.testing.assert_array_equal(got["Ypred_index"], ri[:, 0, :])
    torch.cuda.synchronize()


def test_rank_kernel_many_rows(ctx):
    """More rows than one grid dimension of the rank-by-count launch."""
    rng = np.random.default_rng(9)
    cube = _cube(rng, 70000, 40, 1, ties=True)
    ti, ts = _rank(ctx, cube, None, 0, 35)
    ri, rs = O.class_topk_from_score_cube(cube, None, 35)   # stable argsort ranks
    np.testing.assert_array_equal(ti, ri)
    np.testing.assert_array_equal(ts, rs)


def _model_cube(ctx, m, Xt):
    import ctypes
    from fastpls_amd._lib import as_mat
    inf = m.info()
    keep = []
    X = as_mat(Xt, keep)
    yp = np.empty((Xt.shape[0], inf.m, inf.n_ncomp), dtype=np.float64, order="F")
    ctx.check(ctx.lib.fpls_predict(ctx.h, m.handle, ctypes.byref(X), yp.ctypes.data))
    return yp


@pytest.mark.parametrize("method", ["plssvd", "simpls"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_model_topk_matches_oracle(ctx, method, dtype):
    import fastpls_amd as F
    rng = np.random.default_rng(21)
    C = 7
    X = rng.standard_normal((400, 30)).astype(dtype)
    y = rng.integers(1, C + 1, 400)
    Xt = rng.standard_normal((150, 30)).astype(dtype)
    m = F.pls(X, y, method=method, ncomp=[1, 3, 5], seed=3)
    bias = rng.standard_normal((C, 3)) * 0.1
    res = F.class_predict_topk(m, Xt, bias, top=4)
    # the oracle's ranking of the device's own dense prediction cube
    cube = _model_cube(ctx, m, Xt)
    mY = m.read("mY", (C,))
    ri, rs = O.class_bias_topk_from_cube(cube - mY[None, :, None], O.class_bias_offsets(mY, bias, 3), 4)
    np.testing.assert_allclose(res["Ypred_top_score"], rs, rtol=0, atol=1e-10)
    gap = np.abs(np.diff(rs, axis=1)).min(axis=1)      # (w, slices); near-ties excused
    ok = np.all(res["Ypred_top"] == ri, axis=1)
    assert np.all(ok | (gap < 1e-9))
    # and the full oracle pipeline: fit + compact predict (FP64 budget)
    Xd = X.astype(np.float64)
    if method == "plssvd":
        mod = O.pls_model1(Xd, O.transformy(y), [1, 3, 5], 1, False, "rsvd", 10, 1, 3)
    else:
        mod = O.pls_model2_fast(Xd, O.transformy(y), [1, 3, 5], 1, False, "rsvd", 10, 1, 3)
    oi, os_ = O.pls_class_predict_topk(mod, Xt.astype(np.float64), bias, 4)
    tol = 1e-8 if dtype == np.float64 else 1e-4
    np.testing.assert_allclose(res["Ypred_top_score"], os_, rtol=0,
                               atol=tol * float(np.abs(os_).max()))
    # the first-ranked class equals the oracle's except where its FP64 top-2
    # gap is within the contraction's error (documented near-ties)
    g12 = os_[:, 0, :] - os_[:, 1, :]
    bad = (res["Ypred_top"][:, 0, :] != oi[:, 0, :]) & (g12 > tol * float(np.abs(os_).max()))
    assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"


def test_predict_routing_like_reference_tests(ctx):
    """tests/testthat/test-class-bias-top5.R:1-23 and :60-100 with this
    library's draws: the ranked fast path keeps the argmax labels of the
    raw-score path; top5 gives w x 5 (x slices), capped at the class count."""
    import fastpls_amd as F
    rng = np.random.default_rng(20260511)
    X = rng.standard_normal((96, 10))
    y = rng.choice([f"C{i}" for i in range(1, 7)], 96)
    idx = rng.choice(96, 18, replace=False)
    tr = np.setdiff1d(np.arange(96), idx)
    fit = F.pls(X[tr], y[tr], ncomp=[1, 2, 3], method="plssvd", seed=123)
    fast = F.predict(fit, X[idx])
    full = F.predict(fit, X[idx], raw_scores=True)
    np.testing.assert_array_equal(fast["Ypred"], full["Ypred"])
    np.testing.assert_array_equal(fast["Ypred_index"], np.argmax(full["Yscore"], axis=1) + 1)
    top5 = F.predict(fit, X[idx], top5=True)
    assert top5["Ypred_top"].shape == (18, 5, 3)
    assert top5["Ypred_top_score"].shape == (18, 5, 3)
    assert set(top5["Ypred_top_label"].ravel()) <= set(y)
    # raw-score ranking of the same cube agrees with the fast path's ranks
    np.testing.assert_array_equal(top5["Ypred_top"],
                                  F.predict(fit, X[idx], top=5, raw_scores=True)["Ypred_top"])
    y4 = rng.choice(list("ABCD"), 72)
    fit4 = F.pls(X[:72, :8], y4, ncomp=2, method="plssvd", seed=321)
    assert F.predict(fit4, X[:72, :8], top5=True)["Ypred_top"].shape == (72, 4, 1)
    with pytest.raises(ValueError, match="top must be a positive integer"):
        F.predict(fit4, X[:5, :8], top=0)


def test_lda_top_k_from_scores(ctx):
    """predict(top > 1) on the LDA head ranks its discriminant scores
    (.class_topk_from_score_cube on lda_scores, R/main.R:2609-2627)."""
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=8, per_class_train=60,
                                                 per_class_test=20, p=40, seed=4)
    m = F.pls(Xtr, ytr, method="plssvd", ncomp=[3, 6], classifier="lda")
    base = F.predict(m, Xte)
    r = F.predict(m, Xte, top=3, raw_scores=True)
    ri, rs = O.class_topk_from_score_cube(r["LDA_scores"], None, 3)
    np.testing.assert_array_equal(r["Ypred_top"], ri)
    np.testing.assert_array_equal(r["Ypred_top_score"], rs)
    np.testing.assert_array_equal(r["Ypred_index"], base["Ypred_index"])
