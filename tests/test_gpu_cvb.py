"""GPU parity: batched cross-validation (fpls_cv_batched) and the device
Gaussian streams (fpls_gaussian_streams).

Batched CV is checked against the oracle's per-fit SIMPLS (pls_model2_fast on
the fit's training rows with the fit's seed, pls_predict on its test rows)
and against the fold-by-fold GPU driver.  X is FP32, so the two big
contractions run on FP32 MFMA (FP32 sums over slabs of p, FP64 across
slabs); predictions are held to 1e-4 relative to the FP64 oracle (1e-3 for
log-normal data and for split-BF16x3), class labels must be equal except
rows whose oracle top-2 margin is below 1e-5 of the score scale (near-ties,
counted and reported).
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _ulp_diff(a, b):
    ia = np.asarray(a, dtype=np.float64).view(np.int64)
    ib = np.asarray(b, dtype=np.float64).view(np.int64)
    same_sign = np.sign(a) == np.sign(b)
    d = np.abs(ia - ib)
    return np.where(same_sign, d, np.iinfo(np.int64).max)


@pytest.mark.parametrize("length", [1, 2, 3, 311, 20011])
def test_gaussian_streams_match_libstdcxx(ctx, length):
    import ctypes
    seeds = np.array([1, 2, 7, 12345, 4294967295], dtype=np.uint32)
    out = np.zeros((len(seeds), length))
    ctx.check(ctx.lib.fpls_gaussian_streams(ctx.h, seeds.ctypes.data, len(seeds), length,
                                            out.ctypes.data))
    ref = np.stack([O.gaussian_matrix(length, 1, int(s))[:, 0] for s in seeds])
    d = _ulp_diff(out, ref)
    # integer stream, uniforms, acceptance and pairing are exact; only log's
    # rounding may differ (glibc's is not correctly rounded everywhere)
    assert d.max() <= 4, f"max ulp difference {d.max()}"
    if length >= 311:
        frac = np.mean(d > 0)
        assert frac < 0.01, f"{frac:.4%} of draws differ"


def _masks(rng, n, F, p_train=0.7, p_test=0.2):
    M = np.zeros((n, F), dtype=np.int8)
    for f in range(F):
        r = rng.random(n)
        M[r < p_train, f] = 1
        M[r > 1.0 - p_test, f] = 2
    return M


def _oracle_fits(X, Yd, masks, nc, seeds, scaling=1):
    """Per-fit oracle predictions of the test rows: list of (te, Ypred)."""
    out = []
    for f in range(masks.shape[1]):
        tr = masks[:, f] == 1
        te = masks[:, f] == 2
        mod = O.pls_model2_fast(X[tr].astype(np.float64), Yd[tr], nc, scaling, False, "rsvd", 10,
                                1, int(seeds[f]))
        out.append((te, O.pls_predict(mod, X[te].astype(np.float64))))
    return out


def _near_tie_ok(pred, ref_scores, rel=1e-5):
    """Labels of pred vs argmax of ref_scores, excusing near-ties."""
    s = np.sort(ref_scores, axis=1)
    margin = s[:, -1] - s[:, -2]
    ref = np.argmax(ref_scores, axis=1) + 1
    bad = (pred != ref) & (margin > rel * max(np.abs(s).max(), 1e-300))
    return int(bad.sum()), int(((pred != ref) & ~bad).sum())


def test_cv_batched_regression_masks(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(21)
    n, p, m, nf = 160, 70, 3, 7
    X = (rng.standard_normal((n, p)) + 2.0).astype(np.float32)
    Y = X[:, :6].astype(np.float64) @ rng.standard_normal((6, m)) + 0.2 * rng.standard_normal((n, m))
    masks = _masks(rng, n, nf)
    nc = [1, 2, 4, 6]
    seeds = rng.integers(1, 2**31, nf)
    res = F.cv_batched(X, Y, masks, nc, seeds)
    assert (res["status"] == 1).all()
    for f, (te, yp) in enumerate(_oracle_fits(X, Y, masks, nc, seeds)):
        got = res["Ypred"][te, :, :, f]
        assert rel_err(got, yp) < 1e-4, (f, rel_err(got, yp))
        assert np.all(res["Ypred"][~te, :, :, f] == 0.0)


@pytest.mark.parametrize("scaling", ["centering", "none"])
def test_cv_batched_classification_masks(ctx, scaling):
    import fastpls_amd as F
    rng = np.random.default_rng(22)
    n, p, C, nf = 200, 90, 4, 9
    y = rng.integers(1, C + 1, n).astype(np.int32)
    X = (rng.standard_normal((n, p)) + 0.6 * y[:, None] * (np.arange(p) < 8)).astype(np.float32)
    masks = _masks(rng, n, nf)
    nc = [1, 2, 3, 5, 8]
    seeds = rng.integers(1, 2**31, nf)
    res = F.cv_batched(X, y, masks, nc, seeds, scaling=scaling, return_scores=True)
    Yd = np.zeros((n, C))
    Yd[np.arange(n), y - 1] = 1.0
    sc = {"centering": 1, "none": 3}[scaling]
    ties = 0
    for f, (te, yp) in enumerate(_oracle_fits(X, Yd, masks, nc, seeds, sc)):
        assert rel_err(res["Ypred"][te, :, :, f], yp) < 1e-4
        for s in range(len(nc)):
            bad, t = _near_tie_ok(res["class_pred"][te, s, f], yp[:, :, s])
            assert bad == 0
            ties += t
            assert res["n_correct"][s, f] == int(np.sum(res["class_pred"][te, s, f] == y[te]))
        assert np.all(res["class_pred"][~te, :, f] == 0)
    print("excused near-ties:", ties)


def test_cv_batched_logo_vs_oracle_and_folds(ctx):
    """The scaled-down nested-CV workload (log-normal blocks, 5 imbalanced
    classes, uneven groups) through both GPU engines and the oracle."""
    import fastpls_amd as F
    from tests.test_gpu_cv import _groups
    rng = np.random.default_rng(6)
    n, p = 300, 400
    y = rng.choice(np.arange(1, 6), n, p=[0.4, 0.25, 0.15, 0.12, 0.08]).astype(np.int32)
    shared = rng.standard_normal((n, p // 40)).repeat(40, axis=1)
    Z = np.sqrt(0.6) * shared + np.sqrt(0.4) * rng.standard_normal((n, p))
    Z[:, :20] += 0.8 * y[:, None]
    X = np.exp(Z).astype(np.float32)
    g = _groups(rng, n, 6)
    ncomp = list(range(1, 9))
    bat = F.pls_cv_predict(X, y, g, ncomp=ncomp, method="simpls", return_scores=True,
                           engine="batched")
    seq = F.pls_cv_predict(X, y, g, ncomp=ncomp, method="simpls", return_scores=True,
                           engine="folds")
    ref = O.pls_cv_predict(X.astype(np.float64), y, g, ncomp, method="simpls",
                           classification=True, n_classes=5)
    assert bat["engine"] == "batched"
    np.testing.assert_array_equal(bat["fold"], ref["fold"])
    np.testing.assert_array_equal(bat["status"], ref["status"])
    assert rel_err(bat["Ypred"], ref["Ypred"]) < 1e-3
    assert rel_err(bat["Ypred"], seq["Ypred"]) < 1e-3
    s = np.sort(ref["Ypred"], axis=1)
    margin = s[:, -1, :] - s[:, -2, :]
    bad = (bat["class_pred"] != ref["class_pred"]) & (margin > 1e-5 * np.abs(s).max())
    assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"


def test_cv_batched_degenerate_folds(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(23)
    n, p, C = 60, 20, 3
    y = np.repeat([1, 2, 3], 20).astype(np.int32)
    X = (rng.standard_normal((n, p)) + y[:, None] * 0.3).astype(np.float32)
    M = np.zeros((n, 4), dtype=np.int8)
    M[:40, 0] = 1
    M[40:, 0] = 2                      # ok
    M[:, 1] = 1                        # no test rows -> 2
    M[:, 2] = 2                        # no training rows -> 3
    M[:20, 3] = 1
    M[20:30, 3] = 2                    # one training class -> 4
    res = F.cv_batched(X, y, M, [1, 2], seeds=[1, 2, 3, 4], return_scores=True)
    np.testing.assert_array_equal(res["status"], [1, 2, 3, 4])
    assert np.all(res["class_pred"][:, :, 1] == 0)
    np.testing.assert_array_equal(res["class_pred"][:, 0, 2], y)          # observed labels
    np.testing.assert_array_equal(res["class_pred"][20:30, :, 3], 1)      # the one class
    assert np.all(res["class_pred"][:20, :, 3] == 0)
    np.testing.assert_array_equal(res["n_correct"][:, 2], [n, n])
    np.testing.assert_array_equal(res["n_correct"][:, 3], [0, 0])
    onehot = np.zeros((n, C))
    onehot[np.arange(n), y - 1] = 1.0
    np.testing.assert_array_equal(res["Ypred"][:, :, 0, 2], onehot)


def test_cv_batched_replay_and_new_inputs(ctx):
    """The first call records and captures the launch sequence; later calls
    with the same shapes replay the graph on new masks / seeds."""
    import fastpls_amd as F
    rng = np.random.default_rng(24)
    n, p, C, nf = 180, 64, 3, 6
    y = rng.integers(1, C + 1, n).astype(np.int32)
    X = (rng.standard_normal((n, p)) + 0.5 * y[:, None] * (np.arange(p) < 6)).astype(np.float32)
    import torch
    from fastpls_amd._lib import device_colmajor
    Xd = device_colmajor(X)          # resident X: the plan is keyed on its address
    yd = torch.from_numpy(y).to("cuda:0")
    nc = [1, 2, 5]
    m1 = _masks(rng, n, nf)
    s1 = rng.integers(1, 2**31, nf)
    a = F.cv_batched(Xd, yd, m1, nc, s1, n_classes=C, return_scores=True)
    b = F.cv_batched(Xd, yd, m1, nc, s1, n_classes=C, return_scores=True)
    np.testing.assert_array_equal(a["Ypred"], b["Ypred"])   # replay is bitwise identical
    np.testing.assert_array_equal(a["class_pred"], b["class_pred"])
    m2 = _masks(rng, n, nf, 0.6, 0.3)
    s2 = rng.integers(1, 2**31, nf)
    c = F.cv_batched(Xd, yd, m2, nc, s2, n_classes=C, return_scores=True)
    Yd = np.zeros((n, C))
    Yd[np.arange(n), y - 1] = 1.0
    for f, (te, yp) in enumerate(_oracle_fits(X, Yd, m2, nc, s2)):
        assert rel_err(c["Ypred"][te, :, :, f], yp) < 1e-4


def test_cv_batched_bf16x3(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(25)
    n, p, C, nf = 220, 150, 5, 8
    y = rng.integers(1, C + 1, n).astype(np.int32)
    X = (rng.standard_normal((n, p)) + 0.4 * y[:, None] * (np.arange(p) < 10)).astype(np.float32)
    masks = _masks(rng, n, nf)
    nc = [1, 3, 6]
    seeds = rng.integers(1, 2**31, nf)
    res = F.cv_batched(X, y, masks, nc, seeds, precision="bf16x3", return_scores=True)
    Yd = np.zeros((n, C))
    Yd[np.arange(n), y - 1] = 1.0
    for f, (te, yp) in enumerate(_oracle_fits(X, Yd, masks, nc, seeds)):
        assert rel_err(res["Ypred"][te, :, :, f], yp) < 1e-3
        for s in range(len(nc)):
            bad, _ = _near_tie_ok(res["class_pred"][te, s, f], yp[:, :, s], rel=1e-3)
            assert bad == 0


def test_pls_double_cv_matches_fold_engine(ctx):
    """One run of the double CV: inner accuracies equal those of the
    fold-by-fold driver on each outer-training set (same seeds), and the outer
    predictions equal a fit of the outer-training rows at the chosen ncomp."""
    import fastpls_amd as F
    rng = np.random.default_rng(26)
    n, p, C = 150, 80, 3
    y = rng.choice([1, 2, 3], n, p=[0.5, 0.3, 0.2]).astype(np.int32)
    X = (rng.standard_normal((n, p)) + 0.7 * y[:, None] * (np.arange(p) < 5)).astype(np.float32)
    J, I = 3, 3
    outer = rng.integers(1, J + 1, n)
    inner = rng.integers(1, I + 1, (n, J))
    nc = [1, 2, 3, 4]
    seed, run = 5, 1
    res = F.pls_double_cv(X, y, nc, outer, inner, seed=seed, run=run)
    for j in range(J):
        tr = outer != j + 1
        ref = F.pls_cv_predict(X[tr], y[tr], inner[tr, j], ncomp=nc,
                               seed=seed + 1000 * run + (j + 1), engine="folds")
        acc = np.asarray(ref["metrics"]["metric_value"])
        assert np.max(np.abs(acc - res["inner_accuracy"][j])) <= 2.0 / tr.sum() + 1e-12
