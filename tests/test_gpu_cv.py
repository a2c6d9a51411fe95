"""GPU parity: cross-validated prediction (leave-one-group-out) against the oracle.

FP64 regression: predictions to 1e-8 relative.  Classification: class
predictions equal except rows whose oracle top-2 score margin is below 1e-6
(documented near-ties); status and fold bookkeeping bit-exact.
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _groups(rng, n, k):
    # uneven group sizes (Dirichlet-like) as explicit fold labels
    sizes = rng.dirichlet(np.full(k, 5.0)) * n
    sizes = np.maximum(1, np.round(sizes).astype(int))
    sizes[-1] = n - sizes[:-1].sum()
    g = np.repeat(np.arange(1, k + 1) * 7, sizes)  # arbitrary group ids
    rng.shuffle(g)
    return g


@pytest.mark.parametrize("method", ["simpls", "plssvd"])
def test_cv_regression_fp64(ctx, method):
    import fastpls_amd as F
    rng = np.random.default_rng(5)
    n, p, m = 120, 40, 3
    X = rng.standard_normal((n, p))
    Y = X[:, :5] @ rng.standard_normal((5, m)) + 0.1 * rng.standard_normal((n, m))
    g = _groups(rng, n, 5)
    res = F.pls_cv_predict(X, Y, g, ncomp=[1, 2, 3], method=method)
    ref = O.pls_cv_predict(X, Y, g, [1, 2, 3], method=method)
    np.testing.assert_array_equal(res["fold"], ref["fold"])
    np.testing.assert_array_equal(res["status"], ref["status"])
    assert rel_err(res["Ypred"], ref["Ypred"]) < 1e-8


def test_cv_classification_simpls_fp32(ctx):
    """Scaled-down nested-CV workload: log-normal blocks, 5 imbalanced classes."""
    import fastpls_amd as F
    rng = np.random.default_rng(6)
    n, p = 300, 400
    y = rng.choice(np.arange(1, 6), n, p=[0.4, 0.25, 0.15, 0.12, 0.08]).astype(np.int32)
    shared = rng.standard_normal((n, p // 40)).repeat(40, axis=1)
    Z = np.sqrt(0.6) * shared + np.sqrt(0.4) * rng.standard_normal((n, p))
    Z[:, :20] += 0.8 * y[:, None]
    X = np.exp(Z).astype(np.float32)
    g = _groups(rng, n, 6)
    ncomp = list(range(1, 9))
    res = F.pls_cv_predict(X, y, g, ncomp=ncomp, method="simpls", return_scores=True)
    ref = O.pls_cv_predict(X.astype(np.float64), y, g, ncomp, method="simpls",
                           classification=True, n_classes=5)
    np.testing.assert_array_equal(res["status"], ref["status"])
    s = np.sort(ref["Ypred"], axis=1)
    margin = s[:, -1, :] - s[:, -2, :]
    bad = (res["class_pred"] != ref["class_pred"]) & (margin > 1e-6 * np.abs(s).max())
    assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"
    assert rel_err(res["Ypred"], ref["Ypred"]) < 1e-3


@pytest.mark.parametrize("classifier", ["lda", "cknn"])
def test_cv_classifier_heads(ctx, classifier):
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, _, _ = synth.classes_on_sphere(n_classes=6, per_class_train=60, per_class_test=1,
                                             p=30, seed=31, dtype=np.float64)
    rng = np.random.default_rng(8)
    g = _groups(rng, ytr.size, 4)
    ncomp = [2, 4, 5]
    res = F.pls_cv_predict(Xtr, ytr, g, ncomp=ncomp, method="plssvd", classifier=classifier,
                           candidate_knn_k=5, candidate_top_m=4)
    ref = O.pls_cv_predict(Xtr, ytr, g, ncomp, method="plssvd", classification=True,
                           n_classes=6, classifier=classifier, knn_k=5, top_m=4)
    # labels equal the oracle's except documented near-ties (FP64 margin)
    bad = (res["class_pred"] != ref["class_pred"]) & (ref["margin"] > 1e-7)
    assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"


def test_cv_degenerate_fold_status(ctx):
    """A fold whose training rows hold a single class gets status 4."""
    import fastpls_amd as F
    rng = np.random.default_rng(9)
    X = rng.standard_normal((30, 8))
    y = np.array([1] * 20 + [2] * 10, dtype=np.int32)
    g = np.array([1] * 20 + [2] * 10)   # fold 1 trains on class 2 only, fold 2 on class 1
    res = F.pls_cv_predict(X, y, g, ncomp=[2], method="simpls")
    ref = O.pls_cv_predict(X, y, g, [2], classification=True, n_classes=2)
    np.testing.assert_array_equal(res["status"], [4, 4])
    np.testing.assert_array_equal(res["class_pred"], ref["class_pred"])


def test_optim_pls_cv_picks_best(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(10)
    X = rng.standard_normal((80, 20))
    y = (X[:, 0] + 0.5 * X[:, 1] + 0.3 * rng.standard_normal(80)).reshape(-1, 1)
    res = F.optim_pls_cv(X, y, ncomp=[1, 2, 3, 4], constrain=np.arange(80) % 8)
    assert res["best_metric_name"] == "q2"
    assert res["optim_comp"] == res["ncomp"][int(np.nanargmax(res["metrics"]["metric_value"]))]
