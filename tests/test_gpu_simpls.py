"""GPU parity: SIMPLS-fast fit and prediction against the FP64 oracle.

FP64 inputs run FP64 kernels end to end: R, Q (sign-aligned), B, Yfit and
predictions agree to 1e-8 relative.  FP32 inputs: 1e-4.
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import label_mismatches_excused, align_signs, rel_err

pytestmark = pytest.mark.gpu


def test_simpls_regression_workload_fp64(ctx):
    """The SIMPLS regression workload at full size (n=200, p=2000, m=3, ncomp=10)."""
    import fastpls_amd as F
    from fastpls_amd import synth
    X, Y, Xt, Yt = synth.regression_lowrank()
    m = F.pls(X, Y, Xt, method="simpls", ncomp=10)
    ref = O.pls_model2_fast(X, Y, [10])
    assert m["n_components"] == ref["n_components"] == 10
    assert rel_err(align_signs(m["R"], ref["R"]), ref["R"]) < 1e-8
    assert rel_err(m["B"], ref["B"]) < 1e-8
    yp = m["Ypred"]
    assert rel_err(yp, O.pls_predict(ref, Xt)) < 1e-8


@pytest.mark.parametrize("ncomp", [[1, 3, 6], [4]])
def test_simpls_fit_r2_fp64(ctx, ncomp):
    import fastpls_amd as F
    rng = np.random.default_rng(1)
    X = rng.standard_normal((80, 30)) + 0.5
    Y = X[:, :4] @ rng.standard_normal((4, 3)) + 0.1 * rng.standard_normal((80, 3))
    m = F.pls(X, Y, method="simpls", ncomp=ncomp, fit=True)
    ref = O.pls_model2_fast(X, Y, ncomp, fit=True)
    assert rel_err(m["B"], ref["B"]) < 1e-8
    assert rel_err(m["Yfit"], ref["Yfit"]) < 1e-8
    np.testing.assert_allclose(m["R2Y"], ref["R2Y"], rtol=1e-8)


def test_simpls_pls1_exact_branch_fp64(ctx):
    """m == 1: the reference takes the exact SVD of S (min dim < 6)."""
    import fastpls_amd as F
    rng = np.random.default_rng(2)
    X = rng.standard_normal((60, 25))
    y = X[:, 0] - 2 * X[:, 3] + 0.1 * rng.standard_normal(60)
    m = F.pls(X, y.reshape(-1, 1), method="simpls", ncomp=[1, 2, 5])
    ref = O.pls_model2_fast(X, y.reshape(-1, 1), [1, 2, 5])
    assert rel_err(m["B"], ref["B"]) < 1e-9


def test_simpls_xtx_branch_and_autoscaling_fp64(ctx):
    """K >= 20, p <= 512, n >= 8p: the X^T X shortcut branch."""
    import fastpls_amd as F
    rng = np.random.default_rng(4)
    X = rng.standard_normal((400, 40)) * rng.uniform(0.5, 2.0, 40) + 1.0
    Y = X[:, :5] @ rng.standard_normal((5, 4)) + 0.3 * rng.standard_normal((400, 4))
    ncomp = [5, 20, 22]
    m = F.pls(X, Y, method="simpls", ncomp=ncomp, scaling="autoscaling")
    ref = O.pls_model2_fast(X, Y, ncomp, scaling=2)
    assert rel_err(m["B"], ref["B"]) < 1e-7


def test_simpls_classification_fp32_labels(ctx):
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=5, per_class_train=200,
                                                 per_class_test=50, p=300, seed=9)
    m = F.pls(Xtr, ytr, Xte, method="simpls", ncomp=[2, 4])
    ref = O.pls_model2_fast(Xtr.astype(np.float64), O.transformy(ytr), [2, 4])
    assert rel_err(m["B"], ref["B"]) < 1e-4
    yp_ref = O.pls_predict(ref, Xte.astype(np.float64))
    lab_ref = np.argmax(yp_ref, axis=1) + 1
    # FP32 contractions: rows whose FP64 and device top-2 margins are under
    # 1e-4 are ties (the device's from its own ranked top-2 scores)
    r2 = F.predict(m, Xte, top=2)
    np.testing.assert_array_equal(r2["Ypred_top"][:, 0, :], m["Ypred_index"])
    for a in range(2):
        dev_gap = r2["Ypred_top_score"][:, 0, a] - r2["Ypred_top_score"][:, 1, a]
        _, unexcused = label_mismatches_excused(m["Ypred_index"][:, a], lab_ref[:, a],
                                                yp_ref[:, :, a], 1e-4, dev_gap)
        assert not unexcused, unexcused
