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


def test_simpls_rsvd_workload_full_shape_fp64(ctx):
    """SIMPLS with ncomp = 50 at the full shape of the rSVD-of-X^T Y workload
    (n = 20,000, p = 50,000, m = 100; X rank-80 signal s_i = 50 0.93^i under
    N(0, 1) noise, Y = Y0c G / 10), X resident on the device in FP64 so the
    device runs its FP64 kernels; the oracle draws the same seeded Gaussian
    sketches.  R (sign-aligned), B and the fitted values agree to 1e-8."""
    import torch
    import fastpls_amd as F
    n, p, m0, m, k, r = 20000, 50000, 20, 100, 50, 80
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(20505)
    U = torch.linalg.qr(torch.randn(n, r, generator=g, device=dev, dtype=torch.float64))[0]
    V = torch.linalg.qr(torch.randn(p, r, generator=g, device=dev, dtype=torch.float64))[0]
    s = 50.0 * 0.93 ** torch.arange(r, device=dev, dtype=torch.float64)
    Xt = torch.randn(p, n, generator=g, device=dev, dtype=torch.float64)
    for j0 in range(0, p, 5000):
        Xt[j0:j0 + 5000] += V[j0:j0 + 5000] @ (U * s).T
    X = Xt.t()  # n x p, column-major on the device
    Bt = torch.randn(p, m0, generator=g, device=dev, dtype=torch.float64) / p ** 0.5
    Y0 = X @ Bt + torch.randn(n, m0, generator=g, device=dev, dtype=torch.float64)
    Y0 -= Y0.mean(0)
    Y = ((Y0 @ torch.randn(m0, m, generator=g, device=dev, dtype=torch.float64)) / 10.0)
    Y = Y.t().contiguous().t()
    del U, V, Bt, Y0
    mod = F.pls(X, Y, method="simpls", ncomp=k, fit=True)
    Xh = X.cpu().numpy()
    Yh = Y.cpu().numpy()
    del X, Xt, Y
    torch.cuda.empty_cache()
    ref = O.pls_model2_fast(Xh, Yh, [k], fit=True)
    assert mod["n_components"] == ref["n_components"] == k
    assert rel_err(align_signs(mod["R"], ref["R"]), ref["R"]) < 1e-8
    assert rel_err(mod["B"], ref["B"]) < 1e-8
    assert rel_err(mod["Yfit"], ref["Yfit"]) < 1e-8


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
