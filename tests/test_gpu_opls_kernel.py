"""GPU parity: OPLS filter and kernel matrices (fpls_opls_filter,
fpls_opls_apply, fpls_kernel_matrix, fpls_center_kernel_*) against the FP64
oracle (src/fastPLS.cpp:4314-4483), and pls(method = "opls" / "kernelpls")
mirroring tests/testthat/test-kernel-opls.R.  Everything runs in FP64 on the
device: the filter and the kernels agree with the oracle to 1e-10 relative,
and the fitted predictions to 1e-8 (the rSVD of the inner SIMPLS fit)."""
import numpy as np
import pytest

from oracle import fastpls_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("scaling", ["centering", "autoscaling", "none"])
def test_opls_filter_matches_oracle(ctx, scaling):
    import fastpls_amd as F
    rng = np.random.default_rng(31)
    X = rng.standard_normal((300, 40)) * np.linspace(0.5, 3.0, 40) + 1.0
    Y = np.c_[X[:, :4] @ rng.standard_normal(4), X[:, 5] - X[:, 7]] + 0.2 * rng.standard_normal((300, 2))
    got = F.opls_filter(X, Y, north=3, scaling=scaling)
    sc = {"centering": 1, "autoscaling": 2, "none": 3}[scaling]
    ref = O.opls_filter(X, Y, 3, sc)
    assert got["north"] == ref["north"] == 3
    np.testing.assert_allclose(got["mX"], ref["mX"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(got["vX"], ref["vX"], rtol=1e-13)
    # w_o and p_o follow the sign of the singular vector w (LAPACK's choice
    # in the reference), together: compare sign-aligned columns; the
    # deflation t_o p_o^T and so the filtered X do not depend on it
    sg = np.sign(np.sum(got["W_orth"] * ref["W_orth"], axis=0))
    assert np.all(sg != 0)
    assert rel(got["W_orth"] * sg, ref["W_orth"]) < 1e-10
    assert rel(got["P_orth"] * sg, ref["P_orth"]) < 1e-10
    assert rel(got["X"], ref["X"]) < 1e-10
    # the stored filter on new rows, host and device outputs
    Xn = rng.standard_normal((50, 40))
    want = O.opls_apply_filter(Xn, ref["mX"], ref["vX"], ref["W_orth"], ref["P_orth"])
    assert rel(F.opls_apply_filter(Xn, got["mX"], got["vX"], got["W_orth"], got["P_orth"]),
               want) < 1e-10
    dev = F.opls_apply_filter(F.device_colmajor(Xn.astype(np.float32)), got["mX"], got["vX"],
                              got["W_orth"], got["P_orth"], out="device")
    want32 = O.opls_apply_filter(Xn.astype(np.float32).astype(np.float64), ref["mX"], ref["vX"],
                                 ref["W_orth"], ref["P_orth"])
    assert rel(dev.cpu().numpy(), want32) < 1e-10


def test_opls_filter_labels_and_early_stop(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(32)
    X = rng.standard_normal((240, 25))
    y = rng.choice(["a", "b", "c"], 240)
    got = F.opls_filter(X, y, north=2)
    codes = np.unique(y, return_inverse=True)[1] + 1
    ref = O.opls_filter(X, O.transformy(codes), 2, 1)
    assert got["north"] == 2
    assert rel(got["X"], ref["X"]) < 1e-10
    # a zero predictor matrix stops before the first component
    z = F.opls_filter(np.zeros((40, 6)), rng.standard_normal((40, 2)), north=2, scaling="none")
    assert z["north"] == 0 and z["W_orth"].shape == (6, 0)
    np.testing.assert_array_equal(z["X"], 0.0)


@pytest.mark.parametrize("kernel,kid", [("linear", 1), ("rbf", 2), ("poly", 3)])
def test_kernel_matrix_and_centering(ctx, kernel, kid):
    import fastpls_amd as F
    rng = np.random.default_rng(33 + kid)
    A = rng.standard_normal((210, 17))
    B = rng.standard_normal((95, 17))
    K = F.kernel_matrix(A, B, kernel, gamma=0.07, degree=3, coef0=0.5)
    Kr = O.kernel_matrix(A, B, kid, 0.07, 3, 0.5)
    assert rel(K, Kr) < 1e-12
    Ktr = F.kernel_matrix(A, A, kernel, gamma=0.07, degree=3, coef0=0.5, out="device")
    kc = F.center_kernel_train(Ktr, in_place=True)
    ref = O.center_kernel_train(O.kernel_matrix(A, A, kid, 0.07, 3, 0.5))
    assert rel(kc["K"].cpu().numpy(), ref["K"]) < 1e-10
    np.testing.assert_allclose(kc["col_means"], ref["col_means"], rtol=1e-11, atol=1e-12)
    assert abs(kc["grand_mean"] - ref["grand_mean"]) <= 1e-11 * max(1.0, abs(ref["grand_mean"]))
    Kt = F.center_kernel_test(F.kernel_matrix(B, A, kernel, 0.07, 3, 0.5), kc["col_means"],
                              kc["grand_mean"])
    want = O.center_kernel_test(O.kernel_matrix(B, A, kid, 0.07, 3, 0.5), ref["col_means"],
                                ref["grand_mean"])
    assert rel(Kt, want) < 1e-10


def test_kernel_errors(ctx):
    import fastpls_amd as F
    A = np.zeros((4, 3))
    with pytest.raises(Exception):
        F.kernel_matrix(A, np.zeros((4, 2)), "rbf")
    with pytest.raises(ValueError):
        F.center_kernel_test(np.zeros((3, 4)), np.zeros(5), 0.0)


def test_pls_opls_regression_like_reference_tests(ctx):
    """tests/testthat/test-kernel-opls.R:45-91 with this library's draws."""
    import fastpls_amd as F
    rng = np.random.default_rng(2202)
    X = rng.standard_normal((70, 11))
    Y = rng.standard_normal((70, 2))
    idx = np.arange(10)
    tr = np.arange(10, 70)
    fit = F.pls(X[tr], Y[tr], X[idx], Y[idx], ncomp=[1, 2], method="opls", north=1)
    assert fit["class"][0] == "fastPLSOpls" and fit["pls_method"] == "simpls"
    assert fit["Ypred"].shape == (10, 2, 2)
    ref = O.opls_pls(X[tr], Y[tr], X[idx], [1, 2], north=1)
    assert rel(fit["Ypred"], ref["Ypred"]) < 1e-8
    y1 = rng.standard_normal((64, 1))
    fit1 = F.pls(X[:64, :9][12:], y1[12:], X[:12, :9], y1[:12], ncomp=[1, 2], method="opls",
                 north=1)
    assert fit1["Ypred"].shape == (12, 1, 2)
    assert len(fit1["Q2Y"]) == 2


def test_pls_kernelpls_like_reference_tests(ctx):
    """tests/testthat/test-kernel-opls.R:1-43 with this library's draws."""
    import fastpls_amd as F
    rng = np.random.default_rng(2201)
    X = rng.standard_normal((72, 10))
    y = rng.choice(["A", "B", "C"], 72)
    idx = np.arange(12)
    tr = np.arange(12, 72)
    fit = F.pls(X[tr], y[tr], X[idx], y[idx], ncomp=[1, 2], method="kernelpls", kernel="rbf")
    assert fit["class"][0] == "fastPLSKernel" and fit["pls_method"] == "simpls"
    assert fit["Ypred"].shape == (12, 2) and set(fit["Ypred"].ravel()) <= set(y)
    codes = np.unique(y, return_inverse=True)[1] + 1
    ref = O.kernel_pls(X[tr], codes[tr], X[idx], [1, 2], kernel=2, classes=3)
    # labels from the device kernel fit equal the oracle's argmax (the score
    # gaps are far above the FP64 rounding of the two fits)
    cube = F.predict(fit, X[idx], raw_scores=True)["Yscore"]
    assert rel(cube, ref["Ypred"]) < 1e-8
    np.testing.assert_array_equal(fit["Ypred_index"], np.argmax(ref["Ypred"], axis=1) + 1)
    # poly kernel regression and the linear kernel's direct SIMPLS fit
    Yr = rng.standard_normal((72, 2))
    fp = F.pls(X[tr], Yr[tr], X[idx], ncomp=2, method="kernelpls", kernel="poly", gamma=0.2,
               degree=2, coef0=1.0)
    rp = O.kernel_pls(X[tr], Yr[tr], X[idx], [2], kernel=3, gamma=0.2, degree=2, coef0=1.0)
    assert rel(fp["Ypred"], rp["Ypred"]) < 1e-8
    fl = F.pls(X[tr], Yr[tr], X[idx], ncomp=2, method="kernelpls", kernel="linear")
    assert fl["kernel_linear_direct"] and fl["Ypred"].shape == (12, 2, 1)
    with pytest.raises(ValueError):
        F.pls(X[tr], Yr[tr], ncomp=2, method="kernelpls", kernel="rbf", gamma=-1.0)


@pytest.mark.parametrize("n1,n2,p", [(1000, 700, 300), (333, 129, 2000), (130, 64, 64)])
def test_large_fp64_products_through_the_linear_kernel(ctx, n1, n2, p):
    """X1 X2^T through the 128 x 128-tile FP64 GEMM (ragged edges; the middle
    shape takes split-K partials), against numpy in FP64."""
    import fastpls_amd as F
    rng = np.random.default_rng(n1 + n2 + p)
    A = rng.standard_normal((n1, p))
    B = rng.standard_normal((n2, p))
    K = F.kernel_matrix(A, B, "linear")
    ref = A @ B.T
    assert np.abs(K - ref).max() <= 1e-13 * np.sqrt(p) * np.abs(A).max() * np.abs(B).max()


@pytest.mark.parametrize("method", ["opls", "kernelpls"])
def test_cv_with_opls_and_kernelpls_methods(ctx, method):
    """pls_cv_predict_compiled methods 4 (OPLS filter per training fold, then
    SIMPLS with scaling none) and 5 (the direct SIMPLS core), leave-one-group-
    out, FP64 regression against the oracle's fold loop; and the LDA head on
    OPLS-filtered folds."""
    import fastpls_amd as F
    rng = np.random.default_rng(41)
    n, p, m = 150, 30, 2
    X = rng.standard_normal((n, p)) + 0.5
    Y = X[:, :4] @ rng.standard_normal((4, m)) + 0.2 * rng.standard_normal((n, m))
    g = np.repeat(np.arange(5), 30)
    rng.shuffle(g)
    res = F.pls_cv_predict(X, Y, g, ncomp=[1, 2, 3], method=method, north=2)
    ref = O.pls_cv_predict(X, Y, g, [1, 2, 3], method=method, north=2)
    np.testing.assert_array_equal(res["status"], ref["status"])
    assert rel(res["Ypred"], ref["Ypred"]) < 1e-8
    if method == "opls":
        y = rng.integers(1, 4, n).astype(np.int32)
        X[:, 0] += y
        rc = F.pls_cv_predict(X, y, g, ncomp=[2, 3], method="opls", north=1, classifier="lda")
        rr = O.pls_cv_predict(X, y, g, [2, 3], method="opls", north=1, classification=True,
                              n_classes=3, classifier="lda")
        np.testing.assert_array_equal(rc["status"], rr["status"])
        bad = (rc["class_pred"] != rr["class_pred"]) & (rr["margin"] > 1e-7)
        assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"
