"""Pins of the oracle's OPLS filter and kernel matrices (src/fastPLS.cpp:4314-4483).

The kernels are pinned against independent routines: scipy's pairwise
squared distances for the rbf kernel, an element-by-element loop for the
polynomial kernel.  Kernel centring is pinned against the double-centring
matrix H K H (H = I - 11'/n).  The OPLS filter is pinned by the invariants
that define it: every orthogonal weight has unit norm and is orthogonal to
the predictive weight; for one response column the orthogonal scores are
orthogonal to the centred response; each deflation leaves the filtered X
orthogonal to the removed scores; and applying the stored filter to the
training rows reproduces the filtered X.  A dropped term, a sign error or a
transposed operand breaks one of these.  The reference's tests
(tests/testthat/test-kernel-opls.R) check the fitted wrappers' shapes and
dispatch; those are mirrored in tests/test_gpu_opls_kernel.py.
"""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import fastpls_oracle as O


def test_rbf_kernel_against_pairwise_distances():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((30, 6))
    B = rng.standard_normal((20, 6))
    K = O.kernel_matrix(A, B, 2, 0.3)
    np.testing.assert_allclose(K, np.exp(-0.3 * cdist(A, B, "sqeuclidean")), rtol=1e-12,
                               atol=1e-14)
    # the diagonal of K(A, A) is exactly 1 (only tiny negative d2 are clamped)
    np.testing.assert_allclose(np.diag(O.kernel_matrix(A, A, 2, 0.3)), 1.0, atol=1e-13)


def test_poly_and_linear_kernels_by_loops():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((7, 4))
    B = rng.standard_normal((5, 4))
    K = O.kernel_matrix(A, B, 3, 0.5, degree=3, coef0=1.5)
    for i in range(7):
        for j in range(5):
            d = sum(A[i, q] * B[j, q] for q in range(4))
            assert abs(K[i, j] - (0.5 * d + 1.5) ** 3) < 1e-12 * max(1.0, abs(K[i, j]))
    np.testing.assert_allclose(O.kernel_matrix(A, B, 1, 9.0), A @ B.T, rtol=1e-14)
    with pytest.raises(ValueError):
        O.kernel_matrix(A, B, 4, 1.0)
    with pytest.raises(ValueError):
        O.kernel_matrix(A, B[:, :3], 1, 1.0)


def test_kernel_centering_is_double_centering():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((25, 5))
    K = O.kernel_matrix(X, X, 2, 0.2)
    H = np.eye(25) - np.ones((25, 25)) / 25
    kc = O.center_kernel_train(K)
    np.testing.assert_allclose(kc["K"], H @ K @ H, atol=1e-13)
    np.testing.assert_allclose(kc["K"].sum(axis=0), 0.0, atol=1e-12)
    # the test-side centring of the training kernel is the training centring
    np.testing.assert_allclose(O.center_kernel_test(K, kc["col_means"], kc["grand_mean"]),
                               kc["K"], atol=1e-14)


@pytest.mark.parametrize("scaling", [1, 2, 3])
def test_opls_filter_invariants_single_response(scaling):
    rng = np.random.default_rng(4 + scaling)
    X = rng.standard_normal((80, 12)) + 2.0
    y = X[:, :3] @ np.array([1.0, -0.5, 0.25]) + 0.3 * rng.standard_normal(80)
    f = O.opls_filter(X, y, north=3, scaling=scaling)
    assert f["north"] == 3
    Xp, mX, vX = O.preprocess_train(X, scaling)
    np.testing.assert_allclose(f["mX"], mX)
    np.testing.assert_allclose(f["vX"], vX)
    yc = y - y.mean()
    Xa = Xp.copy()
    for a in range(3):
        wo, po = f["W_orth"][:, a], f["P_orth"][:, a]
        assert abs(np.linalg.norm(wo) - 1.0) < 1e-12
        w = Xa.T @ yc
        w /= np.linalg.norm(w)
        assert abs(w @ wo) < 1e-10                       # orthogonal weight
        t = Xa @ wo
        assert abs(t @ yc) < 1e-8 * np.linalg.norm(t) * np.linalg.norm(yc)   # orthogonal score
        Xa = Xa - np.outer(t, po)
        assert np.abs(Xa.T @ t).max() < 1e-9 * np.linalg.norm(t) ** 2       # deflated
    np.testing.assert_allclose(f["X"], Xa, atol=1e-10)
    # the stored filter applied to the training rows gives the filtered X
    np.testing.assert_allclose(O.opls_apply_filter(X, f["mX"], f["vX"], f["W_orth"],
                                                   f["P_orth"]), f["X"], atol=1e-10)


def test_opls_filter_edge_cases():
    rng = np.random.default_rng(9)
    X = rng.standard_normal((30, 5))
    Y = rng.standard_normal((30, 2))
    f0 = O.opls_filter(X, Y, north=0)
    assert f0["north"] == 0 and f0["W_orth"].shape == (5, 0)
    np.testing.assert_allclose(f0["X"], X - X.mean(axis=0))
    with pytest.raises(ValueError):
        O.opls_filter(X, Y, north=-1)
    # a zero predictor matrix stops before the first component (t't = 0)
    f = O.opls_filter(np.zeros((30, 4)), Y, north=2, scaling=3)
    assert f["north"] == 0
    np.testing.assert_array_equal(f["X"], 0.0)
