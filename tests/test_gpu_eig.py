"""GPU: the small symmetric eigensolver (tridiagonalisation + divide and
conquer) through the exact-SVD entry point (Gram matrix -> eigenpairs).

Checked against LAPACK (numpy) on spectra that stress the divide-and-conquer
merges: random, five-fold repeated singular values (every merge deflates),
rank deficiency (zero eigenvalues), graded values and every order up to 128.
Tolerances: singular values rel 1e-10 where sigma >= 1e-2 sigma_1 (the Gram
matrix squares the condition number, so smaller values carry eps/sigma^2
relative error; 1e-8 for the graded spectrum); subspaces of separated
clusters to sin(theta) < 1e-8; right singular vectors orthonormal to 1e-12.
"""
import numpy as np
import pytest

from tests.helpers import sin_theta

pytestmark = pytest.mark.gpu


def _with_spectrum(rng, rows, s):
    n = len(s)
    U, _ = np.linalg.qr(rng.standard_normal((rows, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (U * s) @ V.T


def _check(A, k, groups, rtol=1e-10):
    import fastpls_amd as F
    got = F.fastsvd(A, k, method="exact")
    U0, s0, Vt0 = np.linalg.svd(A, full_matrices=False)
    kk = got["u"].shape[1]
    big = s0[:kk] >= 1e-2 * s0[0]
    np.testing.assert_allclose(got["d"][big], s0[:kk][big], rtol=rtol)
    V = got["v"]
    assert np.abs(V.T @ V - np.eye(kk)).max() < 1e-12
    for lo, hi in groups:
        assert sin_theta(got["u"][:, lo:hi], U0[:, lo:hi]) < 1e-8, (lo, hi)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 33, 64, 100, 127, 128])
def test_eig_random_orders(ctx, n):
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((max(3 * n, 8), n))
    _check(A, n, [(0, min(n, 5))])


def test_eig_repeated_values(ctx):
    rng = np.random.default_rng(7)
    s = np.repeat([9.0, 7.0, 5.0, 3.0, 2.0, 1.5, 1.0, 0.7, 0.5, 0.3], 10)
    A = _with_spectrum(rng, 400, s)
    _check(A, 100, [(10 * g, 10 * g + 10) for g in range(10)])


def test_eig_rank_deficient(ctx):
    rng = np.random.default_rng(8)
    s = np.concatenate([np.linspace(10.0, 1.0, 60), np.zeros(68)])
    A = _with_spectrum(rng, 300, s)
    import fastpls_amd as F
    got = F.fastsvd(A, 60, method="exact")
    U0, s0, _ = np.linalg.svd(A, full_matrices=False)
    np.testing.assert_allclose(got["d"], s0[:60], rtol=1e-10)
    assert sin_theta(got["u"], U0[:, :60]) < 1e-9


def test_eig_graded(ctx):
    rng = np.random.default_rng(9)
    s = np.logspace(0, -3, 90)
    A = _with_spectrum(rng, 200, s)
    _check(A, 90, [(0, 10), (10, 40)], rtol=1e-8)


def test_eig_wide_and_tall(ctx):
    rng = np.random.default_rng(10)
    for rows, cols in [(50, 120), (128, 128), (700, 96)]:
        A = rng.standard_normal((rows, cols)) @ np.diag(np.linspace(3, 1, cols))
        _check(A, min(rows, cols), [(0, 4)])


# ---- orders above 128: block Jacobi over pairs of <= 64-row blocks, each pair
# diagonalised by the one-workgroup solver

@pytest.mark.parametrize("n", [129, 200, 333])
def test_eig_large_random_orders(ctx, n):
    rng = np.random.default_rng(300 + n)
    A = rng.standard_normal((2 * n, n))
    _check(A, n, [(0, 5), (n // 2, n // 2 + 3)])


def test_eig_large_repeated_and_rank_deficient(ctx):
    rng = np.random.default_rng(9)
    s = np.concatenate([np.repeat([8.0, 5.0, 3.0, 2.0], 40), np.linspace(1.5, 0.5, 60),
                        np.zeros(40)])
    A = _with_spectrum(rng, 400, s)
    import fastpls_amd as F
    got = F.fastsvd(A, 220, method="exact")
    U0, s0, _ = np.linalg.svd(A, full_matrices=False)
    np.testing.assert_allclose(got["d"], s0[:220], rtol=1e-10, atol=1e-12 * s0[0])
    for lo, hi in [(0, 40), (40, 80), (80, 120), (120, 160)]:
        assert sin_theta(got["u"][:, lo:hi], U0[:, lo:hi]) < 1e-8, (lo, hi)
    V = got["v"]
    assert np.abs(V.T @ V - np.eye(220)).max() < 1e-11


def test_eig_large_wide_cross_covariance(ctx):
    """The label-aware shape at reduced size: a wide p x C class-sum matrix
    whose columns sum to zero (rank C - 1)."""
    rng = np.random.default_rng(10)
    p, C = 300, 260
    S = rng.standard_normal((p, C))
    S -= S.mean(axis=1, keepdims=True)
    import fastpls_amd as F
    got = F.fastsvd(S, 150, method="exact")
    U0, s0, _ = np.linalg.svd(S, full_matrices=False)
    np.testing.assert_allclose(got["d"], s0[:150], rtol=1e-10)
    assert sin_theta(got["u"][:, :150], U0[:, :150]) < 1e-8
