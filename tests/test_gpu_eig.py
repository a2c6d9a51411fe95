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


def _c3_class_sums(seed=11, p=1024, C=1000, n=1_200_000):
    """Label-aware class-sum matrix S = X^T Yc (p x C) of the 1000-class
    workload, built from class means without the 1.2M rows: Zipf(1.1) class
    sizes, means on the radius-4 sphere, sample means with N(0, 0.6^2/n_c)
    noise, columns centred by the weighted grand mean (rank C - 1)."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, C + 1) ** 1.1
    cnt = np.maximum(1, np.floor(n * w / w.sum()))
    mu = rng.standard_normal((C, p))
    mu *= 4.0 / np.linalg.norm(mu, axis=1, keepdims=True)
    xbar = mu + 0.6 / np.sqrt(cnt)[:, None] * rng.standard_normal((C, p))
    mX = (cnt[:, None] * xbar).sum(0) / cnt.sum()
    return np.asfortranarray((cnt[:, None] * (xbar - mX)).T)


def test_eig_chip_c3_smallest_kept_factors(ctx):
    """All K = 500 kept factors of the 1000-class label-aware cross-product
    (p = 1024, 999 nonzero singular values, sigma_1 / sigma_500 ~ 900),
    against the FP64 SVD the oracle restates (svd_econ, the reference's
    dgesdd route in src/svd_cpu_exact.cpp:7-38).

    The exact route diagonalises the Gram matrix S^T S, so an eigenvalue
    error |d lambda| <= c eps ||S||^2 becomes a relative singular-value error
    c/2 eps (sigma_1/sigma_k)^2, and the K-column left subspace moves by
    sin(theta) <= c eps sigma_1^2 / (sigma_K^2 - sigma_{K+1}^2) (Davis-Kahan).
    The bounds below take c = sqrt(C) ~ 32 (Gram formation over 1000 terms
    plus a backward-stable eigensolver); LAPACK's own Gram route measures
    c ~ 3 on this matrix.  Also bounded: the rank-K reconstruction error
    against the Eckart-Young optimum, and the orthonormality of U and V."""
    import fastpls_amd as F
    from oracle.fastpls_oracle import svd_econ
    S = _c3_class_sums()
    p, C = S.shape
    K = 500
    eps = np.finfo(np.float64).eps
    c = np.sqrt(C)
    got = F.fastsvd(S, K, method="exact")
    U0, s0, V0 = svd_econ(S)
    d, U, V = got["d"], got["u"], got["v"]
    assert d.shape == (K,) and U.shape == (p, K) and V.shape == (C, K)
    rel = np.abs(d - s0[:K]) / s0[:K]
    bound = c * eps * (s0[0] / s0[:K]) ** 2
    worst = int(np.argmax(rel / bound))
    assert np.all(rel <= bound), (worst, rel[worst], bound[worst])
    # the smallest kept factors themselves, not only the leading ones
    assert rel[K - 50:].max() <= c * eps * (s0[0] / s0[K - 1]) ** 2
    gapK = s0[K - 1] ** 2 - s0[K] ** 2
    assert sin_theta(U, U0[:, :K]) <= c * eps * s0[0] ** 2 / gapK
    assert sin_theta(V, V0[:, :K]) <= c * eps * s0[0] ** 2 / gapK
    # individual vectors, each with its own two-sided gap
    for j in (0, 1, 10, 100, 250, K - 2):
        g = min(s0[j - 1] ** 2 - s0[j] ** 2 if j else np.inf, s0[j] ** 2 - s0[j + 1] ** 2)
        assert sin_theta(U[:, j:j + 1], U0[:, j:j + 1]) <= c * eps * s0[0] ** 2 / g, j
    # S V = U diag(d) and the rank-K residual against the optimum
    assert np.abs(S @ V - U * d).max() <= c * eps * s0[0] * 10
    opt = np.sqrt((s0[K:] ** 2).sum())
    res = np.linalg.norm(S - (U * d) @ V.T)
    assert res <= opt * (1 + 1e-10), (res, opt)
    assert np.abs(U.T @ U - np.eye(K)).max() < c * eps * (s0[0] / s0[K - 1]) ** 2
    assert np.abs(V.T @ V - np.eye(K)).max() < 1e-11
