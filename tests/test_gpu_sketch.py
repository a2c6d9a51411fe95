"""Learned sketch for the rSVD range finder: fpls_sketch_train /
fpls_sketch_eval and the sketch argument of fpls_rsvd_xprod.

The reference has no learned sketch; the range finder it feeds is
src/svd_cpu_rsvd.cpp:95-124 (matrix free: src/fastPLS.cpp:550-623).  The
loss, its gradient and the evaluation are pinned here by independent
computations: subspaces from numpy QR / SVD, principal angles from
scipy.linalg.subspace_angles, and the gradient (through the first Adam step)
by central finite differences of the QR-based loss."""
import numpy as np
import pytest
import scipy.linalg

from oracle import fastpls_oracle as O
from tests.helpers import sin_theta

pytestmark = pytest.mark.gpu


def _data(n, p, m, rank, seed):
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((n, rank)))[0]
    V = np.linalg.qr(rng.standard_normal((p, rank)))[0]
    s = 40.0 * 0.85 ** np.arange(rank)
    X = (U * s) @ V.T + 0.05 * rng.standard_normal((n, p))
    B = rng.standard_normal((p, m)) / np.sqrt(p)
    Y = X @ B + 0.01 * rng.standard_normal((n, m))
    return X.astype(np.float32), Y.astype(np.float32)


def _masks(n, D, frac, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((n, D)) < frac).astype(np.int8)


def _S(X, Y, mask):
    Xd = X[mask.astype(bool)].astype(np.float64)
    Yd = Y[mask.astype(bool)].astype(np.float64)
    return (Xd - Xd.mean(0)).T @ (Yd - Yd.mean(0))


def _qr_loss(S, Om, k):
    """k - ||Q^T U_k||_F^2 with Q = orth(S Om): squared Frobenius sines."""
    Uk = np.linalg.svd(S, full_matrices=False)[0][:, :k]
    Q = np.linalg.qr(S @ Om)[0]
    return k - np.linalg.norm(Q.T @ Uk) ** 2


def _orth_pos(A):
    Q, R = np.linalg.qr(A)
    return Q * np.sign(np.diag(R))


def test_sketch_start_loss_and_orthonormal_start(ctx):
    import fastpls_amd as F
    n, p, m, k, l, D = 600, 700, 24, 5, 9, 6
    X, Y = _data(n, p, m, 12, 1)
    masks = _masks(n, D, 0.8, 2)
    r = F.sketch_train(X, Y, masks, k, l, iters=0, seed=7)
    om0 = _orth_pos(O.gaussian_matrix(m, l, 7))
    # zero steps: Omega is the seed's Gaussian stream, Cholesky-QR orthonormalised
    np.testing.assert_allclose(r["omega"], om0, rtol=0, atol=1e-12)
    ref = np.mean([_qr_loss(_S(X, Y, masks[:, d]), om0, k) for d in range(D)])
    assert r["loss"].shape == (1,)
    np.testing.assert_allclose(r["loss"][0], ref, rtol=1e-4, atol=1e-8)


def test_sketch_first_step_follows_the_gradient(ctx):
    """One Adam step moves each entry by -lr sign(dL/dOmega) before the
    re-orthonormalisation: the step's Omega and loss are rebuilt from the
    finite-difference gradient of the QR-based loss."""
    import fastpls_amd as F
    n, p, m, k, l, D = 500, 400, 16, 4, 7, 4
    X, Y = _data(n, p, m, 10, 3)
    masks = _masks(n, D, 0.75, 4)
    lr = 1e-3
    r = F.sketch_train(X, Y, masks, k, l, iters=1, lr=lr, seed=11)
    Ss = [_S(X, Y, masks[:, d]) for d in range(D)]
    om0 = _orth_pos(O.gaussian_matrix(m, l, 11))

    def L(Om):
        return sum(_qr_loss(S, Om, k) for S in Ss)

    h = 1e-6
    G = np.zeros_like(om0)
    for i in range(m):
        for j in range(l):
            E = np.zeros_like(om0)
            E[i, j] = h
            G[i, j] = (L(om0 + E) - L(om0 - E)) / (2 * h)
    # the loss depends only on range(Omega): the gradient is horizontal
    assert np.abs(om0.T @ G).max() < 1e-5 * np.abs(G).max()
    om1 = _orth_pos(om0 - lr * np.sign(G))
    big = np.abs(G) > 1e-3 * np.abs(G).max()
    assert big.mean() > 0.8
    np.testing.assert_allclose(r["omega"], om1, rtol=0, atol=2.5 * lr * (1 - big.mean()) + 1e-9)
    np.testing.assert_allclose(r["loss"][1], L(om1) / D, rtol=1e-4, atol=1e-9)
    assert r["loss"][1] < r["loss"][0]


def test_sketch_training_beats_gaussian_on_fresh_draws(ctx):
    import fastpls_amd as F
    n, p, m, k, l = 1500, 1200, 40, 8, 14
    X, Y = _data(n, p, m, 30, 5)
    train = _masks(n, 40, 0.8, 6)
    fresh = _masks(n, 10, 0.8, 7)
    r = F.sketch_train(X, Y, train, k, l, iters=150, lr=0.02, seed=3)
    om = r["omega"]
    np.testing.assert_allclose(om.T @ om, np.eye(l), atol=1e-10)
    assert np.all(np.isfinite(r["loss"]))
    assert r["loss"][-1] < 0.25 * r["loss"][0]
    # the final trace entry is the loss at the returned Omega; sin^2 of the
    # largest angle is bounded by the sum of the squared sines
    st_train = F.sketch_eval(X, Y, train, k, omega=om, power=0)
    assert np.mean(st_train ** 2) <= r["loss"][-1] * (1 + 1e-6) + 1e-12
    learned = F.sketch_eval(X, Y, fresh, k, omega=om, power=0)
    gauss = F.sketch_eval(X, Y, fresh, k, l=l, power=0, seed=3)
    assert np.median(learned) < 0.5 * np.median(gauss)


@pytest.mark.parametrize("power", [0, 1, 2])
def test_sketch_eval_matches_principal_angles(ctx, power):
    import fastpls_amd as F
    n, p, m, k, l, D = 700, 900, 20, 4, 8, 3
    X, Y = _data(n, p, m, 12, 9)
    masks = _masks(n, D, 0.7, 10)
    om = O.gaussian_matrix(m, l, 5)
    got = F.sketch_eval(X, Y, masks, k, omega=om, power=power)
    got_seed = F.sketch_eval(X, Y, masks, k, l=l, power=power, seed=5)
    np.testing.assert_array_equal(got, got_seed)
    for d in range(D):
        S = _S(X, Y, masks[:, d])
        Uk = np.linalg.svd(S, full_matrices=False)[0][:, :k]
        Yq = S @ om
        if power == 1:
            Yq = S @ (S.T @ Yq)
        else:
            for _ in range(power):
                Yq = S @ np.linalg.qr(S.T @ Yq)[0]
        ang = scipy.linalg.subspace_angles(Yq, Uk)
        ref = np.sin(ang.max())
        np.testing.assert_allclose(got[d], ref, rtol=2e-3, atol=1e-9)


def test_xprod_with_sketch_argument(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(12)
    n, p, m, k, os_ = 800, 500, 30, 6, 4
    X = rng.standard_normal((n, p)).astype(np.float32)
    Y = (X @ rng.standard_normal((p, m)) * 0.05 + 0.01 * rng.standard_normal((n, m))).astype(np.float32)
    l = k + os_
    # the seed's Gaussian stream passed explicitly gives the same result
    a = F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, rsvd_power=1, seed=4)
    b = F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, rsvd_power=1, seed=4,
                                   omega=O.gaussian_matrix(m, l, 4))
    for key in ("U", "s", "Vt"):
        np.testing.assert_array_equal(a[key], b[key])
    # any sketch: matches the oracle's range finder run with that sketch
    om = np.linalg.qr(rng.standard_normal((m, l)))[0]
    for power in (0, 2):
        got = F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, rsvd_power=power, omega=om)
        ref = O.truncated_rsvd_crossprod(X.astype(np.float64), Y.astype(np.float64), k, os_, power,
                                         1, omega=om)
        np.testing.assert_allclose(got["s"], ref["s"], rtol=2e-4)
        # S formed once, the same range finder on S
        mat = F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, rsvd_power=power, omega=om,
                                         route="materialise")
        np.testing.assert_allclose(mat["s"], ref["s"], rtol=2e-4)
        assert sin_theta(mat["U"], ref["U"]) < 1e-3
        b3 = F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, rsvd_power=power, omega=om,
                                        route="materialise", precision="bf16x3")
        np.testing.assert_allclose(b3["s"], ref["s"], rtol=1e-3)
    with pytest.raises(F.FplsError):
        F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=os_, omega=om[:, :-1])


def test_sketch_argument_errors(ctx):
    import fastpls_amd as F
    X, Y = _data(200, 100, 12, 5, 13)
    masks = _masks(200, 2, 0.8, 14)
    with pytest.raises(F.FplsError):
        F.sketch_train(X, Y, masks, 5, 13)          # l > m
    with pytest.raises(F.FplsError):
        F.sketch_train(X, Y, masks, 6, 5)           # k > l
    with pytest.raises(F.FplsError):
        F.sketch_train(X.astype(np.float64), Y, masks, 2, 4)
    with pytest.raises(F.FplsError):
        F.sketch_eval(X, Y, masks, 5, l=4)
