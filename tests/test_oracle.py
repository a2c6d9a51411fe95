"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong index or sign, a transposed operand) fails it:
published known answers, closed forms, brute force on tiny inputs, textbook
or library routines, and the reference's own tolerance tests restated
(tests/testthat/*.R in the fastPLS tree).
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import fastpls_oracle as O
from tests.helpers import align_signs, rel_err, sin_theta


# ---------------------------------------------------------------- RNG

def test_mt19937_known_answer():
    # The 10000th output of mt19937 default-seeded with 5489 is 4123659995
    # (ISO C++ [rand.predef]).
    assert int(O.mt19937_raw(10000, 5489)[-1]) == 4123659995


def _polar_normals(raw32, count):
    """Marsaglia polar method over 53-bit canonical doubles built from two
    32-bit engine draws (low word first), caching the second variate:
    the algorithm of std::normal_distribution<double> on a 32-bit engine."""
    out = []
    pos = 0
    saved = None
    while len(out) < count:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            u1 = (int(raw32[pos]) + int(raw32[pos + 1]) * 2 ** 32) / 2.0 ** 64
            u2 = (int(raw32[pos + 2]) + int(raw32[pos + 3]) * 2 ** 32) / 2.0 ** 64
            pos += 4
            x, y = 2.0 * u1 - 1.0, 2.0 * u2 - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        saved = x * mult
        out.append(y * mult)
    return np.array(out)


@pytest.mark.parametrize("seed", [1, 7, 20260601])
def test_gaussian_stream_is_the_polar_method(seed):
    g = O.gaussian_matrix(7, 3, seed)  # column-major fill: ravel in F order
    ref = _polar_normals(O.mt19937_raw(400, seed), 21)
    np.testing.assert_array_equal(g.ravel(order="F"), ref)


# ---------------------------------------------------------------- SVD

def _known_svd(rng, m, n, s):
    U, _ = np.linalg.qr(rng.standard_normal((m, len(s))))
    V, _ = np.linalg.qr(rng.standard_normal((n, len(s))))
    return (U * s) @ V.T, U, V


def test_exact_svd_recovers_known_factors():
    rng = np.random.default_rng(0)
    s = np.array([9.0, 5.0, 3.0, 1.0, 0.5])
    A, U, V = _known_svd(rng, 40, 12, s)
    r = O.truncated_svd_exact(A, 5)
    np.testing.assert_allclose(r["s"], s, rtol=1e-12)
    assert rel_err(align_signs(r["U"], U), U) < 1e-12
    assert rel_err(align_signs(r["Vt"].T, V), V) < 1e-12


@pytest.mark.parametrize("power", [0, 1, 2, 3])
def test_rsvd_exact_on_rank_r(power):
    # rank-6 matrix, sketch width 6 + 4 >= rank: the range finder is exact
    rng = np.random.default_rng(1)
    A, U, V = _known_svd(rng, 120, 50, np.array([8, 6, 4, 3, 2, 1.0]))
    r = O.truncated_svd_rsvd(A, 6, oversample=4, power=power, seed=5)
    np.testing.assert_allclose(r["s"], [8, 6, 4, 3, 2, 1.0], rtol=1e-10)
    assert sin_theta(r["U"], U) < 1e-10
    assert rel_err(r["U"] * r["s"] @ r["Vt"], A) < 1e-10


def test_rsvd_reference_tolerance_test():
    # restates tests/testthat/test-svd-rsvd.R:1-33 (80 x 30 Gaussian, k = 6)
    rng = np.random.default_rng(1001)
    A = rng.standard_normal((80, 30))
    ex = O.truncated_svd(A, 6, "exact", 10, 1, 7)
    rs = O.truncated_svd(A, 6, "rsvd", 12, 2, 7)
    assert len(ex["s"]) == len(rs["s"]) == 6
    np.testing.assert_allclose(rs["s"], ex["s"], rtol=1e-2)
    assert np.linalg.norm(ex["U"] @ ex["U"].T - rs["U"] @ rs["U"].T) < 0.5


def test_small_inputs_take_exact_path():
    # restates tests/testthat/test-svd-public-tools.R:43-59: min dim < 6 -> exact
    rng = np.random.default_rng(42)
    A = rng.standard_normal((40, 5))
    out = O.truncated_svd(A, 3, "rsvd", 0, 0, 99)
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    np.testing.assert_allclose(out["s"], s[:3], rtol=1e-8)
    np.testing.assert_allclose(np.abs(out["U"]), np.abs(U[:, :3]), atol=1e-6)


def test_power_iterations_reduce_subspace_error():
    rng = np.random.default_rng(3)
    s = 0.85 ** np.arange(60)
    A, U, V = _known_svd(rng, 300, 80, s)
    errs = [sin_theta(O.truncated_svd_rsvd(A, 5, 5, q, 11)["U"], U[:, :5]) for q in (0, 1, 3)]
    assert errs[0] > errs[1] > errs[2]


def test_crossprod_rsvd_equals_materialised():
    # matrix-free X^T (Y M) / Y^T (X M) sketch == rSVD of S = X^T Y (same seed)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((200, 70))
    Y = X[:, :8] @ rng.standard_normal((8, 40)) + 0.01 * rng.standard_normal((200, 40))
    a = O.truncated_rsvd_crossprod(X, Y, 5, 6, 1, 9)
    b = O.truncated_svd_rsvd(X.T @ Y, 5, 6, 1, 9)
    np.testing.assert_allclose(a["s"], b["s"], rtol=1e-9)
    assert sin_theta(a["U"], b["U"]) < 1e-8


# ---------------------------------------------------------------- SIMPLS

def _krylov_pls1(X, y, K):
    """PLS1 closed form: B_K = Kr (Kr^T Xc^T Xc Kr)^-1 Kr^T Xc^T yc with the
    Krylov basis Kr = [s, (Xc^T Xc) s, ...], s = Xc^T yc."""
    Xc = X - X.mean(0)
    yc = y - y.mean()
    s = Xc.T @ yc
    A = Xc.T @ Xc
    cols = [s]
    for _ in range(K - 1):
        cols.append(A @ cols[-1])
    Kr, _ = np.linalg.qr(np.column_stack(cols))
    return Kr @ np.linalg.solve(Kr.T @ A @ Kr, Kr.T @ s)


@pytest.mark.parametrize("K", [1, 3, 6])
def test_simpls_pls1_equals_krylov_closed_form(K):
    rng = np.random.default_rng(8)
    X = rng.standard_normal((70, 20))
    y = X @ rng.standard_normal(20) + 0.3 * rng.standard_normal(70)
    m = O.pls_model2_fast(X, y.reshape(-1, 1), [K])
    np.testing.assert_allclose(m["B"][:, 0, 0], _krylov_pls1(X, y, K), rtol=1e-8, atol=1e-10)


def test_simpls_invariants():
    rng = np.random.default_rng(9)
    X = rng.standard_normal((150, 40))
    Y = X[:, :6] @ rng.standard_normal((6, 4)) + 0.2 * rng.standard_normal((150, 4))
    K = 8
    m = O.pls_model2_fast(X, Y, [2, K], fit=True, return_ttrain=True)
    Xc = X - X.mean(0)
    V, R, T = m["V"], m["R"], m["Ttrain"]
    np.testing.assert_allclose(V.T @ V, np.eye(K), atol=1e-10)
    np.testing.assert_allclose(T.T @ T, np.eye(K), atol=1e-10)  # |t| = 1, t orthogonal
    np.testing.assert_allclose(Xc @ R, T, atol=1e-10)
    np.testing.assert_allclose(m["B"][:, :, 1], R @ m["Q"].T, atol=1e-10)
    np.testing.assert_allclose(m["Q"], (Y - Y.mean(0)).T @ T, atol=1e-10)


def test_simpls_warm_start_only_component_zero_draws():
    # changing the seed changes component 0's start only; a converged first
    # direction (m = 2, strong signal) makes B nearly seed independent
    rng = np.random.default_rng(10)
    X = rng.standard_normal((100, 30))
    Y = np.column_stack([X[:, 0] * 5, X[:, 1] * 3]) + 0.01 * rng.standard_normal((100, 2))
    a = O.pls_model2_fast(X, Y, [2], seed=1)["B"]
    b = O.pls_model2_fast(X, Y, [2], seed=1)["B"]
    c = O.pls_model2_fast(X, Y, [2], seed=2)["B"]
    np.testing.assert_array_equal(a, b)             # determinism (test-pls-rsvd.R:100-140)
    assert not np.array_equal(a, c)
    assert rel_err(a, c) < 0.1  # two warm-started power steps: close, not converged


# ---------------------------------------------------------------- PLSSVD

def test_plssvd_exact_fit_is_least_squares_on_scores():
    # With the exact SVD of S, D V^T = U^T S = T^T Yc, so W = G^-1 T^T Yc:
    # the fitted values are the least-squares projection of Yc onto span(T).
    rng = np.random.default_rng(12)
    X = rng.standard_normal((90, 25)) + 1.0
    Y = rng.standard_normal((90, 7))
    m = O.pls_model1(X, Y, [3, 5], fit=True)
    Xc, Yc = X - X.mean(0), Y - Y.mean(0)
    for a, k in enumerate([3, 5]):
        T = Xc @ m["R"][:, :k]
        coef, *_ = np.linalg.lstsq(T, Yc, rcond=None)
        np.testing.assert_allclose(m["Yfit"][:, :, a] - Y.mean(0), T @ coef, atol=1e-9)


def test_label_aware_S_equals_dense_onehot():
    rng = np.random.default_rng(13)
    X = rng.standard_normal((200, 15)) + 3.0
    y = rng.integers(1, 6, 200)
    m = O.plssvd_label_aware(X, y, [4])
    Xc = X - X.mean(0)
    Yc = O.transformy(y) - O.transformy(y).mean(0)
    np.testing.assert_allclose(m["S"], Xc.T @ Yc, atol=1e-10)


# ---------------------------------------------------------------- LDA

def test_lda_pooled_covariance_brute_force():
    rng = np.random.default_rng(14)
    n, k, C = 120, 5, 4
    T = rng.standard_normal((n, k))
    y = np.repeat(np.arange(1, C + 1), n // C)
    mod = O.lda_train_prefix(T, y, C, [k], ridge=0.0)[0]
    W = np.zeros((k, k))
    for c in range(1, C + 1):
        D = T[y == c] - T[y == c].mean(0)
        W += D.T @ D
    W /= n - C
    np.testing.assert_allclose(np.linalg.inv(mod["inv_cov"]), W, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(mod["means"][1], T[y == 2].mean(0), rtol=1e-12)


def test_lda_matches_sklearn_on_balanced_classes():
    from sklearn.discriminant_analysis import LinearDiscriminantAnalysis
    rng = np.random.default_rng(15)
    C, per, k = 6, 40, 4
    mu = rng.standard_normal((C, k)) * 2
    y = np.repeat(np.arange(1, C + 1), per)
    T = mu[y - 1] + rng.standard_normal((C * per, k))
    mod = O.lda_train_prefix(T, y, C, [k], ridge=0.0)[0]
    Tt = mu[np.arange(C).repeat(30)] + 1.3 * rng.standard_normal((C * 30, k))
    ours = O.lda_labels_from_scores(Tt @ mod["linear"].T, mod["constants"])
    sk = LinearDiscriminantAnalysis(solver="lsqr").fit(T, y).predict(Tt)
    np.testing.assert_array_equal(ours, sk)


def test_lda_projected_equals_explicit():
    # restates tests/testthat/test-lda-cuda-native.R:1-31
    rng = np.random.default_rng(20260503)
    X = rng.standard_normal((90, 14))
    R = rng.standard_normal((14, 6))
    off = rng.standard_normal(6)
    T = X @ R - off
    y = np.tile(np.arange(1, 6), 18)
    ex = O.lda_train_prefix(T, y, 5, [2, 6], 1e-8)
    pr = O.lda_project_train_prefix(X, R, off, y, 5, [2, 6], 1e-8)
    for a, k in enumerate([2, 6]):
        for f in ("means", "linear", "constants"):
            np.testing.assert_allclose(pr[a][f], ex[a][f], rtol=1e-12, atol=1e-12)
        p1 = O.lda_labels_from_scores(T[:, :k] @ ex[a]["linear"].T, ex[a]["constants"])
        p2 = O.lda_project_predict_labels(X, R[:, :k], off[:k], pr[a])
        np.testing.assert_array_equal(p1, p2)


def test_lda_ties_go_to_lowest_class():
    s = np.array([[1.0, 3.0, 3.0], [2.0, 2.0, 1.0]])
    np.testing.assert_array_equal(O.lda_labels_from_scores(s, np.zeros(3)), [2, 1])


def test_lda_errors():
    T = np.ones((6, 2))
    with pytest.raises(ValueError, match="empty class"):
        O.lda_train_prefix(T + np.arange(6)[:, None], np.array([1, 1, 1, 3, 3, 3]), 3, [2])
    with pytest.raises(ValueError, match="1..n_classes"):
        O.lda_train_prefix(T, np.array([1, 2, 3, 4, 1, 2]), 3, [2])


# ---------------------------------------------------------------- cKNN

def test_cknn_k1_is_nearest_neighbour_in_candidates():
    rng = np.random.default_rng(16)
    C, k = 8, 5
    Ttr = O.candidate_row_l2(rng.standard_normal((60, k)))
    y = rng.integers(1, C + 1, 60)
    y[:C] = np.arange(1, C + 1)
    cent = O.candidate_centroids(Ttr, y, C)
    Tte = O.candidate_row_l2(rng.standard_normal((15, k)))
    r = O.candidate_knn_predict(Tte, Ttr, y, cent, None, top_k=3, top_m=4, knn_k=1, tau=0.2,
                                alpha=0.0, return_candidates=True)
    for i in range(15):
        best = {}
        for c in r["candidates"][i]:
            best[int(c)] = max(float(Ttr[j] @ Tte[i]) for j in np.nonzero(y == c)[0])
        # alpha = 0: score = max similarity; rerank is stable descending
        order = sorted(best, key=lambda c: (-best[c], list(r["candidates"][i]).index(c)))
        np.testing.assert_array_equal(r["top_index"][i], order[:3])
        np.testing.assert_allclose(r["top_score"][i], [best[c] for c in order[:3]], rtol=1e-12)


def test_cknn_candidate_ties_keep_lower_class():
    # identical centroid scores: candidate order is ascending class id
    Tte = np.array([[1.0, 0.0]])
    cent = np.array([[0.0, 1.0], [1.0, 0.0], [1.0, 0.0], [0.5, 0.5]])
    Ttr = np.array([[1.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.6, 0.8]])
    y = np.array([2, 3, 1, 4])
    r = O.candidate_knn_predict(Tte, Ttr, y, cent, None, 2, 3, 1, 0.2, 1.0, return_candidates=True)
    np.testing.assert_array_equal(r["candidates"][0], [2, 3, 4])
    np.testing.assert_array_equal(r["top_index"][0], [2, 3])


def test_cknn_smoothed_score_formula():
    # local = mx + tau log(mean exp((v - mx)/tau)) over the top-k similarities
    rng = np.random.default_rng(17)
    Ttr = O.candidate_row_l2(rng.standard_normal((30, 4)))
    y = np.ones(30, dtype=int)
    y[15:] = 2
    cent = O.candidate_centroids(Ttr, y, 2)
    Tte = O.candidate_row_l2(rng.standard_normal((1, 4)))
    r = O.candidate_knn_predict(Tte, Ttr, y, cent, None, 2, 2, 5, 0.3, 0.5, return_candidates=True)
    for j, c in enumerate(r["candidates"][0]):
        v = np.sort(Ttr[y == c] @ Tte[0])[::-1][:5]
        local = v[0] + 0.3 * np.log(np.mean(np.exp((v - v[0]) / 0.3)))
        assert abs(r["candidate_score"][0, j] - (local + 0.5 * r["candidate_base"][0, j])) < 1e-12
