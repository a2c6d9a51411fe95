"""Pins of the oracle functions that the closed-form tests of test_oracle.py
leave open: the SIMPLS refresh for several responses, the CV fold map and its
R generator, the CV driver's bookkeeping and status codes, label-aware PLSSVD
against the dense one-hot fit, and the candidate-kNN score-space slice rule.

Each check compares the oracle with something independent of it: a textbook
formulation in a different form (golden fixtures, tests/golden/make_golden.py),
values R prints, hand-worked expectations, or a closed form that a wrong power
count, warm start, deflation or slice index would break.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import align_signs, rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _all_equal_tol(target, current):
    """R's all.equal tolerance measure: mean |target - current| / mean |target|."""
    t = np.asarray(target).ravel()
    c = np.asarray(current).ravel()
    return float(np.mean(np.abs(t - c)) / np.mean(np.abs(t)))


def _textbook():
    """The golden-fixture generator's textbook formulations."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


# ---------------------------------------------------------------- SIMPLS refresh

@pytest.mark.parametrize("q", [1, 2, 3])
def test_simpls_refresh_is_warm_started_power_steps(q):
    """For several responses every weight r_a points along (S_a S_a^T)^q r_{a-1}
    (component 0: along (S_0 S_0^T)^q omega, omega the seed's Gaussian column),
    with S_a = (I - V_a V_a^T) S_0 (src/fastPLS.cpp:1117-1189: column 0 of the
    sketch is the previous weight, q unnormalised power steps, then
    normalisation).  Random data keep the singular values close, so one power
    step more or less changes the direction: the wrong count must fail."""
    rng = np.random.default_rng(31)
    X = rng.standard_normal((40, 16))
    Y = rng.standard_normal((40, 4))
    K, seed = 6, 5
    m = O.pls_model2_fast(X, Y, [K], inc_iters=q, seed=seed)
    Xc, Yc = X - X.mean(0), Y - Y.mean(0)
    S0 = Xc.T @ Yc
    R, V = m["R"], m["V"]

    def direction(a, qq):
        Sa = S0 - V[:, :a] @ (V[:, :a].T @ S0)
        z = R[:, a - 1].copy() if a > 0 else O.gaussian_matrix(S0.shape[0], 1, seed + a)[:, 0]
        for _ in range(qq):
            z = Sa @ (Sa.T @ z)
        return z

    def cos(u, v):
        return abs(float(u @ v)) / (np.linalg.norm(u) * np.linalg.norm(v))

    for a in range(K):
        assert cos(direction(a, q), R[:, a]) > 1 - 1e-12, (a, q)
        assert cos(direction(a, q + 1), R[:, a]) < 1 - 1e-7, (a, q)
        if q > 1:
            assert cos(direction(a, q - 1), R[:, a]) < 1 - 1e-7, (a, q)
    # the power count is that of inc_iters, and it changes the fit
    m2 = O.pls_model2_fast(X, Y, [K], inc_iters=q + 1, seed=seed)
    assert rel_err(m2["B"], m["B"]) > 1e-6


def test_simpls_fast_converges_to_textbook_simpls():
    """Golden fixture: SIMPLS in projection form with exact SVDs (de Jong 1993).
    Warm-started power steps converge to its weights, so 30 steps per
    component (S0 scaled to unit norm, so the unnormalised steps stay in range)
    give the textbook coefficients for every ncomp."""
    d = np.load(os.path.join(GOLDEN, "simpls_textbook.npz"))
    K = int(d["K"])
    m = O.pls_model2_fast(d["X"], d["Y"], list(range(1, K + 1)), inc_iters=30)
    for a in range(K):
        assert rel_err(m["B"][:, :, a], d["B"][:, :, a]) < 1e-8, a
    # the reference's default two steps are close but not converged
    m2 = O.pls_model2_fast(d["X"], d["Y"], list(range(1, K + 1)))
    assert 1e-6 < rel_err(m2["B"][:, :, K - 1], d["B"][:, :, K - 1]) < 0.1


def test_simpls_rsvd_seed_tolerance_restated():
    """test-pls-rsvd.R:230-259 restated: SIMPLS with two sketch seeds (the
    'irlba' and 'cpu_rsvd' calls both take the incremental refresh here and
    differ only in the seed of component 0's Gaussian start) agree to
    all.equal tolerance 7e-2, and stay within it of textbook SIMPLS.  The
    reference draws pure noise for X and Y; with two warm-started power steps
    the agreement needs a spectral gap (on pure noise the measure is 0.2-0.8
    over 40 numpy draws), so the restatement draws the same shapes with a
    five-factor signal under unit noise."""
    rng = np.random.default_rng(1234)
    Z = rng.standard_normal((100, 5)) * np.array([8.0, 5.0, 3.0, 1.8, 1.1])
    X = Z @ rng.standard_normal((5, 25)) + rng.standard_normal((100, 25))
    Y = Z @ rng.standard_normal((5, 8)) + rng.standard_normal((100, 8))
    nc = [1, 2, 3, 4, 5]
    a = O.pls_model2_fast(X, Y, nc, seed=1)["B"]
    b = O.pls_model2_fast(X, Y, nc, seed=99)["B"]
    assert _all_equal_tol(a, b) < 7e-2
    ref = _textbook().simpls_textbook(X, Y, 5)
    assert _all_equal_tol(ref, b) < 7e-2


def test_simpls_fast_close_to_classic_on_controlled_task():
    """test-simpls-optimized.R:51-82 restated: on the two-factor regression
    task the accelerated SIMPLS's test RMSE at ncomp 6 is within 0.05 of the
    classic (textbook) SIMPLS's."""
    rng = np.random.default_rng(20260321)
    n = 120
    z1, z2 = rng.standard_normal(n), rng.standard_normal(n)
    X = np.column_stack([np.outer(z1, np.arange(1, 13)) + 0.08 * rng.standard_normal((n, 12)),
                         np.outer(z2, np.arange(1, 13)) + 0.08 * rng.standard_normal((n, 12))])
    Y = np.column_stack([1.5 * z1 - 0.4 * z2 + 0.05 * rng.standard_normal(n),
                         -0.7 * z1 + 1.8 * z2 + 0.05 * rng.standard_normal(n)])
    idx = rng.choice(n, 30, replace=False)
    tr = np.setdiff1d(np.arange(n), idx)
    fast = O.pls_model2_fast(X[tr], Y[tr], [2, 4, 6], seed=123)
    opt = O.pls_predict(fast, X[idx])[:, :, 2]
    Bc = _textbook().simpls_textbook(X[tr], Y[tr], 6)[:, :, 5]
    base = (X[idx] - X[tr].mean(0)) @ Bc + Y[tr].mean(0)
    rmse = lambda P: float(np.sqrt(np.mean((P - Y[idx]) ** 2)))  # noqa: E731
    assert abs(rmse(opt) - rmse(base)) <= 0.05


def test_plssvd_matches_textbook_fixture():
    """Golden fixture (test-plssvd-optimized.R:71-72's 1e-8 restated): with the
    exact SVD, B and Yfit equal least squares on the scores; the rSVD default
    takes the exact path here (l >= min(p, m))."""
    d = np.load(os.path.join(GOLDEN, "plssvd_textbook.npz"))
    for method in ("exact", "rsvd"):
        m = O.pls_model1(d["X"], d["Y"], d["ncomp"], fit=True, svd_method=method)
        assert rel_err(m["B"], d["B"]) < 1e-8
        assert rel_err(m["Yfit"], d["Yfit"]) < 1e-8


def test_predict_latent_equals_stored_coefficients():
    """test-predict-optimized.R:54 restated: prediction through the latent
    factors (no B stored) equals X B to 1e-12, for SIMPLS and PLSSVD."""
    rng = np.random.default_rng(54)
    X = rng.standard_normal((220, 8))
    Y = np.column_stack([0.7 * X[:, 0] - 0.4 * X[:, 1], -0.2 * X[:, 2] + 0.5 * X[:, 3]]) \
        + 0.15 * rng.standard_normal((220, 2))
    Xt = X[:40]
    for fit in (O.pls_model2_fast, O.pls_model1):
        m = fit(X[40:], Y[40:], [1, 3, 5])
        full = O.pls_predict(m, Xt)
        lat = dict(m)
        lat.pop("B", None)
        lat["B_stored"] = False
        np.testing.assert_allclose(O.pls_predict(lat, Xt), full, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- R generator and folds

def test_r_uniform_generator_known_answers():
    """R's set.seed(s); runif(5) (golden fixture r_runif.json)."""
    with open(os.path.join(GOLDEN, "r_runif.json")) as f:
        g = json.load(f)
    for seed, vals in g["runif"].items():
        rng = O.RUniform(int(seed))
        got = [rng.unif_rand() for _ in vals]
        np.testing.assert_allclose(got, vals, rtol=0, atol=0.6e-7)


def test_samplewithoutreplace_is_a_permutation():
    rng = O.RUniform(7)
    for n in (1, 2, 5, 40):
        perm = O.samplewithoutreplace(range(1, n + 1), n, rng)
        assert sorted(perm) == list(range(1, n + 1))


def _lib():
    from fastpls_amd import _lib as L
    return L.load()


def _lib_folds(cons, kfold, labels, C, seed):
    lib = _lib()
    cons = np.ascontiguousarray(np.asarray(cons, dtype=np.int32))
    fold = np.zeros(cons.size, dtype=np.int32)
    nf = ctypes.c_int32()
    lab = None if labels is None else np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
    rc = lib.fpls_cv_folds(cons.ctypes.data, cons.size, int(kfold),
                           None if lab is None else lab.ctypes.data, int(C),
                           int(seed) & 0xFFFFFFFF, fold.ctypes.data, ctypes.byref(nf))
    assert rc == 0
    return fold - 1, nf.value


def test_loocv_constrain_invariants_restated():
    """test-loocv-constrain.R:1-127: leave one constraint group out (kfold
    'loocv' = -1, or numeric kfold >= #groups): one fold per group, one group
    per fold, and folds numbered by the sorted group values."""
    for cons, kfold in ((np.repeat(np.arange(12), 2), -1), (np.repeat(np.arange(12), 2), 24),
                        (np.repeat([101, 203, 307, 409, 503, 601], 3), -1)):
        for fold, nf in (O.cv_folds(cons, kfold, None, 0, 11),
                         _lib_folds(cons, kfold, None, 0, 11)):
            assert nf == np.unique(cons).size
            for g in np.unique(cons):
                assert np.unique(fold[cons == g]).size == 1
            for f in np.unique(fold):
                assert np.unique(cons[fold == f]).size == 1
            np.testing.assert_array_equal(fold, np.searchsorted(np.unique(cons), cons))


@pytest.mark.parametrize("seed", [1, 12, 2026])
def test_stratified_folds_deal_each_class_round_robin(seed):
    """src/fastPLS.cpp:6737-6756: each class's groups (class of the group's
    first row) are permuted and dealt round-robin from fold 0, so class c with
    g_c groups puts floor(g_c / k) + (f < g_c mod k) of them in fold f, whatever
    the permutation; a group's rows share one fold."""
    rng = np.random.default_rng(seed)
    ng, k, C = 37, 5, 3
    cons = np.repeat(rng.permutation(ng) * 3 + 100, rng.integers(1, 4, size=ng))
    labels = rng.integers(1, C + 1, size=cons.size)
    fold, nf = O.cv_folds(cons, k, labels, C, seed)
    assert nf == k
    groups = np.unique(cons)
    first_cls = np.array([labels[np.nonzero(cons == g)[0][0]] for g in groups])
    gfold = np.array([fold[np.nonzero(cons == g)[0][0]] for g in groups])
    for g in groups:
        assert np.unique(fold[cons == g]).size == 1
    for c in range(1, C + 1):
        gc = int((first_cls == c).sum())
        for f in range(k):
            assert int(((first_cls == c) & (gfold == f)).sum()) == gc // k + (f < gc % k)
    # unstratified: the same rule over all groups
    fold_u, _ = O.cv_folds(cons, k, None, 0, seed)
    gf = np.array([fold_u[np.nonzero(cons == g)[0][0]] for g in groups])
    for f in range(k):
        assert int((gf == f).sum()) == ng // k + (f < ng % k)
    # the permutation depends on the seed
    assert not np.array_equal(O.cv_folds(cons, k, None, 0, seed + 1)[0], fold_u)


def test_fold_map_small_case_by_hand():
    """Four singleton groups, kfold 2, unstratified, seed 1: R's first uniforms
    after set.seed(1) are 0.2655087, 0.3721239, 0.5728534, 0.9082078, so
    samplewithoutreplace(1:4) takes positions floor(4 u1) = 1, floor(3 u2) = 1,
    floor(2 u3) = 1, 0 of the remaining values: perm = (2, 3, 4, 1), and group j
    goes to fold (perm[j] - 1) mod 2 = (1, 0, 1, 0)."""
    fold, nf = O.cv_folds([1, 2, 3, 4], 2, None, 0, 1)
    assert nf == 2
    np.testing.assert_array_equal(fold, [1, 0, 1, 0])


def test_library_fold_map_equals_oracle():
    """fpls_cv_folds (host code of the library, no device needed) against the
    oracle's restatement, bit for bit, over random shapes, seeds and kfold."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(2, 80))
        ng = int(rng.integers(1, n + 1))
        cons = (rng.integers(0, ng, size=n) * 5 - 7).astype(np.int32)
        C = int(rng.integers(2, 6))
        labels = rng.integers(1, C + 1, size=n) if rng.random() < 0.6 else None
        kfold = int(rng.integers(-1, 15))
        seed = int(rng.integers(0, 2 ** 31 - 1))
        fo, nfo = O.cv_folds(cons, kfold, labels, C if labels is not None else 0, seed)
        fl, nfl = _lib_folds(cons, kfold, labels, C if labels is not None else 0, seed)
        assert nfo == nfl
        np.testing.assert_array_equal(fl, fo)


# ---------------------------------------------------------------- CV driver

def test_cv_driver_equals_per_fold_fits():
    """pls_cv_predict's bookkeeping: every held-out row is predicted by a fit
    on exactly the other folds' rows, seeded seed + fold (10 stratified folds
    over singleton groups)."""
    rng = np.random.default_rng(44)
    n, p, C = 90, 12, 3
    y = rng.integers(1, C + 1, size=n)
    X = rng.standard_normal((n, p)) + np.eye(C)[y - 1] @ rng.standard_normal((C, p))
    nc = [1, 3]
    out = O.pls_cv_predict(X, y, np.arange(n), nc, "simpls", classification=True, n_classes=C,
                           seed=7, kfold=10)
    fold = out["fold"] - 1
    assert np.unique(fold).size == 10 and np.all(out["status"] == 1)
    Yd = np.eye(C)[y - 1]
    for f in range(10):
        te, tr = fold == f, fold != f
        mod = O.pls_model2_fast(X[tr], Yd[tr], nc, seed=7 + f)
        yp = O.pls_predict(mod, X[te])
        np.testing.assert_allclose(out["Ypred"][te], yp, rtol=1e-12, atol=1e-12)
        np.testing.assert_array_equal(out["class_pred"][te], np.argmax(yp, axis=1) + 1)


def test_cv_status_codes_by_hand():
    """Status 2 (empty test fold), 3 (no training rows) and 4 (one training
    class) with their fallback predictions (src/fastPLS.cpp:6807-6860)."""
    rng = np.random.default_rng(45)
    X = rng.standard_normal((6, 7))
    # three classes, one group each, kfold 2: every class deals its only group
    # to fold 0, so fold 0 holds every row (no training rows: status 3, the
    # observed classes) and fold 1 is empty (status 2)
    y = np.array([1, 1, 2, 2, 3, 3])
    out = O.pls_cv_predict(X, y, [1, 1, 2, 2, 3, 3], [1], classification=True, n_classes=3,
                           kfold=2)
    np.testing.assert_array_equal(out["status"], [3, 2])
    np.testing.assert_array_equal(out["class_pred"][:, 0], y)
    np.testing.assert_array_equal(out["Ypred"][:, :, 0], np.eye(3)[y - 1])
    # leave one group out where a held-out group takes a whole class: the
    # training rows have one class left (status 4: predicted as that class)
    y2 = np.array([1, 1, 1, 2, 2, 2])
    out = O.pls_cv_predict(X, y2, [1, 1, 1, 2, 2, 2], [1], classification=True, n_classes=2)
    np.testing.assert_array_equal(out["status"], [4, 4])
    np.testing.assert_array_equal(out["class_pred"][:, 0], [2, 2, 2, 1, 1, 1])
    np.testing.assert_array_equal(out["Ypred"][:3, :, 0], [[0, 1]] * 3)
    np.testing.assert_array_equal(out["Ypred"][3:, :, 0], [[1, 0]] * 3)


# ---------------------------------------------------------------- label-aware PLSSVD

def test_label_aware_plssvd_equals_dense_onehot_fit():
    """R/main.R:609-624: the label-aware model is the dense PLSSVD of the
    one-hot response up to the 1e-10 ridge of its latent solve: same S, same
    singular factors (sign-aligned), C/W and B = R W within 1e-7."""
    rng = np.random.default_rng(46)
    n, p, C = 300, 20, 6
    y = np.repeat(np.arange(1, C + 1), n // C)
    rng.shuffle(y)
    X = rng.standard_normal((n, p)) + 1.5 * np.eye(C)[y - 1] @ rng.standard_normal((C, p))
    nc = [2, 4, 5]
    la = O.plssvd_label_aware(X, y, nc)
    dn = O.pls_model1(X, np.eye(C)[y - 1], nc, svd_method="exact")
    np.testing.assert_allclose(la["mY"], dn["mY"], rtol=1e-15)
    R = align_signs(la["R"], dn["R"])
    assert rel_err(R, dn["R"][:, :R.shape[1]]) < 1e-10
    sg = np.sign(np.sum(la["R"] * dn["R"][:, :la["R"].shape[1]], axis=0))
    for a, k in enumerate(nc):
        D = np.diag(sg[:k])
        Cl = D @ la["C_latent"][:k, :k, a] @ D
        assert rel_err(Cl, dn["C_latent"][:k, :k, a]) < 1e-7
        B_la = la["R"][:, :k] @ la["W_latent"][:k, :, a]
        assert rel_err(B_la, dn["B"][:, :, a]) < 1e-7


# ---------------------------------------------------------------- cKNN score space

def _slice_model(ncomp, K, method="plssvd"):
    L = len(ncomp)
    Cl = np.zeros((K, K, L))
    for s in range(L):
        Cl[:, :, s] = 10.0 ** s  # slice s marked by the value 10^s
    return {"pls_method": method, "C_latent": Cl, "ncomp": np.array(ncomp)}


def test_candidate_score_space_slice_rule_by_hand():
    """R/main.R:700-730: T[:, :kk] C_latent[:kk, :kk, slice] with the slice of
    the exact ncomp match, else the first larger ncomp, else the last slice;
    the raw scores for non-PLSSVD models, missing or too small C_latent, or a
    non-finite block."""
    rng = np.random.default_rng(47)
    T = rng.standard_normal((7, 5))
    model = _slice_model([2, 5, 3], 5)
    expect_slice = {2: 0, 3: 2, 4: 1, 5: 1, 1: 0, 9: 1}  # worked out from the rule by hand
    for k, s in expect_slice.items():
        kk = min(k, 5)
        got = O.candidate_score_space(model, T, k)
        np.testing.assert_allclose(got, T[:, :kk] @ np.full((kk, kk), 10.0 ** s))
    # no ncomp >= k: the last slice
    m2 = _slice_model([1, 2], 5)
    np.testing.assert_allclose(O.candidate_score_space(m2, T, 4),
                               T[:, :4] @ np.full((4, 4), 10.0))
    # raw scores
    np.testing.assert_array_equal(O.candidate_score_space(_slice_model([2], 5, "simpls"), T, 2),
                                  T[:, :2])
    np.testing.assert_array_equal(O.candidate_score_space({"pls_method": "plssvd"}, T, 3),
                                  T[:, :3])
    small = _slice_model([2, 3], 3)
    np.testing.assert_array_equal(O.candidate_score_space(small, T, 4), T[:, :4])
    bad = _slice_model([2, 3], 5)
    bad["C_latent"][0, 0, 1] = np.nan
    np.testing.assert_array_equal(O.candidate_score_space(bad, T, 3), T[:, :3])


def test_candidate_centroids_empty_class_and_fit_by_loops():
    """R/main.R:686-698 and 795-873 by brute force: centroids are row-normalised
    class means of the row-normalised score space, an empty class gives a zero
    row (count 1, norm 1); the fit keeps the training scores and parameters."""
    rng = np.random.default_rng(48)
    Tn = O.candidate_row_l2(rng.standard_normal((9, 3)))
    y = np.array([1, 1, 3, 3, 3, 4, 4, 1, 4])  # class 2 empty
    cen = O.candidate_centroids(Tn, y, 4)
    for c in range(4):
        rows = [i for i in range(9) if y[i] == c + 1]
        if not rows:
            np.testing.assert_array_equal(cen[c], 0.0)
            continue
        mean = sum(Tn[i] for i in rows) / len(rows)
        np.testing.assert_allclose(cen[c], mean / np.sqrt((mean * mean).sum()), rtol=1e-14)
    # zero / non-finite norms map to 1 (R/main.R:657-662)
    Z = np.array([[0.0, 0.0], [3.0, 4.0]])
    np.testing.assert_allclose(O.candidate_row_l2(Z), [[0, 0], [0.6, 0.8]])
    # fit: Ttrain = Xc R, per k the normalised score space T_k C_k and centroids
    X = rng.standard_normal((40, 6))
    yy = np.repeat([1, 2, 3, 4], 10)
    mod = O.pls_model1(X, np.eye(4)[yy - 1], [2, 3])
    mod["lev"] = ["a", "b", "c", "d"]
    O.fit_candidate_knn(mod, X, yy, knn_k=3)
    ck = mod["candidate_knn"]
    T = (X - mod["mX"]) @ mod["R"][:, :3]
    np.testing.assert_allclose(ck["Ttrain"], T, rtol=1e-10, atol=1e-12)
    for a, k in enumerate([2, 3]):
        S = T[:, :k] @ mod["C_latent"][:k, :k, a]
        Sn = S / np.linalg.norm(S, axis=1, keepdims=True)
        np.testing.assert_allclose(ck["models"][k]["centroids"],
                                   O.candidate_centroids(Sn, yy, 4), rtol=1e-10)
    assert ck["parameters"] == {"knn_k": 3, "tau": 0.2, "alpha": 0.75, "top_m": 20}


def test_plssvd_xprod_route_equals_formed_s():
    """pls_model1_rsvd_xprod_precision ("implicit64", src/fastPLS.cpp:5518-5690)
    against pls_model1 on the formed S with the same Gaussian sketch: the
    range finder sees S Omega either way, so R, s, B and Yfit agree to
    rounding; a missing centring or scaling term in the matrix-free products
    (a_times = Xs^T (Yc M)) would not."""
    rng = np.random.default_rng(41)
    n, p, q = 160, 70, 40
    X = rng.standard_normal((n, p)) @ rng.standard_normal((p, p)) * 0.3 + 2.0
    Y = X[:, :q] * 0.5 + rng.standard_normal((n, q)) + 1.0
    for scaling in (1, 2):
        a = O.pls_model1(X, Y, [2, 5], scaling, True, "rsvd", 6, 1, 3, xprod=True)
        b = O.pls_model1(X, Y, [2, 5], scaling, True, "rsvd", 6, 1, 3)
        assert a["xprod_mode"] == "implicit" and "xprod_mode" not in b
        np.testing.assert_allclose(a["s"], b["s"], rtol=1e-10)
        assert rel_err(align_signs(a["R"], b["R"]), b["R"]) < 1e-9
        assert rel_err(a["B"], b["B"]) < 1e-9
        assert rel_err(a["Yfit"], b["Yfit"]) < 1e-10
    # with one operand centred the other's centring cancels (1^T Yc = 0), but
    # dropping the autoscaling of X in the products changes the result
    Xs, Ys, *_ = O._preprocess(X, Y, 2)
    Xc, *_ = O._preprocess(X, Y, 1)
    bad = O.truncated_rsvd_crossprod(Xc, Ys, 5, 6, 1, 3)
    good = O.truncated_rsvd_crossprod(Xs, Ys, 5, 6, 1, 3)
    assert np.abs(bad["s"] / bad["s"][0] - good["s"] / good["s"][0]).max() > 1e-3
