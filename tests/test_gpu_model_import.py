"""GPU: models in the reference's list format (fpls_model_import) and the
B_stored policy (src/fastPLS.cpp:63-90, 3737-3836).  A model fitted by the FP64
oracle (the reference's algorithm on the CPU) is imported and predicted on the
device: predictions agree with the oracle's pls_predict to FP64 rounding, and
ranked classes with its pls_class_predict_topk.  A device fit exported through
its fields and imported again predicts bit-identically.  The classifier
sub-lists (lda, R/main.R:1244-1250; candidate_knn, R/main.R:861-872) of an
oracle-fitted model are imported onto it and predict with the oracle's labels,
barring near-ties of the oracle's own scores."""
import os

import numpy as np
import pytest

from oracle import fastpls_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def _data(seed, n=200, p=30, m=4):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, p))
    Y = X[:, :m] @ rng.standard_normal((m, m)) + 0.3 * rng.standard_normal((n, m))
    Xt = rng.standard_normal((40, p))
    return X, Y, Xt


def test_import_oracle_simpls_and_predict(ctx):
    import fastpls_amd as F
    X, Y, Xt = _data(1)
    mod = O.pls_model2_fast(X, Y, [1, 3, 5], 1, False, "rsvd", 10, 1, 7)
    mod["pls_method"] = "simpls"
    dm = F.model_from_fields(mod)
    got = F.predict(dm, Xt)["Ypred"]
    assert rel(got, O.pls_predict(mod, Xt)) < 1e-12
    # the latent projection of the imported model
    T = F.predict(dm, Xt, proj=True)["Ttest"]
    ref = ((Xt - mod["mX"]) / mod["vX"]) @ mod["R"]
    assert rel(T, ref) < 1e-12


@pytest.mark.parametrize("drop_w", [False, True])
def test_import_oracle_plssvd(ctx, drop_w):
    """PLSSVD with W_latent, or with only C_latent + Q (W_latent formed on the
    device as C_a[:mc, :mc] Q[:, :mc]^T, the reference's third branch)."""
    import fastpls_amd as F
    X, Y, Xt = _data(2)
    mod = O.pls_model1(X, Y, [1, 2, 4], 2, False, "rsvd", 10, 1, 5)
    fields = dict(mod)
    fields["pls_method"] = "plssvd"
    if drop_w:
        fields.pop("W_latent")
        fields.pop("B", None)
    dm = F.model_from_fields(fields)
    assert rel(F.predict(dm, Xt)["Ypred"], O.pls_predict(mod, Xt)) < 1e-12


def test_import_classifier_ranks_like_the_oracle(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(3)
    X = rng.standard_normal((300, 25))
    y = rng.integers(1, 7, 300)
    Xt = rng.standard_normal((80, 25))
    mod = O.pls_model1(X, O.transformy(y), [2, 5], 1, False, "rsvd", 10, 1, 9)
    fields = dict(mod, pls_method="plssvd", classification=True, lev=list("abcdef"))
    dm = F.model_from_fields(fields)
    res = F.predict(dm, Xt, top=3)
    ri, rs = O.pls_class_predict_topk(mod, Xt, None, 3)
    np.testing.assert_allclose(res["Ypred_top_score"], rs, rtol=0, atol=1e-11)
    gap = np.abs(np.diff(rs, axis=1)).min(axis=1)
    assert np.all(np.all(res["Ypred_top"] == ri, axis=1) | (gap < 1e-9))
    assert set(res["Ypred"].ravel()) <= set("abcdef")


def test_export_import_round_trip_is_bitwise(ctx):
    import fastpls_amd as F
    X, Y, Xt = _data(4)
    for method in ("simpls", "plssvd"):
        m = F.pls(X, Y, ncomp=[1, 2, 3], method=method, seed=11)
        dm = F.model_from_fields(dict(m))
        np.testing.assert_array_equal(F.predict(dm, Xt)["Ypred"], F.predict(m, Xt)["Ypred"])


def test_import_errors(ctx):
    import fastpls_amd as F
    X, Y, Xt = _data(5)
    mod = O.pls_model2_fast(X, Y, [1, 2], 1, False, "rsvd", 10, 1, 7)
    bad = dict(mod, pls_method="simpls")
    bad.pop("Q")
    with pytest.raises(F.FplsError, match="Q"):
        F.model_from_fields(bad)
    with pytest.raises(F.FplsError, match="ncomp exceeds latent rank"):
        F.model_from_fields(dict(mod, pls_method="simpls", ncomp=[1, 99]))
    pm = O.pls_model1(X, Y, [1, 2], 1, False, "rsvd", 10, 1, 5)
    pm = dict(pm, pls_method="plssvd")
    pm.pop("W_latent")
    pm.pop("C_latent")
    with pytest.raises(F.FplsError, match="compact low-rank factors"):
        F.model_from_fields(pm)


def test_b_storage_policy_environment(ctx):
    """FASTPLS_STORE_B / FASTPLS_STORE_B_MAX_MB as should_store_coefficients."""
    import fastpls_amd as F
    X, Y, _ = _data(6)
    saved = {k: os.environ.get(k) for k in ("FASTPLS_STORE_B", "FASTPLS_STORE_B_MAX_MB")}
    try:
        os.environ.pop("FASTPLS_STORE_B_MAX_MB", None)
        os.environ["FASTPLS_STORE_B"] = "never"
        assert not F.pls(X, Y, ncomp=2)["B_stored"]
        os.environ["FASTPLS_STORE_B"] = "Always"
        assert F.pls(X, Y, ncomp=2, method="plssvd")["B_stored"]
        os.environ["FASTPLS_STORE_B"] = "auto"
        assert F.pls(X, Y, ncomp=2)["B_stored"]           # 30 x 4 x 8 bytes << 256 MB
        os.environ["FASTPLS_STORE_B_MAX_MB"] = "0"
        m = F.pls(X, Y, ncomp=2)
        assert not m["B_stored"] and m["compact_prediction"]
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


# ---- classifier heads (R/main.R:1244-1250, 861-872) --------------------------

def _class_data(seed, n=360, p=28, C=6, nt=90):
    rng = np.random.default_rng(seed)
    mu = 1.5 * rng.standard_normal((C, p))
    y = rng.integers(1, C + 1, n)
    y[:C] = np.arange(1, C + 1)
    X = mu[y - 1] + rng.standard_normal((n, p))
    yt = rng.integers(1, C + 1, nt)
    Xt = mu[yt - 1] + rng.standard_normal((nt, p))
    return X, y, Xt


def _oracle_plssvd_lda(X, y, C, ks):
    """A PLSSVD model with its lda list, fitted by the oracle (the reference's
    CPU path: pls_model1 + .attach_lda_classifier's lda_train_prefix_cpp)."""
    mod = O.pls_model1(X, O.transformy(y), ks, 1, False, "rsvd", 10, 1, 9)
    mod["pls_method"] = "plssvd"
    O.attach_latent_projection_cache(mod)
    T = O.latent_scores(mod, X, max(ks))
    models = O.lda_train_prefix(T, y, C, ks, 1e-8)
    mod["lda"] = {"ncomp": ks, "models": {m_["ncomp"]: m_ for m_ in models}, "ridge": 1e-8}
    mod["classification_rule"] = "lda_cpp"
    mod["lev"] = [f"c{i}" for i in range(1, C + 1)]
    return mod


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 2e-3)])
def test_import_oracle_lda_head_predicts_like_the_oracle(ctx, dtype, tol):
    """lda_project_predict_labels_cpp on the imported head: labels equal the
    oracle's except where its top-2 discriminant margin is below tol (FP64
    input: FP64 products end to end; FP32 input: the FP32 MFMA projection)."""
    import fastpls_amd as F
    from tests.helpers import label_mismatches_excused
    C, ks = 6, [2, 4, 5]
    X, y, Xt = _class_data(11, C=C)
    mod = _oracle_plssvd_lda(X, y, C, ks)
    dm = F.model_from_fields(mod)
    assert dm["classification_rule"] == "lda_hip"
    res = F.predict(dm, Xt.astype(dtype), raw_scores=True, tie_log=True)
    for a, k in enumerate(ks):
        lda = mod["lda"]["models"][k]
        lab, sc = O.lda_project_predict_labels(Xt, mod["R_predict"][:, :k], mod["R_offset"],
                                               lda, return_scores=True)
        bad, unexcused = label_mismatches_excused(res["Ypred_index"][:, a], lab, sc, tol,
                                                  res["tie_margin"][:, a])
        assert not unexcused, f"ncomp={k}: non-tie label mismatches at rows {unexcused[:10]}"
        assert len(bad) <= 2
        np.testing.assert_allclose(res["LDA_scores"][:, :, a], sc, rtol=0,
                                   atol=tol * max(1.0, np.abs(sc).max()))
        assert list(res["Ypred"][:, a]) == [mod["lev"][i - 1] for i in res["Ypred_index"][:, a]]
        # the tie log is the device's own winning margin, to the last bit
        srt = np.sort(res["LDA_scores"][:, :, a], axis=1)
        np.testing.assert_array_equal(res["tie_margin"][:, a], srt[:, -1] - srt[:, -2])
    # the imported fields read back unchanged
    for a, k in enumerate(ks):
        lda = mod["lda"]["models"][k]
        np.testing.assert_array_equal(dm.lda_read(a, "linear", (C, k)), lda["linear"])
        np.testing.assert_array_equal(dm.lda_read(a, "constants", (C,)), lda["constants"])
        np.testing.assert_array_equal(dm.lda_read(a, "inv_cov", (k, k)), lda["inv_cov"])
        np.testing.assert_array_equal(dm.lda_read(a, "means", (C, k)), lda["means"])
        np.testing.assert_array_equal(dm.lda_read(a, "priors", (C,)), lda["priors"])


def test_device_lda_head_round_trips_through_import(ctx):
    """A device-fitted LDA head exported with fpls_lda_read and imported onto
    the imported model predicts labels and scores bit-identically (FP64)."""
    import fastpls_amd as F
    C = 5
    X, y, Xt = _class_data(12, C=C)
    m = F.pls(X, y, ncomp=[2, 3], method="plssvd", classifier="lda", seed=3)
    ks = m["lda"]["ncomp"]
    models = {k: {"linear": m.lda_read(a, "linear", (C, k)),
                  "constants": m.lda_read(a, "constants", (C,))} for a, k in enumerate(ks)}
    fields = dict(m)
    fields["lda"] = {"ncomp": ks, "models": models}
    fields["classification_rule"] = "lda_cpp"
    dm = F.model_from_fields(fields)
    a = F.predict(m, Xt, raw_scores=True)
    b = F.predict(dm, Xt, raw_scores=True)
    np.testing.assert_array_equal(a["Ypred_index"], b["Ypred_index"])
    np.testing.assert_array_equal(a["LDA_scores"], b["LDA_scores"])
    with pytest.raises(F.FplsError, match="no such field"):
        dm.lda_read(0, "means", (C, ks[0]))


@pytest.mark.parametrize("method", ["plssvd", "simpls"])
def test_import_oracle_cknn_head_predicts_like_the_oracle(ctx, method):
    """.candidate_knn_predictions on the imported candidate_knn list: the
    ranked classes equal the oracle's except at near-ties of its scores
    (1e-5: FP32 similarities), scores to 1e-5.  PLSSVD maps the scores
    through C_latent (the .candidate_score_space slice rule)."""
    import fastpls_amd as F
    from tests.helpers import topk_mismatches_excused
    C, ks = 6, [2, 3, 5]
    X, y, Xt = _class_data(13, C=C)
    if method == "plssvd":
        mod = O.pls_model1(X, O.transformy(y), ks, 1, False, "rsvd", 10, 1, 9)
    else:
        mod = O.pls_model2_fast(X, O.transformy(y), ks, 1, False, "rsvd", 10, 1, 9)
    mod["pls_method"] = method
    mod["lev"] = list(range(1, C + 1))
    O.fit_candidate_knn(mod, X, y, knn_k=4, tau=0.2, alpha=0.75, top_m=5)
    mod["classification_rule"] = "candidate_knn_cpp"
    dm = F.model_from_fields(mod)
    assert dm["classification_rule"] == "candidate_knn_hip"
    top = 3
    res = F.predict(dm, Xt, top=top, tie_log=True)
    ref = O.candidate_knn_predictions(mod, Xt, top=top)
    for a in range(len(ks)):
        ri, rs = ref[a]["top_index"], ref[a]["top_score"]
        ti = res["Ypred_top"][:, :, a]
        bad = topk_mismatches_excused(ti, ri, rs, 1e-5, res["Ypred_top_score"][:, :, a],
                                      res["tie_margin"][:, :, a])
        assert not bad, f"ncomp={ks[a]}: non-tie top-k mismatches at rows {bad[:10]}"
        np.testing.assert_allclose(res["Ypred_top_score"][:, :, a], rs, rtol=0, atol=1e-5)


def test_import_heads_errors(ctx):
    import fastpls_amd as F
    C, ks = 4, [2, 3]
    X, y, _ = _class_data(14, C=C)
    mod = _oracle_plssvd_lda(X, y, C, ks)
    bad = dict(mod, lda=dict(mod["lda"], ncomp=[3, 2]))
    with pytest.raises(F.FplsError, match="strictly ascending"):
        F.model_from_fields(bad)
    bad = dict(mod, lda=dict(mod["lda"], ncomp=[2, 99],
                             models={2: mod["lda"]["models"][2], 99: mod["lda"]["models"][3]}))
    with pytest.raises(ValueError, match="inconsistent"):
        F.model_from_fields(bad)
    ck = dict(mod)
    ck.pop("lda")
    O.fit_candidate_knn(ck, X, y)
    yc = ck["candidate_knn"]["y_codes"].copy()
    yc[5] = C + 1
    with pytest.raises(F.FplsError, match="outside the training levels"):
        F.model_from_fields(dict(ck, candidate_knn=dict(ck["candidate_knn"], y_codes=yc)))
    par = dict(ck["candidate_knn"]["parameters"], tau=0.0)
    with pytest.raises(F.FplsError, match="candidate_tau"):
        F.model_from_fields(dict(ck, candidate_knn=dict(ck["candidate_knn"], parameters=par)))
