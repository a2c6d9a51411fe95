"""GPU: models in the reference's list format (fpls_model_import) and the
B_stored policy (src/fastPLS.cpp:63-90, 3737-3836).  A model fitted by the FP64
oracle (the reference's algorithm on the CPU) is imported and predicted on the
device: predictions agree with the oracle's pls_predict to FP64 rounding, and
ranked classes with its pls_class_predict_topk.  A device fit exported through
its fields and imported again predicts bit-identically."""
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
