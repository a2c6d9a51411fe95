"""GPU parity: Gaussian response compression (fpls_gaussian_y_*) against the
oracle restatement, and pls(..., gaussian_y = TRUE) mirroring the reference's
tests/testthat/test-gaussian-y.R (flags, default dimension, prediction
shapes, factor predictions)."""
import numpy as np
import pytest

from oracle import fastpls_oracle as O

pytestmark = pytest.mark.gpu


def test_encode_regression_matches_oracle(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(11)
    Y = rng.standard_normal((300, 24)) * 3.0 + 1.5
    X = rng.standard_normal((300, 40))
    got = F.prepare_gaussian_y(Y, X, True, 7, seed=120)
    ref = O.prepare_gaussian_y_regression(Y, 7, 120)
    g = got["gaussian"]
    assert g["task"] == "regression" and g["dim"] == 7 and g["original_dim"] == 24
    np.testing.assert_allclose(g["y_center"], ref["y_center"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(got["Ytrain"], ref["Z"], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(g["decoder"], ref["decoder"], rtol=1e-8, atol=1e-9)
    # decode on the device equals the oracle decode
    Zt = ref["Z"] + 0.01 * rng.standard_normal(ref["Z"].shape)
    np.testing.assert_allclose(F.decode_gaussian_y_matrix(g, Zt),
                               O.decode_gaussian_y_regression(Zt, ref["decoder"], ref["y_center"]),
                               rtol=1e-8, atol=1e-9)


def test_encode_regression_fp32_device(ctx):
    import torch
    import fastpls_amd as F
    rng = np.random.default_rng(12)
    Y = rng.standard_normal((2000, 50)).astype(np.float32)
    dY = F.device_colmajor(Y)
    got = F.prepare_gaussian_y(dY, np.zeros((1, 100)), True, 30, seed=4)
    ref = O.prepare_gaussian_y_regression(Y.astype(np.float64), 30, 4)
    Z = got["Ytrain"]
    assert Z.is_cuda and Z.shape == (2000, 30) and Z.stride() == (1, 2000)
    np.testing.assert_allclose(Z.cpu().numpy(), ref["Z"], rtol=1e-11, atol=1e-11)
    # default dimension: min(ncol(X), 100)
    assert F.gaussian_y_default_dim(np.zeros((3, 150))) == 100
    torch.cuda.synchronize()


def test_class_codes_and_nearest_code(ctx):
    import fastpls_amd as F
    rng = np.random.default_rng(13)
    y = rng.choice(list("abcdef"), 90)
    got = F.prepare_gaussian_y(y, np.zeros((90, 12)), True, 5, seed=121)
    lev, codes_i = np.unique(y, return_inverse=True)
    ref = O.prepare_gaussian_y_classes(codes_i + 1, 6, 5, 121)
    g = got["gaussian"]
    assert g["task"] == "classification" and got["lev"] == list(lev)
    # the codebook agrees to rounding (summation order); the gather is exact
    np.testing.assert_allclose(g["codes"], ref["codes"], rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(got["Ytrain"], g["codes"][codes_i])
    np.testing.assert_allclose(got["Ytrain"], ref["Z"], rtol=1e-14, atol=1e-15)
    Zn = ref["Z"] + 0.4 * rng.standard_normal(ref["Z"].shape)
    scores = F.decode_gaussian_y_matrix(g, Zn)
    np.testing.assert_allclose(scores, O.gaussian_y_class_scores(Zn, ref["codes"]), atol=1e-12)
    np.testing.assert_array_equal(np.argmax(scores, axis=1) + 1,
                                  O.nearest_code_classes(Zn, ref["codes"]))


def test_pls_gaussian_y_regression_like_reference_tests(ctx):
    """tests/testthat/test-gaussian-y.R:3-58 with this library's draws."""
    import fastpls_amd as F
    rng = np.random.default_rng(120)
    X = rng.standard_normal((70, 15))
    Y = rng.standard_normal((70, 24))
    Xt = rng.standard_normal((9, 15))
    Yt = rng.standard_normal((9, 24))
    fit = F.pls(X, Y, Xt, Yt, ncomp=[1, 2, 3], method="simpls", gaussian_y=True,
                gaussian_y_dim=7, seed=120)
    assert fit["gaussian_y_task"] == "regression" and fit["gaussian_y_dim"] == 7
    assert fit["Ypred"].shape == (9, 24, 3)
    assert len(fit["Q2Y"]) == 3 and np.all(np.isfinite(fit["Ypred"]))
    # the decoded prediction equals the oracle pipeline on the same draw
    enc = O.prepare_gaussian_y_regression(Y, 7, 120)
    mod = O.pls_model2_fast(X, enc["Z"], [1, 2, 3], 1, False, "rsvd", 10, 1, 120)
    zp = O.pls_predict(mod, Xt)
    for a in range(3):
        ref = O.decode_gaussian_y_regression(zp[:, :, a], enc["decoder"], enc["y_center"])
        np.testing.assert_allclose(fit["Ypred"][:, :, a], ref, rtol=1e-7, atol=1e-8)
    default_fit = F.pls(X, Y, Xt, Yt, ncomp=2, method="plssvd", gaussian_y=True, seed=120)
    assert default_fit["gaussian_y_dim"] == 15
    assert default_fit["Ypred"].shape == (9, 24, 1)
    small = F.pls(X, Y[:, :3], Xt, Yt[:, :3], ncomp=2, method="plssvd", gaussian_y=True,
                  seed=120)
    assert small["gaussian_y_dim"] == 15 and small["Ypred"].shape == (9, 3, 1)


def test_pls_gaussian_y_classification_like_reference_tests(ctx):
    """tests/testthat/test-gaussian-y.R:60-89 with this library's draws."""
    import fastpls_amd as F
    rng = np.random.default_rng(121)
    X = rng.standard_normal((90, 12))
    y = rng.choice(list("abcdef"), 90)
    Xt = rng.standard_normal((11, 12))
    fit = F.pls(X, y, Xt, ncomp=2, method="simpls", gaussian_y=True, gaussian_y_dim=5,
                seed=121)
    assert fit["gaussian_y_task"] == "classification"
    assert fit["lev"] == sorted(set(y))
    assert fit["Ypred"].shape[0] == 11 and set(fit["Ypred"].ravel()) <= set(y)
    enc = O.prepare_gaussian_y_classes(np.unique(y, return_inverse=True)[1] + 1, 6, 5, 121)
    mod = O.pls_model2_fast(X, enc["Z"], [2], 1, False, "rsvd", 10, 1, 121)
    ref = O.nearest_code_classes(O.pls_predict(mod, Xt)[:, :, 0], enc["codes"])
    np.testing.assert_array_equal(fit["Ypred_index"][:, 0], ref)


def test_gaussian_y_argument_errors(ctx):
    import fastpls_amd as F
    X = np.zeros((10, 4))
    with pytest.raises(ValueError):
        F.prepare_gaussian_y(np.ones((10, 3)), X, True, 0)
    with pytest.raises(ValueError):
        F.pls(X, np.array(list("ab") * 5), ncomp=1, gaussian_y=True, classifier="knn")


def test_gaussian_y_through_opls_and_kernelpls(ctx):
    """tests/testthat/test-gaussian-y.R:89-122 with this library's draws: the
    compressed response through the OPLS filter and kernel PLS; the OPLS
    prediction equals the oracle pipeline (filter on the original response,
    SIMPLS on the encoded one, decode)."""
    import fastpls_amd as F
    rng = np.random.default_rng(122)
    X = rng.standard_normal((55, 10))
    Y = rng.standard_normal((55, 18))
    Xt = rng.standard_normal((8, 10))
    Yt = rng.standard_normal((8, 18))
    of = F.pls(X, Y, Xt, Yt, ncomp=2, method="opls", gaussian_y=True, gaussian_y_dim=6, seed=3)
    kf = F.pls(X, Y, Xt, Yt, ncomp=2, method="kernelpls", kernel="linear", gaussian_y=True,
               gaussian_y_dim=6, seed=3)
    assert of["Ypred"].shape == (8, 18, 1) and kf["Ypred"].shape == (8, 18, 1)
    assert of["gaussian_y"] and kf["gaussian_y"]
    enc = O.prepare_gaussian_y_regression(Y, 6, 3)
    filt = O.opls_filter(X, Y, 1, 1)
    mod = O.pls_model2_fast(filt["X"], enc["Z"], [2], 3, False, "rsvd", 10, 1, 3)
    zt = O.pls_predict(mod, O.opls_apply_filter(Xt, filt["mX"], filt["vX"], filt["W_orth"],
                                                filt["P_orth"]))
    ref = O.decode_gaussian_y_regression(zt[:, :, 0], enc["decoder"], enc["y_center"])
    np.testing.assert_allclose(of["Ypred"][:, :, 0], ref, rtol=1e-7, atol=1e-8)


def test_gaussian_y_with_the_lda_head(ctx):
    """.attach_lda_classifier after a Gaussian-Y classification fit
    (R/main.R:6197-6203): the LDA head trained with the original labels on the
    latent scores of the model of the codes; labels equal the oracle's."""
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, _ = synth.classes_on_sphere(n_classes=6, per_class_train=60,
                                               per_class_test=15, p=30, seed=9)
    Xtr, Xte = Xtr.astype(np.float64), Xte.astype(np.float64)
    fit = F.pls(Xtr, ytr, Xte, ncomp=[2, 4], method="simpls", gaussian_y=True,
                gaussian_y_dim=5, classifier="lda", seed=4)
    assert fit["classification_rule"] == "lda_hip" and fit["gaussian_y"]
    codes = np.unique(ytr, return_inverse=True)[1] + 1
    enc = O.prepare_gaussian_y_classes(codes, 6, 5, 4)
    mod = O.pls_model2_fast(Xtr, enc["Z"], [2, 4], 1, False, "rsvd", 10, 1, 4)
    O.attach_latent_projection_cache(mod)
    Rp = mod["R_predict"]
    lda = O.lda_project_train_prefix(Xtr, Rp[:, :4], mod["R_offset"][:4], codes, 6, [2, 4], 1e-8)
    for s_, kk in enumerate([2, 4]):
        lab, sc = O.lda_project_predict_labels(Xte, Rp[:, :kk], mod["R_offset"][:kk], lda[s_],
                                               return_scores=True)
        srt = np.sort(sc, axis=1)
        tie = (srt[:, -1] - srt[:, -2]) < 1e-8 * np.abs(sc).max()
        assert np.all((fit["Ypred_index"][:, s_] == lab) | tie)


@pytest.mark.parametrize("method", ["simpls", "plssvd"])
def test_nearest_code_ranking_from_the_factors(ctx, method):
    """fpls_gaussian_y_predict_classes (the code scores ranked straight from the
    compact factors) against the decoded score cube ranked by
    .class_topk_from_score_cube (the raw-score path)."""
    import fastpls_amd as F
    rng = np.random.default_rng(17)
    X = rng.standard_normal((400, 20))
    y = rng.choice([f"k{i}" for i in range(9)], 400)
    Xt = rng.standard_normal((120, 20))
    fit = F.pls(X, y, ncomp=[2, 5], method=method, gaussian_y=True, gaussian_y_dim=6, seed=5)
    fast = F.predict(fit, Xt, top=4)
    full = F.predict(fit, Xt, top=4, raw_scores=True)
    np.testing.assert_allclose(fast["Ypred_top_score"], full["Ypred_top_score"], rtol=0,
                               atol=1e-10)
    gap = np.abs(np.diff(full["Ypred_top_score"], axis=1)).min(axis=1)
    assert np.all(np.all(fast["Ypred_top"] == full["Ypred_top"], axis=1) | (gap < 1e-9))
    np.testing.assert_array_equal(fast["Ypred"][gap.min(axis=1) > 1e-9],
                                  full["Ypred"][gap.min(axis=1) > 1e-9])
