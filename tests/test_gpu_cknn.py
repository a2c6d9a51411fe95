"""GPU parity: candidate kNN head against the FP64 oracle.

Similarities run in FP32 on MFMA, smoothing in FP64.  top-k class indices
must match the oracle exactly except where the oracle's own score gap between
the disagreeing positions is below 1e-5 (a documented near-tie); scores agree
to 1e-5 absolute.
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O
from tests.helpers import sin_theta, topk_mismatches_excused

pytestmark = pytest.mark.gpu


def _check_topk(ti, ts, ref, tol=1e-5, gpu_margin=None):
    ri, rs = ref["top_index"], ref["top_score"]
    assert ti.shape == ri.shape
    # excuse only near-ties in both the oracle's and the device's scores
    bad = topk_mismatches_excused(ti, ri, rs, tol, ts, gpu_margin)
    assert not bad, f"rows with non-tie top-k mismatches: {bad[:10]}"
    fin = np.isfinite(rs)
    np.testing.assert_allclose(np.sort(ts, axis=1)[fin.all(1)], np.sort(rs, axis=1)[fin.all(1)],
                               atol=tol)


@pytest.mark.parametrize("C,k,ntr,nte,knn_k,top_m,top", [
    (12, 16, 900, 200, 5, 6, 3),
    (40, 64, 5000, 300, 10, 20, 5),
    (7, 9, 300, 100, 1, 7, 1),          # k = 1: nearest neighbour inside the candidates
])
def test_cknn_scores_match_oracle(ctx, C, k, ntr, nte, knn_k, top_m, top):
    import fastpls_amd as F
    rng = np.random.default_rng(C * 1000 + k)
    mu = rng.standard_normal((C, k))
    ytr = rng.integers(1, C + 1, ntr).astype(np.int32)
    ytr[:C] = np.arange(1, C + 1)
    Ttr = O.candidate_row_l2(mu[ytr - 1] + 0.8 * rng.standard_normal((ntr, k)))
    yte = rng.integers(1, C + 1, nte)
    Tte = O.candidate_row_l2(mu[yte - 1] + 0.8 * rng.standard_normal((nte, k)))
    cen = O.candidate_centroids(Ttr, ytr, C)
    got = F.candidate_knn_predict(Tte, Ttr, ytr, cen, None, top, top_m, knn_k, 0.2, 0.75)
    ref = O.candidate_knn_predict(Tte, Ttr, ytr, cen, None, top, top_m, knn_k, 0.2, 0.75)
    _check_topk(got["top_index"], got["top_score"], ref, gpu_margin=got["tie_margin"])


def test_cknn_large_class_spans_chunks(ctx):
    """One class with > 4096 rows exercises the partial-list merge."""
    import fastpls_amd as F
    rng = np.random.default_rng(3)
    C, k = 3, 8
    sizes = [9000, 500, 50]
    ytr = np.concatenate([np.full(s, c + 1) for c, s in enumerate(sizes)]).astype(np.int32)
    rng.shuffle(ytr)
    mu = rng.standard_normal((C, k))
    Ttr = O.candidate_row_l2(mu[ytr - 1] + rng.standard_normal((ytr.size, k)))
    Tte = O.candidate_row_l2(rng.standard_normal((64, k)))
    cen = O.candidate_centroids(Ttr, ytr, C)
    got = F.candidate_knn_predict(Tte, Ttr, ytr, cen, None, 3, 3, 10, 0.2, 0.75)
    ref = O.candidate_knn_predict(Tte, Ttr, ytr, cen, None, 3, 3, 10, 0.2, 0.75)
    _check_topk(got["top_index"], got["top_score"], ref, gpu_margin=got["tie_margin"])


def test_pls_cknn_pipeline(ctx):
    """pls(method="plssvd", classifier="cknn") then predict(top=5)."""
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=15, per_class_train=120,
                                                 per_class_test=20, p=64, seed=21)
    ncomp = 10
    m = F.pls(Xtr, ytr, method="plssvd", ncomp=ncomp, classifier="cknn", candidate_knn_k=10,
              candidate_tau=0.2, candidate_alpha=0.75, candidate_top_m=8)
    res = F.predict(m, Xte, top=5, tie_log=True)
    ref = O.pls_model1(Xtr.astype(np.float64), O.transformy(ytr), [ncomp])
    ref["lev"] = list(range(1, 16))
    O.fit_candidate_knn(ref, Xtr.astype(np.float64), ytr, 10, 0.2, 0.75, 8)
    rr = O.candidate_knn_predictions(ref, Xte.astype(np.float64), top=5)[0]
    _check_topk(res["Ypred_top"][:, :, 0], res["Ypred_top_score"][:, :, 0], rr, tol=1e-4,
                gpu_margin=res["tie_margin"][:, :, 0])
    assert np.mean(res["Ypred_index"][:, 0] == yte) > 0.9


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_label_aware_cknn_many_classes_large_ncomp(ctx, precision):
    """The large-class-count label-aware PLSSVD + candidate-kNN path at reduced
    size: Zipf class sizes, ncomp above 128 (exact SVD and the latent solve
    take the block-Jacobi eigensolver), in both contraction forms.  Against
    the oracle's label-aware model and candidate-kNN predictions."""
    import fastpls_amd as F
    rng = np.random.default_rng(41)
    C, p, n, nt = 180, 256, 9000, 400
    w = 1.0 / np.arange(1, C + 1) ** 1.1
    w /= w.sum()
    mu = rng.standard_normal((C, p))
    mu *= 4.0 / np.linalg.norm(mu, axis=1, keepdims=True)
    ytr = np.concatenate([np.arange(1, C + 1), rng.choice(np.arange(1, C + 1), n - C, p=w)])
    yte = rng.choice(np.arange(1, C + 1), nt, p=w)
    Xtr = (mu[ytr - 1] + 0.6 * rng.standard_normal((n, p))).astype(np.float32)
    Xte = (mu[yte - 1] + 0.6 * rng.standard_normal((nt, p))).astype(np.float32)
    ytr = ytr.astype(np.int32)
    ncomp = 150
    m = F.pls(Xtr, ytr, method="plssvd", ncomp=ncomp, classifier="cknn", label_aware=True,
              candidate_knn_k=10, candidate_tau=0.2, candidate_alpha=0.75, candidate_top_m=20,
              precision=precision)
    res = F.predict(m, Xte, top=5, tie_log=True)
    ref = O.plssvd_label_aware(Xtr.astype(np.float64), ytr, [ncomp])
    assert sin_theta(m["R"][:, :ncomp], ref["R"][:, :ncomp]) < 1e-4
    ref["lev"] = list(range(1, C + 1))
    O.fit_candidate_knn(ref, Xtr.astype(np.float64), ytr, 10, 0.2, 0.75, 20)
    rr = O.candidate_knn_predictions(ref, Xte.astype(np.float64), top=5)[0]
    _check_topk(res["Ypred_top"][:, :, 0], res["Ypred_top_score"][:, :, 0], rr, tol=1e-3,
                gpu_margin=res["tie_margin"][:, :, 0])


def test_cknn_tie_log(ctx):
    """The device tie log: the rank-1 minus rank-2 reranked score equals the
    gap of its own ranked scores; the candidate-set boundary gap matches the
    FP64 base scores to the FP32 similarity error, and is +inf when every
    class is a candidate."""
    import fastpls_amd as F
    rng = np.random.default_rng(8)
    C, k, ntr, nte = 10, 12, 800, 150
    mu = rng.standard_normal((C, k))
    ytr = rng.integers(1, C + 1, ntr).astype(np.int32)
    ytr[:C] = np.arange(1, C + 1)
    Ttr = O.candidate_row_l2(mu[ytr - 1] + 0.8 * rng.standard_normal((ntr, k)))
    Tte = O.candidate_row_l2(mu[rng.integers(0, C, nte)] + 0.8 * rng.standard_normal((nte, k)))
    cen = O.candidate_centroids(Ttr, ytr, C)
    got = F.candidate_knn_predict(Tte, Ttr, ytr, cen, None, 3, 4, 5, 0.2, 0.75)
    mg = got["tie_margin"]
    np.testing.assert_array_equal(mg[:, 0], got["top_score"][:, 0] - got["top_score"][:, 1])
    base = np.sort(Tte @ cen.T, axis=1)[:, ::-1]
    np.testing.assert_allclose(mg[:, 1], base[:, 3] - base[:, 4], rtol=0, atol=1e-5)
    assert np.all(mg[:, 1] >= 0)
    full = F.candidate_knn_predict(Tte, Ttr, ytr, cen, None, 3, C, 5, 0.2, 0.75)
    assert np.all(np.isinf(full["tie_margin"][:, 1]))
