"""GPU: the projection-tile autotuner.  Every tile height (64, 128, 256 rows
per workgroup) and split-K choice it can pick gives the same latent scores
T = (X - 1 mX^T) R as an FP64 product of the model's own factors (FP32
contractions: rel 2e-6; split-BF16x3: rel 2e-6), and the tuning table records
timed choices per shape (built-in rule when tuning is off)."""
import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _data(seed=3):
    rng = np.random.default_rng(seed)
    C, p = 20, 300
    means = rng.standard_normal((C, p)).astype(np.float32)
    y = rng.integers(1, C + 1, size=6000)
    X = (means[y - 1] + 0.5 * rng.standard_normal((6000, p))).astype(np.float32) + 2.0
    Xt = (means[rng.integers(0, C, size=1500)] + 0.5 * rng.standard_normal((1500, p)))
    return X, y, Xt.astype(np.float32) + 2.0


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_every_projection_tile_agrees(ctx, precision):
    import fastpls_amd as F
    X, y, Xt = _data()
    m = F.pls(X, y.astype(str), method="plssvd", ncomp=19, precision=precision)
    Tref = (Xt.astype(np.float64) - m["mX"]) @ m["R"]
    try:
        for mb in (1, 2, 4):
            for splits in (1, 3):
                ctx.autotune(1, mb, splits)
                T = F.predict(m, Xt, proj=True)["Ttest"]
                assert rel_err(T, Tref) < 2e-6, (mb, splits)
    finally:
        ctx.autotune(1, 0, 0)


def test_autotuner_records_shapes(ctx):
    import fastpls_amd as F
    X, y, Xt = _data(5)
    m = F.pls(X, y.astype(str), method="plssvd", ncomp=19)
    try:
        ctx.autotune(1)
        T1 = F.predict(m, Xt, proj=True)["Ttest"]
        rep = ctx.autotune_report()
        keys = [k for k in rep if k.startswith("proj:f32:centred:M1500:N19")]
        assert keys, rep
        ent = rep[keys[0]]
        assert ent["tuned"] and ent["rows"] in (64, 128, 256) and ent["splits"] >= 1
        assert ent["us"] > 0
        # the cached choice is reused: the same scores bit for bit
        T2 = F.predict(m, Xt, proj=True)["Ttest"]
        assert np.array_equal(T1, T2)
        ctx.autotune(0)
        F.predict(m, Xt, proj=True)
        rep0 = ctx.autotune_report()
        assert all(not v["tuned"] for v in rep0.values())
    finally:
        ctx.autotune(1, 0, 0)
