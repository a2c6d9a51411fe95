"""Seeded synthetic inputs for the benchmark workloads (numpy PCG64).

Each generator returns host arrays; callers move them to HBM.  Seeds are
fixed per workload (20260601 + workload index) so that the GPU runs and the
CPU oracle consume identical inputs.
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 20260601


def _orthonormal(rng, rows, cols):
    q, _ = np.linalg.qr(rng.standard_normal((rows, cols)))
    return q


def regression_lowrank(n=200, p=2000, m=3, n_test=100, rank=20, seed=SEED_BASE + 1,
                       dtype=np.float64):
    """SIMPLS regression workload: X = U diag(s) V^T + 0.1 N(0,1), s_i = 10 0.8^i,
    Y = X B + 0.05 N(0,1), B ~ N(0,1)/sqrt(p); test rows share V."""
    rng = np.random.default_rng(seed)
    U = _orthonormal(rng, n + n_test, rank)
    V = _orthonormal(rng, p, rank)
    s = 10.0 * 0.8 ** np.arange(rank)
    X = (U * s) @ V.T + 0.1 * rng.standard_normal((n + n_test, p))
    B = rng.standard_normal((p, m)) / np.sqrt(p)
    Y = X @ B + 0.05 * rng.standard_normal((n + n_test, m))
    X = X.astype(dtype)
    return (np.asfortranarray(X[:n]), np.asfortranarray(Y[:n]), np.asfortranarray(X[n:]),
            np.asfortranarray(Y[n:]))


def classes_on_sphere(n_classes=100, per_class_train=500, per_class_test=100, p=768,
                      radius=3.0, noise=0.5, seed=SEED_BASE + 2, dtype=np.float32):
    """PLS-DA workload: class means uniform on the radius-`radius` sphere,
    rows = mean + N(0, noise^2 I), balanced labels 1..C (shuffled)."""
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal((n_classes, p))
    mu *= radius / np.linalg.norm(mu, axis=1, keepdims=True)

    def draw(per):
        y = np.repeat(np.arange(1, n_classes + 1, dtype=np.int32), per)
        rng.shuffle(y)
        X = mu[y - 1] + noise * rng.standard_normal((y.size, p))
        return np.asfortranarray(X.astype(dtype)), y

    Xtr, ytr = draw(per_class_train)
    Xte, yte = draw(per_class_test)
    return Xtr, ytr, Xte, yte


def _dirichlet_folds(rng, rows, k, alpha=5.0):
    """Assign `rows` (index array) to k folds with Dirichlet(alpha) sizes,
    rounded, at least one row each; returns fold ids 1..k per row."""
    sizes = np.maximum(1, np.round(rng.dirichlet(np.full(k, alpha)) * rows.size).astype(int))
    while sizes.sum() > rows.size:
        sizes[np.argmax(sizes)] -= 1
    sizes[-1] += rows.size - sizes.sum()
    ids = np.repeat(np.arange(1, k + 1), sizes)
    rng.shuffle(ids)
    return ids


def nested_cv_lognormal(n=5000, p=20000, block=200, rho=0.6, n_outer=10, n_inner=5,
                        props=(0.4, 0.25, 0.15, 0.12, 0.08), seed=SEED_BASE + 4):
    """Double cross-validation workload: X = exp(Z), Z ~ N(0,1) with blocks of
    `block` columns sharing a factor (within-block correlation rho), FP32
    column-major; labels 1..5 with the given proportions; outer folds of
    Dirichlet(5) sizes and, for each outer fold j, inner folds (Dirichlet(5)
    sizes) of the rows outside fold j.  Returns X, y, outer (n), inner (n x
    n_outer; entries of rows inside outer fold j are 0)."""
    rng = np.random.default_rng(seed)
    nb = p // block
    Z = rng.standard_normal((n, p), dtype=np.float32)
    Z *= np.float32(np.sqrt(1.0 - rho))
    shared = rng.standard_normal((n, nb), dtype=np.float32) * np.float32(np.sqrt(rho))
    Z += np.repeat(shared, block, axis=1)
    np.exp(Z, out=Z)
    X = np.asfortranarray(Z)
    del Z
    y = rng.choice(np.arange(1, len(props) + 1), n, p=np.asarray(props)).astype(np.int32)
    outer = _dirichlet_folds(rng, np.arange(n), n_outer)
    inner = np.zeros((n, n_outer), dtype=np.int64)
    for j in range(n_outer):
        tr = np.nonzero(outer != j + 1)[0]
        inner[tr, j] = _dirichlet_folds(rng, tr, n_inner)
    return X, y, outer, inner


def zipf_classes_large(n=1_200_000, n_test=50_000, p=1024, n_classes=1000, zipf=1.1,
                       radius=4.0, noise=0.6, seed=SEED_BASE + 3, chunk=100_000):
    """Large-class-count PLS-DA workload: class sizes proportional to
    Zipf(zipf) (every class present in the training set), class means
    uniform on the radius-`radius` sphere in R^p, rows = mean + N(0, noise^2 I),
    FP32 column-major.  Test rows drawn the same way."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, n_classes + 1) ** zipf
    w /= w.sum()
    mu = rng.standard_normal((n_classes, p))
    mu *= radius / np.linalg.norm(mu, axis=1, keepdims=True)
    mu = mu.astype(np.float32)

    def draw(rows, cover):
        y = rng.choice(np.arange(1, n_classes + 1), rows, p=w).astype(np.int32)
        if cover:
            y[:n_classes] = np.arange(1, n_classes + 1)
            rng.shuffle(y)
        X = np.empty((rows, p), dtype=np.float32, order="F")
        for a in range(0, rows, chunk):
            b = min(rows, a + chunk)
            blk = rng.standard_normal((b - a, p), dtype=np.float32)
            blk *= np.float32(noise)
            blk += mu[y[a:b] - 1]
            X[a:b] = blk
        return X, y

    Xtr, ytr = draw(n, True)
    Xte, yte = draw(n_test, False)
    return Xtr, ytr, Xte, yte
