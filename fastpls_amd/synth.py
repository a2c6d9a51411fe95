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
