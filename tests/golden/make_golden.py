"""Writes the golden fixtures of tests/golden/ (run: python tests/golden/make_golden.py).

The expected outputs here are computed by textbook formulations written
independently of the oracle, so that the oracle and the device are checked
against something other than themselves:

* simpls_textbook.npz -- SIMPLS (de Jong 1993) in projection form: the weight of
  component a is the exact top left singular vector of (I - P_a) S0, with P_a
  the orthogonal projector onto the span of the first a loadings X0^T t (QR of
  the loading matrix), t = X0 r / |X0 r|; B_a = R_a Q_a^T.  The reference's
  SIMPLS-fast (src/fastPLS.cpp:2454-2778) converges to it as its warm-started
  power iterations grow (FASTPLS_FAST_INC_ITERS).
* plssvd_textbook.npz -- PLSSVD with the exact SVD: the fitted values are the
  least-squares projection of Yc onto the span of the scores T_k = Xc U_k
  (src/fastPLS.cpp:4780-4986 computes the same through C = G^-1 diag(s)).

r_runif.json is not made here: it holds the values R prints for
set.seed(s); runif(5), the generator the reference's CV permutations draw from.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def simpls_textbook(X, Y, K):
    X0 = X - X.mean(0)
    Y0 = Y - Y.mean(0)
    S0 = X0.T @ Y0
    p, m = S0.shape
    R = np.zeros((p, K))
    Q = np.zeros((m, K))
    P = np.zeros((p, K))
    Bs = []
    for a in range(K):
        if a == 0:
            Sa = S0
        else:
            Qp, _ = np.linalg.qr(P[:, :a])
            Sa = S0 - Qp @ (Qp.T @ S0)
        r = np.linalg.svd(Sa)[0][:, 0]
        t = X0 @ r
        nt = np.linalg.norm(t)
        r, t = r / nt, t / nt
        R[:, a], P[:, a], Q[:, a] = r, X0.T @ t, Y0.T @ t
        Bs.append(R[:, :a + 1] @ Q[:, :a + 1].T)
    return np.stack(Bs, axis=2)


def plssvd_textbook(X, Y, ks):
    Xc = X - X.mean(0)
    Yc = Y - Y.mean(0)
    U = np.linalg.svd(Xc.T @ Yc, full_matrices=False)[0]
    B, Yfit = [], []
    for k in ks:
        T = Xc @ U[:, :k]
        coef = np.linalg.solve(T.T @ T, T.T @ Yc)
        B.append(U[:, :k] @ coef)
        Yfit.append(T @ coef + Y.mean(0))
    return np.stack(B, axis=2), np.stack(Yfit, axis=2)


def main():
    rng = np.random.default_rng(20260601)
    n, p, m, K = 60, 14, 3, 5
    # well separated latent directions so that the power iterations converge
    Z = rng.standard_normal((n, 4)) * np.array([6.0, 3.5, 2.0, 1.2])
    X = Z @ rng.standard_normal((4, p)) + 0.3 * rng.standard_normal((n, p)) + 2.0
    Y = Z[:, :3] @ np.array([[1.5, -0.3, 0.4], [0.2, 1.1, -0.6], [-0.4, 0.5, 0.9]]) \
        + 0.2 * rng.standard_normal((n, m))
    # scaled so that S0 = X0^T Y0 has unit spectral norm: the reference's power
    # steps do not normalise, and many of them must neither overflow nor
    # underflow (B is unchanged by a common scale of X and Y)
    c = np.sqrt(np.linalg.norm((X - X.mean(0)).T @ (Y - Y.mean(0)), 2))
    X, Y = X / c, Y / c
    np.savez(os.path.join(HERE, "simpls_textbook.npz"), X=X, Y=Y, K=K,
             B=simpls_textbook(X, Y, K))
    X2 = rng.standard_normal((50, 9)) + 1.0
    Y2 = X2[:, :4] @ rng.standard_normal((4, 5)) + 0.1 * rng.standard_normal((50, 5))
    ks = np.array([1, 3, 5])
    B2, F2 = plssvd_textbook(X2, Y2, ks)
    np.savez(os.path.join(HERE, "plssvd_textbook.npz"), X=X2, Y=Y2, ncomp=ks, B=B2, Yfit=F2)
    print("wrote simpls_textbook.npz, plssvd_textbook.npz")


if __name__ == "__main__":
    main()
