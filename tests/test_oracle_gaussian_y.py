"""Pins of the oracle's Gaussian response compression (R/main.R:1945-2142,
src/fastPLS.cpp:4203-4215).

The draw itself (R's set.seed + rnorm in the reference) is parity unpinned:
the oracle and the library both use the seeded MT19937 Gaussian stream.
Everything after the draw is pinned here by what the mathematics fixes: the
ridge decoder against least squares, exact recovery when dim >= ncol(Y), the
codebook's normalisation invariants, and nearest-code decoding against a
brute-force distance argmin.  The reference's own tests
(tests/testthat/test-gaussian-y.R) check flags and shapes; those checks are
mirrored at the pls() level in tests/test_gpu_gaussian_y.py.
"""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import fastpls_oracle as O


def test_default_dim_rule():
    assert O.gaussian_y_default_dim(15) == 15
    assert O.gaussian_y_default_dim(768) == 100
    assert O.gaussian_y_default_dim(15, 7) == 7
    with pytest.raises(ValueError):
        O.gaussian_y_default_dim(15, 0)


def test_ridge_decoder_is_regularised_least_squares():
    rng = np.random.default_rng(1)
    Z = rng.standard_normal((80, 9))
    Yc = rng.standard_normal((80, 13))
    D = O.gaussian_y_ridge_decoder(Z, Yc)
    ls = np.linalg.lstsq(Z, Yc, rcond=None)[0]
    # the 1e-8 relative ridge moves the least-squares solution by O(1e-8)
    np.testing.assert_allclose(D, ls, rtol=0, atol=1e-6 * np.abs(ls).max())
    # and satisfies its own normal equations exactly
    G = Z.T @ Z
    r = 1e-8 * np.mean(np.diag(G))
    np.testing.assert_allclose((G + r * np.eye(9)) @ D, Z.T @ Yc, atol=1e-10)


def test_regression_round_trip_when_dim_covers_the_response():
    rng = np.random.default_rng(2)
    Y = rng.standard_normal((60, 5)) * [1, 2, 3, 4, 5] + 10.0
    enc = O.prepare_gaussian_y_regression(Y, dim=8, seed=3)
    np.testing.assert_allclose(enc["y_center"], Y.mean(axis=0), rtol=1e-14)
    np.testing.assert_allclose(enc["Z"], (Y - Y.mean(axis=0)) @ enc["projection"], atol=1e-12)
    # P (5 x 8) has full row rank: Z determines Yc, so decoding recovers Y
    back = O.decode_gaussian_y_regression(enc["Z"], enc["decoder"], enc["y_center"])
    np.testing.assert_allclose(back, Y, atol=1e-6)
    # projection scale: the stream divided by sqrt(dim)
    np.testing.assert_allclose(enc["projection"] * np.sqrt(8), O.gaussian_matrix(5, 8, 3),
                               rtol=1e-15)


def test_codebook_invariants():
    C, dim = 9, 6
    codes = O.gaussian_y_class_codes(C, dim, seed=5)
    np.testing.assert_allclose(np.linalg.norm(codes, axis=1), 1.0, atol=1e-14)
    # before the row scaling the columns are centred with unit norm: the row
    # scaling multiplies row i by a positive factor, so codes times the row
    # norms of the column-normalised draw gives back centred unit columns
    raw = O.gaussian_matrix(C, dim, 5) / np.sqrt(dim)
    cen = raw - raw.mean(axis=0)
    colnorm = cen / np.linalg.norm(cen, axis=0)
    rows = np.linalg.norm(colnorm, axis=1)
    rebuilt = codes * rows[:, None]
    np.testing.assert_allclose(rebuilt.sum(axis=0), 0.0, atol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(rebuilt, axis=0), 1.0, atol=1e-13)
    # directions: each code points along its centred draw row
    cos = (codes * cen).sum(1) / np.linalg.norm(cen, axis=1)
    assert np.all(cos > 0)


def test_nearest_code_matches_distance_argmin():
    rng = np.random.default_rng(6)
    codes = O.gaussian_y_class_codes(12, 7, seed=8)
    y = rng.integers(1, 13, 200)
    enc = O.prepare_gaussian_y_classes(y, 12, 7, seed=8)
    np.testing.assert_array_equal(enc["Z"], codes[y - 1])
    # exact codes decode to their own class
    np.testing.assert_array_equal(O.nearest_code_classes(enc["Z"], codes), y)
    # noisy codes: first minimum of the Euclidean distance
    Zn = enc["Z"] + 0.3 * rng.standard_normal(enc["Z"].shape)
    ref = np.argmin(cdist(Zn, codes, "sqeuclidean"), axis=1) + 1
    np.testing.assert_array_equal(O.nearest_code_classes(Zn, codes), ref)
    # ties resolve to the first class (index_max)
    twin = np.vstack([codes[0], codes[0]])
    assert O.nearest_code_classes(codes[:1], twin)[0] == 1
