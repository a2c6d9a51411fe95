"""Pins of the oracle's argmax-head class ranking (src/fastPLS.cpp:3737-4113,
R/main.R:377-492).

fill_topk_from_yblock is the reference's insertion loop restated; it is
pinned here against what fixes its result independently: numpy's stable
argsort (a library routine: descending value, ties in class order), argmax
(first maximum) for one rank, and brute force over the special values the
loop treats specially (NaN and -inf never ranked, empty slots hold class 0
and -inf).  The compact prediction path is pinned against the dense
prediction cube (pls_predict) it factorises, and the bias rules against the
error messages and broadcasts of class_bias_offsets / .class_bias_matrix.
The reference's own test (tests/testthat/test-class-bias-top5.R) checks that
top-k prediction keeps the argmax labels and the top-5 shapes; those checks
are mirrored in tests/test_gpu_class_topk.py.
"""
import numpy as np
import pytest

from oracle import fastpls_oracle as O


def _stable_topk(S, k):
    order = np.argsort(-S, axis=1, kind="stable")[:, :k]
    return order + 1, np.take_along_axis(S, order, axis=1)


@pytest.mark.parametrize("k", [1, 2, 5, 9])
def test_insertion_equals_stable_sort_with_ties(k):
    rng = np.random.default_rng(k)
    Y = rng.integers(-3, 4, size=(60, 9)).astype(np.float64)   # many ties
    off = rng.integers(-1, 2, size=9).astype(np.float64)
    idx, sc = O.fill_topk_from_yblock(Y, off, k)
    ref_i, ref_s = _stable_topk(Y + off, k)
    np.testing.assert_array_equal(idx, ref_i)
    np.testing.assert_array_equal(sc, ref_s)


def test_one_rank_is_the_first_maximum():
    rng = np.random.default_rng(3)
    Y = np.round(rng.standard_normal((200, 7)), 1)
    idx, sc = O.fill_topk_from_yblock(Y, np.zeros(7), 1)
    np.testing.assert_array_equal(idx[:, 0], np.argmax(Y, axis=1) + 1)
    np.testing.assert_array_equal(sc[:, 0], Y.max(axis=1))


def test_nan_and_minus_inf_are_never_ranked():
    Y = np.array([[np.nan, 2.0, -np.inf, 1.0],
                  [np.nan, np.nan, -np.inf, -np.inf],
                  [np.inf, np.inf, 0.0, np.nan]])
    idx, sc = O.fill_topk_from_yblock(Y, np.zeros(4), 4)
    np.testing.assert_array_equal(idx, [[2, 4, 0, 0], [0, 0, 0, 0], [1, 2, 3, 0]])
    np.testing.assert_array_equal(sc[0], [2.0, 1.0, -np.inf, -np.inf])
    assert np.all(sc[1] == -np.inf)
    np.testing.assert_array_equal(sc[2], [np.inf, np.inf, 0.0, -np.inf])
    # top_k is clamped to [1, m]
    assert O.fill_topk_from_yblock(Y, np.zeros(4), 0)[0].shape == (3, 1)
    assert O.fill_topk_from_yblock(Y, np.zeros(4), 99)[0].shape == (3, 4)


def test_bias_offsets_rules():
    mY = np.array([0.5, -1.0, 2.0])
    np.testing.assert_array_equal(O.class_bias_offsets(mY, None, 2), np.c_[mY, mY])
    b = np.array([1.0, 2.0, 3.0])
    np.testing.assert_array_equal(O.class_bias_offsets(mY, b, 2), np.c_[mY + b, mY + b])
    B = np.arange(6.0).reshape(3, 2)
    np.testing.assert_array_equal(O.class_bias_offsets(mY, B, 2), mY[:, None] + B)
    with pytest.raises(ValueError, match="one row per class"):
        O.class_bias_offsets(mY, np.ones(4), 2)
    with pytest.raises(ValueError, match="one column or one column per ncomp"):
        O.class_bias_offsets(mY, np.ones((3, 3)), 2)


def test_r_score_cube_ranking_equals_the_insertion():
    """.class_topk_from_score_cube (max.col first / stable order) and
    class_bias_topk_from_cube agree on finite scores."""
    rng = np.random.default_rng(5)
    cube = rng.integers(-2, 3, size=(40, 6, 3)).astype(np.float64)
    bias = rng.standard_normal((6, 3))
    for top in (1, 4):
        a_i, a_s = O.class_topk_from_score_cube(cube, bias, top)
        b_i, b_s = O.class_bias_topk_from_cube(cube, bias, top)
        np.testing.assert_array_equal(a_i, b_i)
        np.testing.assert_array_equal(a_s, b_s)


@pytest.mark.parametrize("method", ["plssvd", "simpls"])
def test_compact_path_ranks_the_prediction_cube(method):
    rng = np.random.default_rng(7)
    X = rng.standard_normal((80, 12))
    y = rng.integers(1, 6, 80)
    Xt = rng.standard_normal((25, 12))
    Y = O.transformy(y)
    if method == "plssvd":
        mod = O.pls_model1(X, Y, [1, 2, 3], 1, False, "rsvd", 10, 1, 1)
    else:
        mod = O.pls_model2_fast(X, Y, [1, 2, 3], 1, False, "rsvd", 10, 1, 1)
    cube = O.pls_predict(mod, Xt)
    ti, ts = O.pls_class_predict_topk(mod, Xt, None, 3, block_size=7)
    for a in range(3):
        np.testing.assert_array_equal(ti[:, 0, a], np.argmax(cube[:, :, a], axis=1) + 1)
        np.testing.assert_allclose(ts[:, :, a], np.take_along_axis(cube[:, :, a], ti[:, :, a] - 1,
                                                                   axis=1), atol=1e-12)
    # a large bias on one class makes it rank first everywhere
    bias = np.zeros(5)
    bias[3] = 1e6
    tb, _ = O.pls_class_predict_topk(mod, Xt, bias, 2)
    assert np.all(tb[:, 0, :] == 4)
    # a bias shared by every class shifts the scores and keeps the ranks
    ts_b = O.pls_class_predict_topk(mod, Xt, np.full(5, 2.5), 3)[1]
    np.testing.assert_allclose(ts_b, ts + 2.5, atol=1e-12)
