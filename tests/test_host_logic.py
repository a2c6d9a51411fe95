"""Host-side argument handling of the R-API mirror (no GPU): the class-bias
matrix rules of .class_bias_matrix (R/main.R:377-415), .resolve_top_k
(R/main.R:368-375) and the error paths of the OPLS / kernel PLS wrappers that
are decided before any device call."""
import numpy as np
import pytest

import fastpls_amd as F


def test_class_bias_matrix_rules():
    lev = ["a", "b", "c"]
    np.testing.assert_array_equal(F.class_bias_matrix(None, lev, 2), np.zeros((3, 2)))
    np.testing.assert_array_equal(F.class_bias_matrix({"c": 3, "a": 1, "b": 2}, lev, 2),
                                  [[1, 1], [2, 2], [3, 3]])
    np.testing.assert_array_equal(F.class_bias_matrix([1.0, 2.0, 3.0], lev, 1), [[1], [2], [3]])
    # a slices x classes matrix is transposed; one column broadcasts
    np.testing.assert_array_equal(F.class_bias_matrix(np.ones((2, 3)), lev, 2), np.ones((3, 2)))
    np.testing.assert_array_equal(F.class_bias_matrix(np.arange(3.0)[:, None], lev, 4),
                                  np.repeat(np.arange(3.0)[:, None], 4, axis=1))
    with pytest.raises(ValueError, match="one numeric value per class"):
        F.class_bias_matrix([1.0, 2.0], lev, 2)
    with pytest.raises(ValueError, match="one numeric value per class"):
        F.class_bias_matrix({"a": 1.0, "b": 2.0}, lev, 2)
    with pytest.raises(ValueError, match="one row per class"):
        F.class_bias_matrix(np.ones((4, 2)), lev, 2)
    with pytest.raises(ValueError, match="one column or one column per ncomp"):
        F.class_bias_matrix(np.ones((3, 3)), lev, 2)


def test_resolve_top_k():
    assert F.resolve_top_k(1) == 1
    assert F.resolve_top_k(3, top5=True) == 5
    assert F.resolve_top_k(7, top5=True) == 7
    for bad in (0, -2):
        with pytest.raises(ValueError, match="top must be a positive integer"):
            F.resolve_top_k(bad)


def test_model_import_field_checks():
    """model_from_fields rejects an unknown method or missing fields before
    any device call (the reference's list must carry R, ncomp, mX, vX, mY)."""
    f = {"R": np.eye(3), "ncomp": [1], "mX": np.zeros(3), "vX": np.ones(3), "mY": np.zeros(2),
         "pls_method": "pcr"}
    with pytest.raises(ValueError, match="pls_method"):
        F.model_from_fields(f)
    for missing in ("R", "mY"):
        g = dict(f, pls_method="simpls")
        g.pop(missing)
        with pytest.raises(ValueError, match=missing):
            F.model_from_fields(g)
    # centring vectors must have one value per predictor (host over-read otherwise)
    g = dict(f, pls_method="simpls", Q=np.zeros((2, 3)), mX=np.zeros(2))
    with pytest.raises(ValueError, match="one value per predictor"):
        F.model_from_fields(g)
    g = dict(f, pls_method="simpls", Q=np.zeros((2, 3)), vX=np.ones(5))
    with pytest.raises(ValueError, match="one value per predictor"):
        F.model_from_fields(g)


def test_cv_kfold_argument_rules():
    """.compiled_cv_kfold_arg / .cv_kfold_int (R/main.R:3982-4009)."""
    cons = np.repeat(np.arange(6), 2)
    assert F.compiled_kfold("loocv", cons) == -1
    assert F.compiled_kfold(" LOO ", cons) == -1
    assert F.compiled_kfold(6, cons) == -1          # kfold >= number of groups
    assert F.compiled_kfold(12, cons) == -1
    assert F.compiled_kfold(3, cons) == 3
    assert F.compiled_kfold(1, cons) == 2           # at least two folds
    assert F.compiled_kfold(-1, cons) == -1
    with pytest.raises(ValueError, match="single integer or 'loocv'"):
        F.compiled_kfold("ten", cons)
    with pytest.raises(ValueError, match="finite integer"):
        F.compiled_kfold(float("nan"), cons)


def test_cv_rejects_missing_values():
    X = np.ones((6, 3))
    X[2, 1] = np.nan
    with pytest.raises(ValueError, match="Missing values are present"):
        F.optim_pls_cv(X, np.arange(6.0), ncomp=1)
    with pytest.raises(ValueError, match="Missing values are present"):
        F.pls_double_cv(X, np.array([1, 2, 1, 2, 1, 2]), ncomp=1,
                        outer_fold=np.array([1, 1, 2, 2, 3, 3]),
                        inner_fold=np.ones((6, 3), dtype=int))


def test_should_use_xprod_default_rule():
    """.should_use_xprod_default (R/main.R:1883-1892): S above 32 MiB (strict),
    or q >= 100 with at most 10 components; non-finite sizes give FALSE."""
    from fastpls_amd.api import should_use_xprod_default as rule
    assert rule(4096, 1024, 50) is False          # exactly 32 MiB
    assert rule(4097, 1024, 50) is True
    assert rule(50000, 100, 10) is True           # the rSVD-of-X^T Y workload
    assert rule(768, 100, 99) is False            # the classification workload
    assert rule(300, 100, 10) is True and rule(300, 99, 10) is False
    assert rule(300, 100, 11) is False
    assert rule(300, 100, [3, 12]) is False       # max(ncomp) decides
    assert rule(float("nan"), 100, 2) is False
