"""CPU checks of the C ABI boundary: the library loads, exports every symbol
include/fastpls_hip.h declares, its struct layouts match the ctypes mirror,
and calls fail loudly (never silently fall back) without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastpls_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fpls_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import fastpls_amd
    lib = fastpls_amd.load()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    from fastpls_amd import _lib
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_struct_sizes():
    from fastpls_amd import _lib
    assert ctypes.sizeof(_lib.Mat) == 40
    assert ctypes.sizeof(_lib.IVec) == 24
    assert ctypes.sizeof(_lib.Opts) == 64
    o = _lib.default_opts()
    assert (o.scaling, o.svd_method, o.oversample, o.power, o.seed, o.inc_iters) == (1, 4, 10, 1, 1, 2)
    assert o.store_b == -1


def test_no_device_fails_loudly():
    import fastpls_amd
    if fastpls_amd.available():
        pytest.skip("a GPU is present")
    with pytest.raises(fastpls_amd.FplsError, match="NO_DEVICE"):
        fastpls_amd.Context(0)
    with pytest.raises(fastpls_amd.FplsError):
        fastpls_amd.pls(np.ones((4, 3)), np.ones((4, 1)), ncomp=1)


def test_backend_other_than_hip_rejected():
    import fastpls_amd
    with pytest.raises(ValueError, match="hip"):
        fastpls_amd.pls(np.ones((4, 3)), np.ones((4, 1)), backend="cpu")
