"""GPU: the allocation-free LDA prediction (fpls_lda_predict_noalloc) against
fpls_lda_predict, through the C ABI.

Both entry points run the same kernels on the same model, so labels must be
identical and the discriminant scores equal to the last bit.  Two calls in a
row, the second with fewer rows, check that the shared scratch array keeps no
state between calls; a row count beyond the scratch limit is rejected on the
host before any work starts (the input tensor is really that large).
"""
import ctypes

import numpy as np
import pytest

from fastpls_amd import synth

pytestmark = pytest.mark.gpu


def _predict_ref(model, X):
    ctx = model.ctx
    from fastpls_amd._lib import as_mat
    keep = []
    Xm = as_mat(X, keep)
    w, L, C = X.shape[0], model.info().n_ncomp, model.info().n_classes
    lab = np.empty((w, L), dtype=np.int32, order="F")
    sc = np.empty((w, C, L), dtype=np.float64, order="F")
    ctx.check(ctx.lib.fpls_lda_predict(ctx.h, model.handle, ctypes.byref(Xm), lab.ctypes.data,
                                       sc.ctypes.data))
    return lab, sc


def _predict_noalloc(model, Xdev, with_scores=True):
    import torch
    ctx = model.ctx
    from fastpls_amd._lib import as_mat
    keep = []
    Xm = as_mat(Xdev, keep)
    w, L, C = Xdev.shape[0], model.info().n_ncomp, model.info().n_classes
    lab = torch.full((L, w), -7, dtype=torch.int32, device=Xdev.device)
    sc = torch.full((L, C, w), float("nan"), dtype=torch.float64, device=Xdev.device)
    rc = ctx.lib.fpls_lda_predict_noalloc(ctx.h, model.handle, ctypes.byref(Xm), lab.data_ptr(),
                                          sc.data_ptr() if with_scores else None)
    ctx.check(rc)
    # device buffers are column-major w x L and w x C x L
    return (lab.cpu().numpy().T.copy(order="F"),
            np.transpose(sc.cpu().numpy(), (2, 1, 0)).copy(order="F"))


@pytest.mark.parametrize("dtype,precision", [(np.float32, "bf16x3"), (np.float32, "fp32"),
                                             (np.float64, "fp32")])
def test_noalloc_matches_lda_predict(ctx, dtype, precision):
    import fastpls_amd as F
    from fastpls_amd._lib import device_colmajor
    Xtr, ytr, Xte, _ = synth.classes_on_sphere(n_classes=12, per_class_train=200,
                                               per_class_test=120, p=96, seed=5)
    Xtr = Xtr.astype(dtype)
    Xte = Xte.astype(dtype)
    model = F.pls(Xtr, ytr, method="plssvd", ncomp=[3, 7, 11], classifier="lda",
                  precision=precision)
    lab0, sc0 = _predict_ref(model, Xte)
    Xd = device_colmajor(Xte)
    lab1, sc1 = _predict_noalloc(model, Xd)
    np.testing.assert_array_equal(lab1, lab0)
    np.testing.assert_array_equal(sc1, sc0)
    # second call, smaller input: nothing of the first call may leak through
    w2 = Xte.shape[0] // 3
    lab0b, sc0b = _predict_ref(model, Xte[:w2])
    lab2, sc2 = _predict_noalloc(model, device_colmajor(Xte[:w2]))
    np.testing.assert_array_equal(lab2, lab0b)
    np.testing.assert_array_equal(sc2, sc0b)
    # labels only
    lab3, _ = _predict_noalloc(model, Xd, with_scores=False)
    np.testing.assert_array_equal(lab3, lab0)
    model.free()


def test_noalloc_rejects_beyond_scratch(ctx):
    import torch
    import fastpls_amd as F
    from fastpls_amd._lib import FplsError, as_mat
    Xtr, ytr, _, _ = synth.classes_on_sphere(n_classes=12, per_class_train=100,
                                             per_class_test=1, p=64, seed=6)
    model = F.pls(Xtr, ytr, method="plssvd", ncomp=11, classifier="lda")
    C, k = 12, 11
    # per row at least 4 k + 12 C bytes of scratch: this many rows cannot fit 64 MiB
    w = (64 << 20) // (4 * k + 12 * C) + 1000
    Xbig = torch.zeros((64, w), dtype=torch.float32, device="cuda:0").t()
    lab = torch.zeros((w,), dtype=torch.int32, device="cuda:0")
    keep = []
    Xm = as_mat(Xbig, keep)
    rc = model.ctx.lib.fpls_lda_predict_noalloc(model.ctx.h, model.handle, ctypes.byref(Xm),
                                                lab.data_ptr(), None)
    assert rc == -1
    msg = model.ctx.lib.fpls_last_error(model.ctx.h).decode()
    assert "scratch limit" in msg
    # a host matrix is refused too
    Xh = np.zeros((10, 64), dtype=np.float32)
    Xm = as_mat(Xh, keep)
    rc = model.ctx.lib.fpls_lda_predict_noalloc(model.ctx.h, model.handle, ctypes.byref(Xm),
                                                lab.data_ptr(), None)
    assert rc == -1
    # the context still works afterwards
    with pytest.raises(FplsError):
        model.ctx.check(rc)
    model.free()
