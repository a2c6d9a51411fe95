"""Shared comparison helpers for the parity tests.

A label (or ranked class list) that differs from the FP64 oracle's is excused
only as a near-tie: the oracle's own margin and the device's margin (its tie
log, when the entry point emits one) are both below the error bound of the
precision mode.  Every excused row is recorded in TIE_LOG, which the test
session prints at its end (and appends to the file named by FPLS_TIE_LOG)."""
import os

import numpy as np

TIE_LOG: list = []


def _log_ties(rows, ref_margin, gpu_margin, tol):
    test = os.environ.get("PYTEST_CURRENT_TEST", "?").split(" ")[0]
    for i in rows:
        TIE_LOG.append({"test": test, "row": int(i), "oracle_margin": float(ref_margin[i]),
                        "gpu_margin": None if gpu_margin is None else float(gpu_margin[i]),
                        "bound": float(tol)})


def align_signs(A, Ref):
    """Flip columns of A to the sign of the matching Ref column (the
    reference test idiom align_signs, tests/testthat/test-pls-rsvd.R:11-20)."""
    A = np.array(A, dtype=np.float64, copy=True)
    for j in range(min(A.shape[1], Ref.shape[1])):
        if float(A[:, j] @ Ref[:, j]) < 0:
            A[:, j] = -A[:, j]
    return A


def sin_theta(U, Uref):
    """Largest principal-angle sine between the column spans."""
    Q1, _ = np.linalg.qr(U)
    Q2, _ = np.linalg.qr(Uref)
    M = Q1 - Q2 @ (Q2.T @ Q1)
    return float(np.linalg.norm(M, 2))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def label_mismatches_excused(labels, ref_labels, scores_ref, tol, gpu_margin=None):
    """Rows whose labels differ are excused only if the reference's top-2
    score margin is below `tol` and, when `gpu_margin` (the device's tie log)
    is given, the device's margin too; excused rows are logged."""
    labels = np.asarray(labels).ravel()
    ref_labels = np.asarray(ref_labels).ravel()
    bad = np.nonzero(labels != ref_labels)[0]
    s = np.sort(scores_ref, axis=1)
    margin = s[:, -1] - s[:, -2]
    gm = None if gpu_margin is None else np.asarray(gpu_margin).ravel()
    ok = [int(i) for i in bad if margin[i] <= tol and (gm is None or gm[i] <= tol)]
    unexcused = [int(i) for i in bad if int(i) not in set(ok)]
    _log_ties(ok, margin, gm, tol)
    return bad, unexcused


def topk_mismatches_excused(ti, ri, rs, tol, ts=None, gpu_margin=None):
    """Rows whose ranked class lists differ from the reference's, excused only
    where the reference's ranked scores hold a gap below `tol` and the
    device's do too: its own ranked scores `ts`, the rank-1 / rank-2 gap and
    the candidate-set boundary gap of its tie log `gpu_margin` (rows x 2).
    Returns the unexcused rows; excused rows are logged."""
    bad, ok, ref_gap, dev_gap = [], [], np.full(ri.shape[0], np.inf), None
    if ts is not None or gpu_margin is not None:
        dev_gap = np.full(ri.shape[0], np.inf)
    for i in range(ri.shape[0]):
        if np.array_equal(ti[i], ri[i]):
            continue
        gaps = np.abs(np.diff(rs[i]))
        ref_gap[i] = np.min(gaps) if gaps.size else np.inf
        if dev_gap is not None:
            cand = []
            if ts is not None and ts.shape[1] > 1:
                cand.append(np.min(np.abs(np.diff(ts[i]))))
            if gpu_margin is not None:
                cand.extend([gpu_margin[i, 0], gpu_margin[i, 1]])
            dev_gap[i] = min(cand) if cand else np.inf
        if ref_gap[i] < tol and (dev_gap is None or dev_gap[i] < tol):
            ok.append(i)
        else:
            bad.append(i)
    _log_ties(ok, ref_gap, dev_gap, tol)
    return bad
