"""Shared comparison helpers for the parity tests."""
import numpy as np


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


def label_mismatches_excused(labels, ref_labels, scores_ref, tol):
    """Rows whose labels differ are excused only if the reference's top-2
    score margin is below `tol` (a documented near-tie)."""
    labels = np.asarray(labels).ravel()
    ref_labels = np.asarray(ref_labels).ravel()
    bad = np.nonzero(labels != ref_labels)[0]
    s = np.sort(scores_ref, axis=1)
    margin = s[:, -1] - s[:, -2]
    unexcused = [int(i) for i in bad if margin[i] > tol]
    return bad, unexcused


def topk_mismatches_excused(ti, ri, rs, tol):
    """Rows whose ranked class lists differ from the reference's, excused only
    where the reference's own ranked scores hold a gap below `tol` (a
    documented near-tie); returns the unexcused rows."""
    bad = []
    for i in range(ri.shape[0]):
        if not np.array_equal(ti[i], ri[i]):
            gaps = np.abs(np.diff(rs[i]))
            if not (gaps.size and np.min(gaps) < tol):
                bad.append(i)
    return bad
