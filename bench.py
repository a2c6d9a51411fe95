#!/usr/bin/env python3
"""Benchmark: PLS fits/s on one MI355X for the PLSSVD + LDA classification
workload (n = 50,000 train / 10,000 test rows, p = 768, 100 balanced classes,
ncomp = 99, FP32 inputs).

One step = one full ``pls(..., method="plssvd", classifier="lda")`` fit plus
``predict`` of the test rows, with X, the labels and Xtest already resident in
HBM (the fitted model stays on the device; the predicted labels come back to
the host).  Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|bf16x3]
                       [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_TRAIN, N_TEST, P, C, NCOMP = 50_000, 10_000, 768, 100, 99
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
METRIC = "PLS fits/s at fixed FP64-oracle error, chosen vs plain HIP; % of HBM/MFMA peak"


def kernel_work(precision: str):
    """Algorithmic FLOPs of the dominant contractions in one step (useful work
    of the reference algorithm: T = Xc U for the training rows and the test
    rows, 2 n p K each)."""
    return {
        "gemm_nn": 2.0 * N_TRAIN * P * NCOMP + 2.0 * N_TEST * P * NCOMP,
    }


def run_gpu(args):
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth

    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=C, per_class_train=N_TRAIN // C,
                                                 per_class_test=N_TEST // C, p=P)
    dXtr = F.device_colmajor(Xtr)
    dXte = F.device_colmajor(Xte)
    dy = torch.from_numpy(ytr).to("cuda:0")
    torch.cuda.synchronize()
    ctx = F.context(0)

    def step():
        m = F.pls(dXtr, dy, dXte, method="plssvd", ncomp=NCOMP, classifier="lda",
                  precision=args.precision, fetch=False)
        labels = m["Ypred_index"]
        m.free()
        return labels

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        labels = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    # kernel timing pass (HIP events on the library stream)
    ctx.timing(True)
    tsteps = 3
    for _ in range(tsteps):
        step()
    names = ["gemm_nn_f32", "gemm_nn_bf16x3", "gemm_tn_f32", "gemm_tn_bf16x3", "sym_eig_jacobi",
             "spd_solve", "col_sums", "lda_argmax", "gemm_f64", "gemm_tn_reduce", "chol_inv"]
    ktimes = {}
    for nm in names:
        n, ms = ctx.kernel_time(nm)
        if n:
            ktimes[nm] = {"launches_per_step": n / tsteps, "ms_per_step": ms / tsteps}
    ctx.timing(False)
    return labels, dt, ktimes, (Xtr, ytr, Xte, yte)


def cpu_baseline(data, labels_gpu):
    """FP64 oracle (numpy/LAPACK restatement of the reference CPU path) on the
    host cores; also reports label agreement with the GPU run."""
    from threadpoolctl import threadpool_limits
    from oracle import fastpls_oracle as O
    Xtr, ytr, Xte, yte = data
    cores = min(16, os.cpu_count() or 1)
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        ref = O.pls_plssvd_lda(Xtr.astype(np.float64), ytr, Xte.astype(np.float64), [NCOMP])
        dt = time.perf_counter() - t0
    agree = float(np.mean(ref["Ypred_labels"][:, 0] == np.asarray(labels_gpu)[:, 0]))
    acc_ref = float(np.mean(ref["Ypred_labels"][:, 0] == yte))
    acc_gpu = float(np.mean(np.asarray(labels_gpu)[:, 0] == yte))
    return {"value": 1.0 / dt, "unit": "fits/s", "cores": cores, "kind": "port",
            "sample": "one full FP64 fit+predict of the same workload (numpy/LAPACK oracle)",
            "seconds": dt}, agree, acc_ref, acc_gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    labels, dt, ktimes, data = run_gpu(args)
    fits_per_s = args.steps / dt
    work = kernel_work(args.precision)
    kname = "gemm_nn_bf16x3" if args.precision == "bf16x3" else "gemm_nn_f32"
    roof = None
    if kname in ktimes:
        ms = ktimes[kname]["ms_per_step"]
        achieved = work["gemm_nn"] / (ms * 1e-3) / 1e12
        peak = BF16_PEAK_TFLOPS if args.precision == "bf16x3" else FP32_PEAK_TFLOPS
        roof = {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None}
    out = {
        "metric": METRIC, "value": fits_per_s, "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "bf16x3",
        "data": "synthetic (seeded; class means on the radius-3 sphere, N(0,0.5^2) noise)",
        "config": {"workload": "plssvd_lda n=50000 ntest=10000 p=768 C=100 ncomp=99",
                   "precision": args.precision, "inputs": "resident in HBM"},
        "roofline": roof, "kernels": ktimes,
    }
    if not args.no_cpu_baseline:
        cb, agree, acc_ref, acc_gpu = cpu_baseline(data, labels)
        out["cpu_baseline"] = cb
        out["oracle_label_agreement"] = agree
        out["accuracy"] = {"gpu": acc_gpu, "oracle": acc_ref}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
