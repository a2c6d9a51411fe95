#!/usr/bin/env python3
"""Benchmark: PLS fits/s on one MI355X for the PLSSVD + LDA classification
workload (n = 50,000 train / 10,000 test rows, p = 768, 100 balanced classes,
ncomp = 99, FP32 inputs).

One step = one full ``pls(..., method="plssvd", classifier="lda")`` fit plus
``predict`` of the test rows, with X, the labels and Xtest already resident in
HBM (the fitted model stays on the device; predicted labels return to the
host).  Two contraction forms are timed: plain FP32 MFMA and split-BF16x3
MFMA.  Each is checked against the FP64 oracle (latent subspace sin(theta),
coefficient error, label agreement); ``value`` is the split-BF16x3 throughput
when it meets sin(theta) <= 1e-4 with labels equal up to near-ties, else the
plain FP32 throughput.  Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
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
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0, "fp64": 78.6}
METRIC = "PLS fits/s at fixed FP64-oracle error, chosen vs plain HIP; % of HBM/MFMA peak"
SIN_THETA_BOUND = 1e-4

# Algorithmic work of the projection kernel per step (useful FLOPs of the
# reference algorithm): T = Xc U for the training rows and Ttest = Xtest_c U,
# 2 n p K each (two launches).  Bytes: the FP32 X read once and the FP32 scores
# T written once, 4 n (p + K) per launch.
PROJ_FLOPS = 2.0 * N_TRAIN * P * NCOMP + 2.0 * N_TEST * P * NCOMP
PROJ_BYTES = 4.0 * (N_TRAIN + N_TEST) * (P + NCOMP)
# Split-BF16x3 issues three BF16 MFMA products per useful product, so its peak
# for useful FLOPs is a third of the BF16 MFMA peak.
USEFUL_PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0 / 3.0}


def measured_traffic(kname, launches):
    """HBM bytes per launch of ``kname`` from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (profiles/hbm_traffic.json, written from those counter
    runs and scaled by a plain copy of known size), or None when absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return float(t["kernels"][kname]["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def calibration_copy(mib):
    """One plain device-to-device copy of ``mib`` MiB: the reference dispatch
    that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def pmc_summary(fetch_csv, write_csv, copy_mib, keys, workload):
    """Per-dispatch HBM bytes from two rocprofv3 counter passes (FETCH_SIZE,
    then WRITE_SIZE; ``*_counter_collection.csv``), each scaled by what the
    same counter reads for the calibration copy of ``copy_mib`` MiB.  Prints a
    markdown table and records ``keys`` ({name: kernel-name prefix}) in
    profiles/hbm_traffic.json."""
    import csv
    import collections

    def load(path, counter):
        per = collections.defaultdict(list)
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name", counter) != counter:
                    continue
                per[row["Kernel_Name"]].append(float(row["Counter_Value"]))
        return per

    fe, wr = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    kib = copy_mib * 1024.0
    cf = max((v for k, vs in fe.items() if "copyBuffer" in k for v in vs), default=0.0)
    cw = max((v for k, vs in wr.items() if "copyBuffer" in k for v in vs), default=0.0)
    if cf <= 0 or cw <= 0:
        raise SystemExit("no calibration copy in the counter passes (run with --copy-mib)")
    fs, ws = kib / cf, kib / cw
    print(f"Calibration: the {copy_mib} MiB copy reads FETCH_SIZE = {cf:.0f} KiB and "
          f"WRITE_SIZE = {cw:.0f} KiB, so FETCH_SIZE is scaled by {fs:.3f} and WRITE_SIZE "
          f"by {ws:.3f}.\n")
    print("| kernel | dispatches | FETCH KiB (raw avg) | WRITE KiB (raw avg) | scaled MB read "
          "| scaled MB written |")
    print("|---|---|---|---|---|---|")
    rows = []
    for k in sorted(set(fe) | set(wr)):
        f_ = fe.get(k, [0.0])
        w_ = wr.get(k, [0.0])
        af, aw = sum(f_) / len(f_), sum(w_) / len(w_)
        rows.append((k, len(f_), af, aw))
    rows.sort(key=lambda r: -(r[2] * fs + r[3] * ws) * r[1])
    for k, cnt, af, aw in rows[:40]:
        print(f"| `{k[:90]}` | {cnt} | {af:.0f} | {aw:.0f} | {af * fs * 1.024e-3:.2f} | "
              f"{aw * ws * 1.024e-3:.2f} |")
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {"kernels": {}}
    for key, prefix in keys.items():
        hit = [r for r in rows if r[0].startswith(prefix)]
        if not hit:
            print(f"\nno dispatch named {prefix!r}")
            continue
        cnt = sum(r[1] for r in hit)
        rd = sum(r[2] * r[1] for r in hit) / cnt * fs * 1024.0
        wb = sum(r[3] * r[1] for r in hit) / cnt * ws * 1024.0
        table["kernels"][key] = {"symbol": prefix, "workload": workload,
                                 "fetch_scale": fs, "write_scale": ws,
                                 "hbm_bytes_per_launch": rd + wb,
                                 "read_bytes_per_launch": rd, "write_bytes_per_launch": wb}
        print(f"\n{key}: {cnt} dispatches, {(rd + wb) / 1e6:.2f} MB per launch "
              f"({rd / 1e6:.2f} read, {wb / 1e6:.2f} written)")
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def sin_theta(U, Uref):
    Q1, _ = np.linalg.qr(U)
    Q2, _ = np.linalg.qr(Uref)
    return float(np.linalg.norm(Q1 - Q2 @ (Q2.T @ Q1), 2))


def run_gpu(args, data):
    import torch
    import fastpls_amd as F

    Xtr, ytr, Xte, yte = data
    dXtr = F.device_colmajor(Xtr)
    dXte = F.device_colmajor(Xte)
    dy = torch.from_numpy(ytr).to("cuda:0")
    torch.cuda.synchronize()
    ctx = F.context(0)
    out = {}
    for prec in ("fp32", "bf16x3"):
        def step(fetch=False):
            m = F.pls(dXtr, dy, dXte, method="plssvd", ncomp=NCOMP, classifier="lda",
                      precision=prec, fetch=fetch, n_classes=C)
            res = (m["Ypred_index"], m.get("R"), m.get("B"))
            m.free()
            return res

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        labels, R, B = step(fetch=True)
        # kernel timing pass: HIP events on the library stream
        ctx.timing(True)
        tsteps = 3
        for _ in range(tsteps):
            step()
        kern = {}
        for nm in ("project_f32", "project_bf16x3", "gemm_nn_f32", "gemm_nn_bf16x3",
                   "gemm_tn_f32", "gemm_tn_bf16x3",
                   "sym_eig_dc", "sym_eig_trd", "sym_eig_pairs", "sym_eig_clorth",
                   "spd_solve", "col_sums", "lda_argmax", "gemm_f64",
                   "gemm_tn_reduce", "label_counts"):
            n, ms = ctx.kernel_time(nm)
            if n:
                kern[nm] = {"launches_per_step": n / tsteps, "ms_per_step": ms / tsteps}
        ctx.timing(False)
        out[prec] = {"fits_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                     "labels": labels, "R": R, "B": B, "kernels": kern,
                     "autotune": {k: v for k, v in ctx.autotune_report().items()
                                  if ("bf16x3" in k) == (prec == "bf16x3")}}
    return out


# ---------------------------------------------------------------- CPU baseline
def host_cpu_info():
    """The host cores this process may use (its affinity set, bounded by
    OMP_NUM_THREADS when that is set) and the CPU model string."""
    try:
        share = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        share = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(share, int(omp)) if omp.isdigit() and int(omp) > 0 else share
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    lscpu = ""
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.strip().startswith("Model name"):
                lscpu = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    return {"nproc": share, "threads_all": threads, "cpu_model": model,
            "lscpu_model": lscpu, "os_cpu_count": os.cpu_count()}


def time_cpu(fn, repeats=5, warmup=1, one_thread=True):
    """Median wall time of ``repeats`` runs of ``fn`` after ``warmup`` runs,
    at all the process's host cores and (``one_thread``) at 1 thread, with the
    BLAS thread pool limited accordingly.  Returns (timings, last result, info)."""
    from threadpoolctl import threadpool_limits
    info = host_cpu_info()
    out = None
    tim = {}
    legs = [("all", info["threads_all"])] + ([("one", 1)] if one_thread else [])
    for label, nt in legs:
        with threadpool_limits(limits=nt):
            for _ in range(warmup):
                out = fn()
            runs = []
            for r in range(repeats):
                t0 = time.perf_counter()
                out = fn()
                runs.append(time.perf_counter() - t0)
                print(f"cpu baseline: {nt} thread(s) run {r + 1}/{repeats} {runs[-1]:.3f} s",
                      file=sys.stderr, flush=True)
        tim[label] = {"threads": nt, "median_s": float(np.median(runs)), "runs_s": runs}
    return tim, out, info


def cpu_baseline_obj(tim, info, units_per_run, unit, sample, kind="port", extra=None):
    """The cpu_baseline object: the all-cores median is ``value``; the 1-thread
    median, the core count and the CPU model are reported beside it."""
    a = tim["all"]
    obj = {"value": units_per_run / a["median_s"], "unit": unit, "cores": a["threads"],
           "kind": kind, "sample": sample, "median_s": a["median_s"],
           "units_per_run": units_per_run, "runs": len(a["runs_s"]), "warmup_runs": 1,
           "nproc": info["nproc"], "cpu_model": info["cpu_model"],
           "lscpu_model": info.get("lscpu_model", "")}
    if "one" in tim:
        o = tim["one"]
        obj["one_thread"] = {"value": units_per_run / o["median_s"], "cores": 1,
                             "median_s": o["median_s"], "runs": len(o["runs_s"])}
    if extra:
        obj.update(extra)
    return obj


def cpu_sample(fn, units_per_run, unit, sample, repeats=5, extra=None):
    """The plan's CPU baseline for one workload: ``fn`` (one bounded sample of
    it) timed by time_cpu -- median of ``repeats`` runs after a warm-up, at all
    the process's cores and at 1 thread -- as a cpu_baseline object whose
    ``value`` is ``units_per_run`` over the all-cores median.  Returns
    (object, fn's last result)."""
    tim, out, info = time_cpu(fn, repeats=repeats)
    desc = (f"{sample}; median of {repeats} runs after one warm-up, at all "
            f"{info['threads_all']} of the process's cores and at 1 thread")
    return cpu_baseline_obj(tim, info, units_per_run, unit, desc, extra=extra), out


def oracle_run(data):
    """FP64 oracle (numpy/LAPACK restatement of the reference CPU path) on the
    host cores, timed as the CPU baseline (median of 5 after a warm-up, all
    cores and 1 thread)."""
    from oracle import fastpls_oracle as O
    Xtr, ytr, Xte, yte = data
    X64, Xt64 = Xtr.astype(np.float64), Xte.astype(np.float64)
    tim, ref, info = time_cpu(lambda: O.pls_plssvd_lda(X64, ytr, Xt64, [NCOMP]))
    _, sc = O.lda_project_predict_labels(Xt64, ref["R_predict"][:, :NCOMP],
                                         ref["R_offset"][:NCOMP], ref["lda"][NCOMP],
                                         return_scores=True)
    return ref, sc, tim, info


# ---------------------------------------------------------------- nested CV
CV_NCOMP = list(range(1, 31))
CV_SEED, CV_RUN = 1, 1


def cv_masks(outer, inner):
    """The inner fits of one run of the double CV: train = outer-train rows
    outside the inner fold, test = the inner fold; seeds as pls.double.cv
    passes them (seed + 1000 run + f, plus the inner fold index)."""
    n = outer.size
    J = int(outer.max())
    cols, seeds, owner = [], [], []
    for j in range(J):
        tr = outer != j + 1
        for fi, iv in enumerate(np.unique(inner[tr, j])):
            mk = np.zeros(n, dtype=np.int8)
            mk[tr] = 1
            mk[tr & (inner[:, j] == iv)] = 2
            cols.append(mk)
            seeds.append(CV_SEED + 1000 * CV_RUN + (j + 1) + fi)
            owner.append(j)
    return np.asfortranarray(np.stack(cols, axis=1)), np.asarray(seeds), np.asarray(owner)


def run_cv(args):
    """50 inner SIMPLS fits (ncomp 1..30) of a 10 x 5 double CV over one
    resident X (5,000 x 20,000 FP32): the batched engine (all fits in one
    captured graph; FP32 and split-BF16x3 MFMA) against the fold-by-fold FP32
    driver (the plain path: each outer fold's inner CV fits its folds one by
    one on a gathered copy of the outer-training rows)."""
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    X, y, outer, inner = synth.nested_cv_lognormal()
    masks, seeds, owner = cv_masks(outer, inner)
    nfit = masks.shape[1]
    dX = F.device_colmajor(X)
    dy = torch.from_numpy(y).to("cuda:0")
    J = int(outer.max())
    dXtr = [F.device_colmajor(X[outer != j + 1]) for j in range(J)]
    torch.cuda.synchronize()
    ctx = F.context(0)
    out = {}

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    for prec in ("fp32", "bf16x3"):
        def step(scores=False):
            return F.cv_batched(dX, dy, masks, CV_NCOMP, seeds, precision=prec, n_classes=5,
                                return_class_pred=scores, return_scores=scores)
        dt = timed(step, args.steps, args.warmup)
        res = step(scores=True)
        ctx.timing(True)
        step()
        kern = {}
        for nm in ("cv_project_f32", "cv_project_bf16x3", "gemm_tn_f32", "gemm_tn_bf16x3",
                   "gemm_tn_reduce", "tn_skinny_f32", "tn_skinny_bf16x3", "tn_skinny_reduce", "cv_treduce", "cv_start", "cv_power", "cv_rr", "cv_vproj",
                   "cv_vorth", "cv_vnorm", "cv_pred", "cv_sinit", "cv_zbuild", "gauss_mt19937",
                   "col_sums", "cv_init"):
            n_, ms = ctx.kernel_time(nm)
            if n_:
                kern[nm] = {"launches_per_step": n_, "ms_per_step": ms}
        ctx.timing(False)
        out["batched_" + prec] = {"fits_per_s": nfit / dt, "ms_per_step": 1e3 * dt,
                                  "res": res, "kernels": kern}

    if args.forms == "batched":
        return X, y, outer, inner, masks, seeds, owner, out

    def plain_step(scores=False):
        res = []
        for j in range(J):
            tr = outer != j + 1
            res.append(F.pls_cv_predict(dXtr[j], y[tr], inner[tr, j], ncomp=CV_NCOMP,
                                        seed=CV_SEED + 1000 * CV_RUN + (j + 1),
                                        return_scores=scores, engine="folds"))
        return res
    dt = timed(plain_step, max(1, args.steps // 10), 1)
    out["plain_fp32"] = {"fits_per_s": nfit / dt, "ms_per_step": 1e3 * dt,
                         "res": plain_step(scores=True)}
    return X, y, outer, inner, masks, seeds, owner, out


def cv_oracle_check(X, y, outer, inner, masks, seeds, out, n_sample):
    """FP64 oracle fits of the first ``n_sample`` inner fits (timed: the CPU
    baseline) and the errors of every GPU form on those fits."""
    from oracle import fastpls_oracle as O
    n = X.shape[0]
    Yd = np.zeros((n, 5))
    Yd[np.arange(n), y - 1] = 1.0
    X64 = X.astype(np.float64)

    def one_fit(f):
        tr = masks[:, f] == 1
        te = masks[:, f] == 2
        mod = O.pls_model2_fast(X64[tr], Yd[tr], CV_NCOMP, 1, False, "rsvd", 10, 1,
                                int(seeds[f]))
        return te, O.pls_predict(mod, X64[te])
    cpu, r0 = cpu_sample(lambda: one_fit(0), 1.0, "fits/s",
                         "inner fit 0 (ncomp 1..30: FP64 numpy oracle SIMPLS fit + predict of "
                         "its held-out rows) of the same workload")
    refs = [r0] + [one_fit(f) for f in range(1, n_sample)]
    del X64
    errors = {}
    for form, r in out.items():
        num = den = 0.0
        nontie = agree = tot = 0
        for f, (te, yp) in enumerate(refs):
            if form.startswith("batched"):
                got = r["res"]["Ypred"][te, :, :, f]
            else:  # fit f is inner fold f of outer fold 0 (plain result j = 0)
                j0 = r["res"][0]
                rows_tr = np.nonzero(outer != 1)[0]
                sel = np.isin(rows_tr, np.nonzero(te)[0])
                got = j0["Ypred"][sel]
            num += float(np.sum((got - yp) ** 2))
            den += float(np.sum(yp ** 2))
            lab = np.argmax(got, axis=1) + 1
            ref_lab = np.argmax(yp, axis=1) + 1
            srt = np.sort(yp, axis=1)
            margin = srt[:, -1] - srt[:, -2]
            diff = lab != ref_lab
            nontie += int((diff & (margin > 1e-3 * np.abs(yp).max())).sum())
            agree += int((~diff).sum())
            tot += diff.size
        errors[form] = {"Ypred_rel_err": (num / den) ** 0.5, "label_agreement": agree / tot,
                        "label_mismatch_non_tie": nontie}
    return errors, cpu


def main_cv(args):
    X, y, outer, inner, masks, seeds, owner, out = run_cv(args)
    errors, cpu = ({}, None)
    if not args.no_cpu_baseline and "plain_fp32" in out:
        errors, cpu = cv_oracle_check(X, y, outer, inner, masks, seeds, out, 2)
    # a form qualifies with no label change outside near-ties and predictions
    # within 1e-3 of the FP64 oracle
    ok = {k: (not errors) or (errors[k]["label_mismatch_non_tie"] == 0
                              and errors[k]["Ypred_rel_err"] <= 1e-3)
          for k in ("batched_bf16x3", "batched_fp32")}
    chosen = "batched_bf16x3" if ok["batched_bf16x3"] else "batched_fp32"
    r = out[chosen]
    sfx = "bf16x3" if chosen.endswith("bf16x3") else "f32"
    n, p = X.shape
    nfit = masks.shape[1]
    # the two passes over X per component: T = Xc R (NN, FP32 result) and
    # P = Xc^T T (TN, FP64 result); algorithmic bytes = X once + the skinny
    # operand + the result; the roofline reports the one that takes longer
    alg = {"cv_project_" + sfx: 4.0 * n * p + 4.0 * p * nfit + 4.0 * n * nfit,
           "tn_skinny_" + sfx: 4.0 * n * p + 4.0 * n * nfit + 8.0 * p * nfit}
    cand = [k for k in alg if k in r["kernels"]]
    roof = None
    if cand:
        kname = max(cand, key=lambda k: r["kernels"][k]["ms_per_step"])
        kt = r["kernels"][kname]
        launches = kt["launches_per_step"]
        avg_s = kt["ms_per_step"] * 1e-3 / launches
        bytes_per = alg[kname]
        flops_per = 2.0 * n * p * nfit
        achieved = bytes_per / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(kname, launches),
                "avg_launch_us": avg_s * 1e6, "launches_per_step": launches,
                "algorithmic_bytes_per_launch": bytes_per,
                "algorithmic_flops_per_launch": flops_per,
                "tflops": flops_per / avg_s / 1e12,
                "other_pass": {k: {"avg_launch_us": 1e3 * r["kernels"][k]["ms_per_step"]
                                   / r["kernels"][k]["launches_per_step"],
                                   "GBps": alg[k] / (1e-3 * r["kernels"][k]["ms_per_step"]
                                                     / r["kernels"][k]["launches_per_step"]) / 1e9}
                               for k in cand if k != kname}}
    res = {
        "metric": METRIC, "value": r["fits_per_s"], "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3" if chosen.endswith("bf16x3") else "f32",
        "data": "synthetic (seeded; log-normal, 200-column blocks with correlation 0.6, "
                "5 classes 0.4/0.25/0.15/0.12/0.08, Dirichlet(5) fold sizes)",
        "config": {"workload": "simpls nested CV: 50 inner fits (10 outer x 5 inner), "
                               "n=5000 p=20000 C=5 ncomp=1..30, FP32 X",
                   "chosen_form": chosen, "inputs": "resident in HBM",
                   "fits_per_step": nfit},
        "plain_fp32_fits_per_s": out["plain_fp32"]["fits_per_s"] if "plain_fp32" in out else None,
        "batched_fp32_fits_per_s": out["batched_fp32"]["fits_per_s"],
        "chosen_vs_plain": (r["fits_per_s"] / out["plain_fp32"]["fits_per_s"]
                            if "plain_fp32" in out else None),
        "roofline": roof,
        "oracle_errors": errors,
        "kernels": {k: v["kernels"] for k, v in out.items() if "kernels" in v},
    }
    if cpu is not None:
        res["cpu_baseline"] = cpu
    print(json.dumps(res))


# rSVD of S = X^T Y with a learned vs a Gaussian sketch: X 20,000 x 50,000
# FP32 (rank-80 part with s_i = 50 * 0.93^i on random orthonormal factors,
# plus N(0, 1)), Y = Y0c G / 10 (20,000 x 100) from Y0 = X B + N(0, 1)
# (2,000 columns, B ~ N(0, 1) / sqrt(p)), k = 50, oversampling 10.
C5_N, C5_P, C5_RANK, C5_M0, C5_M, C5_K, C5_OS = 20_000, 50_000, 80, 2_000, 100, 50, 10
C5_TRAIN, C5_EVAL, C5_FRAC, C5_ITERS, C5_LR = 200, 50, 0.8, 300, 0.02


def c5_data(seed):
    """The workload's X (column-major FP32 on the device) and Y, generated on
    the GPU from one seeded torch generator (orthonormal factors from numpy)."""
    import torch
    dev = "cuda:0"
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((C5_N, C5_RANK)))[0]
    V = np.linalg.qr(rng.standard_normal((C5_P, C5_RANK)))[0]
    s = 50.0 * 0.93 ** np.arange(C5_RANK)
    Us = torch.from_numpy(U * s).to(dev)
    Vd = torch.from_numpy(V).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    Xt = torch.randn(C5_P, C5_N, generator=g, device=dev, dtype=torch.float32)
    for j0 in range(0, C5_P, 5000):  # rows of X^T = columns of X
        Xt[j0:j0 + 5000] += (Vd[j0:j0 + 5000] @ Us.T).float()
    X = Xt.t()  # n x p, column-major
    B = torch.randn(C5_P, C5_M0, generator=g, device=dev, dtype=torch.float32) / C5_P ** 0.5
    Y0 = X @ B + torch.randn(C5_N, C5_M0, generator=g, device=dev, dtype=torch.float32)
    Y0 -= Y0.mean(0)
    G = torch.randn(C5_M0, C5_M, generator=g, device=dev, dtype=torch.float32)
    Y = ((Y0 @ G) / 10.0).t().contiguous().t()
    del B, Y0, G, Us, Vd
    torch.cuda.synchronize()
    return X, Y


def main_sketch(args):
    """Matrix-free rSVD with Gaussian sketches (q = 0..3: the reference
    algorithm, q = 1 its default) against the learned sketch and against S
    formed once; sin(theta) of every form against the FP64 SVD of S, the
    learned sketch trained on fold draws and judged on fresh draws."""
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    X, Y = c5_data(synth.SEED_BASE + 5)
    n, p, m, k = C5_N, C5_P, C5_M, C5_K
    l = k + C5_OS
    # FP64 reference: S = X^T Y in FP64 and its top-k left singular subspace
    X64 = X.double()
    S64 = X64.t() @ Y.double()
    del X64
    S64 = S64.cpu().numpy()
    lam, E = np.linalg.eigh(S64.T @ S64)
    order = np.argsort(-lam)
    sig = np.sqrt(np.maximum(lam[order], 0.0))
    Uex = np.linalg.qr((S64 @ E[:, order[:k]]) / sig[:k])[0]
    del S64

    def sin_dev(U):
        Q = np.linalg.qr(U)[0]
        R = Q - Uex @ (Uex.T @ Q)
        return float(np.sqrt(max(np.linalg.eigvalsh(R.T @ R).max(), 0.0)))

    rng = np.random.default_rng(synth.SEED_BASE + 6)
    train = (rng.random((n, C5_TRAIN)) < C5_FRAC).astype(np.int8)
    fresh = (rng.random((n, C5_EVAL)) < C5_FRAC).astype(np.int8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tr = F.sketch_train(X, Y, train, k, l, iters=C5_ITERS, lr=C5_LR, seed=1)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    om = tr["omega"]
    draws = {}
    for q in range(4):
        st = F.sketch_eval(X, Y, fresh, k, l=l, power=q, seed=1)
        draws[f"gaussian_q{q}"] = {"median": float(np.median(st)), "max": float(st.max())}
    for q in range(2):
        st = F.sketch_eval(X, Y, fresh, k, omega=om, power=q)
        draws[f"learned_q{q}"] = {"median": float(np.median(st)), "max": float(st.max())}

    ctx = F.context(0)
    forms = {}
    kern_names = ("tn_skinny_f32", "tn_skinny_bf16x3", "tn_skinny_reduce", "gemm_nn_f32",
                  "gemm_nn_bf16x3", "gemm_nn_reduce", "pad_b", "gemm_tn_f32", "gemm_tn_bf16x3",
                  "gemm_tn_reduce", "gemm_f64", "gemm_f64_reduce", "sym_eig_dc", "convert",
                  "copy", "fill", "chol_inv", "scale_cols", "scale_rows", "transpose", "symmatu",
                  "div", "norm2", "jacobi_cols")

    def run_form(name, **kw):
        def step(fetch=False):
            return F.truncated_rsvd_crossprod(X, Y, k, rsvd_oversample=C5_OS, seed=1, fetch=fetch,
                                              **kw)
        res = step(fetch=True)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ctx.timing(True)
        step()
        kern = {}
        for nm in kern_names:
            c_, ms = ctx.kernel_time(nm)
            if c_:
                kern[nm] = {"launches_per_step": c_, "ms_per_step": ms}
        ctx.timing(False)
        forms[name] = {"ms": 1e3 * dt, "fits_per_s": 1.0 / dt, "sin_theta": sin_dev(res["U"]),
                       "kernels": kern}

    for q in range(4):
        run_form(f"matrix_free_gaussian_q{q}_fp32", rsvd_power=q)
    run_form("matrix_free_learned_q0_fp32", rsvd_power=0, omega=om)
    run_form("matrix_free_learned_q1_fp32", rsvd_power=1, omega=om)
    run_form("formed_S_gaussian_q1_fp32", rsvd_power=1, route="materialise")
    run_form("formed_S_learned_q0_fp32", rsvd_power=0, omega=om, route="materialise")
    run_form("formed_S_exact_fp32", exact=True)
    run_form("formed_S_exact_bf16x3", exact=True, precision="bf16x3")

    # SIMPLS ncomp = 50 fit of the same data
    def simpls():
        return F.pls(X, Y, ncomp=k, method="simpls", fetch=False)
    simpls()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        simpls()
    torch.cuda.synchronize()
    simpls_ms = 1e3 * (time.perf_counter() - t0) / 2

    plain = forms["matrix_free_gaussian_q1_fp32"]
    ok = [f for f in forms if forms[f]["sin_theta"] <= SIN_THETA_BOUND]
    chosen = max(ok, key=lambda f: forms[f]["fits_per_s"]) if ok else \
        min(forms, key=lambda f: forms[f]["sin_theta"])
    r = forms[chosen]
    kname = "tn_skinny_bf16x3" if "bf16x3" in chosen else "tn_skinny_f32"
    roof = None
    if kname in r["kernels"]:
        kt = r["kernels"][kname]
        launches = kt["launches_per_step"]
        avg_s = kt["ms_per_step"] * 1e-3 / launches
        # the S pass: X read once, the FP32 Y read, the split-K FP32 partials of S
        bytes_per = 4.0 * n * p + 4.0 * n * m
        achieved = bytes_per / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(kname + "_c5", launches),
                "avg_launch_us": avg_s * 1e6, "launches_per_step": launches,
                "algorithmic_bytes_per_launch": bytes_per,
                "tflops": 2.0 * n * p * m / avg_s / 1e12}
    out = {
        "metric": METRIC, "value": r["fits_per_s"], "unit": "rSVD/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3" if "bf16x3" in chosen else "f32",
        "data": "synthetic (seeded, generated on the GPU; rank-80 signal s_i = 50*0.93^i "
                "plus N(0,1); Y = Y0c G / 10 from Y0 = X B + N(0,1))",
        "config": {"workload": f"rsvd of X^T Y, X {n}x{p} FP32, Y {n}x{m}, k={k}, l={l}",
                   "chosen_form": chosen, "inputs": "resident in HBM",
                   "sin_theta_bound": SIN_THETA_BOUND},
        "plain_form": "matrix_free_gaussian_q1_fp32",
        "plain_fits_per_s": plain["fits_per_s"], "plain_sin_theta": plain["sin_theta"],
        "chosen_vs_plain": r["fits_per_s"] / plain["fits_per_s"],
        "forms": forms,
        "sketch": {"train_draws": C5_TRAIN, "eval_draws": C5_EVAL, "row_fraction": C5_FRAC,
                   "iters": C5_ITERS, "lr": C5_LR, "train_seconds": train_s,
                   "loss_first": float(tr["loss"][0]), "loss_last": float(tr["loss"][-1]),
                   "fresh_draw_sin_theta": draws},
        "spectrum": {"s1": float(sig[0]), "s50": float(sig[k - 1]), "s51": float(sig[k]),
                     "s61": float(sig[l]), "s100": float(sig[-1])},
        "simpls_ncomp50_ms": simpls_ms,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = c5_cpu_baseline(X, Y)
    print(json.dumps(out))


C5_CPU_ROWS = 5_000


def c5_cpu_baseline(X, Y):
    """The oracle's matrix-free rSVD (FP64, q = 1: the reference default) of
    S = X^T Y over the first C5_CPU_ROWS rows on the host cores.  Its passes
    over X and Y are linear in the rows and the l x l factorisations are
    row-independent, so the rate is scaled by C5_CPU_ROWS / n (which counts
    the small row-independent part n / C5_CPU_ROWS times: a conservative
    CPU rate)."""
    from oracle import fastpls_oracle as O
    r = min(C5_CPU_ROWS, X.shape[0])
    Xh = X[:r].cpu().numpy().astype(np.float64)
    Yh = Y[:r].cpu().numpy().astype(np.float64)
    cpu, _ = cpu_sample(lambda: O.truncated_rsvd_crossprod(Xh, Yh, C5_K, C5_OS, 1, 1),
                        r / X.shape[0], "rSVD/s",
                        f"one matrix-free rSVD (q=1, FP64 numpy oracle) of X^T Y over the "
                        f"first {r} of the {X.shape[0]} rows, rate scaled by {r}/{X.shape[0]}")
    return cpu


# Label-aware PLSSVD + candidate kNN at 1000 classes: train 1,200,000 x 1024
# FP32 (Zipf(1.1) class sizes, means on the radius-4 sphere, N(0, 0.6^2)),
# 50,000 test rows, ncomp 500, knn_k 10, tau 0.2, alpha 0.75, top_m 20, top-5.
C3_C, C3_K, C3_TOP, C3_SAMPLE = 1000, 500, 5, 500


def c3_pair_work(m, dX, dXt, dy):
    """Similarities the candidate-kNN kernel has to compute: sum over (test
    row, candidate class) pairs of the class size, from the candidate lists
    rebuilt in torch (FP32) from the model's projection, C_latent slice and
    the class centroids of the normalised score space."""
    import torch
    p = dX.shape[1]
    Rp = torch.from_numpy(m.read("R_predict", (p, C3_K))).float().to("cuda:0")
    off = torch.from_numpy(m.read("R_offset", (C3_K,))).float().to("cuda:0")
    Ck = torch.from_numpy(m.read("C_latent", (C3_K, C3_K, 1))[:, :, 0]).float().to("cuda:0")

    def space(X):
        T = (X @ Rp - off) @ Ck
        nr = T.norm(dim=1, keepdim=True)
        nr[(nr == 0) | ~torch.isfinite(nr)] = 1.0
        return T / nr
    Str = space(dX)
    yi = dy.long() - 1
    sums = torch.zeros(C3_C, C3_K, device="cuda:0").index_add_(0, yi, Str)
    counts = torch.bincount(yi, minlength=C3_C)
    cen = sums / counts.clamp_min(1)[:, None].float()
    cen = cen / cen.norm(dim=1, keepdim=True).clamp_min(1e-30)
    del Str
    cand = torch.topk(space(dXt) @ cen.T, 20, dim=1).indices
    return int(counts[cand].sum().item())


def main_cknn(args):
    """pls(method = "plssvd", classifier = "cknn", label-aware) + predict(top = 5)
    on the large-class-count workload; FP32 and split-BF16x3 contractions."""
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.zipf_classes_large()
    dX = F.device_colmajor(Xtr)
    dXt = F.device_colmajor(Xte)
    dy = torch.from_numpy(ytr).to("cuda:0")
    torch.cuda.synchronize()
    ctx = F.context(0)
    names = ("cknn_partial", "cknn_merge", "cknn_rerank", "cknn_topm", "cknn_pairs",
             "cknn_rownorm", "centroid_sums", "centroids", "class_sort", "gemm_tn_bf16x3",
             "gemm_tn_f32", "gemm_tn_reduce", "gemm_nn_bf16x3", "gemm_nn_f32", "project_bf16x3",
             "project_f32", "gemm_rr_bf16x3", "gemm_rr_f32", "sym_eig_dc", "jacobi_cols",
             "jacobi_nearid", "jacobi_offnorm", "jacobi_gather", "gemm_f64", "gemm_f64_reduce",
             "col_sums", "spd_solve", "transpose", "copy", "fill", "eig_chip_col",
             "eig_chip_start", "eig_chip_pairs", "eig_chip_orth", "eig_chip_tfac",
             "eig_chip_sym", "eig_chip_end")
    res = {}
    model_keep = None
    for prec in ("fp32", "bf16x3"):
        def step(fetch=False):
            m = F.pls(dX, dy, method="plssvd", ncomp=C3_K, classifier="cknn", label_aware=True,
                      candidate_knn_k=10, candidate_tau=0.2, candidate_alpha=0.75,
                      candidate_top_m=20, precision=prec, n_classes=C3_C, fetch=fetch)
            r = F.predict(m, dXt, top=C3_TOP)
            return m, r
        for _ in range(args.warmup):
            step()[0].free()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()[0].free()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ctx.timing(True)
        m, r = step()
        kern = {}
        for nm in names:
            c_, ms = ctx.kernel_time(nm)
            if c_:
                kern[nm] = {"launches_per_step": c_, "ms_per_step": ms}
        ctx.timing(False)
        res[prec] = {"fits_per_s": 1.0 / dt, "ms_per_step": 1e3 * dt, "kernels": kern,
                     "top": np.asarray(r["Ypred_top"][:, :, 0]),
                     "top_score": np.asarray(r["Ypred_top_score"][:, :, 0]),
                     "accuracy": float(np.mean(r["Ypred_top"][:, 0, 0] == yte))}
        if model_keep is not None:
            model_keep.free()
        model_keep = m
    work = c3_pair_work(model_keep, dX, dXt, dy)
    model_keep.free()
    errors, cpu = {}, None
    if not args.no_cpu_baseline:
        del dX, dXt
        torch.cuda.empty_cache()
        errors, cpu = c3_oracle_check(Xtr, ytr, Xte, res)
    ok = {p: (not errors) or errors[p]["top5_mismatch_non_tie"] == 0 for p in res}
    chosen = "bf16x3" if ok["bf16x3"] and res["bf16x3"]["fits_per_s"] > res["fp32"]["fits_per_s"] \
        else "fp32"
    r = res[chosen]
    roof = None
    if "cknn_partial" in r["kernels"]:
        kt = r["kernels"]["cknn_partial"]
        launches = kt["launches_per_step"]
        avg_s = kt["ms_per_step"] * 1e-3 / launches
        flops = 2.0 * C3_K * work / launches
        achieved = flops / avg_s / 1e12
        if chosen == "bf16x3":
            # split form: FP16 hi/lo pairs of both operands, three FP16 MFMA
            # products per similarity; peak = the FP16 dense MFMA peak over 3
            peak, note = USEFUL_PEAK_TFLOPS["bf16x3"], (
                "similarities as three FP16 MFMA products of 2^12-scaled hi/lo pairs "
                "(k_cand_gemm_h); achieved = algorithmic FLOPs / time, peak = 2500 / 3")
        else:
            peak, note = PEAK_TFLOPS["fp32"], "similarities on FP32 MFMA (k_cand_gemm)"
        roof = {"bound": "mfma", "kernel": "cknn_partial", "achieved": achieved,
                "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": measured_traffic("cknn_partial_c3_" + chosen, launches),
                "avg_launch_us": avg_s * 1e6, "launches_per_step": launches,
                "algorithmic_flops_per_launch": flops,
                "candidate_pair_rows": work, "note": note}
    out = {
        "metric": METRIC, "value": r["fits_per_s"], "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3" if chosen == "bf16x3" else "f32",
        "data": "synthetic (seeded; Zipf(1.1) class sizes, means on the radius-4 sphere, "
                "N(0, 0.6^2) noise)",
        "config": {"workload": "label-aware plssvd + cknn n=1200000 ntest=50000 p=1024 C=1000 "
                               "ncomp=500 knn_k=10 top_m=20 top=5",
                   "chosen_form": chosen, "inputs": "resident in HBM"},
        "plain_fp32_fits_per_s": res["fp32"]["fits_per_s"],
        "chosen_vs_plain": r["fits_per_s"] / res["fp32"]["fits_per_s"],
        "accuracy": {p: v["accuracy"] for p, v in res.items()},
        "roofline": roof,
        "oracle_errors": errors,
        "kernels": {p: v["kernels"] for p, v in res.items()},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


# Argmax-head class ranking on the same large-class-count data: the label-aware
# PLSSVD model (ncomp 500, 1000 classes) fitted once, then one step ranks the
# 50,000 test rows (top-5 with per-class offsets mY, no class bias).
TK_SAMPLE, TK_CPU_ROWS = 30000, 5000


def main_topk(args):
    """predict(top = 5) of the argmax head (pls_class_predict_topk) on the
    C = 1000 workload; the fit is outside the timed region."""
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.zipf_classes_large()
    dX = F.device_colmajor(Xtr)
    dXt = F.device_colmajor(Xte)
    dy = torch.from_numpy(ytr).to("cuda:0")
    torch.cuda.synchronize()
    ctx = F.context(0)
    m = F.pls(dX, dy, method="plssvd", ncomp=C3_K, label_aware=True, n_classes=C3_C,
              fetch=False)
    del dX
    torch.cuda.empty_cache()
    names = ("project_f32", "project_bf16x3", "gemm_nn_f32", "gemm_nn_reduce", "gemm_f64",
             "gemm_f64_reduce", "class_topk", "class_topk_rank", "class_topk_offsets")

    # one step: the C-ABI call an R binding makes (outputs into host buffers
    # allocated once); label decoration on the host is not part of it
    import ctypes
    from fastpls_amd._lib import as_mat
    keep = []
    Xm = as_mat(dXt, keep)
    n_test = Xte.shape[0]
    ti = np.empty((n_test, C3_TOP, 1), dtype=np.int32, order="F")
    ts = np.empty((n_test, C3_TOP, 1), dtype=np.float64, order="F")

    def step():
        ctx.check(ctx.lib.fpls_class_predict_topk(ctx.h, m.handle, ctypes.byref(Xm), None, 0,
                                                  C3_TOP, ti.ctypes.data, ts.ctypes.data))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    r = {"Ypred_top": ti, "Ypred_top_score": ts}
    ctx.timing(True)
    step()
    kern = {}
    for nm in names:
        c_, ms = ctx.kernel_time(nm)
        if c_:
            kern[nm] = {"launches_per_step": c_, "ms_per_step": ms}
    ctx.timing(False)
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    kt = kern[dom]
    avg_s = kt["ms_per_step"] * 1e-3 / kt["launches_per_step"]
    if dom.startswith("gemm_f64"):
        flops = 2.0 * n_test * C3_K * C3_C / kt["launches_per_step"]
        achieved = flops / avg_s / 1e12
        roof = {"bound": "fp64", "kernel": dom, "achieved": achieved, "peak": PEAK_TFLOPS["fp64"],
                "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp64"]}
    elif dom.startswith("project"):
        flops = 2.0 * n_test * Xte.shape[1] * C3_K / kt["launches_per_step"]
        achieved = flops / avg_s / 1e12
        roof = {"bound": "mfma", "kernel": dom, "achieved": achieved,
                "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["fp32"]}
    else:
        byts = (8.0 * n_test * C3_C + 12.0 * n_test * C3_TOP) / kt["launches_per_step"]
        achieved = byts / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS}
    roof.update({"traffic": measured_traffic(dom + "_topk", kt["launches_per_step"]),
                 "avg_launch_us": avg_s * 1e6, "launches_per_step": kt["launches_per_step"]})
    errors, cpu = {}, None
    if not args.no_cpu_baseline:
        errors, cpu = topk_oracle_check(m, Xte, r)
    out = {
        "metric": "argmax-head top-5 ranked test rows/s", "value": n_test / dt, "unit": "rows/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded; Zipf(1.1) class sizes, means on the radius-4 sphere, "
                "N(0, 0.6^2) noise)",
        "config": {"workload": "class-bias top-5 prediction ntest=50000 p=1024 C=1000 ncomp=500 "
                               "(label-aware plssvd model, FP32 inputs)",
                   "inputs": "resident in HBM", "outputs": "host (top index + score)"},
        "accuracy": float(np.mean(r["Ypred_top"][:, 0, 0] == yte)),
        "roofline": roof, "oracle_errors": errors, "kernels": kern,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    m.free()
    print(json.dumps(out))


def topk_oracle_check(m, Xte, r):
    """The oracle's pls_class_predict_topk on the device model's fields for the
    first TK_SAMPLE test rows, compared with the device's; its first
    TK_CPU_ROWS rows are the timed CPU baseline (per row)."""
    from oracle import fastpls_oracle as O
    inf = m.info()
    p, k, C, L = Xte.shape[1], inf.kmax, inf.m, inf.n_ncomp
    mod = {"R": m.read("R", (p, k)), "W_latent": m.read("W_latent", (k, C, L)),
           "mX": m.read("mX", (p,)), "vX": m.read("vX", (p,)), "mY": m.read("mY", (C,)),
           "ncomp": np.array([inf.ncomp[i] for i in range(L)]), "m": C, "pls_method": "plssvd"}
    Xs = Xte[:TK_CPU_ROWS].astype(np.float64)
    cpu, (ri0, rs0) = cpu_sample(
        lambda: O.pls_class_predict_topk(mod, Xs, None, C3_TOP), float(TK_CPU_ROWS), "rows/s",
        f"FP64 numpy oracle pls_class_predict_topk on the first {TK_CPU_ROWS} test rows "
        "(NumPy BLAS products + the insertion loop)")
    ri1, rs1 = O.pls_class_predict_topk(mod, Xte[TK_CPU_ROWS:TK_SAMPLE].astype(np.float64),
                                        None, C3_TOP)
    ri, rs = np.concatenate([ri0, ri1]), np.concatenate([rs0, rs1])
    ti, ts = r["Ypred_top"][:TK_SAMPLE, :, 0], r["Ypred_top_score"][:TK_SAMPLE, :, 0]
    ri, rs = ri[:, :, 0], rs[:, :, 0]
    bad = 0
    for i in range(ri.shape[0]):
        if not np.array_equal(ti[i], ri[i]):
            gaps = np.abs(np.diff(rs[i]))
            if not (gaps.size and np.min(gaps) < 1e-4 * np.abs(rs[i]).max()):
                bad += 1
    errors = {"top5_equal_rows": float(np.mean(np.all(ti == ri, axis=1))),
              "top5_mismatch_non_tie": bad,
              "max_abs_score_err": float(np.max(np.abs(ts - rs))),
              "max_abs_score": float(np.max(np.abs(rs)))}
    return errors, cpu


# Kernel PLS (rbf, gamma = 1/p) on 100 classes: 10,000 training rows x 768
# (FP32, classes on the sphere), 2,000 test rows, ncomp 20.  One step = the
# training kernel (10,000 x 10,000 FP64) built and centred in HBM, the SIMPLS
# fit on it, the test kernel and the prediction.  The OPLS filter (north 1)
# + SIMPLS + prediction on the 50,000-row classification data is timed beside.
KP_C, KP_TRAIN, KP_TEST, KP_P, KP_K = 100, 100, 20, 768, 20


def main_kernel(args):
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    Xtr, ytr, Xte, yte = synth.classes_on_sphere(n_classes=KP_C, per_class_train=KP_TRAIN,
                                                 per_class_test=KP_TEST, p=KP_P, seed=71)
    dX, dXt = F.device_colmajor(Xtr), F.device_colmajor(Xte)
    torch.cuda.synchronize()
    ctx = F.context(0)
    n, nt = Xtr.shape[0], Xte.shape[0]

    def kstep():
        m = F.pls(dX, ytr, dXt, method="kernelpls", kernel="rbf", ncomp=KP_K, fetch=False)
        lab = m["Ypred_index"]
        m.free()
        return lab
    for _ in range(args.warmup):
        kstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lab = kstep()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.timing(True)
    kstep()
    names = ("kernel_gemm", "kernel_epilogue", "row_sumsq", "row_means", "kernel_center",
             "kernel_colmeans", "col_sums", "opls_prep", "gemm_f64", "gemm_f64_reduce", "gemv_n",
             "gemv_t", "class_topk")
    kern = {}
    for nm in names:
        c_, ms = ctx.kernel_time(nm)
        if c_:
            kern[nm] = {"launches_per_step": c_, "ms_per_step": ms}
    ctx.timing(False)
    kt = kern["kernel_gemm"]
    avg_s = kt["ms_per_step"] * 1e-3 / kt["launches_per_step"]
    flops = 2.0 * (n * n + nt * n) * KP_P / kt["launches_per_step"]
    achieved = flops / avg_s / 1e12
    roof = {"bound": "fp64", "kernel": "kernel_gemm", "achieved": achieved,
            "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp64"],
            "traffic": measured_traffic("kernel_gemm_kp", kt["launches_per_step"]),
            "avg_launch_us": avg_s * 1e6, "launches_per_step": kt["launches_per_step"],
            "algorithmic_flops_per_launch": flops}
    # OPLS on the 50,000-row classification data
    data = synth.classes_on_sphere(n_classes=C, per_class_train=N_TRAIN // C,
                                   per_class_test=N_TEST // C, p=P)
    oX, oXt = F.device_colmajor(data[0]), F.device_colmajor(data[2])

    def ostep():
        m = F.pls(oX, data[1], oXt, method="opls", north=1, ncomp=KP_K, fetch=False)
        lab = m["Ypred_index"]
        m.free()
        return lab
    ostep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        olab = ostep()
    torch.cuda.synchronize()
    odt = (time.perf_counter() - t0) / args.steps
    errors, cpu = {}, None
    if not args.no_cpu_baseline:
        from oracle import fastpls_oracle as O
        X64, Xt64 = Xtr.astype(np.float64), Xte.astype(np.float64)
        cpu, ref = cpu_sample(
            lambda: O.kernel_pls(X64, ytr, Xt64, [KP_K], kernel=2, classes=KP_C), 1.0,
            "fits/s", "one full FP64 kernel-PLS fit + predict of the same workload (numpy oracle)")
        rl = np.argmax(ref["Ypred"][:, :, 0], axis=1) + 1
        s = np.sort(ref["Ypred"][:, :, 0], axis=1)
        margin = s[:, -1] - s[:, -2]
        diff = lab[:, 0] != rl
        errors = {"label_agreement": float(1.0 - diff.mean()),
                  "label_mismatch_non_tie": int((diff & (margin > 1e-6)).sum()),
                  "accuracy": float(np.mean(lab[:, 0] == yte)),
                  "oracle_accuracy": float(np.mean(rl == yte))}
    out = {
        "metric": "kernel PLS (rbf) fits/s", "value": 1.0 / dt, "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded; classes on the sphere)",
        "config": {"workload": f"kernelpls rbf n={n} ntest={nt} p={KP_P} C={KP_C} ncomp={KP_K}",
                   "inputs": "resident in HBM (FP32)"},
        "opls": {"workload": f"opls north=1 n={N_TRAIN} ntest={N_TEST} p={P} C={C} ncomp={KP_K}",
                 "fits_per_s": 1.0 / odt, "ms_per_step": 1e3 * odt,
                 "accuracy": float(np.mean(olab[:, 0] == data[3]))},
        "roofline": roof, "oracle_errors": errors, "kernels": kern,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


# Gaussian response compression (gaussian_y): SIMPLS regression of a 1,000-column
# dense response on 50,000 x 768 FP32 rows (rank-40 signal + noise), the
# response projected to 100 dimensions, ncomp 20, 10,000 test rows decoded back
# to 1,000 columns.  One step = encode + fit + predict + decode.
GY_N, GY_NT, GY_P, GY_M, GY_DIM, GY_K = 50_000, 10_000, 768, 1_000, 100, 20


def gy_data():
    import torch
    g = torch.Generator(device="cuda:0").manual_seed(2024)
    dev = "cuda:0"
    W = torch.randn(GY_P, 40, device=dev, generator=g, dtype=torch.float32)
    Bt = torch.randn(40, GY_M, device=dev, generator=g, dtype=torch.float32) / 8.0

    def rows(n):
        L = torch.randn(n, 40, device=dev, generator=g, dtype=torch.float32)
        X = L @ W.t() + 0.5 * torch.randn(n, GY_P, device=dev, generator=g)
        Y = L @ Bt + 0.3 * torch.randn(n, GY_M, device=dev, generator=g)
        return X.t().contiguous().t(), Y.double().t().contiguous().t()
    X, Y = rows(GY_N)
    Xt, Yt = rows(GY_NT)
    torch.cuda.synchronize()
    return X, Y, Xt, Yt


def main_gaussian(args):
    import torch
    import fastpls_amd as F
    dX, dY, dXt, dYt = gy_data()
    ctx = F.context(0)

    def step():
        m = F.pls(dX, dY, dXt, method="simpls", ncomp=GY_K, gaussian_y=True,
                  gaussian_y_dim=GY_DIM, seed=7, fetch=False)
        yp = m["Ypred"]
        m.free()
        return yp
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        yp = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.timing(True)
    step()
    names = ("gy_encode", "gy_cross", "gy_decode", "gy_center", "gemm_f64", "gemm_f64_reduce",
             "project_f32", "gemm_tn_f32", "gemm_nn_f32", "col_sums", "spd_solve", "sym_eig_dc",
             "gemv_n", "gemv_t")
    kern = {}
    for nm in names:
        c_, ms = ctx.kernel_time(nm)
        if c_:
            kern[nm] = {"launches_per_step": c_, "ms_per_step": ms}
    ctx.timing(False)
    dom = max(("gy_encode", "gy_cross", "gy_decode"), key=lambda k: kern.get(k, {"ms_per_step": 0})["ms_per_step"])
    kt = kern[dom]
    avg_s = kt["ms_per_step"] * 1e-3 / kt["launches_per_step"]
    rows = GY_N if dom != "gy_decode" else GY_NT
    flops = 2.0 * rows * GY_M * GY_DIM
    achieved = flops / avg_s / 1e12
    roof = {"bound": "fp64", "kernel": dom, "achieved": achieved, "peak": PEAK_TFLOPS["fp64"],
            "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp64"],
            "traffic": measured_traffic(dom + "_gy", kt["launches_per_step"]),
            "avg_launch_us": avg_s * 1e6, "launches_per_step": kt["launches_per_step"],
            "algorithmic_flops_per_launch": flops}
    Yt = dYt.cpu().numpy()
    q2 = float(1.0 - ((Yt - yp[:, :, 0]) ** 2).sum() / ((Yt - Yt.mean(axis=0)) ** 2).sum())
    errors, cpu = {}, None
    if not args.no_cpu_baseline:
        from oracle import fastpls_oracle as O
        X, Y, Xt = (dX.double().cpu().numpy(), dY.cpu().numpy(), dXt.double().cpu().numpy())

        def gy_run():
            enc = O.prepare_gaussian_y_regression(Y, GY_DIM, 7)
            mod = O.pls_model2_fast(X, enc["Z"], [GY_K], 1, False, "rsvd", 10, 1, 7)
            zt = O.pls_predict(mod, Xt)[:, :, 0]
            return O.decode_gaussian_y_regression(zt, enc["decoder"], enc["y_center"])
        cpu, ref = cpu_sample(gy_run, 1.0, "fits/s",
                              "the full workload through the FP64 numpy oracle (encode, fit, "
                              "predict, decode)")
        errors = {"Ypred_rel_err": float(np.linalg.norm(yp[:, :, 0] - ref) / np.linalg.norm(ref))}
    out = {
        "metric": "gaussian_y SIMPLS fits/s", "value": 1.0 / dt, "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded; rank-40 signal + noise, generated on the device)",
        "config": {"workload": f"simpls gaussian_y n={GY_N} ntest={GY_NT} p={GY_P} m={GY_M} "
                               f"dim={GY_DIM} ncomp={GY_K}",
                   "inputs": "resident in HBM (X FP32, Y FP64)",
                   "outputs": "host (decoded test predictions, 80 MB)"},
        "Q2Y_test": q2, "roofline": roof, "oracle_errors": errors, "kernels": kern,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


C3_CPU_FRAC = 20


def c3_oracle_check(Xtr, ytr, Xte, res):
    """FP64 oracle: full fit on the training rows and prediction of the first
    C3_SAMPLE test rows, against which each GPU form's top-5 lists are
    compared.  The CPU baseline is a 1/C3_CPU_FRAC sample of the workload --
    the oracle fit on the first n/C3_CPU_FRAC training rows and the
    prediction of the first ntest/C3_CPU_FRAC test rows -- with the rate
    scaled by 1/C3_CPU_FRAC (its row-independent part, the S factorisation,
    is counted C3_CPU_FRAC times: a conservative CPU rate)."""
    from oracle import fastpls_oracle as O
    X64 = Xtr.astype(np.float64)
    t0 = time.perf_counter()
    ref = O.plssvd_label_aware(X64, ytr, [C3_K])
    ref["lev"] = list(range(1, C3_C + 1))
    O.fit_candidate_knn(ref, X64, ytr, 10, 0.2, 0.75, 20)
    t_full = time.perf_counter() - t0
    del X64
    rr = O.candidate_knn_predictions(ref, Xte[:C3_SAMPLE].astype(np.float64), top=C3_TOP)[0]
    errors = {}
    for p, v in res.items():
        ti, ts = v["top"][:C3_SAMPLE], v["top_score"][:C3_SAMPLE]
        ri, rs = rr["top_index"], rr["top_score"]
        bad = 0
        for i in range(ri.shape[0]):
            if not np.array_equal(ti[i], ri[i]):
                gaps = np.abs(np.diff(rs[i]))
                if not (gaps.size and np.min(gaps) < 1e-3):
                    bad += 1
        errors[p] = {"top1_agreement": float(np.mean(ti[:, 0] == ri[:, 0])),
                     "top5_equal_rows": float(np.mean(np.all(ti == ri, axis=1))),
                     "top5_mismatch_non_tie": bad,
                     "max_abs_score_err": float(np.max(np.abs(np.sort(ts, 1) - np.sort(rs, 1))))}
    ns, nts = Xtr.shape[0] // C3_CPU_FRAC, Xte.shape[0] // C3_CPU_FRAC
    Xs, ys, Xts = Xtr[:ns].astype(np.float64), ytr[:ns], Xte[:nts].astype(np.float64)

    def sample():
        mod = O.plssvd_label_aware(Xs, ys, [C3_K])
        mod["lev"] = list(range(1, C3_C + 1))
        O.fit_candidate_knn(mod, Xs, ys, 10, 0.2, 0.75, 20)
        return O.candidate_knn_predictions(mod, Xts, top=C3_TOP)[0]
    cpu, _ = cpu_sample(sample, 1.0 / C3_CPU_FRAC, "fits/s",
                        f"FP64 numpy oracle fit on the first {ns} training rows and prediction "
                        f"of the first {nts} test rows (1/{C3_CPU_FRAC} of the workload), rate "
                        f"scaled by 1/{C3_CPU_FRAC}",
                        extra={"full_fit_seconds_once": t_full})
    return errors, cpu


# ---------------------------------------------------------------- SIMPLS regression
C1_N, C1_P, C1_M, C1_NT, C1_K = 200, 2000, 3, 100, 10
C1_CPU_BATCH = 50
# Algorithmic work per fit (FP64): bytes 8 [n p (1 + 2K) + n m + n_t p] (X read
# for S and twice per component, Y once, Xtest once); flops 2npm + K (4np +
# 2nm + 14pm) + 4p sum_a a + 2pmK + 2 n_t p m.
C1_BYTES = 8.0 * (C1_N * C1_P * (1 + 2 * C1_K) + C1_N * C1_M + C1_NT * C1_P)
C1_FLOPS = (2.0 * C1_N * C1_P * C1_M
            + C1_K * (4.0 * C1_N * C1_P + 2.0 * C1_N * C1_M + 14.0 * C1_P * C1_M)
            + 4.0 * C1_P * (C1_K * (C1_K - 1) / 2) + 2.0 * C1_P * C1_M * C1_K
            + 2.0 * C1_NT * C1_P * C1_M)


def main_simpls_reg(args):
    """SIMPLS regression fit + predict (n = 200, p = 2,000, m = 3, ncomp = 10,
    100 test rows, FP64 inputs as the reference's arma::mat): fits/s with X, Y
    and Xtest resident in HBM.  The chosen form is the FP64 path (the
    reference's own precision); the plain form takes FP32 copies of X and
    Xtest.  Both are checked against the FP64 oracle."""
    import torch
    import fastpls_amd as F
    from fastpls_amd import synth
    from oracle import fastpls_oracle as O
    X, Y, Xt, Yt = synth.regression_lowrank(n=C1_N, p=C1_P, m=C1_M, n_test=C1_NT)
    ctx = F.context(0)
    dev = {"fp64": (F.device_colmajor(X), F.device_colmajor(Xt)),
           "fp32": (F.device_colmajor(X.astype(np.float32)),
                    F.device_colmajor(Xt.astype(np.float32)))}
    dY = F.device_colmajor(Y)
    torch.cuda.synchronize()
    res = {}
    knames = ("gemv_n", "gemv_n_reduce", "gemv_t", "simpls_refresh", "simpls_orth",
              "simpls_normalize", "div_vec", "gemm_f64", "gemm_f64_reduce", "gemm_tn_f32",
              "gemm_tn_reduce", "col_sums", "div", "fill", "project_f32", "gemm_nn_f32",
              "gemm_nn_reduce", "add_row")
    for form, (dX, dXt) in dev.items():
        def step(fetch=False):
            m = F.pls(dX, dY, dXt, method="simpls", ncomp=C1_K, fetch=fetch)
            out = (m["Ypred"], m.get("B"))
            m.free()
            return out
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        yp, B = step(fetch=True)
        ctx.timing(True)
        tsteps = 5
        for _ in range(tsteps):
            step()
        kern = {}
        for nm in knames:
            n, ms = ctx.kernel_time(nm)
            if n:
                kern[nm] = {"launches_per_step": n / tsteps, "ms_per_step": ms / tsteps}
        ctx.timing(False)
        res[form] = {"fits_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                     "Ypred": yp, "B": B, "kernels": kern}
    ref = O.pls_model2_fast(X, Y, [C1_K])
    ref_pred = O.pls_predict(ref, Xt)
    errors = {f: {"B_rel_err": float(np.linalg.norm(r["B"] - ref["B"]) / np.linalg.norm(ref["B"])),
                  "Ypred_rel_err": float(np.linalg.norm(r["Ypred"] - ref_pred)
                                         / np.linalg.norm(ref_pred))}
              for f, r in res.items()}
    chosen = "fp64"
    r = res[chosen]
    # dominant data-parallel kernels: the two passes over X per component
    roof = None
    gk = [k for k in ("gemv_n", "gemv_t") if k in r["kernels"]]
    if gk:
        kn = max(gk, key=lambda k: r["kernels"][k]["ms_per_step"])
        kt = r["kernels"][kn]
        avg_s = kt["ms_per_step"] * 1e-3 / kt["launches_per_step"]
        per_launch = 8.0 * C1_N * C1_P
        roof = {"bound": "hbm", "kernel": kn, "achieved": per_launch / avg_s / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": per_launch / avg_s / 1e9 / HBM_PEAK_GBS,
                "traffic": measured_traffic(kn, kt["launches_per_step"]),
                "avg_launch_us": avg_s * 1e6, "launches_per_step": kt["launches_per_step"],
                "algorithmic_bytes_per_launch": per_launch,
                "fit_algorithmic_bytes": C1_BYTES, "fit_algorithmic_flops": C1_FLOPS,
                "fit_GBps": C1_BYTES / (r["ms_per_step"] * 1e-3) / 1e9,
                "note": "a 3.2 MB X per pass: each launch is latency bound, not HBM bound"}
    out = {
        "metric": METRIC, "value": r["fits_per_s"], "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded; rank-20 X = U diag(10 0.8^i) V^T + 0.1 N(0,1), Y = X B + "
                "0.05 N(0,1))",
        "config": {"workload": f"simpls_reg n={C1_N} p={C1_P} m={C1_M} ntest={C1_NT} "
                               f"ncomp={C1_K}", "chosen_form": chosen,
                   "inputs": "resident in HBM"},
        "plain_fp32_fits_per_s": res["fp32"]["fits_per_s"],
        "chosen_vs_plain": r["fits_per_s"] / res["fp32"]["fits_per_s"],
        "roofline": roof, "oracle_errors": errors,
        "kernels": {f: v["kernels"] for f, v in res.items()},
    }
    if not args.no_cpu_baseline:
        def c1_batch():
            for _ in range(C1_CPU_BATCH):
                r = O.pls_predict(O.pls_model2_fast(X, Y, [C1_K]), Xt)
            return r
        out["cpu_baseline"], _ = cpu_sample(
            c1_batch, float(C1_CPU_BATCH), "fits/s",
            f"{C1_CPU_BATCH} full FP64 SIMPLS fits + predictions of the same workload per run "
            "(numpy/LAPACK oracle)")
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["plssvd_lda", "simpls_reg", "cv", "sketch", "cknn",
                                           "topk", "kernelpls", "gaussian"],
                    default="plssvd_lda",
                    help="plssvd_lda: PLSSVD + LDA fit/predict (default); simpls_reg: SIMPLS "
                         "regression fit/predict (n=200, p=2000); cv: batched nested "
                         "cross-validation of SIMPLS; sketch: rSVD of X^T Y with learned vs "
                         "Gaussian sketches; cknn: label-aware PLSSVD + candidate kNN at 1000 "
                         "classes")
    ap.add_argument("--forms", choices=["all", "batched"], default="all",
                    help="cv workload: 'batched' skips the fold-by-fold plain path (profiling)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU run: summarise two rocprofv3 counter passes (FETCH_SIZE, "
                         "WRITE_SIZE) of a run made with --copy-mib into a per-dispatch table "
                         "and profiles/hbm_traffic.json")
    ap.add_argument("--pmc-key", action="append", default=[], metavar="NAME=KERNEL_PREFIX",
                    help="with --pmc-summary: record the kernels named KERNEL_PREFIX... as NAME")
    ap.add_argument("--pmc-workload", default="", help="with --pmc-summary: workload label")
    args = ap.parse_args()
    if args.pmc_summary:
        keys = dict(kv.split("=", 1) for kv in args.pmc_key)
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_mib or 1024, keys,
                    args.pmc_workload)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_mib > 0:
        calibration_copy(args.copy_mib)
    if args.workload == "simpls_reg":
        main_simpls_reg(args)
        return
    if args.workload == "cv":
        main_cv(args)
        return
    if args.workload == "sketch":
        main_sketch(args)
        return
    if args.workload == "cknn":
        main_cknn(args)
        return
    if args.workload == "topk":
        main_topk(args)
        return
    if args.workload == "kernelpls":
        main_kernel(args)
        return
    if args.workload == "gaussian":
        main_gaussian(args)
        return
    from fastpls_amd import synth
    data = synth.classes_on_sphere(n_classes=C, per_class_train=N_TRAIN // C,
                                   per_class_test=N_TEST // C, p=P)
    res = run_gpu(args, data)
    errors = {}
    cpu = None
    ok = {"fp32": True, "bf16x3": True}
    if not args.no_cpu_baseline:
        ref, sc, tim, info = oracle_run(data)
        s = np.sort(sc, axis=1)
        margin = s[:, -1] - s[:, -2]
        tie_tol = 1e-3 * float(np.abs(sc).max())
        for prec, r in res.items():
            lab = np.asarray(r["labels"])[:, 0]
            diff = lab != ref["Ypred_labels"][:, 0]
            st = sin_theta(r["R"], ref["R"])
            errors[prec] = {
                "sin_theta_R": st,
                "B_rel_err": float(np.linalg.norm(r["B"] - ref["B"]) / np.linalg.norm(ref["B"])),
                "label_agreement": float(1.0 - diff.mean()),
                "label_mismatch_non_tie": int((diff & (margin > tie_tol)).sum()),
                "accuracy": float(np.mean(lab == data[3])),
            }
            ok[prec] = st <= SIN_THETA_BOUND and errors[prec]["label_mismatch_non_tie"] == 0
        cpu = cpu_baseline_obj(
            tim, info, 1.0, "fits/s",
            "full FP64 fit + predict of the same workload (numpy/LAPACK oracle); median of 5 "
            "runs after one warm-up, at all the process's cores and at 1 thread",
            extra={"oracle_accuracy": float(np.mean(ref["Ypred_labels"][:, 0] == data[3]))})
    chosen = "bf16x3" if ok["bf16x3"] else "fp32"
    r = res[chosen]
    kname = "project_bf16x3" if chosen == "bf16x3" else "project_f32"
    roof = None
    if kname in r["kernels"]:
        kt = r["kernels"][kname]
        launches = kt["launches_per_step"]
        avg_s = kt["ms_per_step"] * 1e-3 / launches
        # arithmetic intensity vs the ridge point of the form's useful-FLOP peak
        ai = PROJ_FLOPS / PROJ_BYTES
        ridge = USEFUL_PEAK_TFLOPS[chosen] * 1e12 / (HBM_PEAK_GBS * 1e9)
        if ai < ridge:
            bound, unit, peak = "hbm", "GB/s", HBM_PEAK_GBS
            achieved = PROJ_BYTES / launches / avg_s / 1e9
        else:
            bound, unit, peak = "mfma", "TFLOP/s", PEAK_TFLOPS[chosen]
            achieved = PROJ_FLOPS / launches / avg_s / 1e12
        roof = {"bound": bound, "kernel": kname, "achieved": achieved, "peak": peak,
                "unit": unit, "frac": achieved / peak,
                "traffic": measured_traffic(kname, launches),
                "arith_intensity": ai, "ridge": ridge,
                "tflops": PROJ_FLOPS / launches / avg_s / 1e12,
                "avg_launch_us": avg_s * 1e6, "launches_per_step": launches,
                "algorithmic_flops_per_launch": PROJ_FLOPS / launches,
                "algorithmic_bytes_per_launch": PROJ_BYTES / launches,
                "hbm_GBps": PROJ_BYTES / launches / avg_s / 1e9}
    out = {
        "metric": METRIC, "value": r["fits_per_s"], "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3" if chosen == "bf16x3" else "f32",
        "data": "synthetic (seeded; class means on the radius-3 sphere, N(0,0.5^2) noise)",
        "config": {"workload": "plssvd_lda n=50000 ntest=10000 p=768 C=100 ncomp=99",
                   "chosen_form": chosen, "inputs": "resident in HBM"},
        "plain_fp32_fits_per_s": res["fp32"]["fits_per_s"],
        "chosen_vs_plain": r["fits_per_s"] / res["fp32"]["fits_per_s"],
        "roofline": roof,
        "oracle_errors": errors,
        "kernels": {p: v["kernels"] for p, v in res.items()},
        "autotune": res[chosen].get("autotune"),
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
