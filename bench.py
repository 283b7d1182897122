#!/usr/bin/env python3
"""Benchmark: plain-FP64 BDF integration of the 2^20-cell, 32-species
mass-action batch (48 reversible reactions, rtol 1e-6, atol 1e-10,
t in [0, 1]) as one coupled system on one MI355X, through the HIP N_Vector,
block-diagonal matrix and batched block LU.

A "step" is one full integration t = 0 -> 1 of the whole batch, with the
problem, vectors, matrices and integrator already resident in HBM (each step
resets y to y0 on the device).  value = cell-steps/s = cells * nst / time.

Also reported:
  roofline     the dominant kernel family (largest share of GPU time over the
               timed steps, measured with HIP events on its launch stream):
               algorithmic bytes per launch / average launch duration, vs the
               8 TB/s HBM peak;
  microbench   N_Vector LinearSum, WrmsNorm, DotProdMulti(8) on 2^28 FP64;
  cpu_baseline the CPU oracle (serial vector + block-dense LU + the same
               integrator) on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
M_SPECIES, R_REACTIONS, SEED = 32, 48, 1002
RTOL, ATOL, TOUT = 1e-6, 1e-10, 1.0

# algorithmic bytes per launch of each timed kernel family
FAMILIES = {0: "getrs", 1: "getrf", 2: "linearsum", 3: "wrmsnorm", 4: "dotprodmulti",
            5: "rhs"}


def family_bytes(fam, ncells, m, R):
    n = ncells * m
    if fam == 0:   # read LU 8m^2 + pivots 4m + b 8m, write x 8m per block
        return ncells * (8 * m * m + 4 * m + 16 * m)
    if fam == 1:   # read J 8m^2, write LU 8m^2 + pivots 4m per block
        return ncells * (16 * m * m + 4 * m)
    if fam == 2:
        return 24 * n
    if fam == 3:
        return 16 * n
    if fam == 5:   # read y 8m + rates 16R, write f 8m per cell
        return ncells * (16 * m + 16 * R)
    return None


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(core, cells):
    from sundials_amd import _lib

    orc = _lib.load_oracle()
    p = core.SUNAMD_MassActionCreate(M_SPECIES, R_REACTIONS, cells, SEED, 0.2)
    st = np.zeros(16)
    rc = orc.oracle_massaction_solve_coupled(p, RTOL, ATOL, TOUT, None,
                                             st.ctypes.data_as(C.POINTER(C.c_double)))
    core.SUNAMD_MassActionFree(p)
    if rc != 0:
        return None
    nst, wall = st[0], st[15]
    return {"value": cells * nst / wall, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": "coupled serial oracle on the first %d cells of the same generator "
                      "(nst=%d, %.1f s, %s, nproc=%d)" % (cells, nst, wall, cpu_model(),
                                                           os.cpu_count() or 0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--lu", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--micro-n", type=int, default=1 << 28)
    ap.add_argument("--cpu-cells", type=int, default=1024)
    ap.add_argument("--no-micro", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(cells are independent; see DESIGN.md for how the batch would shard)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)

    from sundials_amd import _lib

    core, hip = _lib.core(), _lib.hip()
    m, R, ncells = M_SPECIES, R_REACTIONS, args.cells
    print("[bench] generating %d cells x %d species" % (ncells, m), file=sys.stderr, flush=True)
    prob = core.SUNAMD_MassActionCreate(m, R, ncells, SEED, 0.2)
    run = hip.SUNAMD_CoupledCreate_Mi355(prob, RTOL, ATOL, 0 if args.lu == "fp64" else 1, 0)
    if not run:
        print("[bench] device setup failed", file=sys.stderr)
        sys.exit(1)
    st = _lib.SolveStats()
    for w in range(args.warmup):
        rc = hip.SUNAMD_CoupledRun_Mi355(run, TOUT, C.byref(st))
        print("[bench] warmup %d: rc=%d nst=%d nni=%d %.3f s" % (w, rc, st.nst, st.nni,
                                                                   st.wall_s),
              file=sys.stderr, flush=True)
        if rc != 0:
            sys.exit(1)

    hip.SUNAMD_ProfilerReset()
    hip.SUNAMD_ProfilerEnable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nst_total = nni_total = 0
    stats = []
    for k in range(args.steps):
        rc = hip.SUNAMD_CoupledRun_Mi355(run, TOUT, C.byref(st))
        if rc != 0:
            print("[bench] run failed rc=%d" % rc, file=sys.stderr)
            sys.exit(1)
        nst_total += st.nst
        nni_total += st.nni
        stats.append(st.as_dict())
        print("[bench] step %d: nst=%d nni=%d %.3f s" % (k, st.nst, st.nni, st.wall_s),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hip.SUNAMD_ProfilerEnable(0)
    elapsed = t1 - t0

    kernels = {}
    for fam, name in FAMILIES.items():
        tot, cnt = C.c_double(), C.c_long()
        hip.SUNAMD_ProfilerGet(fam, C.byref(tot), C.byref(cnt))
        if cnt.value:
            kernels[name] = {"total_ms": tot.value, "launches": cnt.value,
                             "avg_us": 1e3 * tot.value / cnt.value}
    dom = max((f for f in FAMILIES if FAMILIES[f] in kernels and family_bytes(f, ncells, m, R)),
              key=lambda f: kernels[FAMILIES[f]]["total_ms"])
    dk = kernels[FAMILIES[dom]]
    bytes_per = family_bytes(dom, ncells, m, R)
    achieved = bytes_per / (dk["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": FAMILIES[dom], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": bytes_per,
                "avg_launch_us": round(dk["avg_us"], 2)}

    micro = None
    if not args.no_micro:
        ms = np.zeros(5)
        if hip.SUNAMD_BenchVectorOps(args.micro_n, 8, 20, ms.ctypes.data_as(
                C.POINTER(C.c_double))) == 0:
            n = args.micro_n
            micro = {"n": n,
                     "linearsum_gbs": round(24 * n / (ms[0] * 1e-3) / 1e9, 1),
                     "wrmsnorm_gbs": round(16 * n / (ms[1] * 1e-3) / 1e9, 1),
                     "dotprodmulti8_gbs": round(72 * n / (ms[2] * 1e-3) / 1e9, 1),
                     "axpy_inplace_gbs": round(24 * n / (ms[3] * 1e-3) / 1e9, 1),
                     "dotprod_gbs": round(16 * n / (ms[4] * 1e-3) / 1e9, 1),
                     "ms": [round(v, 4) for v in ms]}

    hip.SUNAMD_CoupledFree_Mi355(run)
    core.SUNAMD_MassActionFree(prob)

    cpu = None if args.no_cpu else cpu_baseline(core, args.cpu_cells)

    value = ncells * nst_total / elapsed
    line = {
        "metric": "cell-steps/s (plain FP64 coupled BDF, Newton iters/step reported)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.lu == "fp64" else "f64+f32lu",
        "data": "synthetic",
        "config": {"workload": "mass-action batch, %d cells x %d species, %d reactions, "
                               "coupled BDF, rtol %g atol %g, t in [0,%g], %s block LU"
                               % (ncells, m, R, RTOL, ATOL, TOUT, args.lu),
                   "cells": ncells, "species": m, "reactions": R},
        "newton_iters_per_step": nni_total / max(nst_total, 1),
        "nst": stats[-1]["nst"] if stats else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "microbench": micro,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
