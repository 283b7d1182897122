#!/usr/bin/env python3
"""Benchmarks on one MI355X.

Default workload (the north-star metric): an uneven batch of 65,536
independent stiff mass-action systems, m drawn from {4, 8, 16, 24, 32, 48, 64}
with weights 30/25/15/10/10/5/5 %, R = 1.5 m reversible reactions, rtol 1e-6,
atol 1e-10, t in [0, 1].  Every system is integrated by its own BDF /
modified-Newton / dense-LU integrator with its own step size and order, all
systems in one persistent launch per size bucket.  A "step" is one
integration t = 0 -> 1 of the whole batch with the problem resident in HBM.
value = cell-steps/s = (sum over systems of nst) / time, for the product
form (FP32 Newton-matrix LU, FP64 residuals / updates / norms); the plain
FP64 form is timed beside it, with Newton iterations per step for both and
the FP32 form's final-state error against the FP64 run (which equals the
CPU oracle bit for bit).

  roofline     the longest-running bucket kernel: algorithmic flops per
               launch (counted from the per-system counters) / its launch
               duration (HIP events on its stream), vs the FP64 peak;
  cpu_baseline the CPU oracle (serial vector + dense LU + the same
               integrator, one CVode run per system, OpenMP over systems) on
               the first systems of the same batch.

--workload coupled keeps the earlier benchmark: the 2^20-cell, 32-species
batch as one coupled system through the HIP N_Vector, block-diagonal matrix
and batched block LU, plus the N_Vector microbench at N = 2^28.
--workload imex: ARKStep IMEX + SPGMR + block-Jacobi on the periodic
128^3 x 3 advection-diffusion-reaction problem (the Newton-Krylov path).
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
FP64_PEAK_TFS = 78.6
M_SPECIES, R_REACTIONS, SEED = 32, 48, 1002
RTOL, ATOL, TOUT = 1e-6, 1e-10, 1.0
TIGHT_RTOL, TIGHT_ATOL = 1e-10, 1e-14   # reference solution for accuracy reports

# uneven batch
BATCH_M = (4, 8, 16, 24, 32, 48, 64)
BATCH_W = (30.0, 25.0, 15.0, 10.0, 10.0, 5.0, 5.0)
BATCH_SEED = 1004
TRAIN_SEED = 2004      # independent batch whose plain runs give the training pairs
CALIB_SEED = 3004      # independent batch for the per-bucket on/off choice
BATCH_MXSTEP = 100000
LU_FP64, LU_FP32 = 0, 1

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


def vector_cpu_baseline(n):
    """The oracle's serial vector ops and their OpenMP form on the host cores,
    at the microbench's length, one timed call each after a warm-up."""
    from sundials_amd import _lib

    orc = _lib.load_oracle()
    out = {}
    for threads in (1, min(16, os.cpu_count() or 1)):
        ms = np.zeros(5)
        if orc.oracle_bench_vector_ops(n, 8, 1, threads,
                                       ms.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            return None
        gbs = [24 * n, 16 * n, 72 * n, 24 * n, 16 * n] / (ms * 1e-3) / 1e9
        out["threads_%d" % threads] = {
            "linearsum_gbs": round(gbs[0], 2), "wrmsnorm_gbs": round(gbs[1], 2),
            "dotprodmulti8_gbs": round(gbs[2], 2), "axpy_inplace_gbs": round(gbs[3], 2),
            "dotprod_gbs": round(gbs[4], 2), "ms": [round(v, 2) for v in ms]}
    out["sample"] = "oracle vector ops on the host, N = %d, %s, nproc=%d" % (
        n, cpu_model(), os.cpu_count() or 0)
    return out


def vector_microbench(args, hip):
    """The N_Vector microbench at N = --micro-n on the device (HIP events), with
    copy ceilings measured in the same process and the host baseline."""
    ms = np.zeros(8)
    if hip.SUNAMD_BenchVectorOps(args.micro_n, 8, 20,
                                 ms.ctypes.data_as(C.POINTER(C.c_double))) != 0:
        return None
    n = args.micro_n
    names = ("linearsum", "wrmsnorm", "dotprodmulti8", "axpy_inplace", "dotprod")
    gbs = np.array([24, 16, 72, 24, 16, 16, 16, 16], dtype=np.float64) * n / (ms * 1e-3) / 1e9
    ceiling = float(max(gbs[5], gbs[6], gbs[7]))
    micro = {"n": n}
    for i, k in enumerate(names):
        micro[k + "_gbs"] = round(float(gbs[i]), 1)
    micro["copy_gbs"] = {"chunked_256": round(float(gbs[5]), 1),
                         "chunked_1024": round(float(gbs[6]), 1),
                         "hipMemcpy": round(float(gbs[7]), 1)}
    micro["frac_of_hbm_peak"] = {k: round(float(gbs[i]) / HBM_PEAK_GBS, 3)
                                 for i, k in enumerate(names)}
    micro["frac_of_copy_ceiling"] = {k: round(float(gbs[i]) / ceiling, 3)
                                     for i, k in enumerate(names)}
    micro["ms"] = [round(float(v), 4) for v in ms]
    if not args.no_cpu:
        micro["cpu_baseline"] = vector_cpu_baseline(n)
    return micro


def bench_blocklu(args):
    """--workload blocklu: the block LU factorisation of 2^20 blocks of 32 x 32
    and 2^18 of 64 x 64 (8 GiB of FP64 blocks each) by the FP64 and FP32
    register-row kernels and the FP32 blocked kernel with MFMA trailing
    updates.  value = the fastest FP32 variant's GB/s at m = 32."""
    from sundials_amd import _lib

    hip = _lib.hip()
    out = {}
    for m, nblocks in ((32, 1 << 20), (64, 1 << 18)):
        row = {}
        for variant, name in ((0, "fp64_register"), (1, "fp32_register"), (2, "fp32_mfma")):
            ms = C.c_double()
            if hip.SUNAMD_BenchBlockLU(nblocks, m, variant, 10, C.byref(ms)) != 0:
                print("[bench] LU benchmark failed", file=sys.stderr)
                sys.exit(1)
            wbytes = 8 if variant == 0 else 4
            traffic = nblocks * (8 * m * m + wbytes * m * m + 4 * m)
            row[name] = {"ms": round(ms.value, 4),
                         "gbs": round(traffic / (ms.value * 1e-3) / 1e9, 1),
                         "blocks_per_s": nblocks / (ms.value * 1e-3)}
            print("[bench] LU m=%d %s: %.3f ms, %.0f GB/s" % (m, name, ms.value, row[name]["gbs"]),
                  file=sys.stderr, flush=True)
        out["m%d" % m] = dict(row, blocks=nblocks)
    best = max(out["m32"]["fp32_register"]["gbs"], out["m32"]["fp32_mfma"]["gbs"])
    return {"metric": "block LU factorisation GB/s (algorithmic bytes), FP32 factors",
            "value": best, "unit": "GB/s", "n_gpus": 1, "steps": 10, "warmup": 1,
            "ms_per_step": min(out["m32"]["fp32_register"]["ms"], out["m32"]["fp32_mfma"]["ms"]),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "batched block LU, 2^20 blocks of 32 x 32 and 2^18 of "
                                   "64 x 64, U(-1,1)/m + I"},
            "roofline": {"bound": "hbm", "achieved": best, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(best / HBM_PEAK_GBS, 4), "traffic": None},
            "cpu_baseline": None, "results": out}


def bench_vector(args):
    """--workload vector: the N_Vector microbench alone (LinearSum, WrmsNorm,
    DotProdMulti(8), in-place LinearSum, DotProd at N = 2^28 FP64) against the
    copy ceiling and HBM peak; value = LinearSum GB/s."""
    from sundials_amd import _lib

    micro = vector_microbench(args, _lib.hip())
    if micro is None:
        print("[bench] microbench failed", file=sys.stderr)
        sys.exit(1)
    ms = micro["ms"][0]
    return {"metric": "N_Vector op HBM GB/s (LinearSum z = a x + b y, algorithmic bytes)",
            "value": micro["linearsum_gbs"], "unit": "GB/s", "n_gpus": 1, "steps": 20,
            "warmup": 1, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "N_Vector microbench, N = %d FP64, x, y ~ U(-1,1), "
                                   "w ~ U(1e-3,1), DotProdMulti over 8 vectors" % args.micro_n},
            "roofline": {"bound": "hbm", "achieved": micro["linearsum_gbs"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(micro["linearsum_gbs"] / HBM_PEAK_GBS, 4),
                         "traffic": 24.0 * args.micro_n},
            "cpu_baseline": micro.get("cpu_baseline"),
            "microbench": micro}


def bench_coupled(args):
    import torch
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

    # the same integration with the integrator's fused kernels (error weights,
    # residual, Nordsieck predictor / undo, and the Newton-matrix setup as one
    # pass: copy + ScaleAddI + factorisation): bit-identical results, fewer passes
    fused = None
    y_plain = np.zeros(ncells * m)
    hip.SUNAMD_CoupledGetState_Mi355(run, y_plain.ctypes.data_as(C.POINTER(C.c_double)))
    plain_counts = (st.nst, st.nni, st.nsetups, st.nje)
    if hip.SUNAMD_CoupledSetFusedKernels_Mi355(run, 1) == 0:
        rc = hip.SUNAMD_CoupledRun_Mi355(run, TOUT, C.byref(st))
        if rc == 0:
            y_fused = np.zeros(ncells * m)
            hip.SUNAMD_CoupledGetState_Mi355(run, y_fused.ctypes.data_as(C.POINTER(C.c_double)))
            fused = {"ms": round(1e3 * st.wall_s, 3), "nst": st.nst, "nni": st.nni,
                     "value": ncells * st.nst / st.wall_s,
                     "bitwise_equal_to_unfused": bool(
                         np.array_equal(y_plain, y_fused)
                         and plain_counts == (st.nst, st.nni, st.nsetups, st.nje))}
            del y_fused
        hip.SUNAMD_CoupledSetFusedKernels_Mi355(run, 0)
    del y_plain

    # the same integration device-stepped (one cooperative launch, every
    # integrator decision on the device), tree-ordered norms; and with FP32
    # factors, device-stepped and host-stepped
    device = None
    if hip.SUNAMD_CoupledSetDeviceStepping_Mi355(run, 1) == 0:
        rc = hip.SUNAMD_CoupledRun_Mi355(run, TOUT, C.byref(st))
        if rc == 0:
            shp = [C.c_int(), C.c_int(), C.c_int()]
            hip.SUNAMD_CoupledDeviceLastShape(C.byref(shp[0]), C.byref(shp[1]), C.byref(shp[2]))
            device = {"fp64": {"ms": round(1e3 * st.wall_s, 3), "nst": st.nst, "nni": st.nni,
                               "value": ncells * st.nst / st.wall_s},
                      "launch": {"workgroups": shp[0].value, "threads": shp[1].value,
                                 "lds_bytes": shp[2].value}}
        hip.SUNAMD_CoupledSetDeviceStepping_Mi355(run, 0)
    if args.lu == "fp64":
        run32 = hip.SUNAMD_CoupledCreate_Mi355(prob, RTOL, ATOL, 1, 0)
        if run32:
            hip.SUNAMD_CoupledSetFusedKernels_Mi355(run32, 1)
            # one untimed run first: the FP32 factor buffers are allocated and
            # first touched there, as the FP64 forms' were in their warm-up
            if (hip.SUNAMD_CoupledRun_Mi355(run32, TOUT, C.byref(st)) == 0
                    and hip.SUNAMD_CoupledRun_Mi355(run32, TOUT, C.byref(st)) == 0
                    and device is not None):
                device["fp32_lu_host_stepped_fused"] = {
                    "ms": round(1e3 * st.wall_s, 3), "nst": st.nst, "nni": st.nni,
                    "value": ncells * st.nst / st.wall_s}
            hip.SUNAMD_CoupledSetFusedKernels_Mi355(run32, 0)
            if hip.SUNAMD_CoupledSetDeviceStepping_Mi355(run32, 1) == 0 and device is not None:
                if hip.SUNAMD_CoupledRun_Mi355(run32, TOUT, C.byref(st)) == 0:
                    device["fp32_lu"] = {"ms": round(1e3 * st.wall_s, 3), "nst": st.nst,
                                         "nni": st.nni, "value": ncells * st.nst / st.wall_s}
            hip.SUNAMD_CoupledFree_Mi355(run32)

    micro = None if args.no_micro else vector_microbench(args, hip)

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
        "fused_kernels": fused,
        "device_stepping": device,
        "microbench": micro,
    }
    return line


def batch_flops(cnt, msys, rsys):
    """Algorithmic flops of a batched integration from per-system counters:
    RHS 9R per evaluation, Jacobian 20R per evaluation, Newton matrix m^2+m and
    LU (2/3)m^3 - m^2/2 - m/6 per setup, GETRS 2m^2 plus 10m of vector work per
    Newton iteration, 30m of vector work per step."""
    m = msys.astype(np.float64)
    R = rsys.astype(np.float64)
    nst, nfe, nset, nni, nje = (cnt[:, k].astype(np.float64) for k in (0, 1, 2, 4, 7))
    lu = (2.0 / 3.0) * m ** 3 - 0.5 * m ** 2 - m / 6.0
    return (9 * R * nfe + 20 * R * nje + (m * m + m + lu) * nset + (2 * m * m + 10 * m) * nni
            + 30 * m * nst)


def _oracle_batch(nsys, threads, rtol, atol):
    """The CPU oracle on the first nsys systems of the measured batch (the
    same systems, in the same order and layout, as its first nsys)."""
    from sundials_amd import _lib

    core, orc = _lib.core(), _lib.load_oracle()
    n = len(BATCH_M)
    b = core.SUNAMD_BatchCreate(nsys, BATCH_SEED, n, (C.c_int * n)(*BATCH_M),
                                (C.c_double * n)(*BATCH_W), 1.5, 0.2)
    bc = b.contents
    y = np.zeros(bc.ytotal)
    cnt = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
    wall = C.c_double()
    rc = orc.oracle_batch_solve(b, rtol, atol, TOUT, threads,
                                y.ctypes.data_as(C.POINTER(C.c_double)),
                                cnt.ctypes.data_as(C.POINTER(C.c_long)), C.byref(wall))
    core.SUNAMD_BatchFree(b)
    return rc, y, cnt, wall.value


def batch_cpu_baseline(nsys_sample, threads):
    rc, y, cnt, wall = _oracle_batch(nsys_sample, threads, RTOL, ATOL)
    if rc != 0:
        return None, None, None
    nst = int(cnt[:, 0].sum())
    return ({"value": nst / wall, "unit": "cell-steps/s", "cores": threads, "kind": "port",
             "sample": "CPU oracle (serial vector + dense LU, one CVode run per system, OpenMP "
                       "over systems) on the first %d of the 65,536 systems (%d cell-steps, "
                       "%.1f s, %s, nproc=%d)" % (nsys_sample, nst, wall, cpu_model(),
                                                  os.cpu_count() or 0)}, y, cnt)


def _wrms_per_system(y, y_ref, yoff, msys):
    """Per-system WRMS of y - y_ref with the run's weights 1/(rtol|y_ref| + atol)."""
    w = 1.0 / (RTOL * np.abs(y_ref) + ATOL)
    d2 = ((y - y_ref) * w) ** 2
    csum = np.concatenate(([0.0], np.cumsum(d2)))
    return np.sqrt((csum[yoff + msys] - csum[yoff]) / msys)


def _err_summary(e):
    return {"median": round(float(np.median(e)), 4), "p99": round(float(np.percentile(e, 99)), 4),
            "max": round(float(e.max()), 4)}


def _batch(core, nsys, seed):
    n = len(BATCH_M)
    return core.SUNAMD_BatchCreate(nsys, seed, n, (C.c_int * n)(*BATCH_M),
                                   (C.c_double * n)(*BATCH_W), 1.5, 0.2)


def train_predictor(args, core, hip):
    """Training pairs from plain FP64 runs of an independent batch (one row per
    step, up to --train-pairs rows), a sample of them round-tripped through the
    versioned training-pair log file, then Adam on the GPU with the library's
    MFMA kernels over one pass of the pairs.  Returns (network, info)."""
    import tempfile

    import torch
    from sundials_amd import _lib

    t0 = time.perf_counter()
    bt = _batch(core, args.train_systems, TRAIN_SEED)
    cap = args.train_pairs
    x = torch.zeros((cap, 144), dtype=torch.float32, device="cuda")
    t = torch.zeros((cap, 64), dtype=torch.float32, device="cuda")
    m = torch.zeros((cap, 64), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    B = hip.SUNAMD_BatchedBDFCreate(bt, RTOL, ATOL, LU_FP64)
    hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, BATCH_MXSTEP)
    hip.SUNAMD_BatchedBDFSetTrainingLog(B, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                        C.c_void_p(m.data_ptr()), cap, 1)
    wall = C.c_double()
    if hip.SUNAMD_BatchedBDFRun(B, TOUT, C.byref(wall)) != 0:
        raise RuntimeError("logging run failed")
    n = int(hip.SUNAMD_BatchedBDFGetTrainingLogCount(B))
    hip.SUNAMD_BatchedBDFFree(B)
    core.SUNAMD_BatchFree(bt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the first rows through the on-disk format and back, compared exactly
    ns = min(n, 100000)
    hx, ht, hm = (a[:ns].cpu().numpy() for a in (x, t, m))
    info = _lib.TrainingLogInfo(1, ns, TRAIN_SEED, RTOL, ATOL, 1)
    rx, rt, rm = np.zeros_like(hx), np.zeros_like(ht), np.zeros_like(hm)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "pairs.sunamdtp").encode()
        ok = (core.SUNAMD_TrainingLogWrite(path, hx.ctypes.data, ht.ctypes.data, hm.ctypes.data,
                                           ns, C.byref(info)) == 0 and
              core.SUNAMD_TrainingLogRead(path, 0, ns, rx.ctypes.data, rt.ctypes.data,
                                          rm.ctypes.data) == 0)
        fsize = os.path.getsize(path) if ok else 0
    roundtrip = bool(ok and np.array_equal(hx, rx) and np.array_equal(ht, rt)
                     and np.array_equal(hm, rm))
    if not roundtrip:
        raise RuntimeError("training-pair log round trip failed")
    perm = torch.randperm(n, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    x, t, m = x[:n][perm].contiguous(), t[:n][perm].contiguous(), m[:n][perm].contiguous()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    batch = 1024
    steps = args.train_steps if args.train_steps > 0 else max(1, n // batch)
    M = hip.SUNAMD_MLPCreate(17, batch)
    loss = C.c_float()
    if hip.SUNAMD_MLPTrain(M, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                           C.c_void_p(m.data_ptr()), n, batch, steps, 1e-3, C.byref(loss)) != 0:
        raise RuntimeError("training failed")
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    del x, t, m
    info = {"pairs": n, "train_systems": args.train_systems, "steps": steps, "batch": batch,
            "samples_seen": steps * batch, "loss": float(loss.value),
            "log_s": round(t1 - t0, 2), "train_s": round(t3 - t2, 2),
            "log_file_roundtrip": {"rows": ns, "bytes": fsize, "identical": roundtrip}}
    print("[bench] predictor: %d pairs, %d steps x %d, loss %.4g, train %.1f s"
          % (n, steps, batch, loss.value, t3 - t2), file=sys.stderr, flush=True)
    return M, info


def autotune(core, hip, prec, args_systems, weights=None):
    """Per-bucket launch choices made on an independent calibration batch of
    the measured batch's size (a bucket's choice only shows its effect when
    the bucket fills the GPU as it does in the measured batch):
      1. the kernel's register budget (1, 2, 3 or 4 wavefronts per SIMD:
         occupancy against spilling);
      2. the launch size (2 .. 32 workgroups per CU: how many systems each
         lane group takes from the bucket's queue, i.e. load balance inside
         a wavefront against wavefronts for the GPU to schedule);
      3. given a trained network, whether the learned start pays off.
    Buckets share the GPU, so per-bucket minima do not add up: a per-bucket
    mix is kept only if the whole batch runs faster with it than with the
    best uniform choice.  Returns (waves, caps, learned mask, info)."""
    nb = len(BATCH_M)
    bcal = _batch(core, args_systems, CALIB_SEED)

    def run_cfg(waves, caps, mask):
        B = hip.SUNAMD_BatchedBDFCreate(bcal, RTOL, ATOL, prec)
        hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, BATCH_MXSTEP)
        hip.SUNAMD_BatchedBDFSetOccupancy(B, (C.c_int * nb)(*waves))
        hip.SUNAMD_BatchedBDFSetGridCap(B, (C.c_int * nb)(*caps))
        if weights is not None:
            hip.SUNAMD_BatchedBDFSetLearnedPredictor(B, weights, mask)
        wall = C.c_double()
        for _ in range(2):
            if hip.SUNAMD_BatchedBDFRun(B, TOUT, C.byref(wall)) != 0:
                raise RuntimeError("calibration run failed")
        kms = np.zeros(nb)
        hip.SUNAMD_BatchedBDFGetKernelTimes(B, kms.ctypes.data_as(C.POINTER(C.c_double)))
        hip.SUNAMD_BatchedBDFFree(B)
        return kms, wall.value

    def choose(options, make_cfg):
        runs = {o: run_cfg(*make_cfg([o] * nb)) for o in options}
        per = [min(options, key=lambda o: runs[o][0][k]) for k in range(nb)]
        uni = min(options, key=lambda o: runs[o][1])
        mix_ms, mix_wall = run_cfg(*make_cfg(per))
        if runs[uni][1] <= mix_wall:
            return [uni] * nb, runs[uni], runs, mix_wall
        return per, (mix_ms, mix_wall), runs, mix_wall

    caps0 = [16] * nb
    best_w, _, wruns, wmix = choose((1, 2, 3, 4), lambda w: (w, caps0, 0))
    best_c, (best_ms, best_wall), cruns, cmix = choose((2, 4, 8, 16, 32),
                                                        lambda c: (best_w, c, 0))
    info = {"waves": best_w, "grid_cap_per_cu": best_c,
            "ms_by_waves": {str(w): [round(v, 2) for v in wruns[w][0]] for w in wruns},
            "wall_ms_by_waves": {str(w): round(wruns[w][1] * 1e3, 2) for w in wruns},
            "wall_ms_waves_mix": round(wmix * 1e3, 2),
            "ms_by_grid_cap": {str(c): [round(v, 2) for v in cruns[c][0]] for c in cruns},
            "wall_ms_by_grid_cap": {str(c): round(cruns[c][1] * 1e3, 2) for c in cruns},
            "wall_ms_grid_cap_mix": round(cmix * 1e3, 2),
            "wall_ms_chosen": round(best_wall * 1e3, 2)}
    mask = 0
    if weights is not None:
        lms, _ = run_cfg(best_w, best_c, (1 << nb) - 1)
        for k in range(nb):
            if lms[k] < best_ms[k]:
                mask |= 1 << k
        info["learned_ms"] = [round(v, 2) for v in lms]
    core.SUNAMD_BatchFree(bcal)
    return best_w, best_c, mask, info


IMEX_SEED = 1003
FAM_GETRS, FAM_GETRF, FAM_LINSUM, FAM_WRMS, FAM_ADR_FE, FAM_SCALE = 0, 1, 2, 3, 7, 8
# HBM bytes per launch of the batched integrator's two-wavefront FP32-LU kernels, measured
# with rocprofv3 --pmc (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors a plain 256 MiB
# copy gives); a PMC run cannot happen inside the timed process, so the figures are recorded
# HBM bytes per launch of the streaming legs' dominant kernels at their benchmark sizes
# (6,291,456 unknowns; 2^20 blocks of 6 x 6), measured the same way; equal to the
# algorithmic bytes within 2 percent
MEASURED_STREAM_TRAFFIC = {
    ("kinsol-newton", "N_VLinearSum"): 151.0e6,
    ("kinsol-newton", "N_VScale"): 100.7e6,
    ("kinsol-newton", "k_getrs (6x6 blocks)"): 436.3e6,
    ("kinsol-newton", "k_getrf (6x6 blocks)"): 629.2e6,
    ("kinsol", "N_VLinearSum"): 151.0e6,
    ("kinsol", "N_VScale"): 100.7e6,
    ("imex", "N_VLinearSum"): 151.0e6,
    ("imex", "N_VScale"): 100.7e6,
}


def stream_traffic(workload, kernel, n):
    """Recorded HBM bytes per launch, only at the size they were measured at."""
    return MEASURED_STREAM_TRAFFIC.get((workload, kernel)) if n == 6291456 else None


# the kernel source the MEASURED_TRAFFIC figures were taken at (first 16 hex digits of the
# SHA-256 of bdf_device.h followed by bdf_args.h); a changed kernel reports traffic null
MEASURED_TRAFFIC_SOURCE = "99e45596109fa425"


def batched_kernel_source_hash():
    import hashlib

    h = hashlib.sha256()
    for f in ("bdf_device.h", "bdf_args.h"):
        with open(os.path.join(ROOT, "sundials_amd", "csrc", "batched", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


# HBM bytes per launch of the FP32-LU kernels at three wavefronts per SIMD (8 workgroups
# per CU) and at two (32 workgroups per CU), rocprofv3 --pmc FETCH_SIZE x 2.000 + WRITE_SIZE
# x 1.000 (the factors a plain 256 MiB copy gives) on the 65,536-system batch; at three
# wavefronts almost all of it is register-spill stores, at two it is close to the
# algorithmic traffic (states, rate constants, counters: ~10-20 MB per launch)
MEASURED_TRAFFIC = {
"k_batched_bdf<4, float, 3, false>": 19.3e6,
    "k_batched_bdf<8, float, 3, false>": 34.1e6,
    "k_batched_bdf<16, float, 3, false>": 343.3e6,
    "k_batched_bdf<24, float, 3, false>": 147.0e6,
    "k_batched_bdf<32, float, 3, false>": 1135.8e6,
    "k_batched_bdf<48, float, 3, false>": 3379.8e6,
    "k_batched_bdf<64, float, 3, false>": 6312.7e6,
    "k_batched_bdf<4, float, 2, false>": 8.5e6,
    "k_batched_bdf<8, float, 2, false>": 11.9e6,
    "k_batched_bdf<16, float, 2, false>": 12.6e6,
    "k_batched_bdf<24, float, 2, false>": 11.1e6,
    "k_batched_bdf<32, float, 2, false>": 15.7e6,
    "k_batched_bdf<48, float, 2, false>": 15.1e6,
    "k_batched_bdf<64, float, 2, false>": 16.2e6,
}


def bench_imex(args):
    """ARKStep IMEX (ARK3(2)4L[2]SA) + SPGMR (maxl 20, fused device classical
    Gram-Schmidt) + per-cell block-Jacobi on the advection-diffusion-reaction
    problem, n^3 cells x 3 species; a step is one integration over [0, tf]."""
    import torch
    from sundials_amd import _lib

    core, hip = _lib.core(), _lib.hip()
    PD = C.POINTER(C.c_double)
    n = args.grid
    p = core.SUNAMD_AdrCreate(n, IMEX_SEED)
    ncells = p.contents.ncells
    N = 3 * ncells

    def run(prec32=0):
        o = _lib.AdrOptions(1e-6, 1e-10, args.tf, 20, 2, 0, 0, prec32, 1000000)
        st = np.zeros(_lib.ADR_STATS)
        rc = hip.SUNAMD_AdrSolve_Mi355(p, C.byref(o), st.ctypes.data_as(PD), None)
        if rc != 0:
            print("[bench] imex solve failed rc=%d" % rc, file=sys.stderr)
            sys.exit(1)
        return dict(zip(_lib.ADR_STAT_NAMES, st))

    for w in range(args.warmup):
        st = run()
        print("[bench] warmup %d: %.3f s, %d steps" % (w, st["wall"], st["nst"]), file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    walls, stats = [], None
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats = run()
        walls.append(stats["wall"])
        print("[bench] step %d: %.3f s, %d steps, %d Newton, %d Krylov" % (
            k, stats["wall"], stats["nst"], stats["nni"], stats["nli"]), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the integrations themselves (the solver's own clock, device idle at both
    # ends), as for the other workloads; each call also builds its integrator,
    # vectors and Krylov workspace, which the whole-call time includes
    integ = float(sum(walls))
    value = ncells * stats["nst"] * args.steps / integ

    # FP32 preconditioner factors, timed once
    s32 = run(1)
    # per-family kernel times of one more integration (HIP events on the vector stream)
    hip.SUNAMD_ProfilerReset()
    hip.SUNAMD_ProfilerEnable(1)
    sprof = run()
    hip.SUNAMD_ProfilerEnable(0)
    fams = {FAM_LINSUM: ("N_VLinearSum", 24.0 * N), FAM_WRMS: ("N_VWrmsNorm", 16.0 * N),
            FAM_ADR_FE: ("advection-diffusion RHS (7-point stencil)", 48.0 * ncells),
            FAM_SCALE: ("N_VScale", 16.0 * N)}
    kern = {}
    for f, (name, nbytes) in fams.items():
        tot, cnt = C.c_double(), C.c_long()
        hip.SUNAMD_ProfilerGet(f, C.byref(tot), C.byref(cnt))
        if cnt.value:
            kern[f] = {"name": name, "total_ms": tot.value, "launches": cnt.value,
                       "avg_us": 1e3 * tot.value / cnt.value, "bytes_per_launch": nbytes}
    dom = max(kern, key=lambda f: kern[f]["total_ms"])
    dk = kern[dom]
    achieved = dk["bytes_per_launch"] / (dk["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dk["name"], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": stream_traffic("imex", dk["name"], N),
                "bytes_per_launch": dk["bytes_per_launch"],
                "avg_launch_us": round(dk["avg_us"], 2)}

    cpu = None
    if not args.no_cpu:
        orc = _lib.load_oracle()
        nc = args.cpu_grid
        pc = core.SUNAMD_AdrCreate(nc, IMEX_SEED)
        o = _lib.AdrOptions(1e-6, 1e-10, args.tf, 20, 2, 0, 1, 0, 1000000)
        st = np.zeros(_lib.ADR_STATS)
        t = time.perf_counter()
        rc = orc.oracle_adr_solve(pc, C.byref(o), st.ctypes.data_as(PD), None)
        wall = time.perf_counter() - t
        if rc == 0:
            cs = dict(zip(_lib.ADR_STAT_NAMES, st))
            cpu = {"value": pc.contents.ncells * cs["nst"] / wall, "unit": "cell-steps/s",
                   "cores": 1, "kind": "port",
                   "sample": "CPU oracle (serial vector, the same ARKStep + SPGMR + block-Jacobi "
                             "restatement) on the %d^3 grid over the same [0, %g] (%d steps, "
                             "%.1f s, %s)" % (nc, args.tf, cs["nst"], wall, cpu_model())}
        core.SUNAMD_AdrFree(pc)
    core.SUNAMD_AdrFree(p)
    ms = 1e3 * integ / args.steps
    ms_call = 1e3 * elapsed / args.steps
    return {
        "metric": "cell-steps/s (ARKStep IMEX + SPGMR + block-Jacobi, advection-reaction)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "ms_per_call_with_setup": round(ms_call, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "advection-diffusion-reaction on a periodic %d^3 grid x 3 species, "
                               "ARKStep IMEX ARK3(2)4L[2]SA, SPGMR maxl 20 (classical GS), "
                               "block-Jacobi 3x3, rtol 1e-6 atol 1e-10, t in [0, %g]"
                               % (n, args.tf), "cells": ncells},
        "integrator": {k: int(stats[k]) for k in ("nst", "attempts", "netf", "ncfn", "nfe", "nfi",
                                                  "nni", "nli", "nsetups", "npe", "nps")},
        "fp32_preconditioner": {"ms": round(1e3 * s32["wall"], 3), "nst": int(s32["nst"]),
                                "nli": int(s32["nli"]),
                                "value": ncells * s32["nst"] / s32["wall"]},
        "roofline": roofline,
        "kernels": {kern[f]["name"]: {k: v for k, v in kern[f].items() if k != "name"}
                    for f in kern},
        "cpu_baseline": cpu,
    }


def bench_kinsol(args):
    """KINSOL fixed-point iteration with Anderson acceleration (depth 5,
    CGS2 QR updates) on the replicated analytic map (the IEEE-only variant)
    over 2^21 cells x 3 unknowns, perturbed initial guesses, tolerance 1e-11;
    a step is one KINSol call."""
    import torch
    from sundials_amd import _lib

    hip = _lib.hip()
    ncells = args.kin_cells
    N = 3 * ncells

    def opts(nc, seq=0):
        o = _lib.KinOptions()
        o.ncells, o.variant, o.perturb, o.tol, o.maxiter = nc, 1, 1, 1e-11, 100
        o.m_aa, o.delay_aa, o.orth_aa = 5, 0, 2
        o.damping_fp, o.damping_aa, o.damping_fn, o.depth_fn, o.sequential = 1.0, 1.0, 0, 0, seq
        o.fused = 1
        return o

    PD = C.POINTER(C.c_double)

    def run():
        st = np.zeros(_lib.KIN_STATS)
        rc = hip.SUNAMD_KinSolve_Mi355(C.byref(opts(ncells)), st.ctypes.data_as(PD), None)
        if rc != 0:
            print("[bench] kinsol solve failed rc=%d" % rc, file=sys.stderr)
            sys.exit(1)
        return {"nni": int(st[1]), "nfe": int(st[2]), "fnorm": st[3], "wall": st[4]}

    for w in range(args.warmup):
        st = run()
        print("[bench] warmup %d: %.4f s, %d iterations" % (w, st["wall"], st["nni"]),
              file=sys.stderr, flush=True)
    # the timed region of each step is the KINSol call (its own synchronised
    # wall clock); creating the problem and the solver's workspace is not
    torch.cuda.synchronize()
    iters, elapsed = 0, 0.0
    for k in range(args.steps):
        st = run()
        iters += st["nni"]
        elapsed += st["wall"]
        print("[bench] step %d: %.4f s, %d iterations" % (k, st["wall"], st["nni"]),
              file=sys.stderr, flush=True)
    value = N * iters / elapsed

    hip.SUNAMD_ProfilerReset()
    hip.SUNAMD_ProfilerEnable(1)
    run()
    hip.SUNAMD_ProfilerEnable(0)
    fams = {FAM_LINSUM: ("N_VLinearSum", 24.0 * N), FAM_SCALE: ("N_VScale", 16.0 * N)}
    kern = {}
    for f, (name, nbytes) in fams.items():
        tot, cnt = C.c_double(), C.c_long()
        hip.SUNAMD_ProfilerGet(f, C.byref(tot), C.byref(cnt))
        if cnt.value:
            kern[f] = {"name": name, "total_ms": tot.value, "launches": cnt.value,
                       "avg_us": 1e3 * tot.value / cnt.value, "bytes_per_launch": nbytes}
    dom = max(kern, key=lambda f: kern[f]["total_ms"])
    dk = kern[dom]
    achieved = dk["bytes_per_launch"] / (dk["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dk["name"], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": stream_traffic("kinsol", dk["name"], N),
                "bytes_per_launch": dk["bytes_per_launch"],
                "avg_launch_us": round(dk["avg_us"], 2)}

    cpu = None
    if not args.no_cpu:
        orc = _lib.load_oracle()
        nc = args.kin_cpu_cells
        so = np.zeros(_lib.KIN_STATS)
        t = time.perf_counter()
        rc = orc.oracle_kin_solve(C.byref(opts(nc)), so.ctypes.data_as(PD), None)
        wall = time.perf_counter() - t
        if rc == 0:
            cpu = {"value": 3 * nc * so[1] / so[4], "unit": "unknown-iterations/s", "cores": 1,
                   "kind": "port",
                   "sample": "CPU oracle (serial vector, the same KINSOL restatement) on %d cells "
                             "(%d iterations, KINSol %.2f s of %.2f s with set-up, %s)"
                             % (nc, int(so[1]), so[4], wall, cpu_model())}
    ms = 1e3 * elapsed / args.steps
    return {
        "metric": "unknown-iterations/s (KINSOL fixed point + Anderson acceleration)",
        "value": value,
        "unit": "unknown-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "KINSol KIN_FP, Anderson depth 5 (CGS2), fused scaled norm, the analytic 3-equation "
                               "map (IEEE-only variant) on %d independent cells, tol 1e-11"
                               % ncells, "unknowns": N},
        "solver": {"nni": st["nni"], "nfe": st["nfe"], "fnorm": st["fnorm"]},
        "timed": "KINSol calls (problem set-up and workspace allocation excluded)",
        "roofline": roofline,
        "kernels": {kern[f]["name"]: {k: v for k, v in kern[f].items() if k != "name"}
                    for f in kern},
        "cpu_baseline": cpu,
    }


def bench_kinsol_newton(args):
    """KINSOL modified Newton with the line search on the constrained
    Ferraris-Tronconi cell system (6 unknowns per cell, analytic 6 x 6
    Jacobian blocks, batched block LU in FP64) over 2^20 cells, perturbed
    initial guesses in the middle of the box, fnormtol 1e-10; a step is one
    KINSol call."""
    import torch
    from sundials_amd import _lib

    hip = _lib.hip()
    ncells = args.kinn_cells
    N = 6 * ncells
    M = 6

    def opts(nc, fused=1):
        o = _lib.KinNewtonOptions()
        o.ncells, o.variant, o.guess, o.perturb = nc, 1, 2, 1
        o.strategy, o.msbset, o.fnormtol, o.scsteptol = _lib.KIN_LINESEARCH, 0, 1e-10, 1e-12
        o.maxiter, o.jac, o.sequential, o.fused = 0, 1, 0, fused
        return o

    PD = C.POINTER(C.c_double)

    def run(fused=1):
        st = np.zeros(_lib.KINN_STATS)
        rc = hip.SUNAMD_KinNewtonSolve_Mi355(C.byref(opts(ncells, fused)), st.ctypes.data_as(PD),
                                             None)
        if rc != 0:
            print("[bench] kinsol newton solve failed rc=%d" % rc, file=sys.stderr)
            sys.exit(1)
        return dict(zip(_lib.KINN_STAT_NAMES, st))

    for w in range(args.warmup):
        st = run()
        print("[bench] warmup %d: %.4f s, %d iterations" % (w, st["wall"], st["nni"]),
              file=sys.stderr, flush=True)
    # the timed region of each step is the KINSol call (its own synchronised
    # wall clock); creating the problem, matrix and solver is not
    torch.cuda.synchronize()
    iters, elapsed = 0, 0.0
    for k in range(args.steps):
        st = run()
        iters += int(st["nni"])
        elapsed += st["wall"]
        print("[bench] step %d: %.4f s, %d iterations, %d Jacobians"
              % (k, st["wall"], st["nni"], st["nje"]), file=sys.stderr, flush=True)
    value = N * iters / elapsed
    # the same solve with KINSOL's scaled norms as op chains (the reference's
    # sequence) instead of the fused one-pass kernels; identical iterates
    unf = run(fused=0)
    print("[bench] unfused norms: %.4f s, %d iterations" % (unf["wall"], unf["nni"]),
          file=sys.stderr, flush=True)

    hip.SUNAMD_ProfilerReset()
    hip.SUNAMD_ProfilerEnable(1)
    run()
    hip.SUNAMD_ProfilerEnable(0)
    fams = {FAM_GETRF: ("k_getrf (6x6 blocks)", ncells * (2.0 * M * M * 8 + M * 4)),
            FAM_GETRS: ("k_getrs (6x6 blocks)", ncells * (M * M * 8 + M * 4 + 2.0 * M * 8)),
            FAM_LINSUM: ("N_VLinearSum", 24.0 * N), FAM_SCALE: ("N_VScale", 16.0 * N)}
    kern = {}
    for f, (name, nbytes) in fams.items():
        tot, cnt = C.c_double(), C.c_long()
        hip.SUNAMD_ProfilerGet(f, C.byref(tot), C.byref(cnt))
        if cnt.value:
            kern[f] = {"name": name, "total_ms": tot.value, "launches": cnt.value,
                       "avg_us": 1e3 * tot.value / cnt.value, "bytes_per_launch": nbytes}
    dom = max(kern, key=lambda f: kern[f]["total_ms"])
    dk = kern[dom]
    achieved = dk["bytes_per_launch"] / (dk["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dk["name"], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": stream_traffic("kinsol-newton", dk["name"], N),
                "bytes_per_launch": dk["bytes_per_launch"],
                "avg_launch_us": round(dk["avg_us"], 2)}

    cpu = None
    if not args.no_cpu:
        orc = _lib.load_oracle()
        nc = args.kinn_cpu_cells
        so = np.zeros(_lib.KINN_STATS)
        t = time.perf_counter()
        rc = orc.oracle_kin_newton_solve(C.byref(opts(nc)), so.ctypes.data_as(PD), None)
        wall = time.perf_counter() - t
        if rc == 0:
            cpu = {"value": 6 * nc * so[1] / so[7], "unit": "unknown-iterations/s", "cores": 1,
                   "kind": "port",
                   "sample": "CPU oracle (serial vector, per-block LU, the same KINSOL "
                             "restatement) on %d cells (%d iterations, KINSol %.2f s of %.2f s "
                             "with set-up, %s)" % (nc, int(so[1]), so[7], wall, cpu_model())}
    ms = 1e3 * elapsed / args.steps
    return {
        "metric": "unknown-iterations/s (KINSOL Newton + line search, block LU)",
        "value": value,
        "unit": "unknown-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "KINSol KIN_LINESEARCH, modified Newton (setup every 10 "
                               "iterations), fused scaled norms, constrained Ferraris-Tronconi cells (IEEE-only "
                               "variant) x %d, analytic 6x6 Jacobian blocks, FP64 block LU, "
                               "fnormtol 1e-10" % ncells, "unknowns": N},
        "solver": {k: int(st[k]) for k in ("nni", "nfe", "nje", "nbcf", "nbacktrack")},
        "timed": "KINSol calls (problem, matrix and solver set-up excluded)",
        "fused_kernels": {"on_ms": ms, "off_ms": 1e3 * unf["wall"],
                          "what": "KINSetUseFusedKernels: KINScSNorm / KINScFNorm in one "
                                  "pass each (bit-identical)"},
        "roofline": roofline,
        "kernels": {kern[f]["name"]: {k: v for k, v in kern[f].items() if k != "name"}
                    for f in kern},
        "cpu_baseline": cpu,
    }


def bench_batched(args):
    import torch
    from sundials_amd import _lib

    core, hip = _lib.core(), _lib.hip()
    n = len(BATCH_M)
    print("[bench] generating %d systems" % args.systems, file=sys.stderr, flush=True)
    b = core.SUNAMD_BatchCreate(args.systems, BATCH_SEED, n, (C.c_int * n)(*BATCH_M),
                                (C.c_double * n)(*BATCH_W), 1.5, 0.2)
    bc = b.contents
    nb = bc.nbuckets
    bm = np.array([bc.bucket[k].contents.m for k in range(nb)])
    bR = np.array([bc.bucket[k].contents.R for k in range(nb)])
    sb = np.ctypeslib.as_array(bc.sys_bucket, shape=(bc.nsys,)).copy()
    msys, rsys = bm[sb], bR[sb]

    def make(prec, waves, caps):
        B = hip.SUNAMD_BatchedBDFCreate(b, RTOL, ATOL, prec)
        if not B:
            print("[bench] device setup failed", file=sys.stderr)
            sys.exit(1)
        hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, BATCH_MXSTEP)
        hip.SUNAMD_BatchedBDFSetOccupancy(B, (C.c_int * nb)(*waves))
        hip.SUNAMD_BatchedBDFSetGridCap(B, (C.c_int * nb)(*caps))
        return B

    def results(B):
        y = np.zeros(bc.ytotal)
        cnt = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
        rc = hip.SUNAMD_BatchedBDFGetResults(B, y.ctypes.data_as(C.POINTER(C.c_double)),
                                             cnt.ctypes.data_as(C.POINTER(C.c_long)))
        if rc != 0:
            print("[bench] some systems failed", file=sys.stderr)
            sys.exit(1)
        return y, cnt

    def run(B, tag):
        wall = C.c_double()
        rc = hip.SUNAMD_BatchedBDFRun(B, TOUT, C.byref(wall))
        if rc != 0:
            print("[bench] %s run failed rc=%d" % (tag, rc), file=sys.stderr)
            sys.exit(1)
        return wall.value

    # plain FP64 form (register budget tuned per bucket): warm up, then time one integration
    fixed = None if args.waves == "auto" else [int(args.waves)] * nb
    fixed_cap = [args.grid_cap] * nb if args.grid_cap > 0 else [16] * nb
    if fixed is None:
        w64, c64cap, _, tune64 = autotune(core, hip, LU_FP64, args.systems)
    else:
        w64, c64cap, tune64 = fixed, fixed_cap, {"waves": fixed, "grid_cap_per_cu": fixed_cap,
                                                 "fixed": True}
    print("[bench] fp64 waves per bucket: %s, workgroups per CU %s" % (w64, c64cap),
          file=sys.stderr, flush=True)
    B64 = make(LU_FP64, w64, c64cap)
    run(B64, "fp64 warmup")
    t64 = run(B64, "fp64")
    y64, c64 = results(B64)
    hip.SUNAMD_BatchedBDFFree(B64)
    nst64 = int(c64[:, 0].sum())
    print("[bench] plain fp64: %.3f s, %d cell-steps" % (t64, nst64), file=sys.stderr, flush=True)

    # product form: FP32 Newton-matrix LU + learned initial iterate where it pays
    learned = None
    mlp = None
    mask = 0
    weights = None
    if not args.no_learned:
        mlp, learned = train_predictor(args, core, hip)
        weights = hip.SUNAMD_MLPDeviceWeights(mlp)
    if fixed is None:
        w32, c32cap, mask, tune32 = autotune(core, hip, LU_FP32, args.systems, weights)
    else:
        w32, c32cap, mask, tune32 = fixed, fixed_cap, 0, {"waves": fixed,
                                                          "grid_cap_per_cu": fixed_cap,
                                                          "fixed": True}
    print("[bench] fp32 waves per bucket: %s, workgroups per CU %s, learned mask %d"
          % (w32, c32cap, mask), file=sys.stderr, flush=True)
    if learned is not None:
        learned["bucket_mask"] = mask
        learned["buckets_on"] = [int(BATCH_M[k]) for k in range(len(BATCH_M)) if mask >> k & 1]
    B = make(LU_FP32, w32, c32cap)
    if mlp is not None:
        hip.SUNAMD_BatchedBDFSetLearnedPredictor(B, hip.SUNAMD_MLPDeviceWeights(mlp), mask)
    for w in range(args.warmup):
        tw = run(B, "warmup")
        print("[bench] warmup %d: %.3f s" % (w, tw), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    walls = []
    for k in range(args.steps):
        walls.append(run(B, "step %d" % k))
        print("[bench] step %d: %.3f s" % (k, walls[-1]), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    y32, c32 = results(B)
    kms = np.zeros(nb)
    hip.SUNAMD_BatchedBDFGetKernelTimes(B, kms.ctypes.data_as(C.POINTER(C.c_double)))
    if learned is not None:
        tried, kept = C.c_long(), C.c_long()
        hip.SUNAMD_BatchedBDFGetLearnedStats(B, C.byref(tried), C.byref(kept))
        learned["steps_tried"], learned["guesses_kept"] = tried.value, kept.value
    hip.SUNAMD_BatchedBDFFree(B)
    nst32 = int(c32[:, 0].sum())
    # the learned start in every bucket of the measured batch (whatever the
    # autotuner kept), beside the timed product form: per bucket kernel time,
    # Newton iterations per step and guesses kept
    learned_all, y_la = None, None
    if mlp is not None:
        BL = make(LU_FP32, w32, c32cap)
        hip.SUNAMD_BatchedBDFSetLearnedPredictor(BL, hip.SUNAMD_MLPDeviceWeights(mlp),
                                                 (1 << nb) - 1)
        run(BL, "learned warmup")
        t_la = run(BL, "learned, every bucket")
        y_la, c_la = results(BL)
        kms_la = np.zeros(nb)
        hip.SUNAMD_BatchedBDFGetKernelTimes(BL, kms_la.ctypes.data_as(C.POINTER(C.c_double)))
        tr = np.zeros(nb, dtype=np.int64)
        kp = np.zeros(nb, dtype=np.int64)
        hip.SUNAMD_BatchedBDFGetLearnedStatsByBucket(BL, tr.ctypes.data_as(C.POINTER(C.c_long)),
                                                     kp.ctypes.data_as(C.POINTER(C.c_long)))
        hip.SUNAMD_BatchedBDFFree(BL)
        per = {}
        for k in range(nb):
            sel = sb == k
            per["m%d" % bm[k]] = {
                "kernel_ms": round(float(kms_la[k]), 3),
                "kernel_ms_without": round(float(kms[k]), 3),
                "newton_iters_per_step": round(float(c_la[sel, 4].sum() / c_la[sel, 0].sum()), 4),
                "newton_iters_per_step_without": round(float(c32[sel, 4].sum() / c32[sel, 0].sum()),
                                                       4),
                "steps": int(c_la[sel, 0].sum()), "steps_without": int(c32[sel, 0].sum()),
                "guesses_tried": int(tr[k]), "guesses_kept": int(kp[k])}
        nst_la = int(c_la[:, 0].sum())
        learned_all = {"ms": round(1e3 * t_la, 3), "value": nst_la / t_la,
                       "newton_iters_per_step": float(c_la[:, 4].sum() / max(nst_la, 1)),
                       "speedup_vs_plain_fp64": (nst_la / t_la) / (nst64 / t64),
                       "buckets": per}
        print("[bench] learned start in every bucket: %.3f s, %.4f Newton iters/step"
              % (t_la, learned_all["newton_iters_per_step"]), file=sys.stderr, flush=True)
        hip.SUNAMD_MLPFree(mlp)

    # roofline: the longest bucket kernel of the last timed run
    fl = batch_flops(c32, msys, rsys)
    kernels = {}
    for k in range(nb):
        sel = sb == k
        kernels["m%d" % bm[k]] = {"systems": int(sel.sum()), "ms": round(float(kms[k]), 3),
                                  "gflop": round(float(fl[sel].sum()) / 1e9, 3)}
    dom = int(np.argmax(kms))
    dflop = float(fl[sb == dom].sum())
    achieved = dflop / (kms[dom] * 1e-3) / 1e12
    learn_dom = bool(mask & (1 << dom)) or w32[dom] >= 4
    kname = "k_batched_bdf<%d, float, %d, %s>" % (bm[dom], w32[dom],
                                                 "true" if learn_dom else "false")
    src = batched_kernel_source_hash()
    traffic_ok = args.systems == 65536 and src == MEASURED_TRAFFIC_SOURCE
    roofline = {"bound": "valu", "kernel": kname,
                "achieved": round(achieved, 4), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFS, 5),
                "traffic": MEASURED_TRAFFIC.get(kname) if traffic_ok else None,
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this benchmark "
                                  "at kernel source %s, scaled by a 256 MiB copy (x2.000 / "
                                  "x1.000); null when the kernel source (now %s) or the batch "
                                  "size differs" % (MEASURED_TRAFFIC_SOURCE, src),
                "flops_per_launch": dflop, "launch_ms": round(float(kms[dom]), 3),
                "note": "VALU issue- and latency-bound persistent integrator (no MFMA: per-lane "
                        "FP64 scalar control logic, register-row LU); algorithmic flops priced "
                        "against the FP64 vector peak"}

    # accuracy of the FP32-LU form against the FP64 form (= CPU oracle bit for bit)
    yoff = np.ctypeslib.as_array(bc.sys_yoff, shape=(bc.nsys,)).copy()
    werr = _wrms_per_system(y32, y64, yoff, msys)
    # relative error of components above the absolute tolerance
    relmax = float(np.max(np.abs(y32 - y64) / np.maximum(np.abs(y64), ATOL)))

    # the CPU oracle on the first cpu_systems systems: the timed CPU baseline,
    # the FP64 leg's parity (every counter and every final value bit for bit),
    # and a reference solution at 10^4x tighter tolerances for the accuracy of
    # both forms
    cpu, parity, accuracy = None, None, None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        ns = min(args.cpu_systems, args.systems)
        cpu, ycpu, ccpu = batch_cpu_baseline(ns, threads)
        if cpu is None:
            print("[bench] CPU oracle failed", file=sys.stderr)
            sys.exit(1)
        nyc = ycpu.size
        same_c = bool(np.array_equal(c64[:ns], ccpu))
        same_y = bool(np.array_equal(y64[:nyc], ycpu))
        parity = {"systems": ns, "counters_equal": same_c, "final_states_bitwise_equal": same_y,
                  "oracle": "CPU oracle (serial vector + dense LU), rtol %g atol %g" % (RTOL, ATOL)}
        print("[bench] fp64 leg vs CPU oracle on %d systems: counters %s, states %s"
              % (ns, "equal" if same_c else "DIFFER", "bitwise equal" if same_y else "DIFFER"),
              file=sys.stderr, flush=True)
        if not (same_c and same_y):
            print("[bench] the plain FP64 leg does not reproduce the oracle", file=sys.stderr)
            sys.exit(1)
        rc_t, yt, ct, wall_t = _oracle_batch(ns, threads, TIGHT_RTOL, TIGHT_ATOL)
        if rc_t == 0:
            yo, ms_ = yoff[:ns], msys[:ns]
            e64 = _wrms_per_system(y64[:nyc], yt, yo, ms_)
            e32 = _wrms_per_system(y32[:nyc], yt, yo, ms_)
            if learned_all is not None:
                learned_all["error_vs_tight_reference"] = _err_summary(
                    _wrms_per_system(y_la[:nyc], yt, yo, ms_))
            worst = np.argsort(-np.maximum(e64, e32))[:5]
            accuracy = {
                "reference": "CPU oracle at rtol %g atol %g on the same %d systems (%.1f s)"
                             % (TIGHT_RTOL, TIGHT_ATOL, ns, wall_t),
                "weights": "1/(%g |y_ref| + %g), per-system WRMS" % (RTOL, ATOL),
                "plain_fp64": _err_summary(e64), "fp32_lu": _err_summary(e32),
                "fp32_lu_vs_fp64_on_sample": _err_summary(werr[:ns]),
                "worst_systems": [{"system": int(s_), "m": int(msys[s_]),
                                   "fp64": round(float(e64[s_]), 3),
                                   "fp32_lu": round(float(e32[s_]), 3),
                                   "fp32_lu_vs_fp64": round(float(werr[s_]), 3)}
                                  for s_ in worst]}
            print("[bench] error vs tight reference: fp64 %s, fp32-lu %s"
                  % (accuracy["plain_fp64"], accuracy["fp32_lu"]), file=sys.stderr, flush=True)
    core.SUNAMD_BatchFree(b)

    value = nst32 * args.steps / elapsed
    v64 = nst64 / t64
    line = {
        "metric": "cell-steps/s & Newton iters/step, learned+FP32-LU vs plain FP64, same tol",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32lu",
        "data": "synthetic",
        "config": {"workload": "uneven batch: %d independent mass-action systems, m in %s "
                               "(weights %s %%), R = 1.5 m, per-system BDF, rtol %g atol %g, "
                               "t in [0,%g], FP32 Newton-matrix LU"
                               % (args.systems, list(BATCH_M), [int(x) for x in BATCH_W],
                                  RTOL, ATOL, TOUT),
                   "systems": args.systems,
                   "predictor": "learned where calibrated faster, else polynomial"
                                if learned is not None else "polynomial"},
        "newton_iters_per_step": c32[:, 4].sum() / max(nst32, 1),
        "cell_steps": nst32,
        "plain_fp64": {"value": v64, "ms": 1e3 * t64, "cell_steps": nst64,
                       "newton_iters_per_step": c64[:, 4].sum() / max(nst64, 1)},
        "speedup_vs_plain_fp64": value / v64,
        "error_vs_fp64": {"max_weighted_rms": float(werr.max()),
                          "median_weighted_rms": float(np.median(werr)),
                          "max_rel": relmax},
        "fp64_parity_vs_cpu_oracle": parity,
        "error_vs_tight_reference": accuracy,
        "learned": learned,
        "learned_every_bucket": learned_all,
        "autotune": {"fp64": tune64, "fp32": tune32},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    return line


PHASE_NAMES = ("acquire", "guard", "predict", "newton", "rhs", "jac", "getrf", "getrs", "norm",
               "errtest", "complete", "prepare", "pow", "learn", "output", "idle")


def bench_phases(args):
    """Where a step's time goes: the uneven batch integrated by the phase-clock
    build of the batched integrator (two wavefronts per SIMD, learned start
    off), FP64 and FP32-LU forms.  Per bucket: shader-clock ticks per step
    spent in each phase by the group that owns a system (for m > 32 the group
    is the whole wavefront), summed over the bucket's groups and divided by
    its steps."""
    from sundials_amd import _lib

    core, hp = _lib.core(), _lib.hip_phase()
    b = _batch(core, args.systems, BATCH_SEED)
    bc = b.contents
    nb = bc.nbuckets
    bm = [bc.bucket[k].contents.m for k in range(nb)]
    sb = np.ctypeslib.as_array(bc.sys_bucket, shape=(bc.nsys,)).copy()
    out = {}
    for prec, tag in ((LU_FP64, "fp64"), (LU_FP32, "fp32_lu")):
        B = hp.SUNAMD_BatchedBDFCreate(b, RTOL, ATOL, prec)
        hp.SUNAMD_BatchedBDFSetMaxNumSteps(B, BATCH_MXSTEP)
        hp.SUNAMD_BatchedBDFSetOccupancy(B, (C.c_int * nb)(*([2] * nb)))
        wall = C.c_double()
        for _ in range(2):
            if hp.SUNAMD_BatchedBDFRun(B, TOUT, C.byref(wall)) != 0:
                raise RuntimeError("phase run failed")
        NV = 2 * 16 + 2
        raw = np.zeros(nb * NV, dtype=np.uint64)
        if hp.SUNAMD_BatchedBDFGetPhaseTicks(
                B, raw.ctypes.data_as(C.POINTER(C.c_ulonglong))) != nb:
            raise RuntimeError("phase ticks unavailable")
        raw = raw.reshape(nb, NV).astype(np.float64)
        kms = np.zeros(nb)
        hp.SUNAMD_BatchedBDFGetKernelTimes(B, kms.ctypes.data_as(C.POINTER(C.c_double)))
        y = np.zeros(bc.ytotal)
        cnt = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
        hp.SUNAMD_BatchedBDFGetResults(B, y.ctypes.data_as(C.POINTER(C.c_double)),
                                       cnt.ctypes.data_as(C.POINTER(C.c_long)))
        hp.SUNAMD_BatchedBDFFree(B)
        buckets = {}
        for k in range(nb):
            sel = sb == k
            nst = float(cnt[sel, 0].sum())
            ticks, entries = raw[k, :16], raw[k, 16:32]
            mark = raw[k, 32] / max(raw[k, 33], 1.0)   # ticks of one phase change
            # each entry's own clock read lands in the phase entered: take it off
            corrected = np.maximum(ticks - entries * mark, 0.0)
            per_step = corrected / max(nst, 1.0)
            buckets["m%d" % bm[k]] = {
                "systems": int(sel.sum()), "steps": int(nst), "kernel_ms": round(float(kms[k]), 2),
                "newton_iters_per_step": round(float(cnt[sel, 4].sum()) / max(nst, 1.0), 4),
                "setups_per_step": round(float(cnt[sel, 2].sum()) / max(nst, 1.0), 4),
                "ticks_per_phase_change": round(float(mark), 1),
                "phase_changes_per_step": round(float(entries.sum()) / max(nst, 1.0), 2),
                "ticks_per_step": {PHASE_NAMES[p]: round(float(per_step[p]), 1)
                                   for p in range(16)},
                "raw_ticks_per_step": round(float(ticks.sum()) / max(nst, 1.0), 1),
                "total_ticks_per_step": round(float(per_step.sum()), 1)}
            print("[phases] %s m=%2d: %7.0f ticks/step (raw %.0f, %.1f changes x %.0f)  %s" % (
                tag, bm[k], per_step.sum(), ticks.sum() / max(nst, 1.0),
                entries.sum() / max(nst, 1.0), mark,
                " ".join("%s %.0f" % (PHASE_NAMES[p], per_step[p]) for p in range(16)
                         if per_step[p] >= 0.02 * per_step.sum())), file=sys.stderr, flush=True)
        out[tag] = {"wall_ms": round(wall.value * 1e3, 2), "buckets": buckets}
    core.SUNAMD_BatchFree(b)
    return {"metric": "phase clock per step, batched BDF (profiling aid)", "value": None,
            "unit": "shader-clock ticks per step", "n_gpus": 1, "data": "synthetic",
            "config": {"workload": "uneven batch: %d independent mass-action systems, m in %s, "
                                   "two wavefronts per SIMD, learned start off"
                                   % (args.systems, list(BATCH_M))},
            "clock": "s_memtime (shader clock) read at each phase change; ticks of one lane "
                     "group per step of its system, summed over a bucket's groups / its steps, "
                     "less the measured cost of one phase change per entry",
            "results": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["batched", "coupled", "imex", "kinsol", "kinsol-newton",
                             "phases", "vector", "blocklu"],
                    default="batched")
    ap.add_argument("--kin-cells", type=int, default=1 << 21, help="kinsol: cells (3 unknowns)")
    ap.add_argument("--kin-cpu-cells", type=int, default=1 << 21, help="kinsol: CPU-oracle cells")
    ap.add_argument("--kinn-cells", type=int, default=1 << 20,
                    help="kinsol-newton: cells (6 unknowns)")
    ap.add_argument("--kinn-cpu-cells", type=int, default=1 << 20,
                    help="kinsol-newton: CPU-oracle cells")
    ap.add_argument("--grid", type=int, default=128, help="imex: cells per dimension")
    ap.add_argument("--tf", type=float, default=0.01, help="imex: final time")
    ap.add_argument("--cpu-grid", type=int, default=64, help="imex: CPU-oracle grid")
    ap.add_argument("--systems", type=int, default=65536)
    ap.add_argument("--cpu-systems", type=int, default=32768)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--lu", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--micro-n", type=int, default=1 << 28)
    ap.add_argument("--cpu-cells", type=int, default=1024)
    ap.add_argument("--no-micro", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--waves", choices=["auto", "1", "2", "3", "4"], default="auto",
                    help="wavefronts per SIMD for every bucket (auto: calibrated per bucket; "
                         "fixed values also keep the learned start off)")
    ap.add_argument("--grid-cap", type=int, default=0,
                    help="with fixed --waves: workgroups per CU for every bucket (0: 16)")
    ap.add_argument("--train-systems", type=int, default=32768,
                    help="systems of the independent batch whose plain runs are logged")
    ap.add_argument("--train-pairs", type=int, default=10000000,
                    help="training pairs kept (one per logged step)")
    ap.add_argument("--train-steps", type=int, default=0,
                    help="Adam steps of 1024 pairs (0: one pass over the pairs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the systems are "
              "independent, so a batch would shard by system ranges with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.workload == "batched":
        line = bench_batched(args)
    elif args.workload == "imex":
        line = bench_imex(args)
    elif args.workload == "kinsol":
        line = bench_kinsol(args)
    elif args.workload == "kinsol-newton":
        line = bench_kinsol_newton(args)
    elif args.workload == "phases":
        line = bench_phases(args)
    elif args.workload == "vector":
        line = bench_vector(args)
    elif args.workload == "blocklu":
        line = bench_blocklu(args)
    else:
        line = bench_coupled(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
