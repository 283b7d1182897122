"""What replacing the reference's SUNRpowerR (libm pow, sundials_math.h:235)
with the correctly rounded sunamd_pow changes in the CVODE restatement.

Test infrastructure, CPU only.  Runs the same oracle workloads with two builds
of the CVODE restatement and its oracle -- the standard one
(sundials_amd/lib + oracle/_build) and the one built by `make -C oracle
libm-pow` (oracle/_build/libm_pow, SUNRpowerR = pow) -- each in its own
process (both libraries carry the same soname), and prints per workload how
many counters and final states differ.

    make -C oracle libm-pow && python oracle/pow_study.py [--json out.json]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILDS = {
    "correctly_rounded": (os.path.join(ROOT, "sundials_amd", "lib", "libsunamd_core.so"),
                          os.path.join(ROOT, "oracle", "_build", "liboracle.so")),
    "libm": (os.path.join(ROOT, "oracle", "_build", "libm_pow", "libsunamd_core.so"),
             os.path.join(ROOT, "oracle", "_build", "libm_pow", "liboracle.so")),
}
M_VALUES = (4, 8, 16, 24, 32, 48, 64)
WEIGHTS = (30.0, 25.0, 15.0, 10.0, 10.0, 5.0, 5.0)
COUNTERS = ("nst", "nfe", "nsetups", "netf", "nni", "nnf", "ncfn", "nje", "qlast")


def _run_one(build, out_path, nthreads):
    sys.path.insert(0, ROOT)
    from sundials_amd._lib import Batch, MassAction  # struct layouts only; loads nothing

    core_path, orc_path = BUILDS[build]
    core = C.CDLL(core_path, mode=C.RTLD_GLOBAL)
    orc = C.CDLL(orc_path, mode=C.RTLD_GLOBAL)
    PD, P = C.POINTER(C.c_double), C.POINTER
    core.SUNAMD_MassActionCreate.restype = P(MassAction)
    core.SUNAMD_MassActionCreate.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_uint64,
                                             C.c_double]
    core.SUNAMD_MassActionFree.argtypes = [P(MassAction)]
    core.SUNAMD_BatchCreate.restype = P(Batch)
    core.SUNAMD_BatchCreate.argtypes = [C.c_int64, C.c_uint64, C.c_int, P(C.c_int), PD,
                                        C.c_double, C.c_double]
    core.SUNAMD_BatchFree.argtypes = [P(Batch)]
    orc.oracle_roberts_run.argtypes = [C.c_int, PD]
    orc.oracle_massaction_solve_coupled.argtypes = [P(MassAction), C.c_double, C.c_double,
                                                    C.c_double, PD, PD]
    orc.oracle_massaction_solve_per_cell.argtypes = [P(MassAction), C.c_double, C.c_double,
                                                     C.c_double, C.c_int, PD, PD, PD]
    orc.oracle_batch_solve.argtypes = [P(Batch), C.c_double, C.c_double, C.c_double, C.c_int,
                                       PD, P(C.c_long), PD]
    res = {}

    out = np.zeros(20 * 12)
    assert orc.oracle_roberts_run(12, out.ctypes.data_as(PD)) == 0
    o = out.reshape(12, 20)
    # nst netf ncfn nfe nni nnf nsetups nje per output, then y1..y3
    res["cvRoberts_dns"] = {"counters": o[:, [5, 6, 7, 13, 14, 15, 16, 17]].astype(int).tolist(),
                            "y": o[:, 1:4].tolist()}

    # configuration 1 as one coupled system and as independent cells
    p = core.SUNAMD_MassActionCreate(8, 16, 4096, 1001, 0.2)
    y = np.zeros(4096 * 8)
    st = np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, 1e-6, 1e-10, 1.0, y.ctypes.data_as(PD),
                                               st.ctypes.data_as(PD)) == 0
    res["config1_coupled"] = {"counters": [[int(st[i]) for i in (0, 1, 2, 3, 4, 5, 6, 7, 8)]],
                              "y": [y.tolist()]}
    y = np.zeros(4096 * 8)
    st = np.zeros((4096, 16))
    assert orc.oracle_massaction_solve_per_cell(p, 1e-6, 1e-10, 1.0, nthreads,
                                                y.ctypes.data_as(PD), st.ctypes.data_as(PD),
                                                None) == 0
    res["config1_per_cell"] = {"counters": st[:, :9].astype(int).tolist(),
                               "y": y.reshape(4096, 8).tolist()}
    core.SUNAMD_MassActionFree(p)

    # uneven batches: the 700-system parity batch and the benchmark's CPU sample
    for name, nsys, seed in (("batch700_seed4004", 700, 4004),
                             ("batch32768_seed4242", 32768, 4242)):
        b = core.SUNAMD_BatchCreate(nsys, seed, 7, (C.c_int * 7)(*M_VALUES),
                                    (C.c_double * 7)(*WEIGHTS), 1.5, 0.2)
        bc = b.contents
        y = np.zeros(bc.ytotal)
        cnt = np.zeros((bc.nsys, 10), dtype=np.int64)
        assert orc.oracle_batch_solve(b, 1e-6, 1e-10, 1.0, nthreads, y.ctypes.data_as(PD),
                                      cnt.ctypes.data_as(P(C.c_long)), None) == 0
        ys = []
        for k in range(bc.nsys):
            m = bc.bucket[bc.sys_bucket[k]].contents.m
            ys.append(y[bc.sys_yoff[k]:bc.sys_yoff[k] + m].tolist())
        res[name] = {"counters": cnt[:, :9].tolist(), "y": ys}
        core.SUNAMD_BatchFree(b)
    with open(out_path, "w") as f:
        json.dump(res, f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", choices=sorted(BUILDS))
    ap.add_argument("--out")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--json", help="write the comparison here")
    a = ap.parse_args()
    if a.child:
        _run_one(a.child, a.out, a.threads)
        return
    for build, (core_path, orc_path) in BUILDS.items():
        if not (os.path.exists(core_path) and os.path.exists(orc_path)):
            sys.exit("missing %s build: %s / %s" % (build, core_path, orc_path))
    runs = {}
    for build in BUILDS:
        path = os.path.join("/tmp", "pow_study_%s_%d.json" % (build, os.getpid()))
        subprocess.run([sys.executable, __file__, "--child", build, "--out", path,
                        "--threads", str(a.threads)], check=True)
        with open(path) as f:
            runs[build] = json.load(f)
        os.unlink(path)
    cr, lm = runs["correctly_rounded"], runs["libm"]
    table = {}
    print("%-22s %8s %14s %14s %12s %10s %s" % ("workload", "systems", "counters_diff",
                                                  "states_diff", "max_rel_y", "max_wrms",
                                                  "sum of counters"))
    for name in cr:
        a_c, b_c = np.array(cr[name]["counters"]), np.array(lm[name]["counters"])
        ndiff = int(np.sum(np.any(a_c != b_c, axis=1)))
        ystates, rel, wrms = 0, 0.0, 0.0
        for ya, yb in zip(cr[name]["y"], lm[name]["y"]):
            ya, yb = np.array(ya), np.array(yb)
            ystates += int(np.any(ya != yb))
            rel = max(rel, float(np.max(np.abs(ya - yb) / np.maximum(np.abs(ya), 1e-300))))
            if name != "cvRoberts_dns":  # rtol 1e-6, atol 1e-10 weighted RMS difference
                w = 1.0 / (1e-6 * np.abs(ya) + 1e-10)
                wrms = max(wrms, float(np.sqrt(np.mean(((ya - yb) * w) ** 2))))
        if name == "cvRoberts_dns":
            # per output: nst netf ncfn nfe nni nnf nsetups nje
            sums = {"final": {"correctly_rounded": a_c[-1].tolist(), "libm": b_c[-1].tolist()}}
        else:
            sums = {c: [int(a_c[:, i].sum()), int(b_c[:, i].sum())]
                    for i, c in enumerate(COUNTERS)}
        table[name] = {"systems": int(a_c.shape[0]), "systems_with_counter_change": ndiff,
                       "states_changed": ystates, "max_rel_state_change": rel,
                       "max_weighted_rms_change": wrms, "sums": sums}
        print("%-22s %8d %14d %14d %12.3g %10.3g %s" % (name, a_c.shape[0], ndiff, ystates, rel,
                                                        wrms, json.dumps(sums)))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
