"""Helpers for the ARKStep IMEX tests: the advection-diffusion-reaction
problem (include/sunamd/imex_adr.h) integrated on the CPU oracle or the
MI355X path, and the Dahlquist driver of the oracle."""
import ctypes as C

import numpy as np

from sundials_amd import _lib

SEED = 1003
MODIFIED_GS, CLASSICAL_GS = 1, 2
PD = C.POINTER(C.c_double)


def make_problem(n, seed=SEED):
    p = _lib.core().SUNAMD_AdrCreate(n, seed)
    assert p, "SUNAMD_AdrCreate failed"
    return p


def options(rtol=1e-5, atol=1e-9, tf=0.1, maxl=20, gstype=MODIFIED_GS, controller=0,
            sequential=1, prec_fp32=0, mxsteps=100000):
    return _lib.AdrOptions(rtol, atol, tf, maxl, gstype, controller, sequential, prec_fp32,
                           mxsteps)


def _run(fn, p, o):
    st = np.zeros(_lib.ADR_STATS)
    y = np.zeros(3 * p.contents.ncells)
    rc = fn(p, C.byref(o), st.ctypes.data_as(PD), y.ctypes.data_as(PD))
    return rc, dict(zip(_lib.ADR_STAT_NAMES, st)), y


def run_oracle(p, o):
    return _run(_lib.load_oracle().oracle_adr_solve, p, o)


def run_device(p, o):
    return _run(_lib.hip().SUNAMD_AdrSolve_Mi355, p, o)


def dahlquist(lam_e, lam_i, h, tf, rtol=1e-10, atol=1e-14, controller=None, flags=0, order=3):
    st = np.zeros(_lib.ADR_STATS)
    y = np.zeros(1)
    rc = _lib.load_oracle().oracle_ark_dahlquist(
        lam_e, lam_i, h, tf, rtol, atol, controller.encode() if controller else None, flags,
        order, st.ctypes.data_as(PD), y.ctypes.data_as(PD))
    return rc, dict(zip(_lib.ADR_STAT_NAMES, st)), y[0]


def tables(order=3):
    core = _lib.core()
    out = {}
    for part, name in ((0, "E"), (1, "I")):
        A, b, c, d = np.zeros(49), np.zeros(7), np.zeros(7), np.zeros(7)
        ns, q, pp = C.c_int(), C.c_int(), C.c_int()
        assert core.SUNAMD_ARKTable(order, part, C.byref(ns), A.ctypes.data_as(PD),
                                    b.ctypes.data_as(PD), c.ctypes.data_as(PD),
                                    d.ctypes.data_as(PD), C.byref(q), C.byref(pp)) == 0
        s = ns.value
        out[name] = (A[:s * s].reshape(s, s), b[:s], c[:s], d[:s], q.value, pp.value)
    return out


def kpr(ncalls=3):
    """The reference's ARKStep logging test on the KPR problem (IMEX,
    Newton, SPGMR, difference-quotient Jv, ARK_ONE_STEP) run on the oracle:
    rows (t, u, v) and the statistics named as ARKodePrintAllStats prints
    them."""
    out, st = np.zeros(3 * (ncalls + 1)), np.zeros(18)
    rc = _lib.load_oracle().oracle_ark_kpr(ncalls, out.ctypes.data_as(PD), st.ctypes.data_as(PD))
    names = ["Steps", "Step attempts", "Error test fails", "NLS step fails",
             "Initial step size", "Last step size", "Current step size",
             "Explicit RHS fn evals", "Implicit RHS fn evals", "NLS iters", "NLS fails",
             "LS setups", "LS RHS fn evals", "LS iters", "LS fails", "Jac-times evals",
             "Prec setup evals", "Prec solves"]
    return rc, out.reshape(ncalls + 1, 3), dict(zip(names, st))
