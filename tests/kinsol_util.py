"""Helpers for the KINSOL fixed-point tests: options from a published run of
the reference example, and the oracle / device solves."""
import ctypes as C

import numpy as np

from sundials_amd import _lib


def options(run=None, **kw):
    o = _lib.KinOptions()
    o.ncells, o.variant, o.perturb = 1, 0, 0
    o.tol, o.maxiter = 100 * np.sqrt(np.finfo(float).eps), 30
    o.m_aa, o.delay_aa, o.orth_aa = 0, 0, 0
    o.damping_fp, o.damping_aa, o.damping_fn, o.depth_fn, o.sequential = 1.0, 1.0, 0, 0, 0
    if run is not None:
        for k in ("maxiter", "m_aa", "delay_aa", "orth_aa", "damping_fp", "damping_aa"):
            setattr(o, k, run[k])
        o.damping_fn, o.depth_fn = int(run["damping_fn"]), int(run["depth_fn"])
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _solve(fn, o):
    st = np.zeros(_lib.KIN_STATS)
    u = np.zeros(3 * o.ncells)
    rc = fn(C.byref(o), st.ctypes.data_as(C.POINTER(C.c_double)),
            u.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, {"flag": int(st[0]), "nni": int(st[1]), "nfe": int(st[2]), "fnorm": st[3],
                "wall": st[4]}, u


def run_oracle(o):
    return _solve(_lib.load_oracle().oracle_kin_solve, o)


def run_device(o):
    return _solve(_lib.hip().SUNAMD_KinSolve_Mi355, o)
