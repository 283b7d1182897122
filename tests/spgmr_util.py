"""The reference's SPGMR unit-test scenarios
(test/unit_tests/sunlinsol/spgmr/serial/test_sunlinsol_spgmr_serial.c): a
tridiag(-1, 5, -1) system of size 100 with a random solution in [1, 2],
solved with maxl = 100 at tol 1e-13 in six set-ups -- no scaling, row
scaling s1, column scaling s2 (entries 1 + 1000 U), each without and with
the Jacobi preconditioner -- from a zero and from a non-zero initial guess.
Runs on either vector implementation: the serial oracle vector with host
callbacks, or the MI355X vector with the device callbacks."""
import ctypes as C

import numpy as np

from sundials_amd import _lib
from helpers import cpu_vec, cpu_get, gpu_vec, gpu_get

N, MAXL, TOL = 100, 100, 1e-13
PREC_NONE, PREC_LEFT, PREC_RIGHT = 0, 1, 2
MODIFIED_GS, CLASSICAL_GS = 1, 2


def _fnptr(f):
    return C.cast(f, C.c_void_p)


class SerialBackend:
    def __init__(self, ctx):
        self.orc = _lib.load_oracle()
        self.ctx = ctx
        self.atimes = _fnptr(self.orc.oracle_tridiag_atimes)
        self.psetup = _fnptr(self.orc.oracle_tridiag_psetup)
        self.psolve = _fnptr(self.orc.oracle_tridiag_psolve)

    def vec(self, vals):
        return cpu_vec(self.orc, self.ctx, vals)

    def get(self, v):
        return cpu_get(v)

    def apply(self, data, x, b):
        assert self.orc.oracle_tridiag_atimes(C.byref(data), x, b) == 0


class DeviceBackend:
    def __init__(self, ctx, order):
        self.hip = _lib.hip()
        self.ctx = ctx
        self.order = order
        self.atimes = _fnptr(self.hip.SUNAMD_TridiagATimes_Mi355)
        self.psetup = _fnptr(self.hip.SUNAMD_TridiagPSetup_Mi355)
        self.psolve = _fnptr(self.hip.SUNAMD_TridiagPSolve_Mi355)

    def vec(self, vals):
        v = gpu_vec(self.ctx, vals)
        assert self.hip.N_VSetReductionOrder_Mi355(v, self.order) == 0
        return v

    def get(self, v):
        return gpu_get(v)

    def apply(self, data, x, b):
        assert self.hip.SUNAMD_TridiagATimes_Mi355(C.byref(data), x, b) == 0


def scenario_scalings(n, seed):
    rng = np.random.default_rng(seed)
    one = np.ones(n)
    s = [1.0 + 1000.0 * rng.random(n) for _ in range(4)]
    # (s1, s2, preconditioned)
    return [(one, one, False), (one, one, True), (s[0], one, False), (s[1], one, True),
            (one, s[2], False), (one, s[3], True)]


def run_scenarios(be, gstype, pretype, n=N, maxl=MAXL, tol=TOL, seed=11, max_restarts=0,
                  cgs_fn=None):
    """Returns one record per solve: (scenario, zeroguess, flag, iters, resnorm,
    solution, exact)."""
    core = _lib.core()
    rng = np.random.default_rng(seed)
    xhat = 1.0 + rng.random(n)
    xv = be.vec(np.zeros(n))
    bv = be.vec(np.zeros(n))
    yv = be.vec(np.zeros(n))
    dv = be.vec(np.full(n, 5.0))
    s1v = be.vec(np.ones(n))
    s2v = be.vec(np.ones(n))
    data = _lib.TridiagData(n, dv, s1v, s2v)
    LS = core.SUNLinSol_SPGMR(xv, pretype, maxl, be.ctx)
    assert LS
    out = []
    try:
        assert core.SUNLinSolSetATimes(LS, C.byref(data), be.atimes) == 0
        assert core.SUNLinSolSetPreconditioner(LS, C.byref(data), be.psetup, be.psolve) == 0
        assert core.SUNLinSolSetScalingVectors(LS, s1v, s2v) == 0
        assert core.SUNLinSolSetZeroGuess(LS, 0) == 0
        assert core.SUNLinSolInitialize(LS) == 0
        assert core.SUNLinSol_SPGMRSetGSType(LS, gstype) == 0
        assert core.SUNLinSol_SPGMRSetMaxRestarts(LS, max_restarts) == 0
        if cgs_fn is not None:
            assert core.SUNLinSol_SPGMRSetClassicalGSFn(LS, cgs_fn) == 0
        for si, (s1, s2, prec) in enumerate(scenario_scalings(n, seed + 1)):
            for v, vals in ((s1v, s1), (s2v, s2), (xv, xhat / s2)):
                _overwrite(be, v, vals)
            be.apply(data, xv, bv)
            assert core.SUNLinSol_SPGMRSetPrecType(LS, pretype if prec else PREC_NONE) == 0
            assert core.SUNLinSolSetup(LS, None) == 0
            for zg in (True, False):
                if zg:
                    core.N_VConst(0.0, yv)
                else:
                    core.N_VAddConst(xv, np.sqrt(np.finfo(float).eps), yv)
                assert core.SUNLinSolSetZeroGuess(LS, 1 if zg else 0) == 0
                flag = core.SUNLinSolSolve(LS, None, yv, bv, tol)
                out.append((si, zg, flag, core.SUNLinSolNumIters(LS), core.SUNLinSolResNorm(LS),
                            be.get(yv), be.get(xv)))
    finally:
        core.SUNLinSolFree(LS)
        for v in (xv, bv, yv, dv, s1v, s2v):
            core.N_VDestroy(v)
    return out


def _overwrite(be, v, vals):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    if isinstance(be, SerialBackend):
        C.memmove(_lib.core().N_VGetArrayPointer(v), vals.ctypes.data, 8 * len(vals))
    else:
        assert be.hip.N_VCopyFromHostArray_Mi355(
            v, vals.ctypes.data_as(C.POINTER(C.c_double))) == 0


def solution_ok(x, y, tol):
    """check_vector of the reference test: every entry within SUNRCompareTol."""
    core = _lib.core()
    return all(core.SUNRCompareTol(float(a), float(b), tol) == 0 for a, b in zip(x, y))
