"""The host integrator's initial-correction hook (CVodeSetInitialCorrectionFn,
core/cvode_bdf.c; the reference always starts the Newton iteration from a
zero correction, cvode.c:3095-3109), on the CPU with the serial vector and
the dense LU, on the reference's Robertson problem (cvRoberts_dns settings).

- A hook that declines (non-zero return) changes nothing: every counter and
  the state at every output are bit-identical to the run without a hook.
- A hook that proposes the zero correction changes nothing either (its use is
  counted): after one iteration from zero the correction equals the update,
  so the correction norm taken explicitly equals the reference's shortcut.
- On a linear problem with an exact Newton matrix, a hook that proposes the
  converged correction of an earlier run (recorded step by step in one-step
  mode) makes every Newton iteration converge at the first update, and the
  run retraces the earlier one's step sizes (to 1e-7): the error test has to use the
  norm of the correction itself, not of the last (rounding-size) update --
  taking the update's norm would accept every step with no error and
  multiply h by the largest growth factor.
- A hook that proposes a poor correction (30 % of the predicted state) costs
  Newton iterations but not accuracy: the outputs stay within the
  tolerance-scaled distance of the plain run.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib

RhsFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)
JacFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                    C.c_void_p, C.c_void_p, C.c_void_p)
InitFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                     C.c_void_p)
CV_BDF, CV_NORMAL, CV_ONE_STEP = 2, 1, 2
PD = C.POINTER(C.c_double)


def _arr(core, v, n=3):
    return np.ctypeslib.as_array(core.N_VGetArrayPointer(v), shape=(n,))


@pytest.fixture(scope="module")
def libs():
    core, orc = _lib.core(), _lib.load_oracle()
    for name, res, args in (
            ("CVodeCreate", C.c_void_p, [C.c_int, C.c_void_p]),
            ("CVodeInit", C.c_int, [C.c_void_p, RhsFn, C.c_double, C.c_void_p]),
            ("CVodeSVtolerances", C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
            ("CVodeSetLinearSolver", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
            ("CVodeSetJacFn", C.c_int, [C.c_void_p, JacFn]),
            ("CVodeSetInitialCorrectionFn", C.c_int, [C.c_void_p, InitFn]),
            ("CVodeSetMaxNumSteps", C.c_int, [C.c_void_p, C.c_long]),
            ("CVodeSetLSetupFrequency", C.c_int, [C.c_void_p, C.c_long]),
            ("CVodeSStolerances", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
            ("CVode", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, PD, C.c_int]),
            ("CVodeGetNumInitialCorrectionsUsed", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeGetNumSteps", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeGetNumNonlinSolvIters", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeGetNumErrTestFails", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeGetNumRhsEvals", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeGetNumLinSolvSetups", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
            ("CVodeFree", None, [C.POINTER(C.c_void_p)])):
        f = getattr(core, name)
        f.restype, f.argtypes = res, args
    orc.SUNDenseMatrix_Cols.restype = C.POINTER(PD)
    orc.SUNDenseMatrix_Cols.argtypes = [C.c_void_p]
    return core, orc


LIN = np.array([[-1.0, 0.5, 0.0], [0.0, -10.0, 1.0], [0.2, 0.0, -100.0]])


def _run(libs, hook=None, itask=CV_NORMAL, touts=(0.4, 4.0, 40.0, 400.0), record=None,
         problem="robertson", msbp=0):
    """Robertson as in cvRoberts_dns (rtol 1e-4, atol 1e-8 / 1e-14 / 1e-6,
    analytic Jacobian), or the linear y' = LIN y from y = (1, 1, 1) (rtol
    1e-6, atol 1e-10); msbp: linear-solver setup frequency (0: default);
    returns (outputs, counters, hook uses).  In one-step mode `record`
    receives (t_n, y_n) after every step up to the last tout."""
    core, orc = libs
    ctx = _lib.new_context()
    y = orc.N_VNew_Serial(3, ctx)
    at = orc.N_VNew_Serial(3, ctx)
    linear = problem == "linear"
    _arr(core, y)[:] = (1.0, 1.0, 1.0) if linear else (1.0, 0.0, 0.0)
    _arr(core, at)[:] = (1e-8, 1e-14, 1e-6)

    def f(t, yv, dv, ud):
        a, d = _arr(core, yv), _arr(core, dv)
        if linear:
            d[:] = LIN @ a
            return 0
        yd1 = -0.04 * a[0] + 1.0e4 * a[1] * a[2]
        yd3 = 3.0e7 * a[1] * a[1]
        d[0], d[2], d[1] = yd1, yd3, -yd1 - yd3
        return 0

    def jac(t, yv, fy, J, ud, t1, t2, t3):
        a = _arr(core, yv)
        c = orc.SUNDenseMatrix_Cols(J)
        if linear:
            for j in range(3):
                for i in range(3):
                    c[j][i] = LIN[i, j]
            return 0
        c[0][0], c[0][1], c[0][2] = -0.04, 0.04, 0.0
        c[1][0], c[1][1], c[1][2] = 1.0e4 * a[2], -1.0e4 * a[2] - 6.0e7 * a[1], 6.0e7 * a[1]
        c[2][0], c[2][1], c[2][2] = 1.0e4 * a[1], -1.0e4 * a[1], 0.0
        return 0

    cbs = [RhsFn(f), JacFn(jac)]
    cv = core.CVodeCreate(CV_BDF, ctx)
    A = orc.SUNDenseMatrix(3, 3, ctx)
    L = orc.SUNLinSol_Dense(y, A, ctx)
    assert core.CVodeInit(cv, cbs[0], 0.0, y) == 0
    if linear:
        assert core.CVodeSStolerances(cv, 1e-6, 1e-10) == 0
    else:
        assert core.CVodeSVtolerances(cv, 1e-4, at) == 0
    assert core.CVodeSetLSetupFrequency(cv, msbp) == 0
    assert core.CVodeSetLinearSolver(cv, L, A) == 0
    assert core.CVodeSetJacFn(cv, cbs[1]) == 0
    assert core.CVodeSetMaxNumSteps(cv, 100000) == 0
    if hook is not None:
        cbs.append(InitFn(lambda t, h, g, yp, ac, ud: hook(t, h, g, _arr(core, yp),
                                                           _arr(core, ac))))
        assert core.CVodeSetInitialCorrectionFn(cv, cbs[-1]) == 0
    outs = []
    t = C.c_double(0.0)
    if itask == CV_NORMAL:
        for tout in touts:
            assert core.CVode(cv, tout, y, C.byref(t), CV_NORMAL) == 0
            outs.append((t.value, _arr(core, y).copy()))
    else:
        while t.value < touts[-1]:
            assert core.CVode(cv, touts[-1], y, C.byref(t), CV_ONE_STEP) == 0
            outs.append((t.value, _arr(core, y).copy()))
            if record is not None:
                record.append((t.value, _arr(core, y).copy()))
    cnt = {}
    for name in ("Steps", "NonlinSolvIters", "ErrTestFails", "RhsEvals", "LinSolvSetups"):
        v = C.c_long()
        assert getattr(core, "CVodeGetNum" + name)(cv, C.byref(v)) == 0
        cnt[name] = v.value
    used = C.c_long()
    assert core.CVodeGetNumInitialCorrectionsUsed(cv, C.byref(used)) == 0
    p = C.c_void_p(cv)
    core.CVodeFree(C.byref(p))
    core.SUNLinSolFree(L)
    core.SUNMatDestroy(A)
    core.N_VDestroy(y)
    core.N_VDestroy(at)
    _lib.free_context(ctx)
    return outs, cnt, used.value


def _same(a, b):
    return all(ta == tb and np.array_equal(ya, yb) for (ta, ya), (tb, yb) in zip(a, b))


def test_declining_hook_changes_nothing(libs):
    plain, c0, _ = _run(libs)
    outs, c1, used = _run(libs, hook=lambda t, h, g, yp, ac: 1)
    assert used == 0 and c1 == c0 and _same(outs, plain)


def test_zero_proposal_changes_nothing(libs):
    plain, c0, _ = _run(libs)

    def zero(t, h, g, yp, ac):
        ac[:] = 0.0
        return 0

    outs, c1, used = _run(libs, hook=zero)
    assert used >= c0["Steps"] and c1 == c0 and _same(outs, plain)


def test_converged_proposal_keeps_the_step_sequence(libs):
    # Linear problem, linear-solver setup at every step: the Newton matrix is
    # exact, so one iteration from any start lands on the corrector's
    # solution.  Run 1 (plain, one step at a time) records y_n after every
    # step and where the hook is called; run 2 proposes y_n - y_pred, the
    # correction run 1 converged to, at every call that matches run 1's call
    # of the same index, so its Newton iterations converge at once with an
    # update of rounding size.
    rec, calls1 = [], []

    def spy(t, h, g, yp, ac):
        calls1.append((t, h))
        return 1

    kw = dict(itask=CV_ONE_STEP, touts=(2.0,), problem="linear", msbp=1)
    _, c1, _ = _run(libs, hook=spy, record=rec, **kw)
    ts = np.array([t for (t, _) in rec])
    calls2, hits = [], [0]

    def oracle_guess(t, h, g, yp, ac):
        calls2.append((t, h))
        k = len(calls2) - 1
        if k < len(calls1) and abs(calls1[k][0] - t) <= 1e-7 * abs(t) \
                and abs(calls1[k][1] - h) <= 1e-7 * abs(h):
            j = int(np.argmin(np.abs(ts - t)))
            if abs(ts[j] - t) <= 1e-7 * abs(t):
                ac[:] = rec[j][1] - yp
                hits[0] += 1
                return 0
        return 1

    _, c2, used = _run(libs, hook=oracle_guess, **kw)
    # every step of run 2 retraces run 1 (same step sizes to 1e-7; the extra
    # update refines run 1's corrections by ~1e-9): its error test measured
    # the correction, not the update -- an error estimate of rounding size
    # would have grown h by the largest factor (x10) at once
    # (a call for an attempt that run 1 then rejected has no accepted y_n:
    # the hook declines there)
    assert len(calls2) == len(calls1) and used == hits[0] == len(rec), \
        (used, len(rec), len(calls1), len(calls2))
    for (t1, h1), (t2, h2) in zip(calls1, calls2):
        assert t2 == pytest.approx(t1, rel=1e-7, abs=0.0)
        assert h2 == pytest.approx(h1, rel=1e-7, abs=0.0)
    growth = [calls1[k + 1][1] / calls1[k][1] for k in range(5, len(calls1) - 1)]
    assert max(growth) < 5.0
    # and each of its Newton iterations converged at the first update
    assert c2["Steps"] == c1["Steps"] and c2["ErrTestFails"] == c1["ErrTestFails"]
    assert c2["NonlinSolvIters"] < c1["NonlinSolvIters"]


def test_poor_proposal_keeps_accuracy(libs):
    plain, c0, _ = _run(libs)

    def poor(t, h, g, yp, ac):
        ac[:] = 0.3 * yp
        return 0

    outs, c1, used = _run(libs, hook=poor)
    assert used > 0
    assert c1["NonlinSolvIters"] > c0["NonlinSolvIters"]
    for (ta, ya), (tb, yb) in zip(outs, plain):
        assert ta == tb
        w = 1.0 / (1e-4 * np.abs(yb) + np.array([1e-8, 1e-14, 1e-6]))
        assert np.sqrt(np.mean(((ya - yb) * w) ** 2)) < 20.0
