"""SUNNonlinSol_NewtonSetNegatedLSolve (no reference counterpart), on the CPU
with the serial vector: with the option on, the Newton solver hands its
linear-solve callback the residual itself instead of its negation
(sunnonlinsol_newton.c:239 skipped), for a callback that solves J x = -b.

A modified Newton iteration on a small nonlinear system is run twice -- the
reference loop with a callback solving J x = b, and the option on with a
callback solving J x = -b (negating first) -- and the two runs see the same
residuals, updates and corrections bit for bit: the callback of the second run
receives r where the first receives -r.  A callback that ignores the option
(solving J x = b for the un-negated residual) walks uphill and fails to
converge, so the option really removes the loop's negation.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib

SysFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
LSolveFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
CTestFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                      C.c_void_p)
SUN_NLS_CONTINUE = 901
SUN_ERR_ARG_INCOMPATIBLE = -9998

A = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.25], [0.5, 0.25, 2.0]])
C0 = np.array([1.0, -2.0, 0.5])


@pytest.fixture(scope="module")
def libs():
    core, orc = _lib.core(), _lib.load_oracle()
    sigs = (
        ("SUNNonlinSol_Newton", C.c_void_p, [C.c_void_p, C.c_void_p]),
        ("SUNNonlinSol_NewtonSetNegatedLSolve", C.c_int, [C.c_void_p, C.c_int]),
        ("SUNNonlinSolSetSysFn", C.c_int, [C.c_void_p, SysFn]),
        ("SUNNonlinSolSetLSolveFn", C.c_int, [C.c_void_p, LSolveFn]),
        ("SUNNonlinSolSetConvTestFn", C.c_int, [C.c_void_p, CTestFn, C.c_void_p]),
        ("SUNNonlinSolSetMaxIters", C.c_int, [C.c_void_p, C.c_int]),
        ("SUNNonlinSolInitialize", C.c_int, [C.c_void_p]),
        ("SUNNonlinSolSolve", C.c_int,
         [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p]),
        ("SUNNonlinSolFree", C.c_int, [C.c_void_p]),
    )
    for name, res, args in sigs:
        f = getattr(core, name)
        f.restype, f.argtypes = res, args
    return core, orc


def _arr(core, v):
    return np.ctypeslib.as_array(C.cast(core.N_VGetArrayPointer(v), C.POINTER(C.c_double)),
                                 shape=(3,))


def _run(core, orc, negated, callback_negates):
    ctx = _lib.new_context()
    y0 = orc.N_VNew_Serial(3, ctx)
    ycor = orc.N_VNew_Serial(3, ctx)
    w = orc.N_VNew_Serial(3, ctx)
    _arr(core, y0)[:] = [0.3, -0.7, 0.2]
    _arr(core, ycor)[:] = 0.0
    _arr(core, w)[:] = 1.0
    seen = {"b": [], "ycor": []}

    def sys(yc, F, mem):
        y = _arr(core, y0) + _arr(core, yc)
        _arr(core, F)[:] = A @ y + 0.1 * y ** 3 - C0
        return 0

    def lsolve(b, mem):
        bv = _arr(core, b)
        seen["b"].append(bv.copy())
        rhs = -bv if callback_negates else bv
        bv[:] = np.linalg.solve(A, rhs)
        return 0

    def ctest(nls, yc, dl, tol, ewt, data):
        seen["ycor"].append(_arr(core, yc).copy())
        with np.errstate(over="ignore", invalid="ignore"):
            nrm = float(np.sqrt(np.mean(_arr(core, dl) ** 2)))
        return 0 if nrm < tol else SUN_NLS_CONTINUE

    cbs = (SysFn(sys), LSolveFn(lsolve), CTestFn(ctest))
    NLS = core.SUNNonlinSol_Newton(y0, ctx)
    assert NLS
    assert core.SUNNonlinSolSetSysFn(NLS, cbs[0]) == 0
    assert core.SUNNonlinSolSetLSolveFn(NLS, cbs[1]) == 0
    assert core.SUNNonlinSolSetConvTestFn(NLS, cbs[2], None) == 0
    assert core.SUNNonlinSolSetMaxIters(NLS, 12) == 0
    assert core.SUNNonlinSolInitialize(NLS) == 0
    if negated:
        assert core.SUNNonlinSol_NewtonSetNegatedLSolve(NLS, 1) == 0
    rc = core.SUNNonlinSolSolve(NLS, y0, ycor, w, 1e-12, 0, None)
    core.SUNNonlinSolFree(NLS)
    for v in (y0, ycor, w):
        core.N_VDestroy(v)
    _lib.free_context(ctx)
    return rc, seen


def test_negated_lsolve_same_iterates(libs):
    core, orc = libs
    rc_ref, ref = _run(core, orc, negated=False, callback_negates=False)
    rc_neg, neg = _run(core, orc, negated=True, callback_negates=True)
    assert rc_ref == 0 and rc_neg == 0
    assert len(ref["b"]) == len(neg["b"]) >= 3
    for b_ref, b_neg in zip(ref["b"], neg["b"]):
        # the reference loop hands over -r, the negated one r
        assert np.array_equal(b_ref, -b_neg)
    for y_ref, y_neg in zip(ref["ycor"], neg["ycor"]):
        assert np.array_equal(y_ref, y_neg)


def test_negated_lsolve_really_skips_the_negation(libs):
    core, orc = libs
    rc, seen = _run(core, orc, negated=True, callback_negates=False)
    assert rc != 0
    assert core.SUNNonlinSol_NewtonSetNegatedLSolve(None, 1) == SUN_ERR_ARG_INCOMPATIBLE
