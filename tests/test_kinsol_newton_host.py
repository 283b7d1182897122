"""KINSOL Newton iteration (core/kinsol_newton.c) on the CPU (serial vector):

- the reference example kinFerTron_dns: all 8 published runs (two initial
  guesses x exact / modified Newton x with / without line search, one KINSOL
  memory as in the example, dense difference-quotient Jacobian, inequality
  constraints) give the published nni / nfe / nje / nfeD and the printed
  solution strings (tests/golden/kin_fertron_dns.json);
- the line search: on F(u) = atan(u) from u0 where the full Newton step
  overshoots, one iteration ends at u0 + lambda p where lambda comes from
  minimising the interpolants of phi(lambda) = |F(u0 + lambda p)|^2 / 2
  (quadratic through phi(0), phi'(0), phi(1); then cubics through the last
  two trials), clipped to [0.1, 0.5] of the previous trial, until the
  sufficient-decrease condition phi(lambda) <= phi(0) + 1e-4 lambda phi'(0)
  holds, then bisection towards the previous trial while the step is too
  short for phi(lambda) >= phi(0) + 0.9 lambda phi'(0) (Dennis & Schnabel,
  Algorithm A6.3.1mod) -- the interpolants are fitted here by solving their
  interpolation conditions; the backtrack count and nfe (one evaluation per
  trial) agree exactly;
- the constraint cut: a step that leaves the feasible set is scaled by 0.9
  times the largest feasible fraction over the constrained components only;
- the test problem's deterministic sin / exp (variant 1) agree with the math
  library to a few ulp and its analytic Jacobian with central differences;
  the device twin (variant 1, analytic block Jacobian, per-block LU) retraces
  the example's counts;
- input errors (infeasible initial guess, Picard, constraints with KIN_FP,
  constraint values beyond +-2, an iterative linear solver, fused norms on
  serial vectors) return the reference's codes.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from sundials_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))

SysFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
JacFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                    C.c_void_p)


def _golden():
    return json.load(open(os.path.join(HERE, "golden", "kin_fertron_dns.json")))


def test_kin_fertron_matches_published_runs():
    g = _golden()
    x = np.zeros(16)
    st = np.zeros(8 * _lib.KINN_STATS)
    assert _lib.load_oracle().oracle_kin_fertron_example(dp(x), dp(st)) == 0
    assert len(g["runs"]) == 8
    for r, run in enumerate(g["runs"]):
        s = st[_lib.KINN_STATS * r:_lib.KINN_STATS * (r + 1)]
        assert int(s[0]) == 0, run
        assert (int(s[1]), int(s[2]), int(s[3]), int(s[4])) == \
            (run["nni"], run["nfe"], run["nje"], run["nfeD"]), run
        assert ("%10.6g" % x[2 * r]).strip() == run["x1"], run
        assert ("%10.6g" % x[2 * r + 1]).strip() == run["x2"], run


def _newton_opts(**kw):
    o = _lib.KinNewtonOptions()
    o.ncells, o.variant, o.guess, o.perturb, o.strategy, o.msbset = 1, 1, 2, 0, 0, 1
    o.fnormtol, o.scsteptol, o.maxiter, o.jac, o.sequential = 1e-5, 1e-5, 0, 1, 0
    o.fused = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def oracle_newton(o):
    st = np.zeros(_lib.KINN_STATS)
    u = np.zeros(6 * o.ncells)
    rc = _lib.load_oracle().oracle_kin_newton_solve(C.byref(o), dp(st), dp(u))
    return rc, st, u


@pytest.mark.parametrize("guess,strategy,msbset,ref_run", [
    (1, 0, 1, 0), (1, 1, 1, 1), (1, 0, 0, 2), (1, 1, 0, 3),
    (2, 0, 1, 4), (2, 1, 1, 5)])
def test_device_twin_retraces_example(guess, strategy, msbset, ref_run):
    """Variant 1 + analytic block Jacobian (what the MI355X path computes):
    same iteration and evaluation counts as the published dense-DQ runs, the
    same printed solution (the two modified-Newton runs of guess 2 stop at a
    point whose 6th digit depends on the Jacobian's rounding and are left out)."""
    run = _golden()["runs"][ref_run]
    rc, st, u = oracle_newton(_newton_opts(guess=guess, strategy=strategy, msbset=msbset))
    assert rc == 0
    assert (int(st[1]), int(st[2]), int(st[3])) == (run["nni"], run["nfe"], run["nje"])
    assert int(st[4]) == 0  # no difference-quotient evaluations with the analytic Jacobian
    assert ("%10.6g" % u[0]).strip() == run["x1"]
    assert ("%10.6g" % u[1]).strip() == run["x2"]


def test_cell_arithmetic_variants_and_jacobian():
    orc = _lib.load_oracle()
    orc.oracle_kft_f.restype = None
    orc.oracle_kft_jac.restype = None
    rng = np.random.default_rng(3)
    for _ in range(200):
        x1 = rng.uniform(-1.0, 2.0)
        x2 = rng.uniform(-7.0, 7.0)
        u = np.array([x1, x2, x1 - 0.25, x1 - 1.0, x2 - 1.5, x2 - 2 * 3.1415926])
        f0, f1 = np.zeros(6), np.zeros(6)
        orc.oracle_kft_f(0, dp(u), dp(f0))
        orc.oracle_kft_f(1, dp(u), dp(f1))
        # the deterministic sin / exp agree with the math library to a few ulp
        assert np.allclose(f1, f0, rtol=0, atol=8e-16 * (1 + np.exp(2 * x1)))
        J = np.zeros(36)
        orc.oracle_kft_jac(1, dp(u), dp(J))
        J = J.reshape(6, 6).T  # column-major block
        h = 1e-6
        for j in range(6):
            up, um = u.copy(), u.copy()
            up[j] += h
            um[j] -= h
            fp, fm = np.zeros(6), np.zeros(6)
            orc.oracle_kft_f(0, dp(up), dp(fp))
            orc.oracle_kft_f(0, dp(um), dp(fm))
            assert np.allclose(J[:, j], (fp - fm) / (2 * h), rtol=1e-6,
                               atol=1e-7 * (1 + np.exp(2 * x1)))


def test_batched_twin_cells_converge_to_the_known_roots():
    for guess, root in ((1, (0.29945, 2.83693)), (2, (0.5, 3.1415926))):
        for strategy, msbset in ((0, 1), (1, 0)):
            rc, st, u = oracle_newton(_newton_opts(ncells=256, perturb=1, guess=guess,
                                                   strategy=strategy, msbset=msbset,
                                                   fnormtol=1e-10, scsteptol=1e-12))
            assert rc == 0
            x = u.reshape(-1, 6)
            assert np.allclose(x[:, 0], root[0], atol=2e-5)
            assert np.allclose(x[:, 1], root[1], atol=2e-5)
            # constraints l >= 0, L <= 0 hold at the solution
            assert (x[:, 2] >= 0).all() and (x[:, 4] >= 0).all()
            assert (x[:, 3] <= 0).all() and (x[:, 5] <= 0).all()


# ---------------------------------------------------------------------------
# small systems through the C API with Python callbacks
# ---------------------------------------------------------------------------
class _Kin:
    """KINSOL memory on serial vectors with a dense matrix and solver."""

    def __init__(self, n, F, J, constraints=None, maxiter=1, strategy=1):
        core, orc = _lib.core(), _lib.load_oracle()
        for name, res, args in (
                ("KINCreate", C.c_void_p, [C.c_void_p]),
                ("KINInit", C.c_int, [C.c_void_p, SysFn, C.c_void_p]),
                ("KINSetLinearSolver", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
                ("KINSetJacFn", C.c_int, [C.c_void_p, JacFn]),
                ("KINSetConstraints", C.c_int, [C.c_void_p, C.c_void_p]),
                ("KINSetNumMaxIters", C.c_int, [C.c_void_p, C.c_long]),
                ("KINSetMaxSetupCalls", C.c_int, [C.c_void_p, C.c_long]),
                ("KINSetFuncNormTol", C.c_int, [C.c_void_p, C.c_double]),
                ("KINSol", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
                ("KINGetNumFuncEvals", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
                ("KINGetNumBacktrackOps", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
                ("KINFree", None, [C.POINTER(C.c_void_p)])):
            fn = getattr(core, name)
            fn.restype, fn.argtypes = res, args
        self.core, self.n = core, n
        self.ctx = _lib.new_context() if hasattr(_lib, "new_context") else self._ctx()
        arr = lambda v: np.ctypeslib.as_array(core.N_VGetArrayPointer(v), shape=(n,))
        self.arr = arr

        def f_cb(uv, fv, ud):
            arr(fv)[:] = F(arr(uv).copy())
            return 0

        def j_cb(uv, fu, Jm, ud, t1, t2):
            a = np.ctypeslib.as_array(orc.SUNDenseMatrix_Data(Jm), shape=(n * n,))
            a[:] = np.asarray(J(arr(uv).copy()), float).ravel(order="F")
            return 0

        self.f_cb, self.j_cb = SysFn(f_cb), JacFn(j_cb)
        self.u = orc.N_VNew_Serial(n, self.ctx)
        self.s = orc.N_VNew_Serial(n, self.ctx)
        arr(self.s)[:] = 1.0
        self.k = C.c_void_p(core.KINCreate(self.ctx))
        assert core.KINInit(self.k, self.f_cb, self.u) == 0
        self.A = orc.SUNDenseMatrix(n, n, self.ctx)
        self.LS = orc.SUNLinSol_Dense(self.u, self.A, self.ctx)
        assert core.KINSetLinearSolver(self.k, self.LS, self.A) == 0
        assert core.KINSetJacFn(self.k, self.j_cb) == 0
        if constraints is not None:
            self.c = orc.N_VNew_Serial(n, self.ctx)
            arr(self.c)[:] = constraints
            assert core.KINSetConstraints(self.k, self.c) == 0
        assert core.KINSetNumMaxIters(self.k, maxiter) == 0
        assert core.KINSetMaxSetupCalls(self.k, 1) == 0
        assert core.KINSetFuncNormTol(self.k, 1e-300) == 0
        self.strategy = strategy

    @staticmethod
    def _ctx():
        ctx = C.c_void_p()
        assert _lib.core().SUNContext_Create(0, C.byref(ctx)) == 0
        return ctx

    def solve(self, u0):
        self.arr(self.u)[:] = u0
        flag = self.core.KINSol(self.k, self.u, self.strategy, self.s, self.s)
        nfe, nbk = C.c_long(), C.c_long()
        self.core.KINGetNumFuncEvals(self.k, C.byref(nfe))
        self.core.KINGetNumBacktrackOps(self.k, C.byref(nbk))
        return flag, self.arr(self.u).copy(), nfe.value, nbk.value

    def close(self):
        self.core.KINFree(C.byref(self.k))


def _expected_line_search(u0, scsteptol):
    """Dennis & Schnabel's backtracking (Algorithm A6.3.1mod) on
    phi(l) = atan(u0 + l p)^2 / 2: interpolants fitted by solving their
    interpolation conditions until the alpha condition (sufficient decrease,
    1e-4) holds; then, if the step is too short for the beta condition (0.9),
    bisection between the last two trial lengths, falling back to the last
    length that met the alpha condition.  Returns (u1, backtracks, fallback)."""
    F0 = math.atan(u0)
    p = -F0 / (1.0 / (1.0 + u0 * u0))
    phi = lambda l: 0.5 * math.atan(u0 + l * p) ** 2
    phi0, dphi0 = 0.5 * F0 * F0, -F0 * F0
    alpha_ok = lambda l, v: v <= phi0 + 1e-4 * l * dphi0
    beta_ok = lambda l, v: v >= phi0 + 0.9 * l * dphi0
    lmin = scsteptol / (abs(p) / (1.0 + abs(u0)))
    trials = [(1.0, phi(1.0))]
    while not alpha_ok(*trials[-1]):
        lam, val = trials[-1]
        if len(trials) == 1:
            # q(l) = a l^2 + dphi0 l + phi0 through (lam, val)
            a = np.linalg.solve([[lam * lam]], [val - phi0 - dphi0 * lam])[0]
            new = -dphi0 / (2 * a)
        else:
            (l1, v1), (l2, v2) = trials[-1], trials[-2]
            M = np.array([[l1 ** 3, l1 ** 2], [l2 ** 3, l2 ** 2]])
            a, b = np.linalg.solve(M, [v1 - phi0 - dphi0 * l1, v2 - phi0 - dphi0 * l2])
            roots = np.roots([3 * a, 2 * b, dphi0])
            new = max(r.real for r in roots if abs(r.imag) == 0 and 6 * a * r.real + 2 * b > 0)
        new = max(min(new, 0.5 * lam), 0.1 * lam)
        trials.append((new, phi(new)))
    lam, val = trials[-1]
    nbk, fallback = len(trials) - 1, 0
    if not beta_ok(lam, val) and lam < 1.0:
        lprev = trials[-2][0]
        lo, width = min(lam, lprev), abs(lprev - lam)
        while True:
            inc = 0.5 * width
            lam = lo + inc
            val = phi(lam)
            nbk += 1
            if not alpha_ok(lam, val):
                width = inc
            elif not beta_ok(lam, val):
                lo, width = lam, width - inc
            if alpha_ok(lam, val) and (beta_ok(lam, val) or width < lmin):
                break
        if not beta_ok(lam, val) or (width < lmin and not alpha_ok(lam, val)):
            lam, fallback = lo, 1
    return u0 + lam * p, nbk, fallback


@pytest.mark.parametrize("u0", [3.0, 10.0, 40.0])
def test_line_search_follows_the_interpolants(u0):
    K = _Kin(1, lambda u: np.arctan(u), lambda u: [[1.0 / (1.0 + u[0] ** 2)]])
    try:
        flag, u, nfe, nbk = K.solve([u0])
    finally:
        K.close()
    exp_u, exp_bk, fallback = _expected_line_search(u0, 2.220446049250313e-16 ** (2 / 3))
    assert flag == -6  # KIN_MAXITER_REACHED after the one iteration allowed
    assert nbk == exp_bk and nbk >= 1
    assert nfe == 2 + nbk + fallback  # initial F, full step, one per trial
    assert u[0] == pytest.approx(exp_u, rel=1e-12, abs=1e-12)


def test_constraint_cut_uses_constrained_components_only():
    # F = (u0 + 1, u1 + 7), u = (1, 1): Newton step p = (-2, -8).  Component 0
    # must stay > 0, component 1 is free: the free component's ratio 1/8 must
    # not count, so the step is scaled by 0.9 * (1 / 2).
    K = _Kin(2, lambda u: np.array([u[0] + 1.0, u[1] + 7.0]),
             lambda u: [[1.0, 0.0], [0.0, 1.0]], constraints=[2.0, 0.0], strategy=0)
    try:
        flag, u, nfe, _ = K.solve([1.0, 1.0])
    finally:
        K.close()
    assert flag == -6 and nfe == 2
    assert u[0] == pytest.approx(1.0 - 0.45 * 2.0, abs=1e-15)
    assert u[1] == pytest.approx(1.0 - 0.45 * 8.0, abs=1e-15)


def test_error_paths():
    """Input errors are reported with the reference's return codes."""
    core = _lib.core()
    core.KINSetUseFusedKernels.restype = C.c_int
    core.KINSetUseFusedKernels.argtypes = [C.c_void_p, C.c_int]
    core.KINSetConstraints.restype = C.c_int
    core.KINSetConstraints.argtypes = [C.c_void_p, C.c_void_p]
    # a solve whose initial guess violates the constraints: KIN_ILL_INPUT
    K = _Kin(1, lambda u: np.arctan(u), lambda u: [[1.0 / (1.0 + u[0] ** 2)]],
             constraints=[1.0])
    try:
        flag, _, _, _ = K.solve([-1.0])
        assert flag == -2
        # Picard is not part of this library; KIN_FP refuses constraints
        assert core.KINSol(K.k, K.u, 2, K.s, K.s) == -2
        assert core.KINSol(K.k, K.u, 3, K.s, K.s) == -2
        # fused norms need MI355X vectors
        assert core.KINSetUseFusedKernels(K.k, 1) == -2
        assert core.KINSetUseFusedKernels(K.k, 0) == 0
        # a constraints vector with an entry beyond +-2 is rejected
        orc = _lib.load_oracle()
        bad = orc.N_VNew_Serial(1, K.ctx)
        K.arr(bad)[:] = 3.0
        assert core.KINSetConstraints(K.k, bad) == -2
        # after removing the constraints the same start converges
        assert core.KINSetConstraints(K.k, None) == 0
        core.KINSetNumMaxIters(K.k, 50)
        core.KINSetFuncNormTol(K.k, 1e-12)
        flag, u, _, _ = K.solve([-1.0])
        assert flag in (0, 2) and abs(u[0]) < 1e-10
        # an iterative linear solver is refused by KINSetLinearSolver
        LS2 = core.SUNLinSol_SPGMR(K.u, 0, 5, K.ctx)
        assert core.KINSetLinearSolver(K.k, LS2, None) == -3
        core.SUNLinSolFree(LS2)
    finally:
        K.close()
