"""SPGMR and its Gram-Schmidt / Givens-QR kernels on the host (serial
oracle vector).

- The reference's SPGMR unit-test scenarios pass with the reference's own
  criterion (every entry of the solution within SUNRCompareTol(10 tol)).
- Both Gram-Schmidt variants produce an orthonormal basis and the Arnoldi
  coefficients of the projection (checked against numpy on random vectors).
- SUNQRfact / SUNQRsol solve the Hessenberg least-squares problem that
  numpy's lstsq solves, both for a fresh factorisation and column by column.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib
from helpers import cpu_vec, cpu_get, vec_array
from spgmr_util import SerialBackend, run_scenarios, solution_ok, TOL, MODIFIED_GS, \
    CLASSICAL_GS, PREC_LEFT, PREC_RIGHT


@pytest.fixture(scope="module")
def ctx():
    c = _lib.new_context()
    yield c
    _lib.free_context(c)


@pytest.mark.parametrize("gstype", [MODIFIED_GS, CLASSICAL_GS])
@pytest.mark.parametrize("pretype", [PREC_LEFT, PREC_RIGHT])
def test_spgmr_reference_scenarios(ctx, gstype, pretype):
    recs = run_scenarios(SerialBackend(ctx), gstype, pretype)
    assert len(recs) == 12
    for (si, zg, flag, iters, res, y, x) in recs:
        assert flag == 0, (si, zg, flag)
        assert 0 < iters <= 100 or (not zg and iters >= 0)
        assert solution_ok(x, y, 10 * TOL), (si, zg, np.max(np.abs(x - y) / np.abs(x)))


def test_spgmr_restarts_and_failure_flags(ctx):
    # maxl 3 cannot reach 1e-13 without restarts: with none the residual is
    # only reduced (SUNLS_RES_REDUCED = 801); with enough restarts it converges
    be = SerialBackend(ctx)
    short = run_scenarios(be, MODIFIED_GS, PREC_LEFT, maxl=3, max_restarts=0)
    assert all(r[2] == 801 for r in short)
    many = run_scenarios(be, MODIFIED_GS, PREC_LEFT, maxl=3, max_restarts=200)
    assert all(r[2] == 0 for r in many)
    for (si, zg, flag, iters, res, y, x) in many:
        assert solution_ok(x, y, 10 * TOL)


def _h_matrix(rows, cols):
    H = np.zeros((rows, cols))
    ptrs = (C.POINTER(C.c_double) * rows)(
        *[H[i].ctypes.data_as(C.POINTER(C.c_double)) for i in range(rows)])
    return H, ptrs


@pytest.mark.parametrize("gs", ["mgs", "cgs"])
def test_gram_schmidt_orthonormal_basis(ctx, gs):
    core, orc = _lib.core(), _lib.load_oracle()
    rng = np.random.default_rng(3)
    n, k = 200, 12
    A = rng.standard_normal((n, k + 1))
    vs = [cpu_vec(orc, ctx, A[:, 0] / np.linalg.norm(A[:, 0]))]
    H, hp = _h_matrix(k + 1, k)
    stemp = np.zeros(k + 2)
    vtemp = (C.c_void_p * (k + 2))()
    for j in range(1, k + 1):
        vs.append(cpu_vec(orc, ctx, A[:, j]))
        arr = vec_array(vs)
        nrm = C.c_double()
        w = A[:, j].copy()
        Q = np.stack([cpu_get(v) for v in vs[:-1]], axis=1)
        if gs == "mgs":
            assert core.SUNModifiedGS(arr, hp, j, k, C.byref(nrm)) == 0
        else:
            assert core.SUNClassicalGS(arr, hp, j, k, C.byref(nrm),
                                       stemp.ctypes.data_as(C.POINTER(C.c_double)), vtemp) == 0
        # projection coefficients and the remaining norm
        assert np.allclose(H[:j, j - 1], Q.T @ w, rtol=1e-12, atol=1e-12)
        r = w - Q @ (Q.T @ w)
        assert abs(nrm.value - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(w)
        core.N_VScale(1.0 / nrm.value, vs[-1], vs[-1])
    V = np.stack([cpu_get(v) for v in vs], axis=1)
    assert np.abs(V.T @ V - np.eye(k + 1)).max() < 1e-13
    for v in vs:
        core.N_VDestroy(v)


def test_gram_schmidt_reorthogonalises_nearly_dependent_vector(ctx):
    # v[k] = v[0] + 1e-12 * e: the first pass leaves a vector 1e-12 long, far
    # below 1/1000 of the input's norm, so the second pass runs.  Classical GS
    # then removes the remaining v[0] component relative to the new vector;
    # modified GS only applies corrections that are visible next to 1000 h,
    # so its result is orthogonal relative to the input's length
    core, orc = _lib.core(), _lib.load_oracle()
    n = 64
    rng = np.random.default_rng(9)
    q0 = rng.standard_normal(n)
    q0 /= np.linalg.norm(q0)
    e = rng.standard_normal(n)
    for gs in ("mgs", "cgs"):
        vs = [cpu_vec(orc, ctx, q0), cpu_vec(orc, ctx, q0 + 1e-12 * e)]
        H, hp = _h_matrix(2, 1)
        nrm = C.c_double()
        if gs == "mgs":
            core.SUNModifiedGS(vec_array(vs), hp, 1, 5, C.byref(nrm))
        else:
            stemp = np.zeros(3)
            core.SUNClassicalGS(vec_array(vs), hp, 1, 5, C.byref(nrm),
                                stemp.ctypes.data_as(C.POINTER(C.c_double)), (C.c_void_p * 3)())
        w = cpu_get(vs[1])
        if gs == "cgs":
            assert abs(np.dot(w, q0)) < 1e-14 * np.linalg.norm(w)
        else:
            assert abs(np.dot(w, q0)) < 1e-15
        assert abs(H[0, 0] - np.dot(q0, q0 + 1e-12 * e)) < 1e-15
        assert abs(nrm.value - np.linalg.norm(w)) < 1e-6 * np.linalg.norm(w)
        for v in vs:
            core.N_VDestroy(v)


@pytest.mark.parametrize("n", [1, 2, 5, 20])
def test_qr_least_squares(n):
    core = _lib.core()
    rng = np.random.default_rng(n)
    Hh = np.triu(rng.standard_normal((n + 1, n)), -1)  # upper Hessenberg
    g = np.zeros(n + 1)
    g[0] = 2.5
    y_ref = np.linalg.lstsq(Hh, g, rcond=None)[0]
    # fresh factorisation (job 0) and column-by-column updates agree with lstsq
    for mode in ("fresh", "incremental"):
        H, hp = _h_matrix(n + 1, n)
        H[:] = Hh
        q = np.zeros(2 * n)
        qp = q.ctypes.data_as(C.POINTER(C.c_double))
        if mode == "fresh":
            assert core.SUNQRfact(n, hp, qp, 0) == 0
        else:
            for j in range(n):
                assert core.SUNQRfact(j + 1, hp, qp, j) == 0
        b = g.copy()
        assert core.SUNQRsol(n, hp, qp, b.ctypes.data_as(C.POINTER(C.c_double))) == 0
        # a fresh factorisation (job 0) applies only k-1 earlier rotations to
        # column k, as the reference does; SPGMR uses it for the first column
        # only, where it is exact
        if mode == "incremental" or n == 1:
            assert np.allclose(b[:n], y_ref, rtol=1e-10, atol=1e-12), mode
            # |last rotated rhs entry| is the least-squares residual norm
            assert abs(abs(b[n]) - np.linalg.norm(Hh @ y_ref - g)) < 1e-10
