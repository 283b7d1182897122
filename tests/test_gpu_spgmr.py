"""SPGMR on the MI355X vector against the same solver on the serial oracle
vector, over the reference's SPGMR unit-test scenarios (tridiag(-1, 5, -1),
N = 100, maxl = 100, tol 1e-13, row / column scaling, Jacobi
preconditioning, zero and non-zero initial guesses).

With sequential-order reductions the device run equals the serial run bit
for bit: return flag, iteration count, residual norm and solution, for both
Gram-Schmidt variants (modified: DotProd + LinearSum; classical: the fused
DotProdMulti + LinearCombination kernels).  With the default tree-order
reductions every solve still meets the reference test's own criterion.
"""
import numpy as np
import pytest

from sundials_amd import _lib
from spgmr_util import SerialBackend, DeviceBackend, run_scenarios, solution_ok, TOL, \
    MODIFIED_GS, CLASSICAL_GS, PREC_LEFT, PREC_RIGHT

pytestmark = pytest.mark.gpu

REDUCE_TREE, REDUCE_SEQUENTIAL = 0, 1


@pytest.fixture(scope="module")
def ctx():
    import torch

    assert torch.cuda.is_available()
    c = _lib.new_context()
    yield c
    _lib.free_context(c)


@pytest.mark.parametrize("gstype", [MODIFIED_GS, CLASSICAL_GS])
@pytest.mark.parametrize("pretype", [PREC_LEFT, PREC_RIGHT])
def test_spgmr_device_bitwise_sequential(ctx, gstype, pretype):
    ref = run_scenarios(SerialBackend(ctx), gstype, pretype)
    dev = run_scenarios(DeviceBackend(ctx, REDUCE_SEQUENTIAL), gstype, pretype)
    for r, d in zip(ref, dev):
        assert r[:4] == d[:4], (r[:4], d[:4])
        assert r[4] == d[4]
        assert np.array_equal(r[5], d[5]), r[:2]


@pytest.mark.parametrize("gstype", [MODIFIED_GS, CLASSICAL_GS])
def test_spgmr_device_tree_order_meets_reference_criterion(ctx, gstype):
    for (si, zg, flag, iters, res, y, x) in run_scenarios(DeviceBackend(ctx, REDUCE_TREE),
                                                          gstype, PREC_RIGHT):
        assert flag == 0, (si, zg, flag)
        assert solution_ok(x, y, 10 * TOL), (si, zg)


def test_spgmr_device_restarts(ctx):
    ref = run_scenarios(SerialBackend(ctx), CLASSICAL_GS, PREC_LEFT, maxl=3, max_restarts=60)
    dev = run_scenarios(DeviceBackend(ctx, REDUCE_SEQUENTIAL), CLASSICAL_GS, PREC_LEFT, maxl=3,
                        max_restarts=60)
    for r, d in zip(ref, dev):
        assert r[:4] == d[:4]
        assert np.array_equal(r[5], d[5])


def _device_gs():
    import ctypes as C

    return C.cast(_lib.hip().SUNAMD_ClassicalGS_Mi355, C.c_void_p)


def test_device_gs_step_builds_orthonormal_basis(ctx):
    """The fused device classical GS step: projections equal V^T w, the new
    norm equals the residual's, and the basis it builds is orthonormal."""
    import ctypes as C
    from helpers import gpu_vec, gpu_get, vec_array

    core, hip = _lib.core(), _lib.hip()
    rng = np.random.default_rng(4)
    n, k = 100003, 20
    A = rng.standard_normal((n, k + 1))
    vs = [gpu_vec(ctx, A[:, 0] / np.linalg.norm(A[:, 0]))]
    H = np.zeros((k + 1, k))
    hp = (C.POINTER(C.c_double) * (k + 1))(
        *[H[i].ctypes.data_as(C.POINTER(C.c_double)) for i in range(k + 1)])
    stemp = np.zeros(k + 2)
    vtemp = (C.c_void_p * (k + 2))()
    for j in range(1, k + 1):
        vs.append(gpu_vec(ctx, A[:, j]))
        Q = np.stack([gpu_get(v) for v in vs[:-1]], axis=1)
        w = A[:, j]
        nrm = C.c_double()
        assert hip.SUNAMD_ClassicalGS_Mi355(vec_array(vs), hp, j, k, C.byref(nrm),
                                            stemp.ctypes.data_as(C.POINTER(C.c_double)),
                                            vtemp) == 0
        assert np.allclose(H[:j, j - 1], Q.T @ w, rtol=1e-11, atol=1e-11)
        r = w - Q @ (Q.T @ w)
        assert abs(nrm.value - np.linalg.norm(r)) <= 1e-11 * np.linalg.norm(w)
        core.N_VScale(1.0 / nrm.value, vs[-1], vs[-1])
    V = np.stack([gpu_get(v) for v in vs], axis=1)
    assert np.abs(V.T @ V - np.eye(k + 1)).max() < 1e-12
    for v in vs:
        core.N_VDestroy(v)


def test_spgmr_with_device_gs_step(ctx):
    # tree order: the reference test's criterion; sequential order: the step
    # defers to SUNClassicalGS, so the run equals the serial one bit for bit
    for (si, zg, flag, iters, res, y, x) in run_scenarios(
            DeviceBackend(ctx, REDUCE_TREE), CLASSICAL_GS, PREC_LEFT, cgs_fn=_device_gs()):
        assert flag == 0, (si, zg, flag)
        assert solution_ok(x, y, 10 * TOL), (si, zg)
    ref = run_scenarios(SerialBackend(ctx), CLASSICAL_GS, PREC_RIGHT)
    dev = run_scenarios(DeviceBackend(ctx, REDUCE_SEQUENTIAL), CLASSICAL_GS, PREC_RIGHT,
                        cgs_fn=_device_gs())
    for r, d in zip(ref, dev):
        assert r[:4] == d[:4]
        assert np.array_equal(r[5], d[5])
