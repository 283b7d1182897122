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
