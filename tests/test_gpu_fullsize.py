"""Full-size checks at the benchmark shapes, through properties that do not
depend on size.

- N_Vector at N = 2^28: LinearSum bit for bit against numpy's elementwise
  a*x + b*y (the same two products and one sum per element); WrmsNorm and
  DotProd against numpy sums to a relative 1e-12 (tree order against numpy's
  pairwise order), and exactly on a constant fill whose partial sums are all
  exact (every term 1/16: the sum 2^24 in any order, the norm 1/4).
- Block LU on 2^20 blocks of 32 x 32 (8 GiB): factors, pivots and one solve
  bit for bit against the oracle's dense GETRF / GETRS on sampled blocks
  (first, last and spread), and the one-pass Newton-matrix setup equal to the
  three separate calls over the whole 8 GiB.
- Coupled BDF on 2^20 cells x 32 species (the coupled benchmark's shape),
  tiled from 128 base cells, host-stepped and device-stepped: cells interact
  only through the norms, so every tile ends bit-identical to every other,
  and the step count and the error against a 1000x tighter reference
  solution of the base cells stay where the oracle's own run of the base
  cells puts them.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from sundials_amd import _lib
from helpers import dp, gpu_get, gpu_vec

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[fullsize] " + msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def libs():
    import torch

    assert torch.cuda.is_available()
    return _lib.core(), _lib.hip(), _lib.load_oracle()


@pytest.fixture
def ctx():
    c = _lib.new_context()
    yield c
    _lib.free_context(c)


def test_vector_ops_2p28(libs, ctx):
    core, hip, _ = libs
    n = 1 << 28
    rng = np.random.default_rng(28)
    _say("vector: generating 2 x 2^28 values")
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    gx, gy = gpu_vec(ctx, x), gpu_vec(ctx, y)
    gz = hip.N_VNew_Mi355(n, ctx)
    assert gz
    _say("vector: LinearSum")
    core.N_VLinearSum(0.3, gx, -0.7, gy, gz)
    assert np.array_equal(gpu_get(gz), 0.3 * x + (-0.7) * y)
    _say("vector: DotProd / WrmsNorm")
    assert core.N_VDotProd(gx, gy) == pytest.approx(float(np.dot(x, y)), rel=1e-12)
    w = np.abs(y) + 1e-3
    gw = gpu_vec(ctx, w)
    ref = float(np.sqrt(np.mean((x * w) ** 2)))
    assert core.N_VWrmsNorm(gx, gw) == pytest.approx(ref, rel=1e-12)
    # every term (1/2 * 1/2)^2 = 1/16: all partial sums exact in any order
    core.N_VConst(0.5, gx)
    core.N_VConst(0.5, gw)
    assert core.N_VWrmsNorm(gx, gw) == 0.25
    assert core.N_VDotProd(gx, gw) == float(n) / 4.0
    for v in (gx, gy, gz, gw):
        core.N_VDestroy(v)


def test_block_lu_2p20x32(libs, ctx):
    core, hip, orc = libs
    nb, m = 1 << 20, 32
    n = nb * m
    rng = np.random.default_rng(2032)
    _say("block LU: generating 2^20 blocks of 32 x 32")
    # block-major column-major storage directly: entries U(-1, 1) / m, plus
    # the identity, scaled per block
    flat = rng.uniform(-1.0, 1.0, nb * m * m) / m
    blocks = flat.reshape(nb, m, m)          # [block][column][row]
    blocks[:, np.arange(m), np.arange(m)] += 1.0
    y = gpu_vec(ctx, np.zeros(n))
    stream = hip.N_VGetStream_Mi355(y)

    def matrix(values):
        M = hip.SUNMatrix_Mi355BlockDense(nb, m, stream, ctx)
        h = hip.SUNMatrix_Mi355BlockDense_HostData(M)
        C.memmove(h, dp(values), 8 * values.size)
        assert hip.SUNMatrix_Mi355BlockDense_CopyToDevice(M) == 0
        return M

    def values(M):
        assert hip.SUNMatrix_Mi355BlockDense_CopyFromDevice(M) == 0
        h = hip.SUNMatrix_Mi355BlockDense_HostData(M)
        return np.ctypeslib.as_array(h, shape=(nb * m * m,)).copy()

    MA = matrix(flat)
    LS = hip.SUNLinSol_Mi355BlockLU(y, MA, 0, ctx)
    assert core.SUNLinSolInitialize(LS) == 0
    _say("block LU: factor + solve")
    assert core.SUNLinSolSetup(LS, MA) == 0
    b = rng.uniform(-1, 1, n)
    gb, gx = gpu_vec(ctx, b), gpu_vec(ctx, np.zeros(n))
    assert core.SUNLinSolSolve(LS, MA, gx, gb, 0.0) == 0
    x = gpu_get(gx)
    F = values(MA)
    piv = np.zeros(n, np.int32)
    assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
    sample = sorted(set([0, 1, nb - 1, nb - 2] + list(rng.integers(0, nb, 60))))
    for k in sample:
        a = flat[k * m * m:(k + 1) * m * m].copy()
        p = np.zeros(m, np.int64)
        assert orc.oracle_block_getrf(dp(a), m, p.ctypes.data_as(C.POINTER(C.c_int64))) == 0
        assert np.array_equal(F[k * m * m:(k + 1) * m * m], a), ("factor", k)
        assert np.array_equal(piv[k * m:(k + 1) * m], p), ("pivots", k)
        bb = b[k * m:(k + 1) * m].copy()
        orc.oracle_block_getrs(dp(a), m, p.ctypes.data_as(C.POINTER(C.c_int64)), dp(bb))
        assert np.array_equal(x[k * m:(k + 1) * m], bb), ("solve", k)
    del F
    _say("block LU: one-pass Newton-matrix setup against the three calls")
    c = -0.0123
    MJ = matrix(flat)                       # the saved Jacobian
    assert hip.SUNLinSol_Mi355BlockLUSetupScaled(LS, MA, MJ, c, 0) == 0
    fused = values(MA)
    piv_f = np.zeros(n, np.int32)
    assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv_f.ctypes.data_as(C.POINTER(C.c_int))) == 0
    assert core.SUNMatCopy(MJ, MA) == 0
    assert core.SUNMatScaleAddI(c, MA) == 0
    assert core.SUNLinSolSetup(LS, MA) == 0
    assert np.array_equal(values(MA), fused)
    assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
    assert np.array_equal(piv, piv_f)
    core.SUNLinSolFree(LS)
    for M in (MA, MJ):
        core.SUNMatDestroy(M)
    for v in (y, gb, gx):
        core.N_VDestroy(v)


def test_coupled_2p20x32_tiled(libs):
    core, hip, orc = libs
    m, R, nbase, ncells = 32, 48, 128, 1 << 20
    base = core.SUNAMD_MassActionCreate(m, R, nbase, 1002, 0.2)
    big = core.SUNAMD_MassActionTile(base, ncells)
    assert base and big
    _say("coupled: 2^20 cells x 32 species, fused kernels, tree reductions")
    r = hip.SUNAMD_CoupledCreate_Mi355(big, 1e-6, 1e-10, 0, 0)
    assert r
    assert hip.SUNAMD_CoupledSetFusedKernels_Mi355(r, 1) == 0
    st = _lib.SolveStats()
    assert hip.SUNAMD_CoupledRun_Mi355(r, 1.0, C.byref(st)) == 0
    y = np.zeros(ncells * m)
    assert hip.SUNAMD_CoupledGetState_Mi355(r, dp(y)) == 0
    hip.SUNAMD_CoupledFree_Mi355(r)
    _say("coupled: nst=%d nni=%d; oracle on the 128 base cells" % (st.nst, st.nni))
    tiles = y.reshape(-1, nbase * m)
    assert all(np.array_equal(tiles[0], t) for t in tiles[1:])
    y_cpu, ost = np.zeros(nbase * m), np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(base, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    y_ref = np.zeros(nbase * m)
    assert orc.oracle_massaction_solve_coupled(base, 1e-9, 1e-13, 1.0, dp(y_ref), None) == 0
    core.SUNAMD_MassActionFree(big)
    core.SUNAMD_MassActionFree(base)
    assert st.flag == 0
    assert abs(st.nst - ost[0]) <= 0.1 * ost[0]
    err = lambda v: np.sqrt(np.mean(((v - y_ref) / (1e-6 * np.abs(y_ref) + 1e-10)) ** 2))
    e_gpu, e_cpu = err(tiles[0]), err(y_cpu)
    assert e_cpu <= 20.0 and e_gpu <= 20.0, (e_gpu, e_cpu)
    # the same with every integrator decision on the device (one launch)
    _say("coupled: the same, device-stepped")
    base_again = core.SUNAMD_MassActionCreate(m, R, nbase, 1002, 0.2)
    big = core.SUNAMD_MassActionTile(base_again, ncells)
    r = hip.SUNAMD_CoupledCreate_Mi355(big, 1e-6, 1e-10, 0, 0)
    assert r and hip.SUNAMD_CoupledSetDeviceStepping_Mi355(r, 1) == 0
    sd = _lib.SolveStats()
    assert hip.SUNAMD_CoupledRun_Mi355(r, 1.0, C.byref(sd)) == 0
    assert hip.SUNAMD_CoupledGetState_Mi355(r, dp(y)) == 0
    hip.SUNAMD_CoupledFree_Mi355(r)
    core.SUNAMD_MassActionFree(big)
    core.SUNAMD_MassActionFree(base_again)
    _say("coupled device-stepped: nst=%d nni=%d" % (sd.nst, sd.nni))
    tiles = y.reshape(-1, nbase * m)
    assert all(np.array_equal(tiles[0], t) for t in tiles[1:])
    assert abs(sd.nst - ost[0]) <= 0.1 * ost[0]
    assert err(tiles[0]) <= 20.0, err(tiles[0])
