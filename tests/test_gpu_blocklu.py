"""Parity of the block-diagonal SUNMatrix and batched LU against the oracle.

FP64: factors, pivots and solutions are bit-identical to the reference
GETRF / GETRS applied block by block (oracle_block_getrf / _getrs), on
matrices with forced pivoting (rand/cols + anti-identity,
test_sunlinsol_dense.c:93-135) and with tied pivot candidates.  FP32: the
solve residual |A x - b| is at single-precision level.  Matrix ops
(Zero, Copy, ScaleAddI, ScaleAdd, Matvec) are bit-identical to the dense
matrix's element formulas.  A zero pivot makes setup return
SUNLS_LUFACT_FAIL with LastFlag = first failing global column + 1.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib
from helpers import dp, gpu_get, gpu_vec

pytestmark = pytest.mark.gpu

SUNLS_LUFACT_FAIL = 808


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


def _blocks(rng, nb, m, ties=True):
    A = rng.uniform(0, 1, (nb, m, m)) / m
    A[:, np.arange(m), m - 1 - np.arange(m)] += 1.0
    A *= rng.uniform(0.5, 2.0, (nb, 1, 1))
    A[::3] *= -1.0
    if ties and m > 2:
        A[1::4, :, 0] = A[1::4, 0:1, 0]          # whole first column tied
        A[2::5, 1:, 1] = -A[2::5, 1:2, 1]       # ties of opposite sign
    return A


def _to_blockmajor(A):
    # (nb, m, m) row-major numpy -> block-contiguous column-major
    return np.ascontiguousarray(np.transpose(A, (0, 2, 1))).ravel()


def _from_blockmajor(a, nb, m):
    return np.transpose(a.reshape(nb, m, m), (0, 2, 1))


def _gpu_matrix(hip, ctx, y, A):
    nb, m, _ = A.shape
    M = hip.SUNMatrix_Mi355BlockDense(nb, m, hip.N_VGetStream_Mi355(y), ctx)
    h = hip.SUNMatrix_Mi355BlockDense_HostData(M)
    flat = _to_blockmajor(A)
    C.memmove(h, dp(flat), 8 * flat.size)
    assert hip.SUNMatrix_Mi355BlockDense_CopyToDevice(M) == 0
    return M


def _gpu_matrix_values(hip, M, nb, m):
    assert hip.SUNMatrix_Mi355BlockDense_CopyFromDevice(M) == 0
    h = hip.SUNMatrix_Mi355BlockDense_HostData(M)
    return np.ctypeslib.as_array(h, shape=(nb * m * m,)).copy()


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 11, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("nb", [1, 37])
def test_fp64_factor_and_solve_bitwise(libs, ctx, m, nb):
    core, hip, orc = libs
    rng = np.random.default_rng(1000 * m + nb)
    A = _blocks(rng, nb, m)
    n = nb * m
    b = rng.uniform(-1, 1, n)
    y = gpu_vec(ctx, np.zeros(n))
    M = _gpu_matrix(hip, ctx, y, A)
    LS = hip.SUNLinSol_Mi355BlockLU(y, M, 0, ctx)
    assert LS
    assert core.SUNLinSolInitialize(LS) == 0
    assert core.SUNLinSolSetup(LS, M) == 0
    F = _gpu_matrix_values(hip, M, nb, m)
    piv = np.zeros(n, np.int32)
    assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
    gb = gpu_vec(ctx, b)
    gx = gpu_vec(ctx, np.zeros(n))
    assert core.SUNLinSolSolve(LS, M, gx, gb, 0.0) == 0
    x = gpu_get(gx)
    # in-place solve (x == b) gives the same answer
    assert core.SUNLinSolSolve(LS, M, gb, gb, 0.0) == 0
    assert np.array_equal(gpu_get(gb), x)
    for k in range(nb):
        a = _to_blockmajor(A[k:k + 1]).copy()
        p = np.zeros(m, np.int64)
        assert orc.oracle_block_getrf(dp(a), m, p.ctypes.data_as(C.POINTER(C.c_int64))) == 0
        assert np.array_equal(F[k * m * m:(k + 1) * m * m], a), ("factor", k)
        assert np.array_equal(piv[k * m:(k + 1) * m], p), ("pivots", k)
        bb = b[k * m:(k + 1) * m].copy()
        orc.oracle_block_getrs(dp(a), m, p.ctypes.data_as(C.POINTER(C.c_int64)), dp(bb))
        assert np.array_equal(x[k * m:(k + 1) * m], bb), ("solve", k)
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(M)
    for v in (y, gb, gx):
        core.N_VDestroy(v)


def _setup_run(core, hip, ctx, J, prec, save_to_J, fused, c):
    """The Newton-matrix setup from J: fused (one call) or as the three
    separate calls; returns (flag, what A holds, pivots, solution of one
    right-hand side, what the saved-Jacobian matrix holds)."""
    nb, m, _ = J.shape
    n = nb * m
    y = gpu_vec(ctx, np.zeros(n))
    # save_to_J: A holds the fresh Jacobian; else the saved matrix holds it
    MA = _gpu_matrix(hip, ctx, y, J if save_to_J else np.zeros_like(J))
    MJ = _gpu_matrix(hip, ctx, y, np.zeros_like(J) if save_to_J else J)
    LS = hip.SUNLinSol_Mi355BlockLU(y, MA, prec, ctx)
    assert core.SUNLinSolInitialize(LS) == 0
    if fused:
        flag = hip.SUNLinSol_Mi355BlockLUSetupScaled(LS, MA, MJ, c, save_to_J)
    else:
        assert core.SUNMatCopy(MA, MJ) == 0 if save_to_J else core.SUNMatCopy(MJ, MA) == 0
        assert core.SUNMatScaleAddI(c, MA) == 0
        flag = core.SUNLinSolSetup(LS, MA)
    a = _gpu_matrix_values(hip, MA, nb, m)
    j = _gpu_matrix_values(hip, MJ, nb, m)
    piv = np.zeros(n, np.int32)
    assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
    x = None
    if flag == 0:
        b = np.random.default_rng(5).uniform(-1, 1, n)
        gb, gx = gpu_vec(ctx, b), gpu_vec(ctx, np.zeros(n))
        assert core.SUNLinSolSolve(LS, MA, gx, gb, 0.0) == 0
        x = gpu_get(gx)
        for v in (gb, gx):
            core.N_VDestroy(v)
    core.SUNLinSolFree(LS)
    for M in (MA, MJ):
        core.SUNMatDestroy(M)
    core.N_VDestroy(y)
    return flag, a, piv, x, j


@pytest.mark.parametrize("m", [1, 3, 4, 8, 10, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("save_to_J", [0, 1])
def test_fused_newton_matrix_setup(libs, ctx, m, prec, save_to_J):
    """SUNLinSol_Mi355BlockLUSetupScaled (copy + ScaleAddI + factorisation in
    one pass) leaves what the three separate calls leave, bit for bit: the
    solver matrix (FP64: the factors; FP32: c J + I), the saved Jacobian,
    the pivots and the solution of a right-hand side; in FP64 that is also
    the oracle's GETRF of c J + I block by block."""
    core, hip, orc = libs
    nb = 67
    rng = np.random.default_rng(17 * m + 3 * prec + save_to_J)
    J = rng.uniform(-1, 1, (nb, m, m)) * 40.0 / m
    c = -0.0371
    ref = _setup_run(core, hip, ctx, J, prec, save_to_J, False, c)
    got = _setup_run(core, hip, ctx, J, prec, save_to_J, True, c)
    assert ref[0] == 0 and got[0] == 0
    for k, (r, g) in enumerate(zip(ref[1:], got[1:])):
        assert np.array_equal(r, g), ("A", "pivots", "solution", "saved J")[k]
    assert np.array_equal(got[4], _to_blockmajor(J))
    if prec == 0:
        for k in range(0, nb, 11):
            a = _to_blockmajor(J[k:k + 1]).copy() * c
            a[::m + 1] += 1.0
            p = np.zeros(m, np.int64)
            assert orc.oracle_block_getrf(dp(a), m, p.ctypes.data_as(C.POINTER(C.c_int64))) == 0
            assert np.array_equal(got[1][k * m * m:(k + 1) * m * m], a), k


def test_fused_setup_declines(libs, ctx):
    """Sizes without a register kernel, and J == A, decline before doing
    anything (SUN_ERR_ARG_INCOMPATIBLE); a zero pivot reports like setup."""
    core, hip, orc = libs
    SUN_ERR_ARG_INCOMPATIBLE = -9998  # SUN_ERR_MINIMUM (-10000) + 2
    for m in (11, 20):
        J = np.ones((3, m, m))
        y = gpu_vec(ctx, np.zeros(3 * m))
        MA, MJ = _gpu_matrix(hip, ctx, y, J), _gpu_matrix(hip, ctx, y, 2 * J)
        LS = hip.SUNLinSol_Mi355BlockLU(y, MA, 0, ctx)
        assert hip.SUNLinSol_Mi355BlockLUSetupScaled(LS, MA, MJ, 1.0, 0) == SUN_ERR_ARG_INCOMPATIBLE
        assert hip.SUNLinSol_Mi355BlockLUSetupScaled(LS, MA, MA, 1.0, 0) == SUN_ERR_ARG_INCOMPATIBLE
        assert np.array_equal(_gpu_matrix_values(hip, MA, 3, m), _to_blockmajor(J))
        core.SUNLinSolFree(LS)
        for M in (MA, MJ):
            core.SUNMatDestroy(M)
        core.N_VDestroy(y)
    # c J + I singular in block 1: c = -1, J = I there
    m = 8
    J = np.tile(np.eye(m) * 3.0, (4, 1, 1))
    J[1] = np.eye(m)
    flag, a, piv, x, j = _setup_run(core, hip, ctx, J, 0, 0, True, -1.0)
    assert flag == SUNLS_LUFACT_FAIL


@pytest.mark.parametrize("m", [1, 3, 8, 12, 24, 32, 48, 64])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("scale", [1.0, 0.8125, 1.3719])
def test_fused_scaled_solve(libs, ctx, m, prec, scale):
    """SUNLinSol_Mi355BlockLUSolveScaled (b solved in place, written as
    scale * x) leaves b as CVODE's calls do -- N_VConst(0, x), SUNLinSolSolve,
    N_VScale(1, x, b), N_VScale(scale, b, b) -- bit for bit; in FP64 that is
    the oracle's GETRS of each block times scale."""
    core, hip, orc = libs
    nb = 71
    n = nb * m
    rng = np.random.default_rng(31 * m + prec)
    J = rng.uniform(-1, 1, (nb, m, m)) * 4.0 / m + np.eye(m)
    y = gpu_vec(ctx, np.zeros(n))
    MA = _gpu_matrix(hip, ctx, y, J)
    LS = hip.SUNLinSol_Mi355BlockLU(y, MA, prec, ctx)
    assert core.SUNLinSolInitialize(LS) == 0
    assert core.SUNLinSolSetup(LS, MA) == 0
    b = rng.uniform(-1, 1, n)
    gb, gx, gr = gpu_vec(ctx, b), gpu_vec(ctx, np.ones(n)), gpu_vec(ctx, b)
    core.N_VConst(0.0, gx)
    assert core.SUNLinSolSolve(LS, MA, gx, gr, 0.0) == 0
    core.N_VScale(1.0, gx, gr)
    if scale != 1.0:
        core.N_VScale(scale, gr, gr)
    assert hip.SUNLinSol_Mi355BlockLUSolveScaled(LS, MA, gb, scale) == 0
    got, ref = gpu_get(gb), gpu_get(gr)
    assert np.array_equal(got, ref)
    if prec == 0:
        F = _gpu_matrix_values(hip, MA, nb, m)
        piv = np.zeros(n, np.int32)
        assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
        for k in range(0, nb, 7):
            bb = b[k * m:(k + 1) * m].copy()
            p = piv[k * m:(k + 1) * m].astype(np.int64)
            orc.oracle_block_getrs(dp(F[k * m * m:(k + 1) * m * m].copy()), m,
                                   p.ctypes.data_as(C.POINTER(C.c_int64)), dp(bb))
            assert np.array_equal(got[k * m:(k + 1) * m], scale * bb), k
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(MA)
    for v in (y, gb, gx, gr):
        core.N_VDestroy(v)


def test_fused_scaled_solve_declines(libs, ctx):
    """Sizes without a register kernel decline before doing anything."""
    core, hip, orc = libs
    SUN_ERR_ARG_INCOMPATIBLE = -9998
    m, nb = 11, 5
    J = np.tile(np.eye(m) * 2.0, (nb, 1, 1))
    y = gpu_vec(ctx, np.zeros(nb * m))
    MA = _gpu_matrix(hip, ctx, y, J)
    LS = hip.SUNLinSol_Mi355BlockLU(y, MA, 0, ctx)
    assert core.SUNLinSolInitialize(LS) == 0
    assert core.SUNLinSolSetup(LS, MA) == 0
    b = np.arange(nb * m, dtype=np.float64)
    gb = gpu_vec(ctx, b)
    assert hip.SUNLinSol_Mi355BlockLUSolveScaled(LS, MA, gb, 1.0) == SUN_ERR_ARG_INCOMPATIBLE
    assert np.array_equal(gpu_get(gb), b)
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(MA)
    for v in (y, gb):
        core.N_VDestroy(v)


@pytest.mark.parametrize("m", [3, 8, 16, 32, 48, 64])
def test_fp32_solve_residual(libs, ctx, m):
    core, hip, orc = libs
    nb = 129
    rng = np.random.default_rng(m)
    A = _blocks(rng, nb, m, ties=False)
    n = nb * m
    xt = rng.uniform(-1, 1, n)
    b = np.einsum("kij,kj->ki", A, xt.reshape(nb, m)).ravel()
    y = gpu_vec(ctx, np.zeros(n))
    M = _gpu_matrix(hip, ctx, y, A)
    LS = hip.SUNLinSol_Mi355BlockLU(y, M, 1, ctx)
    assert core.SUNLinSolSetup(LS, M) == 0
    # FP32 mode leaves the FP64 matrix untouched
    assert np.array_equal(_gpu_matrix_values(hip, M, nb, m), _to_blockmajor(A))
    gb, gx = gpu_vec(ctx, b), gpu_vec(ctx, np.zeros(n))
    assert core.SUNLinSolSolve(LS, M, gx, gb, 0.0) == 0
    x = gpu_get(gx)
    err = np.abs(x - xt).max()
    assert err < 2e-4 * m, err
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(M)
    for v in (y, gb, gx):
        core.N_VDestroy(v)


@pytest.mark.parametrize("m", [32, 64])
def test_fp32_mfma_factorisation(libs, ctx, m):
    """The blocked FP32 factorisation with MFMA trailing updates
    (SUNLinSol_Mi355BlockLUSetMFMA) against the register-row FP32 kernel on
    the same matrices (forced pivoting, no near-ties): the same pivot rows,
    factors that agree to single precision, and solutions at the same
    single-precision level; a zero pivot is reported the same way."""
    core, hip, orc = libs
    nb = 97
    rng = np.random.default_rng(100 + m)
    A = _blocks(rng, nb, m, ties=False)
    n = nb * m
    xt = rng.uniform(-1, 1, n)
    b = np.einsum("kij,kj->ki", A, xt.reshape(nb, m)).ravel()
    y = gpu_vec(ctx, np.zeros(n))
    M = _gpu_matrix(hip, ctx, y, A)
    out = {}
    for mf in (0, 1):
        LS = hip.SUNLinSol_Mi355BlockLU(y, M, 1, ctx)
        assert hip.SUNLinSol_Mi355BlockLUSetMFMA(LS, mf) == 0
        assert core.SUNLinSolSetup(LS, M) == 0
        piv = np.zeros(n, dtype=np.int32)
        assert hip.SUNLinSol_Mi355BlockLU_GetPivots(LS, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
        gb, gx = gpu_vec(ctx, b), gpu_vec(ctx, np.zeros(n))
        assert core.SUNLinSolSolve(LS, M, gx, gb, 0.0) == 0
        out[mf] = (piv, gpu_get(gx))
        core.SUNLinSolFree(LS)
        for v in (gb, gx):
            core.N_VDestroy(v)
    assert np.array_equal(out[0][0], out[1][0])
    err_reg = np.abs(out[0][1] - xt).max()
    err_mfma = np.abs(out[1][1] - xt).max()
    print("m=%d: max |x - x_true| register %.3g, MFMA %.3g" % (m, err_reg, err_mfma))
    assert err_mfma < 2e-4 * m
    assert np.abs(out[1][1] - out[0][1]).max() < 2e-4 * m
    # the switch applies to FP32 solvers of m = 32 or 64 only
    LS64 = hip.SUNLinSol_Mi355BlockLU(y, M, 0, ctx)
    assert hip.SUNLinSol_Mi355BlockLUSetMFMA(LS64, 1) != 0
    core.SUNLinSolFree(LS64)
    # zero pivot: the first failing global column, as the register kernel reports it
    A[5, :, 3] = 0.0
    A[9, :, 1] = 0.0
    M2 = _gpu_matrix(hip, ctx, y, A)
    LS = hip.SUNLinSol_Mi355BlockLU(y, M2, 1, ctx)
    assert hip.SUNLinSol_Mi355BlockLUSetMFMA(LS, 1) == 0
    assert core.SUNLinSolSetup(LS, M2) == SUNLS_LUFACT_FAIL
    assert core.SUNLinSolLastFlag(LS) == 5 * m + 3 + 1
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(M2)
    core.SUNMatDestroy(M)
    core.N_VDestroy(y)


@pytest.mark.parametrize("prec", [0, 1])
def test_zero_pivot_reported(libs, ctx, prec):
    core, hip, orc = libs
    nb, m = 6, 4
    rng = np.random.default_rng(3)
    A = _blocks(rng, nb, m, ties=False)
    A[4, :, 2] = 0.0          # block 4, column 2 singular
    A[5, :, 1] = 0.0          # block 5, column 1 singular
    y = gpu_vec(ctx, np.zeros(nb * m))
    M = _gpu_matrix(hip, ctx, y, A)
    LS = hip.SUNLinSol_Mi355BlockLU(y, M, prec, ctx)
    assert core.SUNLinSolSetup(LS, M) == SUNLS_LUFACT_FAIL
    # first zero pivot in the global column order: block 4, column 2
    assert core.SUNLinSolLastFlag(LS) == 4 * m + 2 + 1
    core.SUNLinSolFree(LS)
    core.SUNMatDestroy(M)
    core.N_VDestroy(y)


def test_matrix_ops_bitwise(libs, ctx):
    core, hip, orc = libs
    nb, m = 33, 5
    rng = np.random.default_rng(8)
    A = rng.uniform(-1, 1, (nb, m, m))
    B = rng.uniform(-1, 1, (nb, m, m))
    y = gpu_vec(ctx, np.zeros(nb * m))
    MA = _gpu_matrix(hip, ctx, y, A)
    MB = _gpu_matrix(hip, ctx, y, B)
    c = -0.3125
    assert core.SUNMatScaleAddI(c, MA) == 0
    ref = _to_blockmajor(A) * c
    diag = (np.arange(nb * m * m) % (m * m)) % (m + 1) == 0
    ref[diag] = ref[diag] + 1.0
    assert np.array_equal(_gpu_matrix_values(hip, MA, nb, m), ref)
    assert core.SUNMatScaleAdd(2.5, MA, MB) == 0
    ref = 2.5 * ref + _to_blockmajor(B)
    assert np.array_equal(_gpu_matrix_values(hip, MA, nb, m), ref)
    assert core.SUNMatCopy(MB, MA) == 0
    assert np.array_equal(_gpu_matrix_values(hip, MA, nb, m), _to_blockmajor(B))
    xs = rng.uniform(-1, 1, nb * m)
    gx, gy = gpu_vec(ctx, xs), gpu_vec(ctx, np.zeros(nb * m))
    assert core.SUNMatMatvec(MA, gx, gy) == 0
    ref = np.zeros(nb * m)
    for k in range(nb):  # column order of the dense matvec
        for j in range(m):
            ref[k * m:(k + 1) * m] = ref[k * m:(k + 1) * m] + B[k, :, j] * xs[k * m + j]
    assert np.array_equal(gpu_get(gy), ref)
    assert core.SUNMatZero(MA) == 0
    assert not _gpu_matrix_values(hip, MA, nb, m).any()
    for M in (MA, MB):
        core.SUNMatDestroy(M)
    for v in (y, gx, gy):
        core.N_VDestroy(v)
