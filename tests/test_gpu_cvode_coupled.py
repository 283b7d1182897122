"""Coupled BDF integration on the HIP drop-ins against the CPU oracle.

The same integrator code drives both runs; only the N_Vector, the block
matrix / LU and the problem kernels differ (HIP vs serial + dense LU + C).
Parity bar: identical step, Newton-iteration, setup, Jacobian and failure
counts, final state within 1e-10 relative (the bar of the plain FP64 path).
Also: the mass-action RHS and Jacobian kernels equal the CPU evaluation bit
for bit, and a large tiled batch (every cell a copy of a small base set)
reproduces the base set's oracle run -- a size-independent check at a
problem size the oracle cannot integrate in test time.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib
from helpers import dp

pytestmark = pytest.mark.gpu

PM = C.POINTER(_lib.MassAction)


@pytest.fixture(scope="module")
def libs():
    import torch

    assert torch.cuda.is_available()
    return _lib.core(), _lib.hip(), _lib.load_oracle()


def _prob(core, m, R, ncells, seed):
    p = core.SUNAMD_MassActionCreate(m, R, ncells, seed, 0.2)
    assert p
    return p


def _dev_array(n):
    import torch

    return torch.zeros(n, dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("m,R,ncells", [(4, 6, 33), (8, 16, 257), (32, 48, 65), (64, 96, 7),
                                        (3, 5, 1001), (24, 36, 130), (48, 72, 9),
                                        (4, 6, 600000)])
def test_rhs_and_jacobian_bitwise(libs, m, R, ncells):
    """Device RHS and block Jacobian bit for bit against the serial evaluation:
    power-of-two and other sizes (idle lanes in a group), cell counts that
    are not a multiple of the groups per workgroup, and 600,000 cells (more
    than one cell per group: the workgroups stride over the cells)."""
    import torch

    core, hip, orc = libs
    p = _prob(core, m, R, ncells, 77 + m)
    pc = p.contents
    y = np.ctypeslib.as_array(pc.y0, shape=(ncells * m,)).copy()
    y *= np.random.default_rng(m).uniform(0.5, 1.5, y.size)
    f_ref = np.zeros(ncells * m)
    J_ref = np.zeros(ncells * m * m)
    core.SUNAMD_MassActionRhsHost(p, dp(y), dp(f_ref), 0, ncells)
    core.SUNAMD_MassActionJacHost(p, dp(y), dp(J_ref), 0, ncells)
    stream = torch.cuda.current_stream().cuda_stream
    dev = hip.SUNAMD_MassActionToDevice(p, C.c_void_p(stream))
    assert dev
    ty = torch.from_numpy(y).cuda()
    tf = _dev_array(ncells * m)
    tJ = torch.full((ncells * m * m,), 7.0, dtype=torch.float64, device="cuda")
    assert hip.SUNAMD_MassActionRhsDevArrays(dev, C.c_void_p(ty.data_ptr()),
                                             C.c_void_p(tf.data_ptr())) == 0
    assert hip.SUNAMD_MassActionJacDevArrays(dev, C.c_void_p(ty.data_ptr()),
                                             C.c_void_p(tJ.data_ptr())) == 0
    torch.cuda.synchronize()
    assert np.array_equal(tf.cpu().numpy(), f_ref)
    assert np.array_equal(tJ.cpu().numpy(), J_ref)
    # mass is conserved by every reaction: columns of J sum to ~0
    hip.SUNAMD_MassActionDevFree(dev)
    core.SUNAMD_MassActionFree(p)


SEQ, TREE = 1, 0


def _solve_both(libs, p, n, prec=0, order=SEQ, rtol=1e-6, atol=1e-10):
    core, hip, orc = libs
    y_gpu, y_cpu = np.zeros(n), np.zeros(n)
    st = _lib.SolveStats()
    assert hip.SUNAMD_MassActionSolveCoupled_Mi355(p, rtol, atol, 1.0, prec, order, dp(y_gpu),
                                                   C.byref(st)) == 0
    ost = np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, rtol, atol, 1.0, dp(y_cpu), dp(ost)) == 0
    return y_gpu, st, y_cpu, ost


def _counts(st):
    return (st.nst, st.nfe, st.nsetups, st.netf, st.nni, st.nnf, st.ncfn, st.nje)


def _wrms_err(y, yref, rtol=1e-6, atol=1e-10):
    return np.sqrt(np.mean(((y - yref) / (rtol * np.abs(yref) + atol)) ** 2))


@pytest.mark.parametrize("m,R,ncells,seed", [(8, 16, 4096, 1001), (8, 16, 64, 7),
                                              (32, 48, 128, 1002), (3, 4, 1, 5)])
def test_coupled_fp64_parity_sequential(libs, m, R, ncells, seed):
    """Reductions in the serial vector's order: the whole integration is
    bit-identical to the oracle (every counter and every final value)."""
    core = libs[0]
    p = _prob(core, m, R, ncells, seed)
    y_gpu, st, y_cpu, ost = _solve_both(libs, p, ncells * m, order=SEQ)
    core.SUNAMD_MassActionFree(p)
    assert st.flag == 0 and ost[14] == 0
    assert _counts(st) == tuple(int(v) for v in ost[:8])
    assert (st.qlast, st.qcur) == (int(ost[8]), int(ost[9]))
    assert (st.hinused, st.hlast, st.tcur) == (ost[10], ost[11], ost[13])
    assert np.array_equal(y_gpu, y_cpu)
    # each cell's mass (sum 1) is conserved to solver tolerance
    sums = y_gpu.reshape(ncells, m).sum(axis=1)
    assert np.abs(sums - 1.0).max() < 1e-6


@pytest.mark.parametrize("m,R,ncells,seed", [(8, 16, 1024, 1001), (32, 48, 128, 1002)])
def test_coupled_fp64_tree_reductions(libs, m, R, ncells, seed):
    """Full-bandwidth tree reductions: the step controller of this stiff
    problem amplifies ulp-level norm differences into a different step
    sequence (the oracle itself moves by ~1-6% in nst when its WRMS is summed
    pairwise instead of left to right), so the run is checked for accuracy:
    its error against a reference solution computed at 1000x tighter
    tolerances is no larger than twice the oracle's own error."""
    core, orc = libs[0], libs[2]
    p = _prob(core, m, R, ncells, seed)
    y_gpu, st, y_cpu, ost = _solve_both(libs, p, ncells * m, order=TREE)
    y_ref = np.zeros(ncells * m)
    assert orc.oracle_massaction_solve_coupled(p, 1e-9, 1e-13, 1.0, dp(y_ref), None) == 0
    core.SUNAMD_MassActionFree(p)
    assert st.flag == 0
    assert abs(st.nst - ost[0]) <= 0.1 * ost[0]
    e_gpu, e_cpu = _wrms_err(y_gpu, y_ref), _wrms_err(y_cpu, y_ref)
    assert e_gpu <= 2.0 * e_cpu + 0.1, (e_gpu, e_cpu)


def test_coupled_tiled_matches_base(libs):
    """2^15 cells tiled from 128 base cells: every tile integrates like the
    base set (WRMS norms of the tiled vector equal the base set's up to
    rounding), and its global error against a reference solution of the 128
    base cells at 1000x tighter tolerances stays at the level of the
    tolerance.  The bound is 20 rtol-weighted units: the oracle's own error on
    tilings of the same cells ranges from 0.85 to 5.7 units (2 to 256 tiles,
    sequential sums; the rounding of the norms alone changes its step
    sequence), so comparing one run's error with another's is not stable."""
    core = libs[0]
    hip, orc = libs[1], libs[2]
    m, R = 8, 16
    base = _prob(core, m, R, 128, 1001)
    big = core.SUNAMD_MassActionTile(base, 1 << 15)
    n_big = (1 << 15) * m
    y_gpu = np.zeros(n_big)
    st = _lib.SolveStats()
    assert hip.SUNAMD_MassActionSolveCoupled_Mi355(big, 1e-6, 1e-10, 1.0, 0, TREE, dp(y_gpu),
                                                   C.byref(st)) == 0
    y_cpu = np.zeros(128 * m)
    ost = np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(base, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    y_ref = np.zeros(128 * m)
    assert orc.oracle_massaction_solve_coupled(base, 1e-9, 1e-13, 1.0, dp(y_ref), None) == 0
    core.SUNAMD_MassActionFree(big)
    core.SUNAMD_MassActionFree(base)
    tiles = y_gpu.reshape(-1, 128 * m)
    # every tile is bit-identical (cells never interact except through norms)
    assert all(np.array_equal(tiles[0], t) for t in tiles[1:])
    assert abs(st.nst - ost[0]) <= 0.1 * ost[0]
    e_gpu, e_cpu = _wrms_err(tiles[0], y_ref), _wrms_err(y_cpu, y_ref)
    assert e_cpu <= 20.0 and e_gpu <= 20.0, (e_gpu, e_cpu)


def test_coupled_fp32_lu_accuracy(libs):
    """FP32 block LU inside the FP64 Newton iteration: final state within
    rtol of the FP64 oracle (modified Newton acts as iterative refinement)."""
    core = libs[0]
    m, R, ncells = 8, 16, 512
    p = _prob(core, m, R, ncells, 1001)
    hip, orc = libs[1], libs[2]
    y32 = np.zeros(ncells * m)
    st = _lib.SolveStats()
    assert hip.SUNAMD_MassActionSolveCoupled_Mi355(p, 1e-6, 1e-10, 1.0, 1, TREE, dp(y32),
                                                   C.byref(st)) == 0
    y_cpu, ost = np.zeros(ncells * m), np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    y_ref = np.zeros(ncells * m)
    assert orc.oracle_massaction_solve_coupled(p, 1e-9, 1e-13, 1.0, dp(y_ref), None) == 0
    core.SUNAMD_MassActionFree(p)
    assert st.flag == 0
    e32, e_cpu = _wrms_err(y32, y_ref), _wrms_err(y_cpu, y_ref)
    assert e32 <= 2.0 * e_cpu + 0.1, (e32, e_cpu)


@pytest.mark.parametrize("order", [SEQ, TREE])
@pytest.mark.parametrize("m,R,ncells", [(8, 16, 257), (3, 4, 33)])
def test_coupled_fused_kernels_change_nothing(libs, order, m, R, ncells):
    """CVODE's fused error-weight and residual kernels (the counterparts of
    cvode_fused_gpu.cpp) and the fused Nordsieck predictor / undo form every
    element with the roundings of the op sequences they replace: with them
    on, the integration is bit-identical to the unfused run and, in
    sequential order, to the CPU oracle (every counter, the step sizes and
    the final state).  3 x 33 = 99 unknowns exercises the odd tail element of
    the paired kernels."""
    core, hip, orc = libs
    p = _prob(core, m, R, ncells, 1001)
    n = ncells * m
    y_cpu, ost = np.zeros(n), np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    runs = []
    for fused in (0, 1):
        r = hip.SUNAMD_CoupledCreate_Mi355(p, 1e-6, 1e-10, 0, order)
        assert r
        assert hip.SUNAMD_CoupledSetFusedKernels_Mi355(r, fused) == 0
        st = _lib.SolveStats()
        assert hip.SUNAMD_CoupledRun_Mi355(r, 1.0, C.byref(st)) == 0
        y = np.zeros(n)
        assert hip.SUNAMD_CoupledGetState_Mi355(r, dp(y)) == 0
        runs.append((_counts(st), st.launches, y, (st.hinused, st.hlast, st.tcur)))
        hip.SUNAMD_CoupledFree_Mi355(r)
    core.SUNAMD_MassActionFree(p)
    (c0, l0, y0, h0), (c1, l1, y1, h1) = runs
    assert c0 == c1 and h0 == h1
    assert np.array_equal(y0, y1)
    assert l1 < l0  # fewer launches
    if order == SEQ:
        assert c1 == tuple(int(v) for v in ost[:8])
        assert h1 == (ost[10], ost[11], ost[13])
        assert np.array_equal(y1, y_cpu)


def _run_coupled(hip, p, n, order, prec=0, fused=0, devstep=0, tout=1.0, runs=1):
    r = hip.SUNAMD_CoupledCreate_Mi355(p, 1e-6, 1e-10, prec, order)
    assert r
    if fused:
        assert hip.SUNAMD_CoupledSetFusedKernels_Mi355(r, 1) == 0
    if devstep:
        assert hip.SUNAMD_CoupledSetDeviceStepping_Mi355(r, 1) == 0
    out = []
    for _ in range(runs):
        st = _lib.SolveStats()
        rc = hip.SUNAMD_CoupledRun_Mi355(r, tout, C.byref(st))
        y = np.zeros(n)
        assert hip.SUNAMD_CoupledGetState_Mi355(r, dp(y)) == 0
        out.append((rc, st, y))
    hip.SUNAMD_CoupledFree_Mi355(r)
    return out


@pytest.mark.parametrize("m,R,ncells,seed", [(8, 16, 4096, 1001), (3, 4, 33, 5),
                                              (32, 48, 128, 1002), (16, 24, 300, 9),
                                              (24, 36, 100, 7), (64, 96, 7, 3),
                                              (48, 72, 50, 4)])
def test_device_stepping_bitwise_sequential(libs, m, R, ncells, seed):
    """Device stepping: the whole integration in one cooperative launch, with
    Newton convergence, Jacobian reuse, the error test and the step-size and
    order choices decided on the device.  With WRMS norms in the serial
    order the run equals the host-stepped run and the CPU oracle bit for
    bit -- every counter, the step sizes and the final state -- and a second
    run of the same runner repeats it exactly.  (8 x 4096 spreads the cells
    over 128 workgroups: the norms' partials are read across the grid.)"""
    core, hip, orc = libs
    p = _prob(core, m, R, ncells, seed)
    n = ncells * m
    (rh, sh, yh), = _run_coupled(hip, p, n, SEQ)
    d1, d2 = _run_coupled(hip, p, n, SEQ, devstep=1, runs=2)
    y_cpu, ost = np.zeros(n), np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    core.SUNAMD_MassActionFree(p)
    assert rh == 0 and sh.flag == 0
    for rd, sd, yd in (d1, d2):
        assert rd == 0 and sd.flag == 0
        assert _counts(sd) == _counts(sh) == tuple(int(v) for v in ost[:8])
        assert (sd.qlast, sd.qcur) == (sh.qlast, sh.qcur)
        assert (sd.hinused, sd.hlast, sd.hcur, sd.tcur) == (sh.hinused, sh.hlast, sh.hcur,
                                                            sh.tcur)
        assert np.array_equal(yd, yh)
        assert np.array_equal(yd, y_cpu)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("m,R,ncells,seed", [(8, 16, 1024, 1001), (32, 48, 128, 1002)])
def test_device_stepping_tree_accuracy(libs, prec, m, R, ncells, seed):
    """Device stepping with tree-ordered norms (the order is fixed by the
    launch shape, so runs repeat exactly) and FP64 or FP32 factors: the error
    against a reference solution at 1000x tighter tolerances is within twice
    the oracle's own error (+0.1), the bar of the host-stepped tree test."""
    core, hip, orc = libs
    p = _prob(core, m, R, ncells, seed)
    n = ncells * m
    d1, d2 = _run_coupled(hip, p, n, TREE, prec=prec, devstep=1, runs=2)
    y_cpu, ost = np.zeros(n), np.zeros(16)
    assert orc.oracle_massaction_solve_coupled(p, 1e-6, 1e-10, 1.0, dp(y_cpu), dp(ost)) == 0
    y_ref = np.zeros(n)
    assert orc.oracle_massaction_solve_coupled(p, 1e-9, 1e-13, 1.0, dp(y_ref), None) == 0
    core.SUNAMD_MassActionFree(p)
    (r1, s1, y1), (r2, s2, y2) = d1, d2
    assert r1 == 0 and r2 == 0
    assert _counts(s1) == _counts(s2) and np.array_equal(y1, y2)
    assert abs(s1.nst - ost[0]) <= 0.1 * ost[0]
    e_dev, e_cpu = _wrms_err(y1, y_ref), _wrms_err(y_cpu, y_ref)
    assert e_dev <= 2.0 * e_cpu + 0.1, (e_dev, e_cpu)


def test_device_stepping_unsupported_size(libs):
    """Sizes without a register-LU kernel (here m = 9) are not offered by the
    device-stepped integrator: the run returns -100 without launching."""
    core, hip, _ = libs
    p = _prob(core, 9, 14, 7, 3)
    r = hip.SUNAMD_CoupledCreate_Mi355(p, 1e-6, 1e-10, 0, TREE)
    assert r
    assert hip.SUNAMD_CoupledSetDeviceStepping_Mi355(r, 1) == 0
    st = _lib.SolveStats()
    assert hip.SUNAMD_CoupledRun_Mi355(r, 1.0, C.byref(st)) == -100
    hip.SUNAMD_CoupledFree_Mi355(r)
    core.SUNAMD_MassActionFree(p)
