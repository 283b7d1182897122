"""Parity of the HIP N_Vector against the oracle serial vector.

Elementwise ops (LinearSum in all its special cases, Const, Prod, Div,
Scale, Abs, Inv, AddConst, Compare, InvTest, ConstrMask) and the fused /
array ops (LinearCombination, ScaleAddMulti, ScaleVectorArray) must be
bit-identical: the kernels evaluate the same per-element formula with no
fused multiply-add.  Reductions are tree-ordered on the GPU, so sums are
checked to a relative 1e-12 (min/max bit-exact), and exactly on the
test_nvector.c constant-fill known answers, where every partial sum is exact.
Sizes cover empty, 1, odd tails and multi-workgroup grids.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib
from helpers import cpu_get, cpu_vec, dp, gpu_get, gpu_vec, vec_array

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 63, 64, 1001, 65537]


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


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    w = rng.uniform(1e-3, 1, n)
    return x, y, w


@pytest.mark.parametrize("n", SIZES)
def test_linearsum_all_branches_bitwise(libs, ctx, n):
    core, hip, orc = libs
    x0, y0, _ = _data(n, n)
    cases = [(1, 1), (1, -1), (-1, 1), (-1, -1), (1, 0.3), (0.3, 1), (-1, 0.7),
             (0.7, -1), (2.5, 2.5), (2.5, -2.5), (0.3, 0.7), (0.0, 2.0)]
    for a, b in cases:
        for alias in ("none", "x", "y"):
            g = [gpu_vec(ctx, x0), gpu_vec(ctx, y0)]
            c = [cpu_vec(orc, ctx, x0), cpu_vec(orc, ctx, y0)]
            gz = {"none": gpu_vec(ctx, np.zeros(n)), "x": g[0], "y": g[1]}[alias]
            cz = {"none": cpu_vec(orc, ctx, np.zeros(n)), "x": c[0], "y": c[1]}[alias]
            core.N_VLinearSum(a, g[0], b, g[1], gz)
            core.N_VLinearSum(a, c[0], b, c[1], cz)
            assert np.array_equal(gpu_get(gz), cpu_get(cz)), (a, b, alias, n)
            for v in {id(v): v for v in g + [gz]}.values():
                core.N_VDestroy(v)
            for v in {id(v): v for v in c + [cz]}.values():
                core.N_VDestroy(v)


@pytest.mark.parametrize("n", SIZES)
def test_elementwise_bitwise(libs, ctx, n):
    core, hip, orc = libs
    x0, y0, w0 = _data(n, 100 + n)
    x0[::7] = 0.0  # zeros for InvTest / Compare
    gx, gy, gz = gpu_vec(ctx, x0), gpu_vec(ctx, w0), gpu_vec(ctx, np.zeros(n))
    cx, cy, cz = cpu_vec(orc, ctx, x0), cpu_vec(orc, ctx, w0), cpu_vec(orc, ctx, np.zeros(n))
    ops = [
        lambda X, Y, Z: core.N_VConst(0.37, Z),
        lambda X, Y, Z: core.N_VProd(X, Y, Z),
        lambda X, Y, Z: core.N_VDiv(X, Y, Z),
        lambda X, Y, Z: core.N_VScale(-3.1, X, Z),
        lambda X, Y, Z: core.N_VScale(1.0, X, Z),
        lambda X, Y, Z: core.N_VScale(-1.0, X, Z),
        lambda X, Y, Z: core.N_VAbs(X, Z),
        lambda X, Y, Z: core.N_VInv(Y, Z),
        lambda X, Y, Z: core.N_VAddConst(X, 0.125, Z),
        lambda X, Y, Z: core.N_VCompare(0.5, X, Z),
    ]
    for op in ops:
        op(gx, gy, gz)
        op(cx, cy, cz)
        assert np.array_equal(gpu_get(gz), cpu_get(cz))
    # InvTest: same mask result and same values where x != 0
    core.N_VConst(-9.0, gz)
    core.N_VConst(-9.0, cz)
    assert core.N_VInvTest(gx, gz) == core.N_VInvTest(cx, cz) == (0 if n > 0 else 1)
    assert np.array_equal(gpu_get(gz), cpu_get(cz))
    # ConstrMask with all constraint kinds
    cons = np.tile([0.0, 1.0, -1.0, 2.0, -2.0], n // 5 + 1)[:n]
    gc, cc = gpu_vec(ctx, cons), cpu_vec(orc, ctx, cons)
    assert core.N_VConstrMask(gc, gx, gz) == core.N_VConstrMask(cc, cx, cz)
    assert np.array_equal(gpu_get(gz), cpu_get(cz))
    # MinQuotient
    assert core.N_VMinQuotient(gx, gy) == core.N_VMinQuotient(cx, cy)
    for v in (gx, gy, gz, gc, cx, cy, cz, cc):
        core.N_VDestroy(v)


@pytest.mark.parametrize("n", SIZES)
def test_reductions(libs, ctx, n):
    core, hip, orc = libs
    x0, y0, w0 = _data(n, 200 + n)
    idm = (np.arange(n) % 3 == 0).astype(float)
    g = [gpu_vec(ctx, a) for a in (x0, y0, w0, idm)]
    c = [cpu_vec(orc, ctx, a) for a in (x0, y0, w0, idm)]
    pairs = [
        (lambda v: core.N_VDotProd(v[0], v[1]), 1e-12),
        (lambda v: core.N_VWrmsNorm(v[0], v[2]), 1e-12),
        (lambda v: core.N_VWrmsNormMask(v[0], v[2], v[3]), 1e-12),
        (lambda v: core.N_VWL2Norm(v[0], v[2]), 1e-12),
        (lambda v: core.N_VL1Norm(v[0]), 1e-12),
        (lambda v: core.N_VMaxNorm(v[0]), 0.0),
        (lambda v: core.N_VMin(v[1]), 0.0),
    ]
    for f, tol in pairs:
        a, b = f(g), f(c)
        if tol == 0.0:
            assert a == b
        else:
            assert a == pytest.approx(b, rel=tol, abs=1e-300)
    for v in g + c:
        core.N_VDestroy(v)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 20001])
def test_sequential_reductions_bitwise(libs, ctx, n):
    """SUNAMD_REDUCE_SEQUENTIAL sums left to right like the serial vector,
    so every sum-type reduction is bit-identical."""
    core, hip, orc = libs
    x0, y0, w0 = _data(n, 300 + n)
    idm = (np.arange(n) % 3 == 0).astype(float)
    g = [gpu_vec(ctx, a) for a in (x0, y0, w0, idm)]
    for v in g:
        assert hip.N_VSetReductionOrder_Mi355(v, 1) == 0
    c = [cpu_vec(orc, ctx, a) for a in (x0, y0, w0, idm)]
    fs = [lambda v: core.N_VDotProd(v[0], v[1]),
          lambda v: core.N_VWrmsNorm(v[0], v[2]),
          lambda v: core.N_VWrmsNormMask(v[0], v[2], v[3]),
          lambda v: core.N_VWL2Norm(v[0], v[2]),
          lambda v: core.N_VL1Norm(v[0])]
    for f in fs:
        assert f(g) == f(c)
    # clones inherit the order; DotProdMulti over up to 40 vectors
    Ys = [np.random.default_rng(k).uniform(-1, 1, n) for k in range(40)]
    gY = []
    for a in Ys:
        v = core.N_VClone(g[0])
        assert hip.N_VCopyFromHostArray_Mi355(v, dp(np.ascontiguousarray(a))) == 0
        gY.append(v)
    cY = [cpu_vec(orc, ctx, a) for a in Ys]
    dg, dc = np.zeros(40), np.zeros(40)
    assert core.N_VDotProdMulti(40, g[0], vec_array(gY), dp(dg)) == 0
    assert core.N_VDotProdMulti(40, c[0], vec_array(cY), dp(dc)) == 0
    assert np.array_equal(dg, dc)
    for v in g + c + gY + cY:
        core.N_VDestroy(v)


@pytest.mark.parametrize("n", [1, 1000, 65536])
def test_known_answers_exact(libs, ctx, n):
    """test_nvector.c closed forms hold exactly on the device."""
    core, hip, orc = libs
    x = gpu_vec(ctx, np.full(n, -0.5))
    w = gpu_vec(ctx, np.full(n, 0.5))
    assert core.N_VWrmsNorm(x, w) == 0.25
    x2 = gpu_vec(ctx, np.full(n, 2.0))
    Y = [gpu_vec(ctx, np.full(n, v)) for v in (0.5, -0.5, 1.0)]
    d = np.zeros(3)
    assert core.N_VDotProdMulti(3, x2, vec_array(Y), dp(d)) == 0
    assert list(d) == [n, -n, 2 * n]
    for v in [x, w, x2] + Y:
        core.N_VDestroy(v)


@pytest.mark.parametrize("n", [1, 3, 1001, 65537])
@pytest.mark.parametrize("nvec", [1, 2, 3, 6, 8, 9, 17])
def test_fused_ops_bitwise(libs, ctx, n, nvec):
    core, hip, orc = libs
    rng = np.random.default_rng(n * 31 + nvec)
    Xs = [rng.uniform(-1, 1, n) for _ in range(nvec)]
    coef = rng.uniform(-2, 2, nvec)
    coef[0] = 1.0 if nvec % 2 else coef[0]
    gX = [gpu_vec(ctx, a) for a in Xs]
    cX = [cpu_vec(orc, ctx, a) for a in Xs]
    # LinearCombination into a separate vector and in place on X[0]
    gz, cz = gpu_vec(ctx, np.zeros(n)), cpu_vec(orc, ctx, np.zeros(n))
    cd = np.ascontiguousarray(coef)
    assert core.N_VLinearCombination(nvec, dp(cd), vec_array(gX), gz) == 0
    assert core.N_VLinearCombination(nvec, dp(cd), vec_array(cX), cz) == 0
    assert np.array_equal(gpu_get(gz), cpu_get(cz))
    assert core.N_VLinearCombination(nvec, dp(cd), vec_array(gX), gX[0]) == 0
    assert core.N_VLinearCombination(nvec, dp(cd), vec_array(cX), cX[0]) == 0
    assert np.array_equal(gpu_get(gX[0]), cpu_get(cX[0]))
    # ScaleAddMulti: Z != Y and Z == Y
    xs = rng.uniform(-1, 1, n)
    gx, cx = gpu_vec(ctx, xs), cpu_vec(orc, ctx, xs)
    gZ = [gpu_vec(ctx, np.zeros(n)) for _ in range(nvec)]
    cZ = [cpu_vec(orc, ctx, np.zeros(n)) for _ in range(nvec)]
    assert core.N_VScaleAddMulti(nvec, dp(cd), gx, vec_array(gX), vec_array(gZ)) == 0
    assert core.N_VScaleAddMulti(nvec, dp(cd), cx, vec_array(cX), vec_array(cZ)) == 0
    for a, b in zip(gZ, cZ):
        assert np.array_equal(gpu_get(a), cpu_get(b))
    gYa, cYa = vec_array(gX), vec_array(cX)
    assert core.N_VScaleAddMulti(nvec, dp(cd), gx, gYa, gYa) == 0
    assert core.N_VScaleAddMulti(nvec, dp(cd), cx, cYa, cYa) == 0
    for a, b in zip(gX, cX):
        assert np.array_equal(gpu_get(a), cpu_get(b))
    # ScaleVectorArray in place
    assert core.N_VScaleVectorArray(nvec, dp(cd), gYa, gYa) == 0
    assert core.N_VScaleVectorArray(nvec, dp(cd), cYa, cYa) == 0
    for a, b in zip(gX, cX):
        assert np.array_equal(gpu_get(a), cpu_get(b))
    for v in gX + cX + gZ + cZ + [gz, cz, gx, cx]:
        core.N_VDestroy(v)


@pytest.mark.parametrize("nvec", [1, 5, 8, 9, 32, 33, 40])
def test_dotprodmulti(libs, ctx, nvec):
    core, hip, orc = libs
    n = 50001
    rng = np.random.default_rng(nvec)
    xs = rng.uniform(-1, 1, n)
    Ys = [rng.uniform(-1, 1, n) for _ in range(nvec)]
    gx, cx = gpu_vec(ctx, xs), cpu_vec(orc, ctx, xs)
    gY = [gpu_vec(ctx, a) for a in Ys]
    cY = [cpu_vec(orc, ctx, a) for a in Ys]
    dg, dc = np.zeros(nvec), np.zeros(nvec)
    assert core.N_VDotProdMulti(nvec, gx, vec_array(gY), dp(dg)) == 0
    assert core.N_VDotProdMulti(nvec, cx, vec_array(cY), dp(dc)) == 0
    np.testing.assert_allclose(dg, dc, rtol=1e-12, atol=1e-12)
    for v in gY + cY + [gx, cx]:
        core.N_VDestroy(v)


def test_clone_ops_and_empty(libs, ctx):
    core, hip, orc = libs
    v = gpu_vec(ctx, np.arange(10.0))
    w = core.N_VClone(v)
    core.N_VScale(2.0, v, w)
    assert np.array_equal(gpu_get(w), 2 * np.arange(10.0))
    assert core.N_VGetLength(w) == 10
    e = hip.N_VNew_Mi355(0, ctx)
    assert e and core.N_VGetLength(e) == 0
    core.N_VConst(1.0, e)
    for x in (v, w, e):
        core.N_VDestroy(x)
