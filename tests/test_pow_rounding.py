"""The x^y of the step-size controller (shared by the host and device
integrators) is correctly rounded: checked against 200-bit mpmath on the
arguments the controller produces (x = 6 dsm or 10 dup over many decades,
y = 1/L for L = 2..6) and on a general sample.

Its fast path returns only when its error bound decides the rounding: the
bound is checked against the 200-bit error of the fast approximation on
samples from every regime (controller arguments, x over the whole exponent
range, x near 1 with large |y|, results near the overflow and underflow
limits), and the two paths agree bit for bit wherever the fast one returns."""
import ctypes as C

import numpy as np
import mpmath
import pytest

from sundials_amd import _lib

mpmath.mp.prec = 200


def _cr(x, y):
    return float(mpmath.power(mpmath.mpf(x), mpmath.mpf(y)))


@pytest.mark.parametrize("L", [2, 3, 4, 5, 6])
def test_pow_correctly_rounded_controller_arguments(L):
    core = _lib.core()
    rng = np.random.default_rng(L)
    xs = 10.0 ** rng.uniform(-14, 8, 4000)
    y = 1.0 / L
    bad = [(x, core.SUNAMD_Pow(x, y), _cr(x, y)) for x in xs if core.SUNAMD_Pow(x, y) != _cr(x, y)]
    assert not bad, bad[:3]


def test_pow_general_arguments_and_special_cases():
    core = _lib.core()
    rng = np.random.default_rng(99)
    xs = 10.0 ** rng.uniform(-30, 30, 3000)
    ys = rng.uniform(-8, 8, 3000)
    ok = [core.SUNAMD_Pow(x, y) == _cr(x, y) for x, y in zip(xs, ys)]
    assert all(ok), sum(not o for o in ok)
    assert core.SUNAMD_Pow(2.0, 10.0) == 1024.0
    assert core.SUNAMD_Pow(1.0, 0.37) == 1.0
    assert core.SUNAMD_Pow(7.0, 0.0) == 1.0
    assert core.SUNAMD_Pow(0.25, 0.5) == 0.5


def _fast_parts(core, x, y):
    r, hi, lo, bd = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    sc = C.c_int()
    ok = core.SUNAMD_PowFastParts(x, y, C.byref(r), C.byref(hi), C.byref(lo), C.byref(bd),
                                  C.byref(sc))
    return ok, r.value, hi.value, lo.value, bd.value, sc.value


def _samples(rng, n):
    yield 10.0 ** rng.uniform(-14, 8, n), 1.0 / rng.integers(2, 7, n)
    yield 10.0 ** rng.uniform(-300, 300, n), rng.uniform(-2, 2, n)
    yield 1.0 + rng.uniform(-1e-3, 1e-3, n), rng.uniform(-500, 500, n)
    yield 1.0 + rng.uniform(-1e-12, 1e-12, n), rng.uniform(-8, 8, n)
    yield rng.uniform(0.5, 2.0, n), 2.0 ** rng.uniform(-60, -2, n)
    yield np.exp(rng.uniform(-700, 700, n) / 1.3), np.full(n, 1.3)


def test_pow_fast_path_error_within_bound():
    core = _lib.core()
    rng = np.random.default_rng(2024)
    worst, decided = 0.0, 0
    for xs, ys in _samples(rng, 1500):
        for x, y in zip(xs, ys):
            x, y = float(x), float(y)
            ok, r, hi, lo, bd, sc = _fast_parts(core, x, y)
            if bd == 0.0:
                assert not ok  # result out of the fast path's range
                continue
            v = mpmath.power(mpmath.mpf(x), mpmath.mpf(y)) * mpmath.mpf(2) ** (-sc)
            err = abs(mpmath.mpf(hi) + mpmath.mpf(lo) - v)
            worst = max(worst, float(err / mpmath.mpf(bd)))
            if ok:
                decided += 1
                assert r == _cr(x, y), (x, y)
    # the bound is 8x the analysed error; observed errors stay far below it
    assert worst < 0.125, worst
    assert decided > 7000


def test_pow_paths_agree():
    core = _lib.core()
    rng = np.random.default_rng(7)
    n, fast = 0, 0
    for xs, ys in _samples(rng, 20000):
        for x, y in zip(xs, ys):
            x, y = float(x), float(y)
            ok, r, _, _, _, _ = _fast_parts(core, x, y)
            if ok:
                fast += 1
                assert r == core.SUNAMD_PowExact(x, y), (x, y)
            n += 1
    assert fast > 0.8 * n


def test_pow_special_values_follow_c99():
    """Zero, negative, infinite and NaN arguments take the C99 special values
    (Annex F.9.4.4), as the C library's pow (numpy.power) does."""
    import math

    core = _lib.core()
    inf, nan = float("inf"), float("nan")
    xs = [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, -3.7, inf, -inf, nan, 1e-310, -1e-310]
    ys = [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 0.5, -0.5, 2.5, 1e300, -1e300, inf, -inf,
          nan, 2.0 ** 60, 2.0 ** 53 + 2]
    with np.errstate(all="ignore"):
        for x in xs:
            for y in ys:
                got = core.SUNAMD_Pow(x, y)
                want = float(np.power(np.float64(x), np.float64(y)))
                if math.isnan(want):
                    assert math.isnan(got), (x, y, got)
                else:
                    assert got == want and math.copysign(1, got) == math.copysign(1, want), \
                        (x, y, got, want)
