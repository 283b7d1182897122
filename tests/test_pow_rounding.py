"""The x^y of the step-size controller (shared by the host and device
integrators) is correctly rounded: checked against 200-bit mpmath on the
arguments the controller produces (x = 6 dsm or 10 dup over many decades,
y = 1/L for L = 2..6) and on a general sample."""
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
