"""ARKStep IMEX + SPGMR + block-Jacobi on the MI355X (advection-diffusion-
reaction problem, include/sunamd/imex_adr.h) against the CPU oracle (the
same integrator over the serial vector with the problem's CPU twin).

- reductions in serial order: every counter (steps, attempts, error-test and
  convergence failures, fe / fi evaluations, Newton and Krylov iterations,
  setups, preconditioner evaluations and solves, Jacobian-vector products)
  equals the oracle's and the final state is identical bit for bit, with
  modified Gram-Schmidt (DotProd + LinearSum per basis vector) and with the
  fused device classical Gram-Schmidt step;
- device-wide tree reductions (production mode): the step sequence may
  differ by rounding, the final state agrees to the integration tolerance;
- FP32 preconditioner factors: the Krylov solver absorbs them, the state
  agrees to the integration tolerance.
"""
import numpy as np
import pytest

from arkode_util import CLASSICAL_GS, MODIFIED_GS, make_problem, options, run_device, run_oracle
from sundials_amd import _lib

pytestmark = pytest.mark.gpu

COUNTERS = ("nst", "attempts", "netf", "ncfn", "nfe", "nfi", "nni", "nnf", "nsetups", "nli",
            "npe", "nps", "njtimes", "ncfl")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch

    assert torch.cuda.is_available()
    _lib.hip()


@pytest.mark.parametrize("n,gstype,tf", [(8, MODIFIED_GS, 0.1), (12, CLASSICAL_GS, 0.1),
                                         (6, MODIFIED_GS, 1.0)])
def test_imex_sequential_bitwise_vs_oracle(n, gstype, tf):
    p = make_problem(n)
    try:
        o = options(rtol=1e-6, atol=1e-10, tf=tf, gstype=gstype, sequential=1)
        rc_o, st_o, y_o = run_oracle(p, o)
        rc_d, st_d, y_d = run_device(p, o)
        assert rc_o == 0 and rc_d == 0
        bad = [(k, st_d[k], st_o[k]) for k in COUNTERS if st_d[k] != st_o[k]]
        assert not bad, bad
        assert np.array_equal(y_d, y_o)
        assert st_d["hlast"] == st_o["hlast"]
    finally:
        _lib.core().SUNAMD_AdrFree(p)


def _werr(y, yref, rtol=1e-6, atol=1e-10):
    w = 1.0 / (rtol * np.abs(yref) + atol)
    return np.sqrt(np.mean(((y - yref) * w) ** 2))


@pytest.mark.parametrize("gstype", [MODIFIED_GS, CLASSICAL_GS])
def test_imex_tree_reductions_accuracy(gstype):
    p = make_problem(24)
    try:
        o = options(rtol=1e-6, atol=1e-10, tf=0.05, gstype=gstype, sequential=0)
        rc_d, st_d, y_d = run_device(p, o)
        rc_o, st_o, y_o = run_oracle(p, options(rtol=1e-6, atol=1e-10, tf=0.05,
                                                gstype=gstype))
        assert rc_d == 0 and rc_o == 0
        # a different summation order changes the step sequence at most slightly
        assert abs(st_d["nst"] - st_o["nst"]) <= max(2, 0.1 * st_o["nst"])
        assert _werr(y_d, y_o) < 1.0
    finally:
        _lib.core().SUNAMD_AdrFree(p)


def test_imex_fp32_preconditioner():
    p = make_problem(16)
    try:
        o64 = options(rtol=1e-6, atol=1e-10, tf=0.05, sequential=0)
        o32 = options(rtol=1e-6, atol=1e-10, tf=0.05, sequential=0, prec_fp32=1)
        rc64, st64, y64 = run_device(p, o64)
        rc32, st32, y32 = run_device(p, o32)
        assert rc64 == 0 and rc32 == 0
        assert _werr(y32, y64) < 1.0
        # the Krylov iterations per Newton iteration stay close
        assert st32["nli"] / st32["nni"] < 1.5 * st64["nli"] / st64["nni"] + 0.5
        print("fp64 prec: nst %d nli %d; fp32 prec: nst %d nli %d" %
              (st64["nst"], st64["nli"], st32["nst"], st32["nli"]))
    finally:
        _lib.core().SUNAMD_AdrFree(p)
