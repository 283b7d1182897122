"""KINSOL Newton iteration (KIN_NONE / KIN_LINESEARCH) on the MI355X vector
with the block-diagonal matrix and batched LU (include/sunamd/kin_newton.h:
cell functions and 6 x 6 Jacobian blocks from device kernels) against the CPU
oracle (the same restatement over the serial vector, per-block LU).

- one cell, the reference example's two initial guesses: the device retraces
  the published kinFerTron_dns counts (nni / nfe / nje) and printed solution;
- 1, 257 and 4096 perturbed cells, both guesses, exact and modified Newton,
  with and without the line search, with and without the fused scaled norms
  (KINSetUseFusedKernels), reductions in serial order: flag, every counter
  and every final value identical to the oracle bit for bit;
- device-wide tree reductions on 2^18 cells: converged, counts within one of
  the oracle's, every cell at its root, the constraints satisfied; the fused
  norms give the identical run.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from sundials_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch

    assert torch.cuda.is_available()
    _lib.hip()


def _opts(**kw):
    o = _lib.KinNewtonOptions()
    o.ncells, o.variant, o.guess, o.perturb, o.strategy, o.msbset = 1, 1, 2, 0, 0, 1
    o.fnormtol, o.scsteptol, o.maxiter, o.jac, o.sequential = 1e-5, 1e-5, 0, 1, 1
    o.fused = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _solve(fn, o):
    st = np.zeros(_lib.KINN_STATS)
    u = np.zeros(6 * o.ncells)
    print("kinsol newton:", {k: getattr(o, k) for k, _ in o._fields_}, flush=True)
    rc = fn(C.byref(o), dp(st), dp(u))
    return rc, st, u


def device(o):
    return _solve(_lib.hip().SUNAMD_KinNewtonSolve_Mi355, o)


def oracle(o):
    return _solve(_lib.load_oracle().oracle_kin_newton_solve, o)


def _golden():
    return json.load(open(os.path.join(HERE, "golden", "kin_fertron_dns.json")))["runs"]


@pytest.mark.parametrize("ref_run", [0, 1, 2, 3, 4, 5])
def test_fertron_on_device_retraces_published_run(ref_run):
    run = _golden()[ref_run]
    o = _opts(guess=run["guess"], strategy=int(run["linesearch"]),
              msbset=1 if run["exact"] else 0)
    rc, st, u = device(o)
    assert rc == 0
    assert (int(st[1]), int(st[2]), int(st[3])) == (run["nni"], run["nfe"], run["nje"])
    assert ("%10.6g" % u[0]).strip() == run["x1"]
    assert ("%10.6g" % u[1]).strip() == run["x2"]


@pytest.mark.parametrize("ncells", [1, 257, 4096])
@pytest.mark.parametrize("guess", [1, 2])
@pytest.mark.parametrize("strategy,msbset", [(0, 1), (1, 1), (0, 0), (1, 0)])
@pytest.mark.parametrize("fused", [0, 1])
def test_device_equals_oracle_bitwise(ncells, guess, strategy, msbset, fused):
    o = _opts(ncells=ncells, perturb=1, guess=guess, strategy=strategy, msbset=msbset,
              fnormtol=1e-10, scsteptol=1e-12, fused=fused)
    rc_d, st_d, u_d = device(o)
    rc_o, st_o, u_o = oracle(_opts(ncells=ncells, perturb=1, guess=guess, strategy=strategy,
                                   msbset=msbset, fnormtol=1e-10, scsteptol=1e-12))
    assert rc_d == rc_o == 0
    assert np.array_equal(st_d[:7], st_o[:7])
    assert np.array_equal(u_d, u_o)


@pytest.mark.parametrize("guess,root", [(1, (0.29945, 2.83693)), (2, (0.5, 3.1415926))])
def test_tree_reductions_large(guess, root):
    o = _opts(ncells=1 << 18, perturb=1, guess=guess, strategy=1, msbset=0,
              fnormtol=1e-10, scsteptol=1e-12, sequential=0)
    rc_d, st_d, u_d = device(o)
    # the fused scaled norms change nothing, also with tree reductions
    o.fused = 1
    rc_f, st_f, u_f = device(o)
    assert rc_f == rc_d and np.array_equal(st_f[:7], st_d[:7]) and np.array_equal(u_f, u_d)
    rc_o, st_o, _ = oracle(o)
    assert rc_d == rc_o == 0
    assert abs(int(st_d[1]) - int(st_o[1])) <= 1
    x = u_d.reshape(-1, 6)
    assert np.allclose(x[:, 0], root[0], atol=2e-5)
    assert np.allclose(x[:, 1], root[1], atol=2e-5)
    assert (x[:, 2] >= 0).all() and (x[:, 4] >= 0).all()
    assert (x[:, 3] <= 0).all() and (x[:, 5] <= 0).all()


def test_bad_options_rejected_before_any_launch():
    for bad in (dict(variant=0), dict(jac=0), dict(guess=3), dict(strategy=3), dict(ncells=0)):
        rc, _, _ = device(_opts(**bad))
        assert rc == -100
