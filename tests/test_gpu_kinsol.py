"""KINSOL fixed-point iteration with Anderson acceleration on the MI355X
vector (include/sunamd/kin_fp.h, the map evaluated by a device kernel)
against the CPU oracle (the same restatement over the serial vector).

- the reference example's 11 published runs (one cell, the example's map
  with the device's cos / sin / exp): the same iteration and function-
  evaluation counts as published, the solution within 1e-9 of the oracle's;
- the IEEE-only map (variant 1) on 4096 perturbed cells with reductions in
  serial order: counts and every final value identical to the oracle bit for
  bit, for the plain, damped, accelerated (each QR update), delayed and
  depth-function configurations, with and without the fused scaled norm
  (KINSetUseFusedKernels);
- device-wide tree reductions on 2^18 cells: the iteration count within one
  of the oracle's, the solution to the tolerance.
"""
import json
import os

import numpy as np
import pytest

from kinsol_util import options, run_device, run_oracle
from sundials_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch

    assert torch.cuda.is_available()
    _lib.hip()


def _runs():
    return json.load(open(os.path.join(HERE, "golden", "kin_analytic_fp.json")))["runs"]


@pytest.mark.parametrize("run", _runs(), ids=lambda r: r["file"])
def test_kin_analytic_on_device_matches_published_counts(run):
    rc_d, st_d, u_d = run_device(options(run))
    rc_o, st_o, u_o = run_oracle(options(run))
    assert rc_d == 0 and rc_o == 0
    assert (st_d["nni"], st_d["nfe"]) == (run["nni"], run["nfe"])
    assert np.abs(u_d - u_o).max() < 1e-9


CONFIGS = [dict(), dict(damping_fp=0.5), dict(m_aa=3), dict(m_aa=3, orth_aa=1),
           dict(m_aa=3, orth_aa=2, damping_fn=1), dict(m_aa=4, orth_aa=3, damping_aa=0.7),
           dict(m_aa=3, delay_aa=2, depth_fn=1)]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: ",".join("%s=%s" % kv for kv in c.items())
                         or "plain")
@pytest.mark.parametrize("fused", [0, 1])
def test_kin_cells_sequential_bitwise_vs_oracle(cfg, fused):
    o = options(ncells=4096, perturb=1, variant=1, tol=1e-11, maxiter=60, **cfg)
    rc_o, st_o, u_o = run_oracle(o)
    o.sequential, o.fused = 1, fused
    rc_d, st_d, u_d = run_device(o)
    assert rc_o == rc_d
    assert (st_d["nni"], st_d["nfe"]) == (st_o["nni"], st_o["nfe"])
    assert st_d["fnorm"] == st_o["fnorm"]
    assert np.array_equal(u_d, u_o)


@pytest.mark.parametrize("orth", [0, 2])
def test_kin_cells_tree_reductions(orth):
    o = options(ncells=1 << 18, perturb=1, variant=1, tol=1e-11, maxiter=60, m_aa=5,
                orth_aa=orth)
    rc_o, st_o, u_o = run_oracle(o)
    rc_d, st_d, u_d = run_device(o)
    assert rc_o == 0 and rc_d == 0
    assert abs(st_d["nni"] - st_o["nni"]) <= 1
    assert np.abs(u_d - u_o).max() < 1e-10
