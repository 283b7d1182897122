"""The learned Newton initial iterate on the MI355X.

- MLP forward and gradient from the device kernels (MFMA bf16 products, FP32
  accumulation) against the float64 CPU reference that rounds to bf16 at the
  same points (tests/mlp_ref.py, itself pinned by finite differences).
- Adam training on device-resident data reduces the loss of a teacher
  regression.
- End to end: training pairs logged by plain FP64 runs of one batch train the
  network; on a held-out batch the integrator with the learned start (and
  the zero start as fallback) completes every system with the final state
  within the integration tolerance of the plain run, and the effect on
  Newton iterations per step is measured and printed.
"""
import ctypes as C

import numpy as np
import pytest

from sundials_amd import _lib
from mlp_ref import NPARAM, IN, IN_PAD, OUT, forward, loss_grad, unpack
from batch_util import RTOL, ATOL, make_batch, free_batch, run_device, run_oracle, \
    system_slices, MXSTEP

pytestmark = pytest.mark.gpu
PF = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def hip():
    import torch

    assert torch.cuda.is_available()
    return _lib.hip()


def _params(M):
    p = np.zeros(NPARAM, dtype=np.float32)
    assert _lib.hip().SUNAMD_MLPGetParams(M, p.ctypes.data_as(PF)) == 0
    return p


def _inputs(rng, B):
    x = np.zeros((B, IN_PAD), dtype=np.float32)
    x[:, :IN] = rng.uniform(-1.5, 0.5, (B, IN))
    return x


def test_mlp_forward_matches_reference(hip):
    M = hip.SUNAMD_MLPCreate(7, 256)
    assert M
    try:
        rng = np.random.default_rng(1)
        x = _inputs(rng, 64)
        y = np.zeros((64, OUT), dtype=np.float32)
        assert hip.SUNAMD_MLPForward(M, x.ctypes.data_as(PF), 64, y.ctypes.data_as(PF)) == 0
        y_ref, _ = forward(_params(M).astype(np.float64), x)
        scale = np.abs(y_ref).max()
        assert np.abs(y - y_ref).max() <= 2e-3 * scale
    finally:
        hip.SUNAMD_MLPFree(M)


def test_integrator_network_matches_reference(hip):
    """The network as the batched integrator evaluates it inside its kernels
    (one wavefront per 16 rows, bf16 weights as 16 x 16 x 32 MFMA operands)
    against the float64 reference that rounds at the same points, and
    against the training-side forward pass; 40 rows, so the last wavefront's
    block is partly empty."""
    M = hip.SUNAMD_MLPCreate(11, 256)
    assert M
    try:
        rng = np.random.default_rng(4)
        x = _inputs(rng, 48)
        yw = np.zeros((48, OUT), dtype=np.float32)
        assert hip.SUNAMD_MLPForwardWave(hip.SUNAMD_MLPDeviceWeights(M), x[:40].ctypes.data_as(PF),
                                         40, yw.ctypes.data_as(PF)) == 0
        y_ref, _ = forward(_params(M).astype(np.float64), x)
        scale = np.abs(y_ref).max()
        err = np.abs(yw[:40] - y_ref[:40]).max()
        print("in-kernel network vs reference: max |dy| %.3g (scale %.3g)" % (err, scale))
        assert err <= 2e-3 * scale
        yf = np.zeros((48, OUT), dtype=np.float32)
        assert hip.SUNAMD_MLPForward(M, x.ctypes.data_as(PF), 48, yf.ctypes.data_as(PF)) == 0
        assert np.abs(yw[:40] - yf[:40]).max() <= 2e-3 * scale
        assert np.all(yw[40:] == 0.0)  # rows beyond the input are not written
    finally:
        hip.SUNAMD_MLPFree(M)


def test_mlp_gradient_matches_reference(hip):
    M = hip.SUNAMD_MLPCreate(11, 256)
    assert M
    try:
        rng = np.random.default_rng(2)
        B = 128
        x = _inputs(rng, B)
        t = rng.standard_normal((B, OUT)).astype(np.float32)
        mask = (rng.random((B, OUT)) < 0.6).astype(np.float32)
        loss = C.c_float()
        g = np.zeros(NPARAM, dtype=np.float32)
        assert hip.SUNAMD_MLPLossGrad(M, x.ctypes.data_as(PF), t.ctypes.data_as(PF),
                                      mask.ctypes.data_as(PF), B, C.byref(loss),
                                      g.ctypes.data_as(PF)) == 0
        l_ref, g_ref = loss_grad(_params(M).astype(np.float64), x, t, mask)
        assert abs(loss.value - l_ref) <= 1e-4 * abs(l_ref)
        gd, gr = unpack(g), unpack(g_ref)
        for name in gr:
            err = np.abs(gd[name] - gr[name]).max()
            assert err <= 1e-2 * np.abs(gr[name]).max() + 1e-7, (name, err)
        # W1 columns of the padding inputs get no gradient
        assert np.all(gd["W1"][:, IN:] == 0.0)
    finally:
        hip.SUNAMD_MLPFree(M)


def test_mlp_training_reduces_loss(hip):
    import torch

    M = hip.SUNAMD_MLPCreate(3, 512)
    T = hip.SUNAMD_MLPCreate(99, 512)  # teacher network
    assert M and T
    try:
        rng = np.random.default_rng(5)
        n = 8192
        x = _inputs(rng, n)
        t = np.zeros((n, OUT), dtype=np.float32)
        for s in range(0, n, 512):
            assert hip.SUNAMD_MLPForward(T, x[s:s + 512].ctypes.data_as(PF), 512,
                                         t[s:s + 512].ctypes.data_as(PF)) == 0
        t *= 10.0
        mask = np.ones((n, OUT), dtype=np.float32)
        xd, td, md = (torch.from_numpy(a).cuda() for a in (x, t, mask))
        torch.cuda.synchronize()
        l0, l1 = C.c_float(), C.c_float()
        assert hip.SUNAMD_MLPTrain(M, C.c_void_p(xd.data_ptr()), C.c_void_p(td.data_ptr()),
                                   C.c_void_p(md.data_ptr()), n, 512, 10, 1e-3, C.byref(l0)) == 0
        assert hip.SUNAMD_MLPTrain(M, C.c_void_p(xd.data_ptr()), C.c_void_p(td.data_ptr()),
                                   C.c_void_p(md.data_ptr()), n, 512, 1500, 1e-3, C.byref(l1)) == 0
        print("teacher regression loss: %.4g -> %.4g" % (l0.value, l1.value))
        assert np.isfinite(l1.value) and l1.value < 0.2 * l0.value
    finally:
        hip.SUNAMD_MLPFree(M)
        hip.SUNAMD_MLPFree(T)


def _log_pairs(hip, b, cap, stride):
    import torch

    x = torch.zeros((cap, IN_PAD), dtype=torch.float32, device="cuda")
    t = torch.zeros((cap, OUT), dtype=torch.float32, device="cuda")
    m = torch.zeros((cap, OUT), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    B = hip.SUNAMD_BatchedBDFCreate(b, RTOL, ATOL, 0)
    assert B
    try:
        hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, MXSTEP)
        assert hip.SUNAMD_BatchedBDFSetTrainingLog(B, C.c_void_p(x.data_ptr()),
                                                   C.c_void_p(t.data_ptr()),
                                                   C.c_void_p(m.data_ptr()), cap, stride) == 0
        wall = C.c_double()
        assert hip.SUNAMD_BatchedBDFRun(B, 1.0, C.byref(wall)) == 0
        n = hip.SUNAMD_BatchedBDFGetTrainingLogCount(B)
    finally:
        hip.SUNAMD_BatchedBDFFree(B)
    return x[:n], t[:n], m[:n]


def test_learned_predictor_end_to_end(hip):
    import torch

    mv, wv = (8, 16, 32), (1.0, 1.0, 1.0)
    train = make_batch(1536, 5005, m_values=mv, weights=wv)
    test = make_batch(512, 6006, m_values=mv, weights=wv)
    M = hip.SUNAMD_MLPCreate(17, 1024)
    assert M
    try:
        x, t, m = _log_pairs(hip, train, 400000, 1)
        n = x.shape[0]
        assert n > 10000
        # logged rows are finite, masks match the species counts
        assert torch.isfinite(x).all() and torch.isfinite(t).all()
        perm = torch.randperm(n, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
        x, t, m = x[perm].contiguous(), t[perm].contiguous(), m[perm].contiguous()
        torch.cuda.synchronize()
        loss = C.c_float()
        assert hip.SUNAMD_MLPTrain(M, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                   C.c_void_p(m.data_ptr()), n, 1024, 3000, 1e-3,
                                   C.byref(loss)) == 0
        print("training pairs %d, final loss %.4g" % (n, loss.value))

        y64, c64, rc, w64 = run_device(test, 0)
        assert rc == 0
        B = hip.SUNAMD_BatchedBDFCreate(test, RTOL, ATOL, 0)
        try:
            hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, MXSTEP)
            assert hip.SUNAMD_BatchedBDFSetLearnedPredictor(B, hip.SUNAMD_MLPDeviceWeights(M),
                                                            0xff) == 0
            wall = C.c_double()
            assert hip.SUNAMD_BatchedBDFRun(B, 1.0, C.byref(wall)) == 0
            bc = test.contents
            yl = np.zeros(bc.ytotal)
            cl = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
            assert hip.SUNAMD_BatchedBDFGetResults(B, yl.ctypes.data_as(C.POINTER(C.c_double)),
                                                   cl.ctypes.data_as(C.POINTER(C.c_long))) == 0
            tried, kept = C.c_long(), C.c_long()
            assert hip.SUNAMD_BatchedBDFGetLearnedStats(B, C.byref(tried), C.byref(kept)) == 0
        finally:
            hip.SUNAMD_BatchedBDFFree(B)
        assert np.all(cl[:, 9] == 0)
        # accuracy: against a reference at 10^4x tighter tolerances the learned
        # run's errors are distributed like the plain run's
        y_tight, _, rc, _ = run_oracle(test, nthreads=16, rtol=1e-10, atol=1e-14)
        assert rc == 0

        def errs(y):
            out = []
            for (k, c, mm, off) in system_slices(test):
                d = y[off:off + mm] - y_tight[off:off + mm]
                w = 1.0 / (RTOL * np.abs(y_tight[off:off + mm]) + ATOL)
                out.append(np.sqrt(np.mean((d * w) ** 2)))
            return np.array(out)

        el, ep = errs(yl), errs(y64)
        print("error vs tight reference: plain median %.3g p95 %.3g, learned median %.3g p95 %.3g"
              % (np.median(ep), np.percentile(ep, 95), np.median(el), np.percentile(el, 95)))
        assert np.median(el) <= 1.5 * np.median(ep) + 0.05
        assert np.percentile(el, 95) <= 1.5 * np.percentile(ep, 95) + 0.1
        nni_l = cl[:, 4].sum() / cl[:, 0].sum()
        nni_p = c64[:, 4].sum() / c64[:, 0].sum()
        print("held-out: Newton iters/step plain %.4f learned %.4f; guess kept %d of %d steps; "
              "steps plain %d learned %d" % (nni_p, nni_l, kept.value, tried.value,
                                             c64[:, 0].sum(), cl[:, 0].sum()))
        assert tried.value > 0
    finally:
        hip.SUNAMD_MLPFree(M)
        free_batch(train)
        free_batch(test)


def test_learned_start_seven_size_mix(hip):
    """Config 5 at test scale over the whole size mix (m in {4, 8, 16, 24, 32,
    48, 64}, the benchmark's weights): the network is trained on pairs logged
    by plain FP64 runs of one batch and evaluated on a held-out batch in the
    product form (FP32 Newton-matrix LU) with the learned start on in every
    bucket, beside the same form without it.  Printed (the measured gain):
    Newton iterations per step, steps, wall time and guesses kept per bucket.
    Asserted: every system succeeds; the guess is tried in every bucket and
    kept in some steps; Newton iterations per step stay within 5 % of the run
    without it (the zero start is the fallback whenever the guess's residual
    is larger, so a poor network cannot make the iteration diverge); accuracy
    against a 10^4x tighter reference is distributed like the run without
    it."""
    import torch

    train = make_batch(2048, 7007)
    test = make_batch(1400, 8008)
    M = hip.SUNAMD_MLPCreate(19, 1024)
    assert M
    try:
        x, t, m = _log_pairs(hip, train, 1200000, 1)
        n = x.shape[0]
        perm = torch.randperm(n, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
        x, t, m = x[perm].contiguous(), t[perm].contiguous(), m[perm].contiguous()
        torch.cuda.synchronize()
        loss = C.c_float()
        assert hip.SUNAMD_MLPTrain(M, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                   C.c_void_p(m.data_ptr()), n, 1024, 4000, 1e-3,
                                   C.byref(loss)) == 0
        del x, t, m
        print("training pairs %d, final loss %.4g" % (n, loss.value))
        bc = test.contents
        nb = bc.nbuckets
        ms = [bc.bucket[k].contents.m for k in range(nb)]
        sb = np.array([bc.sys_bucket[s_] for s_ in range(bc.nsys)])

        def run(mask):
            B = hip.SUNAMD_BatchedBDFCreate(test, RTOL, ATOL, 1)
            assert B
            try:
                hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, MXSTEP)
                hip.SUNAMD_BatchedBDFSetOccupancy(B, (C.c_int * nb)(*([2] * nb)))
                if mask:
                    assert hip.SUNAMD_BatchedBDFSetLearnedPredictor(
                        B, hip.SUNAMD_MLPDeviceWeights(M), mask) == 0
                wall = C.c_double()
                assert hip.SUNAMD_BatchedBDFRun(B, 1.0, C.byref(wall)) == 0  # warm-up
                assert hip.SUNAMD_BatchedBDFRun(B, 1.0, C.byref(wall)) == 0
                y = np.zeros(bc.ytotal)
                c = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
                assert hip.SUNAMD_BatchedBDFGetResults(B, y.ctypes.data_as(C.POINTER(C.c_double)),
                                                       c.ctypes.data_as(C.POINTER(C.c_long))) == 0
                tried = np.zeros(nb, dtype=np.int64)
                kept = np.zeros(nb, dtype=np.int64)
                assert hip.SUNAMD_BatchedBDFGetLearnedStatsByBucket(
                    B, tried.ctypes.data_as(C.POINTER(C.c_long)),
                    kept.ctypes.data_as(C.POINTER(C.c_long))) == nb
                return y, c, wall.value, tried, kept
            finally:
                hip.SUNAMD_BatchedBDFFree(B)

        yp, cp, wp, _, _ = run(0)
        yl, cl, wl, tried, kept = run((1 << nb) - 1)
        assert np.all(cp[:, 9] == 0) and np.all(cl[:, 9] == 0)
        for k in range(nb):
            sel = sb == k
            print("m=%2d: Newton iters/step plain %.4f learned %.4f, steps %d / %d, guess kept "
                  "%d of %d" % (ms[k], cp[sel, 4].sum() / cp[sel, 0].sum(),
                                cl[sel, 4].sum() / cl[sel, 0].sum(), cp[sel, 0].sum(),
                                cl[sel, 0].sum(), kept[k], tried[k]))
        nni_p = cp[:, 4].sum() / cp[:, 0].sum()
        nni_l = cl[:, 4].sum() / cl[:, 0].sum()
        print("all sizes: Newton iters/step plain %.4f learned %.4f; wall plain %.2f ms learned "
              "%.2f ms; guess kept %d of %d" % (nni_p, nni_l, 1e3 * wp, 1e3 * wl, kept.sum(),
                                               tried.sum()))
        assert np.all(tried > 0) and kept.sum() > 0
        assert nni_l <= nni_p * 1.05
        y_tight, _, rc, _ = run_oracle(test, nthreads=16, rtol=1e-10, atol=1e-14)
        assert rc == 0

        def errs(y):
            out = []
            for (k, c_, mm, off) in system_slices(test):
                d = y[off:off + mm] - y_tight[off:off + mm]
                w = 1.0 / (RTOL * np.abs(y_tight[off:off + mm]) + ATOL)
                out.append(np.sqrt(np.mean((d * w) ** 2)))
            return np.array(out)

        el, ep = errs(yl), errs(yp)
        print("error vs tight reference: plain median %.3g p95 %.3g, learned median %.3g p95 %.3g"
              % (np.median(ep), np.percentile(ep, 95), np.median(el), np.percentile(el, 95)))
        assert np.median(el) <= 1.5 * np.median(ep) + 0.05
        assert np.percentile(el, 95) <= 1.5 * np.percentile(ep, 95) + 0.1
    finally:
        hip.SUNAMD_MLPFree(M)
        free_batch(train)
        free_batch(test)
