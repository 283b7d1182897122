"""Allocation-free bucket integration (SUNAMD_BatchedBDFSolveBucketNoAlloc)
against the allocating entry point (SUNAMD_BatchedBDFCreate / Run /
GetResults) on the same systems.

Both run the same persistent kernel at the same register budget, so the bar
is bit-for-bit equality of every final state and every counter, in FP64 and
FP32-LU form.  Two calls in a row, the second on a smaller bucket (fewer
systems, fewer species, fewer reactions), show that the shared device array
keeps nothing from the previous call that matters; an out-of-range reaction
index is caught on the device (flag -22, state returned unintegrated) and an
over-limit reaction count is refused on the host.
"""
import ctypes as C

import numpy as np
import pytest

from batch_util import RTOL, ATOL, TOUT, MXSTEP, make_batch, free_batch, system_slices
from sundials_amd import _lib

pytestmark = pytest.mark.gpu

LU_FP64, LU_FP32 = 0, 1
WAVES = 2  # the register budget the no-alloc entry point compiles for


def _run_allocating(b, precision):
    hip = _lib.hip()
    B = hip.SUNAMD_BatchedBDFCreate(b, RTOL, ATOL, precision)
    assert B
    try:
        nb = b.contents.nbuckets
        assert hip.SUNAMD_BatchedBDFSetOccupancy(B, (C.c_int * nb)(*([WAVES] * nb))) == 0
        assert hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, MXSTEP) == 0
        wall = C.c_double(0.0)
        assert hip.SUNAMD_BatchedBDFRun(B, TOUT, C.byref(wall)) == 0
        bc = b.contents
        y = np.zeros(bc.ytotal)
        cnt = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
        hip.SUNAMD_BatchedBDFGetResults(B, y.ctypes.data_as(C.POINTER(C.c_double)),
                                        cnt.ctypes.data_as(C.POINTER(C.c_long)))
        return y, cnt
    finally:
        hip.SUNAMD_BatchedBDFFree(B)


def _bucket_tensors(p, dev):
    import torch

    m, R, n = p.m, p.R, p.ncells
    idx = torch.from_numpy(np.ctypeslib.as_array(p.idx, shape=(4 * R,)).copy()).to(dev)
    kf = torch.from_numpy(np.ctypeslib.as_array(p.kf, shape=(n * R,)).copy()).to(dev)
    kb = torch.from_numpy(np.ctypeslib.as_array(p.kb, shape=(n * R,)).copy()).to(dev)
    y0 = torch.from_numpy(np.ctypeslib.as_array(p.y0, shape=(n * m,)).copy()).to(dev)
    return idx, kf, kb, y0


def _run_no_alloc(m, R, n, idx, kf, kb, y0, precision, stream):
    import torch

    hip = _lib.hip()
    yout = torch.full((n * m,), -1.0, dtype=torch.float64, device=y0.device)
    cnt = torch.full((n * _lib.BS_COUNT,), -7, dtype=torch.int64, device=y0.device)
    print("no-alloc call: m=%d R=%d ncells=%d precision=%d" % (m, R, n, precision), flush=True)
    rc = hip.SUNAMD_BatchedBDFSolveBucketNoAlloc(
        m, R, n, idx.data_ptr(), kf.data_ptr(), kb.data_ptr(), y0.data_ptr(), RTOL, ATOL,
        TOUT, MXSTEP, precision, yout.data_ptr(), cnt.data_ptr(), stream.cuda_stream)
    assert rc == 0
    stream.synchronize()
    return yout.cpu().numpy(), cnt.cpu().numpy().reshape(n, _lib.BS_COUNT)


def _expected(b, y, cnt, k):
    """bucket k's systems of the allocating run, in cell order"""
    p = b.contents.bucket[k].contents
    ye = np.zeros(p.ncells * p.m)
    ce = np.zeros((p.ncells, _lib.BS_COUNT), dtype=np.int64)
    for s, (kk, c, m, off) in enumerate(system_slices(b)):
        if kk == k:
            ye[c * m:(c + 1) * m] = y[off:off + m]
            ce[c] = cnt[s]
    return ye, ce


@pytest.mark.parametrize("precision", [LU_FP64, LU_FP32])
def test_no_alloc_equals_allocating_entry_point(precision):
    import torch

    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    # first call on the larger bucket, second on a smaller one
    cases = [(make_batch(300, 8101, m_values=(16,), weights=(1.0,)), "first"),
             (make_batch(90, 8102, m_values=(8,), weights=(1.0,)), "second, smaller")]
    try:
        for b, label in cases:
            y_ref, c_ref = _run_allocating(b, precision)
            assert np.all(c_ref[:, 9] == 0)
            p = b.contents.bucket[0].contents
            tens = _bucket_tensors(p, dev)
            torch.cuda.synchronize()
            y, cnt = _run_no_alloc(p.m, p.R, p.ncells, *tens, precision, stream)
            ye, ce = _expected(b, y_ref, c_ref, 0)
            assert np.array_equal(cnt, ce), "%s call: counters differ" % label
            assert np.array_equal(y, ye), "%s call: states differ" % label
    finally:
        for b, _ in cases:
            free_batch(b)


def test_no_alloc_rejects_bad_input():
    import torch

    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    hip = _lib.hip()
    b = make_batch(20, 8103, m_values=(4,), weights=(1.0,))
    try:
        p = b.contents.bucket[0].contents
        idx, kf, kb, y0 = _bucket_tensors(p, dev)
        n, m, R = p.ncells, p.m, p.R
        yout = torch.zeros(n * m, dtype=torch.float64, device=dev)
        cnt = torch.zeros(n * _lib.BS_COUNT, dtype=torch.int64, device=dev)
        # host checks: too many reactions, a size without a kernel, NULL output
        assert hip.SUNAMD_BatchedBDFSolveBucketNoAlloc(
            m, 1025, n, idx.data_ptr(), kf.data_ptr(), kb.data_ptr(), y0.data_ptr(), RTOL, ATOL,
            TOUT, MXSTEP, LU_FP64, yout.data_ptr(), cnt.data_ptr(), stream.cuda_stream) == -1
        assert hip.SUNAMD_BatchedBDFSolveBucketNoAlloc(
            9, R, n, idx.data_ptr(), kf.data_ptr(), kb.data_ptr(), y0.data_ptr(), RTOL, ATOL,
            TOUT, MXSTEP, LU_FP64, yout.data_ptr(), cnt.data_ptr(), stream.cuda_stream) == -1
        assert hip.SUNAMD_BatchedBDFSolveBucketNoAlloc(
            m, R, n, idx.data_ptr(), kf.data_ptr(), kb.data_ptr(), y0.data_ptr(), RTOL, ATOL,
            TOUT, MXSTEP, LU_FP64, None, cnt.data_ptr(), stream.cuda_stream) == -1
        # device check: a species index equal to m
        bad = idx.clone()
        bad[5] = m
        y, c = _run_no_alloc(m, R, n, bad, kf, kb, y0, LU_FP64, stream)
        assert np.all(c[:, 9] == -22)
        assert np.all(c[:, 0] == 0)
        assert np.array_equal(y, y0.cpu().numpy())
        # the next good call is unaffected by the rejected one
        y, c = _run_no_alloc(m, R, n, idx, kf, kb, y0, LU_FP64, stream)
        y_ref, c_ref = _run_allocating(b, LU_FP64)
        ye, ce = _expected(b, y_ref, c_ref, 0)
        assert np.array_equal(c, ce) and np.array_equal(y, ye)
    finally:
        free_batch(b)
