"""Shared helpers for the uneven-batch tests: batch creation, the device
integrator through the C-ABI, and the CPU oracle run of the same batch."""
import ctypes as C

import numpy as np

from sundials_amd import _lib

M_VALUES = (4, 8, 16, 24, 32, 48, 64)
WEIGHTS = (30.0, 25.0, 15.0, 10.0, 10.0, 5.0, 5.0)
RTOL, ATOL, TOUT = 1e-6, 1e-10, 1.0


def make_batch(nsys, seed, m_values=M_VALUES, weights=WEIGHTS, rfactor=1.5, sigma=0.2):
    core = _lib.core()
    n = len(m_values)
    b = core.SUNAMD_BatchCreate(nsys, seed, n, (C.c_int * n)(*m_values),
                                (C.c_double * n)(*weights), rfactor, sigma)
    assert b, "SUNAMD_BatchCreate failed"
    return b


def free_batch(b):
    _lib.core().SUNAMD_BatchFree(b)


def _out(b):
    bc = b.contents
    y = np.zeros(bc.ytotal)
    cnt = np.zeros((bc.nsys, _lib.BS_COUNT), dtype=np.int64)
    return y, cnt


MXSTEP = 100000  # the oracle's limit is 10^6; no test system comes near either


def run_device(b, precision, rtol=RTOL, atol=ATOL, tout=TOUT, mxstep=MXSTEP):
    """Integrate the batch with the persistent device integrator."""
    hip = _lib.hip()
    B = hip.SUNAMD_BatchedBDFCreate(b, rtol, atol, precision)
    assert B, "SUNAMD_BatchedBDFCreate failed"
    try:
        assert hip.SUNAMD_BatchedBDFSetMaxNumSteps(B, mxstep) == 0
        wall = C.c_double(0.0)
        rc = hip.SUNAMD_BatchedBDFRun(B, tout, C.byref(wall))
        assert rc == 0, "SUNAMD_BatchedBDFRun returned %d" % rc
        y, cnt = _out(b)
        rc = hip.SUNAMD_BatchedBDFGetResults(
            B, y.ctypes.data_as(C.POINTER(C.c_double)), cnt.ctypes.data_as(C.POINTER(C.c_long)))
        return y, cnt, rc, wall.value
    finally:
        hip.SUNAMD_BatchedBDFFree(B)


def run_oracle(b, nthreads=8, rtol=RTOL, atol=ATOL, tout=TOUT):
    """Integrate every system of the batch on the CPU oracle."""
    orc = _lib.load_oracle()
    y, cnt = _out(b)
    wall = C.c_double(0.0)
    rc = orc.oracle_batch_solve(b, rtol, atol, tout, nthreads,
                                y.ctypes.data_as(C.POINTER(C.c_double)),
                                cnt.ctypes.data_as(C.POINTER(C.c_long)), C.byref(wall))
    return y, cnt, rc, wall.value


def system_slices(b):
    """(bucket, cell, m, y-offset) of every system."""
    bc = b.contents
    out = []
    for s in range(bc.nsys):
        k = bc.sys_bucket[s]
        out.append((k, bc.sys_cell[s], bc.bucket[k].contents.m, bc.sys_yoff[s]))
    return out
