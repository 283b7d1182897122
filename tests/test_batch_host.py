"""Uneven batches on the host: the batch generator and the oracle's
per-system driver.

The generator must give the requested size mix (m drawn with the given
weights, R = round(1.5 m)), reproducibly from its seed.  The oracle's batch
driver (every system its own CVode run, results scattered to system order)
must agree with integrating a chosen system on its own through the coupled
driver with a single cell -- the same integrator on a one-block matrix, so
the counters and the state agree exactly.
"""
import ctypes as C

import numpy as np

from sundials_amd import _lib
from batch_util import M_VALUES, WEIGHTS, RTOL, ATOL, TOUT, make_batch, free_batch, \
    run_oracle, system_slices


def test_batch_generator_mix_and_layout():
    b = make_batch(40000, 123)
    try:
        bc = b.contents
        assert bc.nbuckets == len(M_VALUES)
        sb = np.ctypeslib.as_array(bc.sys_bucket, shape=(bc.nsys,))
        frac = np.bincount(sb, minlength=len(M_VALUES)) / bc.nsys
        want = np.array(WEIGHTS) / sum(WEIGHTS)
        assert np.all(np.abs(frac - want) < 0.01), (frac, want)
        ms = np.array([bc.bucket[k].contents.m for k in range(bc.nbuckets)])
        Rs = np.array([bc.bucket[k].contents.R for k in range(bc.nbuckets)])
        assert tuple(ms) == M_VALUES
        assert np.all(Rs == np.floor(1.5 * ms + 0.5))
        # cells of a bucket are numbered in system order; offsets are prefix sums
        cells = np.ctypeslib.as_array(bc.sys_cell, shape=(bc.nsys,))
        yoff = np.ctypeslib.as_array(bc.sys_yoff, shape=(bc.nsys,))
        for k in range(bc.nbuckets):
            sel = np.nonzero(sb == k)[0]
            assert np.array_equal(cells[sel], np.arange(sel.size))
            assert bc.bucket[k].contents.ncells == sel.size
        msys = ms[sb]
        assert yoff[0] == 0 and np.array_equal(yoff[1:], np.cumsum(msys)[:-1])
        assert bc.ytotal == msys.sum()
        # same seed -> same batch
        b2 = make_batch(40000, 123)
        try:
            sb2 = np.ctypeslib.as_array(b2.contents.sys_bucket, shape=(bc.nsys,))
            assert np.array_equal(sb, sb2)
            k = 3
            p1, p2 = bc.bucket[k].contents, b2.contents.bucket[k].contents
            n = p1.ncells * p1.R
            assert np.array_equal(np.ctypeslib.as_array(p1.kf, shape=(n,)),
                                  np.ctypeslib.as_array(p2.kf, shape=(n,)))
        finally:
            free_batch(b2)
    finally:
        free_batch(b)


def test_batch_generator_empty_bucket():
    # a weight of zero leaves a bucket with no systems
    b = make_batch(300, 5, m_values=(4, 8, 16), weights=(1.0, 0.0, 1.0))
    try:
        bc = b.contents
        assert bc.bucket[1].contents.ncells == 0
        sb = np.ctypeslib.as_array(bc.sys_bucket, shape=(bc.nsys,))
        assert not np.any(sb == 1)
        y, cnt, rc, _ = run_oracle(b, nthreads=4)
        assert rc == 0 and np.all(cnt[:, 9] == 0)
    finally:
        free_batch(b)


def _single_system_view(p, c):
    """A one-cell MassAction that aliases cell c of bucket network p."""
    v = _lib.MassAction()
    for name, _ in _lib.MassAction._fields_:
        setattr(v, name, getattr(p, name))
    v.ncells = 1
    dbl = C.sizeof(C.c_double)
    v.kf = C.cast(C.cast(p.kf, C.c_void_p).value + c * p.R * dbl, C.POINTER(C.c_double))
    v.kb = C.cast(C.cast(p.kb, C.c_void_p).value + c * p.R * dbl, C.POINTER(C.c_double))
    v.y0 = C.cast(C.cast(p.y0, C.c_void_p).value + c * p.m * dbl, C.POINTER(C.c_double))
    return v


def test_oracle_batch_driver_matches_single_system_runs():
    orc = _lib.load_oracle()
    b = make_batch(160, 2024)
    try:
        y, cnt, rc, _ = run_oracle(b, nthreads=4)
        assert rc == 0
        assert np.all(cnt[:, 9] == 0)
        slices = system_slices(b)
        picks = {}
        for s, (k, c, m, off) in enumerate(slices):
            picks.setdefault(k, s)  # first system of every bucket
        picks = sorted(picks.values()) + [len(slices) - 1]
        for s in picks:
            k, c, m, off = slices[s]
            v = _single_system_view(b.contents.bucket[k].contents, c)
            ys = np.zeros(m)
            st = np.zeros(16)
            assert orc.oracle_massaction_solve_coupled(
                C.byref(v), RTOL, ATOL, TOUT, ys.ctypes.data_as(C.POINTER(C.c_double)),
                st.ctypes.data_as(C.POINTER(C.c_double))) == 0
            assert list(cnt[s, :8]) == [int(x) for x in st[:8]], s
            assert cnt[s, 8] == int(st[8])
            assert np.array_equal(y[off:off + m], ys), s
    finally:
        free_batch(b)
