"""The training-pair log file (include/sunamd/learned_predictor.h): rows
written from the integrator's logging layout read back bit for bit, with the
header's metadata; partial reads; and the rejections the format promises
(foreign file, other version, truncated file, a mask that is not a prefix of
ones, rows out of range)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from sundials_amd import _lib

NIN, NOUT = 144, 64


def _rows(n, seed=7):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, NIN)).astype(np.float32)
    t = rng.standard_normal((n, NOUT)).astype(np.float32)
    m = rng.integers(1, NOUT + 1, n)
    mask = (np.arange(NOUT)[None, :] < m[:, None]).astype(np.float32)
    t *= mask
    return x, t, mask


def _write(path, x, t, mask, info=None):
    return _lib.core().SUNAMD_TrainingLogWrite(str(path).encode(), x.ctypes.data, t.ctypes.data,
                                               mask.ctypes.data, x.shape[0],
                                               C.byref(info) if info else None)


def _read(path, row0, n):
    x = np.zeros((n, NIN), np.float32)
    t = np.zeros((n, NOUT), np.float32)
    mask = np.zeros((n, NOUT), np.float32)
    rc = _lib.core().SUNAMD_TrainingLogRead(str(path).encode(), row0, n, x.ctypes.data,
                                            t.ctypes.data, mask.ctypes.data)
    return rc, x, t, mask


def test_round_trip_bitwise(tmp_path):
    x, t, mask = _rows(1000)
    x[3, 5] = np.float32("inf")
    x[4, 6] = np.float32(-0.0)
    info = _lib.TrainingLogInfo(0, 0, 2004, 1e-6, 1e-10, 3)
    p = tmp_path / "pairs.bin"
    assert _write(p, x, t, mask, info) == 0
    assert os.path.getsize(p) == 64 + 1000 * (4 + 4 * (NIN + NOUT))
    got = _lib.TrainingLogInfo()
    assert _lib.core().SUNAMD_TrainingLogInfoRead(str(p).encode(), C.byref(got)) == 0
    assert (got.version, got.rows, got.seed, got.rtol, got.atol, got.stride) == \
        (1, 1000, 2004, 1e-6, 1e-10, 3)
    rc, x2, t2, m2 = _read(p, 0, 1000)
    assert rc == 0
    assert x2.tobytes() == x.tobytes() and t2.tobytes() == t.tobytes()
    assert np.array_equal(m2, mask)
    rc, x3, t3, m3 = _read(p, 600, 400)
    assert rc == 0 and np.array_equal(x3, x[600:]) and np.array_equal(m3, mask[600:])
    assert _read(p, 600, 401)[0] == -7


def test_rejections(tmp_path):
    x, t, mask = _rows(10)
    p = tmp_path / "a.bin"
    assert _write(p, x, t, mask) == 0
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "bad.bin"
    b = bytearray(raw)
    b[0:8] = b"NOTALOG!"
    open(bad, "wb").write(b)
    assert _read(bad, 0, 1)[0] == -5
    b = bytearray(raw)
    b[8:12] = struct.pack("<I", 2)
    open(bad, "wb").write(b)
    assert _read(bad, 0, 1)[0] == -6
    open(bad, "wb").write(raw[:-10])
    assert _read(bad, 0, 10)[0] == -3
    b = bytearray(raw)
    b[64:68] = struct.pack("<I", 65)
    open(bad, "wb").write(b)
    assert _read(bad, 0, 1)[0] == -8
    mask2 = mask.copy()
    mask2[0, :] = 0.0
    mask2[0, 3] = 1.0
    assert _write(tmp_path / "c.bin", x, t, mask2) == -4
    assert _lib.core().SUNAMD_TrainingLogInfoRead(str(tmp_path / "none.bin").encode(),
                                                  C.byref(_lib.TrainingLogInfo())) == -2
