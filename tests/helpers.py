"""Shared helpers for the parity tests: build HIP and serial vectors from
numpy arrays and read them back."""
import ctypes as C

import numpy as np

from sundials_amd import _lib

PD = C.POINTER(C.c_double)


def dp(a):
    return a.ctypes.data_as(PD)


def gpu_vec(ctx, vals):
    hip = _lib.hip()
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    v = hip.N_VNew_Mi355(len(vals), ctx)
    assert v, "N_VNew_Mi355 failed"
    if len(vals):
        assert hip.N_VCopyFromHostArray_Mi355(v, dp(vals)) == 0
    return v


def gpu_get(v):
    n = _lib.core().N_VGetLength(v)
    out = np.zeros(n)
    if n:
        assert _lib.hip().N_VCopyToHostArray_Mi355(v, dp(out)) == 0
    return out


def cpu_vec(orc, ctx, vals):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    v = orc.N_VNew_Serial(len(vals), ctx)
    if len(vals):
        C.memmove(_lib.core().N_VGetArrayPointer(v), dp(vals), 8 * len(vals))
    return v


def cpu_get(v):
    core = _lib.core()
    n = core.N_VGetLength(v)
    if n == 0:
        return np.zeros(0)
    return np.ctypeslib.as_array(core.N_VGetArrayPointer(v), shape=(n,)).copy()


def vec_array(vs):
    return (C.c_void_p * len(vs))(*vs)
