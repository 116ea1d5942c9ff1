"""ctypes loader of oracle/libcpu_fp32.so, the FP32 host implementation of
the BSSN scheme (TEST INFRASTRUCTURE ONLY: bench.py's CPU baseline and the
tests).  See am_cpu_fp32.c."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "libcpu_fp32.so"
_P = C.c_void_p
_SIGS = {
    "cpu32_set_threads": (None, [C.c_int]),
    "cpu32_max_threads": (C.c_int, []),
    "cpu32_bssn_stage": (None, [_P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                C.c_int, _P, C.c_float, C.c_int]),
    "cpu32_equal_ghosts": (None, [_P, C.c_int, _P, _P, C.c_int, C.c_int, C.c_int]),
    "cpu32_rk4_step": (None, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _P,
                              C.c_float, C.c_int]),
    "cpu32_derivatives": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float]),
}
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import oracle
            oracle.build()
        L = C.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def rk4_step(yn, a, b, acc, bnd, tiles, S, P, q, inv_dx, dt, params, sigma=0.0, upwind=False):
    """One RK4 step of the listed tiles (lists of 24 float32 padded arrays;
    the result lands in b).  tiles: int32 array; bnd: uint64 boundaries."""
    import numpy as np
    prm = np.ascontiguousarray(params, np.float64)
    lib().cpu32_rk4_step(ptrs(yn), ptrs(a), ptrs(b), ptrs(acc), bnd.ctypes.data, tiles.ctypes.data, len(tiles), S, P,
                         q, inv_dx, dt, prm.ctypes.data, sigma, 1 if upwind else 0)
