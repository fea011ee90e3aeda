"""ctypes access to the C/OpenMP restatement of numba.py:197-207 (``wtilde_cpu.c``).

TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The library is built by
``make -C oracle``; both functions are pinned against ``restated.w_tilde_curvature_interferometer_from``
in tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle_wtilde.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C oracle`")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.oracle_w_tilde.restype = ctypes.c_int
        lib.oracle_w_tilde.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                                               ctypes.c_void_p]
        lib.oracle_w_tilde_rows.restype = ctypes.c_int
        lib.oracle_w_tilde_rows.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                                                    ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
        _lib = lib
    return _lib


def default_threads():
    return min(16, len(os.sched_getaffinity(0)))


def _inputs(noise_map_real, uv_wavelengths, grid_radians_slim):
    n = np.ascontiguousarray(noise_map_real, np.float64)
    uv = np.ascontiguousarray(uv_wavelengths, np.float64)
    g = np.ascontiguousarray(grid_radians_slim, np.float64)
    if uv.shape != (n.size, 2) or g.ndim != 2 or g.shape[1] != 2:
        raise ValueError("shapes: noise_map_real [K], uv_wavelengths [K, 2], grid_radians_slim [M, 2]")
    return n, uv, g


def w_tilde(noise_map_real, uv_wavelengths, grid_radians_slim, threads=None):
    """The full M x M matrix (parallel over visibilities with per-thread accumulators, as numba's prange)."""
    n, uv, g = _inputs(noise_map_real, uv_wavelengths, grid_radians_slim)
    M = g.shape[0]
    out = np.empty((M, M))
    rc = load().oracle_w_tilde(n.ctypes.data, uv.ctypes.data, g.ctypes.data, n.size, M, threads or default_threads(),
                               out.ctypes.data)
    if rc:
        raise MemoryError("oracle_w_tilde: per-thread accumulators could not be allocated")
    return out


def w_tilde_rows(noise_map_real, uv_wavelengths, grid_radians_slim, rows, threads=None):
    """Rows `rows` of the matrix, each element summed in visibility order."""
    n, uv, g = _inputs(noise_map_real, uv_wavelengths, grid_radians_slim)
    r = np.ascontiguousarray(rows, np.int64).reshape(-1)
    M = g.shape[0]
    out = np.empty((r.size, M))
    rc = load().oracle_w_tilde_rows(n.ctypes.data, uv.ctypes.data, g.ctypes.data, n.size, M, r.ctypes.data, r.size,
                                    threads or default_threads(), out.ctypes.data)
    if rc:
        raise ValueError("row index out of range")
    return out
