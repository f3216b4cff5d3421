"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product path never does.
"""
import ctypes
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(_ROOT, "oracle", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_coeff_count.restype = ctypes.c_long
        L.oracle_coeff_count.argtypes = [ctypes.c_int]
        L.oracle_coeff_index.restype = ctypes.c_long
        L.oracle_coeff_index.argtypes = [ctypes.c_int] * 3
        L.oracle_seeded_coeffs.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, dp]
        L.oracle_quadrature_weights_d.argtypes = [ctypes.c_int, dp]
        L.oracle_wigner_d_d.argtypes = [ctypes.c_int] * 4 + [dp]
        L.oracle_fft2_layers.argtypes = [ctypes.c_int] * 3 + [dp, ctypes.c_int, ctypes.c_int]
        L.oracle_dsoft.restype = ctypes.c_int
        L.oracle_dsoft.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.oracle_idsoft.restype = ctypes.c_int
        L.oracle_idsoft.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def coeff_count(B):
    return int(lib().oracle_coeff_count(B))


def coeff_index(l, M, Mp):
    return int(lib().oracle_coeff_index(l, M, Mp))


def seeded_coeffs(B, seed, power=1.0):
    """Packed complex128 coefficients from the SplitMix64 generator."""
    out = np.zeros(2 * coeff_count(B), dtype=np.float64)
    lib().oracle_seeded_coeffs(B, seed, power, _p(out))
    return out.view(np.complex128)


def quadrature_weights(B):
    w = np.zeros(2 * B, dtype=np.float64)
    lib().oracle_quadrature_weights_d(B, _p(w))
    return w


def wigner_d(B, M, Mp, weighted=False):
    """(B - max|M|,|M'|) x 2B table exactly as the reference builds it
    (including its -1 base sign), rounded to double."""
    R = B - max(abs(M), abs(Mp))
    out = np.zeros(R * 2 * B, dtype=np.float64)
    lib().oracle_wigner_d_d(B, M, Mp, 1 if weighted else 0, _p(out))
    return out.reshape(2 * B, R).T.copy()  # column-major -> [row, col]


def fft2_layers(arr, sign, threads=1):
    """In-place layer-wise 2D DFT of a complex128 array shaped (lays, n0, n1)."""
    a = np.array(arr, dtype=np.complex128, copy=True, order="C")
    lays, n0, n1 = a.shape
    f = a.view(np.float64).reshape(-1)
    lib().oracle_fft2_layers(n0, n1, lays, _p(f), sign, threads)
    return a


def dsoft(B, grid, threads=None):
    """grid: complex128 array of N^3 elements in reference layout."""
    g = np.ascontiguousarray(grid, dtype=np.complex128).reshape(-1)
    assert g.size == (2 * B) ** 3
    out = np.zeros(coeff_count(B), dtype=np.complex128)
    t = threads or lib().oracle_max_threads()
    rc = lib().oracle_dsoft(B, _p(g.view(np.float64)), _p(out.view(np.float64)), t)
    assert rc == 0
    return out


def idsoft(B, coeffs, grid=None, threads=None):
    """Returns the synthesised grid; grid (if given) supplies the entries the
    inverse never writes (row or column index B), as in the reference."""
    c = np.ascontiguousarray(coeffs, dtype=np.complex128).reshape(-1)
    assert c.size == coeff_count(B)
    N = 2 * B
    if grid is None:
        g = np.zeros(N ** 3, dtype=np.complex128)
    else:
        g = np.array(grid, dtype=np.complex128, copy=True).reshape(-1)
    t = threads or lib().oracle_max_threads()
    rc = lib().oracle_idsoft(B, _p(c.view(np.float64)), _p(g.view(np.float64)), t)
    assert rc == 0
    return g
