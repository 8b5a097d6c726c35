"""ctypes binding of oracle/liboracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        i, d = C.c_int, C.c_double
        lib.orc_Mss.argtypes = [_dp, _dp, _dp, i, i, i, d, d, i]
        lib.orc_Mbb.argtypes = [_dp, _dp, _dp, i, i, i, d, d, i]
        lib.orc_Msb.argtypes = [_dp, _dp, _dp, i, _dp, i, i, i, d, d, d, i]
        lib.orc_Mbs.argtypes = [_dp, _dp, _dp, i, _dp, i, i, i, d, d, d, i]
        lib.orc_barrier.argtypes = [_dp, _dp, _dp, i, _dp, i, i, d, d, d, i]
        lib.orc_ss_block.argtypes = [_dp, _dp, _dp, d, d]
        for f in (lib.orc_Mss, lib.orc_Mbb, lib.orc_Msb, lib.orc_Mbs,
                  lib.orc_barrier, lib.orc_ss_block):
            f.restype = None
        _lib = lib
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _c(a, cols):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, cols))


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


def Mss(xs, fs, a, mu=1.0, rows=None, nthreads=None):
    xs, fs = _c(xs, 3), _c(fs, 6)
    n = xs.shape[0]
    i0, i1 = rows if rows else (0, n)
    V = np.zeros((n, 6))
    load().orc_Mss(_p(V), _p(fs), _p(xs), n, i0, i1, a, mu, nthreads or nthreads_default())
    return V


def Mbb(xb, fb, a, mu=1.0, rows=None, nthreads=None):
    xb, fb = _c(xb, 3), _c(fb, 3)
    n = xb.shape[0]
    i0, i1 = rows if rows else (0, n)
    V = np.zeros((n, 3))
    load().orc_Mbb(_p(V), _p(fb), _p(xb), n, i0, i1, a, mu, nthreads or nthreads_default())
    return V


def Msb(xs, xb, fb, a_s, a_b, mu=1.0, V=None, rows=None, nthreads=None):
    """V_s += M_sb F_b (V defaults to zeros)."""
    xs, xb, fb = _c(xs, 3), _c(xb, 3), _c(fb, 3)
    ns, nb = xs.shape[0], xb.shape[0]
    V = np.zeros((ns, 6)) if V is None else np.ascontiguousarray(V, dtype=np.float64).copy()
    i0, i1 = rows if rows else (0, ns)
    load().orc_Msb(_p(V), _p(fb), _p(xs), ns, _p(xb), nb, i0, i1, a_s, a_b, mu,
                   nthreads or nthreads_default())
    return V


def Mbs(xb, xs, fs, a_s, a_b, mu=1.0, rows=None, nthreads=None):
    xb, xs, fs = _c(xb, 3), _c(xs, 3), _c(fs, 6)
    ns, nb = xs.shape[0], xb.shape[0]
    V = np.zeros((nb, 3))
    i0, i1 = rows if rows else (0, nb)
    load().orc_Mbs(_p(V), _p(fs), _p(xb), nb, _p(xs), ns, i0, i1, a_s, a_b, mu,
                   nthreads or nthreads_default())
    return V


def grand(xs, xb, fs, fb, a_s, a_b, mu=1.0, seg_rows=None, blob_rows=None, nthreads=None):
    """(V_s, V_b) = all four reference blocks."""
    vs = Mss(xs, fs, a_s, mu, rows=seg_rows, nthreads=nthreads)
    vs = Msb(xs, xb, fb, a_s, a_b, mu, V=vs, rows=seg_rows, nthreads=nthreads)
    vb = Mbs(xb, xs, fs, a_s, a_b, mu, rows=blob_rows, nthreads=nthreads)
    vb += Mbb(xb, fb, a_b, mu, rows=blob_rows, nthreads=nthreads)
    return vs, vb


def barrier(xs, xb, fs, nseg, a_s, a_b, K, nthreads=None):
    """Returns (f_segs + barrier, f_blobs_repulsion)."""
    xs, xb = _c(xs, 3), _c(xb, 3)
    fs = _c(fs, 6).copy()
    frep = np.zeros((xb.shape[0], 3))
    load().orc_barrier(_p(fs), _p(frep), _p(xs), xs.shape[0], _p(xb), xb.shape[0],
                       nseg, a_s, a_b, K, nthreads or nthreads_default())
    return fs, frep


def ss_block(xi, xj, a, mu=1.0):
    out = np.zeros(36)
    load().orc_ss_block(_p(out), _p(np.ascontiguousarray(xi, float)),
                        _p(np.ascontiguousarray(xj, float)), a, mu)
    return out.reshape(6, 6)
