"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (cwb_oracle.c, wdm_oracle.c): the CPU
restatement of PycWB's Numba likelihood sky scan and of cWB's WDM forward
transform.  Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg use it, and only as the checker.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

vp = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        lib.oracle_calculate_dpf.restype = C.c_double
        lib.oracle_calculate_dpf.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_double,
                                             C.c_double, vp, C.c_int64, vp, C.c_int]
        lib.oracle_calculate_dpf_f64.restype = C.c_double
        lib.oracle_calculate_dpf_f64.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                                 C.c_double, C.c_double, vp, C.c_int64, C.c_int]
        lib.oracle_scan_sky.restype = None
        lib.oracle_scan_sky.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                        vp, vp, C.c_double, C.c_double, C.c_double, vp,
                                        C.c_int64, vp, vp, vp, vp, C.c_int]
        lib.oracle_scan_sky_f64.restype = None
        lib.oracle_scan_sky_f64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                            vp, vp, vp, C.c_double, C.c_double, vp, C.c_int64,
                                            vp, vp, vp, C.c_int]
        lib.wdm_oracle_cos_size.restype = C.c_int
        lib.wdm_oracle_cos_size.argtypes = [C.c_int]
        lib.wdm_oracle_fourier.restype = C.c_int
        lib.wdm_oracle_fourier.argtypes = [C.c_int, vp]
        lib.wdm_oracle_filter.restype = C.c_int
        lib.wdm_oracle_filter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int]
        lib.wdm_oracle_t2w.restype = C.c_int64
        lib.wdm_oracle_t2w.argtypes = [vp, C.c_int64, C.c_int, vp, C.c_int, C.c_int, vp, vp]
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(vp)


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def calculate_dpf(FP, FX, rms, n_sky, n_ifo, gamma_regulator, En, valid, threads=0,
                  return_quality=False):
    """dpf.py:7-34, Numba-faithful mixed precision. rms is (n_pix, n_ifo)."""
    FP, FX, rms = _c(FP, np.float32), _c(FX, np.float32), _c(rms, np.float32)
    valid = _c(valid, np.int64)
    q = np.zeros(valid.size)
    r = load().oracle_calculate_dpf(n_ifo, rms.shape[0], n_sky, _p(FP), _p(FX), _p(rms),
                                    gamma_regulator, En, _p(valid), valid.size, _p(q), threads)
    return (r, q) if return_quality else r


def calculate_dpf_f64(FP, FX, rms, n_sky, n_ifo, gamma_regulator, En, valid, threads=0):
    FP, FX, rms = _c(FP, np.float64), _c(FX, np.float64), _c(rms, np.float64)
    valid = _c(valid, np.int64)
    return load().oracle_calculate_dpf_f64(n_ifo, rms.shape[0], n_sky, _p(FP), _p(FX), _p(rms),
                                           gamma_regulator, En, _p(valid), valid.size, threads)


def scan_sky(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, REG, netCC, delta_regulator,
             En, valid, threads=0, return_best=False):
    """sky_scan.py:17-192, mixed precision.  Returns (l_max, maps(11,S), stat_max)."""
    FP, FX, rms = _c(FP, np.float32), _c(FX, np.float32), _c(rms, np.float32)
    td00, td90 = _c(td00, np.float32), _c(td90, np.float32)
    ml, REG, valid = _c(ml, np.int32), _c(REG, np.float32), _c(valid, np.int64)
    maps = np.zeros((11, n_sky), dtype=np.float32)
    best = np.zeros(n_sky, dtype=np.float32)
    lm = np.zeros(1, dtype=np.int64)
    sm = np.zeros(1, dtype=np.float64)
    load().oracle_scan_sky(n_ifo, n_pix, n_sky, td00.shape[0], _p(FP), _p(FX), _p(rms),
                           _p(td00), _p(td90), _p(ml), _p(REG), netCC, delta_regulator, En,
                           _p(valid), valid.size, _p(maps), _p(best), _p(lm), _p(sm), threads)
    if return_best:
        return int(lm[0]), maps, float(sm[0]), best
    return int(lm[0]), maps, float(sm[0])


def scan_sky_f64(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, REG64, netCC, En, valid,
                 threads=0):
    """float64 yardstick of the scan; REG64 = (delta*sqrt(2), regulator) unrounded."""
    FP, FX, rms = _c(FP, np.float32), _c(FX, np.float32), _c(rms, np.float32)
    td00, td90 = _c(td00, np.float32), _c(td90, np.float32)
    ml, REG64, valid = _c(ml, np.int32), _c(REG64, np.float64), _c(valid, np.int64)
    maps = np.zeros((11, n_sky), dtype=np.float64)
    lm = np.zeros(1, dtype=np.int64)
    sm = np.zeros(1, dtype=np.float64)
    load().oracle_scan_sky_f64(n_ifo, n_pix, n_sky, td00.shape[0], _p(FP), _p(FX), _p(rms),
                               _p(td00), _p(td90), _p(ml), _p(REG64), netCC, En, _p(valid),
                               valid.size, _p(maps), _p(lm), _p(sm), threads)
    return int(lm[0]), maps, float(sm[0])


def wdm_fourier(beta):
    n = load().wdm_oracle_cos_size(beta)
    out = np.zeros(n)
    load().wdm_oracle_fourier(beta, _p(out))
    return out


def wdm_filter(M, K, beta, precision, max_taps=400000):
    buf = np.zeros(max_taps)
    n = load().wdm_oracle_filter(M, K, beta, precision, _p(buf), max_taps)
    if n < 0:
        raise ValueError("filter longer than max_taps")
    return buf[:n].copy()


def wdm_t2w(x, M, filt, MM=-1):
    """Returns (map00, map90) with shape (N, M+1) each, float64."""
    x = _c(x, np.float64)
    filt = _c(filt, np.float64)
    stride = M if MM <= 0 else MM
    L = x.size
    N = (L + (stride - L % stride if L % stride else 0)) // stride
    m00 = np.zeros((N, M + 1))
    m90 = np.zeros((N, M + 1))
    n = load().wdm_oracle_t2w(_p(x), L, M, _p(filt), filt.size, MM, _p(m00), _p(m90))
    assert n == N
    return m00, m90
