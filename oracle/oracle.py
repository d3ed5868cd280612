"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so``: the CPU restatement of PycWB's Numba
likelihood sky scan and best-sky statistics (cwb_oracle.c), cWB's WDM forward
transform (wdm_oracle.c), the continuous-direction statistic (grad_oracle.c),
the TD filters and amplitudes (td_oracle.c) and the sub-network cut
(subnet_oracle.c).  Only tests, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg use it, and only as the checker.
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


def dpf_terms(FP, FX, rms, dirs, threads=0):
    """Per-pixel float32 (fp, fx, ni) of dpf_np_loops_vec for the listed
    directions, each (len(dirs), P)."""
    FP, FX, rms = _c(FP, np.float32), _c(FX, np.float32), _c(rms, np.float32)
    dirs = _c(dirs, np.int64)
    P, n = rms.shape
    out = [np.zeros((dirs.size, P), np.float32) for _ in range(3)]
    fn = load().oracle_dpf_terms
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_int] + [vp] * 4 + [C.c_int64] + [vp] * 3 + [C.c_int]
    fn(n, P, _p(FP), _p(FX), _p(rms), _p(dirs), dirs.size, _p(out[0]), _p(out[1]), _p(out[2]),
       threads)
    return tuple(out)


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


SKYSTAT_SCALARS = ("Gn", "Ec", "Dc", "Rc", "Eh", "Es", "Np", "Em", "Lm", "norm", "cc", "rho",
                   "xrho", "Lo", "Eo", "N_pix_effective")


def self_xtalk(n_pix):
    """Self-only cross-talk CSR: each pixel its own neighbour, coefficients (1,0,0,1)."""
    lookup = np.stack([np.arange(n_pix), np.arange(1, n_pix + 1)], axis=1).astype(np.int32)
    xt = np.zeros((n_pix, 8), dtype=np.float32)
    xt[:, 0] = np.arange(n_pix)
    xt[:, 4] = 1.0
    xt[:, 7] = 1.0
    return lookup, xt


def sky_statistics(l, n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, REG, En, lookup=None,
                   xt=None, xgb_rho_mode=False):
    """compute_statistics_at_sky_position (sky_statistics.py:26-290)."""
    lib = load()
    fn = lib.oracle_sky_statistics
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64] + [vp] * 7 + [C.c_double] + \
        [vp, vp, C.c_int, vp] + [vp] * 13
    if lookup is None:
        lookup, xt = self_xtalk(n_pix)
    FP, FX, rms = _c(FP, np.float32), _c(FX, np.float32), _c(rms, np.float32)
    td00, td90 = _c(td00, np.float32), _c(td90, np.float32)
    ml, REG = _c(ml, np.int32), _c(REG, np.float32)
    lookup, xt = _c(lookup, np.int32), _c(xt, np.float32)
    scal = np.zeros(16)
    arr = {k: np.zeros((n_ifo, n_pix)) for k in ("pd", "pD", "ps", "pS")}
    arr.update({k: np.zeros((n_ifo, n_pix), np.float32) for k in ("nul00", "nul90")})
    arr.update({k: np.zeros(n_pix, np.float32) for k in ("pixel_mask", "gn", "ec", "ee", "EE")})
    arr["S_snr"] = np.zeros(n_ifo)
    pol = np.zeros((4, n_pix), np.float32)
    fn(n_ifo, n_pix, n_sky, td00.shape[0], int(l), _p(FP), _p(FX), _p(rms), _p(td00), _p(td90),
       _p(ml), _p(REG), En, _p(lookup), _p(xt), int(bool(xgb_rho_mode)), _p(scal),
       _p(arr["pd"]), _p(arr["pD"]), _p(arr["ps"]), _p(arr["pS"]), _p(arr["nul00"]),
       _p(arr["nul90"]), _p(arr["pixel_mask"]), _p(arr["gn"]), _p(arr["ec"]), _p(arr["ee"]),
       _p(arr["EE"]), _p(arr["S_snr"]), _p(pol))
    out = dict(zip(SKYSTAT_SCALARS, scal.tolist()))
    out.update(arr)
    out.update({"p00_POL_r": pol[0], "p00_POL_a": pol[1], "p90_POL_r": pol[2],
                "p90_POL_a": pol[3]})
    return out


def packet_norm(p, q, xtalks, lookup, mk, q_E):
    lib = load()
    fn = lib.oracle_packet_norm
    fn.restype = None
    fn.argtypes = [vp] * 5 + [C.c_int, C.c_int] + [vp] * 5
    p, q = _c(p, np.float64), _c(q, np.float64)
    n, P = p.shape
    xt = _c(xtalks, np.float32).reshape(-1, 8)
    lk = _c(lookup, np.int32)
    mk, q_E = _c(mk, np.float32), _c(q_E, np.float32)
    det, norm, rn, qn = np.zeros(n), np.zeros(n), np.zeros(P), np.zeros((n, P))
    fn(_p(p), _p(q), _p(lk), _p(xt) if xt.size else vp(), _p(mk), n, P, _p(q_E), _p(det),
       _p(norm), _p(rn), _p(qn))
    return det, norm, rn, qn


def xtalk_sum(p, q, xtalks, lookup, mk):
    lib = load()
    fn = lib.oracle_xtalk_sum
    fn.restype = C.c_double
    fn.argtypes = [vp] * 5 + [C.c_int, C.c_int]
    p, q = _c(p, np.float64), _c(q, np.float64)
    n, P = p.shape
    xt = _c(xtalks, np.float32).reshape(-1, 8)
    lk = _c(lookup, np.int32)
    mk = _c(mk, np.float32)
    return fn(_p(p), _p(q), _p(lk), _p(xt) if xt.size else vp(), _p(mk), n, P)


# ---- continuous-direction sky statistic (grad_oracle.c) ----

CS_STEP = 1e-30   # complex step


def direction_terms(geo, theta, phi):
    """(tau, fp, fx) float64 of the n detectors in ``geo`` ((n, 12) float64)."""
    lib = load()
    fn = lib.oracle_direction_terms
    fn.restype = None
    fn.argtypes = [C.c_int, vp, C.c_double, C.c_double, vp, vp, vp]
    geo = _c(geo, np.float64).reshape(-1, 12)
    n = geo.shape[0]
    tau, fp, fx = np.zeros(n), np.zeros(n), np.zeros(n)
    fn(n, _p(geo), theta, phi, _p(tau), _p(fp), _p(fx))
    return tau, fp, fx


class SkyObjective:
    """AA(theta, phi; amplitudes) of one cluster in float64, with complex-step
    derivatives.  rms (P, n) float32, td (P, n, 2K+1, 2) float32, geo (n, 12)."""

    def __init__(self, rms, td, geo, reg0, reg1, En):
        self.rms = _c(rms, np.float32)
        self.td = _c(td, np.float32)
        self.P, self.n, nd, _ = self.td.shape
        self.K = (nd - 1) // 2
        self.geo = _c(geo, np.float64).reshape(self.n, 12)
        self.reg0, self.reg1, self.En = float(reg0), float(reg1), float(En)
        lib = load()
        self._fn = lib.oracle_sky_objective
        self._fn.restype = None
        self._fn.argtypes = ([C.c_int] * 3 + [vp] * 3 + [C.c_double] * 7 + [C.c_int64] +
                             [C.c_double] * 2 + [vp])
        self._at = lib.oracle_sky_objective_at
        self._at.restype = C.c_double
        self._at.argtypes = [C.c_int] * 3 + [vp] * 5 + [C.c_double] * 3

    def _call(self, th, thi, ph, phi_, idx=-1, dre=0.0, dim=0.0):
        out = np.zeros(2)
        self._fn(self.n, self.P, self.K, _p(self.rms), _p(self.td), _p(self.geo), self.reg0,
                 self.reg1, self.En, th, thi, ph, phi_, idx, dre, dim, _p(out))
        return out

    def value(self, theta, phi, td_index=-1, td_delta=0.0):
        return float(self._call(theta, 0.0, phi, 0.0, td_index, td_delta)[0])

    def grad(self, theta, phi):
        """(dAA/dtheta, dAA/dphi) by complex step."""
        gt = self._call(theta, CS_STEP, phi, 0.0)[1] / CS_STEP
        gp = self._call(theta, 0.0, phi, CS_STEP)[1] / CS_STEP
        return np.array([gt, gp])

    def grad_td(self, theta, phi, td_index):
        """dAA/d(td flat float entry) by complex step."""
        return float(self._call(theta, 0.0, phi, 0.0, td_index, 0.0, CS_STEP)[1] / CS_STEP)

    def at(self, tau, fp, fx):
        """AA at given delays (taps) and patterns."""
        tau, fp, fx = (_c(a, np.float64) for a in (tau, fp, fx))
        return float(self._at(self.n, self.P, self.K, _p(self.rms), _p(self.td), _p(tau),
                              _p(fp), _p(fx), self.reg0, self.reg1, self.En))


# ---- time-delay amplitudes (td_oracle.c) ----

def td_fourier(beta, which):
    """which 0: Cos2 table, 1: SinCos table (cosine-series coefficients)."""
    lib = load()
    lib.td_oracle_size.restype = C.c_int
    lib.td_oracle_size.argtypes = [C.c_int, C.c_int]
    n = lib.td_oracle_size(beta, which)
    out = np.zeros(n)
    lib.td_oracle_fourier.restype = C.c_int
    lib.td_oracle_fourier.argtypes = [C.c_int, C.c_int, vp]
    lib.td_oracle_fourier(beta, which, _p(out))
    return out


def td_filters(M, K, beta, n_coeffs, L):
    """(T0, Tx) float64 (2J+1, 2 n_coeffs+1), J = M L."""
    lib = load()
    fn = lib.td_oracle_filters
    fn.restype = None
    fn.argtypes = [C.c_int] * 5 + [vp, vp]
    J = M * L
    T0 = np.zeros((2 * J + 1, 2 * n_coeffs + 1))
    Tx = np.zeros_like(T0)
    fn(M, K, beta, n_coeffs, L, _p(T0), _p(Tx))
    return T0, Tx


def td_vectors(pixel_indices, p00, p90, T0, Tx, M, n_coeffs, K, J):
    """(P, 4K+2) float32 rows [a00(-K..K), a90(-K..K)]."""
    lib = load()
    fn = lib.td_oracle_vectors
    fn.restype = None
    fn.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                   vp]
    pix = _c(pixel_indices, np.int32)
    p00, p90 = _c(p00, np.float32), _c(p90, np.float32)
    T0, Tx = _c(T0, np.float64), _c(Tx, np.float64)
    out = np.zeros((pix.size, 4 * K + 2), dtype=np.float32)
    fn(_p(pix), pix.size, _p(p00), _p(p90), p00.shape[0], _p(T0), _p(Tx), M, n_coeffs, K, J,
       _p(out))
    return out


# ---- sub-network cut (subnet_oracle.c) ----

def _subnet_common(FP, FX, rms, td00, td90, ml):
    FP, FX = _c(FP, np.float32), _c(FX, np.float32)
    rms = _c(rms, np.float32)
    td00, td90 = _c(td00, np.float32), _c(td90, np.float32)
    ml = _c(ml, np.int32)
    return FP, FX, rms, td00, td90, ml


def sky_loc_from_td(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, En, e2or, subcut):
    """optimze_sky_loc_from_td: (l_max, stat, Em, Am, lm, Vm, suball, EE).
    FP/FX (S, n), rms (P, n), td00/td90 (nd, n, P), ml (n, S); En float32."""
    lib = load()
    fn = lib.subnet_oracle_sky_loc_from_td
    fn.restype = None
    fn.argtypes = [C.c_int] * 3 + [vp] * 5 + [C.c_int, vp, C.c_float, C.c_double, C.c_double,
                                              vp]
    FP, FX, rms, td00, td90, ml = _subnet_common(FP, FX, rms, td00, td90, ml)
    out = np.zeros(8)
    fn(n_ifo, n_pix, n_sky, _p(FP), _p(FX), _p(rms), _p(td00), _p(td90), td00.shape[0], _p(ml),
       float(np.float32(En)), float(e2or), float(subcut), _p(out))
    return _sky_tuple(out)


def _sky_tuple(out):
    return (int(out[0]), float(out[1]), float(out[2]), float(out[3]), int(out[4]), int(out[5]),
            float(out[6]), float(out[7]))


def sky_loc(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, td_energy, ml, En, e2or, subcut):
    """optimze_sky_loc (the variant with a per-detector energy table)."""
    lib = load()
    fn = lib.subnet_oracle_sky_loc
    fn.restype = None
    fn.argtypes = [C.c_int] * 3 + [vp] * 6 + [C.c_int, vp, C.c_float, C.c_double, C.c_double,
                                              vp]
    FP, FX, rms, td00, td90, ml = _subnet_common(FP, FX, rms, td00, td90, ml)
    te = _c(td_energy, np.float32)
    out = np.zeros(8)
    fn(n_ifo, n_pix, n_sky, _p(FP), _p(FX), _p(rms), _p(td00), _p(td90), _p(te), td00.shape[0],
       _p(ml), float(np.float32(En)), float(e2or), float(subcut), _p(out))
    return _sky_tuple(out)


def mra_from_td(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, En, e2or, xtalks, lookup,
                l_max, components=False):
    """mra_statistics_from_td: (submra, rHo, Eo, Lo, Ls, m); with components=True
    also (amp, AMP, n_extracted)."""
    lib = load()
    fn = lib.subnet_oracle_mra_from_td
    fn.restype = C.c_int
    fn.argtypes = [C.c_int] * 3 + [vp] * 5 + [C.c_int, vp, C.c_float, C.c_double, vp, vp,
                                              C.c_int, vp, vp, vp]
    FP, FX, rms, td00, td90, ml = _subnet_common(FP, FX, rms, td00, td90, ml)
    xt = _c(xtalks, np.float32).reshape(-1, 8)
    lk = _c(lookup, np.int32).reshape(-1, 2)
    out = np.zeros(6)
    amp = np.zeros((n_ifo, n_pix), np.float32)
    AMP = np.zeros((n_ifo, n_pix), np.float32)
    k = fn(n_ifo, n_pix, n_sky, _p(FP), _p(FX), _p(rms), _p(td00), _p(td90), td00.shape[0],
           _p(ml), float(np.float32(En)), float(e2or), _p(xt), _p(lk), int(l_max), _p(out),
           _p(amp), _p(AMP))
    res = (float(out[0]), float(out[1]), float(out[2]), float(out[3]), float(out[4]),
           int(out[5]))
    return (res, amp, AMP, k) if components else res


def reduced_likelihood(Fp0, Fx0, rms, v00, v90):
    """Lo = sum over pixels of sse_like_ps at the DPF of (Fp0, Fx0, rms); (m, n) inputs."""
    lib = load()
    fn = lib.subnet_oracle_reduced_likelihood
    fn.restype = C.c_float
    fn.argtypes = [vp] * 5 + [C.c_int, C.c_int]
    Fp0, Fx0 = _c(Fp0, np.float32), _c(Fx0, np.float32)
    rms, v00, v90 = _c(rms, np.float32), _c(v00, np.float32), _c(v90, np.float32)
    m, n = rms.shape
    return float(fn(_p(Fp0), _p(Fx0), _p(rms), _p(v00), _p(v90), m, n))


# ---- cross-talk lookups (xtalk_oracle.c) ----

def _catalog(layers, table, coeff):
    layers = _c(layers, np.int32)
    table = _c(table, np.int32)
    coeff = _c(coeff, np.float32).reshape(-1, 5)
    return layers, table, coeff, int(layers.size), int(table.shape[2] - 1)


def xtalk_get(nlayer1, indx1, nlayer2, indx2, layers, table, coeff):
    """getXTalk: (c1, c2, c3, c4) float32, 3s when the pair has no entry."""
    lib = load()
    fn = lib.xtalk_oracle_get
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int64, vp, C.c_int, C.c_int, vp, vp, vp]
    layers, table, coeff, nres, maxl = _catalog(layers, table, coeff)
    out = np.zeros(4, np.float32)
    if fn(int(nlayer1), int(indx1), int(nlayer2), int(indx2), _p(layers), nres, maxl,
          _p(table), _p(coeff), _p(out)) != 0:
        raise ValueError("resolution not found in the catalog layers")
    return out


def _pixels_call(name, nlayer, time, layers, table, coeff, extra=()):
    lib = load()
    fn = getattr(lib, name)
    fn.restype = C.c_int64
    layers, table, coeff, nres, maxl = _catalog(layers, table, coeff)
    nl = _c(nlayer, np.int32)
    tm = _c(time, np.int64)
    P = nl.size
    lookup = np.zeros((P, 2), np.int32)
    cap = max(1, 64 * P)
    while True:
        xt = np.zeros((cap, 8), np.float32)
        args = [P, _p(nl), _p(tm), _p(layers), nres, maxl, _p(table), _p(coeff)]
        types = [C.c_int, vp, vp, vp, C.c_int, C.c_int, vp, vp]
        if extra:
            args.append(coeff.shape[0])
            types.append(C.c_int64)
        fn.argtypes = types + [vp, vp, C.c_int64]
        n = fn(*args, _p(lookup), _p(xt), cap)
        if n >= 0:
            return lookup, xt[:n].copy()
        if cap > 4 * P * P + 64:
            raise ValueError("resolution not found in the catalog layers")
        cap *= 4


def xtalk_pixels_pairwise(nlayer, time, layers, table, coeff):
    """getXTalk_pixels_numba: (lookup (P, 2) int32, xt (m, 8) float32)."""
    return _pixels_call("xtalk_oracle_pixels_pairwise", nlayer, time, layers, table, coeff)


def xtalk_pixels_fast(nlayer, time, layers, table, coeff):
    """getXTalk_pixels_fast: (lookup (P, 2) int32, xt (m, 8) float32)."""
    return _pixels_call("xtalk_oracle_pixels_fast", nlayer, time, layers, table, coeff,
                        extra=(True,))


def wdm_w2t(map_complex, M, filt, n_out, quad=False):
    """WDM<T>::w2t (quad=False: the 00 coefficients, the real parts) or w2tQ
    (quad=True: the 90 coefficients, the imaginary parts) of a complex
    (M+1, N) map; returns n_out float64 samples."""
    lib = load()
    fn = lib.wdm_oracle_w2t
    fn.restype = C.c_int
    fn.argtypes = [vp, C.c_int64, C.c_int, vp, C.c_int, C.c_int, C.c_int64, vp]
    mp = np.asarray(map_complex)
    col = np.ascontiguousarray((mp.imag if quad else mp.real).T, dtype=np.float64)
    filt = _c(filt, np.float64)
    out = np.zeros(int(n_out))
    fn(_p(col), col.shape[0], int(M), _p(filt), filt.size, 1 if quad else 0, int(n_out), _p(out))
    return out


# ---- network-pixel selection (select_oracle.c) ----

def select_pixels(energy, shift_bins, veto, edge_bins, ib, ie, eo):
    """One lag of select_network_pixels: dict with live_mask, mask, frequency,
    time, energy, pix_det_energy, pix_det_index (energy (n_ifo, n_freq, n_time))."""
    lib = load()
    fn = lib.select_oracle
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int, C.c_int, C.c_int64, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                   C.c_double, vp, vp, vp, vp, vp, vp, vp, C.c_int64]
    E = _c(energy, np.float64)
    n_ifo, n_freq, n_time = E.shape
    sh = _c(shift_bins, np.int64)
    vt = None if veto is None else _c(veto, np.int16)
    cap = n_freq * n_time
    live = np.zeros(n_time, np.uint8)
    mask = np.zeros((n_freq, n_time), np.uint8)
    fr = np.zeros(cap, np.int64)
    tm = np.zeros(cap, np.int64)
    val = np.zeros(cap)
    de = np.zeros((cap, n_ifo))
    di = np.zeros((cap, n_ifo), np.int64)
    n = fn(n_ifo, n_freq, n_time, _p(E), _p(sh), None if vt is None else _p(vt), int(edge_bins),
           int(ib), int(ie), float(eo), _p(live), _p(mask), _p(fr), _p(tm), _p(val), _p(de),
           _p(di), cap)
    return {"live_mask": live.astype(bool), "mask": mask.astype(bool), "frequency": fr[:n],
            "time": tm[:n], "energy": val[:n], "pix_det_energy": de[:n],
            "pix_det_index": di[:n]}


# ---- time-delay max-energy maps (maxenergy_oracle.c) ----

def max_energy_params(M1, T, pattern, edge, wavelet_rate, f_low, f_high, df):
    """_compute_packet_energy_params: (jb, je, mL, mH, mean, p[9])."""
    lib = load()
    fn = lib.maxenergy_oracle_params
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                   C.c_double, vp, vp, vp, vp, vp, vp]
    jb, je = C.c_int64(), C.c_int64()
    mL, mH = C.c_int(), C.c_int()
    mean = C.c_double()
    p = np.zeros(9, np.int64)
    fn(int(M1), int(T), int(pattern), float(edge), int(wavelet_rate), float(f_low),
       float(f_high), float(df), C.byref(jb), C.byref(je), C.byref(mL), C.byref(mH),
       C.byref(mean), _p(p))
    return jb.value, je.value, mL.value, mH.value, mean.value, p


def max_energy(ts, M, filt, max_delay, downsample, pattern, edge, wavelet_rate, f_low, f_high,
               df):
    """(M+1, T) maximum packet energy over the time-delayed copies of ts."""
    lib = load()
    fn = lib.maxenergy_oracle
    fn.restype = C.c_int64
    fn.argtypes = [vp, C.c_int64, C.c_int, vp, C.c_int, C.c_int64, C.c_int, C.c_int,
                   C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, vp]
    x = _c(ts, np.float64)
    filt = _c(filt, np.float64)
    L = x.size
    T = (L + (M - L % M if L % M else 0)) // M
    out = np.zeros((T, M + 1))
    fn(_p(x), L, int(M), _p(filt), filt.size, int(max_delay), int(downsample), int(pattern),
       float(edge), int(wavelet_rate), float(f_low), float(f_high), float(df), _p(out))
    return out.T.copy()


# ---- connected-component clustering (cluster_oracle.c) ----

def cluster_labels(freq, time, n_freq, n_time, kf=1, kt=1):
    """_label_components_grid: 1-based labels in pixel order of the roots."""
    lib = load()
    fn = lib.cluster_oracle_labels
    fn.restype = None
    fn.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    f, t = _c(freq, np.int64), _c(time, np.int64)
    out = np.zeros(f.size, np.int64)
    fn(_p(f), _p(t), f.size, int(n_freq), int(n_time), int(kf), int(kt), _p(out))
    return out


def np_sum(x):
    """numpy's add.reduce of a contiguous float64 array, restated."""
    lib = load()
    fn = lib.cluster_oracle_np_sum
    fn.restype = C.c_double
    fn.argtypes = [vp, C.c_int64]
    a = _c(x, np.float64)
    return fn(_p(a), a.size)


def subnet_subrho(asnr, noise_rms, n_sub):
    """_subnet_subrho_numba on (n_pix, n_ifo) arrays: (subnet, subrho)."""
    lib = load()
    fn = lib.cluster_oracle_subnet
    fn.restype = None
    fn.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_double, vp, vp]
    a, v = _c(asnr, np.float64), _c(noise_rms, np.float64)
    sn, sr = C.c_double(), C.c_double()
    fn(_p(a), _p(v), a.shape[0], a.shape[1], float(n_sub), C.byref(sn), C.byref(sr))
    return sn.value, sr.value


def cluster_run(freq, time, energy, det_energy, n_freq, n_time, kf, kt, rate, n_sub):
    """cluster_pixels for one payload: dict with labels, members, cluster_off,
    energy, c_time, c_freq, subnet, subrho (per cluster)."""
    lib = load()
    fn = lib.cluster_oracle_run
    fn.restype = C.c_int64
    fn.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                   C.c_double, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp]
    f, t = _c(freq, np.int64), _c(time, np.int64)
    e = _c(energy, np.float64)
    de = _c(det_energy, np.float64)
    n = f.size
    n_ifo = de.shape[1] if de.ndim == 2 else 0
    lab, mem = np.zeros(n, np.int64), np.zeros(n, np.int64)
    off = np.zeros(n + 1, np.int64)
    cap = max(n, 1)
    ce, ct, cf, cs, cr = (np.zeros(cap) for _ in range(5))
    nc = fn(_p(f), _p(t), _p(e), _p(de), n, int(n_ifo), int(n_freq), int(n_time), int(kf),
            int(kt), float(rate), float(n_sub), _p(lab), _p(mem), _p(off), _p(ce), _p(ct),
            _p(cf), _p(cs), _p(cr))
    return {"labels": lab, "members": mem, "cluster_off": off[:nc + 1], "energy": ce[:nc],
            "c_time": ct[:nc], "c_freq": cf[:nc], "subnet": cs[:nc], "subrho": cr[:nc]}


# ---- threshold statistics (threshold_oracle.c) ----

def threshold_stats(maps, time_major, begin, count, hi):
    """compute_threshold's array work on (n_ifo, n_freq, n_time) maps: dict with
    values, counts (> 1e-4, > 1e-3), sorted, sum and log_sum of the values
    above 1e-3 (numpy order), n_pos."""
    lib = load()
    E = _c(maps, np.float64)
    n_ifo, n_freq, n_time = E.shape
    fv = lib.threshold_oracle_values
    fv.restype = None
    fv.argtypes = [vp, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int64, C.c_int64, C.c_double,
                   vp, vp]
    vals = np.zeros(max(count, 1))
    counts = np.zeros(2, np.int64)
    fv(_p(E), n_ifo, n_freq, n_time, int(time_major), int(begin), int(count), float(hi),
       _p(vals), _p(counts))
    vals = vals[:count]
    fs = lib.threshold_oracle_sorted
    fs.restype = None
    fs.argtypes = [vp, C.c_int64, vp]
    srt = np.zeros(max(count, 1))
    fs(_p(np.ascontiguousarray(vals)), int(count), _p(srt))
    fp = lib.threshold_oracle_positive_sums
    fp.restype = C.c_int64
    fp.argtypes = [vp, C.c_int64, vp, vp]
    s, ls = C.c_double(), C.c_double()
    m = fp(_p(np.ascontiguousarray(vals)), int(count), C.byref(s), C.byref(ls))
    return {"values": vals, "counts": counts, "sorted": srt[:count], "sum": s.value,
            "log_sum": ls.value, "n_pos": m}
