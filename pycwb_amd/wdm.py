"""WDM (Wilson-Daubechies-Meyer) forward transform with the wdm_wavelet object surface.

``WDM(M, K, beta_order, precision)`` builds the filter of cwb-core/WDM.cc
(WDM<T>::getFilter, :498-548, and the length rule of the constructor,
:108-173) from its definition: the band edge of the basis in the Fourier
domain is ``cos(pi/2 * I_x(beta, beta))`` (regularised incomplete beta), whose
cosine-series coefficients (the table FourierCoefficients.icc holds) are
computed here by Gauss-Legendre quadrature.  The filter is built once on the
host in float64; ``t2w`` runs the HIP kernel ``pycwb_wdm_t2w`` (float64) on
the GPU.

Output convention (coherence_native/tf_batch_generation.py:186-189): a complex
(M+1, N) array with phase 00 in the real and phase 90 in the imaginary part,
``dt = M / fs`` and ``df = fs / (2M)``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
import torch

from . import _lib
from .device import lib, ptr, require_gpu, stream_ptr

# Number of Fourier terms per beta order used by the reference's coefficient
# tables (Cos, Cos2, SinCos of cwb-core/FourierCoefficients.icc).
COS_SIZE = {2: 401, 3: 501, 4: 151, 5: 126, 6: 81, 7: 61}
COS2_SIZE = {2: 101, 3: 61, 4: 46, 5: 36, 6: 36, 7: 31}
SINCOS_SIZE = {2: 601, 3: 801, 4: 151, 5: 151, 6: 81, 7: 81}


def _cosine_series(beta: int, n: int, which: str) -> np.ndarray:
    """a_0 = int_0^1 f, a_j = sqrt(2) int_0^1 f cos(j pi x) for f = g (which="cos"),
    g^2 ("cos2") or sin(pi nu) = 2 g sqrt(1 - g^2) ("sincos"), with
    g = cos(pi/2 nu), nu = I_x(beta, beta)."""
    x, w = np.polynomial.legendre.leggauss(2048 if which != "cos" else 1024)
    x, w = 0.5 * (x + 1.0), 0.5 * w
    m = 2 * beta - 1
    nu = sum(math.comb(m, k) * x ** k * (1 - x) ** (m - k) for k in range(beta, m + 1))
    if which == "cos":
        f = np.cos(0.5 * np.pi * nu)
    elif which == "cos2":
        f = np.cos(0.5 * np.pi * nu) ** 2
    else:
        f = np.sin(np.pi * nu)
    j = np.arange(n)[:, None]
    a = (np.cos(j * np.pi * x[None, :]) * (w * f)[None, :]).sum(axis=1)
    a[1:] *= math.sqrt(2.0)
    # nu(1 - x) = 1 - nu(x): g^2 - 1/2 is odd and sin(pi nu) even about x = 1/2,
    # so their even (j > 0) and odd coefficients vanish exactly
    if which == "cos2":
        a[2::2] = 0.0
    elif which == "sincos":
        a[1::2] = 0.0
    return a


@lru_cache(maxsize=None)
def edge_fourier(beta: int) -> np.ndarray:
    """Cosine-series coefficients of the band edge g = cos(pi/2 I_x(beta, beta))."""
    return _cosine_series(beta, COS_SIZE[beta], "cos")


@lru_cache(maxsize=None)
def edge_fourier_cos2(beta: int) -> np.ndarray:
    """Cosine-series coefficients of g^2 (the reference's Cos2 table)."""
    return _cosine_series(beta, COS2_SIZE[beta], "cos2")


@lru_cache(maxsize=None)
def edge_fourier_sincos(beta: int) -> np.ndarray:
    """Cosine-series coefficients of 2 g sqrt(1 - g^2) (the reference's SinCos table)."""
    return _cosine_series(beta, SINCOS_SIZE[beta], "sincos")


def _td_integral1(M: int, K: int, beta: int, n: int, L: int) -> np.ndarray:
    """WDM<T>::getTDFilter1 (cwb-core/WDM.cc:551-594): n*L samples, step tau/L."""
    four = edge_fourier_cos2(beta).copy()
    nf = four.size
    B = math.pi / K
    A = (K - M) * math.pi / 2.0 / K / M
    K2 = float(K) * K
    gNorm = 2.0 * M / math.pi
    out = np.empty(n * L)
    out[0] = (A + four[0] * B) * gNorm
    four[0] /= math.sqrt(2.0)
    i = np.arange(1, n * L)
    di = i / float(L)
    jj = np.arange(nf)
    with np.errstate(divide="ignore", invalid="ignore"):
        terms = di[:, None] / (di[:, None] ** 2 - (jj[None, :] ** 2) * K2) * four[None, :]
    terms = np.where(jj[None, :] * K * L == i[:, None], 0.0, terms)
    odd = (jj & 1).astype(bool)
    s_even = np.cumsum(np.where(odd[None, :], 0.0, terms), axis=1)[:, -1]
    s_odd = np.cumsum(np.where(odd[None, :], terms, 0.0), axis=1)[:, -1]
    intAB = np.zeros(i.size)
    sel = (i % (K * L) == 0) & (i // (K * L) < nf)
    intAB[sel] = four[i[sel] // (K * L)] * B / 2 * np.cos(di[sel] * A)
    intAB += 2 * (s_even * np.sin(di * B / 2) * np.cos(di * np.pi / (2 * M))
                  - s_odd * np.sin(di * np.pi / (2 * M)) * np.cos(di * B / 2))
    out[1:] = gNorm * (np.sin(di * A) / di + math.sqrt(2.0) * intAB)
    return out


def _td_integral2(M: int, K: int, beta: int, n: int, L: int) -> np.ndarray:
    """WDM<T>::getTDFilter2 (cwb-core/WDM.cc:596-636): n*L samples, step tau/L."""
    four = edge_fourier_sincos(beta).copy()
    nf = four.size
    B = math.pi / K
    K2 = float(K) * K
    gNorm = M / math.pi
    out = np.empty(n * L)
    out[0] = four[0] * B * gNorm
    four[0] /= math.sqrt(2.0)
    i = np.arange(1, n * L)
    di = i / float(L)
    jj = np.arange(0, nf, 2)  # the reference runs j = 0, 2, .. <= nf (nf is odd)
    with np.errstate(divide="ignore", invalid="ignore"):
        terms = di[:, None] / (di[:, None] ** 2 - (jj[None, :] ** 2) * K2) * four[jj][None, :]
    terms = np.where(jj[None, :] * K * L == i[:, None], 0.0, terms)
    ssum = np.cumsum(terms, axis=1)[:, -1] * (2 * np.sin(di * B / 2))
    sel = (i % (2 * K * L) == 0) & (i // (K * L) <= nf)
    idx = i[sel] // (K * L)
    sign = np.where((i[sel] // (2 * K * L)) & 1, -1.0, 1.0)
    ssum[sel] += sign * four[np.minimum(idx, nf - 1)] * B / 2
    out[1:] = gNorm * math.sqrt(2.0) * ssum
    return out


@dataclass
class TDFilterBank:
    """Time-delay filter tables (the wdm_wavelet TDFilterBank fields PycWB reads).

    T0, Tx: float64 (2J+1, 2 n_coeffs+1), row i + J for delay i in [-J, J]
    (steps of tau / L), column j + n_coeffs for time-bin offset j.
    """
    T0: np.ndarray
    Tx: np.ndarray
    n_coeffs: int
    M: int
    L: int

    @property
    def max_delay(self) -> int:
        return self.M * self.L


def td_filters(M: int, K: int, beta: int, n_coeffs: int, L: int) -> TDFilterBank:
    """WDM<T>::setTDFilter (cwb-core/WDM.cc:638-700) as dense tables."""
    n = M * (n_coeffs + 1) + 1
    f1 = _td_integral1(M, K, beta, n, L)
    f2 = _td_integral2(M, K, beta, n, L)
    J = M * L
    T0 = np.zeros((2 * J + 1, 2 * n_coeffs + 1))
    Tx = np.zeros((2 * J + 1, 2 * n_coeffs + 1))
    c = n_coeffs
    for i in range(-J, J + 1):
        r = i + J
        T0[r, c] = f1[abs(i)]
        Tx[r, c] = f2[abs(i)] / 2.0
        for j in range(1, n_coeffs + 1):
            T0[r, c + j] = f1[i + J * j]
            T0[r, c - j] = f1[J * j - i]
            Tx[r, c + j] = f2[i + J * j] / 2.0
            Tx[r, c - j] = f2[J * j - i] / 2.0
        for j in range(1, n_coeffs + 1):  # bring the i^l factor to C_l form
            if j % 4 == 1:
                Tx[r, c - j] *= -1
            elif j % 4 == 2:
                Tx[r, c + j] *= -1
                Tx[r, c - j] *= -1
            elif j % 4 == 3:
                Tx[r, c + j] *= -1
    return TDFilterBank(T0=T0, Tx=Tx, n_coeffs=int(n_coeffs), M=int(M), L=int(L))


def wdm_filter(M: int, K: int, beta: int, precision: int, chunk: int = 4096) -> np.ndarray:
    """Filter taps h[0..m_H) (WDM.cc:108-173, 498-548), float64."""
    beta = min(max(int(beta), 2), 7)
    four = edge_fourier(beta).copy()
    nf = four.size
    B = math.pi / K
    A = (K - M) * math.pi / 2.0 / K / M
    K2 = float(K) * K
    gNorm = math.sqrt(2.0 * M) / math.pi
    f0 = (A + four[0] * B) * gNorm
    four[0] /= math.sqrt(2.0)
    taps = [np.array([f0])]
    residual = 1.0 - f0 * f0
    prec = 10.0 ** (-precision)
    M2 = 2 * M
    N = 1
    jj = np.arange(nf)
    start = 1
    while True:
        i = np.arange(start, start + chunk, dtype=np.float64)
        ii = i.astype(np.int64)
        with np.errstate(divide="ignore", invalid="ignore"):
            terms = i[:, None] / (i[:, None] ** 2 - (jj[None, :] ** 2) * K2) * four[None, :]
        skip = (ii[:, None] % K == 0) & (jj[None, :] == ii[:, None] // K)
        terms = np.where(skip, 0.0, terms)
        odd = (jj & 1).astype(bool)
        # sequential sums over j, as the reference's loop
        s_even = np.cumsum(np.where(odd[None, :], 0.0, terms), axis=1)[:, -1]
        s_odd = np.cumsum(np.where(odd[None, :], terms, 0.0), axis=1)[:, -1]
        intAB = np.zeros_like(i)
        sel = (ii % K == 0) & (ii // K < nf)
        intAB[sel] = four[ii[sel] // K] * B / 2 * np.cos(i[sel] * A)
        intAB = intAB + 2 * (s_even * np.sin(i * B / 2) * np.cos(i * np.pi / (2 * M))
                             - s_odd * np.sin(i * np.pi / (2 * M)) * np.cos(i * B / 2))
        h = gNorm * (np.sin(i * A) / i + math.sqrt(2.0) * intAB)
        taps.append(h)
        for v in h:  # constructor loop: residual -= 2 f[N]^2; N++; while (...)
            residual -= 2 * v * v
            N += 1
            if not (residual > prec or (N - 1) % M2 or N // M2 < 3):
                return np.concatenate(taps)[:N]
        start += chunk


@dataclass
class TimeFrequencyMap:
    """Minimal container with the fields of pycwb's TimeFrequencyMap."""
    data: np.ndarray
    dt: float
    df: float
    start_time: float
    len_timeseries: int

    @property
    def end_time(self):
        return self.start_time + self.data.shape[1] * self.dt

    @property
    def start_freq(self):
        return 0.0

    @property
    def end_freq(self):
        return (self.data.shape[0] - 1) * self.df

    def prepare_td_inputs(self, td_filters, device=None):
        """Padded quadrature planes and TD filters on the device
        (time_frequency_map.py:482-521); see pycwb_amd.tdamp."""
        from .tdamp import TDBatchInputs

        return TDBatchInputs.from_map(self.data, td_filters, device)


class WDM:
    """WDM transform with the surface of wdm_wavelet.WDM used by PycWB."""

    def __init__(self, M: int, K: int | None = None, beta_order: int = 6, precision: int = 10):
        self.M = int(M)
        self.K = int(K if K is not None else M)
        self.beta_order = int(beta_order)
        self.precision = int(precision)
        if self.M < 1 or self.M & (self.M - 1) or self.M > 1024:
            raise ValueError("M must be a power of two in [1, 1024]")
        self.filter = wdm_filter(self.M, self.K, self.beta_order, self.precision)
        self.m_H = int(self.filter.size)
        self._dev = {}

    def set_td_filter(self, TDSize, upTDF):
        """Time-delay filters with TDSize coefficients per side and upsampling
        upTDF (WDM.cc:638-700); stored as ``self.td_filters``."""
        self.td_filters = td_filters(self.M, self.K, self.beta_order, int(TDSize), int(upTDF))
        return self.td_filters

    def _device(self, device):
        if device not in self._dev:
            M2 = 2 * self.M
            j = np.arange(M2)
            tw = np.stack([np.cos(2 * np.pi * j / M2), -np.sin(2 * np.pi * j / M2)], axis=1)
            self._dev[device] = (torch.from_numpy(self.filter).to(device),
                                 torch.from_numpy(np.ascontiguousarray(tw)).to(device))
        return self._dev[device]

    def t2w_device(self, x: torch.Tensor, MM: int = -1, stream=None) -> torch.Tensor:
        """Batched transform on the GPU.

        x: (n_series, L) or (L,) float64 device tensor.  Returns complex128
        (n_series, M+1, N) (a transposed view of the bin-major kernel output).
        """
        dev = x.device
        X = x.reshape(1, -1) if x.dim() == 1 else x
        X = X.to(torch.float64).contiguous()
        S_, L = X.shape
        if L <= self.m_H:
            raise ValueError(f"series of {L} samples is not longer than the filter ({self.m_H})")
        stride = self.M if MM <= 0 else int(MM)
        N = (L + (stride - L % stride if L % stride else 0)) // stride
        filt, tw = self._device(dev)
        out = torch.empty((S_, N, self.M + 1, 2), dtype=torch.float64, device=dev)
        rc = lib().pycwb_wdm_t2w(ptr(X), L, L, S_, self.M, stride, ptr(filt), self.m_H, ptr(tw),
                                 1 if MM < 0 else 0, ptr(out), stream_ptr(stream))
        _lib.check(rc, "pycwb_wdm_t2w")
        return torch.view_as_complex(out).transpose(1, 2)

    def t2w_shifted_device(self, ts: torch.Tensor, shifts: torch.Tensor, stream=None):
        """Quadrature transforms (like t2w_device(x, MM=-1)) of shifted copies of
        the device series ts, copy s shifted by shifts[s] (int64 device tensor;
        k > 0: x[i] = ts[i + k] for i < L - k, k < 0: x[i] = ts[i + k] for
        i >= -k).  The copies are never written out.  None when the filter is
        not of the form m_H = 12 M + 1 that the shifted path supports."""
        if self.m_H - 1 != 12 * self.M or self.M < 2:
            return None
        dev = ts.device
        X = ts.reshape(-1).to(torch.float64).contiguous()
        L = int(X.numel())
        if L <= self.m_H:
            raise ValueError(f"series of {L} samples is not longer than the filter ({self.m_H})")
        sh = shifts.to(device=dev, dtype=torch.int64).contiguous()
        n = int(sh.numel())
        N = (L + (self.M - L % self.M if L % self.M else 0)) // self.M
        filt, tw = self._device(dev)
        out = torch.empty((n, N, self.M + 1, 2), dtype=torch.float64, device=dev)
        rc = lib().pycwb_wdm_t2w_shifted(ptr(X), L, ptr(sh), n, self.M, ptr(filt), self.m_H,
                                         ptr(tw), 1, ptr(out), stream_ptr(stream))
        _lib.check(rc, "pycwb_wdm_t2w_shifted")
        return torch.view_as_complex(out).transpose(1, 2)

    def w2t_device(self, tf: torch.Tensor, n_out: int, quad: bool = False,
                   stream=None) -> torch.Tensor:
        """Inverse transform on the GPU (WDM.cc:1536-1710): tf complex128
        (n_series, M+1, N) or (M+1, N) -- t2w_device's output -- to
        (n_series, n_out) float64 from the 00 (quad=False, w2t) or the 90
        (quad=True, w2tQ) coefficients."""
        T = tf if tf.dim() == 3 else tf.unsqueeze(0)
        S_, M1, N = T.shape
        if M1 != self.M + 1:
            raise ValueError(f"map has {M1} layers, the transform {self.M + 1}")
        if self.M < 2:
            raise ValueError("the inverse needs M >= 2")
        n_out = int(n_out)
        if not (n_out <= self.M * N < 2 * n_out):
            raise ValueError("inverse is not defined for the up-sampled map "
                             f"(M N = {self.M * N}, {n_out} samples)")
        dev = T.device
        bm = torch.view_as_real(T.to(torch.complex128).transpose(1, 2).contiguous())
        filt, tw = self._device(dev)
        work = torch.empty((S_, N, 2 * self.M), dtype=torch.float64, device=dev)
        out = torch.empty((S_, n_out), dtype=torch.float64, device=dev)
        rc = lib().pycwb_wdm_w2t(ptr(bm), N, S_, self.M, ptr(filt), self.m_H, ptr(tw),
                                 1 if quad else 0, ptr(work), ptr(out), n_out,
                                 stream_ptr(stream))
        _lib.check(rc, "pycwb_wdm_w2t")
        return out

    def _inverse(self, tf_map, quad):
        data = tf_map.data if hasattr(tf_map, "data") else tf_map
        n_out = int(getattr(tf_map, "len_timeseries", 0) or 0)
        dev = require_gpu()
        T = data if isinstance(data, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(np.asarray(data, np.complex128)))
        T = T.to(dev)
        if not n_out:
            n_out = self.M * T.shape[-1]
        return self.w2t_device(T, n_out, quad)[0].cpu().numpy()

    def w2t(self, tf_map) -> np.ndarray:
        """Time series from the 00 coefficients of a TimeFrequencyMap (or a
        complex (M+1, N) array); len_timeseries samples (WDM.cc:1536-1630)."""
        return self._inverse(tf_map, False)

    def w2tQ(self, tf_map) -> np.ndarray:
        """Time series from the 90 coefficients (WDM.cc:1635-1710)."""
        return self._inverse(tf_map, True)

    def t2w(self, x, sample_rate: float = 1.0, t0: float = 0.0, MM: int = -1):
        """One series -> TimeFrequencyMap with complex (M+1, N) data."""
        dev = require_gpu()
        xt = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.asarray(x, np.float64))
        tf = self.t2w_device(xt.to(dev).reshape(-1), MM)[0]
        data = tf.cpu().numpy()
        if MM >= 0:
            data = data.real.copy()
        return TimeFrequencyMap(data=data, dt=self.M / float(sample_rate),
                                df=float(sample_rate) / (2.0 * self.M), start_time=float(t0),
                                len_timeseries=int(xt.numel()))
