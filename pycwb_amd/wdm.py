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

# Number of Fourier terms per beta order used by the reference's coefficient table.
COS_SIZE = {2: 401, 3: 501, 4: 151, 5: 126, 6: 81, 7: 61}


@lru_cache(maxsize=None)
def edge_fourier(beta: int) -> np.ndarray:
    """a_0 = int_0^1 g, a_j = sqrt(2) int_0^1 g cos(j pi x), g = cos(pi/2 I_x(beta, beta))."""
    n = COS_SIZE[beta]
    x, w = np.polynomial.legendre.leggauss(1024)
    x, w = 0.5 * (x + 1.0), 0.5 * w
    m = 2 * beta - 1
    nu = sum(math.comb(m, k) * x ** k * (1 - x) ** (m - k) for k in range(beta, m + 1))
    g = np.cos(0.5 * np.pi * nu)
    j = np.arange(n)[:, None]
    a = (np.cos(j * np.pi * x[None, :]) * (w * g)[None, :]).sum(axis=1)
    a[1:] *= math.sqrt(2.0)
    return a


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
        raise NotImplementedError("time-delay filters are not part of this package")

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
