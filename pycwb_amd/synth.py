"""Seeded synthetic clusters for the likelihood workloads.

A cluster is ``P`` time-frequency pixels seen by ``n`` detectors: per pixel and
detector a table of 2K+1 delayed amplitudes in two phases (00, 90), float32,
unit-variance Gaussian noise, plus a coherent component on a subset of
"signal" pixels.  The coherent component is projected from one sky direction
``l0`` of the grid: ``v_i = r_i (F+_i h+ + Fx_i hx)`` placed at delay tap
``ml[i, l0] + K``, scaled so ``sqrt(sum_i v00_i^2 + v90_i^2) = snr``.

Noise weights follow the reference's normalisation
(pycwb/modules/likelihoodWP/pixel_data.py:60-62): ``r_ij = (1/s_ij) /
sqrt(sum_i 1/s_ij^2)`` in float64, stored float32, with noise rms ``s`` either 1
(``r = 1/sqrt(n)``) or drawn |N(1, 0.1)|.

Seeds: base seed 1000 + workload index; cluster ``c`` uses
``base * 100000 + c``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import sky

NETWORK_3 = ("H1", "L1", "V1")
NETWORK_4 = ("H1", "L1", "V1", "K1")
NETWORK_5 = ("H1", "L1", "V1", "K1", "I1")


@dataclass
class Network:
    ifos: tuple
    order: int
    K: int
    td_rate: float
    ml: np.ndarray      # (n, S) int32
    FP_t: np.ndarray    # (S, n) float32
    FX_t: np.ndarray    # (S, n) float32
    phi: np.ndarray     # (S,) float64
    lat: np.ndarray     # (S,) float64

    @property
    def n_ifo(self):
        return len(self.ifos)

    @property
    def n_sky(self):
        return self.ml.shape[1]


def make_network(ifos=NETWORK_3, order=7, K=64, td_rate=None) -> Network:
    """Sky tables for a network; the delay rate keeps max |ml| <= K."""
    ref = ifos[0]
    if td_rate is None:
        dmax = sky.max_network_delay(ifos, ref)
        td_rate = math.floor(K / dmax) if dmax > 0 else 1.0
    ml, FP, FX = sky.compute_sky_delay_and_patterns(list(ifos), ref, td_rate, K, order)
    phi, lat = sky.build_sky_directions(ml.shape[1], order)
    return Network(tuple(ifos), order, K, float(td_rate), ml, FP.T.astype(np.float32),
                   FX.T.astype(np.float32), phi, lat)


def noise_weights(sigma: np.ndarray) -> np.ndarray:
    """(n, P) noise rms -> (P, n) float32 weights (pixel_data.py:60-62)."""
    inv = 1.0 / sigma.astype(np.float64)
    norm = 1.0 / np.sqrt(np.sum(inv ** 2, axis=0))
    return (inv * norm[np.newaxis, :]).astype(np.float32).T.copy()


def make_cluster(net: Network, n_pix: int, seed: int, n_signal=None, snr: float = 3.0,
                 random_rms: bool = True, l0=None):
    """One cluster in the reference layout.

    Returns dict with td00, td90 (2K+1, n, P) float32, rms (P, n) float32,
    l0 (injection direction) and the signal pixel indices.
    """
    rng = np.random.default_rng(seed)
    n, K = net.n_ifo, net.K
    nd = 2 * K + 1
    td00 = rng.standard_normal((nd, n, n_pix), dtype=np.float32)
    td90 = rng.standard_normal((nd, n, n_pix), dtype=np.float32)
    sigma = np.abs(rng.normal(1.0, 0.1, size=(n, n_pix))) if random_rms else np.ones((n, n_pix))
    rms = noise_weights(sigma)
    if n_signal is None:
        n_signal = max(1, round(n_pix * 200 / 1024))
    n_signal = min(n_signal, n_pix)
    if l0 is None:
        l0 = int(rng.integers(net.n_sky))
    sig = np.sort(rng.choice(n_pix, size=n_signal, replace=False))
    fp = net.FP_t[l0].astype(np.float64)
    fx = net.FX_t[l0].astype(np.float64)
    taps = net.ml[:, l0] + K
    for j in sig:
        h = rng.standard_normal(4)
        r = rms[j].astype(np.float64)
        v00 = r * (fp * h[0] + fx * h[1])
        v90 = r * (fp * h[2] + fx * h[3])
        e = math.sqrt(float(np.sum(v00 ** 2 + v90 ** 2)))
        if e > 0:
            scale = snr / e
            for i in range(n):
                td00[taps[i], i, j] += np.float32(v00[i] * scale)
                td90[taps[i], i, j] += np.float32(v90[i] * scale)
    return {"td00": td00, "td90": td90, "rms": rms, "l0": l0, "signal": sig}


def cluster_sizes(n_clusters: int, seed: int, lo: int = 16, hi: int = 4096) -> np.ndarray:
    """Pixel counts log-uniform in [lo, hi]: P = round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(lo), math.log(hi), size=n_clusters)
    return np.clip(np.rint(np.exp(u)).astype(np.int64), lo, hi)


def packed_td(td00: np.ndarray, td90: np.ndarray) -> np.ndarray:
    """(2K+1, n, P) planes -> (P, n, 2K+1, 2) float32 pairs (the GPU layout)."""
    return np.ascontiguousarray(np.stack([td00, td90], axis=-1).transpose(2, 1, 0, 3))
