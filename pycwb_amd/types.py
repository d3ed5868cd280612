"""Minimal cluster containers with the field names the likelihood stage uses.

The GPU facade duck-types its inputs: a PycWB ``Cluster`` (with its
``PixelArrays``) works unchanged.  These light classes carry the same fields for
standalone use and tests:

* ``PixelArrays`` -- pycwb/types/pixel_arrays.py:48-104 (SoA pixels; TD amplitudes
  in CSR rows ``pixel * n_ifo + ifo`` of length 2*(2K+1): phase 00 then 90;
  ``from_arrays``, :159-216);
* ``ClusterMeta`` / ``Cluster`` / ``FragmentCluster`` -- pycwb/types/network_cluster.py;
* ``Config`` -- the parameters the likelihood reads, with the schema defaults of
  pycwb/constants/user_parameters_schema.py.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class PixelArrays:
    time: np.ndarray
    frequency: np.ndarray
    layers: np.ndarray
    rate: np.ndarray
    core: np.ndarray
    likelihood: np.ndarray
    null: np.ndarray
    noise_rms: np.ndarray       # (n_ifo, n_pix) float32
    wave: np.ndarray
    w_90: np.ndarray
    asnr: np.ndarray
    a_90: np.ndarray
    td_amp_flat: np.ndarray     # float32
    td_amp_offsets: np.ndarray  # int32 (n_pix * n_ifo + 1,)
    _n_ifo: int = 0
    pixel_index: np.ndarray = None  # (n_ifo, n_pix) int32

    @classmethod
    def from_arrays(cls, *, time, frequency, layers, rate, noise_rms, pixel_index, n_ifo,
                    core=None, likelihood=None, null=None, wave=None, w_90=None, asnr=None,
                    a_90=None) -> "PixelArrays":
        """pixel_arrays.py:159-216 (without TD amplitudes)."""
        n = len(time)
        zp = lambda: np.zeros(n, dtype=np.float32)
        zi = lambda: np.zeros((n_ifo, n), dtype=np.float32)
        f32 = lambda a, z: np.asarray(a, dtype=np.float32) if a is not None else z()
        return cls(time=np.asarray(time, dtype=np.int32),
                   frequency=np.asarray(frequency, dtype=np.int32),
                   layers=np.asarray(layers, dtype=np.int32),
                   rate=np.asarray(rate, dtype=np.float32),
                   core=np.asarray(core, dtype=bool) if core is not None else np.zeros(n, bool),
                   likelihood=f32(likelihood, zp), null=f32(null, zp),
                   noise_rms=np.asarray(noise_rms, dtype=np.float32),
                   wave=f32(wave, zi), w_90=f32(w_90, zi), asnr=f32(asnr, zi),
                   a_90=f32(a_90, zi), td_amp_flat=np.zeros(0, dtype=np.float32),
                   td_amp_offsets=np.zeros(n * n_ifo + 1, dtype=np.int32), _n_ifo=n_ifo,
                   pixel_index=np.asarray(pixel_index, dtype=np.int32))

    @classmethod
    def from_td(cls, td00: np.ndarray, td90: np.ndarray, noise_rms: np.ndarray,
                time=None, frequency=None, layers=None) -> "PixelArrays":
        """Build from reference-layout planes td00/td90 (2K+1, n_ifo, n_pix)."""
        nd, n, P = td00.shape
        dense = np.concatenate([td00.transpose(2, 1, 0), td90.transpose(2, 1, 0)], axis=2)
        z = lambda dt: np.zeros(P, dtype=dt)
        zz = lambda: np.zeros((n, P), dtype=np.float32)
        return cls(time=np.arange(P, dtype=np.int32) if time is None else np.asarray(time),
                   frequency=z(np.int32) if frequency is None else np.asarray(frequency),
                   layers=z(np.int32) if layers is None else np.asarray(layers),
                   rate=z(np.float32), core=z(bool), likelihood=z(np.float32),
                   null=z(np.float32), noise_rms=np.asarray(noise_rms, np.float32),
                   wave=zz(), w_90=zz(), asnr=zz(), a_90=zz(),
                   td_amp_flat=np.ascontiguousarray(dense, np.float32).ravel(),
                   td_amp_offsets=(np.arange(P * n + 1) * 2 * nd).astype(np.int32), _n_ifo=n)

    def __len__(self) -> int:
        return len(self.time)

    @property
    def n_ifo(self) -> int:
        return self._n_ifo

    def has_td_amp(self) -> bool:
        return len(self.td_amp_flat) > 0

    def td_amp_dense(self) -> np.ndarray:
        """(n_pix, n_ifo, tsize) float32 (pixel_arrays.py:227-245)."""
        P, n = len(self.time), self._n_ifo
        rows = P * n
        if rows == 0 or len(self.td_amp_flat) == 0:
            return np.zeros((P, n, 0), dtype=np.float32)
        sizes = self.td_amp_offsets[1:rows + 1] - self.td_amp_offsets[:rows]
        tsize = int(sizes[0])
        if not np.all(sizes == tsize):
            raise ValueError("td_amp vectors have non-uniform lengths; cannot form a dense array")
        return self.td_amp_flat[:rows * tsize].reshape(P, n, tsize)

    def populate_noise_rms(self, nRMS: list) -> None:
        """Per-pixel noise floor from the detectors' TF noise maps
        (pixel_arrays.py:367-399): map row round(f * nf / layers), column
        time // layers (both clipped to the map), |value| as float32, kept only
        where positive."""
        freq = self.frequency.astype(np.int64)
        layers = self.layers.astype(np.int64)
        tbin = np.where(layers > 0, self.time.astype(np.int64) // np.maximum(layers, 1), 0)
        for i in range(self._n_ifo):
            arr = np.asarray(nRMS[i].data, dtype=np.float64)
            nf, nt = arr.shape
            fb = np.where(layers > 0,
                          np.round(freq * nf / np.maximum(layers, 1)).astype(np.int64), 0)
            vals = np.abs(arr[np.clip(fb, 0, nf - 1), np.clip(tbin, 0, nt - 1)]).astype(
                np.float32)
            self.noise_rms[i] = np.where(vals > 0.0, vals, self.noise_rms[i])

    def set_waveform_data(self, wave, w_90, asnr, a_90, core_mask, energy_plus, energy_cross):
        """pixel_arrays.py:264-289."""
        self.wave[:] = np.asarray(wave, dtype=np.float32)
        self.w_90[:] = np.asarray(w_90, dtype=np.float32)
        self.asnr[:] = np.asarray(asnr, dtype=np.float32)
        self.a_90[:] = np.asarray(a_90, dtype=np.float32)
        self.core[:] = np.asarray(core_mask) > 0
        self.likelihood[:] = np.where(
            np.asarray(core_mask) > 0,
            -(np.asarray(energy_plus, np.float32) + np.asarray(energy_cross, np.float32)) / 2.0,
            0.0)
        self.null[:] = 0.0


@dataclass
class ClusterMeta:
    energy: float = 0.0
    energy_sky: float = 0.0
    like_net: float = 0.0
    net_ecor: float = 0.0
    norm_cor: float = 0.0
    net_null: float = 0.0
    net_ed: float = 0.0
    g_noise: float = 0.0
    like_sky: float = 0.0
    sky_cc: float = 0.0
    net_cc: float = 0.0
    sky_chi2: float = 0.0
    sub_net: float = 0.0
    sub_net2: float = 0.0
    sky_stat: float = 0.0
    net_rho: float = 0.0
    net_rho2: float = 0.0
    theta: float = 0.0
    phi: float = 0.0
    iota: float = 0.0
    psi: float = 0.0
    ellipticity: float = 0.0
    c_time: float = 0.0
    c_freq: float = 0.0
    g_net: float = 0.0
    a_net: float = 0.0
    i_net: float = 0.0
    norm: float = 0.0
    ndof: float = 0.0
    l_max: int = 0


@dataclass
class Cluster:
    pixel_arrays: PixelArrays = None
    cluster_meta: ClusterMeta = field(default_factory=ClusterMeta)
    cluster_status: int = 0
    cluster_id: int = 0
    cluster_time: float = 0.0
    cluster_freq: float = 0.0
    sky_time_delay: list = field(default_factory=list)


@dataclass
class FragmentCluster:
    clusters: list = field(default_factory=list)
    rate: float = 0.0
    start: float = 0.0
    stop: float = 0.0
    bpp: float = 0.0
    shift: float = 0.0
    f_low: float = 0.0
    f_high: float = 0.0
    n_pix: int = 0
    run: int = 0
    pair: bool = False
    subnet_threshold: float = 0.0

    @staticmethod
    def _status_ok(event_status):
        if event_status is not None and (not isinstance(event_status, int) or event_status > 1
                                         or event_status < -2):
            raise ValueError('event_status must be -2, -1, 0, 1 or None')

    def event_count(self, event_status=None) -> int:
        """network_cluster.py:396-410: clusters with status < 1, or == event_status."""
        self._status_ok(event_status)
        if event_status is None:
            return sum(1 for c in self.clusters if c.cluster_status < 1)
        return sum(1 for c in self.clusters if c.cluster_status == event_status)

    def pixel_count(self, event_status=None) -> int:
        """network_cluster.py:412-426."""
        self._status_ok(event_status)
        if event_status is None:
            return sum(len(c.pixel_arrays) for c in self.clusters if c.cluster_status < 1)
        return sum(len(c.pixel_arrays) for c in self.clusters
                   if c.cluster_status == event_status)

    def remove_rejected(self) -> None:
        """network_cluster.py:447-451."""
        self.clusters = [c for c in self.clusters if c.cluster_status < 1]

    def select(self, name: str, threshold: float) -> np.ndarray:
        """network_cluster.py:453-487: rejects (status 1) the not-yet-rejected
        clusters whose subrho (net_rho) or subnet (sub_net) is below threshold
        (a NaN value is not below it); returns the values."""
        key = str(name).strip().lower()
        if key == "subrho":
            values = np.array([c.cluster_meta.net_rho for c in self.clusters], dtype=float)
        elif key == "subnet":
            values = np.array([c.cluster_meta.sub_net for c in self.clusters], dtype=float)
        else:
            raise ValueError(f"Unsupported selection '{name}'. Supported: subrho, subnet")
        for c, v in zip(self.clusters, values):
            if c.cluster_status <= 0 and v < threshold:
                c.cluster_status = 1
        return values


@dataclass
class Config:
    """Likelihood parameters with the reference schema defaults."""
    nIFO: int = 3
    ifo: tuple = ("H1", "L1", "V1")
    refIFO: str = "H1"
    Acore: float = math.sqrt(2.0)
    gamma: float = 0.5
    delta: float = 0.5
    netCC: float = 0.5
    netRHO: float = 4.0
    xgb_rho_mode: bool = False
    healpix: int = 7
    TDSize: int = 12
    upTDF: int = 4
    rateANA: int = 1024
    TDRate: float | None = None
    max_delay: float = 0.0
    precision: int = 0
    nRES: int = 1
    sky_mask: dict | None = None
    pattern: int = 10

    def __post_init__(self):
        if self.TDRate is None:
            self.TDRate = self.rateANA * self.upTDF
