"""Connected-component clustering of the selected pixels on the GPU with the
signature of PycWB's coherence_native.clustering (clustering.py:18-157).

  cluster_pixels_device  every lag of a selection in one stream-ordered call
                         (pycwb_cluster_pixels): labels, the clusters' pixel
                         lists, energy, c_time, c_freq, subnet, subrho
  cluster_pixels         one payload of select_network_pixels ->
                         FragmentCluster, as the reference builds it
  cluster_pixels_lags    one FragmentCluster per lag of a device selection

n_sub = 2 iGamma^-1(n_ifo - 1, 0.314) is the reference's host constant
(veto_threshold.py:64-73, scipy's gammainccinv).
"""

from __future__ import annotations

import ctypes as C
import threading

import numpy as np
import torch

from .. import _lib
from ..device import lib, ptr, require_gpu, stream_ptr
from ..types import Cluster, ClusterMeta, FragmentCluster, PixelArrays
from .veto_threshold import _igamma_inv_upper


_maps: dict = {}
_maps_lock = threading.Lock()


def _index_map(dev: torch.device, cells: int, stream) -> torch.Tensor:
    """The device's cached int32 index map (grow-only), all -1 between calls:
    pycwb_cluster_map_init fills a new one, every clustering call restores the
    cells it touched.  Calls that share it must not overlap (one stream)."""
    with _maps_lock:
        m = _maps.get(dev.index)
        if m is None or m.numel() < cells:
            _maps.pop(dev.index, None)
            m = torch.empty(cells, dtype=torch.int32, device=dev)
            _lib.check(lib().pycwb_cluster_map_init(ptr(m), cells, stream_ptr(stream)),
                       "pycwb_cluster_map_init")
            _maps[dev.index] = m
        return m


def _n_sub(n_ifo: int) -> float:
    return 2.0 * _igamma_inv_upper(float(n_ifo - 1), 0.314) if n_ifo > 1 else 0.0


def cluster_pixels_device(sel: dict, kt: int = 1, kf: int = 1, rate: float = 0.0,
                          stream=None) -> dict:
    """Clusters every lag of ``sel`` (device tensors: ``lag_off_dev`` or host
    ``lag_off`` (n_lag + 1), ``frequency``, ``time``, ``energy``,
    ``pix_det_energy`` (n, n_ifo); ints ``n_freq``, ``n_time``).  Returns device
    ``label``, ``members``, ``cluster_off`` and per-cluster ``energy``,
    ``c_time``, ``c_freq``, ``subnet``, ``subrho`` (capacity n), and host
    ``lag_cluster_off`` (n_lag + 1) after one synchronisation."""
    lo_h = np.asarray(sel["lag_off"], dtype=np.int64)
    n_lag = lo_h.size - 1
    if n_lag < 1 or lo_h[0] != 0 or np.any(np.diff(lo_h) < 0):
        raise ValueError("lag_off must start at 0 and not decrease")
    n = int(lo_h[-1])
    if any(int(sel[k].shape[0]) < n for k in ("frequency", "time", "energy", "pix_det_energy")):
        raise ValueError("pixel arrays shorter than lag_off[-1]")
    dev = sel["frequency"].device if isinstance(sel["frequency"], torch.Tensor) else require_gpu()
    lo_d = sel.get("lag_off_dev")
    if lo_d is None:
        lo_d = torch.from_numpy(np.ascontiguousarray(lo_h)).to(dev)
    n_freq, n_time = int(sel["n_freq"]), int(sel["n_time"])
    kt, kf = int(max(1, kt)), int(max(1, kf))
    de = sel["pix_det_energy"]
    n_ifo = int(de.shape[1]) if de.dim() == 2 else 0
    rate = float(rate)
    d = _lib.ClusterDesc()
    d.n_lag, d.lag_off, d.n_pix = n_lag, ptr(lo_d), n
    d.freq, d.time, d.energy = ptr(sel["frequency"]), ptr(sel["time"]), ptr(sel["energy"])
    d.det_energy, d.noise_rms, d.n_ifo = ptr(de), None, n_ifo
    d.n_freq, d.n_time, d.kf, d.kt = n_freq, n_time, kf, kt
    d.dt = 1.0 / rate if rate > 0.0 else 1.0
    d.half_rate = rate / 2
    d.n_sub = _n_sub(n_ifo)
    cap = max(n, 1)
    i64 = lambda k: torch.empty(k, dtype=torch.int64, device=dev)
    f64 = lambda: torch.empty(cap, dtype=torch.float64, device=dev)
    out = {"label": i64(cap), "lag_cluster_off_dev": i64(n_lag + 1), "cluster_off": i64(cap + 1),
           "members": i64(cap), "energy": f64(), "c_time": f64(), "c_freq": f64(),
           "subnet": f64(), "subrho": f64()}
    d.label, d.lag_cluster_off, d.cluster_off = ptr(out["label"]), \
        ptr(out["lag_cluster_off_dev"]), ptr(out["cluster_off"])
    d.members, d.c_energy, d.c_time, d.c_freq = ptr(out["members"]), ptr(out["energy"]), \
        ptr(out["c_time"]), ptr(out["c_freq"])
    d.c_subnet, d.c_subrho = ptr(out["subnet"]), ptr(out["subrho"])
    nb = C.c_size_t()
    _lib.check(lib().pycwb_cluster_work_bytes(C.byref(d), C.byref(nb)), "pycwb_cluster_work_bytes")
    work = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=dev)
    d.work, d.work_bytes = ptr(work), int(nb.value)
    stream = stream or torch.cuda.current_stream()
    if n > 0:
        d.index_map = ptr(_index_map(dev, max(n_lag * n_freq * n_time, 1), stream))
    try:
        _lib.check(lib().pycwb_cluster_pixels(C.byref(d), stream_ptr(stream)),
                   "pycwb_cluster_pixels")
    except Exception:
        with _maps_lock:            # the map may not have been restored
            _maps.pop(dev.index, None)
        raise
    out["work"] = work
    out["index_map"] = d.index_map
    stream.synchronize()
    out["lag_cluster_off"] = out["lag_cluster_off_dev"].cpu().numpy()
    out["n_clusters"] = int(out["lag_cluster_off"][-1])
    out["lag_off"] = lo_h
    return out


def _fragment(payload: dict, f, t, energy, det_energy, det_index, members, offsets, stats,
              layers: int, rate: float) -> FragmentCluster:
    """Cluster objects of one lag (clustering.py:100-157); members are
    lag-local pixel indices grouped by cluster."""
    n_ifo = int(det_energy.shape[1]) if det_energy.ndim == 2 and det_energy.shape[1] > 0 else 0
    pix_asnr = np.sqrt(det_energy) if n_ifo > 0 else np.empty((len(f), 0))
    ce, ct, cf, cs, cr = stats
    clusters = []
    for c in range(len(offsets) - 1):
        idx = members[offsets[c]:offsets[c + 1]]
        fg, tg = f[idx], t[idx]
        ng = idx.size
        pa = PixelArrays.from_arrays(
            time=tg * layers + fg, frequency=fg, layers=np.full(ng, layers, dtype=np.int32),
            rate=np.full(ng, rate, dtype=np.float32), core=np.ones(ng, dtype=bool),
            likelihood=energy[idx], null=np.zeros(ng, dtype=np.float32),
            noise_rms=np.ones((n_ifo, ng), dtype=np.float32),
            pixel_index=det_index[idx].T if n_ifo > 0 else np.zeros((0, ng), dtype=np.int32),
            asnr=pix_asnr[idx].T if n_ifo > 0 else np.zeros((0, ng), dtype=np.float32),
            n_ifo=n_ifo)
        meta = ClusterMeta(energy=float(ce[c]), like_net=float(ce[c]), sub_net=float(cs[c]),
                           net_rho=float(cr[c]), c_time=float(ct[c]), c_freq=float(cf[c]))
        clusters.append(Cluster(pixel_arrays=pa, cluster_meta=meta))
    return FragmentCluster(
        rate=float(payload.get("rate", 0.0)), start=float(payload.get("start", 0.0)),
        stop=float(payload.get("stop", 0.0)), bpp=1.0, shift=0.0,
        f_low=float(payload.get("f_low", 0.0)), f_high=float(payload.get("f_high", 0.0)),
        n_pix=int(sum(len(c.pixel_arrays) for c in clusters)), run=0, pair=False,
        subnet_threshold=0.0, clusters=clusters)


def _payload_shape(payload: dict) -> tuple[int, int]:
    mask = payload.get("mask")
    if mask is not None:
        return tuple(np.asarray(mask).shape)
    if payload.get("shape") is not None:
        return tuple(int(v) for v in payload["shape"])
    raise ValueError("cluster_pixels needs the payload's mask (or shape)")


def cluster_pixels(pixel_candidates: dict, kt: int = 1, kf: int = 1) -> FragmentCluster:
    """clustering.py:18-157: one payload of select_network_pixels."""
    if pixel_candidates is None:
        raise ValueError("cluster_pixels requires pixel_candidates")
    dev = require_gpu()
    n_freq, n_time = _payload_shape(pixel_candidates)
    f = np.asarray(pixel_candidates.get("frequency", []), dtype=np.int64)
    t = np.asarray(pixel_candidates.get("time", []), dtype=np.int64)
    de = np.asarray(pixel_candidates.get("pix_det_energy", np.empty((0, 0))), dtype=np.float64)
    di = np.asarray(pixel_candidates.get("pix_det_index", np.empty((0, 0), np.int64)),
                    dtype=np.int64)
    e = np.asarray(pixel_candidates.get("energy", []), dtype=np.float64)
    layers = int(pixel_candidates.get("layers", 1))
    rate = float(pixel_candidates.get("rate", 0.0))
    n = f.size
    if de.ndim != 2 or de.shape[0] != n:
        de = de.reshape(n, -1) if de.size else np.zeros((n, 0))
    h2d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sel = {"lag_off": np.array([0, n], dtype=np.int64), "frequency": h2d(f), "time": h2d(t),
           "energy": h2d(e), "pix_det_energy": h2d(de), "n_freq": n_freq, "n_time": n_time}
    out = cluster_pixels_device(sel, kt=kt, kf=kf, rate=rate)
    nc = out["n_clusters"]
    members = out["members"][:n].cpu().numpy()
    offsets = out["cluster_off"][:nc + 1].cpu().numpy()
    stats = [out[k][:nc].cpu().numpy() for k in ("energy", "c_time", "c_freq", "subnet",
                                                   "subrho")]
    return _fragment(pixel_candidates, f, t, e, de, di, members, offsets, stats, layers, rate)


def cluster_pixels_lags(sel: dict, cache: dict, kt: int = 1, kf: int = 1) -> list:
    """One FragmentCluster per lag of a device selection
    (select_network_pixels_device) with its selection cache."""
    rate = float(cache["rate"])
    out = cluster_pixels_device(sel, kt=kt, kf=kf, rate=rate)
    lo, lco = out["lag_off"], out["lag_cluster_off"]
    n, nc = int(lo[-1]), out["n_clusters"]
    host = lambda k, m: out[k][:m].cpu().numpy()
    members, offsets = host("members", n), host("cluster_off", nc + 1)
    stats = [host(k, nc) for k in ("energy", "c_time", "c_freq", "subnet", "subrho")]
    f, t, e = (sel[k].cpu().numpy()[:n] for k in ("frequency", "time", "energy"))
    de, di = sel["pix_det_energy"].cpu().numpy()[:n], sel["pix_det_index"].cpu().numpy()[:n]
    payload = {"rate": rate, "start": cache["start"], "stop": cache["stop"],
               "f_low": cache["f_low"], "f_high": cache["f_high"]}
    frags = []
    for l in range(lo.size - 1):
        a, b, ca, cb = int(lo[l]), int(lo[l + 1]), int(lco[l]), int(lco[l + 1])
        off = offsets[ca:cb + 1] - offsets[ca]
        mem = members[offsets[ca]:offsets[cb]] - a
        frags.append(_fragment(payload, f[a:b], t[a:b], e[a:b], de[a:b], di[a:b], mem, off,
                               [s[ca:cb] for s in stats], int(sel["n_freq"]), rate))
    return frags
