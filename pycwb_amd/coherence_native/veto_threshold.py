"""PycWB's coherence_native.veto_threshold on the GPU: compute_threshold
(veto_threshold.py:98-222) with its array work (detector sum, clip, counts,
order statistics, sums of the positive values) in pycwb_threshold_stats, and
the host constant _igamma_inv_upper (:64-73)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch
from scipy.special import gammaincc, gammainccinv

from .. import _lib
from ..device import lib, ptr, require_gpu, stream_ptr


def _igamma_inv_upper(shape: float, p: float) -> float:
    """veto_threshold.py:64-73: x with Q(shape, x) = p (scipy's gammainccinv),
    p clipped to [1e-12, 1 - 1e-12] and shape to >= 1e-12."""
    p = float(np.clip(p, 1.0e-12, 1.0 - 1.0e-12))
    s = float(max(shape, 1.0e-12))
    return float(gammainccinv(s, p))


def _real64(data) -> np.ndarray:
    arr = np.asarray(data)
    if np.iscomplexobj(arr):
        arr = arr.real
    return np.asarray(arr, dtype=np.float64)


def threshold_stats_device(maps: torch.Tensor, time_major: bool, begin: int, count: int,
                           hi: float, want_sorted: bool, want_sums: bool, stream=None) -> dict:
    """pycwb_threshold_stats on device maps (n_ifo, n_freq, n_time) float64:
    ``counts`` (host: above 1e-4, above 1e-3), device ``values`` and
    ``sorted``, and host ``sum`` / ``log_sum`` of the values above 1e-3 (the
    chunk sums added to 0.0 in order)."""
    maps = maps.contiguous()
    n_ifo, n_freq, n_time = (int(v) for v in maps.shape)
    dev = maps.device
    d = _lib.ThresholdDesc()
    d.maps, d.n_ifo, d.n_freq, d.n_time = ptr(maps), n_ifo, n_freq, n_time
    d.time_major, d.begin, d.count, d.hi = 1 if time_major else 0, int(begin), int(count), \
        float(hi)
    cap = max(int(count), 1)
    values = torch.empty(cap, dtype=torch.float64, device=dev)
    counts = torch.empty(2, dtype=torch.int64, device=dev)
    srt = torch.empty(cap, dtype=torch.float64, device=dev) if want_sorted else None
    n_chunks = (int(count) + 8191) // 8192
    sums = torch.empty(max(2 * n_chunks, 2), dtype=torch.float64, device=dev) \
        if want_sums else None
    d.values, d.counts, d.sorted, d.chunk_sums = ptr(values), ptr(counts), ptr(srt), ptr(sums)
    nb = C.c_size_t()
    _lib.check(lib().pycwb_threshold_work_bytes(C.byref(d), C.byref(nb)),
               "pycwb_threshold_work_bytes")
    work = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=dev)
    d.work, d.work_bytes = ptr(work), int(nb.value)
    stream = stream or torch.cuda.current_stream()
    _lib.check(lib().pycwb_threshold_stats(C.byref(d), stream_ptr(stream)),
               "pycwb_threshold_stats")
    stream.synchronize()
    c = counts.cpu().numpy()
    out = {"counts": (int(c[0]), int(c[1])), "values": values[:count], "work": work,
           "sorted": None if srt is None else srt[:count]}
    if want_sums:
        m = int(c[1])
        ch = sums[:2 * ((m + 8191) // 8192)].cpu().numpy()
        s = ls = 0.0
        for k in range(ch.size // 2):
            s += float(ch[2 * k])
            ls += float(ch[2 * k + 1])
        out["sum"], out["log_sum"] = s, ls
    return out


def compute_threshold(tf_maps, bpp: float, alp: float | None = None,
                      edge: float | None = None) -> float:
    """veto_threshold.py:98-222: the pixel energy threshold.  The detector sum,
    clip, counts, order statistics and the sums of the positive values run on
    the GPU; the Gamma fits on those scalars are the reference's host
    arithmetic."""
    if not tf_maps or not hasattr(tf_maps[0], "data"):
        raise ValueError("compute_threshold requires TF maps")
    dev = require_gpu()
    arrays = [_real64(t.data) for t in tf_maps]
    two_d = arrays[0].ndim == 2
    arrays = [a if a.ndim == 2 else a.reshape(1, -1) for a in arrays]
    if any(a.shape != arrays[0].shape for a in arrays[1:]):
        raise ValueError("all detector TF maps must have the same shape")
    stack = torch.from_numpy(np.ascontiguousarray(np.stack(arrays))).to(dev)
    dt = None
    if alp is None and two_d and edge is not None:
        n_time = arrays[0].shape[1]
        crops = set()
        for t in tf_maps:
            e = 0
            if hasattr(t, "dt") and t.dt > 0:
                e = int(max(0, round(float(edge) / float(t.dt))))
                if not (e > 0 and n_time > 2 * e):
                    e = 0
            crops.add(e)
        if len(crops) > 1:
            raise ValueError("detector TF maps are cropped to different shapes")
        dt = float(tf_maps[0].dt) if getattr(tf_maps[0], "dt", 0) and tf_maps[0].dt > 0 \
            else None
    return compute_threshold_device(stack, bpp, alp=alp, edge=edge,
                                    wavelet_rate=getattr(tf_maps[0], "wavelet_rate", None),
                                    dt=dt, two_d=two_d)


def compute_threshold_device(stack: torch.Tensor, bpp: float, alp: float | None = None,
                             edge: float | None = None, wavelet_rate=None, dt=None,
                             two_d: bool = True) -> float:
    """compute_threshold on maps already on the device: ``stack`` float64
    (n_ifo, layers, time bins) (a 1-D map as one row, ``two_d`` False);
    ``wavelet_rate`` for the shape (alp) path, ``dt`` for the edge crop of the
    no-shape path."""
    n_ifo, n_freq, n_time = (int(v) for v in stack.shape)
    stack = stack.to(torch.float64).contiguous()
    hi = n_ifo * 100.0
    if alp is not None:
        M = n_freq if two_d else 1
        size = n_freq * n_time
        nL = int(float(edge or 0.0) * wavelet_rate * M)
        nR = size - nL - 1
        start, stop, _ = slice(nL, nR).indices(size)
        count = max(0, stop - start)
        r = threshold_stats_device(stack, True, start, count, hi, False, True)
        n_pos = r["counts"][1]
        if n_pos == 0:
            return 0.0
        avr = float(np.float64(r["sum"]) / n_pos)
        bbb = float(np.float64(r["log_sum"]) / n_pos)
        alp_fit = np.log(avr) - bbb
        alp_fit = (3 - alp_fit + np.sqrt((alp_fit - 3) * (alp_fit - 3) + 24 * alp_fit)) / \
            (12 * alp_fit)
        bpp_corr = float(bpp) * alp_fit / float(alp)
        return avr * _igamma_inv_upper(alp_fit, bpp_corr) / alp_fit / 2.0

    crop = 0
    if two_d and edge is not None and dt is not None and dt > 0:
        e = int(max(0, round(float(edge) / float(dt))))
        crop = e if (e > 0 and n_time > 2 * e) else 0
    ncol = n_time - 2 * crop
    n_total = n_freq * ncol
    if n_total == 0:
        return 0.0
    r = threshold_stats_device(stack, False, crop, n_total, hi, True, False)
    if r["counts"][1] == 0:
        return 0.0
    fff = float(np.int64(r["counts"][0]) / n_total)
    if fff <= 0.0:
        return 0.0
    k_val = int(float(bpp) * fff * n_total)
    k_med = int(0.2 * fff * n_total)
    srt = r["sorted"]
    val = float(srt[max(0, n_total - k_val - 1)] if k_val > 0 else srt[-1])
    med = float(srt[max(0, n_total - k_med - 1)] if k_med > 0 else srt[-1])
    m = 1.0
    p00 = 0.0
    while p00 < 0.2:
        p00 = float(gammaincc(n_ifo * m, med))
        m += 0.01
    if m > 1.01:
        m -= 0.01
    return 0.3 * (_igamma_inv_upper(n_ifo * m, float(bpp)) + val) + n_ifo * np.log(m)
