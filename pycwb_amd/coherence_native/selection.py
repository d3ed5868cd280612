"""Network-pixel selection on the GPU with the signatures of PycWB's
coherence_native.selection (selection.py:14-243).

  _shift_bins_from_lag_shifts  (:14-28)   per-detector lag shifts -> bins
  _build_selection_cache       (:31-122)  lag-invariant inputs; the energy
                                          stack is kept on the device
  select_network_pixels        (:125-243) one lag, the reference's payload
  select_network_pixels_lags              every lag of a segment in one pass
                                          (pycwb_select_count / _emit)
  select_network_pixels_device            the same, the pixels left on the
                                          device (clustering's input)
  selection_cache_device                  the cache of maps already on the
                                          device

The energy maps are the real parts of the detectors' TF maps (the
max-energy maps the coherence setup produces; veto_threshold.py:75-95).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..device import lib, ptr, require_gpu, stream_ptr


def _shift_bins_from_lag_shifts(lag_shifts, n_ifo: int, rate: float) -> np.ndarray:
    """selection.py:14-28: circular TF-bin shifts relative to the smallest."""
    if lag_shifts is None:
        sec = np.zeros(n_ifo, dtype=float)
    else:
        sec = np.asarray(lag_shifts, dtype=float)
        if sec.size != n_ifo:
            raise ValueError("lag_shifts size mismatch with number of detectors")
    ref = float(np.min(sec)) if sec.size else 0.0
    return np.asarray([int((float(s) - ref) * rate + 0.001) for s in sec], dtype=np.int64)


def _energy_array(tf_map) -> np.ndarray:
    arr = np.asarray(tf_map.data)
    if np.iscomplexobj(arr):
        arr = arr.real
    return np.asarray(arr, dtype=np.float64)


def _build_selection_cache(tf_maps, edge: float = 0.0, lag_shifts_by_lag=None,
                           device=None) -> dict:
    """selection.py:31-122; ``arrays_stack`` is a float64 device tensor."""
    if not tf_maps or not hasattr(tf_maps[0], "data"):
        raise ValueError("select_network_pixels requires TF maps")
    arrays = [_energy_array(t) for t in tf_maps]
    if not all(a.ndim == 2 for a in arrays):
        raise ValueError("python get_network_pixels expects 2D TF arrays")
    n_freq, n_time = arrays[0].shape
    if any(a.shape != (n_freq, n_time) for a in arrays[1:]):
        raise ValueError("all detector TF maps must have the same shape")
    device = device or require_gpu()
    tf0 = tf_maps[0]
    stack = torch.from_numpy(np.ascontiguousarray(np.stack(arrays))).to(device)
    return selection_cache_device(
        stack, float(tf0.dt), float(getattr(tf0, "df", 0.0) or 0.0), edge=edge,
        f_low=getattr(tf0, "f_low", 0.0), f_high=getattr(tf0, "f_high", None),
        start=float(getattr(tf0, "start", 0.0)), stop=float(getattr(tf0, "stop", 0.0)),
        lag_shifts_by_lag=lag_shifts_by_lag,
        f_high_out=float(getattr(tf0, "f_high", (n_freq - 1) * float(getattr(tf0, "df", 0.0)
                                                                      or 0.0)) or 0.0))


def selection_cache_device(stack: torch.Tensor, dt: float, df: float, edge: float = 0.0,
                           f_low=0.0, f_high=None, start: float = 0.0, stop: float = 0.0,
                           lag_shifts_by_lag=None, f_high_out=None) -> dict:
    """The selection cache of maps already on the device: ``stack`` float64
    (n_ifo, n_freq, n_time) energy maps; the band, edge and lag bins as
    _build_selection_cache computes them (selection.py:31-122); ``f_high_out``
    is the payload's f_high (the band edge when not given)."""
    n_ifo, n_freq, n_time = (int(v) for v in stack.shape)
    if dt <= 0.0:
        raise ValueError("tf_map.dt must be positive")
    rate = 1.0 / dt
    edge_bins = int(max(0, float(edge) * rate + 0.001))
    f_low = float(f_low or 0.0)
    f_high = float(f_high) if f_high is not None else float((n_freq - 1) * df)
    ib, ie = 1, n_freq
    if df > 0:
        for idx_f, freq in enumerate(np.arange(n_freq, dtype=np.float64) * df):
            if freq <= f_high:
                ie = idx_f
            if freq <= f_low:
                ib = idx_f + 1
    ie = min(ie, n_freq - 1)
    ib = max(ib, 1)
    shifts = None
    if lag_shifts_by_lag is not None:
        rows = list(lag_shifts_by_lag)
        shifts = (np.vstack([_shift_bins_from_lag_shifts(r, n_ifo, rate) for r in rows])
                  if rows else np.empty((0, n_ifo), dtype=np.int64))
    return {
        "arrays_stack": stack.to(torch.float64).contiguous(),
        "n_ifo": n_ifo, "n_freq": n_freq, "n_time": n_time, "dt": dt, "rate": rate,
        "edge_bins": edge_bins, "valid_start": edge_bins, "valid_stop": n_time - edge_bins,
        "nn_valid": n_time - 2 * edge_bins, "ib": ib, "ie": ie,
        "start": start, "stop": stop, "f_low": f_low,
        "f_high": f_high if f_high_out is None else float(f_high_out),
        "shift_bins_by_lag": shifts,
    }


def select_network_pixels_device(cache: dict, energy_threshold: float, shift_bins: np.ndarray,
                                 veto=None, stream=None) -> dict:
    """Every lag of ``shift_bins`` ((n_lag, n_ifo) bins) in one GPU pass; the
    selected pixels stay on the device, lag-major: ``lag_off`` (n_lag + 1)
    host int64 ranges (and ``lag_off_dev``), ``frequency``, ``time``,
    ``energy``, ``pix_det_energy`` (n, n_ifo), ``pix_det_index`` (n, n_ifo),
    ``flags`` (n_lag, n_freq, n_time) uint8, ``live`` (n_lag, n_time) and the
    support maps ``combined``."""
    n_ifo, n_freq, n_time = cache["n_ifo"], cache["n_freq"], cache["n_time"]
    sh = np.ascontiguousarray(np.asarray(shift_bins, dtype=np.int64).reshape(-1, n_ifo))
    n_lag = sh.shape[0]
    if n_lag < 1 or n_lag > 65535:
        raise ValueError("between 1 and 65535 lags per call")
    if (sh < 0).any():
        raise ValueError("shift bins must be non-negative")
    if n_freq < 3 or n_time < 5 or cache["nn_valid"] < 1:
        raise ValueError("TF maps too small for the selection window")
    E = cache["arrays_stack"]
    dev = E.device
    has_veto = veto is not None and len(veto) == n_time
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sh_d = t(sh)
    veto_d = t(np.asarray(veto, dtype=np.int16)) if has_veto else None
    comb = torch.empty((n_lag, n_freq, n_time), dtype=torch.float64, device=dev)
    flags = torch.empty((n_lag, n_freq, n_time), dtype=torch.uint8, device=dev)
    live = torch.empty((n_lag, n_time), dtype=torch.uint8, device=dev)
    row_off = torch.empty(n_lag * n_freq + 1, dtype=torch.int64, device=dev)
    lag_off = torch.empty(n_lag + 1, dtype=torch.int64, device=dev)
    d = _lib.SelectDesc()
    d.n_ifo, d.n_freq, d.n_time, d.energy = n_ifo, n_freq, n_time, ptr(E)
    d.n_lag, d.shift_bins, d.veto = n_lag, ptr(sh_d), ptr(veto_d)
    d.edge_bins, d.ib, d.ie, d.eo = cache["edge_bins"], cache["ib"], cache["ie"], \
        float(energy_threshold)
    d.combined, d.flags, d.live, d.row_off, d.lag_off = ptr(comb), ptr(flags), ptr(live), \
        ptr(row_off), ptr(lag_off)
    stream = stream or torch.cuda.current_stream()
    st = stream_ptr(stream)
    _lib.check(lib().pycwb_select_count(C.byref(d), st), "pycwb_select_count")
    stream.synchronize()
    lo = lag_off.cpu().numpy()
    total = int(lo[-1])
    cap = max(total, 1)
    fr = torch.empty(cap, dtype=torch.int64, device=dev)
    tm = torch.empty(cap, dtype=torch.int64, device=dev)
    val = torch.empty(cap, dtype=torch.float64, device=dev)
    de = torch.empty((cap, n_ifo), dtype=torch.float64, device=dev)
    di = torch.empty((cap, n_ifo), dtype=torch.int64, device=dev)
    d.capacity, d.freq, d.time, d.value, d.det_energy, d.det_index = total, ptr(fr), ptr(tm), \
        ptr(val), ptr(de), ptr(di)
    _lib.check(lib().pycwb_select_emit(C.byref(d), st), "pycwb_select_emit")
    return {"lag_off": lo, "lag_off_dev": lag_off, "frequency": fr, "time": tm, "energy": val,
            "pix_det_energy": de, "pix_det_index": di, "flags": flags, "live": live,
            "combined": comb,
            "n_pix": total, "n_lag": n_lag, "n_ifo": n_ifo, "n_freq": n_freq,
            "n_time": n_time}


def _to_host(tensors, stream):
    """Device tensors to host arrays through page-locked buffers, every copy
    queued on ``stream`` behind the work that wrote them, one wait at the end."""
    out = []
    with torch.cuda.stream(stream):
        for a in tensors:
            h = torch.empty(a.shape, dtype=a.dtype, pin_memory=True)
            h.copy_(a, non_blocking=True)
            out.append(h)
    stream.synchronize()
    return [h.numpy() for h in out]


def select_network_pixels_lags(cache: dict, energy_threshold: float, shift_bins: np.ndarray,
                               veto=None, stream=None, with_mask: bool = True) -> list[dict]:
    """Every lag of ``shift_bins`` ((n_lag, n_ifo) bins) in one GPU pass; one
    payload per lag (the keys of select_network_pixels, and ``shape``, the
    (n_freq, n_time) of the maps).  ``with_mask=False`` leaves out the dense
    (n_freq, n_time) selection masks (``mask`` is None), which are most of
    the bytes brought back to the host."""
    stream = stream or torch.cuda.current_stream()
    sel = select_network_pixels_device(cache, energy_threshold, shift_bins, veto, stream)
    n_lag, lo = sel["n_lag"], sel["lag_off"]
    keys = ["frequency", "time", "energy", "pix_det_energy", "pix_det_index", "live"] + \
        (["flags"] if with_mask else [])
    host = dict(zip(keys, _to_host([sel[k] for k in keys], stream)))
    fr_h, tm_h, val_h = host["frequency"], host["time"], host["energy"]
    de_h, di_h = host["pix_det_energy"], host["pix_det_index"]
    flags_h = host["flags"] if with_mask else None
    live_h = host["live"]
    out = []
    for l in range(n_lag):
        a, b = int(lo[l]), int(lo[l + 1])
        lm = live_h[l].astype(bool)
        out.append({
            "mask": flags_h[l].astype(bool) if with_mask else None,
            "time": tm_h[a:b], "frequency": fr_h[a:b],
            "energy": val_h[a:b], "pix_det_energy": de_h[a:b], "pix_det_index": di_h[a:b],
            "rate": cache["rate"], "layers": sel["n_freq"], "start": cache["start"],
            "stop": cache["stop"], "f_low": cache["f_low"], "f_high": cache["f_high"],
            "live_mask": lm, "live_samples": int(np.sum(lm)),
            "shape": (sel["n_freq"], sel["n_time"]),
        })
    return out


def select_network_pixels(tf_maps, lag_index: int, energy_threshold: float, lag_shifts=None,
                          veto=None, edge: float = 0.0, selection_cache=None) -> dict:
    """selection.py:125-243 for one lag."""
    if selection_cache is None:
        selection_cache = _build_selection_cache(tf_maps, edge=edge)
    sh = selection_cache.get("shift_bins_by_lag")
    if sh is not None and 0 <= int(lag_index) < len(sh):
        shift = np.asarray(sh[int(lag_index)], dtype=np.int64)
    else:
        shift = _shift_bins_from_lag_shifts(lag_shifts, selection_cache["n_ifo"],
                                            selection_cache["rate"])
    return select_network_pixels_lags(selection_cache, energy_threshold, shift[None, :], veto)[0]
