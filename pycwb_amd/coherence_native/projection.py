"""Time-delay max-energy maps on the GPU with the signatures of PycWB's
coherence_native projection / time-delay modules.

  _compute_packet_energy_params  (time_delay_packet.py:8-64)
  time_delay_max_energy_numba    (time_delay_numba.py:578-679), pattern path:
      the series and its copies shifted by k = d, 2d, ... <= max_delay samples
      to both sides go through the quadrature t2w in batches
      (pycwb_wdm_t2w_shifted, which reads each copy straight from the series;
      pycwb_max_energy_shift + pycwb_wdm_t2w for other filter lengths), each
      map's packet energy is folded into the pixel maximum
      (pycwb_max_energy_packet), and the lowest layers are zeroed
      (pycwb_max_energy_finish)
  gamma_to_gauss_device          (types/time_frequency_map.py:143-246,
      TimeFrequencyMap.Gamma2Gauss): counts, order statistics, sums and the
      transform on the GPU (pycwb_g2g_phase), the Gamma fit on the host;
      gamma_to_gauss_device_many does several maps with shared host reads;
      gamma_to_gauss is the same on host arrays (used when a histogram is
      requested)
  max_energy                     (projection.py:63-124)

The complex path (pattern 0), which the reference delegates to JAX, is not
provided.
"""

from __future__ import annotations

import copy
import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..device import lib, ptr, require_gpu, stream_ptr


def _compute_packet_energy_params(M, T, pattern, edge, wavelet_rate, f_low, f_high, df):
    """time_delay_packet.py:8-64: (jb, je, mL, mH, mean, p[9])."""
    J = M * T
    jb = int(edge * float(wavelet_rate) / 4.0) * M
    jb = max(jb, 4 * M)
    je = J - jb
    jb, je = max(jb, 0), min(je, J)
    mL, mH = max(int(f_low / df + 0.1), 0), min(int(f_high / df + 0.1), M - 1)
    pattern = abs(int(pattern))
    mean = 1.0
    if pattern in (1, 3, 4):
        mean, mL, mH = 3.0, mL + 1, mH - 1
    elif pattern == 2:
        mean = 3.0
    elif pattern in (5, 6):
        mean, mL, mH = 5.0, mL + 2, mH - 2
    elif pattern in (7, 8):
        mean, mL, mH = 5.0, mL + 1, mH - 1
    elif pattern == 9:
        mean, mL, mH = 9.0, mL + 1, mH - 1
    offsets = {1: [1, -1], 2: [M, -M], 3: [M + 1, -M - 1], 4: [-M + 1, M - 1],
               5: [M + 1, -M - 1, 2 * M + 2, -2 * M - 2],
               6: [-M + 1, M - 1, -2 * M + 2, 2 * M - 2], 7: [1, -1, M, -M],
               8: [M + 1, -M + 1, M - 1, -M - 1],
               9: [1, -1, M, -M, M + 1, M - 1, -M + 1, -M - 1]}
    p = np.zeros(9, dtype=np.int64)
    q = offsets.get(pattern, [])
    p[1:1 + len(q)] = q
    return jb, je, mL, mH, mean, p


def delay_shifts(max_delay: int, downsample: int, size: int) -> np.ndarray:
    """0, then +k, -k for k = d, 2d, ... <= max_delay, k < size (left, right)."""
    ks = []
    k = int(downsample)
    while k <= max_delay and k < size:
        ks += [k, -k]
        k += int(downsample)
    return np.asarray([0] + ks, dtype=np.int64)


def max_energy_device(ts: torch.Tensor, wdm, max_delay: int, downsample: int, pattern: int,
                      edge: float, wavelet_rate: int, f_low: float, f_high: float, df: float,
                      batch: int = 16, stream=None) -> torch.Tensor:
    """(T, M+1) float64 device map: the pixel-wise maximum packet energy over
    the shifted copies of ts (a float64 device series)."""
    if int(pattern) == 0:
        raise NotImplementedError("the complex max-energy path (pattern 0) is not provided")
    if int(downsample) <= 0:
        raise ValueError("downsample must be >= 1")
    ts = ts.to(torch.float64).contiguous().reshape(-1)
    L = int(ts.numel())
    M = int(wdm.M)
    M1 = M + 1
    T = (L + (M - L % M if L % M else 0)) // M
    jb, je, mL, mH, mean, p = _compute_packet_energy_params(M1, T, pattern, edge, wavelet_rate,
                                                            f_low, f_high, df)
    shifts = delay_shifts(int(max_delay), int(downsample), L)
    dev = ts.device
    cur = torch.empty((T, M1), dtype=torch.float64, device=dev)
    st = stream_ptr(stream)
    d = _lib.PacketDesc()
    d.T, d.M1, d.jb, d.je, d.mL, d.mH, d.mean = T, M1, jb, je, mL, mH, float(mean)
    for k in range(9):
        d.p[k] = int(p[k])
    d.cur = ptr(cur)
    shifts_dev = torch.from_numpy(shifts).to(dev)   # one upload for every batch
    for b0 in range(0, shifts.size, batch):
        sh = shifts_dev[b0:b0 + batch]
        n = int(sh.numel())
        # the transforms read the shifted copies straight from ts where the
        # filter allows it; otherwise the copies are written out first
        maps = wdm.t2w_shifted_device(ts, sh, stream=stream) \
            if hasattr(wdm, "t2w_shifted_device") else None
        if maps is None:
            xs = torch.empty((n, L), dtype=torch.float64, device=dev)
            _lib.check(lib().pycwb_max_energy_shift(ptr(ts), L, ptr(sh), n, ptr(xs), st),
                       "pycwb_max_energy_shift")
            maps = wdm.t2w_device(xs, MM=-1, stream=stream)         # (n, M+1, T) view
            del xs
        mr = torch.view_as_real(maps.transpose(1, 2))              # (n, T, M+1, 2)
        if not mr.is_contiguous():
            mr = mr.contiguous()
        d.maps, d.n_maps, d.init = ptr(mr), n, 1 if b0 == 0 else 0
        _lib.check(lib().pycwb_max_energy_packet(C.byref(d), st), "pycwb_max_energy_packet")
        del maps, mr, sh
    _lib.check(lib().pycwb_max_energy_finish(ptr(cur), T, M1,
                                             1 if abs(int(pattern)) in (5, 6, 9) else 0, st),
               "pycwb_max_energy_finish")
    return cur


def gamma_to_gauss(data: np.ndarray, edge: float, wavelet_rate: int, hist=None):
    """TimeFrequencyMap.Gamma2Gauss (time_frequency_map.py:143-246) on an
    (M, T) energy map: (transformed map, ALP), or (map, 0.0) on failure."""
    original = np.asarray(data)
    shape = original.shape
    flat_2d = np.asarray(original.real if np.iscomplexobj(original) else original, np.float64)
    flat = flat_2d.T.ravel() if flat_2d.ndim == 2 else flat_2d.ravel()
    if flat.size < 4:
        return data, 0.0
    nn_all = int(flat.size)
    M = int(shape[0]) if len(shape) == 2 else 1
    nL = int(float(edge or 0.0) * wavelet_rate * M)
    nL = max(0, min(nL, nn_all - 2))
    nR = nn_all - nL - 1
    if nR <= nL:
        return data, 0.0
    region = flat[nL:nR]
    ws_region = flat[nL:nR + 1]
    wavecount_1 = int(np.sum(flat > 0.001))
    fff = (nR - nL) * wavecount_1 / float(nn_all)
    split = max(nL, min(nR - int(0.5 * fff), nR))
    med = float(np.partition(ws_region, split - nL)[split - nL])
    if med <= 0.0:
        return data, 0.0
    valid = region[(region > 0.01) & (region < 20.0 * med)]
    if valid.size == 0:
        return data, 0.0
    aaa, bbb, count = float(np.sum(valid)), float(np.sum(np.log(valid))), float(valid.size)
    alp = np.log(aaa / count) - bbb / count
    if alp <= 0.0:
        return data, 0.0
    alp = (3.0 - alp + np.sqrt((alp - 3.0) * (alp - 3.0) + 24.0 * alp)) / (12.0 * alp)
    avr = med * (3.0 * alp + 0.2) / (3.0 * alp - 0.8)
    ALP = med * alp / avr
    amp = flat * alp / avr
    out = np.zeros_like(flat)
    pos = amp >= ALP
    if np.any(pos):
        out[pos] = amp[pos] - ALP * (1.0 + np.log(amp[pos] / ALP))
    if hist is not None:
        hist.extend(out[nL:nR].tolist())
    if out[nL:nR].size == 0:
        return data, 0.0
    ws2 = out[nL:nR + 1]
    wavecount_2 = int(np.sum(out[nL:nn_all - nL] > 1.0e-5))
    split = max(nL, min(nR - int(0.3173 * wavecount_2), nR))
    qv = float(np.partition(ws2, split - nL)[split - nL])
    if qv <= 0.0:
        return data, 0.0
    out *= 1.0 / qv
    if hist is not None:
        hist.extend(np.sqrt(np.clip(out[nL:nR], 0.0, None)).tolist())
    if len(shape) == 2:
        M2, T2 = shape
        return out.reshape(T2, M2).T, float(ALP)
    return out, float(ALP)


def gamma_to_gauss_device(energy: torch.Tensor, edge: float, wavelet_rate: int, stream=None):
    """Gamma2Gauss (time_frequency_map.py:143-246) of a (T, M+1) time-major
    float64 device map: (transformed (T, M+1) device map, ALP), or (energy,
    0.0) where the reference gives up.  The counts, order statistics, the sums
    of the valid values and the transform run on the GPU (pycwb_g2g_phase);
    the Gamma fit on those scalars is the reference's host arithmetic.  The
    device log differs from NumPy's in the last bit for some values, so ALP
    and the map agree with the host version to about 1e-15 relative."""
    return gamma_to_gauss_device_many([energy], edge, wavelet_rate, stream)[0]


def gamma_to_gauss_device_many(energies, edge: float, wavelet_rate: int, stream=None):
    """gamma_to_gauss_device of several maps (the detectors of one level), their
    phases interleaved: each phase runs for every map still going, then one
    read to the host serves them all.  Every map goes through the same kernels
    and the same host arithmetic as a call of its own, so each result equals
    the single-map call's; only the host round trips are shared."""
    stream = stream or torch.cuda.current_stream()
    st = stream_ptr(stream)
    res = [None] * len(energies)
    jobs = []
    for idx, energy in enumerate(energies):
        T, M1 = (int(v) for v in energy.shape)
        flat = energy.to(torch.float64).contiguous().reshape(-1)
        n = int(flat.numel())
        if n < 4:
            res[idx] = (energy, 0.0)
            continue
        nL = int(float(edge or 0.0) * wavelet_rate * M1)
        nL = max(0, min(nL, n - 2))
        nR = n - nL - 1
        if nR <= nL:
            res[idx] = (energy, 0.0)
            continue
        dev = flat.device
        win = nR - nL + 1
        n_chunks = max(1, (nR - nL + 8191) // 8192)
        j = dict(idx=idx, energy=energy, T=T, M1=M1, flat=flat, n=n, nL=nL, nR=nR,
                 count=torch.zeros(2, dtype=torch.int64, device=dev),
                 srt=torch.empty(win, dtype=torch.float64, device=dev),
                 n_valid=torch.zeros(1, dtype=torch.int64, device=dev),
                 chunk=torch.empty(2 * n_chunks, dtype=torch.float64, device=dev),
                 out=torch.empty_like(flat))
        d = _lib.G2gDesc()
        d.flat, d.n, d.nL, d.nR = ptr(flat), n, nL, nR
        d.count, d.sorted, d.n_valid, d.chunk_sums, d.out = ptr(j["count"]), ptr(j["srt"]), \
            ptr(j["n_valid"]), ptr(j["chunk"]), ptr(j["out"])
        nb = C.c_size_t()
        _lib.check(lib().pycwb_g2g_work_bytes(C.byref(d), C.byref(nb)), "pycwb_g2g_work_bytes")
        j["work"] = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=dev)
        d.work, d.work_bytes = ptr(j["work"]), int(nb.value)
        j["d"] = d
        jobs.append(j)

    def phase(j, k):
        _lib.check(lib().pycwb_g2g_phase(C.byref(j["d"]), k, st), "pycwb_g2g_phase")

    def fail(j):
        res[j["idx"]] = (j["energy"], 0.0)

    def host(vals):  # one device-to-host read of one scalar per job
        stream.synchronize()
        return torch.stack(vals).cpu().tolist()

    for j in jobs:
        phase(j, 1)
    if jobs:
        for j, c in zip(jobs, host([j["count"][0] for j in jobs])):
            fff = (j["nR"] - j["nL"]) * int(c) / float(j["n"])
            j["split"] = max(j["nL"], min(j["nR"] - int(0.5 * fff), j["nR"]))
        meds = host([j["srt"][j["split"] - j["nL"]] for j in jobs])
        live = []
        for j, med in zip(jobs, meds):
            med = float(med)
            if med <= 0.0:
                fail(j)
                continue
            j["med"] = med
            j["d"].lo, j["d"].hi = 0.01, 20.0 * med
            phase(j, 2)
            live.append(j)
        jobs = live
    if jobs:
        nv = [int(m) for m in host([j["n_valid"][0] for j in jobs])]
        live = [(j, m) for j, m in zip(jobs, nv) if m > 0]
        for j, m in zip(jobs, nv):
            if m == 0:
                fail(j)
        jobs = []
        if live:
            parts = [j["chunk"][:2 * ((m + 8191) // 8192)] for j, m in live]
            stream.synchronize()
            allch = torch.cat(parts).cpu().numpy()
            off = 0
            for (j, m), part in zip(live, parts):
                ch = allch[off:off + part.numel()]
                off += part.numel()
                aaa = bbb = 0.0
                for k in range(ch.size // 2):
                    aaa += float(ch[2 * k])
                    bbb += float(ch[2 * k + 1])
                cnt = float(m)
                alp = np.log(aaa / cnt) - bbb / cnt
                if alp <= 0.0:
                    fail(j)
                    continue
                med = j["med"]
                alp = (3.0 - alp + np.sqrt((alp - 3.0) * (alp - 3.0) + 24.0 * alp)) / (12.0 * alp)
                avr = med * (3.0 * alp + 0.2) / (3.0 * alp - 0.8)
                ALP = med * alp / avr
                j["ALP"] = ALP
                j["d"].alp, j["d"].avr, j["d"].ALP = float(alp), float(avr), float(ALP)
                phase(j, 3)
                jobs.append(j)
    if jobs:
        for j, c in zip(jobs, host([j["count"][1] for j in jobs])):
            j["split2"] = max(j["nL"], min(j["nR"] - int(0.3173 * int(c)), j["nR"]))
        qvs = host([j["srt"][j["split2"] - j["nL"]] for j in jobs])
        for j, qv in zip(jobs, qvs):
            qv = float(qv)
            if qv <= 0.0:
                fail(j)
                continue
            j["d"].scale = 1.0 / qv
            phase(j, 4)
            res[j["idx"]] = (j["out"].view(j["T"], j["M1"]), float(j["ALP"]))
    return res


def _series_of(tf_map, wdm, device):
    ts = getattr(tf_map, "ts_data", None)
    if ts is not None:
        return torch.from_numpy(np.ascontiguousarray(np.asarray(ts, np.float64))).to(device)
    n = getattr(tf_map, "len_timeseries", None)
    data = np.asarray(tf_map.data)
    if not n:
        n = wdm.M * data.shape[1]
    return wdm.w2t_device(torch.from_numpy(np.ascontiguousarray(data.astype(np.complex128)))
                          .to(device), int(n))[0]


def time_delay_max_energy_numba(tf_map, dt, downsample=1, pattern=0, hist=None, mode=None):
    """time_delay_numba.py:578-679 (pattern != 0): (new_tf_map, alp).  tf_map
    carries data (M+1, T) complex, dt, df, edge, f_low, f_high, wavelet (M,
    m_H, filter) and ts_data or len_timeseries."""
    if abs(int(pattern)) == 0:
        raise NotImplementedError("time_delay_max_energy_numba: pattern=0 (complex path) is "
                                  "not provided; use pattern != 0")
    if int(downsample) <= 0:
        raise ValueError("downsample must be >= 1")
    if not np.isfinite(dt):
        raise ValueError("dt must be finite")
    wdm = tf_map.wavelet
    dev = require_gpu()
    n_freq = int(np.asarray(tf_map.data).shape[0])
    rate = float(2.0 * float(tf_map.df) * (n_freq - 1))
    f_low = 0.0 if getattr(tf_map, "f_low", None) is None else float(tf_map.f_low)
    f_high = (float(tf_map.df) * (n_freq - 1) if getattr(tf_map, "f_high", None) is None
              else float(tf_map.f_high))
    wavelet_rate = int(1.0 / float(tf_map.dt))
    ts = _series_of(tf_map, wdm, dev)
    cur = max_energy_device(ts, wdm, int(rate * abs(float(dt))), int(downsample), int(pattern),
                            float(getattr(tf_map, "edge", 0.0) or 0.0), wavelet_rate, f_low,
                            f_high, float(tf_map.df))
    edge = float(getattr(tf_map, "edge", 0.0) or 0.0)
    new = copy.copy(tf_map)
    if hist is None:
        out, alp = gamma_to_gauss_device(cur, edge, wavelet_rate)
        new.data = out.t().contiguous().cpu().numpy()
    else:   # the histogram of the reference's diagnostics is filled on the host
        new.data, alp = gamma_to_gauss(cur.t().contiguous().cpu().numpy(), edge, wavelet_rate,
                                       hist)
    return new, alp


def max_energy(tf_map, max_delay, up_n, pattern, f_low=None, f_high=None, hist=None,
               backend="numba"):
    """projection.py:63-124: the max-energy map and its Gamma-to-Gauss pivot."""
    if hasattr(tf_map, "bandpass"):
        tf_map.bandpass(f_low=f_low, f_high=f_high)
    else:
        if f_low is not None:
            tf_map.f_low = f_low
        if f_high is not None:
            tf_map.f_high = f_high
    return time_delay_max_energy_numba(tf_map, max_delay, downsample=up_n, pattern=pattern,
                                       hist=hist)
