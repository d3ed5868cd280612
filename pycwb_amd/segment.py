"""One analysis segment end to end on the GPU: the 600 s, five-detector case.

The chain of PycWB's native segment processor for the stages built here
(pycwb/workflow/subflow/process_job_segment_native.py: coherence setup
coherence_native/setup.py:100-211, per-lag pipeline coherence_native/
pipeline.py:81-172, then the likelihood, likelihoodWP/likelihood.py:58-135):
per WDM level, t2w of every detector's whitened series, time-delay
max-energy maps and their Gamma-to-Gauss, the pixel-energy threshold, network
pixel selection and clustering; then the TD amplitudes of every cluster of
every level, one batched regulator + sky scan + argmax, and optionally the
gradient refinement of every cluster's direction.  Every map stays on the
device; the host does the reference's own scalar fits.  Super-clustering
across levels (the defragmentation between clustering and likelihood) is not
built, so every level's clusters go to the likelihood.

Also here: the synthetic input of that case -- seeded coloured Gaussian noise
(PSD shape (f_a/f)^4 + 1 + (f/f_b)^2) with sine-Gaussian and white-noise-burst
injections added as strain, whitened by the same PSD to unit variance.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import sky, synth

LEVELS = (8, 16, 32, 64, 128, 256, 512)


def psd_shape(f, f_a: float = 40.0, f_b: float = 300.0, f_min: float = 8.0):
    """One-sided noise PSD shape (arbitrary units): (f_a/f)^4 + 1 + (f/f_b)^2,
    held at its f_min value below f_min."""
    f = np.maximum(np.asarray(f, dtype=np.float64), f_min)
    return (f_a / f) ** 4 + 1.0 + (f / f_b) ** 2


def _whiten_short(h, rate, pad=4096):
    """Whitened copy of a short strain snippet (zero-padded FFT filter)."""
    n = h.size + 2 * pad
    H = np.fft.rfft(np.concatenate([np.zeros(pad), h, np.zeros(pad)]))
    f = np.fft.rfftfreq(n, 1.0 / rate)
    return np.fft.irfft(H / np.sqrt(psd_shape(f)), n)[pad:pad + h.size]


def _projections(ifos, phi, lat, t_rel_fn, hp_fn, hx_fn, rate, i0, i1):
    """Detector strains of a source at Earth-fixed (phi, lat): delay
    n.(R_i - R_0)/c relative to the first detector, F+ and Fx at psi = 0."""
    dets = [sky.Detector(n) for n in ifos]
    nh = np.array([math.cos(lat) * math.cos(phi), math.cos(lat) * math.sin(phi), math.sin(lat)])
    t = np.arange(i0, i1) / rate
    out = []
    for d in dets:
        shift = float(nh @ (d.vertex - dets[0].vertex) / sky.C_LIGHT)
        fp, fx = d.antenna_pattern_gha(np.array([-phi]), np.array([lat]))
        tt = t_rel_fn(t + shift)
        out.append(float(fp[0]) * hp_fn(tt) + float(fx[0]) * hx_fn(tt))
    return out


def make_injections(ifos, rate, L, n_inj, seed, wnb_fraction=0.5):
    """Injection list: times uniform in the series' middle, f0 uniform in
    50-1000 Hz, network SNR uniform in 8-40, directions uniform on the sphere.
    Sine-Gaussians (Q = 9, linearly polarised) and white-noise bursts (two
    independent band-limited polarisations, bandwidth 50-250 Hz, Gaussian
    envelope of 10-50 ms)."""
    rng = np.random.default_rng(seed)
    inj = []
    for k in range(n_inj):
        kind = "wnb" if rng.uniform() < wnb_fraction else "sg"
        inj.append({"kind": kind, "time": rng.uniform(20.0, L / rate - 20.0),
                    "f0": rng.uniform(50.0, 1000.0), "snr": rng.uniform(8.0, 40.0),
                    "phi": rng.uniform(0.0, 2 * math.pi),
                    "lat": math.asin(rng.uniform(-1.0, 1.0)),
                    "bw": rng.uniform(50.0, 250.0), "dur": rng.uniform(0.01, 0.05),
                    "seed": int(rng.integers(1 << 31))})
    inj.sort(key=lambda d: d["time"])
    return inj


def _strain(ifos, rate, d):
    """Per-detector strain snippets (start index, arrays) of one injection."""
    t0, f0 = d["time"], d["f0"]
    if d["kind"] == "sg":
        tau = 9.0 / (math.sqrt(2.0) * math.pi * f0)
        half = 8 * tau + 0.05
        i0, i1 = int((t0 - half) * rate), int((t0 + half) * rate) + 1
        env = lambda tt: np.exp(-(tt / tau) ** 2)
        hp = lambda tt: env(tt) * np.sin(2 * math.pi * f0 * tt)
        hx = lambda tt: np.zeros_like(tt)
    else:
        dur, bw = d["dur"], d["bw"]
        half = 5 * dur + 0.05
        i0, i1 = int((t0 - half) * rate), int((t0 + half) * rate) + 1
        n = i1 - i0 + 2048
        g = np.random.default_rng(d["seed"])
        freqs = np.fft.rfftfreq(n, 1.0 / rate)
        band = (freqs >= max(f0 - bw / 2, 20.0)) & (freqs <= f0 + bw / 2)
        pols = []
        for _ in range(2):
            spec = (g.normal(size=freqs.size) + 1j * g.normal(size=freqs.size)) * band
            pols.append(np.fft.irfft(spec, n))
        tgrid = (np.arange(n) - 1024 + i0) / rate - t0
        env = lambda tt: np.exp(-(tt / dur) ** 2)
        hp = lambda tt: env(tt) * np.interp(tt, tgrid, pols[0])
        hx = lambda tt: env(tt) * np.interp(tt, tgrid, pols[1])
    hs = _projections(ifos, d["phi"], d["lat"], lambda t: t - t0, hp, hx, rate, i0, i1)
    return i0, hs


def synthetic_segment(ifos=synth.NETWORK_5, seconds=600, rate=4096, n_inj=200, seed=11,
                      wnb_fraction=0.5):
    """Whitened (n, L) float64 series with injections, and the injection list."""
    n, L = len(ifos), int(seconds * rate)
    rng = np.random.default_rng(seed)
    f = np.fft.rfftfreq(L, 1.0 / rate)
    asd = np.sqrt(psd_shape(f))
    inj = make_injections(ifos, rate, L, n_inj, seed + 1, wnb_fraction)
    strains = [_strain(ifos, rate, d) for d in inj]
    x = np.empty((n, L))
    for i in range(n):
        # coloured noise: white Gaussian noise shaped by the ASD
        xi = np.fft.irfft(np.fft.rfft(rng.normal(size=L)) * asd, L)
        for d, (i0, hs) in zip(inj, strains):
            # scale each injection to its network SNR in the whitened data
            if "scale" not in d:
                e = sum(float(np.sum(_whiten_short(h, rate) ** 2)) for h in hs)
                d["scale"] = d["snr"] / math.sqrt(e) if e > 0 else 0.0
            xi[i0:i0 + hs[i].size] += d["scale"] * hs[i]
        # whitening by the same PSD, normalised to unit noise variance
        w = np.fft.irfft(np.fft.rfft(xi) / asd, L)
        sigma = 1.4826 * float(np.median(np.abs(w - np.median(w))))
        x[i] = w / sigma
    return x, inj


def run(xd, ifos, rate, stream, order=8, up_tdf=4, td_size=12, bpp=0.001, seg_edge=2.0,
        max_delay_s=0.04, pattern=3, levels=LEVELS, log=None):
    """The GPU chain on whitened device series ``xd`` (n, L) float64.

    Returns a dict with the per-level summaries, the likelihood ClusterBatch
    and network tables, the batch's results, per-stage wall times (device
    synchronised at each stage's end) and the cluster times."""
    from .batch import BatchedSkyScan, ClusterBatch, net_tables
    from .coherence_native.clustering import cluster_pixels_device
    from .coherence_native.projection import gamma_to_gauss_device_many, max_energy_device
    from .coherence_native.selection import select_network_pixels_device, selection_cache_device
    from .coherence_native.veto_threshold import compute_threshold_device
    from .tdamp import TDBatchInputs
    from .wdm import WDM

    dev = xd.device
    n = len(ifos)
    seconds = xd.shape[1] / rate
    stages = {}
    add = lambda k, dt: stages.__setitem__(k, stages.get(k, 0.0) + dt)
    td_rate = rate * up_tdf
    K = int(math.ceil(sky.max_network_delay(ifos, ifos[0]) * td_rate)) + 1
    net = synth.make_network(ifos, order=order, K=K, td_rate=td_rate)
    nt = net_tables([net], device=dev)
    lv_out = []
    for M in levels:
        t = time.perf_counter()
        w = WDM(M, M, 6, 10)
        w.set_td_filter(td_size, up_tdf)
        add("setup", time.perf_counter() - t)
        t = time.perf_counter()
        tf = w.t2w_device(xd)                                       # (n, M+1, N)
        torch.cuda.synchronize()
        add("t2w", time.perf_counter() - t)
        dt_bin, df = M / rate, rate / 2.0 / M
        t = time.perf_counter()
        energy = [max_energy_device(xd[i], w, int(max_delay_s * rate), 4, pattern, seg_edge,
                                    int(round(1 / dt_bin)), 32.0, 2000.0, df)
                  for i in range(n)]
        torch.cuda.synchronize()
        add("max_energy", time.perf_counter() - t)
        t = time.perf_counter()
        maps, alp = [], 0.0
        for g, a in gamma_to_gauss_device_many(energy, seg_edge, int(round(1 / dt_bin)),
                                               stream=stream):
            maps.append(g)
            alp += a
        stack = torch.stack([g.t() for g in maps]).contiguous()      # (n, M+1, N)
        torch.cuda.synchronize()
        add("gamma_to_gauss", time.perf_counter() - t)
        t = time.perf_counter()
        Eo = compute_threshold_device(stack, bpp, alp=alp / n, edge=seg_edge,
                                      wavelet_rate=int(round(1 / dt_bin)))
        torch.cuda.synchronize()
        add("threshold", time.perf_counter() - t)
        t = time.perf_counter()
        cache = selection_cache_device(stack, dt_bin, df, edge=seg_edge, f_low=32.0,
                                       f_high=2000.0, stop=float(seconds))
        sel = select_network_pixels_device(cache, Eo, np.zeros((1, n), np.int64), stream=stream)
        torch.cuda.synchronize()
        add("selection", time.perf_counter() - t)
        t = time.perf_counter()
        cl = cluster_pixels_device(sel, kt=2, kf=3, rate=cache["rate"], stream=stream)
        nc = cl["n_clusters"]
        np_ = int(sel["n_pix"])
        off = cl["cluster_off"][:nc + 1].cpu().numpy()
        mem = cl["members"][:np_]
        pix = (sel["time"][mem] * (M + 1) + sel["frequency"][mem]).to(torch.int32)
        ctime = cl["c_time"][:nc].cpu().numpy()
        torch.cuda.synchronize()
        add("clustering", time.perf_counter() - t)
        lv_out.append({"M": M, "w": w, "tf": tf, "pix": pix, "sizes": np.diff(off),
                       "ctime": ctime, "Eo": Eo, "selected": np_, "n_time": tf.shape[2]})
        if log:
            log(f"segment level M={M}: Eo={Eo:.3f}, {np_} pixels selected, {nc} clusters")
    # one likelihood batch: every cluster of every level
    t = time.perf_counter()
    sizes = np.concatenate([lv["sizes"] for lv in lv_out]).astype(np.int64)
    C_ = sizes.size
    nd = 2 * K + 1
    rms_off = np.concatenate([[0], np.cumsum(sizes * n)[:-1]]).astype(np.int64)
    td_off = np.concatenate([[0], np.cumsum(sizes * n * nd)[:-1]]).astype(np.int64)
    td = torch.empty(2 * int(sizes.sum()) * n * nd, dtype=torch.float32, device=dev)
    rms = torch.full((int(sizes.sum()) * n,), 1.0 / math.sqrt(n), dtype=torch.float32,
                     device=dev)
    c0 = 0
    for lv in lv_out:
        bank = lv["w"].td_filters
        if lv["sizes"].size:
            for i in range(n):
                inp = TDBatchInputs.from_device(lv["tf"][i], bank)
                inp.extract_td_into(lv["pix"], K, td, int(td_off[c0]), i, n, stream=stream)
        c0 += lv["sizes"].size
        lv["tf"] = None
    batch = ClusterBatch(n_pix=sizes.astype(np.int32), net=np.zeros(C_, np.int32),
                         K=np.full(C_, K, np.int32), nifo=np.full(C_, n, np.int32),
                         rms_off=rms_off, td_off=td_off, rms=rms, td=td)
    torch.cuda.synchronize()
    add("td_amplitudes", time.perf_counter() - t)
    valid = torch.arange(net.n_sky, dtype=torch.int64, device=dev)
    scan = BatchedSkyScan(batch, nt, valid, 0.5 * 0.5 * 2 / 3, 0.5, 0.5, block=64)
    scan.launch(stream)                                            # warm (graph build)
    torch.cuda.synchronize()
    t = time.perf_counter()
    scan.launch(stream)
    torch.cuda.synchronize()
    add("likelihood_batch", time.perf_counter() - t)
    return {"levels": lv_out, "net": net, "nets": nt, "batch": batch, "scan": scan,
            "valid": valid, "results": scan.results(), "K": K, "td_rate": td_rate,
            "stage_s": stages, "ctime": np.concatenate([lv["ctime"] for lv in lv_out]),
            "sizes": sizes}


def recovered(inj, ctimes, l_max, net, ifos, td_rate, tol_t=0.1, tol_delay=1e-3) -> int:
    """Injections with a cluster within ``tol_t`` whose argmax direction gives
    the injected detector delays to ``tol_delay`` seconds."""
    ref = sky.Detector(ifos[0]).vertex
    vert = [sky.Detector(d).vertex - ref for d in ifos]
    found = 0
    for d in inj:
        near = np.nonzero(np.abs(ctimes - d["time"]) < tol_t)[0]
        if near.size == 0:
            continue
        nh = np.array([math.cos(d["lat"]) * math.cos(d["phi"]),
                       math.cos(d["lat"]) * math.sin(d["phi"]), math.sin(d["lat"])])
        dl = np.array([nh @ v / sky.C_LIGHT for v in vert])
        rec = (net.ml[:, l_max[near]] - net.ml[0, l_max[near]]) / td_rate
        if np.any(np.all(np.abs(rec - dl[:, None]) < tol_delay, axis=0)):
            found += 1
    return found
