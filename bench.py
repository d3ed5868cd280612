"""Benchmark: sky x cluster-pixel evaluations/s of the batched GPU likelihood.

Workload (one step): 4,096 synthetic clusters, pixel counts log-uniform in
[16, 4096], half on the 3-detector network (H1, L1, V1) and half on the
4-detector network (+K1), delayed amplitudes with K=64 (129 taps), HEALPix
order 7 (196,608 directions), float32 amplitudes.  One step = DPF regulator +
sky scan + argmax of every cluster, one hipGraph launch of the persistent
work-queue kernels.  Inputs are resident in HBM before timing starts.

Also measured in the same run: the plain family (one cluster per launch)
looped over the same clusters, the roofline of the dominant kernel (the scan
stage) from HIP events on its stream, and the CPU baseline: the oracle's
OpenMP restatement of PycWB's Numba scan on a bounded sample.  Side legs
report under their own keys: "refinement" (512 three-detector clusters, grid
argmax then 20 projected-gradient steps of the continuous sky statistic) and
"wdm" (forward WDM transform of 5 x 600 s at 4096 Hz at M = 8 .. 512),
"tdamp" (TD-amplitude extraction of 200,000 pixels x 129 delays), "subnet"
(sub-network cut of 4,096 superclusters' 300 loudest pixels at order 4),
"xtalk" (cross-talk CSRs of 4,096 clusters from a 7-resolution catalog),
"select" (network-pixel selection of 128 lags of 5 detectors' 600 s maps),
"maxenergy" (time-delay max-energy map of a 600 s series, 87 shifted copies)
"cluster" (connected-component clustering of the 128 lags' selected
pixels), "threshold" (the pixel-energy threshold's array statistics of the
same maps) and "e2e" (5 detectors' 600 s with 200 injections through 7 WDM
levels of max-energy, threshold, selection and clustering into one likelihood
batch at HEALPix order 8).

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector/matrix peak (AMD spec)
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scan_flops(batch, n_valid):
    """Algorithmic flops of the scan stage: (78 n + 129) per evaluation."""
    return float(np.sum(batch.n_pix.astype(np.float64) * (78.0 * batch.nifo + 129.0))) * n_valid


def required_scan_flops(batch, tested, active):
    """Flops the scan needs on this data: the energy test (4n + 4) on the
    evaluations whose energy it computes (``tested``: those its tile energy
    bound does not rule out) and the rest of the chain (74n + 125) on the
    evaluations above the energy threshold (``active``); the others contribute
    exactly nothing."""
    n = batch.nifo.astype(np.float64)
    return float(np.sum(tested.astype(np.float64) * (4.0 * n + 4.0))) + \
        float(np.sum(active.astype(np.float64) * (74.0 * n + 125.0)))


def regulator_flops(batch, n_valid):
    return float(np.sum(batch.n_pix.astype(np.float64) * (24.0 * batch.nifo + 25.0))) * n_valid


def scan_bytes(batch, nets, n_valid):
    """Unique HBM bytes of the scan stage: every cluster's amplitude pairs and
    noise weights once, the sky tables of each network once."""
    td = float(np.sum(batch.n_pix.astype(np.float64) * batch.nifo * (2 * batch.K + 1) * 8))
    rms = float(np.sum(batch.n_pix.astype(np.float64) * batch.nifo * 4))
    sky = float(sum(n * n_valid * 12 for n in nets.nifo))
    return td + rms + sky


def cpu_info() -> dict:
    """CPU model, physical cores of the machine and the cores this process may use."""
    model, phys = None, set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name" and model is None:
                    model = v
            if cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "physical_cores": len(phys) or None, "usable_cpus": usable}


def _time_oracle(o, net, cl, P, threads, repeats):
    """Median over ``repeats`` runs (after one warm-up) of the oracle's regulator +
    scan of one cluster over every direction; returns (seconds, evals)."""
    valid = np.arange(net.n_sky, dtype=np.int64)
    gamma = 0.5 * 0.5 * 2 / 3
    times = []
    for k in range(repeats + 1):
        t0 = time.perf_counter()
        reg = o.calculate_dpf(net.FP_t, net.FX_t, cl["rms"], net.n_sky, 3, gamma, 12.0, valid,
                              threads=threads)
        REG = np.array([np.float32(0.5 * np.sqrt(2)), reg, 0.0], np.float32)
        o.scan_sky(3, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"], cl["td90"],
                   net.ml, REG, 0.5, 0.5, 12.0, valid, threads=threads)
        if k:
            times.append(time.perf_counter() - t0)
    return float(np.median(times)), float(net.n_sky) * P


def cpu_baseline(threads, order=7, P=1024, repeats=5):
    """The oracle (C/OpenMP restatement of PycWB's Numba regulator + sky scan,
    not PycWB itself) on the plain configuration's cluster: three detectors,
    P = 1,024, K = 64, every direction of HEALPix order ``order`` -- the same
    per-evaluation work as the GPU step.  Median of ``repeats`` runs after a
    warm-up, on ``threads`` threads, and on one thread over order 5 (the
    one-thread rate does not depend on the number of directions)."""
    from oracle import oracle as o
    from pycwb_amd import synth

    net = synth.make_network(synth.NETWORK_3, order=order, K=64)
    cl = synth.make_cluster(net, P, seed=1001 * 100000)
    med, evals = _time_oracle(o, net, cl, P, threads, repeats)
    net1 = synth.make_network(synth.NETWORK_3, order=min(order, 5), K=64)
    cl1 = synth.make_cluster(net1, P, seed=1001 * 100000)
    med1, evals1 = _time_oracle(o, net1, cl1, P, 1, repeats)
    info = cpu_info()
    sample = (f"regulator + scan of one three-detector cluster, P={P}, K=64, HEALPix order "
              f"{order} ({net.n_sky} directions): median of {repeats} runs after a warm-up, "
              f"{med:.2f} s on {threads} threads; one thread: order {min(order, 5)}, median "
              f"{med1:.2f} s")
    return {"value": evals / med, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": sample, "one_thread_evals_per_s": evals1 / med1,
            "cpu_model": info["model"], "physical_cores": info["physical_cores"],
            "usable_cpus": info["usable_cpus"],
            "note": "C restatement of PycWB's Numba likelihood (oracle/), gcc -O2, OpenMP"}


def oracle_parity(batch, nets, nt, rb, lm_gpu, gamma, delta, netCC, reg0, n_check=6):
    """l_max bit-exact and Lm relative error against the oracle for the
    n_check smallest clusters of each network (whole sky each)."""
    from oracle import oracle as o

    picks = []
    for k in range(len(nets)):
        ids = np.nonzero(batch.net == k)[0]
        picks += ids[np.argsort(batch.n_pix[ids], kind="stable")][:n_check // len(nets)].tolist()
    valid = np.arange(nets[0].n_sky, dtype=np.int64)
    l_ok, lm_err = True, 0.0
    for c in picks:
        net = nets[batch.net[c]]
        td = batch.cluster_td(c).cpu().numpy()
        td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
        td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
        rms = batch.cluster_rms(c).cpu().numpy()
        En = nt.En[batch.net[c]]
        reg = o.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, net.n_ifo, gamma, En, valid)
        REG = np.array([reg0, reg, 0.0], np.float32)
        l_o, _, _ = o.scan_sky(net.n_ifo, int(batch.n_pix[c]), net.n_sky, net.FP_t, net.FX_t,
                               rms, td00, td90, net.ml, REG, netCC, delta, En, valid)
        st = o.sky_statistics(l_o, net.n_ifo, int(batch.n_pix[c]), net.n_sky, net.FP_t,
                              net.FX_t, rms, td00, td90, net.ml, REG, En)
        l_ok &= int(rb["l_max"][c]) == l_o
        lm_err = max(lm_err, abs(lm_gpu[c] - st["Lm"]) / max(abs(st["Lm"]), 1e-30))
    return {"oracle_clusters": len(picks), "oracle_l_max_bit_exact": bool(l_ok),
            "oracle_Lm_max_rel_err": lm_err}


def profiled_value(key, clusters, order, K):
    """A figure of the default workload from the committed rocprofv3 summary
    (``key: value`` line): the scan stage's HBM bytes per launch
    (``traffic_bytes_per_scan_stage``, FETCH_SIZE and WRITE_SIZE runs each
    scaled by a 1 GiB copy) or its VALU wave-instructions per launch
    (``valu_insts_per_scan_stage``, SQ_INSTS_VALU); None for other workloads
    or without the file."""
    path = os.path.join(ROOT, "profiles", "batched_step_4096clusters_order7.md")
    if (clusters, order, K) != (4096, 7, 64):
        return None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def valu_ceiling(stream, blocks=2048, iters=20000, reps=3):
    """Wave-instructions per second of the VALU probe kernel (eight waves per
    SIMD of independent float32 FMA chains), best of ``reps``."""
    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import lib

    out = torch.zeros(blocks, dtype=torch.float32, device="cuda")
    best = 0.0
    for _ in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _lib.check(lib().pycwb_valu_probe(out.data_ptr(), blocks, iters, stream.cuda_stream),
                   "pycwb_valu_probe")
        b.record(stream)
        torch.cuda.synchronize()
        best = max(best, blocks * 4 * 64.0 * iters / (a.elapsed_time(b) * 1e-3))
    return best


def refinement_leg(n_clusters, order, K, steps, stream, log):
    """Gradient refinement: 3-detector clusters, grid argmax, then projected
    gradient ascent of the continuous statistic (forward + angle gradient per
    iteration), plus one amplitude-gradient pass."""
    import torch

    from pycwb_amd import synth
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch_smooth
    from pycwb_amd.skygrad import SkyGradient, angular_distance, refine_sky

    dev = torch.device("cuda", 0)
    net = synth.make_network(synth.NETWORK_3, order=order, K=K)
    sizes = synth.cluster_sizes(n_clusters, seed=1003)
    ids = np.zeros(n_clusters, dtype=np.int64)
    b = synth_batch_smooth([net], sizes, ids, seed=1003 * 100000, snr=6.0, tap_sigma=2.0,
                           device=dev)
    nt = net_tables([net], device=dev)
    valid = torch.arange(net.n_sky, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    log(f"refinement workload ready: {n_clusters} clusters, {int(sizes.sum())} pixels")
    bs = BatchedSkyScan(b, nt, valid, 0.5 * 0.5 * 2 / 3, 0.5, 0.5, block=64)
    bs.launch(stream)
    r = bs.results()
    log("refinement: grid argmax done")
    eng = SkyGradient(b, nt, [net], r["REG1"], bs.desc.reg0)
    th0 = torch.from_numpy(np.pi / 2 - net.lat[r["l_max"]]).to(dev)
    ph0 = torch.from_numpy(net.phi[r["l_max"]]).to(dev)
    refine_sky(eng, th0, ph0, steps=1)  # warm
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    out = refine_sky(eng, th0, ph0, steps=steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    # kernel alone: forward + angle gradient, and with the amplitude gradient
    k = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = 10
    k[0].record(stream)
    for _ in range(reps):
        eng.run(out["theta"], out["phi"])
    k[1].record(stream)
    for _ in range(reps):
        eng.run(out["theta"], out["phi"], want_td=True)
    k[2].record(stream)
    torch.cuda.synchronize()
    fwd_ms = k[0].elapsed_time(k[1]) / reps
    td_ms = k[1].elapsed_time(k[2]) / reps
    v0, v = out["value0"].cpu().numpy(), out["value"].cpu().numpy()
    grid = r["sky_stat_max"]
    lm = lm_criterion(b, nt, [net], bs, r, th0, ph0, out)
    th_t, ph_t = b.truth[:, 0], b.truth[:, 1]
    err0 = angular_distance(th0.cpu().numpy(), ph0.cpu().numpy(), th_t, ph_t)
    err1 = angular_distance(out["theta"].cpu().numpy(), out["phi"].cpu().numpy(), th_t, ph_t)
    res = {"clusters": n_clusters, "healpix_order": order, "K": K, "steps": steps,
           "pixels": int(sizes.sum()), "ms_total": round(ms, 3),
           "ms_per_iteration": round(ms / (steps + 1), 3),
           "kernel_ms_angles": round(fwd_ms, 4), "kernel_ms_with_amplitudes": round(td_ms, 4),
           "frac_improved": float(np.mean(v > v0)),
           "frac_above_grid_statistic": float(np.mean(v > grid)),
           "median_gain": float(np.median((v - v0) / np.maximum(np.abs(v0), 1e-12))),
           "median_error_deg_grid": float(np.degrees(np.median(err0))),
           "median_error_deg_refined": float(np.degrees(np.median(err1))),
           "frac_closer_to_truth": float(np.mean(err1 < err0)),
           "Lm": lm,
           "data": "tap-correlated noise (sigma 2 taps), off-grid injections, SNR 6",
           "dtype": "f64 (angles), f32 output (amplitudes)"}
    log(f"refinement: {res}")
    bs.close()
    return res


def lm_criterion(batch, nt, nets_host, bs, r, th0, ph0, out):
    """The refinement's success criterion on Lm = Em - Np - Gn of the best-sky
    statistics: at the refined direction (continuous delays, interpolated
    amplitudes) against the grid argmax, both as the pipeline evaluates it
    (integer delay taps) and with the same interpolation at the grid
    direction."""
    import torch

    from pycwb_amd.skystats import SkyStatistics, statistics_at_directions

    st = SkyStatistics(batch, nt)
    st.run(bs.l_max, bs.REG1, bs.desc.reg0)
    lm_grid = st.scalars()["Lm"]
    reg1 = r["REG1"]
    lm_ref = statistics_at_directions(batch, nets_host, out["theta"].cpu().numpy(),
                                      out["phi"].cpu().numpy(), reg1, bs.desc.reg0,
                                      nt.En)["Lm"]
    lm_gi = statistics_at_directions(batch, nets_host, th0.cpu().numpy(), ph0.cpu().numpy(),
                                     reg1, bs.desc.reg0, nt.En)["Lm"]
    torch.cuda.synchronize()
    return {"frac_refined_above_grid_Lm": float(np.mean(lm_ref > lm_grid)),
            "frac_refined_above_grid_interpolated_Lm": float(np.mean(lm_ref > lm_gi)),
            "median_Lm_grid": float(np.median(lm_grid)),
            "median_Lm_refined": float(np.median(lm_ref)),
            "note": "Lm = Em - Np - Gn of the best-sky statistics; grid = integer taps at "
                    "the grid argmax; interpolated = the refinement's linear delay "
                    "interpolation at the grid direction"}


def wdm_leg(stream, log, seconds=600, rate=4096, n_det=5):
    """WDM t2w of n_det series of seconds*rate samples at 7 resolutions
    (M = 8 .. 512), HIP events per level; algorithmic bytes = input read once
    plus the complex128 map written once.  Then the inverse w2t of the same
    maps (the same bytes, read and written the other way) and its round-trip
    error away from the edges."""
    import torch

    from pycwb_amd.wdm import WDM

    dev = torch.device("cuda", 0)
    L = seconds * rate
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    x = torch.randn(n_det, L, generator=g, device=dev, dtype=torch.float64)
    levels = []
    tot_ms, tot_ims, tot_bytes = 0.0, 0.0, 0
    for M in (8, 16, 32, 64, 128, 256, 512):
        w = WDM(M, M, 6, 10)
        w.t2w_device(x)  # warm (filter upload)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20   # back-to-back calls: the mean is the kernel's time plus small launch gaps
        e0.record(stream)
        for _ in range(reps):
            y = w.t2w_device(x, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        N = y.shape[-1]
        nbytes = n_det * (L * 8 + N * (M + 1) * 16)
        # inverse (00 phase) of the same maps
        z = w.w2t_device(y, L, stream=stream)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            z = w.w2t_device(y, L, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ims = e0.elapsed_time(e1) / reps
        err = float((z - x)[:, w.m_H:-w.m_H].abs().max() / x.abs().max())
        levels.append({"M": M, "m_H": int(w.m_H), "ms": round(ms, 4),
                       "gbs": nbytes / (ms * 1e-3) / 1e9, "w2t_ms": round(ims, 4),
                       "w2t_gbs": nbytes / (ims * 1e-3) / 1e9, "roundtrip_rel_err": err})
        tot_ms += ms
        tot_ims += ims
        tot_bytes += nbytes
        del y, z
    res = {"series": n_det, "samples": L, "levels": levels, "ms_total": round(tot_ms, 3),
           "gbs": tot_bytes / (tot_ms * 1e-3) / 1e9,
           "frac_hbm": tot_bytes / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "w2t_ms_total": round(tot_ims, 3),
           "w2t_gbs": tot_bytes / (tot_ims * 1e-3) / 1e9}
    log(f"wdm: {res}")
    return res


def tdamp_leg(stream, log, M=64, n_pixels=200000, K=64, nc=12):
    """TD-amplitude extraction: one detector's WDM map of 600 s at 4096 Hz
    (M = 64: 38,400 time bins x 65 layers), n_pixels pixels, delays -K..K,
    TD filters with nc coefficients per side.  Work per (pixel, delay): two
    phases x three layer sums x (2 nc + 1) multiply-adds in float64."""
    import torch

    from pycwb_amd.tdamp import TDBatchInputs
    from pycwb_amd.wdm import td_filters

    dev = torch.device("cuda", 0)
    n_time = 600 * 4096 // M
    rng = np.random.default_rng(11)
    data = rng.normal(size=(M + 1, n_time)) + 1j * rng.normal(size=(M + 1, n_time))
    inp = TDBatchInputs.from_map(data, td_filters(M, M, 6, nc, 1), dev)
    pix = rng.choice(n_time * (M + 1), size=n_pixels, replace=False).astype(np.int32)
    out = inp.extract_td_vecs(pix, K)   # warm
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record(stream)
    for _ in range(reps):
        out = inp.extract_td_vecs(pix, K, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    amps = n_pixels * (2 * K + 1)
    flops = amps * 2 * 3 * (2 * nc + 1) * 2
    res = {"M": M, "pixels": n_pixels, "K": K, "n_coeffs": nc, "ms": round(ms, 4),
           "amplitudes_per_s": amps / (ms * 1e-3), "tflops_f64": flops / (ms * 1e-3) / 1e12,
           "frac_f64_peak": flops / (ms * 1e-3) / 1e12 / 78.6,
           "out_gbs": out.numel() * 4 / (ms * 1e-3) / 1e9}
    log(f"tdamp: {res}")
    return res


def subnet_leg(stream, log, n_clusters=4096, loud=300, order=4, K=64, cpu=True):
    """Sub-network cut of the superclusters of a lag: the LOUD = 300 loudest
    pixels of 4,096 clusters (sizes log-uniform[16, 4096], capped at 300), half
    H1L1V1 and half H1L1V1K1, K = 64, HEALPix order 4 (3,072 directions),
    Acore = sqrt(2), subcut 0.33, subnet 0.7.  Times the sky loop (one launch
    per detector count) with HIP events, then the MRA statistics of the
    clusters with suball > subnet (self-only cross-talk, CSR upload included).
    The CPU baseline is the oracle's single-threaded sky loop (the reference's
    loop is serial per cluster) on the first 32 clusters, whose GPU results must
    match it bit for bit."""
    import torch

    from pycwb_amd import synth
    from pycwb_amd.batch import synth_batch
    from pycwb_amd.skystats import self_xtalk
    from pycwb_amd.xtalk import self_xtalk_batch
    from pycwb_amd.super_cluster_native.sub_net_cut import (SubnetBatch,
                                                            calculate_e2or_from_acore, needs_mra,
                                                            sky_tuple, subnet_sky_tables)

    dev = torch.device("cuda", 0)
    nets = [synth.make_network(synth.NETWORK_3, order=order, K=K),
            synth.make_network(synth.NETWORK_4, order=order, K=K)]
    sizes = np.minimum(synth.cluster_sizes(n_clusters, seed=2024), loud)
    ids = np.arange(n_clusters) % 2
    batch = synth_batch(nets, sizes, ids, seed=2024 * 1000, device=dev)
    acore, subcut, subnet = math.sqrt(2.0), 0.33, 0.7
    En = [np.float32(2 * acore * acore * nt.n_ifo) for nt in nets]
    e2or = [calculate_e2or_from_acore(acore, nt.n_ifo) for nt in nets]
    Es = [np.float32(2 * e) for e in e2or]
    tables = subnet_sky_tables([(nt.FP_t, nt.FX_t, nt.ml) for nt in nets], En, Es, dev)
    sb = SubnetBatch(batch, tables, subcut, dev)
    S = nets[0].n_sky
    evals = float(np.sum(sizes.astype(np.float64))) * S
    keep = sb.launch_sky(stream)
    torch.cuda.synchronize()
    reps = 3
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record(stream)
    for _ in range(reps):
        keep = sb.launch_sky(stream)
    e[1].record(stream)
    torch.cuda.synchronize()
    del keep
    sky_ms = e[0].elapsed_time(e[1]) / reps
    sky = sb.sky_d.view(-1, 8)[:n_clusters].cpu().numpy()
    need = [c for c in range(n_clusters) if needs_mra(sky_tuple(sky[c]), subnet)]
    csrs = [self_xtalk(int(sizes[c])) for c in need]
    t0 = time.perf_counter()
    h = sb.prepare_mra(need, csrs)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    e[2].record(stream)
    sb.launch_mra(h, stream)
    e[3].record(stream)
    torch.cuda.synchronize()
    mra_ms = e[2].elapsed_time(e[3])
    mra_host = sb.mra_d.view(-1, 8)[need].cpu().numpy()
    del h
    # the same from the batch's device CSRs (as the cross-talk builder returns them)
    sb.mra_d.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dcsr = self_xtalk_batch(sizes, dev)
    h = sb.prepare_mra(need, dcsr)
    sb.launch_mra(h, stream)
    torch.cuda.synchronize()
    dev_s = time.perf_counter() - t0
    mra_dev_equal = bool(np.array_equal(sb.mra_d.view(-1, 8)[need].cpu().numpy(), mra_host))
    del h, dcsr
    nd = 2 * K + 1
    td_bytes = float(np.sum(sizes.astype(np.float64) * batch.nifo * nd * 8))
    res = {"clusters": n_clusters, "loud": loud, "healpix_order": order, "n_sky": S,
           "pixels": int(sizes.sum()), "evals": evals, "sky_ms": round(sky_ms, 3),
           "evals_per_s": evals / (sky_ms * 1e-3), "mra_clusters": len(need),
           "mra_ms": round(mra_ms, 3), "mra_host_prep_ms": round(prep_s * 1e3, 3),
           "mra_device_csr_wall_ms": round(dev_s * 1e3, 3),
           "mra_device_csr_equal": mra_dev_equal,
           "td_bytes": td_bytes, "td_gbs": td_bytes / (sky_ms * 1e-3) / 1e9,
           "frac_hbm": td_bytes / (sky_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
    log(f"subnet: sky loop {sky_ms:.2f} ms ({evals / (sky_ms * 1e-3):.3e} evals/s), "
        f"MRA of {len(need)} clusters {mra_ms:.2f} ms (host CSR preparation "
        f"{prep_s * 1e3:.1f} ms; from device CSRs {dev_s * 1e3:.1f} ms wall, equal "
        f"{mra_dev_equal})")
    if cpu:
        from oracle import oracle as ora

        sample = list(range(32))
        cpu_s = 0.0
        outs = []
        for c in sample:
            k, P = int(ids[c]), int(sizes[c])
            td = batch.cluster_td(c).cpu().numpy()
            rms = batch.cluster_rms(c).cpu().numpy()
            td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
            td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
            t1 = time.perf_counter()
            outs.append(ora.sky_loc_from_td(nets[k].n_ifo, P, S, nets[k].FP_t, nets[k].FX_t, rms,
                                            td00, td90, nets[k].ml, En[k], e2or[k], subcut))
            cpu_s += time.perf_counter() - t1
        ev = float(np.sum(sizes[sample].astype(np.float64))) * S
        res["cpu_baseline"] = {"value": ev / cpu_s, "unit": "evals/s", "cores": 1,
                               "kind": "port",
                               "sample": f"sky loop of clusters 0-31 ({int(ev):d} evals), "
                                         "oracle, one thread"}
        res["parity_sample_bit_exact"] = all(sky_tuple(sky[c]) == o for c, o in zip(sample, outs))
        log(f"subnet cpu baseline {ev / cpu_s:.3e} evals/s (1 thread); parity "
            f"{res['parity_sample_bit_exact']}")
    return res


def xtalk_leg(stream, log, n_clusters=4096, cpu=True):
    """Cross-talk CSRs of 4,096 clusters (sizes log-uniform[16, 4096], pixels in
    compact time-frequency blocks over 7 resolutions, 8..512 layers) from a
    synthetic catalog with the real one's bucket structure.  Times
    XTalk.build_batch end to end (uploads, hash, catalog walk, scan, host read
    of the total, ordered fill); the CPU baseline is the oracle's restatement
    of getXTalk_pixels_fast (one thread) on the first 64 clusters, whose CSRs
    must match the GPU's bit for bit."""
    import torch

    from pycwb_amd import synth
    from pycwb_amd.xtalk import synthetic_catalog

    layers = [8, 16, 32, 64, 128, 256, 512]
    cat = synthetic_catalog(layers, seed=7)
    rng = np.random.default_rng(7)
    sizes = synth.cluster_sizes(n_clusters, seed=7)
    clusters = []
    for P in sizes:
        r = rng.integers(0, len(layers), size=P)
        L = np.asarray(layers)[r]
        t0 = int(rng.integers(100, 10000))
        band = float(rng.uniform(0.1, 0.8))
        ts = t0 * (512 // L) + rng.integers(0, 4 * (512 // L) + 1)
        fs = np.minimum(L, (band * L).astype(np.int64) + rng.integers(0, L // 8 + 2))
        clusters.append(((L + 1).astype(np.int32), (ts * (L + 1) + fs).astype(np.int64)))
    cat.build_batch(clusters[:8], stream)             # warm (catalog upload)
    torch.cuda.synchronize()
    # the batch's pixels laid end to end, as a clustering stage hands them over
    flat_l = np.concatenate([c[0] for c in clusters])
    flat_t = np.concatenate([c[1] for c in clusters])
    t = time.perf_counter()
    csr = cat.build_batch_flat(flat_l, flat_t, sizes, stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    csr_list = cat.build_batch(clusters, stream)
    torch.cuda.synchronize()
    ms_list = (time.perf_counter() - t) * 1e3
    same = bool(np.array_equal(csr.xt_off, csr_list.xt_off) and
                torch.equal(csr.lookup, csr_list.lookup) and torch.equal(csr.xt, csr_list.xt))
    ld, td = torch.from_numpy(flat_l).to(stream.device), torch.from_numpy(flat_t).to(stream.device)
    torch.cuda.synchronize()
    t = time.perf_counter()
    cat.build_batch_flat(ld, td, sizes, stream)
    torch.cuda.synchronize()
    ms_dev = (time.perf_counter() - t) * 1e3
    n_pix = int(sizes.sum())
    res = {"clusters": n_clusters, "pixels": n_pix, "entries": int(csr.xt_off[-1]),
           "ms": round(ms, 3), "pixels_per_s": n_pix / (ms * 1e-3),
           "ms_device_inputs": round(ms_dev, 3), "ms_cluster_list": round(ms_list, 3),
           "list_equals_flat": same,
           "note": "ms: build_batch_flat from host arrays (uploads, hash, catalog walk, scan, "
                   "host read of the total, ordered fill); ms_device_inputs: the same from "
                   "device tensors; ms_cluster_list: build_batch on a list of 4096 per-cluster "
                   "arrays"}
    log(f"xtalk: CSR of {n_pix} pixels ({res['entries']} entries) in {ms:.1f} ms "
        f"(device inputs {ms_dev:.1f} ms, cluster list {ms_list:.1f} ms, equal {same})")
    if cpu:
        from oracle import oracle as ora

        sample = list(range(64))
        t = time.perf_counter()
        outs = [ora.xtalk_pixels_fast(clusters[c][0], clusters[c][1], cat.layers,
                                      cat.lookup_table, cat.coeff) for c in sample]
        cpu_s = time.perf_counter() - t
        px = int(sizes[sample].sum())
        ok = True
        for c, (lk, xt) in zip(sample, outs):
            g_lk, g_xt = csr.cluster(c)
            ok &= bool(np.array_equal(g_lk, lk) and np.array_equal(g_xt, xt))
        res["cpu_baseline"] = {"value": px / cpu_s, "unit": "pixels/s", "cores": 1,
                               "kind": "port", "sample": f"clusters 0-63 ({px} pixels), oracle"}
        res["parity_sample_bit_exact"] = ok
        log(f"xtalk cpu baseline {px / cpu_s:.3e} pixels/s (1 thread); parity {ok}")
    return res


def _select_workload(M=64, seconds=600, rate=4096, n_det=5, n_lag=128):
    """5 detectors' 600 s energy maps at M = 64 (65 x 38,400 pixels): unit-mean
    exponential noise plus 200 coherent blobs, 128 lags, Eo = 2.3 n."""
    import torch

    dev = torch.device("cuda", 0)
    n_freq, n_time = M + 1, seconds * rate // M
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    E = -torch.log(torch.rand((n_det, n_freq, n_time), generator=g, device=dev,
                              dtype=torch.float64))
    rng = np.random.default_rng(3)
    for _ in range(200):
        f0, t0 = int(rng.integers(1, n_freq - 4)), int(rng.integers(0, n_time - 8))
        E[:, f0:f0 + 3, t0:t0 + 6] += float(rng.uniform(2, 10))
    edge = 2 * rate // M
    cache = {"arrays_stack": E.contiguous(), "n_ifo": n_det, "n_freq": n_freq,
             "n_time": n_time, "dt": M / rate, "rate": rate / M, "edge_bins": edge,
             "valid_start": edge, "valid_stop": n_time - edge, "nn_valid": n_time - 2 * edge,
             "ib": 1, "ie": n_freq - 1, "start": 0.0, "stop": float(seconds), "f_low": 0.0,
             "f_high": rate / 2.0, "shift_bins_by_lag": None}
    shifts = np.zeros((n_lag, n_det), np.int64)
    shifts[:, 1:] = (np.arange(n_lag)[:, None] * np.arange(1, n_det)[None, :] * 7) % \
        (n_time - 2 * edge)
    return cache, shifts, 2.3 * n_det, edge


def select_leg(stream, log, M=64, seconds=600, rate=4096, n_det=5, n_lag=128, cpu=True):
    """Network-pixel selection of one resolution (M = 64: 65 x 38,400 pixels) of
    5 detectors' 600 s energy maps for 128 lags in one pass (count, host read
    of the total, emit; the sparse payload comes back, not the dense masks).
    Energy maps: unit-mean exponential noise plus 200 coherent blobs; threshold
    Eo = 2.3 n (about 1 % of the summed pixels above it).  CPU baseline: the
    oracle's single-lag loop (one thread) on 4 lags, whose payloads must match
    the GPU's bit for bit."""
    import torch

    from pycwb_amd.coherence_native.selection import select_network_pixels_lags

    cache, shifts, eo, edge = _select_workload(M, seconds, rate, n_det, n_lag)
    n_freq, n_time = cache["n_freq"], cache["n_time"]
    select_network_pixels_lags(cache, eo, shifts[:2], stream=stream)     # warm
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = select_network_pixels_lags(cache, eo, shifts, stream=stream, with_mask=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3
    pix = n_freq * n_time * n_lag
    sel = int(sum(o["time"].size for o in out))
    from pycwb_amd.coherence_native.selection import select_network_pixels_device

    t = time.perf_counter()
    dsel = select_network_pixels_device(cache, eo, shifts, stream=stream)
    stream.synchronize()
    ms_dev = (time.perf_counter() - t) * 1e3
    del dsel
    res = {"M": M, "detectors": n_det, "lags": n_lag, "pixels_per_lag": n_freq * n_time,
           "selected": sel, "ms": round(ms, 3), "lag_pixels_per_s": pix / (ms * 1e-3),
           "ms_device_payload": round(ms_dev, 3),
           "note": "ms: select_network_pixels_lags with the payload brought to the host; "
                   "ms_device_payload: the same pass with the payload left on the device"}
    log(f"select: {n_lag} lags x {n_freq * n_time} pixels in {ms:.1f} ms ({ms_dev:.1f} ms with "
        f"the payload on the device), {sel} selected")
    if cpu:
        from oracle import oracle as ora

        Eh = cache["arrays_stack"].cpu().numpy()
        t = time.perf_counter()
        outs = [ora.select_pixels(Eh, shifts[l], None, edge, 1, n_freq - 1, eo)
                for l in range(4)]
        cpu_s = time.perf_counter() - t
        ok = all(np.array_equal(outs[l]["time"], out[l]["time"]) and
                 np.array_equal(outs[l]["frequency"], out[l]["frequency"]) and
                 np.array_equal(outs[l]["energy"], out[l]["energy"]) and
                 np.array_equal(outs[l]["pix_det_index"], out[l]["pix_det_index"])
                 for l in range(4))
        res["cpu_baseline"] = {"value": 4 * n_freq * n_time / cpu_s, "unit": "lag-pixels/s",
                               "cores": 1, "kind": "port", "sample": "lags 0-3, oracle"}
        res["parity_sample_bit_exact"] = bool(ok)
        log(f"select cpu baseline {4 * n_freq * n_time / cpu_s:.3e} lag-pixels/s; parity {ok}")
    return res


def cluster_leg(stream, log, M=64, seconds=600, rate=4096, n_det=5, n_lag=128, cpu=True,
                reps=5):
    """Connected-component clustering (kf = kt = 1) of the pixels the selection
    leg's workload selects, all 128 lags in one stream-ordered call (index map
    of 128 x 65 x 38,400 cells, initialised once and restored by each call;
    union-find, root scan, member sort, per-cluster sums and subnet/subrho),
    timed with HIP events over ``reps`` calls on a selection already resident
    on the device.  CPU baseline: the oracle's
    cluster_pixels (one thread) on lags 0-3, whose clusters must match the
    GPU's bit for bit."""
    import ctypes as C

    import torch

    from pycwb_amd import _lib
    from pycwb_amd.coherence_native.clustering import cluster_pixels_device, _n_sub
    from pycwb_amd.coherence_native.selection import select_network_pixels_device
    from pycwb_amd.device import lib, ptr, stream_ptr

    cache, shifts, eo, edge = _select_workload(M, seconds, rate, n_det, n_lag)
    n_freq, n_time = cache["n_freq"], cache["n_time"]
    sel = select_network_pixels_device(cache, eo, shifts, stream=stream)
    stream.synchronize()
    out = cluster_pixels_device(sel, kt=1, kf=1, rate=cache["rate"], stream=stream)  # warm
    n, nc = int(sel["n_pix"]), out["n_clusters"]
    # the same call through the C ABI, timed on its stream
    d = _lib.ClusterDesc()
    d.n_lag, d.lag_off, d.n_pix = n_lag, ptr(sel["lag_off_dev"]), n
    d.freq, d.time, d.energy = ptr(sel["frequency"]), ptr(sel["time"]), ptr(sel["energy"])
    d.det_energy, d.noise_rms, d.n_ifo = ptr(sel["pix_det_energy"]), None, n_det
    d.n_freq, d.n_time, d.kf, d.kt = n_freq, n_time, 1, 1
    d.dt, d.half_rate, d.n_sub = 1.0 / cache["rate"], cache["rate"] / 2, _n_sub(n_det)
    d.work, d.work_bytes = ptr(out["work"]), out["work"].numel()
    d.index_map = out["index_map"]
    d.label, d.lag_cluster_off = ptr(out["label"]), ptr(out["lag_cluster_off_dev"])
    d.cluster_off, d.members = ptr(out["cluster_off"]), ptr(out["members"])
    d.c_energy, d.c_time, d.c_freq = ptr(out["energy"]), ptr(out["c_time"]), ptr(out["c_freq"])
    d.c_subnet, d.c_subrho = ptr(out["subnet"]), ptr(out["subrho"])
    st = stream_ptr(stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        _lib.check(lib().pycwb_cluster_pixels(C.byref(d), st), "pycwb_cluster_pixels")
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    t = time.perf_counter()
    cluster_pixels_device(sel, kt=1, kf=1, rate=cache["rate"], stream=stream)
    wall = (time.perf_counter() - t) * 1e3
    sizes = np.diff(out["cluster_off"][:nc + 1].cpu().numpy())
    res = {"M": M, "detectors": n_det, "lags": n_lag, "pixels": n, "clusters": nc,
           "largest_cluster": int(sizes.max()) if nc else 0, "ms": round(ms, 3),
           "wall_ms": round(wall, 3), "pixels_per_s": n / (ms * 1e-3),
           "map_cells": n_lag * n_freq * n_time}
    log(f"cluster: {n} pixels of {n_lag} lags -> {nc} clusters in {ms:.2f} ms "
        f"(call with allocation and sync {wall:.1f} ms)")
    if cpu:
        from oracle import oracle as ora

        lo, lco = out["lag_off"], out["lag_cluster_off"]
        host = {k: sel[k].cpu().numpy() for k in ("frequency", "time", "energy",
                                                  "pix_det_energy")}
        lab = out["label"].cpu().numpy()
        mem = out["members"].cpu().numpy()
        off = out["cluster_off"].cpu().numpy()
        stats = {k: out[k].cpu().numpy() for k in ("energy", "c_time", "c_freq", "subnet",
                                                   "subrho")}
        ok, n_cpu = True, 0
        t = time.perf_counter()
        runs = [ora.cluster_run(host["frequency"][lo[l]:lo[l + 1]],
                                host["time"][lo[l]:lo[l + 1]], host["energy"][lo[l]:lo[l + 1]],
                                host["pix_det_energy"][lo[l]:lo[l + 1]], n_freq, n_time, 1, 1,
                                cache["rate"], _n_sub(n_det)) for l in range(4)]
        cpu_s = time.perf_counter() - t
        for l, o in enumerate(runs):
            a, b, ca, cb = lo[l], lo[l + 1], lco[l], lco[l + 1]
            n_cpu += b - a
            ok = ok and np.array_equal(lab[a:b], o["labels"]) and \
                np.array_equal(mem[off[ca]:off[cb]] - a, o["members"]) and \
                np.array_equal(off[ca:cb + 1] - off[ca], o["cluster_off"]) and \
                all(np.array_equal(stats[k][ca:cb], o[k], equal_nan=True)
                    for k in ("energy", "c_time", "c_freq", "subnet", "subrho"))
        res["cpu_baseline"] = {"value": n_cpu / cpu_s, "unit": "pixels/s", "cores": 1,
                               "kind": "port", "sample": "lags 0-3, oracle"}
        res["parity_sample_bit_exact"] = bool(ok)
        log(f"cluster cpu baseline {n_cpu / cpu_s:.3e} pixels/s; parity {ok}")
    return res


def threshold_leg(stream, log, M=64, seconds=600, rate=4096, n_det=5, cpu=True, reps=5):
    """Array work of the pixel-energy threshold on the selection workload's
    maps (5 detectors, 65 x 38,400 pixels, 2 s edges): the no-shape path
    (edge-cropped detector sum, clip, counts, radix sort of 2.48e6 values) and
    the shape path (time-major region, clip, counts, compaction of the values
    above 1e-3, numpy-order sums of them and of their logs), each one
    pycwb_threshold_stats call timed with HIP events over ``reps`` calls.
    CPU baseline: the oracle's restatement of both (one thread) on the same
    maps; counts, order statistics and sums must equal the GPU's."""
    import ctypes as C

    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import lib, ptr, stream_ptr

    cache, _, _, edge = _select_workload(M, seconds, rate, n_det, 1)
    E = cache["arrays_stack"]
    n_freq, n_time = cache["n_freq"], cache["n_time"]
    dev = E.device
    st = stream_ptr(stream)
    res = {"M": M, "detectors": n_det, "pixels": n_freq * n_time}
    outs = {}
    for name, tm, begin, count in (
            ("no_shape", 0, edge, n_freq * (n_time - 2 * edge)),
            ("shape", 1, edge * n_freq, n_freq * n_time - 2 * edge * n_freq - 1)):
        d = _lib.ThresholdDesc()
        d.maps, d.n_ifo, d.n_freq, d.n_time = ptr(E), n_det, n_freq, n_time
        d.time_major, d.begin, d.count, d.hi = tm, begin, count, n_det * 100.0
        vals = torch.empty(count, dtype=torch.float64, device=dev)
        counts = torch.empty(2, dtype=torch.int64, device=dev)
        srt = torch.empty(count, dtype=torch.float64, device=dev) if tm == 0 else None
        sums = torch.empty(2 * ((count + 8191) // 8192), dtype=torch.float64, device=dev) \
            if tm == 1 else None
        d.values, d.counts, d.sorted, d.chunk_sums = ptr(vals), ptr(counts), ptr(srt), ptr(sums)
        nb = C.c_size_t()
        _lib.check(lib().pycwb_threshold_work_bytes(C.byref(d), C.byref(nb)), "work bytes")
        work = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
        d.work, d.work_bytes = ptr(work), int(nb.value)
        _lib.check(lib().pycwb_threshold_stats(C.byref(d), st), "pycwb_threshold_stats")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            _lib.check(lib().pycwb_threshold_stats(C.byref(d), st), "pycwb_threshold_stats")
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        c = counts.cpu().numpy()
        outs[name] = (tm, begin, count, c, srt, sums)
        res[name] = {"values": count, "ms": round(ms, 4), "values_per_s": count / (ms * 1e-3),
                     "above_1e-3": int(c[1]),
                     "read_gbs": n_det * count * 8 / (ms * 1e-3) / 1e9}
        log(f"threshold {name}: {count} values in {ms:.3f} ms")
    if cpu:
        from oracle import oracle as ora

        Eh = E.cpu().numpy()
        ok, cpu_s, n_cpu = True, 0.0, 0
        for name, (tm, begin, count, c, srt, sums) in outs.items():
            t = time.perf_counter()
            o = ora.threshold_stats(Eh, tm, begin, count, n_det * 100.0)
            cpu_s += time.perf_counter() - t
            n_cpu += count
            ok = ok and list(o["counts"]) == [int(c[0]), int(c[1])]
            if srt is not None:
                ok = ok and np.array_equal(srt.cpu().numpy(), o["sorted"])
            if sums is not None:
                ch = sums.cpu().numpy()
                s = ls = 0.0
                for k in range((int(c[1]) + 8191) // 8192):
                    s += float(ch[2 * k])
                    ls += float(ch[2 * k + 1])
                ok = ok and s == o["sum"] and abs(ls - o["log_sum"]) <= 1e-13 * abs(o["log_sum"])
        res["cpu_baseline"] = {"value": n_cpu / cpu_s, "unit": "values/s", "cores": 1,
                               "kind": "port", "sample": "both paths, whole maps, oracle"}
        res["parity_bit_exact"] = bool(ok)
        log(f"threshold cpu baseline {n_cpu / cpu_s:.3e} values/s; parity {ok}")
    return res


def e2e_leg(stream, log, seconds=600, rate=4096, n_inj=200, order=8, cpu=True,
            refine_steps=20, cpu_rates=None):
    """The end-to-end case: 5 detectors (H1 L1 V1 K1 I1), 600 s of coloured
    Gaussian noise at 4096 Hz whitened to unit variance, with 200 injections
    (sine-Gaussians and white-noise bursts, f0 50-1000 Hz, network SNR 8-40),
    7 WDM levels (M = 8 .. 512), then one likelihood batch at HEALPix order 8
    (786,432 directions) and the gradient refinement of every cluster
    (pycwb_amd.segment).  Reports the wall time per stage, the roofline
    fraction of the stages with an algorithmic count (t2w: HBM bytes; TD
    amplitudes: float64 flops; likelihood: the scan's required flops), the
    likelihood batch against the plain one-cluster-per-launch GPU path on the
    same clusters and against the host CPU path (the oracle, 16 threads, on two
    clusters of the batch; ``cpu_rates`` carries the other stages' measured
    single-thread oracle rates), and the injections recovered."""
    import torch

    from pycwb_amd import segment, synth
    from pycwb_amd.batch import PlainLoop
    from pycwb_amd.skygrad import SkyGradient, refine_sky

    dev = torch.device("cuda", 0)
    ifos = synth.NETWORK_5
    n = len(ifos)
    t0 = time.perf_counter()
    x, inj = segment.synthetic_segment(ifos, seconds, rate, n_inj, seed=11)
    xd = torch.from_numpy(x).to(dev)
    torch.cuda.synchronize()
    data_s = time.perf_counter() - t0
    seg = segment.run(xd, ifos, rate, stream, order=order, log=log)
    stages = dict(seg["stage_s"])
    stages["data"] = data_s
    net, nt, batch, scan, res = seg["net"], seg["nets"], seg["batch"], seg["scan"], \
        seg["results"]
    sizes, K, td_rate = seg["sizes"], seg["K"], seg["td_rate"]
    C_ = sizes.size
    S = net.n_sky
    evals = int(sizes.sum()) * S
    # per-stage kernel times of the likelihood batch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    scan.launch_stage(0, stream)
    ev[0].record(stream)
    scan.launch_stage(1, stream)
    ev[1].record(stream)
    scan.launch_stage(2, stream)
    ev[2].record(stream)
    scan.launch_stage(3, stream)
    torch.cuda.synchronize()
    reg_ms, scan_ms = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    active = scan.active_pairs()
    tested = scan.tested_pairs()
    # the plain GPU path on the same clusters
    pl = PlainLoop(batch, nt, seg["valid"], 0.5 * 0.5 * 2 / 3, 0.5, 0.5)
    pl.launch(stream)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    pl.launch(stream)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - tp
    rp = pl.results()
    plain_same = bool(np.array_equal(rp["l_max"], res["l_max"]))
    # gradient refinement of every cluster from its grid argmax
    eng = SkyGradient(batch, nt, [net], res["REG1"], scan.desc.reg0)
    th0 = torch.from_numpy(np.pi / 2 - net.lat[res["l_max"]]).to(dev)
    ph0 = torch.from_numpy(net.phi[res["l_max"]]).to(dev)
    torch.cuda.synchronize()
    tr = time.perf_counter()
    ref = refine_sky(eng, th0, ph0, steps=refine_steps)
    torch.cuda.synchronize()
    stages["refinement"] = time.perf_counter() - tr
    v0, v1 = ref["value0"].cpu().numpy(), ref["value"].cpu().numpy()
    lm = lm_criterion(batch, nt, [net], scan, res, th0, ph0, ref)
    # parity sample and the CPU path on it
    parity, cpu_rate = None, None
    if cpu:
        from oracle import oracle as ora

        picks = [int(np.argmin(np.abs(sizes - 10))), int(np.argmin(np.abs(sizes - 40)))]
        vh = np.arange(S, dtype=np.int64)
        En = nt.En[0]
        parity, cpu_evals, cpu_s = [], 0, 0.0
        for c in picks:
            P = int(sizes[c])
            rms_c = batch.cluster_rms(c).cpu().numpy()
            td_c = batch.cluster_td(c).cpu().numpy()                  # (P, n, 2K+1, 2)
            t00 = np.ascontiguousarray(td_c[..., 0].transpose(2, 1, 0))
            t90 = np.ascontiguousarray(td_c[..., 1].transpose(2, 1, 0))
            tc = time.perf_counter()
            reg_o = ora.calculate_dpf(net.FP_t, net.FX_t, rms_c, S, n, 0.5 * 0.5 * 2 / 3, En,
                                      vh, threads=16)
            REG = np.array([np.float32(0.5 * np.sqrt(2)), res["REG1"][c], 0.0], np.float32)
            l_o, _, _ = ora.scan_sky(n, P, S, net.FP_t, net.FX_t, rms_c, t00, t90, net.ml, REG,
                                     0.5, 0.5, En, vh, threads=16)
            cpu_s += time.perf_counter() - tc
            cpu_evals += P * S
            parity.append({"cluster": c, "pixels": P,
                           "regulator_bit_exact": bool(reg_o == res["regulator"][c]),
                           "l_max_bit_exact": bool(int(l_o) == int(res["l_max"][c]))})
        cpu_rate = cpu_evals / cpu_s
        log(f"e2e parity sample: {parity}")
    scan.close()
    found = segment.recovered(inj, seg["ctime"], res["l_max"], net, ifos, td_rate)
    found_sg = segment.recovered([d for d in inj if d["kind"] == "sg"], seg["ctime"],
                                 res["l_max"], net, ifos, td_rate)
    # roofline of the stages with an algorithmic count
    levels = seg["levels"]
    L = x.shape[1]
    t2w_bytes = float(sum(n * (8 * L + 16 * (lv["M"] + 1) * lv["n_time"]) for lv in levels))
    td_flops = float(sizes.sum()) * n * (2 * K + 1) * 300.0
    nd_frac = float(active.sum()) / float(evals)
    req = float(tested.sum()) * (4.0 * n + 4.0) + float(active.sum()) * (74.0 * n + 125.0)
    roof = {
        "t2w": {"bound": "hbm", "achieved_gbs": t2w_bytes / stages["t2w"] / 1e9,
                "frac": t2w_bytes / stages["t2w"] / 1e9 / PEAK_HBM_GBS,
                "note": "wall time including launches"},
        "td_amplitudes": {"bound": "fp64", "achieved_tflops": td_flops / stages["td_amplitudes"]
                          / 1e12, "frac": td_flops / stages["td_amplitudes"] / 1e12 / 78.6},
        "likelihood_scan": {"bound": "valu", "achieved_tflops": req / (scan_ms * 1e-3) / 1e12,
                            "frac": req / (scan_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS},
    }
    cpu_path = None
    if cpu_rate:
        cpu_path = {"likelihood_evals_per_s": cpu_rate, "threads": 16,
                    "likelihood_batch_s_estimated": evals / cpu_rate,
                    "gpu_over_cpu": (evals / stages["likelihood_batch"]) / cpu_rate}
        if cpu_rates:
            cpu_path["other_stage_rates_1_thread"] = cpu_rates
    total = sum(v for k, v in stages.items() if k not in ("data", "setup"))
    out = {"detectors": n, "seconds": seconds, "rate": rate, "injections": n_inj,
           "injection_kinds": {"sg": sum(d["kind"] == "sg" for d in inj),
                               "wnb": sum(d["kind"] == "wnb" for d in inj)},
           "noise": "coloured (PSD (40/f)^4 + 1 + (f/300)^2), whitened to unit variance",
           "levels": [{"M": lv["M"], "Eo": round(lv["Eo"], 4), "selected": lv["selected"],
                       "clusters": int(lv["sizes"].size)} for lv in levels],
           "clusters": C_, "pixels": int(sizes.sum()), "healpix_order": order, "K": K,
           "likelihood_evals": evals,
           "likelihood_evals_per_s": evals / stages["likelihood_batch"],
           "plain_gpu": {"s": round(plain_s, 4), "evals_per_s": evals / plain_s,
                         "batched_over_plain": plain_s / stages["likelihood_batch"],
                         "l_max_equal": plain_same},
           "cpu_path": cpu_path,
           "stage_s": {k: round(v, 4) for k, v in stages.items()},
           "likelihood_stage_ms": {"regulator": round(reg_ms, 2), "scan": round(scan_ms, 2)},
           "likelihood_active_fraction": round(nd_frac, 4),
           "roofline": roof,
           "refinement": {"steps": refine_steps, "frac_improved": float(np.mean(v1 > v0)),
                          "Lm": lm},
           "parity_sample": parity,
           "wall_s_without_data_and_setup": round(total, 3),
           "injections_found_with_delays": found,
           "sine_gaussians_found_with_delays": found_sg}
    log(f"e2e: {C_} clusters ({int(sizes.sum())} px) from 7 levels in {total:.2f} s "
        f"(stages {out['stage_s']}); {found}/{n_inj} injections recovered with their delays")
    return out


def maxenergy_leg(stream, log, M=64, seconds=600, rate=4096, max_delay_s=0.042, ds=4,
                  pattern=3, cpu=True):
    """Time-delay max-energy map of one detector's 600 s series at M = 64:
    the series and its copies shifted by 4, 8, ..., 172 samples to both sides
    (87 quadrature t2w, batched 16 at a time), packet energy (pattern 3) and
    the pixel maximum.  CPU baseline: the oracle's loop (direct-DFT t2w, one
    thread) on the first 30 s of the same series, max delay 12 samples, whose
    map the GPU reproduces to 1e-10 of its maximum."""
    import torch

    from pycwb_amd.coherence_native.projection import delay_shifts, max_energy_device
    from pycwb_amd.wdm import WDM

    dev = torch.device("cuda", 0)
    L = seconds * rate
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    x = torch.randn(L, generator=g, device=dev, dtype=torch.float64)
    w = WDM(M, M, 6, 10)
    md = int(rate * max_delay_s)
    n_maps = int(delay_shifts(md, ds, L).size)
    args = (w, md, ds, pattern, 1.0, rate // M, 32.0, 1024.0, rate / (2.0 * M))
    max_energy_device(x, *args, stream=stream)   # warm: filter upload, device buffers cached
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    cur = max_energy_device(x, *args, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    res = {"M": M, "samples": L, "maps": n_maps, "ms": round(ms, 3),
           "samples_per_s": L / (ms * 1e-3), "map_samples_per_s": L * n_maps / (ms * 1e-3)}
    log(f"max-energy: {n_maps} maps of {L} samples in {ms:.1f} ms")
    if cpu:
        from oracle import oracle as ora

        Ls = 30 * rate
        xs = x[:Ls].cpu().numpy()
        sargs = (w, 12, ds, pattern, 1.0, rate // M, 32.0, 1024.0, rate / (2.0 * M))
        g_small = max_energy_device(x[:Ls], *sargs).cpu().numpy()
        t = time.perf_counter()
        o = ora.max_energy(xs, M, w.filter, 12, ds, pattern, 1.0, rate // M, 32.0, 1024.0,
                           rate / (2.0 * M))
        cpu_s = time.perf_counter() - t
        n_small = int(delay_shifts(12, ds, Ls).size)
        ok = bool(np.abs(g_small - o.T).max() <= 1e-10 * np.abs(o).max())
        res["cpu_baseline"] = {"value": Ls * n_small / cpu_s, "unit": "map-samples/s",
                               "cores": 1, "kind": "port",
                               "sample": f"first 30 s, {n_small} maps, oracle (direct-DFT t2w)"}
        res["parity_sample"] = ok
        log(f"max-energy cpu baseline {Ls * n_small / cpu_s:.3e} map-samples/s; parity {ok}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--clusters", type=int, default=4096)
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--K", type=int, default=64)
    # 64-direction workgroups, 20 per CU: the fastest launch shape measured on
    # the 4,096-cluster order-7 workload (the autotuner's 384-cluster
    # calibration batch can prefer another); --block 0 runs the autotuner
    ap.add_argument("--block", type=int, default=64, help="0: autotune")
    ap.add_argument("--blocks-per-cu", type=int, default=20)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parity-clusters", type=int, default=6)
    ap.add_argument("--refine-clusters", type=int, default=512)
    ap.add_argument("--refine-steps", type=int, default=20)
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-wdm", action="store_true")
    ap.add_argument("--no-tdamp", action="store_true")
    ap.add_argument("--no-subnet", action="store_true")
    ap.add_argument("--no-xtalk", action="store_true")
    ap.add_argument("--no-select", action="store_true")
    ap.add_argument("--no-maxenergy", action="store_true")
    ap.add_argument("--no-cluster", action="store_true")
    ap.add_argument("--no-threshold", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-pixel-classes", action="store_true",
                    help="regulator sweep with cluster-wide term bounds only")
    ap.add_argument("--only", choices=("refine", "wdm", "tdamp", "subnet", "xtalk", "select",
                                       "maxenergy", "cluster", "threshold", "e2e"), default=None,
                    help="run one side leg alone and print its JSON (for profiling)")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="first copy this many GiB device-to-device (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "the likelihood path runs on one GPU; --gpus must be 1 "
                          "(sharding by cluster is not implemented)", "n_gpus": args.gpus}))
        log("multi-GPU runs are not implemented (clusters would shard by rank); use --gpus 1")
        return 2

    import torch

    from pycwb_amd import autotune, synth
    from pycwb_amd.batch import BatchedSkyScan, PlainLoop, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calib_gib > 0:
        n = int(args.copy_calib_gib * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n * 4} bytes read and written")
        del src, dst
    if args.only == "refine":
        print(json.dumps({"refinement": refinement_leg(
            args.refine_clusters, args.order, args.K, args.refine_steps,
            torch.cuda.current_stream(), log)}), flush=True)
        return 0
    if args.only == "wdm":
        print(json.dumps({"wdm": wdm_leg(torch.cuda.current_stream(), log)}), flush=True)
        return 0
    if args.only == "tdamp":
        print(json.dumps({"tdamp": tdamp_leg(torch.cuda.current_stream(), log)}), flush=True)
        return 0
    if args.only == "maxenergy":
        print(json.dumps({"maxenergy": maxenergy_leg(torch.cuda.current_stream(), log,
                                                     cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "select":
        print(json.dumps({"select": select_leg(torch.cuda.current_stream(), log,
                                               cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "e2e":
        print(json.dumps({"e2e": e2e_leg(torch.cuda.current_stream(), log,
                                         cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "threshold":
        print(json.dumps({"threshold": threshold_leg(torch.cuda.current_stream(), log,
                                                     cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "cluster":
        print(json.dumps({"cluster": cluster_leg(torch.cuda.current_stream(), log,
                                                 cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "xtalk":
        print(json.dumps({"xtalk": xtalk_leg(torch.cuda.current_stream(), log,
                                             cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.only == "subnet":
        print(json.dumps({"subnet": subnet_leg(torch.cuda.current_stream(), log,
                                               cpu=not args.no_cpu)}), flush=True)
        return 0
    t0 = time.perf_counter()
    nets = [synth.make_network(synth.NETWORK_3, order=args.order, K=args.K),
            synth.make_network(synth.NETWORK_4, order=args.order, K=args.K)]
    sizes = synth.cluster_sizes(args.clusters, seed=1002)
    net_ids = np.arange(args.clusters) % 2
    batch = synth_batch(nets, sizes, net_ids, seed=1002 * 100000, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    n_valid = int(valid.numel())
    evals = batch.evals(n_valid)
    log(f"workload ready in {time.perf_counter() - t0:.1f} s: {args.clusters} clusters, "
        f"{int(sizes.sum())} pixels, {evals:.3e} evals, td {batch.td.numel() * 4 / 1e9:.2f} GB")

    gamma, delta, netCC = 0.5 * 0.5 * 2 / 3, 0.5, 0.5
    if args.block:
        block = args.block
        grid = None if not args.blocks_per_cu else \
            torch.cuda.get_device_properties(0).multi_processor_count * args.blocks_per_cu
        tuned = {"block": block, "grid": grid, "source": "arguments"}
    else:
        tuned = autotune.tune_batched(nets, nt, valid, gamma, delta, netCC, log=log)
        block, grid = tuned["block"], tuned["grid"]
    bs = BatchedSkyScan(batch, nt, valid, gamma, delta, netCC, block=block, grid=grid,
                        pixel_classes=not args.no_pixel_classes)
    stats = SkyStatistics(batch, nt)   # self-only cross-talk
    reg0 = float(np.float32(delta * np.sqrt(2)))
    log(f"batched: block={bs.block} grid={bs.grid} ({tuned.get('source')})")

    def step():
        bs.launch(stream)                              # regulator + scan + argmax (hipGraph)
        stats.run(bs.l_max, bs.REG1, reg0, stream=stream)  # Lm, Ec, rho, ... at l_max

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    rb = bs.results()
    lm_gpu = stats.scalars()["Lm"]
    log(f"batched step {ms:.1f} ms -> {evals / ms * 1e3:.4e} evals/s")

    # dominant kernel: the scan stage, timed alone on the same stream
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    bs.launch_stage(0, stream)
    e[0].record(stream)
    bs.launch_stage(1, stream)
    e[1].record(stream)
    bs.launch_stage(2, stream)
    e[2].record(stream)
    bs.launch_stage(3, stream)
    torch.cuda.synchronize()
    reg_ms = e[0].elapsed_time(e[1])
    scan_ms = e[1].elapsed_time(e[2])
    rs = bs.results()
    active = bs.active_pairs()
    tested = bs.tested_pairs()
    sf = scan_flops(batch, n_valid)
    rf = required_scan_flops(batch, tested, active)
    achieved = rf / (scan_ms * 1e-3) / 1e12
    act_frac = float(active.sum()) / float(batch.evals(n_valid))
    tested_frac = float(tested.sum()) / float(batch.evals(n_valid))
    ceiling = valu_ceiling(stream)
    valu_insts, valu_src = profiled_value("valu_insts_per_scan_stage", args.clusters, args.order,
                                          args.K)
    issue = None if valu_insts is None else valu_insts / (scan_ms * 1e-3)
    log(f"scan stage {scan_ms:.1f} ms ({achieved:.2f} TFLOP/s required work, "
        f"{sf / (scan_ms * 1e-3) / 1e12:.2f} effective), active fraction {act_frac:.4f}, "
        f"energy-tested fraction {tested_frac:.4f}, regulator {reg_ms:.1f} ms, VALU ceiling "
        f"{ceiling:.3e} wave-instr/s")

    plain = None
    parity = {"batched_vs_plain_l_max": None}
    if not args.no_plain:
        pl = PlainLoop(batch, nt, valid, gamma, delta, netCC)
        pl.launch(stream)  # warm
        torch.cuda.synchronize()
        ep0, ep1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ep0.record(stream)
        pl.launch(stream)
        ep1.record(stream)
        torch.cuda.synchronize()
        pms = ep0.elapsed_time(ep1)
        rp = pl.results()
        plain = {"ms_per_step": round(pms, 3), "evals_per_s": evals / pms * 1e3}
        parity = {
            "batched_vs_plain_l_max": bool(np.array_equal(rb["l_max"], rp["l_max"])),
            "batched_vs_plain_stat": bool(np.array_equal(rb["sky_stat_max"], rp["sky_stat_max"])),
            "stage_vs_graph_l_max": bool(np.array_equal(rb["l_max"], rs["l_max"])),
        }
        log(f"plain loop {pms:.1f} ms -> {evals / pms * 1e3:.4e} evals/s; parity {parity}")

    # parity against the oracle on the smallest clusters (full sky each)
    if not args.no_cpu:
        parity.update(oracle_parity(batch, nets, nt, rb, lm_gpu, gamma, delta, netCC, reg0,
                                    n_check=args.parity_clusters))
        log(f"oracle parity: {parity}")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, cpu_info()["usable_cpus"])
        cpu = cpu_baseline(threads, order=args.order)
        log(f"cpu baseline {cpu['value']:.4e} evals/s on {threads} threads "
            f"({cpu['one_thread_evals_per_s']:.4e} on one), {cpu['cpu_model']}")

    refinement = None
    if not args.no_refine:
        refinement = refinement_leg(args.refine_clusters, args.order, args.K, args.refine_steps,
                                    stream, log)
    wdm = None if args.no_wdm else wdm_leg(stream, log)
    tdamp = None if args.no_tdamp else tdamp_leg(stream, log)
    subnet = None if args.no_subnet else subnet_leg(stream, log, cpu=not args.no_cpu)
    xtalk = None if args.no_xtalk else xtalk_leg(stream, log, cpu=not args.no_cpu)
    select = None if args.no_select else select_leg(stream, log, cpu=not args.no_cpu)
    maxen = None if args.no_maxenergy else maxenergy_leg(stream, log, cpu=not args.no_cpu)
    clus = None if args.no_cluster else cluster_leg(stream, log, cpu=not args.no_cpu)
    thr = None if args.no_threshold else threshold_leg(stream, log, cpu=not args.no_cpu)
    rates = {k: v["cpu_baseline"] for k, v in (("selection", select), ("max_energy", maxen),
                                                 ("clustering", clus), ("threshold", thr))
             if v and v.get("cpu_baseline")}
    e2e = None if args.no_e2e else e2e_leg(stream, log, cpu=not args.no_cpu, cpu_rates=rates)

    traffic, traffic_src = profiled_value("traffic_bytes_per_scan_stage", args.clusters,
                                          args.order, args.K)
    out = {
        "metric": "sky x cluster-pixel evals/s, batched vs plain GPU at Lm rel.err<=1e-5",
        "value": evals / ms * 1e3,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded Gaussian noise + coherent injections, generated on device)",
        "config": {"workload": "batched likelihood: 4096 clusters, P log-uniform[16,4096], "
                               "half 3-det half 4-det, K=64, HEALPix order 7",
                   "clusters": args.clusters, "healpix_order": args.order,
                   "n_sky": n_valid, "pixels": int(sizes.sum()), "evals": evals,
                   "block": bs.block, "grid": bs.grid},
        "plain": plain,
        "speedup_vs_plain": (plain and (evals / ms * 1e3) / plain["evals_per_s"]),
        "parity": parity,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_batch_persistent<N,BLOCK,1> (scan stage)",
                     "kernel_ms": round(scan_ms, 3), "required_flops": rf,
                     "required_flops_rule": "(4n+4) per energy-tested evaluation + (74n+125) "
                                            "per evaluation above the energy threshold",
                     "nominal_flops": sf,
                     "effective_tflops": sf / (scan_ms * 1e-3) / 1e12,
                     "effective_note": "(78n+129) flops on every evaluation / scan time: a "
                                       "throughput, not a roofline (exact skips)",
                     "active_fraction": act_frac, "energy_tested_fraction": tested_frac,
                     "valu_issue": {"insts_per_launch": valu_insts, "source": valu_src,
                                    "wave_insts_per_s": issue,
                                    "ceiling_wave_insts_per_s": ceiling,
                                    "frac": None if issue is None else issue / ceiling,
                                    "ceiling_kernel": "pycwb_valu_probe (float32 FMA chains, "
                                                      "8 waves per SIMD)",
                                    "note": "frac above 1: the stage issues VALU instructions "
                                            "faster than a stream of v_fma_f32 (one per SIMD "
                                            "every 4 cycles): it is bound by instruction "
                                            "issue, and its instruction count is the lever"},
                     "algorithmic_bytes": scan_bytes(batch, nt, n_valid),
                     "hbm_gbs": scan_bytes(batch, nt, n_valid) / (scan_ms * 1e-3) / 1e9},
        "regulator": {"ms": round(reg_ms, 3),
                      "nominal_tflops": regulator_flops(batch, n_valid) / (reg_ms * 1e-3) / 1e12,
                      "nominal_note": "(24n+25) flops per pixel of every direction / regulator "
                                      "time: a throughput, not a roofline (the term bounds "
                                      "decide most directions before their last pixel)"},
        "cpu_baseline": cpu,
        "refinement": refinement,
        "wdm": wdm,
        "tdamp": tdamp,
        "subnet": subnet,
        "xtalk": xtalk,
        "select": select,
        "maxenergy": maxen,
        "cluster": clus,
        "threshold": thr,
        "e2e": e2e,
    }
    print(json.dumps(out), flush=True)
    bs.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
