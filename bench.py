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
OpenMP restatement of PycWB's Numba scan on a bounded sample.

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


def required_scan_flops(batch, n_valid, active):
    """Flops the scan needs on this data: the energy test (4n + 4) on every
    evaluation and the rest of the chain (74n + 125) on the evaluations above
    the energy threshold; the others contribute exactly nothing."""
    n = batch.nifo.astype(np.float64)
    every = float(np.sum(batch.n_pix.astype(np.float64) * (4.0 * n + 4.0))) * n_valid
    return every + float(np.sum(active.astype(np.float64) * (74.0 * n + 125.0)))


def regulator_flops(batch, n_valid):
    return float(np.sum(batch.n_pix.astype(np.float64) * (24.0 * batch.nifo + 25.0))) * n_valid


def scan_bytes(batch, nets, n_valid):
    """Unique HBM bytes of the scan stage: every cluster's amplitude pairs and
    noise weights once, the sky tables of each network once."""
    td = float(np.sum(batch.n_pix.astype(np.float64) * batch.nifo * (2 * batch.K + 1) * 8))
    rms = float(np.sum(batch.n_pix.astype(np.float64) * batch.nifo * 4))
    sky = float(sum(n * n_valid * 12 for n in nets.nifo))
    return td + rms + sky


def cpu_baseline(threads, order=7, P=1024, repeats=2):
    """Oracle (C/OpenMP restatement of PycWB's Numba scan) on one plain-config
    cluster: n=3, P=1,024, K=64, all order-7 directions."""
    from oracle import oracle as o
    from pycwb_amd import synth

    net = synth.make_network(synth.NETWORK_3, order=order, K=64)
    cl = synth.make_cluster(net, P, seed=1001 * 100000)
    valid = np.arange(net.n_sky, dtype=np.int64)
    REG = np.array([0.5 * np.sqrt(2), 1.0, 0.0], np.float32)
    best = math.inf
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.scan_sky(3, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"], cl["td90"],
                   net.ml, REG, 0.5, 0.5, 12.0, valid, threads=threads)
        best = min(best, time.perf_counter() - t0)
    evals = float(net.n_sky) * P
    return evals / best, f"one n=3 cluster, P={P}, K=64, order {order} ({net.n_sky} dirs), " \
                         f"best of {repeats}, {best:.2f} s"


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--clusters", type=int, default=4096)
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--K", type=int, default=64)
    ap.add_argument("--block", type=int, default=0, help="0: autotune")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parity-clusters", type=int, default=6)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("multi-GPU runs are not part of this build (one GPU per box); use --gpus 1")
        return 2

    import torch

    from pycwb_amd import autotune, synth
    from pycwb_amd.batch import BatchedSkyScan, PlainLoop, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
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
        tuned = {"block": block, "grid": grid, "source": "cli"}
    else:
        tuned = autotune.tune_batched(nets, nt, valid, gamma, delta, netCC, log=log)
        block, grid = tuned["block"], tuned["grid"]
    bs = BatchedSkyScan(batch, nt, valid, gamma, delta, netCC, block=block, grid=grid)
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
    sf = scan_flops(batch, n_valid)
    rf = required_scan_flops(batch, n_valid, active)
    achieved = rf / (scan_ms * 1e-3) / 1e12
    act_frac = float(active.sum()) / float(batch.evals(n_valid))
    log(f"scan stage {scan_ms:.1f} ms ({achieved:.2f} TFLOP/s required work, "
        f"{sf / (scan_ms * 1e-3) / 1e12:.2f} nominal), active fraction {act_frac:.4f}, "
        f"regulator {reg_ms:.1f} ms")

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
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, sample = cpu_baseline(threads, order=args.order)
        cpu = {"value": v, "unit": "evals/s", "cores": threads, "kind": "port", "sample": sample}
        log(f"cpu baseline {v:.4e} evals/s on {threads} threads")

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
                     "traffic": None, "kernel": "k_batch_persistent<N,BLOCK,1> (scan stage)",
                     "kernel_ms": round(scan_ms, 3), "required_flops": rf,
                     "nominal_flops": sf,
                     "nominal_tflops": sf / (scan_ms * 1e-3) / 1e12,
                     "active_fraction": act_frac,
                     "algorithmic_bytes": scan_bytes(batch, nt, n_valid),
                     "hbm_gbs": scan_bytes(batch, nt, n_valid) / (scan_ms * 1e-3) / 1e9},
        "regulator": {"ms": round(reg_ms, 3),
                      "tflops": regulator_flops(batch, n_valid) / (reg_ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    bs.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
