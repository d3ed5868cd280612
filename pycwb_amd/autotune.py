"""Autotuner for the batched sky scan's launch shape.

Tunes the two shape parameters of the persistent work-queue kernels:
  block : sky directions per workgroup (64, 128, 256) -- the sky-tile size;
  grid  : persistent workgroups (waves per SIMD of resident work).
It times the regulator and scan stages together (the persistent kernels
the shape affects) with HIP events on a seeded calibration batch drawn from
the same cluster-size distribution (384 clusters by default: with fewer, the
few largest clusters' tails decide the timing and favour shapes that are
slower on a full batch), keeps the fastest shape (among shapes
within 2 % of the fastest, the largest grid: more resident waves hide the
regulator's latency better on the full batch), and caches the choice per
(device, detector counts, K, directions) in
``$XDG_CACHE_HOME/pycwb_amd/autotune.json``.
"""

from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import synth

CANDIDATES = {64: (4, 8, 12, 16, 20), 128: (2, 4, 6, 8, 10), 256: (1, 2, 3, 4, 5)}


def _cache_path():
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "pycwb_amd", "autotune.json")


def _key(nets, n_valid):
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{props.name}|{props.multi_processor_count}|n={sorted(set(nets.nifo))}|" \
           f"K={sorted(set(nets.K))}|S={n_valid}"


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d):
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


def time_stages(bs, reps: int = 2) -> float:
    """Best-of-reps time of the regulator + scan stages (ms)."""
    stream = torch.cuda.current_stream()
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        bs.launch_stage(0, stream)
        a.record(stream)
        bs.launch_stage(1, stream)
        bs.launch_stage(2, stream)
        b.record(stream)
        bs.launch_stage(3, stream)
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune_batched(net_list, nets, valid, gamma, delta, netCC, n_calib: int = 384,
                 use_cache: bool = True, log=None) -> dict:
    """Returns {"block", "grid", "source", "ms"} for the batched regulator + scan."""
    from .batch import BatchedSkyScan, synth_batch

    key = _key(nets, int(valid.numel()))
    cache = _load() if use_cache else {}
    if key in cache:
        r = dict(cache[key])
        r["source"] = "cache"
        return r
    dev = valid.device
    sizes = synth.cluster_sizes(n_calib, seed=4242)
    ids = np.arange(n_calib) % len(net_list)
    calib = synth_batch(net_list, sizes, ids, seed=4242, device=dev)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    results = []
    for block, per_cu in CANDIDATES.items():
        for k in per_cu:
            grid = cus * k
            bs = BatchedSkyScan(calib, nets, valid, gamma, delta, netCC, block=block, grid=grid,
                                use_graph=False)
            bs.launch_stage(0)
            bs.launch_stage(1)
            bs.launch_stage(2)
            bs.launch_stage(3)
            ms = time_stages(bs)
            bs.close()
            results.append((ms, block, grid))
            if log:
                log(f"autotune block={block} grid={grid}: regulator+scan {ms:.2f} ms")
    fastest = min(r[0] for r in results)
    ms, block, grid = max((r for r in results if r[0] <= 1.02 * fastest),
                          key=lambda r: (r[2] * r[1], -r[0]))
    r = {"block": block, "grid": grid, "ms": ms, "source": "tuned",
         "table": [[b, g, m] for m, b, g in results]}
    cache[key] = {k: v for k, v in r.items() if k != "source"}
    if use_cache:
        _store(cache)
    return r
