"""Batched sky scan: many clusters, one hipGraph of persistent work-queue kernels.

Replaces the per-cluster loop of ``evaluate_fragment_clusters`` /
``likelihood_wrapper`` (pycwb/modules/likelihoodWP/likelihood.py:58-135) up to
the sky scan: regulator, scan and argmax of every cluster run on the GPU with
no host round trip.  ``PlainLoop`` runs the same clusters one per launch family
(the plain kernels), on the same device buffers, for comparison.

Device layout of a batch (all clusters concatenated):
  rms  float32, cluster c at rms_off[c]:  (P_c, n_c)
  td   float32 pairs, cluster c at 2*td_off[c]: (P_c, n_c, 2K_c+1, 2)
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, sky, synth
from .device import lib, ptr, require_gpu, stream_ptr


@dataclass
class NetTablesDev:
    """Sky tables of all networks in one batch, concatenated."""
    n_sky: int
    nifo: list
    En: list
    off: torch.Tensor          # int64 (n_net,)  element offset per network
    FP: torch.Tensor           # float32 flat: per network (S, n)
    FX: torch.Tensor
    ml: torch.Tensor           # int32 flat: per network (n, S)
    nifo_d: torch.Tensor       # int32 (n_net,)
    En_d: torch.Tensor         # float64 (n_net,)
    K: list = field(default_factory=list)
    ml_host: list = field(default_factory=list)   # per network (n, S) int32
    side_host: list = field(default_factory=list)  # per network (S,) bool or None

    def sky_order(self, valid_h: np.ndarray, device) -> torch.Tensor:
        """(n_net, n_valid) valid-list positions in delay-tuple order per network."""
        key = ("order", valid_h.tobytes())
        if getattr(self, "_order_key", None) != key:
            sides = self.side_host or [None] * len(self.ml_host)
            o = np.stack([sky.delay_order(m, valid_h, sd) for m, sd in zip(self.ml_host, sides)])
            self._order = torch.from_numpy(np.ascontiguousarray(o)).to(device)
            self._order_key = key
        return self._order

    def finite(self) -> bool:
        return bool(torch.isfinite(self.FP).all()) and bool(torch.isfinite(self.FX).all())


def net_tables(nets, acore: float = math.sqrt(2.0), device=None) -> NetTablesDev:
    device = device or require_gpu()
    S = nets[0].n_sky
    if any(n.n_sky != S for n in nets):
        raise ValueError("all networks of a batch must share the sky grid")
    offs, fp, fx, ml = [], [], [], []
    o = 0
    for n in nets:
        offs.append(o)
        fp.append(np.ascontiguousarray(n.FP_t, np.float32).ravel())
        fx.append(np.ascontiguousarray(n.FX_t, np.float32).ravel())
        ml.append(np.ascontiguousarray(n.ml, np.int32).ravel())
        o += S * n.n_ifo
    En = [2 * acore * acore * n.n_ifo for n in nets]
    t = lambda a, dt: torch.from_numpy(np.concatenate(a)).to(device=device, dtype=dt)
    return NetTablesDev(
        n_sky=S, nifo=[n.n_ifo for n in nets], En=En,
        off=torch.tensor(offs, dtype=torch.int64, device=device),
        FP=t(fp, torch.float32), FX=t(fx, torch.float32), ml=t(ml, torch.int32),
        nifo_d=torch.tensor([n.n_ifo for n in nets], dtype=torch.int32, device=device),
        En_d=torch.tensor(En, dtype=torch.float64, device=device), K=[n.K for n in nets],
        ml_host=[np.ascontiguousarray(n.ml, np.int32) for n in nets],
        side_host=[sky.mirror_side(n.ifos, n.phi, n.lat) for n in nets])


@dataclass
class ClusterBatch:
    n_pix: np.ndarray          # int32 (C,)
    net: np.ndarray            # int32 (C,)
    K: np.ndarray              # int32 (C,)
    nifo: np.ndarray           # int32 (C,)
    rms_off: np.ndarray        # int64 (C,)
    td_off: np.ndarray         # int64 (C,) in float pairs
    rms: torch.Tensor          # float32 flat
    td: torch.Tensor           # float32 flat (pairs)
    l0: np.ndarray = None      # injection directions (synthetic batches)
    truth: np.ndarray = None   # (C, 2) off-grid injection (theta, phi), smooth batches

    @property
    def n_clusters(self):
        return int(self.n_pix.size)

    def cluster_rms(self, c):
        n, P = int(self.nifo[c]), int(self.n_pix[c])
        o = int(self.rms_off[c])
        return self.rms[o:o + P * n].view(P, n)

    def cluster_td(self, c):
        """(P, n, 2K+1, 2) view of cluster c."""
        n, P, K = int(self.nifo[c]), int(self.n_pix[c]), int(self.K[c])
        o = 2 * int(self.td_off[c])
        return self.td[o:o + P * n * (2 * K + 1) * 2].view(P, n, 2 * K + 1, 2)

    def evals(self, n_valid: int) -> int:
        return int(np.sum(self.n_pix.astype(np.int64)) * n_valid)


def synth_batch(nets, sizes, net_ids, seed: int, snr: float = 3.0, device=None,
                signal_fraction: float = 200 / 1024) -> ClusterBatch:
    """Synthetic batch generated on the device (noise) plus host-drawn injections.

    Noise: unit Gaussian on every amplitude.  Noise weights from |N(1, 0.1)|
    rms.  Each cluster has round(P * signal_fraction) signal pixels carrying a
    coherent component from its own uniform-sphere grid direction.
    """
    device = device or require_gpu()
    sizes = np.asarray(sizes, dtype=np.int64)
    net_ids = np.asarray(net_ids, dtype=np.int64)
    C_ = sizes.size
    nifo = np.array([nets[k].n_ifo for k in net_ids], dtype=np.int64)
    K = np.array([nets[k].K for k in net_ids], dtype=np.int64)
    nd = 2 * K + 1
    rms_len = sizes * nifo
    td_len = sizes * nifo * nd
    rms_off = np.concatenate([[0], np.cumsum(rms_len)[:-1]])
    td_off = np.concatenate([[0], np.cumsum(td_len)[:-1]])
    total_td = int(td_len.sum())
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    td = torch.randn(2 * total_td, generator=g, device=device, dtype=torch.float32)
    sigma = (torch.randn(int(rms_len.sum()), generator=g, device=device, dtype=torch.float64)
             * 0.1 + 1.0).abs()
    # r_ij = (1/s_ij) / sqrt(sum_i 1/s_ij^2), float64 -> float32 (pixel_data.py:60-62)
    inv = 1.0 / sigma
    pix_id = torch.from_numpy(np.repeat(np.arange(int(sizes.sum())), np.repeat(nifo, sizes)))
    pix_id = pix_id.to(device)
    ss = torch.zeros(int(sizes.sum()), dtype=torch.float64, device=device)
    ss.index_add_(0, pix_id, inv * inv)
    rms = (inv / torch.sqrt(ss)[pix_id]).to(torch.float32)
    rms_h = rms.cpu().numpy()

    rng = np.random.default_rng(seed)
    l0 = np.zeros(C_, dtype=np.int64)
    idx_all, val_all = [], []
    for c in range(C_):
        net = nets[net_ids[c]]
        n, P, k = int(nifo[c]), int(sizes[c]), int(K[c])
        l0[c] = rng.integers(net.n_sky)
        ns = max(1, round(P * signal_fraction))
        sig = rng.choice(P, size=min(ns, P), replace=False)
        h = rng.standard_normal((sig.size, 4))
        fp = net.FP_t[l0[c]].astype(np.float64)
        fx = net.FX_t[l0[c]].astype(np.float64)
        r = rms_h[rms_off[c]: rms_off[c] + P * n].reshape(P, n)[sig].astype(np.float64)
        v00 = r * (fp[None, :] * h[:, :1] + fx[None, :] * h[:, 1:2])
        v90 = r * (fp[None, :] * h[:, 2:3] + fx[None, :] * h[:, 3:4])
        e = np.sqrt(np.sum(v00 ** 2 + v90 ** 2, axis=1, keepdims=True))
        scale = np.where(e > 0, snr / np.where(e > 0, e, 1), 0)
        v00, v90 = v00 * scale, v90 * scale
        taps = net.ml[:, l0[c]].astype(np.int64) + k
        base = td_off[c] + (sig[:, None] * n + np.arange(n)[None, :]) * (2 * k + 1) + taps[None, :]
        idx_all.append((2 * base).ravel())
        val_all.append(v00.ravel())
        idx_all.append((2 * base + 1).ravel())
        val_all.append(v90.ravel())
    idx = torch.from_numpy(np.concatenate(idx_all)).to(device)
    val = torch.from_numpy(np.concatenate(val_all).astype(np.float32)).to(device)
    td.index_add_(0, idx, val)
    return ClusterBatch(n_pix=sizes.astype(np.int32), net=net_ids.astype(np.int32),
                        K=K.astype(np.int32), nifo=nifo.astype(np.int32),
                        rms_off=rms_off.astype(np.int64), td_off=td_off.astype(np.int64),
                        rms=rms, td=td, l0=l0)


def synth_batch_smooth(nets, sizes, net_ids, seed: int, snr: float = 6.0, tap_sigma: float = 2.0,
                       device=None, signal_fraction: float = 200 / 1024) -> ClusterBatch:
    """Synthetic batch whose delayed amplitudes vary smoothly with the delay.

    A td table samples one pixel's amplitude at successive time shifts, so a
    real one is correlated along the tap axis over the pixel's duration.  Here
    the noise is white Gaussian smoothed along the taps by a Gaussian of
    sigma tap_sigma / sqrt(2) (unit variance per tap, correlation
    exp(-d^2 / (2 tap_sigma^2)) between taps d apart), and each cluster's
    signal comes from a direction drawn uniformly on the sphere (off the
    grid): detector i carries r_i (F+_i h+ + Fx_i hx) * exp(-(k - u_i)^2 /
    (2 tap_sigma^2)) at tap k, u_i = K + continuous delay.  All clusters must
    share K.  ``truth`` holds the injected (theta, phi).
    """
    from .skygrad import network_geometry

    device = device or require_gpu()
    sizes = np.asarray(sizes, dtype=np.int64)
    net_ids = np.asarray(net_ids, dtype=np.int64)
    Ks = {nets[k].K for k in set(net_ids.tolist())}
    if len(Ks) != 1:
        raise ValueError("synth_batch_smooth needs one K for every cluster")
    K = Ks.pop()
    nd = 2 * K + 1
    b = synth_batch(nets, sizes, net_ids, seed, snr=0.0, device=device, signal_fraction=0.0)
    rows = b.td.numel() // (2 * nd)
    hw = int(math.ceil(4 * tap_sigma))
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1)
    xs = np.arange(-hw, hw + 1, dtype=np.float64)
    ker = np.exp(-xs ** 2 / (tap_sigma ** 2))        # sigma_n = tap_sigma / sqrt(2)
    ker = ker / math.sqrt(float(np.sum(ker ** 2)))
    td = b.td.view(rows, nd, 2)
    chunk = 1 << 18
    for r0 in range(0, rows, chunk):
        r1 = min(rows, r0 + chunk)
        w = torch.randn(r1 - r0, nd + 2 * hw, 2, generator=g, device=device)
        sm = torch.zeros(r1 - r0, nd, 2, device=device)
        for k in range(ker.size):  # Gaussian smoothing along the taps
            sm.add_(w[:, k:k + nd], alpha=float(ker[k]))
        td[r0:r1] = sm
    rms_h = b.rms.cpu().numpy()
    rng = np.random.default_rng(seed)
    truth = np.zeros((sizes.size, 2))
    geos = {}
    taps = np.arange(nd, dtype=np.float64)
    idx_all, val_all = [], []
    for c in range(sizes.size):
        net = nets[net_ids[c]]
        n, P = net.n_ifo, int(sizes[c])
        if net_ids[c] not in geos:
            geos[net_ids[c]] = network_geometry(net.ifos, net.td_rate)[:n]
        geo = geos[net_ids[c]]
        th, ph = math.acos(rng.uniform(-1.0, 1.0)), rng.uniform(0.0, 2 * math.pi)
        truth[c] = th, ph
        nvec = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        u = np.clip(geo[:, :3] @ nvec + K, 0.0, nd - 1.0)
        fp, fx = np.zeros(n), np.zeros(n)
        for i in range(n):
            x = np.array([math.sin(ph), -math.cos(ph), 0.0])
            y = np.array([-math.cos(ph) * math.cos(th), -math.sin(ph) * math.cos(th),
                          math.sin(th)])
            D = geo[i, 3:].reshape(3, 3)
            fp[i] = x @ D @ x - y @ D @ y
            fx[i] = x @ D @ y + y @ D @ x
        ns = max(1, round(P * signal_fraction))
        sig = rng.choice(P, size=min(ns, P), replace=False)
        h = rng.standard_normal((sig.size, 4))
        r = rms_h[b.rms_off[c]: b.rms_off[c] + P * n].reshape(P, n)[sig].astype(np.float64)
        v00 = r * (fp[None, :] * h[:, :1] + fx[None, :] * h[:, 1:2])
        v90 = r * (fp[None, :] * h[:, 2:3] + fx[None, :] * h[:, 3:4])
        e = np.sqrt(np.sum(v00 ** 2 + v90 ** 2, axis=1, keepdims=True))
        scale = np.where(e > 0, snr / np.where(e > 0, e, 1), 0)
        v00, v90 = v00 * scale, v90 * scale
        prof = np.exp(-(taps[None, :] - u[:, None]) ** 2 / (2 * tap_sigma ** 2))  # (n, nd)
        base = b.td_off[c] + (sig[:, None] * n + np.arange(n)[None, :]) * nd      # (s, n)
        idx = base[:, :, None] + np.arange(nd)[None, None, :]
        idx_all += [(2 * idx).ravel(), (2 * idx + 1).ravel()]
        val_all += [(v00[:, :, None] * prof[None]).ravel(), (v90[:, :, None] * prof[None]).ravel()]
    if idx_all:
        b.td.index_add_(0, torch.from_numpy(np.concatenate(idx_all)).to(device),
                        torch.from_numpy(np.concatenate(val_all).astype(np.float32)).to(device))
    b.l0 = None
    b.truth = truth
    return b


def lpt_order(batch: ClusterBatch):
    """Clusters grouped by detector count; inside a group, largest work first."""
    cost = batch.n_pix.astype(np.float64) * (78 * batch.nifo + 129 + 24 * batch.nifo + 25)
    groups = []
    order = []
    for n in sorted(set(batch.nifo.tolist())):
        ids = np.nonzero(batch.nifo == n)[0]
        ids = ids[np.argsort(-cost[ids], kind="stable")]
        groups.append((len(order), len(ids), n))
        order.extend(ids.tolist())
    return np.asarray(order, dtype=np.int32), groups


def default_grid(block: int, blocks_per_cu: int | None = None) -> int:
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    cus = props.multi_processor_count
    if blocks_per_cu is None:
        blocks_per_cu = max(1, 1280 // block)
    return cus * blocks_per_cu


class BatchedSkyScan:
    """Regulator + scan + argmax of a ClusterBatch as one hipGraph launch."""

    def __init__(self, batch: ClusterBatch, nets: NetTablesDev, valid: torch.Tensor,
                 gamma_regulator: float, delta_regulator: float, netCC: float,
                 block: int = 256, grid: int | None = None, use_graph: bool = True,
                 flags: int | None = None, ordered: bool = True, pixel_classes: bool = True):
        dev = batch.td.device
        self.batch, self.nets, self.valid = batch, nets, valid
        C_ = batch.n_clusters
        order, groups = lpt_order(batch)
        self.order = torch.from_numpy(order).to(dev)
        self.d_npix = torch.from_numpy(batch.n_pix).to(dev)
        self.d_net = torch.from_numpy(batch.net).to(dev)
        self.d_K = torch.from_numpy(batch.K).to(dev)
        self.d_rms_off = torch.from_numpy(batch.rms_off).to(dev)
        self.d_td_off = torch.from_numpy(batch.td_off).to(dev)
        self.tickets = torch.zeros(_lib.TICKETS_PER_GROUP * len(groups), dtype=torch.int32,
                                   device=dev)
        self.counts = torch.zeros(C_, dtype=torch.int32, device=dev)
        self.best = torch.zeros(C_, dtype=torch.int64, device=dev)
        self.REG1 = torch.zeros(C_, dtype=torch.float32, device=dev)
        self.l_max = torch.zeros(C_, dtype=torch.int64, device=dev)
        self.stat_max = torch.zeros(C_, dtype=torch.float64, device=dev)
        self.reg_f64 = torch.zeros(C_, dtype=torch.float64, device=dev)
        self.wstats = torch.zeros(C_ * _lib.WSTATS_PER_CLUSTER, dtype=torch.float64, device=dev)
        d = _lib.BatchDesc()
        d.n_clusters = C_
        d.n_sky = nets.n_sky
        d.sky_valid_indices = ptr(valid)
        d.n_valid = int(valid.numel())
        d.order = ptr(self.order)
        d.n_groups = len(groups)
        for g, (first, count, n) in enumerate(groups):
            d.group_first[g], d.group_count[g], d.group_nifo[g] = first, count, n
        d.n_pix, d.net, d.K = ptr(self.d_npix), ptr(self.d_net), ptr(self.d_K)
        d.rms_off, d.td_off = ptr(self.d_rms_off), ptr(self.d_td_off)
        d.rms, d.td = ptr(batch.rms), ptr(batch.td)
        d.net_nifo, d.net_off, d.net_En = ptr(nets.nifo_d), ptr(nets.off), ptr(nets.En_d)
        d.FP, d.FX, d.ml = ptr(nets.FP), ptr(nets.FX), ptr(nets.ml)
        self.sky_order = nets.sky_order(valid.cpu().numpy(), dev) if ordered else None
        d.sky_order = ptr(self.sky_order)
        if flags is None:
            finite = nets.finite() and bool(torch.isfinite(batch.rms).all())
            flags = _lib.DEFAULT_FLAGS if finite else _lib.REGULATOR_EARLY_EXIT
        d.flags = int(flags)
        self.flags = d.flags
        d.gamma_regulator = float(gamma_regulator)
        d.reg0 = float(np.float32(delta_regulator * np.sqrt(2)))
        d.netCC = float(netCC)
        d.tickets, d.counts, d.best = ptr(self.tickets), ptr(self.counts), ptr(self.best)
        d.REG1, d.l_max = ptr(self.REG1), ptr(self.l_max)
        d.sky_stat_max, d.reg_f64 = ptr(self.stat_max), ptr(self.reg_f64)
        d.block = int(block)
        d.grid = int(grid or default_grid(block))
        d.use_graph = 1 if use_graph else 0
        d.weight_stats = ptr(self.wstats)
        if pixel_classes:
            # regulator sweep bounds per class of pixels with similar noise weights
            ent = np.array([int(lib().pycwb_class_rem_entries(int(p))) for p in
                            np.unique(batch.n_pix)], np.int64)
            per = dict(zip(np.unique(batch.n_pix).tolist(), ent.tolist()))
            n_ent = np.array([per[int(p)] for p in batch.n_pix], np.int64)
            rem_off = np.concatenate([[0], np.cumsum(n_ent)[:-1]]).astype(np.int64)
            # one byte per noise weight: cluster c's class bytes start at rms_off[c]
            self.pclass = torch.empty(max(int(batch.rms.numel()), 1), dtype=torch.uint8,
                                      device=dev)
            self.cstats = torch.empty(C_ * _lib.CLASS_STATS_PER_CLUSTER, dtype=torch.float64,
                                      device=dev)
            self.crem = torch.empty(max(int(n_ent.sum()), 16), dtype=torch.int16, device=dev)
            self.crem_off = torch.from_numpy(rem_off).to(dev)
            d.pixel_class, d.class_stats = ptr(self.pclass), ptr(self.cstats)
            d.class_rem, d.class_rem_off = ptr(self.crem), ptr(self.crem_off)
            nb = int(lib().pycwb_class_bound_scratch_bytes(d.block, d.grid))
            self.cbound = torch.empty(max(nb // 4, 1), dtype=torch.int32, device=dev)
            d.class_bounds, d.class_bounds_bytes = ptr(self.cbound), nb
        self.desc = d
        self.block, self.grid = d.block, d.grid
        h = C.c_void_p()
        _lib.check(lib().pycwb_batch_create(C.byref(d), C.byref(h)), "pycwb_batch_create")
        self.handle = h

    def launch(self, stream=None) -> None:
        _lib.check(lib().pycwb_batch_launch(self.handle, stream_ptr(stream)), "pycwb_batch_launch")

    def launch_noalloc(self, stream=None) -> None:
        """The same kernels through pycwb_batch_run_noalloc (library-held
        scratch, no allocation, no graph)."""
        _lib.check(lib().pycwb_batch_run_noalloc(C.byref(self.desc), stream_ptr(stream)),
                   "pycwb_batch_run_noalloc")

    def launch_stage(self, stage: int, stream=None) -> None:
        _lib.check(lib().pycwb_batch_launch_stage(self.handle, int(stage), stream_ptr(stream)),
                   "pycwb_batch_launch_stage")

    def results(self) -> dict:
        torch.cuda.current_stream().synchronize()
        return {"l_max": self.l_max.cpu().numpy(), "sky_stat_max": self.stat_max.cpu().numpy(),
                "REG1": self.REG1.cpu().numpy(), "regulator": self.reg_f64.cpu().numpy()}

    def active_pairs(self) -> np.ndarray:
        """Per cluster, the (direction, pixel) pairs above the energy threshold
        (those the scan evaluates in full).  Overwrites the regulator counts, so
        call it after reading results()."""
        self.launch_stage(0)
        self.launch_stage(4)
        torch.cuda.current_stream().synchronize()
        return self.counts.cpu().numpy().astype(np.uint32).astype(np.int64)

    def tested_pairs(self) -> np.ndarray:
        """Per cluster, the (direction, pixel) pairs whose energy the scan
        evaluates (the rest are ruled out by the tile energy bound without
        loading their amplitudes).  Overwrites the regulator counts and runs
        the scan once more with the regulator of the last launch."""
        self.launch_stage(0)
        self.launch_stage(5)
        torch.cuda.current_stream().synchronize()
        return self.counts.cpu().numpy().astype(np.uint32).astype(np.int64)

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            lib().pycwb_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainLoop:
    """The same clusters, one plain launch family (regulator, scan) each."""

    def __init__(self, batch: ClusterBatch, nets: NetTablesDev, valid: torch.Tensor,
                 gamma_regulator: float, delta_regulator: float, netCC: float,
                 flags: int | None = None, ordered: bool = True):
        dev = batch.td.device
        self.batch, self.nets, self.valid = batch, nets, valid
        C_ = batch.n_clusters
        self.gamma, self.netCC = float(gamma_regulator), float(netCC)
        self.reg0 = float(np.float32(delta_regulator * np.sqrt(2)))
        self.ws = torch.zeros(256, dtype=torch.uint8, device=dev)
        self.REG = torch.zeros((C_, 3), dtype=torch.float32, device=dev)
        self.reg_f64 = torch.zeros(C_, dtype=torch.float64, device=dev)
        self.l_max = torch.zeros(C_, dtype=torch.int64, device=dev)
        self.stat_max = torch.zeros(C_, dtype=torch.float64, device=dev)
        self.order = nets.sky_order(valid.cpu().numpy(), dev) if ordered else None
        if flags is None:
            finite = nets.finite() and bool(torch.isfinite(batch.rms).all())
            flags = _lib.DEFAULT_FLAGS if finite else _lib.REGULATOR_EARLY_EXIT
        self.flags = int(flags)
        offs = nets.off.cpu().numpy()
        self.net_views = []
        S = nets.n_sky
        for k, n in enumerate(nets.nifo):
            o = int(offs[k])
            self.net_views.append((nets.FP[o:o + S * n], nets.FX[o:o + S * n],
                                   nets.ml[o:o + S * n]))

    def launch(self, stream=None) -> None:
        L = lib()
        b, nets = self.batch, self.nets
        S, nv = nets.n_sky, int(self.valid.numel())
        st = stream_ptr(stream)
        vp = ptr(self.valid)
        ws, wsn = ptr(self.ws), self.ws.numel()
        for c in range(b.n_clusters):
            k = int(b.net[c])
            n, P, K = int(b.nifo[c]), int(b.n_pix[c]), int(b.K[c])
            FP, FX, ml = self.net_views[k]
            rms = b.rms.data_ptr() + 4 * int(b.rms_off[c])
            td = b.td.data_ptr() + 8 * int(b.td_off[c])
            REG = self.REG.data_ptr() + 12 * c
            rc = L.pycwb_dpf_regulator(n, P, S, ptr(FP), ptr(FX), rms, self.gamma,
                                       nets.En[k], self.reg0, vp, nv, self.flags, REG,
                                       self.reg_f64.data_ptr() + 8 * c, ws, wsn, st)
            _lib.check(rc, "pycwb_dpf_regulator")
            order = None if self.order is None else self.order[k].data_ptr()
            rc = L.pycwb_sky_scan_plain(n, P, S, 2 * K + 1, ptr(FP), ptr(FX), ptr(ml), rms, td,
                                        REG, self.netCC, nets.En[k], vp, nv, order, self.flags,
                                        None, self.l_max.data_ptr() + 8 * c,
                                        self.stat_max.data_ptr() + 8 * c, ws, wsn, st)
            _lib.check(rc, "pycwb_sky_scan_plain")

    def results(self) -> dict:
        torch.cuda.current_stream().synchronize()
        return {"l_max": self.l_max.cpu().numpy(), "sky_stat_max": self.stat_max.cpu().numpy(),
                "REG1": self.REG[:, 1].cpu().numpy(), "regulator": self.reg_f64.cpu().numpy()}
