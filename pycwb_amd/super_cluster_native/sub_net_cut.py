"""GPU sub-network cut with the call signatures of PycWB's super_cluster_native.

Drop-in for pycwb/modules/super_cluster_native/sub_net_cut.py:
  optimze_sky_loc_from_td        (:404-517)  -> pycwb_subnet_sky  (HIP)
  mra_statistics_from_td         (:659-724)  -> pycwb_subnet_mra  (HIP)
  sub_net_cut                    (:22-62)
  sub_net_cut_from_pixel_arrays  (:65-113)
  the decision of _sub_net_cut_prepared_packets (:206-267) and the result
  dictionary of _format_subnet_result (:270-294).
``SubnetBatch`` evaluates many clusters in one launch per detector count; it
backs ``apply_subnet_cut`` (utils.py), which runs every cluster of a lag as one
batch.

Flow per cluster (as in the reference): the sky loop on every cluster; the
cross-talk CSR is fetched and the multi-resolution statistics run only for
clusters whose ``suball`` exceeds ``subnet`` (``min(suball, submra) > subnet``
cannot pass otherwise).

Cross-talk: ``xtalk`` is any object with ``get_xtalk_pixels(pixels, check)``
or ``get_xtalk_pixels_from_arrays(layers, times, check)`` returning the CSR
``(lookup (P, 2) int32, xtalk (m, 8) float32)`` of pycwb/modules/xtalk; a row
must not list a neighbour twice.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from .. import sky as sky_grid
from ..batch import ClusterBatch
from ..device import lib, ptr, require_gpu, stream_ptr

SKY_FIELDS = ("l_max", "stat", "Em", "Am", "lm", "Vm", "suball", "EE")
MRA_FIELDS = ("submra", "rHo", "Eo", "Lo", "Ls", "m")


def calculate_e2or_from_acore(acore: float, n_ifo: int) -> float:
    """Sub-network energy threshold from Acore (pycwb/types/detector.py:1188-1218):
    p = Q(n, Acore^2 n), e2or = Q^-1(n - 1, p) with Q the regularised upper
    incomplete gamma function (cWB's network::setAcore)."""
    from scipy.special import gammaincc, gammainccinv

    n_ifo = int(max(1, n_ifo))
    if n_ifo < 2:
        return float(max(0.0, acore))
    prob = float(gammaincc(float(n_ifo), float(acore) * float(acore) * float(n_ifo)))
    prob = float(np.clip(prob, 1.0e-12, 1.0 - 1.0e-12))
    return float(gammainccinv(float(n_ifo - 1), prob))


def _add_timing(timing, key, elapsed):
    if timing is not None:
        timing[key] = timing.get(key, 0.0) + float(elapsed)


def _format_subnet_result(subnet_pass, subrho_pass, subthr_pass, suball, submra, rHo, Em,
                          subnet, subrho, subnorm, Eo=0.0) -> dict:
    """sub_net_cut.py:270-294."""
    return {
        "subnet_passed": subnet_pass,
        "subrho_passed": subrho_pass,
        "subthr_passed": subthr_pass,
        "subnet_condition": f"min(suball = {suball:.4f}, submra = {submra:.4f}) > "
                            f"subnet = {subnet:.4f}",
        "subrho_condition": f"rho = {rHo:.4f} > subrho = {subrho:.4f}",
        "subthr_condition": f"Em = {Em:.4f} > (subnorm = {subnorm:.4f} * Eo = {Eo:.4f})",
    }


def needs_mra(sky: tuple, subnet: float) -> bool:
    """The reference skips cross-talk and MRA when suball <= subnet (:254-259)."""
    return not (sky[6] <= subnet)


def subnet_decision(sky: tuple, mra: tuple | None, subnet: float, subrho: float,
                    subnorm: float) -> dict:
    """The pass/fail logic of _sub_net_cut_prepared_packets (:254-267) from the
    sky-loop tuple and the MRA tuple (None when skipped)."""
    l_max, stat, Em, Am, lm, Vm, suball, EE = sky
    if not needs_mra(sky, subnet):
        return _format_subnet_result(False, True, True, suball, 0.0, 0.0, Em, subnet, subrho,
                                     subnorm, 0.0)
    submra, rHo, Eo, Lo, Ls, m = mra
    return _format_subnet_result(min(suball, submra) > subnet, rHo > subrho, Em > subnorm * Eo,
                                 suball, submra, rHo, Em, subnet, subrho, subnorm, Eo)


@dataclass
class SubnetSky:
    """Sky tables of the sub-network cut's grid, all networks concatenated."""
    n_sky: int
    nifo: list
    off: torch.Tensor     # int64 (n_net,)
    FP: torch.Tensor      # float32 flat, per network (S, n)
    FX: torch.Tensor
    ml: torch.Tensor      # int32 flat, per network (n, S)
    ml_host: list
    En: torch.Tensor      # float32 (n_net,): float32(2 Acore^2 n)
    Es: torch.Tensor      # float32 (n_net,): float32(2 e2or)
    order: torch.Tensor = None   # int32 (n_net, S): directions sorted by delay tuple


def subnet_sky_tables(tables, En, Es, device=None, ordered: bool = True) -> SubnetSky:
    """tables: list of (FP (S, n), FX (S, n), ml (n, S)) per network; En, Es:
    per-network float32 thresholds.  ordered: lanes visit directions sorted by
    delay tuple (neighbouring lanes gather neighbouring taps); the results do
    not depend on it."""
    device = device or require_gpu()
    S = int(np.asarray(tables[0][0]).shape[0])
    offs, fp, fx, ml, nifo, mlh = [], [], [], [], [], []
    o = 0
    for FP, FX, m in tables:
        FP = np.ascontiguousarray(FP, np.float32)
        FX = np.ascontiguousarray(FX, np.float32)
        m = np.ascontiguousarray(m, np.int32)
        n = FP.shape[1]
        if FP.shape != (S, n) or FX.shape != (S, n) or m.shape != (n, S):
            raise ValueError(f"sky tables must be FP/FX (S, n) and ml (n, S); got "
                             f"{FP.shape}, {FX.shape}, {m.shape}")
        offs.append(o)
        fp.append(FP.ravel())
        fx.append(FX.ravel())
        ml.append(m.ravel())
        nifo.append(n)
        mlh.append(m)
        o += S * n
    t = lambda a, dt: torch.from_numpy(np.concatenate(a)).to(device=device, dtype=dt)
    return SubnetSky(n_sky=S, nifo=nifo, off=torch.tensor(offs, dtype=torch.int64, device=device),
                     FP=t(fp, torch.float32), FX=t(fx, torch.float32), ml=t(ml, torch.int32),
                     ml_host=mlh,
                     En=torch.tensor(np.asarray(En, np.float32), device=device),
                     Es=torch.tensor(np.asarray(Es, np.float32), device=device),
                     order=None if not ordered else torch.from_numpy(np.ascontiguousarray(
                         np.stack([sky_grid.delay_order(m, np.arange(S)) for m in mlh]),
                         np.int32)).to(device))


def cluster_batch_from_host(rms_list, td_list, net_ids, device=None) -> ClusterBatch:
    """rms_list[c]: (P, n) float32; td_list[c]: (P, n, 2K+1, 2) float32 {00, 90}."""
    device = device or require_gpu()
    n_pix, nifo, K, rms_len, td_len = [], [], [], [], []
    for r, t in zip(rms_list, td_list):
        P, n = r.shape
        if t.ndim != 4 or t.shape[:2] != (P, n) or t.shape[3] != 2 or t.shape[2] % 2 != 1:
            raise ValueError(f"td must be (P, n, 2K+1, 2) matching rms (P, n); got {t.shape}")
        n_pix.append(P)
        nifo.append(n)
        K.append((t.shape[2] - 1) // 2)
        rms_len.append(P * n)
        td_len.append(P * n * t.shape[2])
    rms_off = np.concatenate([[0], np.cumsum(rms_len)[:-1]]).astype(np.int64)
    td_off = np.concatenate([[0], np.cumsum(td_len)[:-1]]).astype(np.int64)
    rms = np.concatenate([np.ascontiguousarray(r, np.float32).ravel() for r in rms_list]
                         + [np.zeros(1, np.float32)])
    td = np.concatenate([np.ascontiguousarray(t, np.float32).ravel() for t in td_list]
                        + [np.zeros(2, np.float32)])
    return ClusterBatch(n_pix=np.asarray(n_pix, np.int32),
                        net=np.asarray(net_ids, np.int32).reshape(-1),
                        K=np.asarray(K, np.int32), nifo=np.asarray(nifo, np.int32),
                        rms_off=rms_off, td_off=td_off,
                        rms=torch.from_numpy(rms).to(device),
                        td=torch.from_numpy(td).to(device))


def _check_csr(lookup, xt, P):
    lookup = np.ascontiguousarray(lookup, np.int32).reshape(-1, 2)
    xt = np.ascontiguousarray(xt, np.float32).reshape(-1, 8)
    if lookup.shape[0] != P:
        raise ValueError(f"cross-talk lookup has {lookup.shape[0]} rows for {P} pixels")
    s, e = lookup[:, 0], lookup[:, 1]
    if P and (s.min() < 0 or (e < s).any() or e.max() > xt.shape[0]):
        raise ValueError("cross-talk lookup rows out of range")
    q = xt[:, 0]
    if xt.shape[0] and (q.min() < 0 or q.max() >= P or (q != np.floor(q)).any()):
        raise ValueError("cross-talk neighbour index out of range")
    counts = (e - s).astype(np.int64)
    rows = np.repeat(np.arange(P, dtype=np.int64), counts)
    if rows.size:
        first = np.cumsum(counts) - counts
        entry = np.arange(rows.size, dtype=np.int64) - np.repeat(first, counts) + \
            np.repeat(s.astype(np.int64), counts)
        cols = q[entry].astype(np.int64)
        if np.unique(rows * P + cols).size != rows.size:
            raise ValueError("a cross-talk row lists a neighbour twice")
    return lookup, xt


class SubnetBatch:
    """Sub-network cut of the clusters of a ClusterBatch on the subnet sky grid."""

    def __init__(self, batch: ClusterBatch, sky: SubnetSky, subcut: float, device=None):
        self.device = device or require_gpu()
        self.batch, self.skyt, self.subcut = batch, sky, float(subcut)
        nC = batch.n_clusters
        for c in range(nC):
            n, K = int(batch.nifo[c]), int(batch.K[c])
            k = int(batch.net[c])
            if not 0 <= k < len(sky.nifo) or sky.nifo[k] != n:
                raise ValueError(f"cluster {c}: network {k} does not have {n} detectors")
            if np.abs(sky.ml_host[k]).max(initial=0) > K:
                raise ValueError(f"cluster {c}: sky delays exceed the {2 * K + 1} TD taps")
        d = self.device
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(d)
        i64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int64)).to(d)
        self.n_pix_d, self.net_d, self.K_d = i32(batch.n_pix), i32(batch.net), i32(batch.K)
        self.rms_off_d, self.td_off_d = i64(batch.rms_off), i64(batch.td_off)
        self.groups = {}
        for n in sorted(set(int(x) for x in batch.nifo)):
            self.groups[n] = np.flatnonzero(batch.nifo == n).astype(np.int32)
        self.n_tiles = int(lib().pycwb_subnet_tiles(sky.n_sky))
        self.sky_d = torch.zeros(max(nC, 1) * 8, dtype=torch.float64, device=d)
        self.mra_d = torch.zeros(max(nC, 1) * 8, dtype=torch.float64, device=d)

    def _desc(self, n, which_d, n_which):
        b, s = self.batch, self.skyt
        desc = _lib.SubnetDesc()
        desc.n_which, desc.which, desc.n_ifo, desc.n_sky = n_which, ptr(which_d), n, s.n_sky
        desc.n_pix, desc.net, desc.K = ptr(self.n_pix_d), ptr(self.net_d), ptr(self.K_d)
        desc.rms_off, desc.td_off = ptr(self.rms_off_d), ptr(self.td_off_d)
        desc.rms, desc.td = ptr(b.rms), ptr(b.td)
        desc.net_off, desc.net_En, desc.net_Es = ptr(s.off), ptr(s.En), ptr(s.Es)
        desc.FP, desc.FX, desc.ml = ptr(s.FP), ptr(s.FX), ptr(s.ml)
        desc.subcut = self.subcut
        desc.sky, desc.mra = ptr(self.sky_d), ptr(self.mra_d)
        desc.sky_order = ptr(s.order)
        return desc

    def launch_sky(self, stream=None) -> list:
        """Enqueue the sky loop of every cluster on ``stream`` (default: the
        current stream) without synchronising; returns the temporaries that
        must stay alive until the stream has run them."""
        st = stream_ptr(stream)
        keep = []
        for n, idx in self.groups.items():
            for a in range(0, idx.size, 65535):
                part = idx[a:a + 65535]
                which = torch.from_numpy(part).to(self.device)
                tiles = torch.empty(part.size * self.n_tiles * 8, dtype=torch.float64,
                                    device=self.device)
                desc = self._desc(n, which, part.size)
                desc.tiles = ptr(tiles)
                _lib.check(lib().pycwb_subnet_sky(C.byref(desc), st), "pycwb_subnet_sky")
                keep.append((which, tiles))
        return keep

    def sky_results(self, stream=None) -> np.ndarray:
        """(C, 8) float64: the sky-loop tuple of every cluster."""
        stream = stream or torch.cuda.current_stream()
        keep = self.launch_sky(stream)
        stream.synchronize()
        del keep
        return self.sky_d.view(-1, 8)[:self.batch.n_clusters].cpu().numpy()

    def prepare_mra(self, clusters, csrs) -> dict:
        """Validate and upload the cross-talk CSRs of the given clusters
        (csrs[k] = (lookup, xtalk) of clusters[k]), or take the device CSRs of
        the whole batch as pycwb_amd.xtalk.XTalk.build_batch returns them
        (an XTalkCSR, used in place: its rows were built valid); the returned
        handle feeds launch_mra."""
        clusters = np.asarray(clusters, np.int64).reshape(-1)
        b, d = self.batch, self.device
        if clusters.size and (clusters.min() < 0 or clusters.max() >= b.n_clusters):
            raise ValueError("cluster index out of range")
        lk_off = np.zeros(b.n_clusters, np.int64)
        xt_off = np.zeros(b.n_clusters, np.int64)
        scr_off = np.zeros(b.n_clusters, np.int64)
        t = lambda a: torch.from_numpy(a).to(d)
        scratch_of = {}

        def scratch(n, P):
            if (n, P) not in scratch_of:
                scratch_of[(n, P)] = int(lib().pycwb_subnet_mra_scratch(n, P))
            return scratch_of[(n, P)]

        if hasattr(csrs, "pix_off") and hasattr(csrs, "xt_off"):
            if csrs.pix_off.size != b.n_clusters + 1 or \
                    np.any(np.diff(csrs.pix_off) != b.n_pix):
                raise ValueError("the device CSRs must cover the batch's clusters in order")
            lk_off[:] = csrs.pix_off[:-1]
            xt_off[:] = csrs.xt_off[:-1]
            lk_d, xt_d = csrs.lookup, csrs.xt
            so = 0
            for c in clusters:
                scr_off[c] = so
                so += scratch(int(b.nifo[c]), int(b.n_pix[c]))
        else:
            lk_parts, xt_parts = [], []
            lo = xo = so = 0
            for c, (lookup, xt) in zip(clusters, csrs):
                P, n = int(b.n_pix[c]), int(b.nifo[c])
                lookup, xt = _check_csr(lookup, xt, P)
                lk_parts.append(lookup.ravel())
                xt_parts.append(xt.ravel())
                lk_off[c], xt_off[c], scr_off[c] = lo, xo, so
                lo += P
                xo += xt.shape[0]
                so += scratch(n, P)
            lk_d = t(np.concatenate(lk_parts + [np.zeros(2, np.int32)]))
            xt_d = t(np.concatenate(xt_parts + [np.zeros(8, np.float32)]))
        nifo = b.nifo[clusters]
        groups = {n: t(clusters[nifo == n].astype(np.int32))
                  for n in sorted(set(int(x) for x in nifo))}
        return {"clusters": clusters, "groups": groups, "lk": lk_d, "xt": xt_d,
                "lk_off": t(lk_off), "xt_off": t(xt_off), "scr_off": t(scr_off),
                "scratch": torch.empty(max(so, 1), dtype=torch.float32, device=d)}

    def launch_mra(self, h: dict, stream=None) -> None:
        """Enqueue the MRA statistics of a prepared handle (keep the handle alive
        until the stream has run it)."""
        st = stream_ptr(stream)
        for n, which in h["groups"].items():
            desc = self._desc(n, which, int(which.numel()))
            desc.lk_off, desc.lookup = ptr(h["lk_off"]), ptr(h["lk"])
            desc.xt_off, desc.xt = ptr(h["xt_off"]), ptr(h["xt"])
            desc.scratch, desc.scr_off = ptr(h["scratch"]), ptr(h["scr_off"])
            _lib.check(lib().pycwb_subnet_mra(C.byref(desc), st), "pycwb_subnet_mra")

    def mra_results(self, clusters, csrs, stream=None) -> np.ndarray:
        """Multi-resolution statistics of the given clusters at their sky-loop
        l_max (run sky_results first); csrs[k] = (lookup, xtalk) of clusters[k].
        Returns (len(clusters), 8) float64."""
        clusters = np.asarray(clusters, np.int64).reshape(-1)
        if clusters.size == 0:
            return np.zeros((0, 8))
        h = self.prepare_mra(clusters, csrs)
        stream = stream or torch.cuda.current_stream()
        self.launch_mra(h, stream)
        stream.synchronize()
        del h
        return self.mra_d.view(-1, 8).cpu().numpy()[clusters]


def sky_tuple(row) -> tuple:
    return (int(row[0]), float(row[1]), float(row[2]), float(row[3]), int(row[4]), int(row[5]),
            float(row[6]), float(row[7]))


def mra_tuple(row) -> tuple:
    return (float(row[0]), float(row[1]), float(row[2]), float(row[3]), float(row[4]),
            int(row[5]))


def _pack_td(td00, td90):
    """Reference layout (nd, n, P) x 2 -> (P, n, nd, 2)."""
    td00 = np.asarray(td00, np.float32)
    td90 = np.asarray(td90, np.float32)
    if td00.shape != td90.shape or td00.ndim != 3:
        raise ValueError("td00/td90 must both be (2K+1, n_ifo, n_pix)")
    if td00.shape[0] % 2 != 1:
        raise ValueError(f"TD tables must hold 2K+1 delays, got {td00.shape[0]}")
    return np.ascontiguousarray(np.stack([td00, td90], axis=-1).transpose(2, 1, 0, 3))


def _single(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, En, e2or, subcut):
    dev = require_gpu()
    FP = np.asarray(FP, np.float32)
    FX = np.asarray(FX, np.float32)
    ml = np.asarray(ml, np.int32)
    if FP.shape != (n_sky, n_ifo) or FX.shape != (n_sky, n_ifo) or ml.shape != (n_ifo, n_sky):
        raise ValueError("FP/FX must be (n_sky, n_ifo) and ml (n_ifo, n_sky)")
    rms = np.asarray(rms, np.float32)
    if rms.shape != (n_pix, n_ifo):
        raise ValueError(f"rms must be ({n_pix}, {n_ifo}), got {rms.shape}")
    packed = _pack_td(td00, td90)
    if packed.shape[:2] != (n_pix, n_ifo):
        raise ValueError("td00/td90 do not match (n_ifo, n_pix)")
    b = cluster_batch_from_host([rms], [packed], [0], dev)
    s = subnet_sky_tables([(FP, FX, ml)], [np.float32(En)], [np.float32(2 * e2or)], dev)
    return SubnetBatch(b, s, subcut, dev)


def optimze_sky_loc_from_td(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml,
                            network_energy_threshold, e2or, subcut):
    """sub_net_cut.py:404-517 on the GPU: (l_max, stat, Em, Am, lm, Vm, suball, EE).
    FP/FX (n_sky, n_ifo) float32, rms (n_pix, n_ifo), td00/td90 (2K+1, n_ifo,
    n_pix), ml (n_ifo, n_sky) with |ml| <= K."""
    sb = _single(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, network_energy_threshold,
                 e2or, subcut)
    return sky_tuple(sb.sky_results()[0])


def mra_statistics_from_td(n_ifo, n_pix, FP, FX, rms, td00, td90, ml,
                           network_energy_threshold, e2or, subcut, xtalks, xtalks_lookup, l_max):
    """sub_net_cut.py:659-724 on the GPU: (submra, rHo, Eo, Lo, Ls, m)."""
    n_sky = int(np.asarray(FP).shape[0])
    sb = _single(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, network_energy_threshold,
                 e2or, subcut)
    if not 0 <= int(l_max) < n_sky:
        raise ValueError(f"l_max {l_max} outside the {n_sky}-direction grid")
    sb.sky_d[0] = float(l_max)
    return mra_tuple(sb.mra_results([0], [(xtalks_lookup, xtalks)])[0])


# ---------------------------------------------------------------------------
# cluster-level entry points

def cluster_inputs(pixels, rows=None):
    """(rms (P, n) float32, td (P, n, 2K+1, 2) float32) of the selected rows of a
    PixelArrays (_load_selected_pixel_arrays, sub_net_cut.py:119-140)."""
    n = int(getattr(pixels, "_n_ifo", 0) or pixels.n_ifo)
    rows = np.arange(len(pixels), dtype=np.int64) if rows is None else \
        np.asarray(rows, dtype=np.int64)
    inv = 1.0 / np.asarray(pixels.noise_rms)[:, rows].astype(np.float64)
    rms_pix = 1.0 / np.sqrt(np.sum(inv ** 2, axis=0))
    rms = np.ascontiguousarray((inv * rms_pix[np.newaxis, :]).T, dtype=np.float32)
    P = rows.size
    offsets = np.asarray(pixels.td_amp_offsets)
    flat = np.asarray(pixels.td_amp_flat, np.float32)
    n_rows = len(pixels) * n
    if n_rows == 0 or flat.size == 0:
        return rms, np.zeros((P, n, 0, 2), np.float32)
    sizes = offsets[1:n_rows + 1] - offsets[:n_rows]
    tsize = int(sizes[0])
    if not np.all(sizes == tsize):
        raise ValueError("td_amp vectors have non-uniform lengths; subnet cut requires dense "
                         "TD amplitudes")
    t2 = tsize // 2
    start = offsets[(rows[:, None] * n + np.arange(n)[None, :]).ravel()].astype(np.int64)
    idx = start[:, None] + np.arange(t2)[None, :]
    td = np.stack([flat[idx], flat[idx + t2]], axis=-1).reshape(P, n, t2, 2)
    return rms, np.ascontiguousarray(td)


def _xtalk_csr(xtalk, pixels, layers, times):
    if pixels is not None and hasattr(xtalk, "get_xtalk_pixels"):
        return xtalk.get_xtalk_pixels(pixels, True)
    if hasattr(xtalk, "get_xtalk_pixels_from_arrays"):
        return xtalk.get_xtalk_pixels_from_arrays(layers, times, True)
    raise TypeError("xtalk must provide get_xtalk_pixels(pixels, check) or "
                    "get_xtalk_pixels_from_arrays(layers, times, check)")


def _prepared_tables(ml, FP, FX, arrays_prepared):
    ml_p = np.ascontiguousarray(ml, dtype=np.int32)
    if arrays_prepared:
        return ml_p, np.ascontiguousarray(FP, np.float32), np.ascontiguousarray(FX, np.float32)
    return ml_p, np.ascontiguousarray(np.asarray(FP).T, np.float32), \
        np.ascontiguousarray(np.asarray(FX).T, np.float32)


def sub_net_cut_from_pixel_arrays(pixels, pixel_indices, ml, FP, FX, acor, e2or, n_ifo, n_sky,
                                  subnet, subcut, subnorm, subrho, xtalk,
                                  arrays_prepared: bool = False, timing: dict | None = None):
    """sub_net_cut.py:65-113 (one cluster; see ``apply_subnet_cut`` for batches)."""
    t_total = time.perf_counter()
    rows = np.arange(len(pixels), dtype=np.int64) if pixel_indices is None else \
        np.asarray(pixel_indices, dtype=np.int64)
    if rows.size == 0:
        _add_timing(timing, "total", time.perf_counter() - t_total)
        return _format_subnet_result(False, False, False, 0.0, 0.0, 0.0, 0.0, subnet, subrho,
                                     subnorm)
    t = time.perf_counter()
    rms, td = cluster_inputs(pixels, rows)
    ml_p, FP_p, FX_p = _prepared_tables(ml, FP, FX, arrays_prepared)
    _add_timing(timing, "data_prep", time.perf_counter() - t)
    res = evaluate_subnet_cut([rms], [td], ml_p, FP_p, FX_p, acor, e2or, n_ifo, subnet, subcut,
                              subnorm, subrho,
                              lambda k: _xtalk_csr(xtalk, None, np.asarray(pixels.layers)[rows],
                                                   np.asarray(pixels.time)[rows]),
                              timing=timing)[0]
    _add_timing(timing, "total", time.perf_counter() - t_total)
    return res


def sub_net_cut(pixels, ml, FP, FX, acor, e2or, n_ifo, n_sky, subnet, subcut, subnorm, subrho,
                xtalk, timing: dict | None = None):
    """sub_net_cut.py:22-62 for PixelArrays input (FP/FX (n_ifo, n_sky))."""
    t_total = time.perf_counter()
    if len(pixels) == 0:
        _add_timing(timing, "total", time.perf_counter() - t_total)
        return _format_subnet_result(False, False, False, 0.0, 0.0, 0.0, 0.0, subnet, subrho,
                                     subnorm)
    t = time.perf_counter()
    rms, td = cluster_inputs(pixels)
    ml_p, FP_p, FX_p = _prepared_tables(ml, FP, FX, False)
    _add_timing(timing, "data_prep", time.perf_counter() - t)
    res = evaluate_subnet_cut([rms], [td], ml_p, FP_p, FX_p, acor, e2or, n_ifo, subnet, subcut,
                              subnorm, subrho, lambda k: _xtalk_csr(xtalk, pixels, None, None),
                              timing=timing)[0]
    _add_timing(timing, "total", time.perf_counter() - t_total)
    return res


def evaluate_subnet_cut(rms_list, td_list, ml, FP, FX, acor, e2or, n_ifo, subnet, subcut,
                        subnorm, subrho, csr_of, timing: dict | None = None,
                        details: list | None = None) -> list[dict]:
    """Sub-network cut of many clusters of one network as one GPU batch.

    rms_list[c] (P, n) float32, td_list[c] (P, n, 2K+1, 2) float32; FP/FX
    (n_sky, n), ml (n, n_sky); csr_of(c) returns cluster c's cross-talk CSR and
    is called only for clusters that need the MRA statistics.  Returns the
    reference's result dictionary per cluster; ``details`` (if a list) receives
    (sky tuple, MRA tuple or None) per cluster."""
    En = np.float32(2 * acor * acor * n_ifo)
    Es = np.float32(2 * e2or)
    dev = require_gpu()
    t = time.perf_counter()
    b = cluster_batch_from_host(rms_list, td_list, np.zeros(len(rms_list), np.int32), dev)
    s = subnet_sky_tables([(FP, FX, ml)], [En], [Es], dev)
    sb = SubnetBatch(b, s, subcut, dev)
    sky = [sky_tuple(r) for r in sb.sky_results()]
    _add_timing(timing, "sky_scan", time.perf_counter() - t)
    need = [c for c in range(len(sky)) if needs_mra(sky[c], subnet)]
    mra = {}
    if need:
        t = time.perf_counter()
        csrs = [csr_of(c) for c in need]
        _add_timing(timing, "xtalk", time.perf_counter() - t)
        t = time.perf_counter()
        rows = sb.mra_results(need, csrs)
        mra = {c: mra_tuple(r) for c, r in zip(need, rows)}
        _add_timing(timing, "mra", time.perf_counter() - t)
    out = []
    for c in range(len(sky)):
        out.append(subnet_decision(sky[c], mra.get(c), subnet, subrho, subnorm))
        if details is not None:
            details.append((sky[c], mra.get(c)))
    return out
