"""Best-sky statistics on the GPU for a batch of clusters.

Device form of compute_statistics_at_sky_position
(pycwb/modules/likelihoodWP/sky_statistics.py:26-290): one workgroup per
cluster evaluates Lm, Ec, Np, rho, ... and the reconstructed TF amplitudes at
the cluster's selected direction (HIP kernel ``pycwb_sky_statistics``).

Cross-talk: each cluster carries a CSR ``(lookup (P, 2) int32, xt (m, 8)
float32)`` in the layout of pycwb/modules/xtalk/monster.py:243-271 (column 0
= neighbour pixel index, columns 4..7 = coefficients).  Without a catalog the
self-only CSR is used (each pixel its own neighbour, coefficients 1, 0, 0, 1),
for which every cross-talk sum reduces to the per-detector energy p^2 + q^2.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .device import lib, ptr, stream_ptr

SCALARS = ("Gn", "Ec", "Dc", "Rc", "Eh", "Es", "Np", "Em", "Lm", "norm", "cc", "rho", "xrho",
           "Lo", "Eo", "N_pix_effective")

# float32 scratch blocks, each (n, P) unless noted
_S32 = ("v00", "v90", "p", "q", "pd32", "pD32", "ps32", "pS32", "noise_amplitude_00",
        "noise_amplitude_90", "f", "F")
_S32P = ("et", "pixel_mask", "energy_array_plus", "energy_array_cross", "coherent_energy",
         "gaussian_noise_correction", "fp", "fx", "p00_POL_r", "p00_POL_a", "p90_POL_r",
         "p90_POL_a")
_S64 = ("q_norm", "p_norm", "pd", "pD", "ps", "pS")


def self_xtalk(n_pix: int):
    lookup = np.stack([np.arange(n_pix), np.arange(1, n_pix + 1)], axis=1).astype(np.int32)
    xt = np.zeros((n_pix, 8), dtype=np.float32)
    xt[:, 0] = np.arange(n_pix)
    xt[:, 4] = 1.0
    xt[:, 7] = 1.0
    return lookup, xt


class SkyStatistics:
    """Best-sky statistics of every cluster of a ClusterBatch."""

    def __init__(self, batch, nets, xtalk=None):
        dev = batch.td.device
        self.batch, self.nets = batch, nets
        Cn = batch.n_clusters
        L = lib()
        n, P = batch.nifo.astype(np.int64), batch.n_pix.astype(np.int64)
        s32 = np.array([L.pycwb_sky_stats_scratch32(int(a), int(b)) for a, b in zip(n, P)])
        s64 = np.array([L.pycwb_sky_stats_scratch64(int(a), int(b)) for a, b in zip(n, P)])
        self.s32_off = np.concatenate([[0], np.cumsum(s32)[:-1]]).astype(np.int64)
        self.s64_off = np.concatenate([[0], np.cumsum(s64)[:-1]]).astype(np.int64)
        self.s32 = torch.zeros(int(s32.sum()), dtype=torch.float32, device=dev)
        self.s64 = torch.zeros(int(s64.sum()), dtype=torch.float64, device=dev)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)
        if hasattr(xtalk, "pix_off") and hasattr(xtalk, "xt_off"):
            # device-resident CSRs of the whole batch (pycwb_amd.xtalk.XTalkCSR)
            if xtalk.pix_off.size != Cn + 1 or np.any(np.diff(xtalk.pix_off) != P):
                raise ValueError("cross-talk CSR does not match the batch's clusters")
            self.lk_off = xtalk.pix_off[:-1].astype(np.int64)
            self.xt_off = xtalk.xt_off[:-1].astype(np.int64)
            self.d_lookup, self.d_xt = xtalk.lookup, xtalk.xt
        if xtalk is None:
            lks, xts = zip(*(self_xtalk(int(p)) for p in P))
        elif hasattr(xtalk, "pix_off"):
            lks = xts = None
        else:
            lks, xts = zip(*xtalk)
            for c, (lk, xt) in enumerate(xtalk):
                lk, xt = np.asarray(lk), np.asarray(xt)
                if lk.shape != (P[c], 2):
                    raise ValueError(f"cluster {c}: xtalk lookup must have shape (n_pix, 2)")
                if lk.size and (lk.min() < 0 or lk.max() > len(xt) or
                                np.any(lk[:, 1] < lk[:, 0])):
                    raise ValueError(f"cluster {c}: xtalk lookup rows out of range")
                if len(xt) and (xt[:, 0].min() < 0 or xt[:, 0].max() >= P[c]):
                    raise ValueError(f"cluster {c}: xtalk neighbour index out of range")
        if lks is not None:
            lk_len = np.array([len(a) for a in lks], dtype=np.int64)
            xt_len = np.array([len(a) for a in xts], dtype=np.int64)
            self.lk_off = np.concatenate([[0], np.cumsum(lk_len)[:-1]]).astype(np.int64)
            self.xt_off = np.concatenate([[0], np.cumsum(xt_len)[:-1]]).astype(np.int64)
            lk_all = np.concatenate([np.asarray(a, np.int32).reshape(-1, 2) for a in lks])
            xt_all = np.concatenate([np.asarray(a, np.float32).reshape(-1, 8) for a in xts]) \
                if xt_len.sum() else np.zeros((1, 8), np.float32)
            self.d_lookup, self.d_xt = t(lk_all, torch.int32), t(xt_all, torch.float32)
        self.d_lk_off, self.d_xt_off = t(self.lk_off, torch.int64), t(self.xt_off, torch.int64)
        self.d_s32_off, self.d_s64_off = t(self.s32_off, torch.int64), t(self.s64_off, torch.int64)
        self.d_npix, self.d_nifo = t(batch.n_pix, torch.int32), t(batch.nifo, torch.int32)
        self.d_net, self.d_K = t(batch.net, torch.int32), t(batch.K, torch.int32)
        self.d_rms_off, self.d_td_off = t(batch.rms_off, torch.int64), t(batch.td_off, torch.int64)
        self.scal = torch.zeros((Cn, 16), dtype=torch.float64, device=dev)
        self.S_snr = torch.zeros((Cn, 8), dtype=torch.float64, device=dev)

    def run(self, l_max: torch.Tensor, REG1: torch.Tensor, reg0: float, xgb_rho_mode=False,
            which=None, stream=None) -> None:
        b, nets = self.batch, self.nets
        d = _lib.StatsDesc()
        d.n_clusters = b.n_clusters if which is None else int(which.numel())
        d.which = ptr(which)
        d.n_sky = nets.n_sky
        d.n_pix, d.nifo, d.net, d.K = ptr(self.d_npix), ptr(self.d_nifo), ptr(self.d_net), \
            ptr(self.d_K)
        d.rms_off, d.td_off, d.rms, d.td = ptr(self.d_rms_off), ptr(self.d_td_off), \
            ptr(b.rms), ptr(b.td)
        d.net_off, d.net_En = ptr(nets.off), ptr(nets.En_d)
        d.FP, d.FX, d.ml = ptr(nets.FP), ptr(nets.FX), ptr(nets.ml)
        self._lmax, self._reg = l_max.contiguous(), REG1.contiguous()
        d.l_max, d.REG1, d.reg0 = ptr(self._lmax), ptr(self._reg), float(reg0)
        d.lk_off, d.lookup, d.xt_off, d.xt = ptr(self.d_lk_off), ptr(self.d_lookup), \
            ptr(self.d_xt_off), ptr(self.d_xt)
        d.xgb_rho_mode = 1 if xgb_rho_mode else 0
        d.s32, d.s32_off, d.s64, d.s64_off = ptr(self.s32), ptr(self.d_s32_off), \
            ptr(self.s64), ptr(self.d_s64_off)
        d.scal, d.S_snr = ptr(self.scal), ptr(self.S_snr)
        _lib.check(lib().pycwb_sky_statistics(C.byref(d), stream_ptr(stream)),
                   "pycwb_sky_statistics")

    def scalars(self) -> dict:
        torch.cuda.current_stream().synchronize()
        s = self.scal.cpu().numpy()
        return {k: s[:, i] for i, k in enumerate(SCALARS)}

    def cluster_arrays(self, c: int) -> dict:
        """Per-pixel arrays of cluster c (host copies)."""
        torch.cuda.current_stream().synchronize()
        n, P = int(self.batch.nifo[c]), int(self.batch.n_pix[c])
        out = {}
        o = int(self.s32_off[c])
        for k in _S32:
            out[k] = self.s32[o:o + n * P].view(n, P).cpu().numpy()
            o += n * P
        out["f"] = out["f"].reshape(P, n)
        out["F"] = out["F"].reshape(P, n)
        for k in _S32P:
            out[k] = self.s32[o:o + P].cpu().numpy()
            o += P
        o = int(self.s64_off[c])
        for k in _S64:
            out[k] = self.s64[o:o + n * P].view(n, P).cpu().numpy()
            o += n * P
        out["S_snr"] = self.S_snr[c, :n].cpu().numpy()
        return out


def statistics_at_directions(batch, nets_host, theta, phi, REG1, reg0: float, En,
                             xgb_rho_mode=False) -> dict:
    """Best-sky statistics (Lm = Em - Np - Gn, Ec, rho, ...) of every cluster at
    a continuous direction (theta = colatitude, phi), on the GPU.

    The direction enters as the sky-gradient objective sees it: each detector's
    delay tau_i = (R_i - R_ref) . n / c * rate between the two neighbouring
    taps, linearly interpolated (clamped to the table), and F+/Fx at psi = 0
    with gha = -phi, rounded to float32 like the grid's tables.  The
    interpolated amplitudes become a one-tap cluster table and each cluster a
    one-direction network, so the statistics are those of the best-sky kernel
    (compute_statistics_at_sky_position, sky_statistics.py:26-290) unchanged.
    ``nets_host[batch.net[c]]`` gives each cluster's detectors and tap rate;
    ``En`` its network energy threshold per network."""
    from .batch import ClusterBatch, NetTablesDev
    from .skygrad import network_geometry

    dev = batch.td.device
    Cn = batch.n_clusters
    theta = np.asarray(theta, dtype=np.float64).reshape(Cn)
    phi = np.asarray(phi, dtype=np.float64).reshape(Cn)
    geos = {}
    fps, fxs, parts = [], [], []
    n_arr = batch.nifo.astype(np.int64)
    for c in range(Cn):
        k = int(batch.net[c])
        if k not in geos:
            geos[k] = network_geometry(nets_host[k].ifos, nets_host[k].td_rate)
        n, K = int(n_arr[c]), int(batch.K[c])
        g = geos[k][:n]
        st, ct, sp, cp = math.sin(theta[c]), math.cos(theta[c]), math.sin(phi[c]), \
            math.cos(phi[c])
        nh = np.array([st * cp, st * sp, ct])
        x = np.array([sp, -cp, 0.0])
        y = np.array([-cp * ct, -sp * ct, st])
        D = g[:, 3:].reshape(n, 3, 3)
        fps.append((np.einsum("a,iab,b->i", x, D, x) - np.einsum("a,iab,b->i", y, D, y))
                   .astype(np.float32))
        fxs.append((np.einsum("a,iab,b->i", x, D, y) + np.einsum("a,iab,b->i", y, D, x))
                   .astype(np.float32))
        u = np.clip(g[:, :3] @ nh + K, 0.0, 2 * K - 1e-9)
        i0 = np.floor(u).astype(np.int64)
        w = torch.from_numpy(u - i0).to(dev)[None, :, None]
        tdc = batch.cluster_td(c).to(torch.float64)               # (P, n, 2K+1, 2)
        idx = torch.from_numpy(i0).to(dev)
        a = tdc[:, torch.arange(n, device=dev), idx]              # (P, n, 2)
        b = tdc[:, torch.arange(n, device=dev), idx + 1]
        parts.append(((1.0 - w) * a + w * b).to(torch.float32).reshape(-1))
    P = batch.n_pix.astype(np.int64)
    td_off = np.concatenate([[0], np.cumsum(P * n_arr)[:-1]]).astype(np.int64)
    one = ClusterBatch(n_pix=batch.n_pix, net=np.arange(Cn, dtype=np.int32),
                       K=np.zeros(Cn, np.int32), nifo=batch.nifo, rms_off=batch.rms_off,
                       td_off=td_off, rms=batch.rms, td=torch.cat(parts))
    En_c = np.asarray([En[int(batch.net[c])] for c in range(Cn)], dtype=np.float64)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)
    nets = NetTablesDev(
        n_sky=1, nifo=n_arr.tolist(), En=En_c.tolist(),
        off=t(np.concatenate([[0], np.cumsum(n_arr)[:-1]]), torch.int64),
        FP=t(np.concatenate(fps), torch.float32), FX=t(np.concatenate(fxs), torch.float32),
        ml=torch.zeros(int(n_arr.sum()), dtype=torch.int32, device=dev),
        nifo_d=t(n_arr, torch.int32), En_d=t(En_c, torch.float64))
    st_ = SkyStatistics(one, nets)
    st_.run(torch.zeros(Cn, dtype=torch.int64, device=dev),
            torch.as_tensor(np.asarray(REG1, np.float32), device=dev), reg0, xgb_rho_mode)
    return st_.scalars()
