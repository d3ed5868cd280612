"""Differentiable sky statistic and off-grid sky refinement.

The grid scan (pycwb/modules/likelihoodWP/sky_scan.py:17-192) evaluates the
statistic AA = L * Co only at HEALPix directions, with delays rounded to whole
taps (pycwb/types/detector.py:1299-1303).  This module extends AA to any
direction (theta = colatitude, phi = Earth-fixed longitude):

* delays tau_i = (R_i - R_ref) . n(theta, phi) * rate / c, linearly
  interpolated between the two neighbouring taps of each td table;
* F+ / Fx from the detector response tensor at psi = 0, gha = -phi
  (detector.py:272-378), the same contraction that fills the grid tables;
* the scan's float64 pixel chain with the cluster's regulator fixed and no
  netCC gate.

Branches (energy mask, max(., 0), DPF sign choices, interpolation cell) are
those of the evaluation point.  The HIP kernel returns AA with dAA/dtheta and
dAA/dphi in one pass (float64 tangents), and on request dAA/d(amplitudes) in
float32, recomputing the delayed amplitudes from the td tables rather than
storing them.

``SkyStatistic`` wraps the kernel as a ``torch.autograd.Function`` of
(theta, phi, td); ``refine_sky`` runs projected gradient ascent on the sphere
from the grid argmax.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .batch import ClusterBatch, NetTablesDev
from .device import lib, ptr, stream_ptr
from .sky import C_LIGHT, Detector

MAX_IFO = 5


def network_geometry(ifos, rate: float, ref=None) -> np.ndarray:
    """(8, 12) float64: per detector (R_i - R_ref) * rate / c, then the
    response tensor row-major; unused rows zero."""
    dets = [Detector(i) if isinstance(i, str) else i for i in ifos]
    if not 1 <= len(dets) <= MAX_IFO:
        raise ValueError(f"1..{MAX_IFO} detectors supported, got {len(dets)}")
    ref_name = ref if ref is not None else dets[0].name
    ref_v = next((d.vertex for d in dets if d.name == ref_name), dets[0].vertex)
    g = np.zeros((8, 12))
    for i, d in enumerate(dets):
        g[i, :3] = (d.vertex - ref_v) / C_LIGHT * float(rate)
        g[i, 3:] = d.response.ravel()
    return g


def grid_angles(net, l):
    """(theta, phi) of grid pixel(s) ``l`` of a network's sky tables."""
    l = np.asarray(l)
    return np.pi / 2 - net.lat[l], net.phi[l]


class SkyGradient:
    """AA and its gradients for every cluster of a ClusterBatch.

    ``nets_host`` are the networks the batch's ``net`` indices refer to (each
    with ``ifos`` and ``td_rate``; the first detector is the delay reference);
    ``REG1`` the per-cluster regulator from the grid scan and ``reg0`` the
    float32 delta * sqrt(2) it used.
    """

    def __init__(self, batch: ClusterBatch, nets: NetTablesDev, nets_host, REG1, reg0: float):
        dev = batch.td.device
        self.batch, self.nets = batch, nets
        self.C = batch.n_clusters
        self.geo = torch.from_numpy(np.stack([network_geometry(n.ifos, n.td_rate)
                                              for n in nets_host])).to(dev)
        self.REG1 = torch.as_tensor(REG1, device=dev).to(torch.float64).contiguous()
        if self.REG1.numel() != self.C:
            raise ValueError("REG1 needs one value per cluster")
        self.reg0 = float(reg0)
        self.d_npix = torch.from_numpy(batch.n_pix).to(dev)
        self.d_net = torch.from_numpy(batch.net).to(dev)
        self.d_K = torch.from_numpy(batch.K).to(dev)
        self.d_rms_off = torch.from_numpy(batch.rms_off).to(dev)
        self.d_td_off = torch.from_numpy(batch.td_off).to(dev)
        nifo = np.asarray(batch.nifo)
        if nifo.max(initial=1) > MAX_IFO:
            raise ValueError(f"at most {MAX_IFO} detectors per cluster")
        # one launch per detector count, clusters listed largest first
        self.groups = []
        for n in sorted(set(int(x) for x in nifo)):
            ids = np.nonzero(nifo == n)[0]
            ids = ids[np.argsort(-batch.n_pix[ids], kind="stable")].astype(np.int32)
            self.groups.append((n, torch.from_numpy(ids).to(dev), ids.size))

    def run(self, theta: torch.Tensor, phi: torch.Tensor, td: torch.Tensor | None = None,
            want_td: bool = False, grad_scale: torch.Tensor | None = None, stream=None):
        """Returns (value (C,), grad (C, 2), grad_td or None), float64 / float32."""
        dev = self.batch.td.device
        theta = theta.to(device=dev, dtype=torch.float64).contiguous()
        phi = phi.to(device=dev, dtype=torch.float64).contiguous()
        if theta.numel() != self.C or phi.numel() != self.C:
            raise ValueError("theta and phi need one value per cluster")
        td = self.batch.td if td is None else td
        if td.dtype != torch.float32 or td.numel() != self.batch.td.numel() or td.device != dev:
            raise ValueError("td must match the batch's amplitude table")
        td = td.contiguous()
        value = torch.empty(self.C, dtype=torch.float64, device=dev)
        grad = torch.empty(self.C, 2, dtype=torch.float64, device=dev)
        grad_td = torch.zeros_like(td) if want_td else None
        if grad_scale is not None:
            grad_scale = grad_scale.to(device=dev, dtype=torch.float64).contiguous()
        d = _lib.GradDesc()
        d.n_pix, d.net, d.K = ptr(self.d_npix), ptr(self.d_net), ptr(self.d_K)
        d.rms_off, d.td_off = ptr(self.d_rms_off), ptr(self.d_td_off)
        d.rms, d.td, d.geo = ptr(self.batch.rms), ptr(td), ptr(self.geo)
        d.net_nifo, d.net_En = ptr(self.nets.nifo_d), ptr(self.nets.En_d)
        d.REG1, d.reg0 = ptr(self.REG1), self.reg0
        d.theta, d.phi, d.value, d.grad = ptr(theta), ptr(phi), ptr(value), ptr(grad)
        d.grad_td = ptr(grad_td) if want_td else None
        d.grad_scale = ptr(grad_scale) if grad_scale is not None else None
        s = stream_ptr(stream)
        for n, ids, count in self.groups:
            d.n_ifo, d.clusters, d.n_clusters = n, ptr(ids), count
            _lib.check(lib().pycwb_sky_grad(C.byref(d), s), "pycwb_sky_grad")
        return value, grad, grad_td


class SkyStatistic(torch.autograd.Function):
    """AA(theta, phi, td) per cluster as an autograd function.

    forward returns AA (C,) float64.  backward gives dAA/dtheta and dAA/dphi
    from the forward pass, and, when td requires a gradient, recomputes the
    delayed amplitudes in a second kernel pass that scatters
    grad_out[c] * dAA/d(amplitude) onto the td table (float32).
    """

    @staticmethod
    def forward(ctx, theta, phi, td, engine: SkyGradient):
        value, grad, _ = engine.run(theta.detach(), phi.detach(), td.detach())
        ctx.engine = engine
        ctx.save_for_backward(theta, phi, td, grad)
        return value

    @staticmethod
    def backward(ctx, grad_out):
        theta, phi, td, grad = ctx.saved_tensors
        g_theta = g_phi = g_td = None
        if ctx.needs_input_grad[0]:
            g_theta = (grad_out * grad[:, 0]).to(theta.dtype)
        if ctx.needs_input_grad[1]:
            g_phi = (grad_out * grad[:, 1]).to(phi.dtype)
        if ctx.needs_input_grad[2]:
            _, _, g_td = ctx.engine.run(theta.detach(), phi.detach(), td.detach(), want_td=True,
                                        grad_scale=grad_out.detach())
        return g_theta, g_phi, g_td, None


def sky_statistic(theta, phi, td, engine: SkyGradient):
    return SkyStatistic.apply(theta, phi, td, engine)


def angular_distance(theta1, phi1, theta2, phi2):
    """Great-circle angle (radians) between directions, numpy float64."""
    c = (np.sin(theta1) * np.sin(theta2) * np.cos(np.asarray(phi1) - phi2) +
         np.cos(theta1) * np.cos(theta2))
    return np.arccos(np.clip(c, -1.0, 1.0))


def _unit(theta, phi):
    st = torch.sin(theta)
    return torch.stack([st * torch.cos(phi), st * torch.sin(phi), torch.cos(theta)], dim=-1)


def _angles(v):
    v = v / v.norm(dim=-1, keepdim=True)
    return torch.acos(v[:, 2].clamp(-1.0, 1.0)), torch.remainder(torch.atan2(v[:, 1], v[:, 0]),
                                                                 2 * math.pi)


def _project_cap(v, centre, radius):
    """Nearest point of the spherical cap {angle(v, centre) <= radius}."""
    v = v / v.norm(dim=-1, keepdim=True)
    cosd = (v * centre).sum(-1).clamp(-1.0, 1.0)
    outside = cosd < math.cos(radius)
    t = v - cosd.unsqueeze(-1) * centre
    tn = t.norm(dim=-1, keepdim=True).clamp_min(1e-300)
    on_edge = math.cos(radius) * centre + math.sin(radius) * t / tn
    return torch.where(outside.unsqueeze(-1), on_edge, v)


def refine_sky(engine: SkyGradient, theta0, phi0, steps: int = 20, step: float | None = None,
               radius: float | None = None, pixel: float | None = None):
    """Projected gradient ascent of AA on the sphere from (theta0, phi0).

    Each iteration moves every cluster along its Riemannian gradient
    (dAA/dtheta e_theta + dAA/dphi / sin(theta) e_phi) by its own step length,
    maps back onto the sphere, and projects onto the cap of angular ``radius``
    around the start.  A move is kept only if AA increases; the step then
    grows by 1.5, otherwise it halves (so AA never decreases).  Defaults: the
    pixel size sqrt(4 pi / n_sky) of the sky grid, a first step of half a
    pixel, a radius of 3 pixels.  Returns a dict of float64 tensors: theta,
    phi, value, value0 (AA at the start) and the number of kept moves.
    """
    dev = engine.batch.td.device
    th = torch.as_tensor(theta0, dtype=torch.float64, device=dev).clone()
    ph = torch.as_tensor(phi0, dtype=torch.float64, device=dev).clone()
    if pixel is None:
        pixel = math.sqrt(4 * math.pi / engine.nets.n_sky)
    step = 0.5 * pixel if step is None else step
    radius = 3.0 * pixel if radius is None else radius
    centre = _unit(th, ph)
    v, g, _ = engine.run(th, ph)
    v0 = v.clone()
    alpha = torch.full_like(v, step)
    kept = torch.zeros_like(v, dtype=torch.int32)
    for _ in range(int(steps)):
        st, ct = torch.sin(th), torch.cos(th)
        sp, cp = torch.sin(ph), torch.cos(ph)
        e_t = torch.stack([ct * cp, ct * sp, -st], dim=-1)
        e_p = torch.stack([-sp, cp, torch.zeros_like(sp)], dim=-1)
        gp = torch.where(st > 1e-12, g[:, 1] / st.clamp_min(1e-12), torch.zeros_like(st))
        tang = g[:, :1] * e_t + gp.unsqueeze(-1) * e_p
        tn = tang.norm(dim=-1, keepdim=True)
        move = torch.where(tn > 0, tang / tn.clamp_min(1e-300), torch.zeros_like(tang))
        p = _unit(th, ph) + alpha.unsqueeze(-1) * move
        p = _project_cap(p, centre, radius)
        th_n, ph_n = _angles(p)
        v_n, g_n, _ = engine.run(th_n, ph_n)
        ok = (v_n > v) & torch.isfinite(v_n)
        th = torch.where(ok, th_n, th)
        ph = torch.where(ok, ph_n, ph)
        v = torch.where(ok, v_n, v)
        g = torch.where(ok.unsqueeze(-1), g_n, g)
        kept += ok.to(torch.int32)
        alpha = torch.where(ok, torch.clamp(alpha * 1.5, max=radius), alpha * 0.5)
    return {"theta": th, "phi": ph, "value": v, "value0": v0, "kept": kept}
