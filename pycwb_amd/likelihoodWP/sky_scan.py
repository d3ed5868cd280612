"""GPU sky scan with the call signature of PycWB's ``scan_sky_for_best_fit``.

``scan_sky_for_best_fit`` mirrors pycwb/modules/likelihoodWP/sky_scan.py:17-192
(arguments, 13-tuple result, float32 maps) and runs the HIP kernel
``pycwb_sky_scan_plain``.  ``PlainClusterScan`` is the device-resident form used
by the likelihood facade: regulator and scan enqueue back to back on one stream
and nothing returns to the host until the caller asks.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib, sky
from ..device import as_dev, lib, ptr, require_gpu, stream_ptr

MAP_NAMES = (
    "nAntennaPrior", "nAlignment", "nLikelihood", "nNullEnergy", "nCorrEnergy",
    "nCorrelation", "nSkyStat", "nDisbalance", "nNetIndex", "nEllipticity", "nPolarisation",
)


def check_scan_inputs(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, sky_valid_indices):
    """Host-side shape and range checks (the kernels assume them)."""
    if FP.shape != (n_sky, n_ifo) or FX.shape != (n_sky, n_ifo):
        raise ValueError(f"FP and FX must have shape (n_sky, n_ifo)=({n_sky}, {n_ifo}), "
                         f"got {tuple(FP.shape)} and {tuple(FX.shape)}")
    if tuple(rms.shape) != (n_pix, n_ifo):
        raise ValueError(f"rms must have shape (n_pix, n_ifo)=({n_pix}, {n_ifo}), "
                         f"got {tuple(rms.shape)}")
    if td00.ndim != 3 or tuple(td00.shape[1:]) != (n_ifo, n_pix) or td00.shape != td90.shape:
        raise ValueError("td00/td90 must both have shape (n_delay, n_ifo, n_pix)")
    n_delay = int(td00.shape[0])
    if n_delay % 2 != 1:
        raise ValueError(f"n_delay must be odd (2K+1), got {n_delay}")
    if tuple(ml.shape) != (n_ifo, n_sky):
        raise ValueError(f"ml must have shape (n_ifo, n_sky)=({n_ifo}, {n_sky})")
    if not 1 <= n_ifo <= 8:
        raise ValueError(f"n_ifo must be in [1, 8], got {n_ifo}")
    valid = np.asarray(sky_valid_indices.cpu() if isinstance(sky_valid_indices, torch.Tensor)
                       else sky_valid_indices, dtype=np.int64)
    if valid.size == 0:
        raise ValueError("sky_valid_indices is empty")
    if valid.min() < 0 or valid.max() >= n_sky:
        raise ValueError("sky_valid_indices out of range")
    K = n_delay // 2
    mlh = ml.cpu().numpy() if isinstance(ml, torch.Tensor) else np.asarray(ml)
    if np.abs(mlh).max() > K:
        raise ValueError(f"|ml| exceeds the delay half-width K={K}")
    return n_delay, valid


def scan_flags(*arrays) -> int:
    """Exact-skip flags, enabled when the weights, patterns and regulators are finite."""
    ok = all(bool(torch.isfinite(a).all()) if isinstance(a, torch.Tensor)
             else bool(np.isfinite(np.asarray(a)).all()) for a in arrays)
    return _lib.DEFAULT_FLAGS if ok else _lib.REGULATOR_EARLY_EXIT


def _dense(*tensors) -> None:
    """The kernels index raw pointers: every tensor must be contiguous."""
    for t in tensors:
        if t is not None and not t.is_contiguous():
            raise ValueError("device arrays passed to the C ABI must be contiguous")


class PlainClusterScan:
    """One cluster's regulator + scan on the GPU (one launch family per call)."""

    def __init__(self, n_ifo: int, n_pix: int, n_delay: int, device=None):
        self.device = device or require_gpu()
        self.n_ifo, self.n_pix, self.n_delay = int(n_ifo), int(n_pix), int(n_delay)
        nbytes = lib().pycwb_sky_scan_workspace_bytes(self.n_ifo, self.n_pix, self.n_delay)
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        self.REG = torch.zeros(3, dtype=torch.float32, device=self.device)
        self.reg_f64 = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.l_max = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.stat_max = torch.zeros(1, dtype=torch.float64, device=self.device)

    def load_td(self, td00: torch.Tensor, td90: torch.Tensor, stream=None) -> None:
        _dense(td00, td90)
        rc = lib().pycwb_td_relayout(self.n_ifo, self.n_pix, self.n_delay, ptr(td00), ptr(td90),
                                     ptr(self.ws), self.ws.numel(), stream_ptr(stream))
        _lib.check(rc, "pycwb_td_relayout")

    def regulator(self, FP, FX, rms, n_sky, gamma, En, reg0, valid, stream=None,
                  flags=_lib.DEFAULT_FLAGS) -> None:
        _dense(FP, FX, rms, valid)
        rc = lib().pycwb_dpf_regulator(self.n_ifo, self.n_pix, int(n_sky), ptr(FP), ptr(FX),
                                       ptr(rms), float(gamma), float(En), float(reg0), ptr(valid),
                                       int(valid.numel()), int(flags), ptr(self.REG),
                                       ptr(self.reg_f64), ptr(self.ws), self.ws.numel(),
                                       stream_ptr(stream))
        _lib.check(rc, "pycwb_dpf_regulator")

    def scan(self, FP, FX, ml, rms, n_sky, netCC, En, valid, maps=None, stream=None,
             td_packed=None, sky_order=None, flags=_lib.DEFAULT_FLAGS) -> None:
        _dense(FP, FX, ml, rms, valid, maps, td_packed, sky_order)
        rc = lib().pycwb_sky_scan_plain(self.n_ifo, self.n_pix, int(n_sky), self.n_delay,
                                        ptr(FP), ptr(FX), ptr(ml), ptr(rms), ptr(td_packed),
                                        ptr(self.REG), float(netCC), float(En), ptr(valid),
                                        int(valid.numel()), ptr(sky_order), int(flags),
                                        ptr(maps), ptr(self.l_max), ptr(self.stat_max),
                                        ptr(self.ws), self.ws.numel(), stream_ptr(stream))
        _lib.check(rc, "pycwb_sky_scan_plain")


def scan_sky_for_best_fit(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml, REG, netCC,
                          delta_regulator, network_energy_threshold, sky_valid_indices):
    """Drop-in for ``scan_sky_for_best_fit`` (sky_scan.py:17-192).

    Returns ``(l_max, nAntennaPrior, nAlignment, nLikelihood, nNullEnergy,
    nCorrEnergy, nCorrelation, nSkyStat, nDisbalance, nNetIndex, nEllipticity,
    nPolarisation, sky_stat_max)`` with float32 numpy maps of length n_sky.
    ``REG`` is the float32 triple ``[delta*sqrt(2), regulator, 0]``.
    """
    dev = require_gpu()
    n_delay, valid_h = check_scan_inputs(n_ifo, n_pix, n_sky, FP, FX, rms, td00, td90, ml,
                                         sky_valid_indices)
    del delta_regulator  # the probability map it selects is not part of the result
    FPd, FXd = as_dev(FP, torch.float32, dev), as_dev(FX, torch.float32, dev)
    rmsd = as_dev(rms, torch.float32, dev)
    mld = as_dev(ml, torch.int32, dev)
    t00, t90 = as_dev(td00, torch.float32, dev), as_dev(td90, torch.float32, dev)
    valid = torch.from_numpy(valid_h).to(dev)
    sc = PlainClusterScan(n_ifo, n_pix, n_delay, dev)
    REGh = np.asarray(REG, dtype=np.float32).reshape(-1)[:3]
    sc.REG.copy_(as_dev(REGh, torch.float32, dev))
    sc.load_td(t00, t90)
    maps = torch.empty((len(MAP_NAMES), int(n_sky)), dtype=torch.float32, device=dev)
    order = torch.from_numpy(sky.delay_order(np.asarray(ml.cpu() if isinstance(ml, torch.Tensor)
                                                        else ml), valid_h)).to(dev)
    flags = scan_flags(rmsd, FPd, FXd, REGh)
    sc.scan(FPd, FXd, mld, rmsd, n_sky, netCC, network_energy_threshold, valid, maps,
            sky_order=order, flags=flags)
    torch.cuda.current_stream().synchronize()
    m = maps.cpu().numpy()
    return (int(sc.l_max.item()), *[m[i] for i in range(len(MAP_NAMES))],
            float(sc.stat_max.item()))


find_optimal_sky_localization = scan_sky_for_best_fit
