"""GPU DPF regulator with the call signature of PycWB's ``calculate_dpf``.

Mirrors pycwb/modules/likelihoodWP/dpf.py:7-34: per valid sky direction the
dominant-polarisation-frame quality ``sqrt(sum fx/(ni+1e-9) / (N+0.01))`` is
compared with ``gamma_regulator``; the regulator is
``(S**2 / (N_gamma**2 + 1e-9) - 1) * network_energy_threshold`` in float64.
Runs the HIP kernel ``pycwb_dpf_regulator``.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from ..device import as_dev, require_gpu
from .sky_scan import PlainClusterScan


def calculate_dpf(FP, FX, rms, n_sky, n_ifo, gamma_regulator, network_energy_threshold,
                  sky_valid_indices):
    if tuple(FP.shape) != (n_sky, n_ifo) or tuple(FX.shape) != (n_sky, n_ifo) or \
            rms.shape[1] != n_ifo:
        n1, n2 = FP.shape
        m1, m2 = FX.shape
        p1, p2 = rms.shape
        raise ValueError(
            f"FP and FX must have shape (n_sky, n_ifo) and rms must have shape (n_pix, n_ifo), "
            f"got FP: ({n1}, {n2}), FX: ({m1}, {m2}), rms: ({p1}, {p2})")
    valid = np.asarray(sky_valid_indices.cpu() if isinstance(sky_valid_indices, torch.Tensor)
                       else sky_valid_indices, dtype=np.int64)
    if valid.size == 0 or valid.min() < 0 or valid.max() >= n_sky:
        raise ValueError("sky_valid_indices must be non-empty and within [0, n_sky)")
    dev = require_gpu()
    n_pix = int(rms.shape[0])
    sc = PlainClusterScan(n_ifo, n_pix, 1, dev)
    sc.regulator(as_dev(FP, torch.float32, dev), as_dev(FX, torch.float32, dev),
                 as_dev(rms, torch.float32, dev), n_sky, gamma_regulator,
                 network_energy_threshold, 0.0, torch.from_numpy(valid).to(dev),
                 flags=_lib.REGULATOR_EARLY_EXIT)
    torch.cuda.current_stream().synchronize()
    return float(sc.reg_f64.item())
