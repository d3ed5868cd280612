"""Threshold cuts of the likelihood stage (detection_statistics.py:444-517)."""

from __future__ import annotations

import numpy as np


def get_likelihood_rejection_reason(sky_statistics, network_energy_threshold, netEC_threshold,
                                    net_rho_threshold=None, xgb_rho_mode=False):
    """Return a rejection string if a cut fails, else None.

    2G: Lm > 0, Eo - Eh > 0, Ec * Rc / cc >= netEC, N >= 1.
    XGB: Lm > 0, Eo - Eh > 0, finite rho >= |netRHO|, N >= 1.
    """
    del network_energy_threshold
    Lm, Eo, Eh = sky_statistics.Lm, sky_statistics.Eo, sky_statistics.Eh
    Ec, Rc, cc = sky_statistics.Ec, sky_statistics.Rc, sky_statistics.cc
    rho, N = sky_statistics.rho, sky_statistics.N_pix_effective
    reasons = []
    if Lm <= 0.:
        reasons.append(f"Lm > 0 but Lm = {Lm};")
    if (Eo - Eh) <= 0.:
        reasons.append(f" (Eo - Eh) > 0 but (Eo - Eh) = {Eo - Eh};")
    if not xgb_rho_mode:
        if Ec * Rc / cc < netEC_threshold:
            reasons.append(f" Ec * Rc / cc >= netEC_threshold but Ec * Rc / cc = "
                           f"{Ec * Rc / cc:.4f} < {netEC_threshold:.4f};")
    else:
        if net_rho_threshold is None:
            net_rho_threshold = (netEC_threshold / 2.0) ** 0.5
        if not np.isfinite(rho) or rho < net_rho_threshold:
            reasons.append(f" rho >= |netRHO| but rho = {rho} < {net_rho_threshold};")
    if N < 1:
        reasons.append(f" N < 1 but N = {N};")
    return "".join(reasons) if reasons else None


threshold_cut = get_likelihood_rejection_reason
