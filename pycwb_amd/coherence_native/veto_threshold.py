"""Host helpers of PycWB's coherence_native.veto_threshold used by the GPU
coherence stage."""

from __future__ import annotations

import numpy as np
from scipy.special import gammainccinv


def _igamma_inv_upper(shape: float, p: float) -> float:
    """veto_threshold.py:64-73: x with Q(shape, x) = p (scipy's gammainccinv),
    p clipped to [1e-12, 1 - 1e-12] and shape to >= 1e-12."""
    p = float(np.clip(p, 1.0e-12, 1.0 - 1.0e-12))
    s = float(max(shape, 1.0e-12))
    return float(gammainccinv(s, p))
