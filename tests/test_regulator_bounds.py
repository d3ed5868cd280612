"""Stress test of the regulator's term bounds (no GPU).

The batched and plain regulator kernels decide a direction's DPF-quality count
early from per-direction bounds on every pixel's term t_j = fx_j / (ni_j + 1e-9)
(skyscan.hip: ``regulator_bounds``): with the cluster's noise-weight statistics
(R2 = sum_i r_i^2 per pixel, its range, the per-detector minimum of
r_i^2 / R2), lower bound R2min times the smallest lambda_min over the weight
simplex's vertices, upper bound n R2max min_u max_i (u.v_i)^2 over 32 sampled
unit vectors (used only when fp is bounded away from zero), each widened by
the margin 1e-4 R2max max|v|^2 for the float32 pipeline's deviation from the
exact eigenvalue.  The count is exact only if every float32 term the
reference computes (dpf.py:342-453, restated by the oracle) lies inside those
bounds.  Here that is checked for every direction of HEALPix order 7 and every
pixel of clusters built to strain it: noise weights spread over four decades,
one detector dominating (a near rank-one Gram matrix, fx near 0), noise rms
spanning thirty decades across pixels, and three to five detectors.
"""

import numpy as np
import pytest

from pycwb_amd import synth


def weight_stats(rms):
    """[R2min, R2max, min_j r_ji^2 / R2_j per detector] in float64."""
    r2 = rms.astype(np.float64) ** 2
    R2 = r2.sum(1)
    return R2.min(), R2.max(), (r2 / R2[:, None]).min(0)


def bounds(FP_t, FX_t, rms):
    """The kernel's (t_lb, t_ub) per direction, float64 (t_ub < 0: none)."""
    R2min, R2max, wmin = weight_stats(rms)
    a, b = FP_t.astype(np.float64), FX_t.astype(np.float64)
    n = a.shape[1]
    v2 = a * a + b * b
    v2max = v2.max(1)
    sw = wmin.sum()
    fpw = (wmin[None, :] * v2).sum(1)
    slack = 1e-4 * R2max * v2max
    C = np.full(a.shape[0], np.inf)
    for k in range(n):
        w = wmin.copy()
        w[k] = 1.0 - (sw - wmin[k])
        g11 = (w * a * a).sum(1)
        g12 = (w * a * b).sum(1)
        g22 = (w * b * b).sum(1)
        h = 0.5 * (g11 + g22)
        disc = h * h - (g11 * g22 - g12 * g12)
        C = np.minimum(C, h - np.sqrt(np.maximum(disc, 0.0)))
    t_lb = np.maximum((R2min * C - slack) / (1.0 + 2e-5), 0.0)
    fp_lb = 0.5 * R2min * fpw * (1.0 - 1e-5)
    q = np.arange(32) * (np.pi / 32.0)
    proj = np.cos(q)[None, :, None] * a[:, None, :] + np.sin(q)[None, :, None] * b[:, None, :]
    B = (proj ** 2).max(2).min(1)
    t_ub = np.where(fp_lb * fp_lb >= 1e-3, n * (1.0 + 2e-5) * (R2max * B + slack), -1.0)
    return t_lb, t_ub


def _rms(kind, P, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "spread":           # noise rms log-uniform over four decades
        sigma = np.exp(rng.uniform(np.log(0.01), np.log(100.0), size=(P, n)))
    elif kind == "dominant":       # one detector far quieter than the others
        sigma = np.full((P, n), 1.0)
        sigma[:, 0] = 1e-3 * np.exp(rng.uniform(-1, 1, P))
        sigma[:, 1:] *= np.exp(rng.uniform(-0.5, 0.5, (P, n - 1)))
    elif kind == "extreme":        # noise rms over thirty decades across pixels
        sigma = np.exp(rng.uniform(np.log(1e-15), np.log(1e15), size=(P, 1))) * \
            np.exp(rng.uniform(-0.3, 0.3, (P, n)))
    else:
        raise ValueError(kind)
    inv = 1.0 / sigma
    return (inv / np.sqrt((inv ** 2).sum(1, keepdims=True))).astype(np.float32)


@pytest.mark.parametrize("ifos,kind,seed", [
    (synth.NETWORK_3, "spread", 1), (synth.NETWORK_4, "spread", 2),
    (synth.NETWORK_3, "dominant", 3), (synth.NETWORK_5, "dominant", 4),
    (synth.NETWORK_4, "extreme", 5), (synth.NETWORK_5, "spread", 6),
])
def test_float32_terms_lie_inside_the_bounds(oracle, ifos, kind, seed):
    net = synth.make_network(ifos, order=7, K=8)
    P = 48
    rms = _rms(kind, P, net.n_ifo, seed)
    dirs = np.arange(net.n_sky, dtype=np.int64)
    fp, fx, ni = oracle.dpf_terms(net.FP_t, net.FX_t, rms, dirs)
    t = fx / (ni + np.float32(1e-9))                     # the float32 term of dpf.py:450
    assert t.dtype == np.float32 and np.all(np.isfinite(t))
    t_lb, t_ub = bounds(net.FP_t, net.FX_t, rms)
    t64 = t.astype(np.float64)
    low = t64 < t_lb[:, None]
    assert not low.any(), (kind, int(low.sum()), float((t_lb[:, None] - t64)[low].max()))
    has_ub = t_ub >= 0
    high = (t64 > t_ub[:, None]) & has_ub[:, None]
    assert not high.any(), (kind, int(high.sum()))
    # an upper bound also guarantees every pixel counts in NN (fp > 0)
    assert np.all(fp[has_ub] > 0)
    # the float32 deviation from the exact term stays inside the margin
    slack = 1e-4 * weight_stats(rms)[1] * (net.FP_t.astype(np.float64) ** 2 +
                                           net.FX_t.astype(np.float64) ** 2).max(1)
    r = rms.astype(np.float64)
    A = np.stack([r[None, :, :] * net.FP_t[:, None, :], r[None, :, :] * net.FX_t[:, None, :]],
                 axis=3)                                  # (S, P, n, 2)
    G = np.einsum("spik,spil->spkl", A, A)
    lam = np.linalg.eigvalsh(G)                          # (S, P, 2)
    dev = np.abs(fx.astype(np.float64) - lam[..., 0])
    assert np.all(dev <= slack[:, None]), float((dev / slack[:, None]).max())
