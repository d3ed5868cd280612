"""Stress test of the regulator's term bounds (no GPU).

The batched and plain regulator kernels decide a direction's DPF-quality count
early from per-direction bounds on every pixel's term t_j = fx_j / (ni_j + 1e-9)
(skyscan.hip: ``regulator_bounds``), computed from the cluster's noise-weight
statistics (R2 = sum_i r_i^2 per pixel, its range, the per-detector range of
r_i^2 / R2): the lower bound is R2min times the smallest lambda_min over the
vertices of the weight polytope, divided by an upper bound on ni from the cone
of dominant-frame directions those vertices span; the upper bound is n R2max
times the smaller of min_u max_i (u.v_i)^2 over 32 sampled unit vectors and
the largest u^T M u over the vertices for the box centre's minor eigenvector u
(used only when fp is bounded away from zero); both widened by the margin
1e-4 R2max max|v|^2 for the float32 pipeline's deviation from the exact
eigenvalue.  ``bounds`` below follows the kernel step by step.  The count is
exact only if every float32 term the reference computes (dpf.py:342-453,
restated by the oracle) lies inside those bounds, and the ni bound holds for
the float32 ni.  Here that is checked for every direction of HEALPix order 7
and every pixel of clusters built to strain it: noise weights spread over four
decades, one detector dominating (a near rank-one Gram matrix, fx near 0),
noise rms spanning thirty decades across pixels, weights nearly equal (a
narrow cone, where the ni bound is tightest), and three to five detectors.
"""

import numpy as np
import pytest

from pycwb_amd import synth


def weight_stats(rms):
    """[R2min, R2max, min_j r_ji^2 / R2_j, max_j r_ji^2 / R2_j per detector] in float64."""
    r2 = rms.astype(np.float64) ** 2
    R2 = r2.sum(1)
    W = r2 / R2[:, None]
    return R2.min(), R2.max(), W.min(0), W.max(0)


def _vertices(lo, hi):
    """The kernel's vertex list: the vertices of {lo <= w <= hi, sum w = 1} for up
    to four detectors (n - 1 coordinates at a bound, the last from the sum),
    else (or if none) the n vertices of {w >= lo, sum w = 1}."""
    n = lo.size
    out = []
    if n <= 4:
        for k in range(n):
            others = [i for i in range(n) if i != k]
            for bits in range(1 << (n - 1)):
                w = np.empty(n)
                rest = 0.0
                for bit, i in enumerate(others):
                    w[i] = hi[i] if (bits >> bit) & 1 else lo[i]
                    rest += w[i]
                w[k] = 1.0 - rest
                if lo[k] - 1e-12 <= w[k] <= hi[k] + 1e-12:
                    out.append(w)
    if not out:
        sw = lo.sum()
        for k in range(n):
            w = lo.copy()
            w[k] = 1.0 - (sw - lo[k])
            out.append(w)
    return np.array(out)


def bounds(FP_t, FX_t, rms):
    """The kernel's (t_lb, t_ub) per direction, float64 (t_ub < 0: none)."""
    R2min, R2max, lo, hi = weight_stats(rms)
    a, b = FP_t.astype(np.float64), FX_t.astype(np.float64)
    S, n = a.shape
    A2, B2, AB = a * a, b * b, a * b
    v2 = A2 + B2
    v2max = v2.max(1)
    slack = 1e-4 * R2max * v2max
    # reference: the box centre scaled to the simplex
    wc = (lo + hi) / (lo + hi).sum()
    m11, m12, m22 = A2 @ wc, AB @ wc, B2 @ wc
    dx, dy = m11 - m22, 2 * m12
    dn = np.hypot(dx, dy)
    dx, dy = np.where(dn > 0, dx / np.where(dn > 0, dn, 1), 1.0), np.where(dn > 0, dy / np.where(dn > 0, dn, 1), 0.0)
    h = 0.5 * (m11 + m22)
    lmin = h - np.sqrt(np.maximum(h * h - (m11 * m22 - m12 * m12), 0))
    p1, p2, q1, q2 = m12, lmin - m11, lmin - m22, m12
    np_, nq = p1 * p1 + p2 * p2, q1 * q1 + q2 * q2
    usep = (np_ >= nq) & (np_ > 0)
    useq = ~usep & (nq > 0)
    u1 = np.where(usep, p1 / np.sqrt(np.where(usep, np_, 1)), np.where(useq, q1 / np.sqrt(np.where(useq, nq, 1)), 1.0))
    u2 = np.where(usep, p2 / np.sqrt(np.where(usep, np_, 1)), np.where(useq, q2 / np.sqrt(np.where(useq, nq, 1)), 0.0))
    V = _vertices(lo, hi)
    g11, g12, g22 = A2 @ V.T, AB @ V.T, B2 @ V.T          # (S, nV)
    hh = 0.5 * (g11 + g22)
    C = (hh - np.sqrt(np.maximum(hh * hh - (g11 * g22 - g12 * g12), 0))).min(1)
    Bq = np.maximum((u1 * u1)[:, None] * g11 + (2 * u1 * u2)[:, None] * g12 +
                    (u2 * u2)[:, None] * g22, 0).max(1)
    x, y, tr = g11 - g22, 2 * g12, g11 + g22
    pr = dx[:, None] * x + dy[:, None] * y
    rho = pr.min(1)
    trmin, trmax = tr.min(1), tr.max(1)
    tan = np.where(pr > 0, (dx[:, None] * y - dy[:, None] * x) / np.where(pr > 0, pr, 1), np.nan)
    cone = (rho > 1e-3 * trmax) & np.all(pr > 0, 1)
    kl = np.nanargmin(np.where(cone[:, None], tan, 0), 1)
    kh = np.nanargmax(np.where(cone[:, None], tan, 0), 1)
    idx = np.arange(S)
    el = np.stack([x[idx, kl], y[idx, kl]], 1)
    eh = np.stack([x[idx, kh], y[idx, kh]], 1)
    el /= np.maximum(np.linalg.norm(el, axis=1, keepdims=True), 1e-300)
    eh /= np.maximum(np.linalg.norm(eh, axis=1, keepdims=True), 1e-300)
    delta = np.where(cone, 1e-5 * (trmax / np.where(cone, rho, 1) + 1), 1.0)
    ok = cone & (delta < 0.1)
    sv = np.where(v2 > 0, v2, 1)
    qx, qy = (A2 - B2) / sv, 2 * AB / sv
    inside = (el[:, 0:1] * qy - el[:, 1:2] * qx >= 0) & (qx * eh[:, 1:2] - qy * eh[:, 0:1] >= 0)
    mc = np.minimum(np.maximum(el[:, 0:1] * qx + el[:, 1:2] * qy,
                               eh[:, 0:1] * qx + eh[:, 1:2] * qy) + delta[:, None], 1.0)
    mc = np.where(inside | (v2 <= 0), 1.0, mc)
    lam = 0.5 * (trmin + rho)
    cap = hi[None, :] * v2 * 0.5 * (1 + mc) / np.where(ok, lam, 1)[:, None] * (1 + 1e-5)
    # greedy: largest caps first (ties in index order)
    tot = np.zeros(S)
    for i in range(n):
        ahead = np.zeros(S)
        for k in range(n):
            if k == i:
                continue
            sel = (cap[:, k] > cap[:, i]) | ((cap[:, k] == cap[:, i]) & (k < i))
            ahead += np.where(sel, cap[:, k], 0.0)
        take = np.clip(np.minimum(1.0 - ahead, cap[:, i]), 0, None)
        tot += take * take
    ni_hi = np.where(ok & (cap.sum(1) >= 1.0), np.minimum(tot, 1.0), 1.0)
    den = np.minimum(ni_hi * (1 + 1e-3) + 1e-9, 1 + 2e-5)
    t_lb = np.maximum((R2min * C - slack) / den, 0.0)
    fpw = (lo[None, :] * v2).sum(1)
    fp_lb = 0.5 * R2min * fpw * (1.0 - 1e-5)
    q = np.arange(32) * (np.pi / 32.0)
    proj = np.cos(q)[None, :, None] * a[:, None, :] + np.sin(q)[None, :, None] * b[:, None, :]
    B = np.minimum((proj ** 2).max(2).min(1), Bq)
    t_ub = np.where(fp_lb * fp_lb >= 1e-3, n * (1.0 + 2e-5) * (R2max * B + slack), -1.0)
    return t_lb, t_ub, ni_hi


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
    elif kind == "bench":          # the benchmark's noise: |N(1, 0.1)| per detector
        sigma = np.abs(rng.normal(1.0, 0.1, size=(P, n)))
    elif kind == "narrow":         # nearly equal weights: the narrowest cone of directions
        sigma = 1.0 + 1e-4 * rng.standard_normal((P, n))
    else:
        raise ValueError(kind)
    inv = 1.0 / sigma
    return (inv / np.sqrt((inv ** 2).sum(1, keepdims=True))).astype(np.float32)


@pytest.mark.parametrize("ifos,kind,seed", [
    (synth.NETWORK_3, "spread", 1), (synth.NETWORK_4, "spread", 2),
    (synth.NETWORK_3, "dominant", 3), (synth.NETWORK_5, "dominant", 4),
    (synth.NETWORK_4, "extreme", 5), (synth.NETWORK_5, "spread", 6),
    (synth.NETWORK_3, "bench", 7), (synth.NETWORK_4, "bench", 8),
    (synth.NETWORK_3, "narrow", 9), (synth.NETWORK_4, "narrow", 10),
    (synth.NETWORK_5, "narrow", 11),
])
def test_float32_terms_lie_inside_the_bounds(oracle, ifos, kind, seed):
    net = synth.make_network(ifos, order=7, K=8)
    P = 48
    rms = _rms(kind, P, net.n_ifo, seed)
    dirs = np.arange(net.n_sky, dtype=np.int64)
    fp, fx, ni = oracle.dpf_terms(net.FP_t, net.FX_t, rms, dirs)
    t = fx / (ni + np.float32(1e-9))                     # the float32 term of dpf.py:450
    assert t.dtype == np.float32 and np.all(np.isfinite(t))
    t_lb, t_ub, ni_hi = bounds(net.FP_t, net.FX_t, rms)
    ni64 = ni.astype(np.float64)
    assert np.all(ni64 <= ni_hi[:, None] * (1 + 1e-3)), float((ni64 / ni_hi[:, None]).max())
    if kind in ("bench", "narrow"):
        assert np.mean(ni_hi < 1.0) > 0.1     # the ni bound takes part
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


@pytest.mark.gpu
@pytest.mark.parametrize("ifos,kind,seed", [
    (synth.NETWORK_3, "bench", 21), (synth.NETWORK_4, "bench", 22),
    (synth.NETWORK_3, "narrow", 23), (synth.NETWORK_5, "spread", 24),
    (synth.NETWORK_4, "dominant", 25),
])
def test_kernel_bounds_follow_the_restatement(oracle, ifos, kind, seed):
    """The HIP kernel's weight statistics and term bounds (through the
    diagnostic entry point) against ``bounds`` above, and against the float32
    terms themselves."""
    import ctypes as C

    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import lib, ptr

    print(f"case kernel term bounds {len(ifos)} detectors, {kind}", flush=True)
    dev = torch.device("cuda", 0)
    net = synth.make_network(ifos, order=5, K=8)
    n, S, P = net.n_ifo, net.n_sky, 96
    rms = _rms(kind, P, n, seed)
    d_rms = torch.from_numpy(rms).to(dev)
    d_fp = torch.from_numpy(np.ascontiguousarray(net.FP_t, np.float32)).to(dev)
    d_fx = torch.from_numpy(np.ascontiguousarray(net.FX_t, np.float32)).to(dev)
    ws = torch.zeros(24, dtype=torch.float64, device=dev)
    lb = torch.empty(S, dtype=torch.float64, device=dev)
    ub = torch.empty(S, dtype=torch.float64, device=dev)
    _lib.check(lib().pycwb_regulator_term_bounds(n, P, ptr(d_rms), ptr(d_fp), ptr(d_fx), S,
                                                  ptr(ws), ptr(lb), ptr(ub), None),
               "pycwb_regulator_term_bounds")
    torch.cuda.synchronize()
    ws, lb, ub = ws.cpu().numpy(), lb.cpu().numpy(), ub.cpu().numpy()
    R2min, R2max, lo, hi = weight_stats(rms)
    np.testing.assert_allclose(ws[:2], [R2min, R2max], rtol=1e-15)
    np.testing.assert_allclose(ws[2:2 + n], lo, rtol=1e-15)
    np.testing.assert_allclose(ws[10:10 + n], hi, rtol=1e-15)
    # the kernel computes the same bounds in float32 and keeps them as
    # half-precision values rounded outwards: at most 2^-10 (plus the float32
    # error) looser than the float64 restatement, never tighter than it
    t_lb, t_ub, _ = bounds(net.FP_t, net.FX_t, rms)
    scale = (net.FP_t.astype(np.float64) ** 2 + net.FX_t.astype(np.float64) ** 2).max(1)
    tol = 1e-5 * scale                                # float32 error, far inside the margins
    assert np.all(lb <= t_lb + tol), float((lb - t_lb).max())
    assert np.all(lb >= t_lb * (1 - 2.0 ** -10) - tol), float((t_lb - lb).max())
    both = (ub >= 0) & (t_ub >= 0)
    assert np.mean(both) >= 0.99 * np.mean(t_ub >= 0)
    assert np.all(ub[both] >= t_ub[both] - tol[both])
    assert np.all(ub[both] <= t_ub[both] * (1 + 2.0 ** -10) + tol[both])
    fp, fx, ni = oracle.dpf_terms(net.FP_t, net.FX_t, rms, np.arange(S, dtype=np.int64))
    t = (fx / (ni + np.float32(1e-9))).astype(np.float64)
    assert np.all(t >= lb[:, None])
    assert np.all(t[ub >= 0] <= ub[ub >= 0, None])


@pytest.mark.gpu
@pytest.mark.parametrize("ifos,kind,P,seed", [
    (synth.NETWORK_3, "bench", 2500, 31), (synth.NETWORK_4, "bench", 700, 32),
    (synth.NETWORK_3, "spread", 300, 33), (synth.NETWORK_4, "dominant", 1100, 34),
    (synth.NETWORK_5, "narrow", 260, 35), (synth.NETWORK_3, "bench", 40, 36),
])
def test_pixel_class_bounds_hold(oracle, ifos, kind, P, seed):
    """The batched regulator's pixel classes (through the diagnostic entry
    point): every pixel's weights inside its class's box, the remaining-count
    table equal to the suffix counts of the classes, and every float32 term
    inside its class's bounds for every direction."""
    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import lib, ptr

    print(f"case pixel classes {len(ifos)} detectors, {kind}, P={P}", flush=True)
    dev = torch.device("cuda", 0)
    net = synth.make_network(ifos, order=5, K=8)
    n, S = net.n_ifo, net.n_sky
    rms = _rms(kind, P, n, seed)
    ent = int(lib().pycwb_class_rem_entries(P))
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in
         (("rms", rms), ("fp", net.FP_t.astype(np.float32)), ("fx", net.FX_t.astype(np.float32)))}
    ws = torch.zeros(24, dtype=torch.float64, device=dev)
    pcl = torch.zeros(P, dtype=torch.uint8, device=dev)
    cst = torch.zeros(16 * 24, dtype=torch.float64, device=dev)
    rem = torch.zeros(ent, dtype=torch.int16, device=dev)
    lb = torch.empty(S * 16, dtype=torch.float64, device=dev)
    ub = torch.empty(S * 16, dtype=torch.float64, device=dev)
    _lib.check(lib().pycwb_regulator_class_bounds(n, P, ptr(d["rms"]), ptr(d["fp"]), ptr(d["fx"]),
                                                   S, ptr(ws), ptr(pcl), ptr(cst), ptr(rem),
                                                   ptr(lb), ptr(ub), None),
               "pycwb_regulator_class_bounds")
    torch.cuda.synchronize()
    pcl = pcl.cpu().numpy().astype(np.int64)
    cst = cst.cpu().numpy().reshape(16, 24)
    G = int(cst[0, 19])
    # 1 class below 128 pixels, then 4, 8 and 16 from 128, 512 and 2048 pixels
    expect_G = 1 if P < 128 else 4 if P < 512 else 8 if P < 2048 else 16
    assert G == expect_G
    assert pcl.max() < G
    counts = np.bincount(pcl, minlength=16)
    np.testing.assert_array_equal(cst[:, 18], counts[:16] * (np.arange(16) < G))
    r2 = rms.astype(np.float64) ** 2
    R2 = r2.sum(1)
    W = r2 / R2[:, None]
    for g in range(G):
        m = pcl == g
        if not m.any():
            continue
        assert cst[g, 0] <= R2[m].min() and R2[m].max() <= cst[g, 1]
        assert np.all(cst[g, 2:2 + n] <= W[m].min(0)) and np.all(W[m].max(0) <= cst[g, 10:10 + n])
    if G == 1:
        return
    rem = rem.cpu().numpy().view(np.uint16).astype(np.int64).reshape(-1, 16)
    nq = (P + 15) // 16
    onehot = np.zeros((nq * 16 + 16, 16), np.int64)
    onehot[np.arange(P), pcl] = 1
    suffix = onehot[::-1].cumsum(0)[::-1]
    np.testing.assert_array_equal(rem[:nq + 1], suffix[0:16 * (nq + 1):16])
    lb = lb.cpu().numpy().reshape(S, 16)
    ub = ub.cpu().numpy().reshape(S, 16)
    fp, fx, ni = oracle.dpf_terms(net.FP_t, net.FX_t, rms, np.arange(S, dtype=np.int64))
    t = (fx / (ni + np.float32(1e-9))).astype(np.float64)
    assert np.all(t >= lb[:, pcl])
    u = ub[:, pcl]
    assert np.all((u < 0) | (t <= u))
    assert np.all(fp[u >= 0] > 0)


@pytest.mark.gpu
def test_batched_regulator_with_pixel_classes_mixed_detectors(oracle):
    """A batch mixing three, four and five detectors (so the clusters' noise
    weights start at offsets that are not multiples of their own detector
    counts): the regulator counts with pixel classes equal those without and
    the oracle's."""
    import torch

    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch

    print("case batched regulator, pixel classes, 3/4/5 detectors", flush=True)
    dev = torch.device("cuda", 0)
    nets = [synth.make_network(synth.NETWORK_3, order=4, K=8),
            synth.make_network(synth.NETWORK_4, order=4, K=8),
            synth.make_network(synth.NETWORK_5, order=4, K=8)]
    sizes = np.array([1500, 37, 700, 260, 1100, 90, 64, 2100, 300, 1023], np.int64)
    ids = np.array([1, 0, 2, 1, 0, 2, 0, 1, 2, 0], np.int64)
    b = synth_batch(nets, sizes, ids, seed=515, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    gamma = 0.5 * 0.5 * 2 / 3
    res = {}
    for cls in (True, False):
        bs = BatchedSkyScan(b, nt, valid, gamma, 0.5, 0.5, block=64, grid=64,
                            pixel_classes=cls)
        bs.launch()
        res[cls] = bs.results()
        bs.close()
    np.testing.assert_array_equal(res[True]["regulator"], res[False]["regulator"])
    np.testing.assert_array_equal(res[True]["l_max"], res[False]["l_max"])
    vh = np.arange(nets[0].n_sky, dtype=np.int64)
    for c in (0, 2, 7):
        net = nets[ids[c]]
        rms = b.cluster_rms(c).cpu().numpy()
        reg = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, net.n_ifo, gamma,
                                   nt.En[ids[c]], vh)
        assert reg == res[True]["regulator"][c], c
