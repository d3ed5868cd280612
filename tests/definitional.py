"""Float64 evaluation of the likelihood's sky statistics from their
mathematical definitions (test helper, no GPU).

Independent of the oracle's transcription of the Numba code: the dominant
polarisation frame comes from ``numpy.linalg.eigh`` of each pixel's 2x2 Gram
matrix of the weighted patterns (r F+, r Fx); the reconstructed signal from
the orthogonal projection onto their span (or onto one eigenvector in the
regulated limits); the 00/90 rotations from the principal axes of the 2x2
Gram matrix of a packet's two quadratures.  Used by tests/test_oracle_pins.py
in the limits of the regulator where the reference's formulas reduce to these
closed forms (pycwb/modules/likelihoodWP/sky_stat.py:92-410, dpf.py:342-453,
sky_scan.py:131-179, sky_statistics.py:26-290, packet_ops.py).
"""

from __future__ import annotations

import numpy as np

EPS = 1e-9


def dpf_frame(FP_l, FX_l, rms):
    """Per pixel: fp = largest and fx = smallest eigenvalue of the Gram matrix
    of (r F+, r Fx); f and F = the pattern matrix times the unit eigenvectors
    (|f|^2 = fp, |F|^2 = fx, f . F = 0); ni = sum f^4 / fp^2."""
    P, n = rms.shape
    A = np.stack([rms * FP_l[None, :], rms * FX_l[None, :]], axis=2).astype(np.float64)
    G = np.einsum("pjk,pjl->pkl", A, A)
    w, U = np.linalg.eigh(G)                   # ascending eigenvalues
    fx, fp = w[:, 0], w[:, 1]
    f = np.einsum("pjk,pk->pj", A, U[:, :, 1])
    F = np.einsum("pjk,pk->pj", A, U[:, :, 0])
    ni = (f ** 4).sum(1) / (fp * fp + EPS)
    return A, fp, fx, f, F, ni


def dpf_quality(FP_l, FX_l, rms):
    """sqrt(sum fx/ni / (N + 0.01)) of dpf_np_loops_vec (dpf.py:441-453)."""
    _, fp, fx, _, _, ni = dpf_frame(FP_l, FX_l, rms)
    return np.sqrt(np.sum(fx / (ni + EPS)) / (np.sum(fp > 0) + 0.01))


def delayed(td00, td90, ml, l):
    n = ml.shape[0]
    K = td00.shape[0] // 2
    rows = ml[:, l] + K
    v00 = np.stack([td00[rows[j], j] for j in range(n)]).astype(np.float64)
    v90 = np.stack([td90[rows[j], j] for j in range(n)]).astype(np.float64)
    return v00, v90


def principal_rotation(a00, a90):
    """Rotate each pixel's (00, 90) pair of a (n, P) packet and of other pairs
    into the principal axes of the packet's 2x2 Gram matrix; returns the unit
    eigenvectors (P, 2) of the larger and the smaller eigenvalue."""
    G = np.empty((a00.shape[1], 2, 2))
    G[:, 0, 0] = (a00 * a00).sum(0)
    G[:, 1, 1] = (a90 * a90).sum(0)
    G[:, 0, 1] = G[:, 1, 0] = (a00 * a90).sum(0)
    _, U = np.linalg.eigh(G)
    return U[:, :, 1], U[:, :, 0]


def signal(v00, v90, frame, mask, limit):
    """Reconstructed signal of avx_GW_ps in a limit of the regulator:
    'span'  (REG = [0, -inf]): projection onto span(f, F);
    'plus'  (REG = [0, +inf]): projection onto f only;
    'cross' (REG = [+inf, -inf]): projection onto F only.
    Also the per-pixel coefficient weights alpha, beta of fp, fx."""
    A, fp, fx, f, F, ni = frame
    xp = (v00.T * f).sum(1)
    XP = (v90.T * f).sum(1)
    xx = (v00.T * F).sum(1)
    XX = (v90.T * F).sum(1)
    m = mask.astype(np.float64)
    alpha = m / (fp + EPS) if limit in ("span", "plus") else np.zeros_like(fp)
    beta = m / (fx + EPS) if limit in ("span", "cross") else np.zeros_like(fx)
    s00 = (f.T * (xp * alpha)) + (F.T * (xx * beta))
    s90 = (f.T * (XP * alpha)) + (F.T * (XX * beta))
    return s00, s90, alpha, beta


def scan_direction(FP_l, FX_l, rms, v00, v90, En, limit="span"):
    """The per-direction scalars of scan_sky_for_best_fit (sky_scan.py:131-179)
    in a regulator limit.  Returns a dict keyed like the scan's maps, or None
    for a direction without pixels above the energy threshold."""
    n, P = v00.shape
    e = (v00 ** 2 + v90 ** 2).sum(0)
    mask = e > En
    M = int(mask.sum())
    if M == 0:
        return None
    Eo = 0.5 * e[mask].sum()
    frame = dpf_frame(FP_l, FX_l, rms)
    _, fp, fx, _, _, _ = frame
    s00, s90, alpha, beta = signal(v00, v90, frame, mask, limit)
    mk = mask.astype(np.float64)
    gn = 2.0 * mk * (alpha * fp + beta * fx)          # 2 x (updated mask), masked pixels
    rn = mk * (((s00 - v00) ** 2).sum(0) + ((s90 - v90) ** 2).sum(0))
    u1, u2 = principal_rotation(s00, s90)
    # rotated signal and data: first and second principal component per detector
    r = lambda a00, a90, u: a00 * u[:, 0] + a90 * u[:, 1]
    sp, Sp = r(s00, s90, u1), r(s00, s90, u2)
    xp_, Xp = r(v00, v90, u1), r(v00, v90, u2)
    ss, SS = (sp ** 2).sum(0), (Sp ** 2).sum(0)
    c = ((sp * xp_) ** 2).sum(0) / ((sp * xp_).sum(0) ** 2 + EPS)
    C = ((Sp * Xp) ** 2).sum(0) / ((Sp * Xp).sum(0) ** 2 + EPS)
    ec = mk * (ss * (1 - c) + SS * (1 - C))
    ll = mk * (ss + SS)
    cc = ec / (2 * np.abs(ec) + rn + gn + EPS)
    Cr = 2 * np.sum(ll * cc) / (ll.sum() + EPS)
    Ec = ec.sum()
    No = (gn.sum() + rn.sum()) / 2
    CH = No / (n * M + np.sqrt(M))
    ccf = max(CH, 1.0)
    Co = Ec / (Ec + No * ccf - M * (n - 1))
    L = Eo - No if Eo > 0 else 0.0
    we = e * mk
    pw = (fp * we).sum() / we.sum()
    cw = (fx * we).sum() / we.sum()
    return {"nAntennaPrior": np.sqrt(pw + cw), "nAlignment": np.sqrt(cw / pw),
            "nLikelihood": Eo - No, "nNullEnergy": No, "nCorrEnergy": Ec, "nCorrelation": Co,
            "nSkyStat": L * Co, "nDisbalance": CH, "nNetIndex": ccf, "nEllipticity": Cr,
            "nPolarisation": float(np.sum(ec > EPS)), "Eo": Eo, "M": M,
            "_s00": s00, "_s90": s90, "_mask": mask, "_ec": ec, "_gn": gn, "_e": e}


def packet(a00, a90, mask):
    """avx_packet_ps as geometry: per detector, the 2x2 Gram matrix of the
    masked pixels' (00, 90) amplitudes has eigenvalues a^2 >= A^2; the packet
    is each pixel's coordinates on the principal axes divided by a and A
    (unit vectors), and the packet energy is sum_j (a_j + A_j)^2 / 4."""
    m = mask.astype(np.float64)
    n = a00.shape[0]
    d00, d90 = np.zeros_like(a00), np.zeros_like(a90)
    amp, Amp, U1, U2 = np.zeros(n), np.zeros(n), [], []
    for j in range(n):
        G = np.array([[np.sum(m * a00[j] ** 2), np.sum(m * a00[j] * a90[j])],
                      [np.sum(m * a00[j] * a90[j]), np.sum(m * a90[j] ** 2)]])
        w, U = np.linalg.eigh(G)
        amp[j], Amp[j] = np.sqrt(max(w[1], 0.0)), np.sqrt(max(w[0], 0.0))
        u1, u2 = U[:, 1], U[:, 0]
        d00[j] = m * (a00[j] * u1[0] + a90[j] * u1[1]) / (amp[j] + EPS)
        d90[j] = m * (a00[j] * u2[0] + a90[j] * u2[1]) / (Amp[j] + EPS)
        U1.append(u1)
        U2.append(u2)
    return d00, d90, amp, Amp, np.array(U1), np.array(U2)


def sky_statistics_self_xtalk(FP_l, FX_l, rms, v00, v90, En, limit="span"):
    """compute_statistics_at_sky_position with self-only cross-talk, in a
    regulator limit.  Self-only cross-talk makes every cross-talk sum a plain
    energy: the packet norms are (nearly) 1 per active pixel, the time-domain
    energies are sums of squared reconstructed amplitudes."""
    n, P = v00.shape
    d = scan_direction(FP_l, FX_l, rms, v00, v90, En, limit)
    mask, ec, gn, e = d["_mask"], d["_ec"], d["_gn"], d["_e"]
    s00, s90 = d["_s00"], d["_s90"]
    m = mask.astype(np.float64)
    D00, D90, da, dA, dU1, dU2 = packet(v00, v90, mask)
    S00, S90, sa, sA, sU1, sU2 = packet(s00, s90, mask)
    Eo = np.sum((da + dA) ** 2) / 4
    Lo = np.sum((sa + sA) ** 2) / 4
    # packet norms with each pixel its own only neighbour: e / (e + 1e-12) < 1
    t = D00 ** 2 + D90 ** 2
    qn = np.where(m > 0, t / (t + 1e-12), 0.0)
    qn = np.where(qn >= 1, 0.0, qn)
    pn = np.where(ec > 0, qn, 0.0)
    nx, ns = qn.mean(0), pn.mean(0)
    core = m * (nx > 0)
    sig = m * (ns > 0)
    Ec = np.sum(core * ec)
    SC = np.sum(sig * ec)
    rc = np.where(gn < 2, 1.0, gn / 2)
    RC = np.sum(sig * ec / (rc + EPS))
    Gn = np.sum(sig * gn * nx)
    Eh = np.sum((m - sig) * e) / 2
    Es = 0.0
    Dc = SC - Ec
    Rc = RC / (SC + 0.01)

    def amplitudes(P00, P90, a, A, U1, U2, nrm):
        # each detector's unit principal coordinates, rotated back to the
        # (00, 90) frame and scaled to the mean amplitude (a + A) / 2
        o00 = np.empty_like(P00)
        o90 = np.empty_like(P90)
        for j in range(n):
            k = 0.5 * (a[j] + A[j]) * m * nrm[j]
            o00[j] = k * (P00[j] * U1[j][0] + P90[j] * U2[j][0])
            o90[j] = k * (P00[j] * U1[j][1] + P90[j] * U2[j][1])
        return o00, o90

    pd, pD = amplitudes(D00, D90, da, dA, dU1, dU2, qn)
    ps, pS = amplitudes(S00, S90, sa, sA, sU1, sU2, pn)
    Np_eff = 0.5 * np.sum(qn.sum(0) * m) * 4 / n - 1
    Em = np.sum(m * (pd ** 2 + pD ** 2))
    Np = np.sum(m * ((pd - ps) ** 2 + (pD - pS) ** 2))
    Lm = Em - Np - Gn
    norm = max((Eo - Eh) / Em if Em > 0 else 1e9, 1.0)
    Ec_n, Dc_n = Ec / norm, Dc / norm
    rho = np.sqrt(Ec_n * Rc / 2) if Ec_n > 0 else 0.0
    # avx_pol_ps: radius = norm of the data's projection onto span(f, F); the
    # angle's cosine = its component along the dominant direction over the radius
    A, fp, fx, f, F, ni = dpf_frame(FP_l, FX_l, rms)
    fhat = f / np.sqrt(fp)[:, None]
    pol = {}
    for name, v in (("p00", v00), ("p90", v90)):
        vm = v * m
        Q = np.linalg.qr(A)[0]                                  # (P, n, 2) orthonormal span
        coef = np.einsum("pjk,jp->pk", Q, vm)
        rad = np.sqrt((coef ** 2).sum(1))
        pol[name + "_POL_r"] = rad
        pol[name + "_POL_abscos"] = np.abs((vm.T * fhat).sum(1)) / np.where(rad > 0, rad, 1)
    return {**pol, "Gn": Gn, "Ec": Ec_n, "Dc": Dc_n, "Rc": Rc, "Eh": Eh, "Es": Es, "Np": Np,
            "Em": Em, "Lm": Lm, "norm": norm, "rho": rho, "Lo": Lo, "Eo": Eo,
            "N_pix_effective": Np_eff, "pd": pd, "pD": pD, "ps": ps, "pS": pS}
