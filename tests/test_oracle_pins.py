"""Pins the oracle's sky-scan and best-sky algebra against closed forms (no GPU).

The oracle restates the reference's Numba transcription op by op; a slip
shared by its mixed and float64 modes would pass every cross-mode check.
Here its outputs are compared with ``tests/definitional.py``, which computes
the same statistics from their definitions with other methods (eigen-
decompositions and orthogonal projections, float64), in the limits of the
regulator where the reference's formulas reduce to closed forms:

* REG = [0, -inf]: avx_GW_ps (sky_stat.py:156-189) reconstructs the orthogonal
  projection of each pixel's data onto span(r F+, r Fx), so the null energy is
  |v|^2 - |Pv|^2, the coherent energy the cross-term form of the projected
  data in its principal 00/90 axes, and fp, fx the eigenvalues of the Gram
  matrix of the weighted patterns;
* REG = [0, +inf]: the cross component is regulated away (projection onto the
  dominant direction f only); REG = [+inf, -inf]: the plus component is
  (projection onto F only) -- these fix which term each regulator acts on
  and the sign of the dynamic x-regulator;
* REG[0] finite: the plus coefficient 1/max(fp, sqrt(ni |x_p|^2 / e) REG[0])
  with f, ni from the eigen-decomposition;
* the DPF quality of calculate_dpf (dpf.py:441-453) from the eigenvalues;
* compute_statistics_at_sky_position with self-only cross-talk: packets as
  principal-axis coordinates normalised by the Gram eigenvalues (a, A), packet
  energies sum (a + A)^2 / 4, amplitudes at the mean (a + A)/2, N_eff = 2M - 1,
  Lm = Em - Np - Gn; the polar coordinates of avx_pol_ps: radius = norm of the
  data's projection onto span(r F+, r Fx), |cos angle| = its dominant
  component over the radius.

``oracle_mutant`` plants one transcription slip at a time in the oracle (a
flipped sign, a dropped term or factor, a wrong index or operand in a4, a6,
a7, a9, the per-sky scalars and the packet operations); each must fail a pin.
"""

import ctypes

import numpy as np
import pytest

from pycwb_amd import synth

import definitional as D

MAPS = ["nAntennaPrior", "nAlignment", "nLikelihood", "nNullEnergy", "nCorrEnergy",
        "nCorrelation", "nSkyStat", "nDisbalance", "nNetIndex", "nEllipticity", "nPolarisation"]
BIG = np.float32(1e30)
LIMITS = {"span": [0.0, -BIG], "plus": [0.0, BIG], "cross": [1e6, -BIG]}
SCALE_MAPS = {"nLikelihood", "nNullEnergy", "nCorrEnergy", "nSkyStat"}   # energies: scale Eo
TOL = 2e-4


def _case(n_ifo, seed, order=2, K=6, P=40, snr=7.0):
    net = synth.make_network(synth.NETWORK_5[:n_ifo], order=order, K=K)
    cl = synth.make_cluster(net, P, seed=seed, snr=snr, random_rms=True)
    return net, cl


def _safe_dirs(net, cl, En, margin=1e-3):
    """Directions with no pixel energy within ``margin`` of the threshold (so
    float32 and float64 select the same pixels)."""
    out = []
    for l in range(net.n_sky):
        v00, v90 = D.delayed(cl["td00"], cl["td90"], net.ml, l)
        e = (v00 ** 2 + v90 ** 2).sum(0)
        if np.all(np.abs(e - En) > margin * En):
            out.append(l)
    return out


def _scan_errors(oracle, net, cl, limit, En=12.0, reg0=None):
    """Largest normalised deviation (|oracle - closed form| / tolerance) over
    the maps of every safely masked, ungated direction."""
    n, P = net.n_ifo, cl["td00"].shape[2]
    REG = np.array(LIMITS[limit] + [0.0], np.float32)
    if reg0 is not None:
        REG[0] = reg0
    _, maps, _ = oracle.scan_sky(n, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                                 cl["td90"], net.ml, REG, -1e30, 0.5, En,
                                 np.arange(net.n_sky))
    worst, checked = 0.0, 0
    for l in _safe_dirs(net, cl, En):
        v00, v90 = D.delayed(cl["td00"], cl["td90"], net.ml, l)
        if reg0 is None:
            d = D.scan_direction(net.FP_t[l], net.FX_t[l], cl["rms"], v00, v90, En, limit)
        else:
            d = _scan_regulated(net.FP_t[l], net.FX_t[l], cl["rms"], v00, v90, En, reg0)
        if d is None:
            if np.any(maps[:, l] != 0):
                return np.inf, checked
            continue
        checked += 1
        for k, name in enumerate(MAPS):
            scale = abs(d["Eo"]) if name in SCALE_MAPS else 1.0
            err = abs(float(maps[k, l]) - d[name]) / (TOL * (abs(d[name]) + scale))
            worst = max(worst, err)
    return worst, checked


def _scan_regulated(FP_l, FX_l, rms, v00, v90, En, reg0):
    """scan_direction with an active plus regulator REG[0] and REG[1] = -inf."""
    frame = D.dpf_frame(FP_l, FX_l, rms)
    _, fp, fx, f, F, ni = frame
    e = (v00 ** 2 + v90 ** 2).sum(0)
    mask = e > En
    orig = D.signal

    def regulated(v00_, v90_, frame_, mask_, limit):
        xp = (v00_.T * f).sum(1)
        XP = (v90_.T * f).sum(1)
        xx = (v00_.T * F).sum(1)
        XX = (v90_.T * F).sum(1)
        m = mask_.astype(np.float64)
        et = e * m + 1e-12
        alpha = m / (np.maximum(fp, np.sqrt(ni * (xp ** 2 + XP ** 2) / (et + D.EPS)) * reg0)
                     + D.EPS)
        beta = m / (fx + D.EPS)
        s00 = f.T * (xp * alpha) + F.T * (xx * beta)
        s90 = f.T * (XP * alpha) + F.T * (XX * beta)
        return s00, s90, alpha, beta

    D.signal = regulated
    try:
        return D.scan_direction(FP_l, FX_l, rms, v00, v90, En, "span")
    finally:
        D.signal = orig


def _quality_errors(oracle, net, cl):
    _, q = oracle.calculate_dpf(net.FP_t, net.FX_t, cl["rms"], net.n_sky, net.n_ifo, 0.0, 1.0,
                                np.arange(net.n_sky), return_quality=True)
    want = np.array([D.dpf_quality(net.FP_t[l], net.FX_t[l], cl["rms"])
                     for l in range(net.n_sky)])
    return float(np.max(np.abs(q - want) / (1e-3 * (want + 1e-3))))


STAT_KEYS = ["Gn", "Ec", "Dc", "Rc", "Eh", "Np", "Em", "Lm", "norm", "rho", "Lo", "Eo",
             "N_pix_effective"]


def _stats_errors(oracle, net, cl, limit="span", En=12.0, n_dirs=6):
    n, P = net.n_ifo, cl["td00"].shape[2]
    REG = np.array(LIMITS[limit] + [0.0], np.float32)
    worst, checked = 0.0, 0
    dirs = [l for l in _safe_dirs(net, cl, En)
            if np.sum((sum(a ** 2 for a in D.delayed(cl["td00"], cl["td90"], net.ml, l))
                       ).sum(0) > En) >= 3]
    for l in dirs[:: max(1, len(dirs) // n_dirs)][:n_dirs]:
        st = oracle.sky_statistics(l, n, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                                   cl["td00"], cl["td90"], net.ml, REG, En)
        v00, v90 = D.delayed(cl["td00"], cl["td90"], net.ml, l)
        d = D.sky_statistics_self_xtalk(net.FP_t[l], net.FX_t[l], cl["rms"], v00, v90, En,
                                        limit)
        checked += 1
        scale = d["Eo"]
        for k in STAT_KEYS:
            s = scale if k not in ("Rc", "norm", "rho", "N_pix_effective") else 1.0
            worst = max(worst, abs(st[k] - d[k]) / (TOL * (abs(d[k]) + s)))
        for k in ("pd", "pD", "ps", "pS", "p00_POL_r", "p90_POL_r"):
            amp = np.sqrt(scale)
            worst = max(worst, float(np.max(np.abs(st[k] - d[k]))) / (TOL * amp))
        for ph in ("p00", "p90"):
            live = d[ph + "_POL_r"] > 1e-3 * np.sqrt(scale)
            cos = np.abs(np.cos(st[ph + "_POL_a"].astype(np.float64)))
            worst = max(worst, float(np.max(np.abs(cos - d[ph + "_POL_abscos"])[live],
                                            initial=0.0)) / (10 * TOL))
    return worst, checked


CASES = [(3, 11), (3, 12), (4, 13)]


@pytest.mark.parametrize("n_ifo,seed", CASES)
@pytest.mark.parametrize("limit", ["span", "plus", "cross"])
def test_scan_maps_match_closed_forms(oracle, n_ifo, seed, limit):
    net, cl = _case(n_ifo, seed)
    worst, checked = _scan_errors(oracle, net, cl, limit)
    assert checked >= net.n_sky // 3
    assert worst <= 1.0, worst


@pytest.mark.parametrize("n_ifo,seed", CASES)
def test_scan_maps_with_plus_regulator(oracle, n_ifo, seed):
    net, cl = _case(n_ifo, seed)
    worst, checked = _scan_errors(oracle, net, cl, "span", reg0=0.5 * np.sqrt(2))
    assert checked >= net.n_sky // 3
    assert worst <= 1.0, worst


@pytest.mark.parametrize("n_ifo,seed", CASES)
def test_dpf_quality_is_the_eigenvalue_form(oracle, n_ifo, seed):
    net, cl = _case(n_ifo, seed)
    assert _quality_errors(oracle, net, cl) <= 1.0


@pytest.mark.parametrize("n_ifo,seed", CASES)
@pytest.mark.parametrize("limit", ["span", "plus"])
def test_sky_statistics_match_closed_forms(oracle, n_ifo, seed, limit):
    net, cl = _case(n_ifo, seed)
    worst, checked = _stats_errors(oracle, net, cl, limit)
    assert checked >= 3
    assert worst <= 1.0, worst


def _all_pins(oracle):
    net3, cl3 = _case(3, 11)
    net4, cl4 = _case(4, 13)
    errs = {}
    for limit in LIMITS:
        errs[f"scan-{limit}-3"] = _scan_errors(oracle, net3, cl3, limit)[0]
        errs[f"scan-{limit}-4"] = _scan_errors(oracle, net4, cl4, limit)[0]
    errs["scan-reg0"] = _scan_errors(oracle, net3, cl3, "span", reg0=0.5 * np.sqrt(2))[0]
    errs["quality"] = _quality_errors(oracle, net3, cl3)
    errs["stats-span"] = _stats_errors(oracle, net3, cl3, "span")[0]
    errs["stats-plus"] = _stats_errors(oracle, net4, cl4, "plus")[0]
    return errs


MUTANTS = {
    1: "a7: sign of the dynamic x-regulator",
    2: "a7: sign of the cross term of the signal",
    3: "a9: residual taken against the 90 data",
    4: "a9: incoherent fraction divided by the 90 correlation",
    5: "a4: network index from f^3",
    6: "a4: DPF rotation sign",
    7: "a6: total energy without the 1/2",
    8: "scan scalars: sqrt(M) dropped from the disbalance",
    9: "a9: Gaussian-noise correction without the factor 2",
    10: "a11: second packet amplitude without the -norm term",
    11: "a11: noise_ps normalisation factor 1 instead of 1/2",
    12: "a11: effective pixel count without the -1",
    13: "a11: Lm with +Gn",
}


def test_every_planted_slip_fails_a_pin(oracle):
    lib = oracle.load()
    mut = ctypes.c_int.in_dll(lib, "oracle_mutant")
    assert mut.value == 0
    base = _all_pins(oracle)
    assert max(base.values()) <= 1.0, base
    caught = {}
    try:
        for k, what in MUTANTS.items():
            mut.value = k
            errs = _all_pins(oracle)
            caught[k] = max(errs.values())
            assert caught[k] > 1.0, (k, what, errs)
    finally:
        mut.value = 0
