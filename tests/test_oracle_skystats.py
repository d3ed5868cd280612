"""Pins the oracle's best-sky statistics (no GPU).

Known answers from the reference's packet-op contract tests
(pycwb/modules/likelihoodWP/tests/test_packet_ops_contract.py) and closed forms
of the self-only cross-talk case.
"""

import numpy as np
import pytest

from pycwb_amd import synth


def test_packet_norm_all_masked_clamps_to_two(oracle):
    """test_packet_ops_contract.py:78-94: all pixels masked -> norm == 2 and
    detector_snr = q_E * 2 / 2 = q_E."""
    p = np.array([[1.0, 2.0]])
    q = np.array([[3.0, 4.0]])
    mk = np.array([0.0, 0.0], np.float32)
    det, norm, rn, qn = oracle.packet_norm(p, q, np.zeros((0, 8)), np.array([[0, 0], [0, 0]]),
                                           mk, np.array([5.0], np.float32))
    np.testing.assert_allclose(det, [5.0])
    assert np.all(norm == 2.0)


def test_xtalk_sum_zero_neighbours(oracle):
    """test_packet_ops_contract.py:57-75: active pixel without neighbours -> 0."""
    p = np.array([[1.0, 0.5]])
    q = np.array([[2.0, 1.0]])
    r = oracle.xtalk_sum(p, q, np.zeros((0, 8)), np.array([[0, 0], [0, 0]]),
                         np.array([1.0, 0.0], np.float32))
    assert r == 0.0


def test_xtalk_self_closed_form(oracle):
    """Self-only cross-talk (coefficients 1,0,0,1): sum over active pixels of
    p^2 + q^2 per detector."""
    rng = np.random.default_rng(0)
    p, q = rng.normal(size=(3, 9)), rng.normal(size=(3, 9))
    mk = (rng.uniform(size=9) > 0.4).astype(np.float32)
    lk, xt = oracle.self_xtalk(9)
    r = oracle.xtalk_sum(p, q, xt, lk, mk)
    assert r == pytest.approx(float(np.sum((p ** 2 + q ** 2)[:, mk > 0])), rel=1e-14)
    det, norm, rn, qn = oracle.packet_norm(p, q, xt, lk, mk, np.ones(3, np.float32))
    # e = (p^2+q^2)/(t+1e-12) with t = p^2+q^2 -> just below 1 -> kept
    expect = np.where(mk > 0, (p ** 2 + q ** 2) / (p ** 2 + q ** 2 + 1e-12), 0)
    np.testing.assert_allclose(qn, np.where(expect >= 1, 0, expect), rtol=1e-12)
    np.testing.assert_allclose(rn, np.where(mk > 0, np.sum(p ** 2 + q ** 2, 0), 0), rtol=1e-12)


def _stats_case(seed=3, P=60, snr=8.0):
    net = synth.make_network(synth.NETWORK_3, order=3, K=8)
    cl = synth.make_cluster(net, P, seed=seed, snr=snr)
    return net, cl


def test_sky_statistics_identities(oracle):
    net, cl = _stats_case()
    REG = np.array([0.5 * np.sqrt(2), 2.0, 0], np.float32)
    l = cl["l0"]
    st = oracle.sky_statistics(l, 3, 60, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                               cl["td90"], net.ml, REG, 12.0)
    # Lm = Em - Np - Gn (sky_statistics.py:194)
    assert st["Lm"] == pytest.approx(st["Em"] - st["Np"] - st["Gn"], rel=1e-12)
    # null packet = f32(data) - f32(signal) (avx_loadNULL_ps)
    np.testing.assert_array_equal(st["nul00"], st["pd"].astype(np.float32)
                                  - st["ps"].astype(np.float32))
    # self-only xtalk: Em = sum over active pixels of pd^2 + pD^2
    act = st["pixel_mask"] > 0
    em = float(np.sum((st["pd"] ** 2 + st["pD"] ** 2)[:, act]))
    assert st["Em"] == pytest.approx(em, rel=1e-12)
    # rho (2G) = sqrt(Ec * Rc / 2)
    assert st["rho"] == pytest.approx(np.sqrt(st["Ec"] * st["Rc"] / 2), rel=1e-12)
    # a coherent injection at l0 gives positive likelihood and correlation
    assert st["Lm"] > 0 and st["Ec"] > 0 and st["norm"] >= 1.0
    # masked pixels reconstruct to zero
    assert np.all(st["pd"][:, ~act] == 0) and np.all(st["ps"][:, ~act] == 0)


def test_sky_statistics_xgb_mode(oracle):
    net, cl = _stats_case(seed=4)
    REG = np.array([0.5 * np.sqrt(2), 2.0, 0], np.float32)
    a = oracle.sky_statistics(cl["l0"], 3, 60, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                              cl["td00"], cl["td90"], net.ml, REG, 12.0, xgb_rho_mode=True)
    b = oracle.sky_statistics(cl["l0"], 3, 60, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                              cl["td00"], cl["td90"], net.ml, REG, 12.0, xgb_rho_mode=False)
    assert a["xrho"] == pytest.approx(b["rho"], rel=1e-12)
    pen = a["cc"] if a["cc"] > 1 else (a["Np"] + a["Gn"]) / (a["N_pix_effective"] * 3)
    chi = (a["Np"] + a["Gn"]) / (a["N_pix_effective"] * 3)
    expect = np.sqrt(a["Ec"] / (1 + chi * (max(1.0, chi) - 1))) if a["Ec"] > 0 else 0
    assert a["rho"] == pytest.approx(expect, rel=1e-12)
