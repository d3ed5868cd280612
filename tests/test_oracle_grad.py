"""Continuous-direction sky statistic: pins of the float64 oracle (CPU).

* direction terms at grid directions reproduce the grid tables: rint of the
  continuous delay is ml and the patterns are FP/FX (pycwb/types/
  detector.py:1243-1329 as restated in pycwb_amd.sky);
* at integer delays with the table patterns the statistic equals AA of the
  separately written float64 scan oracle (sky_scan.py:131-141);
* complex-step derivatives agree with central finite differences of the
  oracle's own value, for the angles and for single amplitudes;
* linear interpolation is continuous across tap boundaries.
"""

import numpy as np
import pytest

from pycwb_amd import synth
from pycwb_amd.skygrad import network_geometry


def _cluster(net, P, seed, snr=6.0):
    cl = synth.make_cluster(net, P, seed, snr=snr)
    td = synth.packed_td(cl["td00"], cl["td90"])
    return cl, td


def _objective(oracle, net, cl, td, reg1=0.05, delta=0.5):
    n = net.n_ifo
    geo = network_geometry(net.ifos, net.td_rate)[:n]
    En = 2 * 2.0 * n
    reg0 = float(np.float32(delta * np.sqrt(2)))
    return oracle.SkyObjective(cl["rms"], td, geo, reg0, reg1, En), geo


@pytest.mark.parametrize("ifos", [synth.NETWORK_3, synth.NETWORK_5])
def test_direction_terms_match_grid_tables(oracle, ifos):
    net = synth.make_network(ifos, order=4, K=16)
    geo = network_geometry(net.ifos, net.td_rate)[:net.n_ifo]
    theta, phi = np.pi / 2 - net.lat, net.phi
    for l in range(0, net.n_sky, 7):
        tau, fp, fx = oracle.direction_terms(geo, theta[l], phi[l])
        np.testing.assert_allclose(fp, net.FP_t[l], rtol=0, atol=2e-7)
        np.testing.assert_allclose(fx, net.FX_t[l], rtol=0, atol=2e-7)
        r = np.clip(np.rint(tau), -net.K, net.K).astype(np.int32)
        near_half = np.abs(np.abs(tau - np.floor(tau)) - 0.5) < 1e-9
        assert np.all((r == net.ml[:, l]) | near_half), l


def test_integer_delays_reproduce_scan_oracle(oracle):
    net = synth.make_network(synth.NETWORK_3, order=3, K=12)
    P = 60
    cl, td = _cluster(net, P, seed=11)
    reg1 = 0.07
    obj, _ = _objective(oracle, net, cl, td, reg1=reg1)
    REG64 = np.array([obj.reg0, reg1, 0.0])
    valid = np.arange(net.n_sky, dtype=np.int64)
    _, maps, _ = oracle.scan_sky_f64(net.n_ifo, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                                     cl["td00"], cl["td90"], net.ml, REG64, -1e300, obj.En,
                                     valid)
    checked = 0
    for l in range(net.n_sky):
        if np.any(np.abs(net.ml[:, l]) >= net.K) or maps[6, l] == 0.0:
            continue
        v = obj.at(net.ml[:, l].astype(np.float64), net.FP_t[l].astype(np.float64),
                   net.FX_t[l].astype(np.float64))
        assert v == pytest.approx(maps[6, l], rel=1e-11, abs=1e-11), l
        checked += 1
    assert checked > net.n_sky // 2


def test_complex_step_matches_finite_differences(oracle):
    net = synth.make_network(synth.NETWORK_3, order=5, K=24)
    cl, td = _cluster(net, 80, seed=3, snr=8.0)
    obj, _ = _objective(oracle, net, cl, td)
    rng = np.random.default_rng(0)
    l0 = int(cl["l0"])
    th0, ph0 = np.pi / 2 - net.lat[l0], net.phi[l0]
    h = 1e-6
    good = 0
    n_pts = 20
    for _ in range(n_pts):
        th = th0 + rng.normal(scale=0.02)
        ph = ph0 + rng.normal(scale=0.02)
        g = obj.grad(th, ph)
        fd = np.array([(obj.value(th + h, ph) - obj.value(th - h, ph)) / (2 * h),
                       (obj.value(th, ph + h) - obj.value(th, ph - h)) / (2 * h)])
        scale = np.abs(g).max() + 1e-3
        if np.abs(g - fd).max() <= 1e-5 * scale:
            good += 1
    # a step across a tap boundary or a mask flip is a kink: allow a few
    assert good >= n_pts - 2


def test_amplitude_derivative_matches_finite_differences(oracle):
    net = synth.make_network(synth.NETWORK_4, order=4, K=16)
    cl, td = _cluster(net, 40, seed=9, snr=10.0)
    obj, geo = _objective(oracle, net, cl, td)
    l0 = int(cl["l0"])
    th, ph = np.pi / 2 - net.lat[l0] + 0.013, net.phi[l0] - 0.021
    tau, _, _ = oracle.direction_terms(geo, th, ph)
    nd = 2 * net.K + 1
    rng = np.random.default_rng(4)
    h = 1e-5
    nz = 0
    for _ in range(25):
        j = int(rng.choice(cl["signal"])) if rng.random() < 0.8 else int(rng.integers(40))
        i, ph2 = rng.integers(net.n_ifo), rng.integers(2)
        tap = int(np.floor(tau[i] + net.K)) + rng.integers(2)
        idx = int(2 * ((j * net.n_ifo + i) * nd + tap) + ph2)
        g = obj.grad_td(th, ph, idx)
        fd = (obj.value(th, ph, idx, h) - obj.value(th, ph, idx, -h)) / (2 * h)
        assert g == pytest.approx(fd, rel=1e-5, abs=1e-7), (j, i, tap, ph2)
        nz += g != 0.0
    assert nz > 5
    # an amplitude the interpolation does not read has zero derivative
    i = 0
    far = (int(np.floor(tau[i] + net.K)) + 5) % nd
    idx = int(2 * ((3 * net.n_ifo + i) * nd + far))
    assert obj.grad_td(th, ph, idx) == 0.0


def test_interpolation_continuous_across_taps(oracle):
    net = synth.make_network(synth.NETWORK_3, order=4, K=16)
    cl, td = _cluster(net, 50, seed=21, snr=7.0)
    obj, geo = _objective(oracle, net, cl, td)
    l0 = int(cl["l0"])
    th0, ph0 = np.pi / 2 - net.lat[l0], net.phi[l0]
    # walk in theta until detector 1's delay crosses a tap boundary
    tau0 = oracle.direction_terms(geo, th0, ph0)[0][1]
    dt = 1e-4
    th = th0
    while np.floor(oracle.direction_terms(geo, th + dt, ph0)[0][1]) == np.floor(tau0):
        th += dt
    lo, hi = th, th + dt
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if np.floor(oracle.direction_terms(geo, mid, ph0)[0][1]) == np.floor(tau0):
            lo = mid
        else:
            hi = mid
    a, b = obj.value(lo, ph0), obj.value(hi, ph0)
    assert a == pytest.approx(b, rel=1e-9, abs=1e-9)
