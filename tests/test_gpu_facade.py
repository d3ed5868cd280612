"""The drop-in facade (setup_likelihood / likelihood / likelihood_wrapper) on the GPU."""

import numpy as np
import pytest

from pycwb_amd import synth
from pycwb_amd.types import Cluster, Config, FragmentCluster, PixelArrays

pytestmark = pytest.mark.gpu


def _cluster(net, P, seed, snr=9.0):
    cl = synth.make_cluster(net, P, seed=seed, snr=snr, random_rms=True)
    rng = np.random.default_rng(seed)
    sigma = np.abs(rng.normal(1.0, 0.1, size=(net.n_ifo, P))).astype(np.float32)
    pa = PixelArrays.from_td(cl["td00"], cl["td90"], sigma)
    return Cluster(pixel_arrays=pa, cluster_time=1.5, cluster_freq=100.0), cl, sigma


def _setup(net, config):
    from pycwb_amd.likelihoodWP.likelihood import setup_likelihood

    return setup_likelihood(config, None, net.n_ifo, ml=net.ml, FP=net.FP_t.T.astype(np.float64),
                            FX=net.FX_t.T.astype(np.float64))


def test_likelihood_single_cluster_matches_oracle(oracle):
    from pycwb_amd.likelihoodWP.likelihood import SelfXTalk, likelihood

    net = synth.make_network(synth.NETWORK_3, order=3, K=10)
    config = Config(healpix=3, netRHO=2.0)
    setup = _setup(net, config)
    cluster, cl, sigma = _cluster(net, 120, seed=21)
    res, skymap = likelihood(3, cluster, config, setup=setup, xtalk=SelfXTalk(), cluster_id=1)
    assert res is cluster and cluster.cluster_status == -1
    # oracle chain on the same inputs
    inv = 1.0 / sigma.astype(np.float64)
    rms = (inv / np.sqrt((inv ** 2).sum(0))).astype(np.float32).T
    valid = np.arange(net.n_sky)
    En = setup["network_energy_threshold"]
    reg = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, 3, setup["gamma_regulator"],
                               En, valid)
    REG = np.array([setup["delta_regulator"] * np.sqrt(2), reg, 0], np.float32)
    l_o, maps_o, smax_o = oracle.scan_sky(3, 120, net.n_sky, net.FP_t, net.FX_t, rms,
                                          cl["td00"], cl["td90"], net.ml, REG, setup["netCC"],
                                          setup["delta_regulator"], En, valid)
    assert skymap.l_max == l_o and skymap.sky_stat_max == smax_o
    assert np.array_equal(skymap.nSkyStat.view(np.uint32), maps_o[6].view(np.uint32))
    o = oracle.sky_statistics(l_o, 3, 120, net.n_sky, net.FP_t, net.FX_t, rms, cl["td00"],
                              cl["td90"], net.ml, REG, En)
    m = cluster.cluster_meta
    assert m.like_sky == pytest.approx(o["Em"], rel=1e-6)
    assert m.net_ecor == pytest.approx(o["Ec"], rel=1e-6)
    assert m.net_rho2 == pytest.approx(o["rho"], rel=1e-6)
    assert m.l_max == l_o
    assert cluster.sky_time_delay == [float(v) for v in net.ml[:, l_o]]
    theta = np.degrees(np.pi / 2 - net.lat[l_o])
    assert m.theta == pytest.approx(theta) and m.phi == pytest.approx(np.degrees(net.phi[l_o]))
    assert skymap.nProbability.sum() == pytest.approx(1.0, rel=1e-5)
    np.testing.assert_allclose(cluster.pixel_arrays.wave, o["pd"].astype(np.float32),
                               rtol=1e-6, atol=1e-6)


def test_wrapper_matches_single_cluster_calls():
    from pycwb_amd.likelihoodWP.likelihood import SelfXTalk, likelihood, likelihood_wrapper

    net = synth.make_network(synth.NETWORK_3, order=3, K=10)
    config = Config(healpix=3, netRHO=2.0, TDRate=net.td_rate, TDSize=10, upTDF=1)
    frags = []
    singles = []
    for lag in range(2):
        cls = []
        for k, (P, snr) in enumerate([(60, 10.0), (25, 0.5), (140, 8.0)]):
            c, _, _ = _cluster(net, P, seed=100 * lag + k, snr=snr)
            c2, _, _ = _cluster(net, P, seed=100 * lag + k, snr=snr)
            cls.append(c)
            singles.append((lag, k, c2))
        frags.append(FragmentCluster(cls))
    out = likelihood_wrapper(config, frags, None, None, xtalk=SelfXTalk())
    setup = _setup(net, config)
    expect = [[] for _ in frags]
    for lag, k, c2 in singles:
        r, sm = likelihood(3, c2, config, setup=setup, xtalk=SelfXTalk(), cluster_id=k + 1)
        if r is not None:
            expect[lag].append((r, sm))
    assert [len(x) for x in out] == [len(x) for x in expect]
    assert sum(len(x) for x in out) >= 2
    for a_lag, e_lag in zip(out, expect):
        for (ca, sa), (ce, se) in zip(a_lag, e_lag):
            assert sa.l_max == se.l_max and sa.sky_stat_max == se.sky_stat_max
            assert np.array_equal(sa.nSkyStat, se.nSkyStat)
            assert ca.cluster_meta.like_sky == pytest.approx(ce.cluster_meta.like_sky, rel=1e-9)
            assert ca.cluster_meta.net_rho2 == pytest.approx(ce.cluster_meta.net_rho2, rel=1e-9)
