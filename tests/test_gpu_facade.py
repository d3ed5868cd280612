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


def _oracle_chain(oracle, net, setup, rms, cl, P, ml=None, FP=None, FX=None):
    ml = net.ml if ml is None else ml
    FP = net.FP_t if FP is None else FP
    FX = net.FX_t if FX is None else FX
    n_sky = ml.shape[1]
    valid = np.arange(n_sky)
    En = setup["network_energy_threshold"]
    reg = oracle.calculate_dpf(FP, FX, rms, n_sky, net.n_ifo, setup["gamma_regulator"], En,
                               valid)
    REG = np.array([setup["delta_regulator"] * np.sqrt(2), reg, 0], np.float32)
    l_o, maps_o, smax_o = oracle.scan_sky(net.n_ifo, P, n_sky, FP, FX, rms, cl["td00"],
                                          cl["td90"], ml, REG, setup["netCC"],
                                          setup["delta_regulator"], En, valid)
    return l_o, maps_o, smax_o, REG


def test_nrms_maps_set_the_noise_weights(oracle):
    """likelihood(nRMS=...) takes each pixel's noise floor from the detectors'
    TF noise maps (map row round(f*nf/layers), column time//layers) before the
    noise weights are formed, in the single-cluster and the batched path."""
    from types import SimpleNamespace

    from pycwb_amd.likelihoodWP.likelihood import SelfXTalk, likelihood, likelihood_wrapper

    net = synth.make_network(synth.NETWORK_3, order=3, K=10)
    config = Config(healpix=3, netRHO=2.0, TDRate=net.td_rate, TDSize=10, upTDF=1)
    setup = _setup(net, config)
    P = 90
    rng = np.random.default_rng(77)
    layers = np.full(P, 8, np.int32)
    freq = rng.integers(0, 8, P).astype(np.int32)
    tb = rng.integers(0, 40, P)
    times = (tb * 8 + freq).astype(np.int32)
    nRMS = [SimpleNamespace(data=rng.uniform(0.4, 2.5, size=(9, 40))) for _ in range(3)]
    row = np.clip(np.round(freq * 9 / 8).astype(int), 0, 8)
    sigma = np.stack([np.abs(m.data[row, tb]) for m in nRMS])
    sigma = sigma.astype(np.float32)

    def make():
        cl = synth.make_cluster(net, P, seed=5, snr=9.0)
        pa = PixelArrays.from_td(cl["td00"], cl["td90"], np.ones((3, P), np.float32),
                                 time=times, frequency=freq, layers=layers)
        return Cluster(pixel_arrays=pa, cluster_time=1.5, cluster_freq=100.0), cl

    cluster, cl = make()
    res, sm = likelihood(3, cluster, config, setup=setup, xtalk=SelfXTalk(), nRMS=nRMS)
    assert np.array_equal(cluster.pixel_arrays.noise_rms, sigma)
    inv = 1.0 / sigma.astype(np.float64)
    rms = (inv / np.sqrt((inv ** 2).sum(0))).astype(np.float32).T
    l_o, maps_o, smax_o, _ = _oracle_chain(oracle, net, setup, rms, cl, P)
    assert res is not None and sm.l_max == l_o and sm.sky_stat_max == smax_o
    assert np.array_equal(sm.nSkyStat.view(np.uint32), maps_o[6].view(np.uint32))
    # the unit noise floor gives another sky statistic
    l_1, maps_1, _, _ = _oracle_chain(oracle, net, setup,
                                      np.full((P, 3), 1 / np.sqrt(3), np.float32), cl, P)
    assert not np.array_equal(maps_1[6], maps_o[6])
    # the batched wrapper applies the same maps
    c2, _ = make()
    out = likelihood_wrapper(config, [FragmentCluster([c2])], None, None, nRMS=nRMS,
                             xtalk=SelfXTalk())
    assert np.array_equal(c2.pixel_arrays.noise_rms, sigma)
    assert len(out[0]) == 1 and out[0][0][1].l_max == l_o
    assert out[0][0][1].sky_stat_max == smax_o


def test_big_cluster_uses_the_supercluster_coarse_grid(oracle):
    """Without ``setup``, the coarse big-cluster grid of ``supercluster_setup``
    is forwarded (likelihood.py:212-219): a cluster above nRES * csize pixels is
    scanned on it."""
    from pycwb_amd.likelihoodWP.likelihood import SelfXTalk, likelihood

    fine = synth.make_network(synth.NETWORK_3, order=3, K=10)
    coarse = synth.make_network(synth.NETWORK_3, order=2, K=10)
    config = Config(healpix=3, netRHO=2.0, precision=40, nRES=2)
    sc = {"ml": fine.ml, "FP": fine.FP_t.T.astype(np.float64),
          "FX": fine.FX_t.T.astype(np.float64), "ml_big_cluster": coarse.ml,
          "FP_big_cluster": coarse.FP_t.T.astype(np.float64),
          "FX_big_cluster": coarse.FX_t.T.astype(np.float64), "big_cluster_healpix_order": 2}
    P = 120  # > nRES * csize = 80
    cluster, cl, sigma = _cluster(coarse, P, seed=31)
    res, sm = likelihood(3, cluster, config, xtalk=SelfXTalk(), supercluster_setup=sc)
    assert len(sm.nSkyStat) == coarse.n_sky
    inv = 1.0 / sigma.astype(np.float64)
    rms = (inv / np.sqrt((inv ** 2).sum(0))).astype(np.float32).T
    from pycwb_amd.likelihoodWP.likelihood import setup_likelihood
    setup = setup_likelihood(config, None, 3, ml=fine.ml, FP=sc["FP"], FX=sc["FX"])
    l_o, maps_o, smax_o, _ = _oracle_chain(oracle, coarse, setup, rms, cl, P)
    assert sm.l_max == l_o and sm.sky_stat_max == smax_o
    assert np.array_equal(sm.nSkyStat.view(np.uint32), maps_o[6].view(np.uint32))
