"""GPU parity of the sub-network cut against the oracle, through the C ABI.

Bar: the sky-loop tuple (l_max, stat, Em, Am, lm, Vm, suball, EE) and the MRA
tuple (submra, rHo, Eo, Lo, Ls, m) are bit-exact against the oracle's
Numba-typed restatement, per cluster, in single-cluster calls and in batches
mixing detector counts; plus the reference's own behaviour case
(test_supercluster_optimizations.py:270-312: suball fails, cross-talk and MRA
are never evaluated) and apply_subnet_cut end to end.
"""

import numpy as np
import pytest

from pycwb_amd import synth

pytestmark = pytest.mark.gpu


def _net_cluster(ifos, order, K, P, seed, snr=6.0):
    net = synth.make_network(ifos, order=order, K=K)
    cl = synth.make_cluster(net, P, seed=seed, snr=snr)
    return net, cl


SKY_CASES = [
    # detectors, order, K, P, acore, subcut
    (synth.NETWORK_3, 3, 8, 60, 1.5, -1.0),
    (synth.NETWORK_3, 4, 12, 300, 1.5, 0.33),
    (synth.NETWORK_4, 3, 10, 120, 1.7, 0.1),
    (synth.NETWORK_5, 2, 12, 45, 1.5, -1.0),
    (synth.NETWORK_3[:2], 3, 6, 80, 1.2, 0.0),
    (synth.NETWORK_3, 2, 4, 1, 0.5, -1.0),
]


@pytest.mark.parametrize("ifos,order,K,P,acore,subcut", SKY_CASES)
def test_sky_loop_bit_exact(oracle, ifos, order, K, P, acore, subcut):
    from pycwb_amd.super_cluster_native import calculate_e2or_from_acore, optimze_sky_loc_from_td

    print(f"case subnet sky n={len(ifos)} order={order} K={K} P={P} subcut={subcut}", flush=True)
    net, cl = _net_cluster(ifos, order, K, P, seed=500 + P)
    n = net.n_ifo
    En = np.float32(2 * acore * acore * n)
    e2or = calculate_e2or_from_acore(acore, n)
    g = optimze_sky_loc_from_td(n, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                                cl["td90"], net.ml, En, e2or, subcut)
    o = oracle.sky_loc_from_td(n, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                               cl["td90"], net.ml, En, e2or, subcut)
    assert g == o


def test_sky_loop_no_positive_statistic(oracle):
    from pycwb_amd.super_cluster_native import optimze_sky_loc_from_td

    print("case subnet sky all below threshold", flush=True)
    net, cl = _net_cluster(synth.NETWORK_3, 3, 8, 40, seed=9)
    g = optimze_sky_loc_from_td(3, 40, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                                cl["td90"], net.ml, np.float32(1e6), 1.0, -1.0)
    assert g == (0, 0.0, 0.0, 0.0, 0, 0, 0.0, 0.0)


def _random_xtalk(rng, P, max_nb=6):
    """Self entry plus up to max_nb distinct random neighbours per pixel."""
    lookup = np.zeros((P, 2), np.int32)
    rows = []
    for j in range(P):
        nb = rng.choice(P, size=min(P, int(rng.integers(0, max_nb + 1)) + 1), replace=False)
        nb = np.concatenate([[j], nb[nb != j]])
        r = np.zeros((nb.size, 8), np.float32)
        r[:, 0] = nb
        r[:, 4:8] = rng.normal(scale=0.2, size=(nb.size, 4))
        r[0, 4:8] = [1.0, 0.0, 0.0, 1.0]
        lookup[j] = [sum(x.shape[0] for x in rows), sum(x.shape[0] for x in rows) + nb.size]
        rows.append(r)
    return lookup, np.concatenate(rows)


@pytest.mark.parametrize("kind", ["empty", "self", "random"])
@pytest.mark.parametrize("ifos,P", [(synth.NETWORK_3, 90), (synth.NETWORK_4, 300),
                                    (synth.NETWORK_5, 33)])
def test_mra_bit_exact(oracle, kind, ifos, P):
    from pycwb_amd.skystats import self_xtalk
    from pycwb_amd.super_cluster_native import calculate_e2or_from_acore, mra_statistics_from_td

    print(f"case subnet mra {kind} n={len(ifos)} P={P}", flush=True)
    net, cl = _net_cluster(ifos, 3, 10, P, seed=700 + P)
    n = net.n_ifo
    rng = np.random.default_rng(P)
    if kind == "empty":
        lookup, xt = np.zeros((P, 2), np.int32), np.zeros((0, 8), np.float32)
    elif kind == "self":
        lookup, xt = self_xtalk(P)
    else:
        lookup, xt = _random_xtalk(rng, P)
    En = np.float32(2 * 1.5 * 1.5 * n)
    e2or = calculate_e2or_from_acore(1.5, n)
    for l in (0, int(rng.integers(net.n_sky)), net.n_sky - 1):
        g = mra_statistics_from_td(n, P, net.FP_t, net.FX_t, cl["rms"], cl["td00"], cl["td90"],
                                   net.ml, En, e2or, -1.0, xt, lookup, l)
        o = oracle.mra_from_td(n, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                               cl["td90"], net.ml, En, e2or, xt, lookup, l)
        assert g == o, (l, g, o)


def test_mra_large_cluster_uses_global_energies(oracle):
    """More than 4,096 pixels: the working energies live in global scratch."""
    from pycwb_amd.skystats import self_xtalk
    from pycwb_amd.super_cluster_native import mra_statistics_from_td

    print("case subnet mra P=5000", flush=True)
    P = 5000
    net, cl = _net_cluster(synth.NETWORK_3, 2, 6, P, seed=31)
    lookup, xt = self_xtalk(P)
    En = np.float32(2 * 1.2 * 1.2 * 3)
    g = mra_statistics_from_td(3, P, net.FP_t, net.FX_t, cl["rms"], cl["td00"], cl["td90"],
                               net.ml, En, 1.0, -1.0, xt, lookup, 7)
    o = oracle.mra_from_td(3, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                           cl["td90"], net.ml, En, 1.0, xt, lookup, 7)
    assert g == o


def test_batch_mixed_networks_bit_exact(oracle):
    """Many clusters of 3 and 4 detectors in one SubnetBatch: sky loop, then MRA
    on the clusters that need it, each against the oracle."""
    from pycwb_amd.device import require_gpu
    from pycwb_amd.skystats import self_xtalk
    from pycwb_amd.super_cluster_native.sub_net_cut import (SubnetBatch, cluster_batch_from_host,
                                                            needs_mra, sky_tuple, mra_tuple,
                                                            subnet_sky_tables)

    print("case subnet batch 3+4 detectors", flush=True)
    dev = require_gpu()
    nets = [synth.make_network(synth.NETWORK_3, order=3, K=12),
            synth.make_network(synth.NETWORK_4, order=3, K=12)]
    acore, subcut, subnet = 1.5, 0.2, 0.85
    from pycwb_amd.super_cluster_native import calculate_e2or_from_acore
    En = [np.float32(2 * acore * acore * nt.n_ifo) for nt in nets]
    Es = [np.float32(2 * calculate_e2or_from_acore(acore, nt.n_ifo)) for nt in nets]
    rng = np.random.default_rng(77)
    rms_l, td_l, ids, cls = [], [], [], []
    for c in range(40):
        k = c % 2
        P = int(rng.integers(1, 250))
        cl = synth.make_cluster(nets[k], P, seed=1000 + c, snr=float(rng.uniform(2, 12)))
        rms_l.append(cl["rms"])
        td_l.append(synth.packed_td(cl["td00"], cl["td90"]))
        ids.append(k)
        cls.append(cl)
    b = cluster_batch_from_host(rms_l, td_l, ids, dev)
    s = subnet_sky_tables([(nt.FP_t, nt.FX_t, nt.ml) for nt in nets], En, Es, dev)
    sb = SubnetBatch(b, s, subcut, dev)
    sky = sb.sky_results()
    need = []
    for c, cl in enumerate(cls):
        nt = nets[ids[c]]
        o = oracle.sky_loc_from_td(nt.n_ifo, cl["rms"].shape[0], nt.n_sky, nt.FP_t, nt.FX_t,
                                   cl["rms"], cl["td00"], cl["td90"], nt.ml, En[ids[c]],
                                   float(Es[ids[c]]) / 2, subcut)
        assert sky_tuple(sky[c]) == o, c
        if needs_mra(o, subnet):
            need.append(c)
    assert 0 < len(need) < len(cls)
    csrs = [self_xtalk(cls[c]["rms"].shape[0]) for c in need]
    mra = sb.mra_results(need, csrs)
    for c, row in zip(need, mra):
        nt, cl = nets[ids[c]], cls[c]
        lk, xt = self_xtalk(cl["rms"].shape[0])
        o = oracle.mra_from_td(nt.n_ifo, cl["rms"].shape[0], nt.n_sky, nt.FP_t, nt.FX_t,
                               cl["rms"], cl["td00"], cl["td90"], nt.ml, En[ids[c]],
                               float(Es[ids[c]]) / 2, xt, lk, int(sky[c][0]))
        assert mra_tuple(row) == o, c
    # the batch's self-only cross-talk as device CSRs gives the same rows
    from pycwb_amd.xtalk import self_xtalk_batch

    dcsr = self_xtalk_batch([cl["rms"].shape[0] for cl in cls], dev)
    sb.mra_d.zero_()
    mra_d = sb.mra_results(need, dcsr)
    np.testing.assert_array_equal(mra_d, mra)


class _EmptyXTalk:
    @staticmethod
    def _empty(n_pix):
        return np.zeros((n_pix, 2), dtype=np.int32), np.zeros((0, 8), dtype=np.float32)

    def get_xtalk_pixels(self, pixels, check=True):
        return self._empty(len(pixels))

    def get_xtalk_pixels_from_arrays(self, layers, times, check=True):
        return self._empty(len(layers))


class _RaisingXTalk(_EmptyXTalk):
    def get_xtalk_pixels(self, pixels, check=True):
        raise AssertionError("xtalk should not be evaluated")

    def get_xtalk_pixels_from_arrays(self, layers, times, check=True):
        raise AssertionError("xtalk should not be evaluated")


def _pixel_arrays(td_amp, likelihood):
    """PixelArrays of td_amp (P, 2, 2*nd) rows {00 | 90}, unit noise."""
    from pycwb_amd.types import PixelArrays

    P, n, ts = td_amp.shape
    nd = ts // 2
    td00 = np.ascontiguousarray(td_amp[:, :, :nd].transpose(2, 1, 0))
    td90 = np.ascontiguousarray(td_amp[:, :, nd:].transpose(2, 1, 0))
    pa = PixelArrays.from_td(td00, td90, np.ones((n, P), np.float32),
                             time=16 * np.arange(1, P + 1), layers=np.full(P, 16, np.int32))
    pa.likelihood[:] = likelihood
    return pa


def test_reference_case_suball_fails_skips_xtalk_and_mra():
    """test_supercluster_optimizations.py:270-312 (seed 19): subnet fails,
    subrho and subthr pass, no cross-talk or MRA evaluated."""
    from pycwb_amd.super_cluster_native import sub_net_cut_from_pixel_arrays

    print("case subnet reference flags", flush=True)
    rng = np.random.default_rng(19)
    td_amp = rng.normal(size=(4, 2, 10)).astype(np.float32)
    pa = _pixel_arrays(td_amp, np.array([5.0, 4.0, 3.0, 2.0], np.float32))
    ml = np.array([[0, 1, -1], [0, -1, 1]], dtype=np.int32)
    fp = np.array([[0.8, 0.4, -0.1], [0.2, -0.3, 0.9]], dtype=np.float32)
    fx = np.array([[0.1, -0.7, 0.5], [0.6, 0.2, -0.4]], dtype=np.float32)
    timing = {}
    res = sub_net_cut_from_pixel_arrays(pa, np.arange(len(pa), dtype=np.int64), ml,
                                        fp.T.copy(), fx.T.copy(), 0.1, 0.1, 2, 3, 2.0, -1.0, 0.0,
                                        -1.0, _RaisingXTalk(), arrays_prepared=True,
                                        timing=timing)
    assert res["subnet_passed"] is False
    assert res["subrho_passed"] is True
    assert res["subthr_passed"] is True
    assert timing.get("sky_scan", 0.0) > 0.0
    assert timing.get("xtalk", 0.0) == 0.0
    assert timing.get("mra", 0.0) == 0.0


def test_prepared_path_matches_wrapper_and_oracle(oracle):
    """test_supercluster_optimizations.py:315-380 (seed 11): the top-2 loudest
    pixels through sub_net_cut and sub_net_cut_from_pixel_arrays agree, and
    both follow the oracle's decision."""
    from pycwb_amd.super_cluster_native import sub_net_cut, sub_net_cut_from_pixel_arrays
    from pycwb_amd.super_cluster_native.sub_net_cut import subnet_decision
    from pycwb_amd.super_cluster_native.utils import _top_loudest_indices

    print("case subnet prepared vs wrapper", flush=True)
    rng = np.random.default_rng(11)
    td_amp = rng.normal(size=(6, 2, 10)).astype(np.float32)
    lik = np.array([5.0, 4.0, 4.0, 3.0, 2.0, 1.0], np.float32)
    pa = _pixel_arrays(td_amp, lik)
    top = _top_loudest_indices(pa.likelihood, 2)
    ml = np.array([[0, 1, -1], [0, -1, 1]], dtype=np.int32)
    fp = np.array([[0.8, 0.4, -0.1], [0.2, -0.3, 0.9]], dtype=np.float32)
    fx = np.array([[0.1, -0.7, 0.5], [0.6, 0.2, -0.4]], dtype=np.float32)
    sub = _pixel_arrays(td_amp[top], lik[top])
    args = (0.1, 0.1, 2, 3, -1.0, -1.0, 0.0, -1.0, _EmptyXTalk())
    expected = sub_net_cut(sub, ml, fp, fx, *args)
    actual = sub_net_cut_from_pixel_arrays(pa, top, ml, fp.T.copy(), fx.T.copy(), *args,
                                           arrays_prepared=True)
    assert actual == expected
    td00 = np.ascontiguousarray(td_amp[top][:, :, :5].transpose(2, 1, 0))
    td90 = np.ascontiguousarray(td_amp[top][:, :, 5:].transpose(2, 1, 0))
    rms = np.full((2, 2), np.float32(1 / np.sqrt(2)), np.float32)
    En = np.float32(2 * 0.1 * 0.1 * 2)
    sky = oracle.sky_loc_from_td(2, 2, 3, fp.T.copy(), fx.T.copy(), rms, td00, td90, ml, En,
                                 0.1, -1.0)
    mra = oracle.mra_from_td(2, 2, 3, fp.T.copy(), fx.T.copy(), rms, td00, td90, ml, En, 0.1,
                             np.zeros((0, 8), np.float32), np.zeros((2, 2), np.int32), sky[0])
    assert actual == subnet_decision(sky, mra, -1.0, -1.0, 0.0)


def test_apply_subnet_cut_marks_clusters(oracle):
    from pycwb_amd.skystats import self_xtalk
    from pycwb_amd.super_cluster_native import apply_subnet_cut, calculate_e2or_from_acore
    from pycwb_amd.super_cluster_native.sub_net_cut import cluster_inputs, subnet_decision
    from pycwb_amd.types import Cluster

    class SelfXT:
        def get_xtalk_pixels_from_arrays(self, layers, times, check=True):
            return self_xtalk(len(layers))

    print("case apply_subnet_cut 12 clusters", flush=True)
    net = synth.make_network(synth.NETWORK_3, order=3, K=10)
    acore, subnet, subcut, subnorm, subrho, loud = 1.5, 0.88, 0.1, 0.0, 20.0, 80
    e2or = calculate_e2or_from_acore(acore, 3)
    clusters, expect = [], []
    rng = np.random.default_rng(5)
    from pycwb_amd.types import PixelArrays
    for c in range(12):
        P = int(rng.integers(20, 200))
        cl = synth.make_cluster(net, P, seed=90 + c, snr=float(rng.uniform(3, 15)))
        sigma = np.abs(rng.normal(1.0, 0.1, size=(3, P))).astype(np.float32)
        pa = PixelArrays.from_td(cl["td00"], cl["td90"], sigma, time=np.arange(P),
                                 layers=np.full(P, 8, np.int32))
        pa.likelihood[:] = rng.permutation(P).astype(np.float32)
        clusters.append(Cluster(pixel_arrays=pa))
        rows = np.argsort(-pa.likelihood)[:loud]
        rms, td = cluster_inputs(pa, rows)
        td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
        td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
        En = np.float32(2 * acore * acore * 3)
        sky = oracle.sky_loc_from_td(3, rows.size, net.n_sky, net.FP_t, net.FX_t, rms, td00,
                                     td90, net.ml, En, e2or, subcut)
        mra = None
        if not sky[6] <= subnet:
            lk, xt = self_xtalk(rows.size)
            mra = oracle.mra_from_td(3, rows.size, net.n_sky, net.FP_t, net.FX_t, rms, td00,
                                     td90, net.ml, En, e2or, xt, lk, sky[0])
        r = subnet_decision(sky, mra, subnet, subrho, subnorm)
        expect.append(r["subnet_passed"] and r["subrho_passed"] and r["subthr_passed"])
    kept = apply_subnet_cut(clusters, loud, net.ml, net.FP_t, net.FX_t, acore, e2or, 3,
                            net.n_sky, subnet, subcut, subnorm, subrho, SelfXT(),
                            arrays_prepared=True)
    got = [c.cluster_status == -1 for c in clusters]
    assert got == expect
    assert [id(c) for c in kept] == [id(c) for c, e in zip(clusters, expect) if e]
    assert 0 < sum(expect) < len(expect)
