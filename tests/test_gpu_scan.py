"""GPU parity of the plain sky-scan family against the CPU oracle.

Bar: the selected sky pixel, the gating pattern and every float32 map are
bit-exact against the oracle's Numba-faithful mixed-precision restatement; the
regulator is bit-exact in float64.  Calls go through the C ABI.
"""

import sys

import numpy as np
import pytest

from pycwb_amd import synth

pytestmark = pytest.mark.gpu


def _case(n_ifo, order, K, P, seed, snr=6.0, random_rms=True):
    net = synth.make_network(synth.NETWORK_5[:n_ifo], order=order, K=K)
    cl = synth.make_cluster(net, P, seed=seed, snr=snr, random_rms=random_rms)
    return net, cl


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _assert_maps_equal(gpu_maps, ora_maps):
    g, o = _bits(gpu_maps), _bits(ora_maps)
    bad = np.argwhere(g != o)
    if bad.size:
        m, s = bad[0]
        raise AssertionError(f"{len(bad)} map entries differ; first map {m} sky {s}: "
                             f"gpu {gpu_maps[m, s]!r} oracle {ora_maps[m, s]!r}")


CASES = [
    # n_ifo, order, K, P, netCC, valid stride
    (3, 2, 6, 24, 0.5, 1),
    (3, 3, 8, 100, 0.5, 1),
    (1, 2, 2, 9, -1.0, 1),
    (2, 2, 5, 33, 0.0, 1),
    (4, 3, 10, 70, 0.5, 2),
    (5, 2, 12, 45, 0.3, 1),
    (3, 4, 20, 257, 0.5, 1),
    (3, 1, 0, 1, -1.0, 1),
]


@pytest.mark.parametrize("n_ifo,order,K,P,netCC,stride", CASES)
def test_plain_scan_bit_exact(oracle, n_ifo, order, K, P, netCC, stride):
    from pycwb_amd.likelihoodWP.dpf import calculate_dpf
    from pycwb_amd.likelihoodWP.sky_scan import scan_sky_for_best_fit

    print(f"case n={n_ifo} order={order} K={K} P={P}", flush=True)
    sys.stdout.flush()
    net, cl = _case(n_ifo, order, K, P, seed=100 + P)
    valid = np.arange(0, net.n_sky, stride, dtype=np.int64)
    En = 2 * 2.0 * n_ifo
    gamma = 0.5 * 0.5 * 2 / 3
    reg_g = calculate_dpf(net.FP_t, net.FX_t, cl["rms"], net.n_sky, n_ifo, gamma, En, valid)
    reg_o = oracle.calculate_dpf(net.FP_t, net.FX_t, cl["rms"], net.n_sky, n_ifo, gamma, En,
                                 valid)
    assert reg_g == reg_o
    REG = np.array([0.5 * np.sqrt(2), reg_o, 0.0], dtype=np.float32)
    g = scan_sky_for_best_fit(n_ifo, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                              cl["td90"], net.ml, REG, netCC, 0.5, En, valid)
    l_o, maps_o, smax_o = oracle.scan_sky(n_ifo, P, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                                          cl["td00"], cl["td90"], net.ml, REG, netCC, 0.5, En,
                                          valid)
    assert len(g) == 13
    maps_g = np.stack(g[1:12])
    _assert_maps_equal(maps_g, maps_o)
    assert g[0] == l_o
    assert g[12] == smax_o


def test_plain_scan_ties_and_masks(oracle):
    """Duplicated winning direction: the later index wins on the GPU too."""
    from pycwb_amd.likelihoodWP.sky_scan import scan_sky_for_best_fit

    net, cl = _case(3, 2, 6, 30, seed=7, snr=9.0)
    REG = np.array([0.5 * np.sqrt(2), 1.0, 0.0], dtype=np.float32)
    args = dict(netCC=0.0, delta_regulator=0.5, network_energy_threshold=12.0)
    FP, FX, ml = net.FP_t.copy(), net.FX_t.copy(), net.ml.copy()
    l, *_ = scan_sky_for_best_fit(3, 30, net.n_sky, FP, FX, cl["rms"], cl["td00"], cl["td90"],
                                  ml, REG, sky_valid_indices=np.arange(net.n_sky), **args)
    dst = net.n_sky - 1 if l != net.n_sky - 1 else 0
    FP[dst], FX[dst], ml[:, dst] = FP[l], FX[l], ml[:, l]
    res = scan_sky_for_best_fit(3, 30, net.n_sky, FP, FX, cl["rms"], cl["td00"], cl["td90"],
                                ml, REG, sky_valid_indices=np.arange(net.n_sky), **args)
    assert res[0] == max(l, dst)
    l_o, _, _ = oracle.scan_sky(3, 30, net.n_sky, FP, FX, cl["rms"], cl["td00"], cl["td90"], ml,
                                REG, 0.0, 0.5, 12.0, np.arange(net.n_sky))
    assert res[0] == l_o
    # masking the winner moves l_max elsewhere and zeroes its maps
    valid = np.array([i for i in range(net.n_sky) if i not in (l, dst)], dtype=np.int64)
    res2 = scan_sky_for_best_fit(3, 30, net.n_sky, FP, FX, cl["rms"], cl["td00"], cl["td90"],
                                 ml, REG, sky_valid_indices=valid, **args)
    assert res2[0] in set(valid.tolist())
    assert all(res2[k][l] == 0 and res2[k][dst] == 0 for k in range(1, 12))


def test_plain_scan_rejects_bad_shapes():
    from pycwb_amd.likelihoodWP.sky_scan import scan_sky_for_best_fit

    net, cl = _case(3, 1, 2, 5, seed=1)
    REG = np.zeros(3, np.float32)
    with pytest.raises(ValueError):
        scan_sky_for_best_fit(3, 5, net.n_sky, net.FP_t[:, :2], net.FX_t, cl["rms"], cl["td00"],
                              cl["td90"], net.ml, REG, 0.5, 0.5, 12.0, np.arange(net.n_sky))
    with pytest.raises(ValueError):
        scan_sky_for_best_fit(3, 5, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                              cl["td90"], net.ml, REG, 0.5, 0.5, 12.0, np.array([net.n_sky]))
    bad_ml = net.ml.copy()
    bad_ml[0, 0] = 3
    with pytest.raises(ValueError):
        scan_sky_for_best_fit(3, 5, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                              cl["td90"], bad_ml, REG, 0.5, 0.5, 12.0, np.arange(net.n_sky))


def test_plain_scan_full_order7_cluster(oracle):
    """Order 7 (196,608 directions), one 1,024-pixel cluster, K=64: the
    plain-config workload, bit-exact against the oracle."""
    from pycwb_amd.likelihoodWP.dpf import calculate_dpf
    from pycwb_amd.likelihoodWP.sky_scan import scan_sky_for_best_fit

    print("case order7 P=1024", flush=True)
    net, cl = _case(3, 7, 64, 1024, seed=1001 * 100000)
    valid = np.arange(net.n_sky, dtype=np.int64)
    gamma = 0.5 * 0.5 * 2 / 3
    reg = calculate_dpf(net.FP_t, net.FX_t, cl["rms"], net.n_sky, 3, gamma, 12.0, valid)
    REG = np.array([0.5 * np.sqrt(2), reg, 0.0], dtype=np.float32)
    g = scan_sky_for_best_fit(3, 1024, net.n_sky, net.FP_t, net.FX_t, cl["rms"], cl["td00"],
                              cl["td90"], net.ml, REG, 0.5, 0.5, 12.0, valid)
    l_o, maps_o, smax_o = oracle.scan_sky(3, 1024, net.n_sky, net.FP_t, net.FX_t, cl["rms"],
                                          cl["td00"], cl["td90"], net.ml, REG, 0.5, 0.5, 12.0,
                                          valid)
    _assert_maps_equal(np.stack(g[1:12]), maps_o)
    assert g[0] == l_o and g[12] == smax_o


@pytest.mark.parametrize("ifos,spread", [(synth.NETWORK_3, 0.1), (synth.NETWORK_3, 3.0),
                                         (synth.NETWORK_4, 1.0), (synth.NETWORK_5, 0.5)])
def test_regulator_with_bounds_matches_oracle(oracle, ifos, spread):
    """The regulator's bound-based decisions leave the count unchanged, also
    for strongly unbalanced noise weights (log-normal rms of the given spread)
    and several thresholds on both sides of the typical quality."""
    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import require_gpu
    from pycwb_amd.likelihoodWP.sky_scan import PlainClusterScan

    print(f"case regulator bounds {len(ifos)} det spread {spread}", flush=True)
    net = synth.make_network(ifos, order=4, K=16)
    n = net.n_ifo
    rng = np.random.default_rng(int(spread * 10) + n)
    dev = require_gpu()
    valid = np.arange(net.n_sky)
    for P in (1, 17, 160, 700):
        sigma = np.exp(rng.normal(0.0, spread, size=(n, P)))
        rms = synth.noise_weights(sigma)
        rms_d = torch.from_numpy(rms).to(dev)
        FP = torch.from_numpy(np.ascontiguousarray(net.FP_t)).to(dev)
        FX = torch.from_numpy(np.ascontiguousarray(net.FX_t)).to(dev)
        vd = torch.from_numpy(valid).to(dev)
        for gamma in (0.02, 0.1667, 0.3, 0.6):
            sc = PlainClusterScan(n, P, 33, dev)
            for flags in (_lib.DEFAULT_FLAGS, 0):
                sc.regulator(FP, FX, rms_d, net.n_sky, gamma, 4.0 * n,
                             float(np.float32(0.5 * np.sqrt(2))), vd, flags=flags)
                torch.cuda.synchronize()
                reg_g = float(sc.reg_f64.item())
                reg_o = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, n, gamma,
                                             4.0 * n, valid)
                assert reg_g == reg_o, (P, gamma, flags)
