"""GPU cross-talk CSR builder against the oracle (bit-exact, same entry order).

The oracle restates getXTalk_pixels_fast (pycwb/modules/xtalk/monster.py:240-440),
whose rows are pinned against the all-pairs builder in tests/test_xtalk.py.
Cases: several clusters in one batch (multi-resolution blocks, repeated
pixels -- the later one owns the key, as in the reference's dictionary --,
pixels whose layer is not in the catalog, an empty cluster), one cluster
through get_xtalk_pixels_from_arrays, and the best-sky statistics fed the
device-resident CSR equal to the same CSR passed from the host.
"""

import numpy as np
import pytest

from pycwb_amd import xtalk as X

pytestmark = pytest.mark.gpu


def _cluster(rng, layers, n_per_res, t_span=8, dup=0, foreign=0):
    lay, tim = [], []
    for r, n in enumerate(n_per_res):
        L = int(layers[r])
        span = t_span * (int(layers[-1]) // L)
        ts = rng.integers(10, 10 + span, size=n)
        fs = rng.integers(0, L + 1, size=n)
        lay.append(np.full(n, L + 1))
        tim.append(ts * (L + 1) + fs)
    lay, tim = np.concatenate(lay), np.concatenate(tim)
    if dup:
        k = rng.choice(lay.size, size=dup)
        lay, tim = np.concatenate([lay, lay[k]]), np.concatenate([tim, tim[k]])
    if foreign:
        lay = np.concatenate([lay, np.full(foreign, 7)])      # 6 layers: not in the catalog
        tim = np.concatenate([tim, rng.integers(0, 500, size=foreign)])
    perm = rng.permutation(lay.size)
    return lay[perm].astype(np.int32), tim[perm].astype(np.int64)


def test_batch_matches_oracle(oracle):
    print("case xtalk batch", flush=True)
    rng = np.random.default_rng(12)
    cat = X.synthetic_catalog([4, 8, 16, 32], seed=12)
    clusters = [
        _cluster(rng, cat.layers, (20, 15, 10, 6)),
        _cluster(rng, cat.layers, (3, 0, 0, 2)),
        (np.zeros(0, np.int32), np.zeros(0, np.int64)),
        _cluster(rng, cat.layers, (30, 30, 20, 10), dup=6, foreign=3),
        _cluster(rng, cat.layers, (1, 0, 0, 0)),
        _cluster(rng, cat.layers, (60, 40, 30, 20), t_span=4),
    ]
    csr = cat.build_batch(clusters)
    assert csr.pix_off.tolist() == np.concatenate([[0], np.cumsum([c[0].size for c in clusters])]
                                                  ).tolist()
    total = 0
    for c, (lay, tim) in enumerate(clusters):
        lk_g, xt_g = csr.cluster(c)
        lk_o, xt_o = oracle.xtalk_pixels_fast(lay, tim, cat.layers, cat.lookup_table, cat.coeff)
        np.testing.assert_array_equal(lk_g, lk_o)
        assert xt_g.shape == xt_o.shape
        np.testing.assert_array_equal(xt_g.view(np.uint32), xt_o.view(np.uint32))
        total += xt_o.shape[0]
    assert csr.xt_off[-1] == total > 0


def test_flat_batch_equals_cluster_list():
    """build_batch_flat on the pixels laid end to end, from host arrays and
    from device tensors, gives build_batch's CSR; bad inputs raise."""
    import torch

    print("case xtalk flat batch", flush=True)
    rng = np.random.default_rng(21)
    cat = X.synthetic_catalog([4, 8, 16, 32], seed=21)
    clusters = [_cluster(rng, cat.layers, (12, 9, 5, 3)),
                (np.zeros(0, np.int32), np.zeros(0, np.int64)),
                _cluster(rng, cat.layers, (40, 20, 10, 5), dup=4, foreign=2),
                _cluster(rng, cat.layers, (2, 0, 1, 0))]
    sizes = [c[0].size for c in clusters]
    ref = cat.build_batch(clusters)
    fl = np.concatenate([c[0] for c in clusters])
    ft = np.concatenate([c[1] for c in clusters])
    dev = torch.device("cuda", 0)
    for lay, tim in ((fl, ft), (torch.from_numpy(fl).to(dev), torch.from_numpy(ft).to(dev))):
        got = cat.build_batch_flat(lay, tim, sizes)
        np.testing.assert_array_equal(got.pix_off, ref.pix_off)
        np.testing.assert_array_equal(got.xt_off, ref.xt_off)
        assert torch.equal(got.lookup, ref.lookup) and torch.equal(got.xt, ref.xt)
    with pytest.raises(ValueError):
        cat.build_batch_flat(fl, ft[:-1], sizes)
    with pytest.raises(ValueError):
        cat.build_batch_flat(fl * 0, ft, sizes)          # layers must be >= 1
    with pytest.raises(ValueError):
        cat.build_batch_flat(fl, ft, [])


def test_single_cluster_entry_points(oracle):
    from pycwb_amd.types import PixelArrays

    print("case xtalk single", flush=True)
    rng = np.random.default_rng(4)
    cat = X.synthetic_catalog([8, 16, 32], seed=4)
    lay, tim = _cluster(rng, cat.layers, (25, 20, 15))
    lk, xt = cat.get_xtalk_pixels_from_arrays(lay, tim, True)
    lk_o, xt_o = oracle.xtalk_pixels_fast(lay, tim, cat.layers, cat.lookup_table, cat.coeff)
    np.testing.assert_array_equal(lk, lk_o)
    np.testing.assert_array_equal(xt, xt_o)
    P = lay.size
    pa = PixelArrays.from_td(np.zeros((3, 2, P), np.float32), np.zeros((3, 2, P), np.float32),
                             np.ones((2, P), np.float32), time=tim, layers=lay)
    lk2, xt2 = cat.get_xtalk_pixels(pa)
    np.testing.assert_array_equal(lk2, lk_o)
    np.testing.assert_array_equal(xt2, xt_o)
    with pytest.raises(ValueError):
        cat.get_xtalk_pixels_from_arrays(lay, tim[:-1])


def test_statistics_take_device_csr(oracle):
    """SkyStatistics fed the builder's device CSR gives the scalars it gives
    for the same CSR passed from the host."""
    import torch

    from pycwb_amd import synth
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics

    print("case xtalk -> statistics", flush=True)
    dev = torch.device("cuda", 0)
    nets = [synth.make_network(synth.NETWORK_3, order=3, K=12)]
    sizes = np.array([40, 90, 25, 130])
    batch = synth_batch(nets, sizes, np.zeros(4, np.int64), seed=5, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    bs = BatchedSkyScan(batch, nt, valid, 0.5 * 0.5 * 2 / 3, 0.5, 0.5)
    bs.launch()
    rng = np.random.default_rng(5)
    cat = X.synthetic_catalog([8, 16, 32], seed=5)
    clusters = []
    for P in sizes:
        lay, tim = _cluster(rng, cat.layers, (P // 2, P // 3, P - P // 2 - P // 3))
        clusters.append((lay[:P], tim[:P]))
    csr = cat.build_batch(clusters)
    reg0 = float(np.float32(0.5 * np.sqrt(2)))
    a = SkyStatistics(batch, nt, csr)
    a.run(bs.l_max, bs.REG1, reg0)
    b = SkyStatistics(batch, nt, [csr.cluster(c) for c in range(len(sizes))])
    b.run(bs.l_max, bs.REG1, reg0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.scal.cpu().numpy(), b.scal.cpu().numpy())
    bs.close()
