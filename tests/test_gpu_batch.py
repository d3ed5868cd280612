"""Batched persistent-kernel scan: bit-exact against the plain family and the oracle."""

import numpy as np
import pytest
import torch

from pycwb_amd import synth

pytestmark = pytest.mark.gpu


def _nets(order, K):
    return [synth.make_network(synth.NETWORK_3, order=order, K=K),
            synth.make_network(synth.NETWORK_4, order=order, K=K)]


def _host_cluster(batch, c):
    td = batch.cluster_td(c).cpu().numpy()          # (P, n, nd, 2)
    td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
    td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
    return td00, td90, batch.cluster_rms(c).cpu().numpy()


@pytest.mark.parametrize("block,use_graph", [(256, True), (64, False), (128, True)])
def test_batched_matches_plain_and_oracle(oracle, block, use_graph):
    from pycwb_amd.batch import BatchedSkyScan, PlainLoop, net_tables, synth_batch

    print(f"case batch block={block} graph={use_graph}", flush=True)
    order, K = 3, 10
    nets = _nets(order, K)
    sizes = np.array([1, 5, 17, 64, 33, 200, 9, 120, 2, 77], dtype=np.int64)
    net_ids = np.arange(sizes.size) % 2
    dev = torch.device("cuda", 0)
    b = synth_batch(nets, sizes, net_ids, seed=77, snr=6.0, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(0, nets[0].n_sky, dtype=torch.int64, device=dev)
    gamma, delta, netCC = 0.5 * 0.5 * 2 / 3, 0.5, 0.5
    bs = BatchedSkyScan(b, nt, valid, gamma, delta, netCC, block=block, grid=37,
                        use_graph=use_graph)
    for _ in range(2):  # replays reuse the graph and must clear their counters
        bs.launch()
        rb = bs.results()
    pl = PlainLoop(b, nt, valid, gamma, delta, netCC)
    pl.launch()
    rp = pl.results()
    assert np.array_equal(rb["l_max"], rp["l_max"])
    assert np.array_equal(rb["sky_stat_max"], rp["sky_stat_max"])
    assert np.array_equal(rb["regulator"], rp["regulator"])
    vh = valid.cpu().numpy()
    for c in range(sizes.size):
        net = nets[net_ids[c]]
        td00, td90, rms = _host_cluster(b, c)
        En = nt.En[net_ids[c]]  # 2 * Acore**2 * n with Acore = sqrt(2), as the reference
        reg = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, net.n_ifo, gamma, En, vh)
        assert reg == rb["regulator"][c]
        REG = np.array([delta * np.sqrt(2), reg, 0], np.float32)
        l_o, _, s_o = oracle.scan_sky(net.n_ifo, int(sizes[c]), net.n_sky, net.FP_t, net.FX_t,
                                      rms, td00, td90, net.ml, REG, netCC, delta, En, vh)
        assert rb["l_max"][c] == l_o, c
        assert rb["sky_stat_max"][c] == s_o, c


def test_batched_with_sky_mask(oracle):
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch

    nets = _nets(2, 6)
    sizes = np.array([30, 3, 41], dtype=np.int64)
    dev = torch.device("cuda", 0)
    b = synth_batch(nets, sizes, np.array([0, 1, 0]), seed=5, snr=8.0, device=dev)
    nt = net_tables(nets, device=dev)
    vh = np.arange(1, nets[0].n_sky, 3, dtype=np.int64)
    valid = torch.from_numpy(vh).to(dev)
    bs = BatchedSkyScan(b, nt, valid, 0.2, 0.5, 0.0, block=64, grid=5)
    bs.launch()
    r = bs.results()
    for c, k in enumerate([0, 1, 0]):
        net = nets[k]
        td00, td90, rms = _host_cluster(b, c)
        En = nt.En[k]
        reg = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, net.n_ifo, 0.2, En, vh)
        REG = np.array([0.5 * np.sqrt(2), reg, 0], np.float32)
        l_o, _, s_o = oracle.scan_sky(net.n_ifo, int(sizes[c]), net.n_sky, net.FP_t, net.FX_t,
                                      rms, td00, td90, net.ml, REG, 0.0, 0.5, En, vh)
        assert r["l_max"][c] == l_o and r["sky_stat_max"][c] == s_o
        assert r["l_max"][c] in set(vh.tolist())


@pytest.mark.parametrize("gamma,grid", [(0.1667, 5), (0.12, 10), (0.25, 20), (0.1667, 400)])
def test_regulator_count_over_many_directions(oracle, gamma, grid):
    """The regulator's per-wavefront direction pools at order 5 (12,288
    directions, 192 tiles of 64 per cluster): per-direction sweeps that end
    at very different pixels (gamma near the middle of the quality
    distribution), counts equal to the oracle's full sweep.  The pool size
    follows the batch's work-item count: 8 tiles at grid 5, 4 at grid 10, 2
    at grid 20, 1 (one direction per lane) at grid 400."""
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch

    print(f"case regulator pools gamma={gamma} grid={grid}", flush=True)
    nets = _nets(5, 8)
    sizes = np.array([700, 1, 300, 64], dtype=np.int64)
    ids = np.array([0, 0, 1, 1])
    dev = torch.device("cuda", 0)
    b = synth_batch(nets, sizes, ids, seed=31, snr=4.0, device=dev)
    nt = net_tables(nets, device=dev)
    vh = np.arange(nets[0].n_sky, dtype=np.int64)
    valid = torch.from_numpy(vh).to(dev)
    bs = BatchedSkyScan(b, nt, valid, gamma, 0.5, 0.5, block=64, grid=grid)
    bs.launch()
    r = bs.results()
    for c, k in enumerate(ids):
        net = nets[k]
        rms = b.cluster_rms(c).cpu().numpy()
        reg, q = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, net.n_ifo, gamma,
                                      nt.En[k], vh, return_quality=True)
        assert r["regulator"][c] == reg, c
        if c == 0:  # a mixed sky: the count decides direction by direction
            assert 0.3 < float(np.mean(q > gamma)) < 0.95


def test_exact_skip_and_order_are_bit_identical():
    """The masked-pixel skip, the tile energy bound, the regulator early exit
    and the delay-tuple lane order change nothing in the results."""
    from pycwb_amd import _lib
    from pycwb_amd.batch import BatchedSkyScan, PlainLoop, net_tables, synth_batch

    nets = _nets(3, 12)
    sizes = np.array([3, 40, 250, 17, 90], dtype=np.int64)
    dev = torch.device("cuda", 0)
    b = synth_batch(nets, sizes, np.arange(5) % 2, seed=123, snr=5.0, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    outs = []
    skip_bound = _lib.SKIP_MASKED_PIXELS | _lib.TILE_ENERGY_BOUND
    for flags, ordered, block in [(0, False, 128), (_lib.DEFAULT_FLAGS, True, 128),
                                  (_lib.SKIP_MASKED_PIXELS, False, 128),
                                  (_lib.REGULATOR_EARLY_EXIT, True, 128),
                                  (skip_bound, True, 64), (skip_bound, False, 64),
                                  (_lib.REGULATOR_EARLY_EXIT | _lib.REGULATOR_BOUNDS, True, 64),
                                  (_lib.REGULATOR_EARLY_EXIT | _lib.REGULATOR_BOUNDS, False, 256),
                                  (_lib.DEFAULT_FLAGS, True, 256)]:
        print(f"case flags={flags} ordered={ordered} block={block}", flush=True)
        bs = BatchedSkyScan(b, nt, valid, 0.1667, 0.5, 0.5, block=block, grid=23, flags=flags,
                            ordered=ordered)
        bs.launch()
        outs.append(bs.results())
        pl = PlainLoop(b, nt, valid, 0.1667, 0.5, 0.5, flags=flags, ordered=ordered)
        pl.launch()
        outs.append(pl.results())
    for r in outs[1:]:
        for k in ("l_max", "sky_stat_max", "regulator"):
            assert np.array_equal(r[k], outs[0][k]), k


def test_plain_scan_with_non_finite_amplitudes(oracle):
    """A NaN and an inf delayed amplitude: the energy test keeps those pixels
    on the full path, and every map still equals the oracle's."""
    from pycwb_amd.likelihoodWP.sky_scan import scan_sky_for_best_fit

    net = synth.make_network(synth.NETWORK_3, order=2, K=4)
    cl = synth.make_cluster(net, 20, seed=8, snr=6.0)
    td00 = cl["td00"].copy()
    td00[2, 1, 3] = np.nan
    td00[5, 0, 7] = np.inf
    REG = np.array([0.5 * np.sqrt(2), 1.0, 0], np.float32)
    valid = np.arange(net.n_sky)
    g = scan_sky_for_best_fit(3, 20, net.n_sky, net.FP_t, net.FX_t, cl["rms"], td00, cl["td90"],
                              net.ml, REG, 0.0, 0.5, 12.0, valid)
    l_o, maps_o, s_o = oracle.scan_sky(3, 20, net.n_sky, net.FP_t, net.FX_t, cl["rms"], td00,
                                       cl["td90"], net.ml, REG, 0.0, 0.5, 12.0, valid)
    mg = np.stack(g[1:12])
    both_nan = np.isnan(mg) & np.isnan(maps_o)
    assert np.array_equal(np.where(both_nan, 0, mg).view(np.uint32),
                          np.where(both_nan, 0, maps_o).view(np.uint32))
    assert g[0] == l_o and (g[12] == s_o or (np.isnan(g[12]) and np.isnan(s_o)))
