"""The benchmark configurations at their own sizes on the GPU.

* batched likelihood: HEALPix order 7 (196,608 directions), K = 64, a mixed
  three/four-detector batch drawn from the benchmark's size distribution with
  the largest four-detector cluster (4,096 pixels), at the launch shape the
  autotuner picks: regulator, l_max and sky_stat_max bit-exact and Lm within
  1e-5 of the oracle for every cluster;
* gradient refinement: 512 three-detector clusters at order 7, 20 projected
  gradient steps: the refined statistic rises over its grid-argmax start in at
  least 95 % of clusters; value and angle gradients of 8 clusters against the
  float64 complex-step oracle.
"""

import math

import numpy as np
import pytest
import torch

from pycwb_amd import synth

pytestmark = pytest.mark.gpu

GAMMA, DELTA, NETCC = 0.5 * 0.5 * 2 / 3, 0.5, 0.5


def test_batched_order7_full_shape_matches_oracle(oracle):
    from pycwb_amd import autotune
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics

    print("case batched order 7, K=64, P up to 4096, autotuned shape", flush=True)
    dev = torch.device("cuda", 0)
    nets = [synth.make_network(synth.NETWORK_3, order=7, K=64),
            synth.make_network(synth.NETWORK_4, order=7, K=64)]
    sizes = np.array([4096, 2901, 1210, 733, 301, 97, 40, 16], np.int64)
    ids = np.array([1, 0, 1, 0, 1, 0, 1, 0], np.int64)
    batch = synth_batch(nets, sizes, ids, seed=1002 * 100000 + 7, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    tuned = autotune.tune_batched(nets, nt, valid, GAMMA, DELTA, NETCC)
    bs = BatchedSkyScan(batch, nt, valid, GAMMA, DELTA, NETCC, block=tuned["block"],
                        grid=tuned["grid"])
    bs.launch()
    r = bs.results()
    stats = SkyStatistics(batch, nt)
    stats.run(bs.l_max, bs.REG1, bs.desc.reg0)
    lm_gpu = stats.scalars()["Lm"]
    bs.close()
    vh = np.arange(nets[0].n_sky, dtype=np.int64)
    for c in range(sizes.size):
        net = nets[ids[c]]
        n, P = net.n_ifo, int(sizes[c])
        td = batch.cluster_td(c).cpu().numpy()
        td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
        td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
        rms = batch.cluster_rms(c).cpu().numpy()
        En = nt.En[ids[c]]
        reg = oracle.calculate_dpf(net.FP_t, net.FX_t, rms, net.n_sky, n, GAMMA, En, vh)
        assert reg == r["regulator"][c], (c, reg, r["regulator"][c])
        REG = np.array([bs.desc.reg0, reg, 0.0], np.float32)
        l_o, _, smax = oracle.scan_sky(n, P, net.n_sky, net.FP_t, net.FX_t, rms, td00, td90,
                                       net.ml, REG, NETCC, DELTA, En, vh)
        assert l_o == r["l_max"][c] and smax == r["sky_stat_max"][c], (c, P, n)
        st = oracle.sky_statistics(l_o, n, P, net.n_sky, net.FP_t, net.FX_t, rms, td00, td90,
                                   net.ml, REG, En)
        assert abs(lm_gpu[c] - st["Lm"]) <= 1e-5 * abs(st["Lm"]), (c, lm_gpu[c], st["Lm"])


def test_refinement_512_clusters_order7(oracle):
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch_smooth
    from pycwb_amd.skygrad import SkyGradient, network_geometry, refine_sky

    print("case refinement 512 clusters order 7", flush=True)
    dev = torch.device("cuda", 0)
    net = synth.make_network(synth.NETWORK_3, order=7, K=64)
    sizes = synth.cluster_sizes(512, seed=1003)
    ids = np.zeros(512, np.int64)
    b = synth_batch_smooth([net], sizes, ids, seed=1003 * 100000, snr=6.0, tap_sigma=2.0,
                           device=dev)
    nt = net_tables([net], device=dev)
    valid = torch.arange(net.n_sky, dtype=torch.int64, device=dev)
    bs = BatchedSkyScan(b, nt, valid, GAMMA, DELTA, NETCC, block=64)
    bs.launch()
    r = bs.results()
    eng = SkyGradient(b, nt, [net], r["REG1"], bs.desc.reg0)
    th0 = torch.from_numpy(np.pi / 2 - net.lat[r["l_max"]]).to(dev)
    ph0 = torch.from_numpy(net.phi[r["l_max"]]).to(dev)
    out = refine_sky(eng, th0, ph0, steps=20)
    v0, v = out["value0"].cpu().numpy(), out["value"].cpu().numpy()
    assert np.mean(v > v0) >= 0.95, np.mean(v > v0)
    # value and angle gradients at the refined directions against the oracle
    th, ph = out["theta"].cpu().numpy(), out["phi"].cpu().numpy()
    val, grad, _ = eng.run(out["theta"], out["phi"])
    val, grad = val.cpu().numpy(), grad.cpu().numpy()
    geo = network_geometry(net.ifos, net.td_rate)[:net.n_ifo]
    order = np.argsort(sizes)
    picks = np.concatenate([order[:3], order[250:252], order[-3:]])
    for c in picks:
        obj = oracle.SkyObjective(b.cluster_rms(int(c)).cpu().numpy(),
                                  b.cluster_td(int(c)).cpu().numpy(), geo, eng.reg0,
                                  float(r["REG1"][c]), nt.En[0])
        vo = obj.value(th[c], ph[c])
        go = obj.grad(th[c], ph[c])
        assert val[c] == pytest.approx(vo, rel=1e-10, abs=1e-10), c
        assert np.abs(grad[c] - go).max() <= 1e-8 * (np.abs(go).max() + 1e-6), c
    bs.close()


def test_statistics_at_directions_reduce_to_the_grid():
    """With amplitudes that do not change along the taps, interpolating the
    delays changes nothing: the best-sky statistics at a grid direction's
    continuous delays equal those at its integer taps (patterns computed at
    the direction agree with the grid tables to float32 rounding)."""
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics, statistics_at_directions

    dev = torch.device("cuda", 0)
    nets = [synth.make_network(synth.NETWORK_3, order=4, K=12),
            synth.make_network(synth.NETWORK_5, order=4, K=12)]
    sizes = np.array([60, 17, 230, 41], np.int64)
    ids = np.array([0, 1, 0, 1], np.int64)
    b = synth_batch(nets, sizes, ids, seed=77, snr=9.0, device=dev)
    for c in range(sizes.size):          # copy each pixel's middle tap to every tap
        v = b.cluster_td(c)
        v.copy_(v[:, :, 12:13, :].clone().expand_as(v))
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    bs = BatchedSkyScan(b, nt, valid, GAMMA, DELTA, 0.0, block=64, grid=16)
    bs.launch()
    r = bs.results()
    st = SkyStatistics(b, nt)
    st.run(bs.l_max, bs.REG1, bs.desc.reg0)
    grid = st.scalars()
    th = np.array([np.pi / 2 - nets[k].lat[l] for k, l in zip(ids, r["l_max"])])
    ph = np.array([nets[k].phi[l] for k, l in zip(ids, r["l_max"])])
    cont = statistics_at_directions(b, nets, th, ph, r["REG1"], bs.desc.reg0, nt.En)
    for k in ("Lm", "Em", "Ec", "Np", "rho"):
        np.testing.assert_allclose(cont[k], grid[k], rtol=1e-5, atol=1e-6, err_msg=k)
    bs.close()
