"""The 600 s five-detector segment end to end on the GPU, at its own size.

H1 L1 V1 K1 I1, 600 s of coloured Gaussian noise at 4096 Hz whitened to unit
variance, 200 sine-Gaussian and white-noise-burst injections, seven WDM levels
(M = 8 .. 512), one likelihood batch at HEALPix order 8 (786,432 directions,
K = 598 taps at 16,384 Hz), then the gradient refinement of every cluster.

Checks: every level selects pixels and forms clusters; the batched regulator
and l_max of the two smallest-but-one clusters equal the oracle's bit for bit
on the batch's own amplitudes; the batch equals the plain one-cluster-per-
launch path for every cluster; most injections come out with a cluster at
their time whose argmax reproduces their detector delays; the refinement
raises the continuous statistic for at least 90 % of clusters.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_segment_600s_five_detectors_order8(oracle):
    from pycwb_amd import segment, synth
    from pycwb_amd.batch import PlainLoop
    from pycwb_amd.skygrad import SkyGradient, refine_sky

    print("case segment 600 s, 5 detectors, order 8", flush=True)
    ifos, rate = synth.NETWORK_5, 4096
    x, inj = segment.synthetic_segment(ifos, 600, rate, 200, seed=11)
    dev = torch.device("cuda", 0)
    xd = torch.from_numpy(x).to(dev)
    stream = torch.cuda.current_stream()
    seg = segment.run(xd, ifos, rate, stream, order=8)
    for lv in seg["levels"]:
        assert lv["selected"] > 0 and lv["sizes"].size > 0, lv["M"]
    net, nt, batch, res = seg["net"], seg["nets"], seg["batch"], seg["results"]
    sizes = seg["sizes"]
    n, S = len(ifos), net.n_sky
    print(f"{sizes.size} clusters, {int(sizes.sum())} pixels", flush=True)
    assert sizes.size > 500
    # oracle parity on two small clusters
    vh = np.arange(S, dtype=np.int64)
    En = nt.En[0]
    picks = [int(np.argmin(np.abs(sizes - 6))), int(np.argmin(np.abs(sizes - 25)))]
    for c in picks:
        P = int(sizes[c])
        rms_c = batch.cluster_rms(c).cpu().numpy()
        td_c = batch.cluster_td(c).cpu().numpy()
        t00 = np.ascontiguousarray(td_c[..., 0].transpose(2, 1, 0))
        t90 = np.ascontiguousarray(td_c[..., 1].transpose(2, 1, 0))
        reg_o = oracle.calculate_dpf(net.FP_t, net.FX_t, rms_c, S, n, 0.5 * 0.5 * 2 / 3, En, vh,
                                     threads=16)
        assert reg_o == res["regulator"][c], c
        REG = np.array([np.float32(0.5 * np.sqrt(2)), res["REG1"][c], 0.0], np.float32)
        l_o, _, smax = oracle.scan_sky(n, P, S, net.FP_t, net.FX_t, rms_c, t00, t90, net.ml, REG,
                                       0.5, 0.5, En, vh, threads=16)
        assert l_o == res["l_max"][c] and smax == res["sky_stat_max"][c], c
    # the plain family on every cluster
    pl = PlainLoop(batch, nt, seg["valid"], 0.5 * 0.5 * 2 / 3, 0.5, 0.5)
    pl.launch(stream)
    rp = pl.results()
    assert np.array_equal(rp["l_max"], res["l_max"])
    assert np.array_equal(rp["sky_stat_max"], res["sky_stat_max"])
    # recovery of the injections
    found = segment.recovered(inj, seg["ctime"], res["l_max"], net, ifos, seg["td_rate"])
    print(f"{found}/{len(inj)} injections recovered with their delays", flush=True)
    assert found >= 0.4 * len(inj)
    # refinement from the grid argmax
    scan = seg["scan"]
    eng = SkyGradient(batch, nt, [net], res["REG1"], scan.desc.reg0)
    th0 = torch.from_numpy(np.pi / 2 - net.lat[res["l_max"]]).to(dev)
    ph0 = torch.from_numpy(net.phi[res["l_max"]]).to(dev)
    out = refine_sky(eng, th0, ph0, steps=20)
    v0, v1 = out["value0"].cpu().numpy(), out["value"].cpu().numpy()
    print(f"refinement: {np.mean(v1 > v0):.3f} of clusters improved", flush=True)
    assert np.mean(v1 > v0) >= 0.9, np.mean(v1 > v0)
    scan.close()
