"""GPU likelihood stage with the call signatures of PycWB's likelihoodWP facade.

Drop-in for pycwb/modules/likelihoodWP/likelihood.py:
  setup_likelihood   = prepare_likelihood_inputs   (likelihood_setup.py:81-238)
  likelihood         = evaluate_cluster_likelihood (likelihood.py:138-462)
  likelihood_wrapper = evaluate_fragment_clusters  (likelihood.py:58-135)

Per cluster everything from the regulator to the best-sky statistics runs on
the GPU (HIP kernels through the C ABI), enqueued on one stream with a single
device-to-host transfer at the end.  ``likelihood_wrapper`` evaluates all
clusters of all lags as one batch: one hipGraph launch for the regulator, scan
and argmax of every cluster, one launch for every cluster's best-sky
statistics, and plain-kernel sky maps only for accepted clusters.

Cross-talk: ``xtalk`` is a catalog (``pycwb_amd.xtalk.XTalk``, loaded from
``MRAcatalog`` when ``xtalk`` is None; the batch builds every cluster's CSR in
one GPU pass) or any object with ``get_xtalk_pixels(pixel_arrays, check)``
returning the CSR ``(lookup, xtalk)`` of pycwb/modules/xtalk (``SelfXTalk``
gives the self-only CSR).

Not computed here (waveform-domain quantities of populate_detection_statistics
that need the MRA waveform reconstruction): net_cc, sky_cc, sub_net, like_net,
energy, net_null, net_ed, norm, per-detector waveform SNRs, chirp mass.
"""

from __future__ import annotations

import logging
import math
import time

import numpy as np
import torch

from .. import sky
from ..batch import BatchedSkyScan, ClusterBatch, NetTablesDev
from ..device import require_gpu
from ..skystats import SkyStatistics as _GpuSkyStats
from ..skystats import SCALARS, self_xtalk
from ..xtalk import XTalk
from .detection_statistics import get_likelihood_rejection_reason
from .sky_mask import compute_sky_valid_indices, sky_valid_indices_for_cluster
from .sky_scan import MAP_NAMES, PlainClusterScan, scan_flags
from .typing import SkyMapStatistics, SkyStatistics

logger = logging.getLogger(__name__)


class SelfXTalk:
    """Cross-talk provider with each pixel its own only neighbour."""

    def get_xtalk_pixels(self, pixel_arrays, check=True):
        return self_xtalk(len(pixel_arrays))


def prepare_likelihood_inputs(config, strains, nIFO, ml=None, FP=None, FX=None, ml_big=None,
                              FP_big=None, FX_big=None, big_cluster_healpix_order=None) -> dict:
    """Segment-level constants and sky tables (likelihood_setup.py:81-238)."""
    if config is None:
        raise ValueError("config is required for pure-Python likelihood")
    acor = float(getattr(config, "Acore"))
    gamma = float(getattr(config, "gamma", 0.0))
    delta = float(getattr(config, "delta", 0.0))
    net_rho = float(getattr(config, "netRHO", 0.0))
    netCC = float(getattr(config, "netCC", 0.0))
    xgb = bool(getattr(config, "xgb_rho_mode", False))
    En = 2 * acor * acor * nIFO
    gamma_regulator = gamma * gamma * 2 / 3
    if delta == 0.0:
        delta = 0.00001
    delta_regulator = abs(delta) if abs(delta) < 1 else 1
    if ml is None or FP is None or FX is None:
        up = int(getattr(config, "upTDF", 1))
        rate = float(getattr(config, "TDRate", None) or int(getattr(config, "rateANA")) * up)
        td_size = max(int(getattr(config, "TDSize")) * up,
                      int(getattr(config, "max_delay", 0.0) * rate) + 1)
        ml, FP, FX = sky.compute_sky_delay_and_patterns(
            list(getattr(config, "ifo")), getattr(config, "refIFO"), rate, td_size,
            healpix_order=int(getattr(config, "healpix", 0)) or None)
    ml, FP, FX = np.asarray(ml), np.asarray(FP), np.asarray(FX)
    n_sky = int(ml.shape[1])
    healpix_order = int(getattr(config, "healpix", 0)) if hasattr(config, "healpix") else None
    phi, lat = sky.build_sky_directions(n_sky, healpix_order)
    mask_cfg = getattr(config, "sky_mask", None)
    t_ref = float(strains[0].t0) if strains is not None and len(strains) > 0 and \
        hasattr(strains[0], "t0") else None
    valid = compute_sky_valid_indices(phi, lat, mask_cfg, t_ref=t_ref)
    big = {}
    if ml_big is not None and FP_big is not None and FX_big is not None:
        n_big = int(np.asarray(ml_big).shape[1])
        phi_b, lat_b = sky.build_sky_directions(n_big, big_cluster_healpix_order)
        big = {
            "ml_big_cluster": np.asarray(ml_big),
            "FP_big_cluster_t": np.asarray(FP_big).T.astype(np.float32),
            "FX_big_cluster_t": np.asarray(FX_big).T.astype(np.float32),
            "n_sky_big_cluster": n_big,
            "phi_geo_arr_big_cluster": phi_b,
            "latitude_arr_big_cluster": lat_b,
            "sky_valid_indices_big": compute_sky_valid_indices(phi_b, lat_b, mask_cfg, t_ref),
        }
    out = {
        "network_energy_threshold": En,
        "xgb_rho_mode": xgb,
        "gamma_regulator": gamma_regulator,
        "delta_regulator": delta_regulator,
        "net_rho_threshold": abs(net_rho),
        "netEC_threshold": abs(net_rho) * abs(net_rho) * 2,
        "netCC": netCC,
        "ml": ml, "FP": FP, "FX": FX,
        "FP_t": FP.T.astype(np.float32), "FX_t": FX.T.astype(np.float32),
        "n_sky": n_sky, "healpix_order": healpix_order,
        "phi_geo_arr": phi, "latitude_arr": lat, "ra_arr": phi, "dec_arr": lat,
        "sky_mask_config": mask_cfg, "segment_start_gps": t_ref,
        "sky_valid_indices": valid,
        "ml_big_cluster": None, "FP_big_cluster_t": None, "FX_big_cluster_t": None,
        "n_sky_big_cluster": None, "big_cluster_healpix_order": big_cluster_healpix_order,
        "phi_geo_arr_big_cluster": None, "latitude_arr_big_cluster": None,
        "sky_valid_indices_big": None,
        "_ifos": list(getattr(config, "ifo", []) or []),
    }
    out.update(big)
    return out


def _device_net(setup: dict, big: bool, device) -> tuple:
    """Upload (once per setup and grid) the sky tables; returns (NetTablesDev, valid)."""
    key = "_dev_big" if big else "_dev"
    if key not in setup:
        ml = setup["ml_big_cluster"] if big else setup["ml"]
        FP = setup["FP_big_cluster_t"] if big else setup["FP_t"]
        FX = setup["FX_big_cluster_t"] if big else setup["FX_t"]
        n_ifo, n_sky = ml.shape
        En = float(setup["network_energy_threshold"])
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)
        nets = NetTablesDev(
            n_sky=n_sky, nifo=[n_ifo], En=[En],
            off=torch.zeros(1, dtype=torch.int64, device=device),
            FP=t(FP, torch.float32).reshape(-1), FX=t(FX, torch.float32).reshape(-1),
            ml=t(ml, torch.int32).reshape(-1),
            nifo_d=torch.tensor([n_ifo], dtype=torch.int32, device=device),
            En_d=torch.tensor([En], dtype=torch.float64, device=device),
            K=[int(np.abs(ml).max())], ml_host=[np.ascontiguousarray(ml, np.int32)],
            side_host=[_mirror_side(setup, big)])
        setup[key] = nets
    return setup[key]


def _mirror_side(setup: dict, big: bool):
    """Mirror sheet of each grid direction (scan lane order only), when the
    detector names are known."""
    ifos = setup.get("_ifos")
    if not ifos:
        return None
    phi = setup["phi_geo_arr_big_cluster" if big else "phi_geo_arr"]
    lat = setup["latitude_arr_big_cluster" if big else "latitude_arr"]
    try:
        return sky.mirror_side(list(ifos), np.asarray(phi), np.asarray(lat))
    except ValueError:  # detector not in the built-in table
        return None


def _noise_weights(pixel_arrays) -> np.ndarray:
    """(P, n) float32 weights (pixel_data.py:60-62, transposed as likelihood.py:305)."""
    inv = 1.0 / np.asarray(pixel_arrays.noise_rms).astype(np.float64)
    norm = 1.0 / np.sqrt(np.sum(inv ** 2, axis=0))
    return np.ascontiguousarray((inv * norm[np.newaxis, :]).astype(np.float32).T)


def _packed_td(pixel_arrays, device) -> tuple[torch.Tensor, int]:
    """td_amp_dense (P, n, 2*nd) -> (P, n, nd, 2) {00, 90} pairs on the device."""
    dense = np.asarray(pixel_arrays.td_amp_dense(), dtype=np.float32)
    P, n, ts = dense.shape
    nd = ts // 2
    if nd % 2 != 1:
        raise ValueError(f"TD amplitude tables must hold 2K+1 taps per phase, got {nd}")
    d = torch.from_numpy(np.ascontiguousarray(dense)).to(device)
    packed = torch.stack([d[:, :, :nd], d[:, :, nd:2 * nd]], dim=-1).contiguous()
    return packed, nd


def _single_batch(P, n, K, rms_d, packed) -> ClusterBatch:
    return ClusterBatch(n_pix=np.array([P], np.int32), net=np.zeros(1, np.int32),
                        K=np.array([K], np.int32), nifo=np.array([n], np.int32),
                        rms_off=np.zeros(1, np.int64), td_off=np.zeros(1, np.int64),
                        rms=rms_d.reshape(-1), td=packed.reshape(-1))


def _sky_statistics_from_gpu(st: _GpuSkyStats, c: int) -> SkyStatistics:
    """SkyStatistics of cluster c.  p00_POL/p90_POL come from the kernel
    (avx_pol_ps).  r00_POL/r90_POL are zeros: the reference computes them from
    the first avx_pol_ps call's output arrays, which it leaves uninitialised
    except at the last pixel (packet_ops.py:593-655, sky_statistics.py:246-251),
    so they carry no defined values to reproduce."""
    sc = st.scal[c].cpu().numpy()
    s = dict(zip(SCALARS, sc.tolist()))
    a = st.cluster_arrays(c)
    z = np.zeros_like(a["pixel_mask"])
    f32 = np.float32
    return SkyStatistics(
        Gn=f32(s["Gn"]), Ec=f32(s["Ec"]), Dc=f32(s["Dc"]), Rc=f32(s["Rc"]), Eh=f32(s["Eh"]),
        Es=f32(s["Es"]), Np=f32(s["Np"]), Em=f32(s["Em"]), Lm=f32(s["Lm"]),
        norm=f32(s["norm"]), cc=f32(s["cc"]), rho=f32(s["rho"]), xrho=f32(s["xrho"]),
        Lo=f32(s["Lo"]), Eo=f32(s["Eo"]), N_pix_effective=s["N_pix_effective"],
        energy_array_plus=a["energy_array_plus"], energy_array_cross=a["energy_array_cross"],
        v00=a["v00"], v90=a["v90"], pd=a["pd"], pD=a["pD"], ps=a["ps"], pS=a["pS"],
        pixel_mask=a["pixel_mask"], gaussian_noise_correction=a["gaussian_noise_correction"],
        noise_amplitude_00=a["noise_amplitude_00"], noise_amplitude_90=a["noise_amplitude_90"],
        coherent_energy=a["coherent_energy"],
        p00_POL=(a["p00_POL_r"], a["p00_POL_a"]), p90_POL=(a["p90_POL_r"], a["p90_POL_a"]),
        r00_POL=(z, z), r90_POL=(z, z), S_snr=a["S_snr"], f=a["f"], F=a["F"])


def _write_cluster(cluster, sky_statistics, skymap, l_max, theta, phi, ml_col, xgb):
    pa = cluster.pixel_arrays
    pa.set_waveform_data(wave=np.asarray(sky_statistics.pd, np.float32),
                         w_90=np.asarray(sky_statistics.pD, np.float32),
                         asnr=np.asarray(sky_statistics.ps, np.float32),
                         a_90=np.asarray(sky_statistics.pS, np.float32),
                         core_mask=sky_statistics.pixel_mask,
                         energy_plus=sky_statistics.energy_array_plus,
                         energy_cross=sky_statistics.energy_array_cross)
    m = cluster.cluster_meta
    m.sub_net2 = float(skymap.nCorrelation[l_max])
    m.like_sky = float(sky_statistics.Em)
    m.energy_sky = float(sky_statistics.Eo)
    m.net_ecor = float(sky_statistics.Ec)
    m.norm_cor = float(sky_statistics.Ec * sky_statistics.Rc)
    if xgb:
        m.net_rho = float(sky_statistics.rho)
    else:
        m.net_rho2 = float(sky_statistics.rho)
    m.g_net = float(skymap.nAntennaPrior[l_max])
    m.a_net = float(skymap.nAlignment[l_max])
    m.i_net = 0
    m.ndof = float(sky_statistics.N_pix_effective)
    m.sky_chi2 = float(skymap.nDisbalance[l_max])
    m.g_noise = float(sky_statistics.Gn)
    m.iota = 0.0
    m.psi = 0.0
    m.ellipticity = 0
    m.l_max = int(l_max)
    m.theta = theta
    m.phi = phi
    if getattr(m, "c_time", 0.0) == 0.0:
        m.c_time = getattr(cluster, "cluster_time", 0.0)
    if getattr(m, "c_freq", 0.0) == 0.0:
        m.c_freq = getattr(cluster, "cluster_freq", 0.0)
    cluster.sky_time_delay = [float(v) for v in ml_col]
    cluster.cluster_status = -1


def _select_grid(setup, config, n_pixels):
    precision = int(abs(getattr(config, "precision", 0) or 0))
    csize = precision % 65536
    nres = int(getattr(config, "nRES", 1) or 1)
    return (csize > 0 and n_pixels > nres * csize and setup.get("ml_big_cluster") is not None)


def evaluate_cluster_likelihood(nIFO, cluster, config, MRAcatalog=None, strains=None,
                                cluster_id=None, nRMS=None, setup=None, xtalk=None,
                                supercluster_setup=None):
    """One cluster (likelihood.py:138-462): returns (cluster, SkyMapStatistics)
    or (None, None) when rejected."""
    if xtalk is None:
        if MRAcatalog is None:
            raise ValueError("likelihood(): xtalk or MRAcatalog must be provided. For "
                             "multi-cluster / multi-lag use, call evaluate_fragment_clusters()"
                             " instead.")
        xtalk = XTalk.load(MRAcatalog, dump=True)
    if setup is None:
        sc = supercluster_setup or {}
        sc_ml = sc.get("ml_likelihood", sc.get("ml"))
        sc_fp = sc.get("FP_likelihood", sc.get("FP"))
        sc_fx = sc.get("FX_likelihood", sc.get("FX"))
        if strains is None and sc_ml is None:
            raise ValueError("likelihood(): setup, strains, or supercluster_setup must be "
                             "provided.")
        # the coarse grid of big clusters comes from the supercluster stage too
        # (likelihood.py:212-219)
        setup = prepare_likelihood_inputs(
            config, strains, nIFO, ml=sc_ml, FP=sc_fp, FX=sc_fx,
            ml_big=sc.get("ml_big_cluster"), FP_big=sc.get("FP_big_cluster"),
            FX_big=sc.get("FX_big_cluster"),
            big_cluster_healpix_order=sc.get("big_cluster_healpix_order"))
    if config is None:
        raise ValueError("likelihood(): config must be provided.")
    t_start = time.perf_counter()
    dev = require_gpu()
    pa = cluster.pixel_arrays
    # noise floor of each pixel from the TF noise maps (likelihood.py:260-263):
    # the noise weights below are derived from it
    if nRMS is not None and len(nRMS) == nIFO:
        pa.populate_noise_rms(nRMS)
    n_pix = len(pa)
    big = _select_grid(setup, config, n_pix)
    nets = _device_net(setup, big, dev)
    valid_h = sky_valid_indices_for_cluster(setup, cluster, use_big_grid=big)
    if valid_h is None:
        valid_h = np.arange(nets.n_sky, dtype=np.int64)
    valid_h = np.asarray(valid_h, dtype=np.int64)
    if len(valid_h) == 0:
        logger.info("Cluster-id=%s rejected: sky mask selects no grid directions", cluster_id)
        return None, None
    lat = setup["latitude_arr_big_cluster" if big else "latitude_arr"]
    phi = setup["phi_geo_arr_big_cluster" if big else "phi_geo_arr"]
    ml_h = setup["ml_big_cluster" if big else "ml"]
    lookup, xt = xtalk.get_xtalk_pixels(pa, True)

    rms_h = _noise_weights(pa)
    packed, nd = _packed_td(pa, dev)
    K = nd // 2
    if int(np.abs(ml_h).max()) > K:
        raise ValueError(f"|ml| exceeds the TD half-width K={K}")
    rms_d = torch.from_numpy(rms_h).to(dev)
    valid = torch.from_numpy(valid_h).to(dev)
    En = float(setup["network_energy_threshold"])
    reg0 = float(np.float32(setup["delta_regulator"] * np.sqrt(2)))
    n_sky = nets.n_sky
    scan = PlainClusterScan(nIFO, n_pix, nd, dev)
    flags = scan_flags(rms_d, nets.FP, nets.FX)
    scan.regulator(nets.FP, nets.FX, rms_d, n_sky, setup["gamma_regulator"], En, reg0, valid,
                   flags=flags)
    maps = torch.empty((len(MAP_NAMES), n_sky), dtype=torch.float32, device=dev)
    scan.scan(nets.FP, nets.FX, nets.ml, rms_d, n_sky, setup["netCC"], En, valid, maps,
              td_packed=packed, sky_order=nets.sky_order(valid_h, dev)[0], flags=flags)
    batch = _single_batch(n_pix, nIFO, K, rms_d, packed)
    st = _GpuSkyStats(batch, nets, [(lookup, xt)])
    st.run(scan.l_max, scan.REG[1:2], reg0, setup["xgb_rho_mode"])
    s64 = maps[6].double()
    e = torch.exp(s64 - s64.max())
    prob = (e / e.sum()).float()
    m = maps.cpu().numpy()
    l_max = int(scan.l_max.item())
    stat_max = float(scan.stat_max.item())
    skymap = SkyMapStatistics(l_max, *[m[i] for i in range(len(MAP_NAMES))], stat_max)
    if skymap.sky_stat_max <= 0.0:
        logger.info("Cluster rejected: non-positive sky_stat_max=%.3f", skymap.sky_stat_max)
        return None, None
    skymap.nProbability = prob.cpu().numpy()
    sky_statistics = _sky_statistics_from_gpu(st, 0)
    rejected = get_likelihood_rejection_reason(
        sky_statistics, En, setup["netEC_threshold"],
        net_rho_threshold=setup["net_rho_threshold"], xgb_rho_mode=setup["xgb_rho_mode"])
    if rejected:
        logger.debug("Cluster rejected due to threshold cuts: %s", rejected)
        return None, None
    theta = float(np.clip(np.degrees(float(np.pi / 2.0 - lat[l_max])), 0.0, 180.0))
    phi_deg = float(np.degrees(float(phi[l_max]))) % 360.0
    _write_cluster(cluster, sky_statistics, skymap, l_max, theta, phi_deg, ml_h[:, l_max],
                   setup["xgb_rho_mode"])
    skymap.stage_timings = {"total": time.perf_counter() - t_start}
    return cluster, skymap


def evaluate_fragment_clusters(config, fragment_clusters, strains, MRAcatalog, nRMS=None,
                               xtalk=None):
    """All clusters of all lags as one GPU batch (likelihood.py:58-135).

    Returns ``results[lag] = [(cluster, SkyMapStatistics), ...]`` for the clusters
    accepted in that lag, in cluster order.
    """
    if xtalk is None:
        xtalk = XTalk.load(MRAcatalog, dump=True)
    t0 = time.perf_counter()
    nIFO = int(getattr(config, "nIFO"))
    setup = prepare_likelihood_inputs(config, strains, nIFO)
    dev = require_gpu()
    todo = []   # (lag, k, cluster)
    for lag, frag in enumerate(fragment_clusters):
        for k, cl in enumerate(frag.clusters):
            if cl.cluster_status > 0:
                continue
            cl.cluster_id = k + 1
            todo.append((lag, k, cl))
    results = [[] for _ in fragment_clusters]
    # The batch scans every cluster over the segment's valid set on the fine
    # grid.  Big-grid clusters, and clusters whose mask differs from the
    # segment's at their own time (ICRS masks), take the per-cluster path.
    seg_valid = np.asarray(setup["sky_valid_indices"], dtype=np.int64)
    batchable, single = [], []
    for item in todo:
        cl = item[2]
        big = _select_grid(setup, config, len(cl.pixel_arrays))
        own = None if big else sky_valid_indices_for_cluster(setup, cl)
        same = own is None or np.array_equal(np.asarray(own, dtype=np.int64), seg_valid)
        (batchable if not big and same else single).append(item)
    accepted = {}
    if batchable:
        if nRMS is not None and len(nRMS) == nIFO:
            for _, _, cl in batchable:
                cl.pixel_arrays.populate_noise_rms(nRMS)
        accepted.update(_evaluate_batch(config, setup, batchable, xtalk, nIFO, dev))
    for lag, k, cl in single:
        res = evaluate_cluster_likelihood(nIFO, cl, config, cluster_id=k + 1, nRMS=nRMS,
                                          setup=setup, xtalk=xtalk)
        if res[0] is not None:
            accepted[(lag, k)] = res
    for (lag, k) in sorted(accepted):
        results[lag].append(accepted[(lag, k)])
    logger.info("Likelihood wrapper done: %d accepted cluster(s) across %d lag(s) in %.2f s",
                sum(len(r) for r in results), len(fragment_clusters), time.perf_counter() - t0)
    return results


def _evaluate_batch(config, setup, items, xtalk, nIFO, dev) -> dict:
    nets = _device_net(setup, False, dev)
    valid_h = np.asarray(setup["sky_valid_indices"], dtype=np.int64)
    if len(valid_h) == 0:
        return {}
    valid = torch.from_numpy(valid_h).to(dev)
    ml_h = setup["ml"]
    n_pix, Ks, rms_parts, td_parts, xts = [], [], [], [], []
    for _, _, cl in items:
        pa = cl.pixel_arrays
        rms_parts.append(torch.from_numpy(_noise_weights(pa)).reshape(-1))
        packed, nd = _packed_td(pa, dev)
        td_parts.append(packed.reshape(-1))
        n_pix.append(len(pa))
        Ks.append(nd // 2)
        if not hasattr(xtalk, "build_batch"):
            xts.append(xtalk.get_xtalk_pixels(pa, True))
    if hasattr(xtalk, "build_batch"):   # a catalog: every cluster's CSR in one GPU pass
        xts = xtalk.build_batch([(cl.pixel_arrays.layers, cl.pixel_arrays.time)
                                 for _, _, cl in items], device=dev)
    if int(np.abs(ml_h).max()) > min(Ks):
        raise ValueError("|ml| exceeds the TD half-width of a cluster")
    n_pix = np.asarray(n_pix, np.int64)
    Ks = np.asarray(Ks, np.int64)
    rms_off = np.concatenate([[0], np.cumsum(n_pix * nIFO)[:-1]])
    td_off = np.concatenate([[0], np.cumsum(n_pix * nIFO * (2 * Ks + 1))[:-1]])
    batch = ClusterBatch(n_pix=n_pix.astype(np.int32), net=np.zeros(len(items), np.int32),
                         K=Ks.astype(np.int32), nifo=np.full(len(items), nIFO, np.int32),
                         rms_off=rms_off.astype(np.int64), td_off=td_off.astype(np.int64),
                         rms=torch.cat(rms_parts).to(dev), td=torch.cat(td_parts))
    bs = BatchedSkyScan(batch, nets, valid, setup["gamma_regulator"], setup["delta_regulator"],
                        setup["netCC"])
    bs.launch()
    reg0 = float(np.float32(setup["delta_regulator"] * np.sqrt(2)))
    st = _GpuSkyStats(batch, nets, xts)
    st.run(bs.l_max, bs.REG1, reg0, setup["xgb_rho_mode"])
    res = bs.results()
    scal = st.scal.cpu().numpy()
    En = float(setup["network_energy_threshold"])
    out = {}
    for c, (lag, k, cl) in enumerate(items):
        if res["sky_stat_max"][c] <= 0.0:
            continue
        s = dict(zip(SCALARS, scal[c].tolist()))
        probe = type("S", (), {key: np.float32(v) for key, v in s.items()})()
        probe.N_pix_effective = s["N_pix_effective"]
        if get_likelihood_rejection_reason(probe, En, setup["netEC_threshold"],
                                           net_rho_threshold=setup["net_rho_threshold"],
                                           xgb_rho_mode=setup["xgb_rho_mode"]):
            continue
        # accepted: sky maps from the plain kernels on the same device data
        n_sky = nets.n_sky
        P, K = int(n_pix[c]), int(Ks[c])
        scan = PlainClusterScan(nIFO, P, 2 * K + 1, dev)
        rms_c = batch.cluster_rms(c).contiguous()
        td_c = batch.cluster_td(c).contiguous()
        scan.regulator(nets.FP, nets.FX, rms_c, n_sky, setup["gamma_regulator"], En, reg0, valid,
                       flags=bs.flags)
        maps = torch.empty((len(MAP_NAMES), n_sky), dtype=torch.float32, device=dev)
        scan.scan(nets.FP, nets.FX, nets.ml, rms_c, n_sky, setup["netCC"], En, valid, maps,
                  td_packed=td_c, sky_order=bs.sky_order[0], flags=bs.flags)
        s64 = maps[6].double()
        e = torch.exp(s64 - s64.max())
        m = maps.cpu().numpy()
        l_max = int(res["l_max"][c])
        skymap = SkyMapStatistics(l_max, *[m[i] for i in range(len(MAP_NAMES))],
                                  float(res["sky_stat_max"][c]))
        skymap.nProbability = (e / e.sum()).float().cpu().numpy()
        sky_statistics = _sky_statistics_from_gpu(st, c)
        lat, phi = setup["latitude_arr"], setup["phi_geo_arr"]
        theta = float(np.clip(np.degrees(float(np.pi / 2.0 - lat[l_max])), 0.0, 180.0))
        phi_deg = float(np.degrees(float(phi[l_max]))) % 360.0
        _write_cluster(cl, sky_statistics, skymap, l_max, theta, phi_deg, ml_h[:, l_max],
                       setup["xgb_rho_mode"])
        out[(lag, k)] = (cl, skymap)
    bs.close()
    return out


setup_likelihood = prepare_likelihood_inputs
likelihood = evaluate_cluster_likelihood
likelihood_wrapper = evaluate_fragment_clusters

__all__ = ["setup_likelihood", "likelihood", "likelihood_wrapper", "prepare_likelihood_inputs",
           "evaluate_cluster_likelihood", "evaluate_fragment_clusters", "SelfXTalk"]
