"""Best-sky statistics on the GPU against the oracle (Lm/Ec/Np/rho etc.).

Bar: float32 per-pixel arrays produced by the Numba-typed stages (pixel mask,
coherent energy, packet energies, DPF vectors, delayed data) bit-exact; the
float64 NumPy-stage scalars within 1e-9 relative (the path's stated tolerance
for Lm/Ec/rho is 1e-5 relative).
"""

import numpy as np
import pytest
import torch

from pycwb_amd import synth

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _random_xtalk(rng, P, k=5):
    lookup = np.zeros((P, 2), np.int32)
    rows = []
    for j in range(P):
        nb = np.unique(np.concatenate([[j], rng.integers(0, P, size=rng.integers(0, k))]))
        lookup[j] = (len(rows), len(rows) + len(nb))
        for b in nb:
            c = rng.normal(scale=0.3, size=4).astype(np.float32)
            if b == j:
                c = np.array([1, 0, 0, 1], np.float32)
            rows.append([b, 0, 0, 0, *c])
    return lookup, np.asarray(rows, np.float32)


@pytest.mark.parametrize("use_catalog", [False, True])
def test_sky_statistics_match_oracle(oracle, use_catalog):
    from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch
    from pycwb_amd.skystats import SkyStatistics

    print(f"case skystats catalog={use_catalog}", flush=True)
    nets = [synth.make_network(synth.NETWORK_3, order=3, K=8),
            synth.make_network(synth.NETWORK_4, order=3, K=8)]
    sizes = np.array([40, 7, 130, 1, 66, 300], dtype=np.int64)
    ids = np.array([0, 1, 0, 1, 1, 0])
    dev = torch.device("cuda", 0)
    b = synth_batch(nets, sizes, ids, seed=91, snr=8.0, device=dev)
    nt = net_tables(nets, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    gamma, delta = 0.5 * 0.5 * 2 / 3, 0.5
    bs = BatchedSkyScan(b, nt, valid, gamma, delta, 0.0, block=64, grid=11)
    bs.launch()
    rng = np.random.default_rng(5)
    xtalk = [_random_xtalk(rng, int(P)) for P in sizes] if use_catalog else None
    st = SkyStatistics(b, nt, xtalk)
    reg0 = float(np.float32(delta * np.sqrt(2)))
    st.run(bs.l_max, bs.REG1, reg0)
    sc = st.scalars()
    lm = bs.l_max.cpu().numpy()
    reg1 = bs.REG1.cpu().numpy()
    for c in range(sizes.size):
        net = nets[ids[c]]
        td = b.cluster_td(c).cpu().numpy()
        td00 = np.ascontiguousarray(td[..., 0].transpose(2, 1, 0))
        td90 = np.ascontiguousarray(td[..., 1].transpose(2, 1, 0))
        rms = b.cluster_rms(c).cpu().numpy()
        REG = np.array([reg0, reg1[c], 0], np.float32)
        lk, xt = xtalk[c] if xtalk else (None, None)
        o = oracle.sky_statistics(int(lm[c]), net.n_ifo, int(sizes[c]), net.n_sky, net.FP_t,
                                  net.FX_t, rms, td00, td90, net.ml, REG, nt.En[ids[c]], lk, xt)
        g = st.cluster_arrays(c)
        for k_o, k_g in (("pixel_mask", "pixel_mask"), ("ec", "coherent_energy"),
                         ("ee", "energy_array_plus"), ("EE", "energy_array_cross")):
            assert np.array_equal(o[k_o].view(np.uint32), g[k_g].view(np.uint32)), (c, k_o)
        scale = max(abs(o["Em"]), 1.0)
        for k in ("Gn", "Ec", "Dc", "Rc", "Eh", "Es", "Np", "Em", "Lm", "norm", "cc", "rho",
                  "xrho", "Lo", "Eo", "N_pix_effective"):
            assert abs(sc[k][c] - o[k]) <= RTOL * max(abs(o[k]), 1e-6 * scale), (c, k, sc[k][c],
                                                                                  o[k])
        for k in ("pd", "pD", "ps", "pS"):
            np.testing.assert_allclose(g[k], o[k], rtol=1e-9, atol=1e-12 * scale)
        np.testing.assert_allclose(g["noise_amplitude_00"], o["nul00"], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(g["gaussian_noise_correction"], o["gn"], rtol=1e-6,
                                   atol=1e-7)
        np.testing.assert_allclose(g["S_snr"], o["S_snr"], rtol=1e-9)
        # avx_pol_ps: radii bit-exact (float64 NumPy order, correctly rounded
        # sqrt), angles within one float32 ulp (the device and host atan2)
        for k in ("p00_POL_r", "p90_POL_r"):
            assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), (c, k)
        for k in ("p00_POL_a", "p90_POL_a"):
            np.testing.assert_array_max_ulp(g[k], o[k], maxulp=1)
