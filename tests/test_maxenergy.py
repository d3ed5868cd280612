"""Time-delay max-energy maps (coherence_native time_delay_numba.py:183-288).

CPU pins of the oracle (oracle/maxenergy_oracle.c): the packet parameters of
every pattern agree with the host restatement; the packet energy equals its
closed form from the 2x2 Gram matrix of the packet's (00, 90) amplitudes,
(ee + EE + 2 sqrt(ee EE - ss^2)) / 4 (the squared sum of the singular values,
over 4); the max-energy map with no delay step equals the packet energy of the
unshifted map, and with delays equals the maximum over the maps of shifted
copies built independently in NumPy.
GPU: the batched HIP path against the oracle for every pattern (1e-10 of the
map's maximum: different FFTs), the device Gamma-to-Gauss against the host
one (order statistics and counts exact, the map and ALP to 1e-12: the device
log is not NumPy's), and the TF-map facade.
"""

import numpy as np
import pytest


def _series(L, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(L)
    x = rng.normal(size=L)
    x += 4 * np.sin(2 * np.pi * 0.11 * t) * np.exp(-0.5 * ((t - L / 2) / 150) ** 2)
    return x


@pytest.mark.parametrize("pattern", range(0, 10))
def test_params_host_matches_oracle(oracle, pattern):
    from pycwb_amd.coherence_native.projection import _compute_packet_energy_params

    for M1, T, edge, f_low, f_high, df in ((17, 300, 0.0, 0.0, 512.0, 32.0),
                                           (33, 200, 2.0, 40.0, 300.0, 16.0)):
        h = _compute_packet_energy_params(M1, T, pattern, edge, 64, f_low, f_high, df)
        o = oracle.max_energy_params(M1, T, pattern, edge, 64, f_low, f_high, df)
        assert h[:5] == o[:5]
        np.testing.assert_array_equal(h[5], o[5])


def _packet_numpy(re, im, M1, T, jb, je, mL, mH, mean, p):
    """Closed form from the Gram matrix of the packet (time-major maps)."""
    J = M1 * T
    out = np.zeros(J)
    for j in range(J):
        m = j % M1
        if j < jb or j >= je:
            out[j] = max(re[j], 0.0)
            continue
        if m < mL or m > mH:
            continue
        idx = np.clip(j + p[1:], 0, J - 1)
        w = np.concatenate([np.ones(8), [mean - 8.0]])
        pr = np.concatenate([re[idx], [re[j]]])
        pi = np.concatenate([im[idx], [im[j]]])
        ee, EE, ss = np.sum(w * pr * pr), np.sum(w * pi * pi), np.sum(w * pr * pi)
        if mean == 1.0:
            out[j] = (ee + EE) / 2
            continue
        det = max(ee * EE - ss * ss, 0.0)
        tr = ee + EE
        out[j] = (tr + 2 * np.sqrt(det)) / 4 if tr >= 0 else 0.0
    return out


@pytest.mark.parametrize("pattern", [1, 2, 3, 5, 7, 9])
def test_oracle_no_delay_is_packet_energy_closed_form(oracle, pattern):
    from pycwb_amd.wdm import wdm_filter

    M, L = 16, 2048
    h = wdm_filter(M, M, 6, 10)
    x = _series(L, pattern)
    out = oracle.max_energy(x, M, h, 0, 1, pattern, 0.0, 64, 0.0, 512.0, 32.0)
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    T = m00.shape[0]
    jb, je, mL, mH, mean, p = oracle.max_energy_params(M + 1, T, pattern, 0.0, 64, 0.0, 512.0,
                                                       32.0)
    ref = _packet_numpy(m00.ravel(), m90.ravel(), M + 1, T, jb, je, mL, mH, mean, p)
    ref = ref.reshape(T, M + 1)
    ref[:, 0] = 0.0
    if pattern in (5, 6, 9):
        ref[:, 1] = 0.0
    np.testing.assert_allclose(out.T, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_oracle_delays_are_maximum_over_shifted_copies(oracle):
    from pycwb_amd.wdm import wdm_filter

    M, L, pattern, max_delay, ds = 8, 1024, 3, 6, 2
    h = wdm_filter(M, M, 6, 10)
    x = _series(L, 5)
    out = oracle.max_energy(x, M, h, max_delay, ds, pattern, 0.0, 128, 0.0, 512.0, 64.0)
    best = None
    for k in [0] + [s * k for k in range(ds, max_delay + 1, ds) for s in (1, -1)]:
        y = x.copy()
        if k > 0:
            y[:L - k] = x[k:]
        elif k < 0:
            y[-k:] = x[:L + k]
        m00, m90 = oracle.wdm_t2w(y, M, h, -1)
        T = m00.shape[0]
        prm = oracle.max_energy_params(M + 1, T, pattern, 0.0, 128, 0.0, 512.0, 64.0)
        e = _packet_numpy(m00.ravel(), m90.ravel(), M + 1, T, *prm).reshape(T, M + 1)
        best = e if best is None else np.maximum(best, e)
    best[:, 0] = 0.0
    np.testing.assert_allclose(out.T, best, rtol=1e-9, atol=1e-9 * best.max())
    none = oracle.max_energy(x, M, h, 0, 1, pattern, 0.0, 128, 0.0, 512.0, 64.0)
    assert (out >= none - 1e-12).all() and (out > none + 1e-9).any()


def test_gamma_to_gauss_normalises():
    from pycwb_amd.coherence_native.projection import gamma_to_gauss

    rng = np.random.default_rng(0)
    E = rng.gamma(2.0, 1.0, size=(33, 400))
    out, alp = gamma_to_gauss(E.copy(), 0.0, 64)
    assert alp > 0 and out.shape == E.shape and (out >= 0).all()
    # monotone in the input energy (a pixel-wise transform, then one scale)
    o = np.argsort(E.ravel())
    assert np.all(np.diff(out.ravel()[o]) >= -1e-12)
    # the pivot quantile (31.73 % of the non-zero pixels lie above it) maps to 1
    flat = out.T.ravel()
    assert np.any(np.abs(flat - 1.0) < 1e-12)
    frac = np.sum(flat > 1.0) / np.sum(flat > 0)
    assert abs(frac - 0.3173) < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("pattern", range(1, 10))
def test_gpu_max_energy_matches_oracle(oracle, pattern):
    import torch

    from pycwb_amd.coherence_native.projection import max_energy_device
    from pycwb_amd.wdm import WDM

    print(f"case max-energy pattern {pattern}", flush=True)
    M, L, max_delay, ds = 16, 4096, 12, 3
    w = WDM(M, M, 6, 10)
    x = _series(L, pattern + 10)
    g = max_energy_device(torch.from_numpy(x).to("cuda:0"), w, max_delay, ds, pattern, 1.0, 256,
                          20.0, 100.0, 8.0, batch=4).cpu().numpy()
    o = oracle.max_energy(x, M, w.filter, max_delay, ds, pattern, 1.0, 256, 20.0, 100.0, 8.0)
    np.testing.assert_allclose(g, o.T, rtol=0, atol=1e-10 * np.abs(o).max())


@pytest.mark.gpu
def test_gpu_max_energy_facade():
    from types import SimpleNamespace

    import torch

    from pycwb_amd.coherence_native.projection import (gamma_to_gauss, max_energy,
                                                       max_energy_device)
    from pycwb_amd.wdm import WDM

    print("case max-energy facade", flush=True)
    M, rate = 32, 1024.0
    w = WDM(M, M, 6, 10)
    x = _series(16384, 3)
    tf = w.t2w(x, sample_rate=rate)
    tfm = SimpleNamespace(data=tf.data, dt=tf.dt, df=tf.df, edge=1.0, f_low=None, f_high=None,
                          wavelet=w, len_timeseries=x.size, ts_data=x)
    new, alp = max_energy(tfm, 0.01, 2, 3, f_low=32.0, f_high=400.0)
    cur = max_energy_device(torch.from_numpy(x).to("cuda:0"), w, int(rate * 0.01), 2, 3, 1.0,
                            int(1 / tf.dt), 32.0, 400.0, tf.df)
    ref, alp_ref = gamma_to_gauss(cur.t().contiguous().cpu().numpy(), 1.0, int(1 / tf.dt))
    assert alp_ref > 0
    np.testing.assert_allclose(alp, alp_ref, rtol=1e-12)
    np.testing.assert_allclose(new.data, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # without the stored series the map is reconstructed with w2t first
    tfm2 = SimpleNamespace(data=tf.data, dt=tf.dt, df=tf.df, edge=1.0, f_low=32.0,
                           f_high=400.0, wavelet=w, len_timeseries=x.size, ts_data=None)
    new2, alp2 = max_energy(tfm2, 0.01, 2, 3)
    e = M * int(w.m_H / M + 2) * 2
    np.testing.assert_allclose(new2.data[:, e:-e], new.data[:, e:-e], rtol=1e-3, atol=1e-3)
    with pytest.raises(NotImplementedError):
        max_energy(tfm, 0.01, 2, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,edge,zeros", [((33, 4000), 0.0, 0.0), ((65, 3000), 1.0, 0.4),
                                              ((9, 20000), 2.0, 0.0), ((17, 500), 0.5, 0.9)])
def test_gpu_gamma_to_gauss_matches_host(shape, edge, zeros):
    import torch

    from pycwb_amd.coherence_native.projection import gamma_to_gauss, gamma_to_gauss_device

    print(f"case gamma-to-gauss {shape} edge={edge}", flush=True)
    rng = np.random.default_rng(shape[1])
    E = rng.gamma(1.7, 1.3, size=shape)
    E[rng.random(shape) < zeros] = 0.0
    ref, alp_ref = gamma_to_gauss(E.copy(), edge, 16)
    out, alp = gamma_to_gauss_device(torch.from_numpy(np.ascontiguousarray(E.T)).to("cuda:0"),
                                     edge, 16)
    assert alp_ref > 0
    np.testing.assert_allclose(alp, alp_ref, rtol=1e-12)
    g = out.t().cpu().numpy()
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # a map with nothing above the floor: the reference gives up, so does the device
    z = torch.zeros((40, 9), dtype=torch.float64, device="cuda:0")
    same, a0 = gamma_to_gauss_device(z, 0.0, 16)
    assert a0 == 0.0 and same is z


@pytest.mark.gpu
def test_gpu_gamma_to_gauss_many_equals_single_calls():
    """Several maps with shared host reads give each map's single-call result
    bit for bit, also when one of them gives up (nothing above the floor)."""
    import torch

    from pycwb_amd.coherence_native.projection import (gamma_to_gauss_device,
                                                       gamma_to_gauss_device_many)

    print("case gamma-to-gauss many", flush=True)
    rng = np.random.default_rng(5)
    maps = [torch.from_numpy(np.ascontiguousarray(rng.gamma(1.7, 1.3, size=(m, t)).T))
            .to("cuda:0") for m, t in ((33, 4000), (65, 3000), (9, 20000))]
    maps.insert(1, torch.zeros((40, 9), dtype=torch.float64, device="cuda:0"))
    many = gamma_to_gauss_device_many(maps, 1.0, 16)
    for e, (g, a) in zip(maps, many):
        g1, a1 = gamma_to_gauss_device(e, 1.0, 16)
        assert a == a1
        assert torch.equal(g, g1)
    assert many[1][1] == 0.0 and many[1][0] is maps[1]

