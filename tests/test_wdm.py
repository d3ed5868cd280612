"""WDM filter design and forward transform.

CPU part pins the oracle and the host-side filter design: known answers from
the reference's coefficient table (tests/golden/wdm_fourier_coefficients.json,
from cwb-core/FourierCoefficients.icc), the filter length rule of
cwb-core/WDM.cc:149-171, and transform invariants (orthonormality of each
phase, band localisation of a tone).  The GPU part checks the HIP transform
against the oracle within 1e-6 relative to the map's largest coefficient.
"""

import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    with open(os.path.join(HERE, "golden", "wdm_fourier_coefficients.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("beta", [2, 3, 4, 5, 6, 7])
def test_fourier_coefficients_known_answers(oracle, beta):
    from pycwb_amd.wdm import COS_SIZE, edge_fourier

    fx = _fixture()
    ref = np.array(fx["cos"][str(beta)])
    assert COS_SIZE[beta] == fx["size"][str(beta)]
    np.testing.assert_allclose(oracle.wdm_fourier(beta)[:ref.size], ref, rtol=0, atol=5e-14)
    np.testing.assert_allclose(edge_fourier(beta)[:ref.size], ref, rtol=0, atol=5e-14)


@pytest.mark.parametrize("M,beta,prec", [(4, 6, 10), (8, 6, 10), (32, 4, 8), (64, 6, 10),
                                         (256, 2, 10)])
def test_filter_rule_and_agreement(oracle, M, beta, prec):
    from pycwb_amd.wdm import wdm_filter

    h = wdm_filter(M, M, beta, prec)
    ho = oracle.wdm_filter(M, M, beta, prec)
    assert h.size == ho.size
    np.testing.assert_allclose(h, ho, rtol=0, atol=1e-14)
    m_H = h.size
    assert (m_H - 1) % (2 * M) == 0 and m_H // (2 * M) >= 3
    norm = h[0] ** 2 + 2 * np.sum(h[1:] ** 2)
    assert 1 - norm <= 10.0 ** -prec
    # the previous admissible length (m_H - 2M) would not have passed the rule
    prev = m_H - 2 * M
    norm_prev = h[0] ** 2 + 2 * np.sum(h[1:prev] ** 2)
    assert 1 - norm_prev > 10.0 ** -prec or prev // (2 * M) < 3


@pytest.mark.parametrize("M", [8, 32])
@pytest.mark.parametrize("freq", [0.11, 0.25, 0.3])
def test_oracle_t2w_phase_orthonormality(oracle, M, freq):
    """Away from the half-band edge layers (0 and M, which cWB writes into both
    phases), each phase of the transform is an orthonormal basis: for a
    band-limited pulse in the middle of the series sum |c_00|^2 = sum |c_90|^2
    = sum x^2."""
    from pycwb_amd.wdm import wdm_filter

    h = wdm_filter(M, M, 6, 10)
    L = 256 * M + 8 * h.size
    t = np.arange(L)
    x = np.exp(-0.5 * ((t - L // 2) / (6 * M)) ** 2) * np.cos(2 * np.pi * freq * t)
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    e = np.sum(x ** 2)
    assert np.sum(m00 ** 2) == pytest.approx(e, rel=1e-5)
    assert np.sum(m90 ** 2) == pytest.approx(e, rel=1e-5)


def test_oracle_t2w_tone_localised(oracle):
    """A tone at the centre frequency of layer m puts most energy in layer m."""
    from pycwb_amd.wdm import wdm_filter

    M, m, fs = 16, 5, 1.0
    h = wdm_filter(M, M, 6, 10)
    df = fs / (2 * M)
    t = np.arange(4096)
    x = np.cos(2 * np.pi * m * df * t + 0.3)
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    core = slice(h.size // M + 2, m00.shape[0] - h.size // M - 2)
    E = (m00[core] ** 2 + m90[core] ** 2).sum(axis=0)
    assert E[m] / E.sum() > 0.9


def test_oracle_power_map_matches_quadrature(oracle):
    """MM = 0 power map = |X|^2 of the same folds; here via the quadrature
    maps: 00^2 + 90^2 = 2 |X|^2 away from the edge layers."""
    from pycwb_amd.wdm import wdm_filter

    M = 8
    h = wdm_filter(M, M, 6, 10)
    x = np.random.default_rng(1).normal(size=1000)
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    pw, _ = oracle.wdm_t2w(x, M, h, 0)
    np.testing.assert_allclose((m00 ** 2 + m90 ** 2)[:, 1:M], 2 * pw[:, 1:M], rtol=1e-9,
                               atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("M,L,S", [(2, 3001, 2), (4, 4003, 2), (8, 8192, 3), (16, 5000, 1),
                                   (32, 30000, 2), (64, 20000, 2), (128, 70001, 2),
                                   (256, 65536, 3), (512, 150000, 2), (1024, 200000, 1),
                                   # long enough for blocks that stream several tiles
                                   # through one window ring
                                   (16, 2000003, 1), (256, 1100000, 1), (512, 1100000, 1),
                                   # barely longer than the filter: one partial tile
                                   (8, 100, 3), (128, 1601, 2), (256, 3200, 1),
                                   (512, 6200, 1)])
def test_gpu_t2w_matches_oracle(oracle, M, L, S):
    import torch

    from pycwb_amd.wdm import WDM

    print(f"case t2w M={M} L={L} S={S}", flush=True)
    w = WDM(M, M, 6, 10)
    rng = np.random.default_rng(M + L)
    x = rng.normal(size=(S, L))
    g = w.t2w_device(torch.from_numpy(x).to("cuda:0"), MM=-1).cpu().numpy()
    for s in range(S):
        m00, m90 = oracle.wdm_t2w(x[s], M, w.filter, -1)
        scale = max(np.abs(m00).max(), np.abs(m90).max())
        np.testing.assert_allclose(g[s].real, m00.T, rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(g[s].imag, m90.T, rtol=0, atol=1e-6 * scale)
        err = max(np.abs(g[s].real - m00.T).max(), np.abs(g[s].imag - m90.T).max()) / scale
        assert err < 1e-12  # float64 throughout; observed error is rounding-level
    pw = w.t2w_device(torch.from_numpy(x).to("cuda:0"), MM=0).cpu().numpy()
    p_o, _ = oracle.wdm_t2w(x[0], M, w.filter, 0)
    np.testing.assert_allclose(pw[0].real, p_o.T, rtol=1e-9, atol=1e-12 * p_o.max())


@pytest.mark.gpu
@pytest.mark.parametrize("M,L", [(8, 20003), (64, 50000), (256, 70001), (512, 150000)])
def test_gpu_t2w_shifted_equals_transforms_of_shifted_copies(M, L):
    """pycwb_wdm_t2w_shifted (the copies read straight from the series) gives
    bit for bit the maps of pycwb_wdm_t2w on the copies pycwb_max_energy_shift
    writes out."""
    import torch

    from pycwb_amd import _lib
    from pycwb_amd.device import lib, ptr
    from pycwb_amd.wdm import WDM

    print(f"case t2w shifted M={M} L={L}", flush=True)
    dev = torch.device("cuda", 0)
    w = WDM(M, M, 6, 10)
    ts = torch.from_numpy(np.random.default_rng(M).normal(size=L)).to(dev)
    shifts = torch.tensor([0, 3, -3, 17, -17, 4 * M + 1, -(4 * M + 1), L - 5, -(L - 5)],
                          dtype=torch.int64, device=dev)
    got = w.t2w_shifted_device(ts, shifts)
    xs = torch.empty((shifts.numel(), L), dtype=torch.float64, device=dev)
    _lib.check(lib().pycwb_max_energy_shift(ptr(ts), L, ptr(shifts), shifts.numel(), ptr(xs),
                                            None), "pycwb_max_energy_shift")
    want = w.t2w_device(xs, MM=-1)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(got), torch.view_as_real(want))
    # a filter of another length is declined (the caller writes the copies)
    assert WDM(M, M, 4, 8).t2w_shifted_device(ts, shifts) is None


@pytest.mark.gpu
def test_gpu_wdm_object_surface():
    from pycwb_amd.wdm import WDM

    w = WDM(32, 32, 6, 10)
    x = np.random.default_rng(3).normal(size=4096)
    tf = w.t2w(x, sample_rate=1024.0, t0=10.0, MM=-1)
    assert tf.data.shape == (33, 128) and tf.data.dtype == np.complex128
    assert tf.dt == 32 / 1024.0 and tf.df == 1024.0 / 64
    assert tf.start_time == 10.0 and tf.len_timeseries == 4096


def _series(L, seed=1):
    rng = np.random.default_rng(seed)
    t = np.arange(L)
    return (np.sin(2 * np.pi * 0.07 * t) * np.exp(-0.5 * ((t - L / 2) / (L / 12)) ** 2)
            + np.cos(2 * np.pi * 0.31 * t + 1.0) * np.exp(-0.5 * ((t - L / 3) / (L / 20)) ** 2)
            + 0.3 * rng.normal(size=L))


@pytest.mark.parametrize("M", [4, 8, 16, 64])
def test_oracle_inverse_reconstructs(oracle, M):
    """w2t and w2tQ (WDM.cc:1536-1710) each reconstruct the series from t2w's
    00 and 90 coefficients: both phases of the WDM basis are tight frames.
    Away from the mirror-padded edges the error is the filter's truncation
    (precision 10: residual norm 1e-10, error ~1e-5 of the signal)."""
    from pycwb_amd.wdm import wdm_filter

    h = wdm_filter(M, M, 6, 10)
    L = 64 * M + 3 * h.size
    x = _series(L, seed=M)
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    mp = (m00 + 1j * m90).T
    e = h.size
    scale = np.abs(x).max()
    for quad in (False, True):
        y = oracle.wdm_w2t(mp, M, h, L, quad)
        assert np.abs(y - x)[e:-e].max() < 3e-5 * scale, quad
    # a 90-phase map read as 00 does not reconstruct (the phase rules matter)
    y = oracle.wdm_w2t((m90 + 1j * m00).T, M, h, L, False)
    assert np.abs(y - x)[e:-e].max() > 0.1 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("M,S,L", [(2, 1, None), (8, 3, None), (16, 2, None), (32, 2, None),
                                   (64, 2, None), (128, 2, None), (256, 1, None),
                                   (512, 2, None), (1024, 1, None),
                                   # long enough for blocks that stream several tiles
                                   (8, 1, 2100003), (512, 1, 1100000)])
def test_gpu_inverse_matches_oracle(oracle, M, S, L):
    import torch

    from pycwb_amd.wdm import WDM

    print(f"case w2t M={M} S={S} L={L}", flush=True)
    w = WDM(M, M, 6, 10)
    L = L or 24 * M + 2 * w.m_H + 7
    xs = np.stack([_series(L, seed=k + M) for k in range(S)])
    dev = torch.device("cuda", 0)
    tf = w.t2w_device(torch.from_numpy(xs).to(dev))
    for quad in (False, True):
        y = w.w2t_device(tf, L, quad).cpu().numpy()
        for k in range(S):
            ref = oracle.wdm_w2t(tf[k].cpu().numpy(), M, w.filter, L, quad)
            scale = np.abs(ref).max()
            np.testing.assert_allclose(y[k], ref, rtol=0, atol=1e-12 * scale)
            e = w.m_H
            assert np.abs(y[k] - xs[k])[e:-e].max() < 3e-5 * np.abs(xs[k]).max()
    with pytest.raises(ValueError):
        w.w2t_device(tf, 2 * L)


@pytest.mark.gpu
def test_gpu_inverse_surface():
    from pycwb_amd.wdm import WDM

    print("case w2t surface", flush=True)
    w = WDM(32, 32, 6, 10)
    x = _series(8192, seed=5)
    tf = w.t2w(x, sample_rate=4096.0)
    a, b = w.w2t(tf), w.w2tQ(tf)
    assert a.shape == b.shape == x.shape
    e = w.m_H
    avg = 0.5 * (a + b)        # the conditioning code's combination (whitening.py:110-113)
    assert np.abs(avg - x)[e:-e].max() < 3e-5 * np.abs(x).max()
