"""Time-delay amplitude extraction (the producer of the likelihood's td tables).

CPU: the oracle's Cos2/SinCos coefficients against the reference's table
(cwb-core/FourierCoefficients.icc, tests/golden/wdm_fourier_coefficients.json),
the host TD filters against the oracle's, and the defining property of TD
amplitudes: the amplitude of pixel (n, m) at delay d equals the WDM
coefficient (n, m) of the series delayed by d samples, x[t - d] (up to the
filter truncation at n_coeffs), for inner and edge layers (in the edge layers
the phase a TD amplitude keeps per time bin; the other is exactly zero) and
for delays past one time bin (whole-bin shift with quadrature swap).

GPU: the HIP kernel against the oracle (float32 outputs of float64 sums,
within 2e-6 of the largest amplitude) in both output layouts, and the same
delay property on the GPU path end to end (t2w on the GPU, TD filters, TD
extraction).
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
def test_td_fourier_known_answers(oracle, beta):
    from pycwb_amd.wdm import edge_fourier_cos2, edge_fourier_sincos

    fx = _fixture()
    for key, which, host in (("cos2", 0, edge_fourier_cos2), ("sincos", 1, edge_fourier_sincos)):
        ref = np.array(fx[key][str(beta)])
        size = fx[key + "_size"][str(beta)]
        o = oracle.td_fourier(beta, which)
        h = host(beta)
        assert o.size == size and h.size == size
        np.testing.assert_allclose(o[:ref.size], ref, rtol=0, atol=1e-12)
        np.testing.assert_allclose(h[:ref.size], ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("M,nc,L", [(8, 6, 1), (16, 12, 1), (32, 8, 2), (64, 12, 1)])
def test_td_filters_host_matches_oracle(oracle, M, nc, L):
    from pycwb_amd.wdm import td_filters

    b = td_filters(M, M, 6, nc, L)
    T0, Tx = oracle.td_filters(M, M, 6, nc, L)
    assert b.T0.shape == (2 * M * L + 1, 2 * nc + 1) and b.max_delay == M * L
    np.testing.assert_allclose(b.T0, T0, rtol=0, atol=1e-11)
    np.testing.assert_allclose(b.Tx, Tx, rtol=0, atol=1e-11)
    # zero delay: the same-band filter's even taps are the identity
    J = M * L
    even = b.T0[J, nc % 2::2]
    assert abs(b.T0[J, nc] - 1.0) < 1e-12
    assert np.abs(np.delete(even, nc // 2)).max() < 1e-6


def _signal(L=4096):
    t = np.arange(L)
    return sum(np.cos(2 * np.pi * f * t + p) * np.exp(-0.5 * ((t - L / 2) / (L / 10)) ** 2)
               for f, p in [(0.11, 0.3), (0.23, 1.0), (0.31, 2.0), (0.02, 0.5)])


def _surviving_phases(n, m, M):
    """Edge layers (m = 0, M): t2w writes the layer into both phases
    (cwb-core/WDM.cc:1494-1505) but a TD amplitude keeps only one per time
    bin (td_vector_kernels.py:126-128): 00 at even n, 90 at odd n."""
    edge = (m == 0) | (m == M)
    odd = (n & 1).astype(bool)
    return ~(edge & odd), ~(edge & ~odd)


def _planes(m00, m90, nc):
    pad = ((nc, nc), (0, 0))
    return np.pad(m00, pad).astype(np.float32), np.pad(m90, pad).astype(np.float32)


@pytest.mark.parametrize("M", [8, 16])
def test_oracle_td_amplitude_is_delayed_coefficient(oracle, M):
    from pycwb_amd.wdm import wdm_filter

    nc, K = 12, 3 * M
    h = wdm_filter(M, M, 6, 10)
    x = _signal()
    m00, m90 = oracle.wdm_t2w(x, M, h, -1)
    T0, Tx = oracle.td_filters(M, M, 6, nc, 1)
    p00, p90 = _planes(m00, m90, nc)
    n0 = m00.shape[0] // 2 - 20
    pix = np.array([n * (M + 1) + m for n in range(n0, n0 + 40) for m in range(M + 1)],
                   dtype=np.int32)
    td = oracle.td_vectors(pix, p00, p90, T0, Tx, M, nc, K, M)
    n, m = pix // (M + 1), pix % (M + 1)
    keep00, keep90 = _surviving_phases(n, m, M)
    scale = np.abs(m00).max()
    for d in (-K, -M - 3, -M, -5, 0, 1, 7, M, M + 1, K):
        s00, s90 = oracle.wdm_t2w(np.roll(x, d), M, h, -1)   # x[t - d]
        e00 = np.abs(td[keep00, K + d] - s00[n, m][keep00]).max() / scale
        e90 = np.abs(td[keep90, 3 * K + 1 + d] - s90[n, m][keep90]).max() / scale
        assert e00 < 2e-4 and e90 < 2e-4, (d, e00, e90)
        assert not td[~keep00, K + d].any() and not td[~keep90, 3 * K + 1 + d].any()
    # the opposite direction is far off (the property pins the sign of the delay)
    s00, _ = oracle.wdm_t2w(np.roll(x, -5), M, h, -1)
    assert np.abs(td[:, K + 5] - s00[n, m]).max() / scale > 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("M,nc,L,K", [(8, 6, 1, 20), (16, 12, 1, 16), (32, 8, 2, 40),
                                      (64, 12, 1, 64),
                                      # whole-bin shifts -4 .. 4 (J = 16, K = 64)
                                      (16, 12, 1, 64),
                                      # 401 delays: three chunks of the pixel-major kernel
                                      (16, 12, 1, 200)])
def test_gpu_td_vectors_match_oracle(oracle, M, nc, L, K):
    import torch

    from pycwb_amd.tdamp import TDBatchInputs
    from pycwb_amd.wdm import td_filters

    print(f"case td M={M} nc={nc} L={L} K={K}", flush=True)
    rng = np.random.default_rng(M + K)
    n_time = 300
    data = rng.normal(size=(M + 1, n_time)) + 1j * rng.normal(size=(M + 1, n_time))
    bank = td_filters(M, M, 6, nc, L)
    inp = TDBatchInputs.from_map(data, bank, torch.device("cuda", 0))
    pix = rng.choice(n_time * (M + 1), size=500, replace=False).astype(np.int32)
    pix[:3] = [0, M, (n_time - 1) * (M + 1) + M]   # first and last bins, edge layers
    g = inp.extract_td_vecs(pix, K).cpu().numpy()
    p00 = inp.padded00.cpu().numpy()
    p90 = inp.padded90.cpu().numpy()
    o = oracle.td_vectors(pix, p00, p90, bank.T0, bank.Tx, M, nc, K, M * L)
    np.testing.assert_allclose(g, o, rtol=0, atol=2e-6 * np.abs(o).max())
    # packed layout for the likelihood batch: (P, n_ifo, 2K+1, 2), detector 1 of 3
    nd = 2 * K + 1
    td = torch.zeros(2 * (5 + pix.size * 3 * nd), dtype=torch.float32, device="cuda:0")
    inp.extract_td_into(pix, K, td, td_off=5, ifo=1, n_ifo=3)
    packed = td[10:].view(pix.size, 3, nd, 2).cpu().numpy()
    np.testing.assert_array_equal(packed[:, 1, :, 0], g[:, :nd])
    np.testing.assert_array_equal(packed[:, 1, :, 1], g[:, nd:])
    assert not packed[:, 0].any() and not packed[:, 2].any()
    with pytest.raises(ValueError):
        inp.extract_td_vecs(np.array([n_time * (M + 1)]), K)


@pytest.mark.gpu
def test_gpu_td_chain_delay_property():
    """t2w on the GPU, TD filters, TD extraction: amplitudes at delay d equal
    the GPU transform of the delayed series."""
    from pycwb_amd.wdm import WDM

    print("case td chain", flush=True)
    M, nc, K = 32, 12, 48
    w = WDM(M, M, 6, 10)
    bank = w.set_td_filter(nc, 1)
    x = _signal(8192)
    tf = w.t2w(x)
    inp = tf.prepare_td_inputs(bank)
    n0 = tf.data.shape[1] // 2 - 10
    pix = np.array([n * (M + 1) + m for n in range(n0, n0 + 20) for m in range(M + 1)],
                   dtype=np.int32)
    td = inp.extract_td_vecs(pix, K).cpu().numpy()
    n, m = pix // (M + 1), pix % (M + 1)
    keep00, keep90 = _surviving_phases(n, m, M)
    scale = np.abs(tf.data).max()
    for d in (-K, -33, -1, 0, 5, 32, 47):
        s = w.t2w(np.roll(x, d)).data
        err = max(np.abs(td[keep00, K + d] - s.real[m, n][keep00]).max(),
                  np.abs(td[keep90, 3 * K + 1 + d] - s.imag[m, n][keep90]).max()) / scale
        assert err < 2e-4, (d, err)


@pytest.mark.gpu
def test_gpu_td_inputs_from_device_map_match_host_map():
    """TDBatchInputs.from_device (planes padded on the device from a t2w_device
    output) gives the same planes and amplitudes as from_map on the host copy."""
    import torch

    from pycwb_amd.tdamp import TDBatchInputs
    from pycwb_amd.wdm import WDM

    print("case td inputs from device map", flush=True)
    M, nc, K = 16, 12, 40
    w = WDM(M, M, 6, 10)
    bank = w.set_td_filter(nc, 2)
    x = torch.from_numpy(_signal(4096)).to("cuda:0")
    tf_dev = w.t2w_device(x)[0]                       # (M+1, N) complex128
    a = TDBatchInputs.from_device(tf_dev, bank)
    b = TDBatchInputs.from_map(tf_dev.cpu().numpy(), bank, torch.device("cuda", 0))
    assert torch.equal(a.padded00, b.padded00) and torch.equal(a.padded90, b.padded90)
    pix = np.arange(0, tf_dev.shape[1] * (M + 1), 7, dtype=np.int32)
    assert torch.equal(a.extract_td_vecs(pix, K), b.extract_td_vecs(pix, K))
