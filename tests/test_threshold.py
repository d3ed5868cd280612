"""Pixel-energy threshold (coherence_native/veto_threshold.py:98-222).

CPU pins of the oracle (oracle/threshold_oracle.c): in both of the
reference's orders (edge-cropped frequency rows; the time-major flat region
w[nL:nR]) the clipped detector sum equals NumPy's (maps added in detector
order, np.clip) bit for bit, the counts equal np.count_nonzero, the sorted
values equal np.sort, the sum of the values above 1e-3 equals np.sum of them
bit for bit, and the sum of their logs equals np.sum(np.log(.)) to 1e-13
(NumPy's log and the C library's differ in the last bit for a few values).
GPU: the same quantities against the oracle (values, counts, order and sums
bit-exact; log sums to 1e-13), and compute_threshold against a NumPy
evaluation of the reference's formula on the same maps: exactly for the
no-shape path, to 1e-12 for the shape (alp) path, whose mean log goes
through the device's log.
"""

from types import SimpleNamespace

import numpy as np
import pytest


def _maps(rng, n_ifo, n_freq, n_time, zero_frac=0.3):
    E = rng.gamma(1.5, 1.0, size=(n_ifo, n_freq, n_time))
    E[rng.random(E.shape) < zero_frac] = 0.0
    E[0, 1, :5] = 500.0             # above the clip
    return E


def _numpy_values(E, time_major, begin, count, hi):
    if time_major:
        w = E[0].T.ravel().copy()
        for d in range(1, E.shape[0]):
            w += E[d].T.ravel()
        region = w[begin:begin + count]
    else:
        ncol = count // E.shape[1]
        combined = np.sum([e[:, begin:begin + ncol] for e in E], axis=0)
        region = combined.ravel()
    return np.clip(region, 0.0, hi)


@pytest.mark.parametrize("time_major,begin,cut", [(1, 37, 41), (0, 6, 12), (0, 0, 0),
                                                  (1, 0, 1)])
def test_oracle_matches_numpy(oracle, time_major, begin, cut):
    rng = np.random.default_rng(begin + cut)
    E = _maps(rng, 3, 33, 3000)
    n_freq, n_time = E.shape[1:]
    count = n_freq * n_time - begin - cut if time_major else n_freq * (n_time - begin - cut)
    o = oracle.threshold_stats(E, time_major, begin, count, 300.0)
    ref = _numpy_values(E, time_major, begin, count, 300.0)
    np.testing.assert_array_equal(o["values"], ref)
    assert list(o["counts"]) == [np.count_nonzero(ref > 1e-4), np.count_nonzero(ref > 1e-3)]
    np.testing.assert_array_equal(o["sorted"], np.sort(ref))
    pos = ref[ref > 1e-3]
    assert o["n_pos"] == pos.size > 8192
    assert o["sum"] == np.sum(pos)
    np.testing.assert_allclose(o["log_sum"], np.sum(np.log(pos)), rtol=1e-13)


@pytest.mark.parametrize("time_major", [0, 1])
def test_oracle_signed_zeros_follow_numpy(oracle, time_major):
    # np.sum(axis=0) starts from +0.0, the time-major in-place sum from the
    # first map, and np.clip keeps -0.0
    E = np.zeros((2, 3, 4))
    E[:, 1, 2] = -0.0
    E[0, 0, 1] = -0.0
    E[0, 2, 3] = -5.0
    o = oracle.threshold_stats(E, time_major, 0, 12, 200.0)
    ref = _numpy_values(E, time_major, 0, 12, 200.0)
    np.testing.assert_array_equal(np.signbit(o["values"]), np.signbit(ref))
    np.testing.assert_array_equal(o["values"], ref)
    assert np.signbit(ref).any() == bool(time_major)


def _reference_formula(tf_maps, bpp, alp=None, edge=None):
    """NumPy evaluation of compute_threshold on host arrays (the test's own
    statement of the reference's formula)."""
    from scipy.special import gammaincc, gammainccinv

    def igi(shape, p):
        return float(gammainccinv(float(max(shape, 1e-12)),
                                  float(np.clip(p, 1e-12, 1 - 1e-12))))

    n_ifo = len(tf_maps)
    if alp is not None:
        arr0 = np.asarray(tf_maps[0].data, dtype=np.float64)
        M = arr0.shape[0] if arr0.ndim == 2 else 1
        w = (arr0.T.ravel() if arr0.ndim == 2 else arr0.ravel()).copy()
        for t in tf_maps[1:]:
            a = np.asarray(t.data, dtype=np.float64)
            w += a.T.ravel() if a.ndim == 2 else a.ravel()
        nL = int(float(edge or 0.0) * tf_maps[0].wavelet_rate * M)
        nR = w.size - nL - 1
        region = np.clip(w[nL:nR], 0.0, n_ifo * 100.0)
        pos = region[region > 1e-3]
        if pos.size == 0:
            return 0.0
        avr = float(np.mean(pos))
        bbb = float(np.mean(np.log(pos)))
        a = np.log(avr) - bbb
        a = (3 - a + np.sqrt((a - 3) * (a - 3) + 24 * a)) / (12 * a)
        return avr * igi(a, float(bpp) * a / float(alp)) / a / 2.0
    arrs = []
    for t in tf_maps:
        arr = np.asarray(t.data, dtype=np.float64)
        if arr.ndim == 2 and edge is not None and t.dt > 0:
            e = int(max(0, round(float(edge) / float(t.dt))))
            if e > 0 and arr.shape[1] > 2 * e:
                arr = arr[:, e:-e]
        arrs.append(arr)
    work = np.clip(np.sum(arrs, axis=0).ravel(), 0.0, n_ifo * 100.0)
    if work.size == 0 or not np.any(work > 1e-3):
        return 0.0
    fff = float(np.sum(work > 1e-4) / work.size)
    n = work.size
    s = np.sort(work)
    kv, km = int(float(bpp) * fff * n), int(0.2 * fff * n)
    val = float(s[max(0, n - kv - 1)]) if kv > 0 else float(s[-1])
    med = float(s[max(0, n - km - 1)]) if km > 0 else float(s[-1])
    m, p00 = 1.0, 0.0
    while p00 < 0.2:
        p00 = float(gammaincc(n_ifo * m, med))
        m += 0.01
    if m > 1.01:
        m -= 0.01
    return 0.3 * (igi(n_ifo * m, float(bpp)) + val) + n_ifo * np.log(m)


def _tf(data, dt=1 / 64.0):
    return SimpleNamespace(data=data, dt=dt, wavelet_rate=int(round(1 / dt)))


@pytest.mark.gpu
@pytest.mark.parametrize("time_major,begin,cut", [(1, 37, 41), (0, 6, 12), (0, 0, 0),
                                                  (1, 0, 1)])
def test_gpu_stats_match_oracle(oracle, time_major, begin, cut):
    import torch

    from pycwb_amd.coherence_native.veto_threshold import threshold_stats_device

    print(f"case threshold stats tm={time_major} begin={begin}", flush=True)
    rng = np.random.default_rng(begin + cut)
    E = _maps(rng, 3, 33, 3000)
    n_freq, n_time = E.shape[1:]
    count = n_freq * n_time - begin - cut if time_major else n_freq * (n_time - begin - cut)
    o = oracle.threshold_stats(E, time_major, begin, count, 300.0)
    g = threshold_stats_device(torch.from_numpy(E).to("cuda:0"), bool(time_major), begin, count,
                               300.0, True, True)
    np.testing.assert_array_equal(g["values"].cpu().numpy(), o["values"])
    assert list(g["counts"]) == list(o["counts"])
    np.testing.assert_array_equal(g["sorted"].cpu().numpy(), o["sorted"])
    assert g["sum"] == o["sum"]
    np.testing.assert_allclose(g["log_sum"], o["log_sum"], rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n_ifo,edge,alp", [(0, 2, 0.5, None), (1, 3, 1.0, 2.5),
                                                 (2, 5, None, None), (3, 3, 0.0, 1.7),
                                                 (4, 1, 0.25, 3.0)])
def test_gpu_compute_threshold(seed, n_ifo, edge, alp):
    from pycwb_amd.coherence_native.veto_threshold import compute_threshold

    print(f"case threshold seed={seed} n_ifo={n_ifo} alp={alp}", flush=True)
    rng = np.random.default_rng(seed)
    E = _maps(rng, n_ifo, 65, 4000, zero_frac=0.1)
    maps = [_tf(E[d]) for d in range(n_ifo)]
    for bpp in (0.001, 0.01):
        g = compute_threshold(maps, bpp, alp=alp, edge=edge)
        r = _reference_formula(maps, bpp, alp=alp, edge=edge)
        if alp is None:
            assert g == r
        else:
            np.testing.assert_allclose(g, r, rtol=1e-12)
        assert g > 0


@pytest.mark.gpu
def test_gpu_compute_threshold_edge_cases():
    from pycwb_amd.coherence_native.veto_threshold import compute_threshold

    print("case threshold edge cases", flush=True)
    z = [_tf(np.zeros((9, 100))), _tf(np.zeros((9, 100)))]
    assert compute_threshold(z, 0.001) == 0.0 and compute_threshold(z, 0.001, alp=2.0) == 0.0
    rng = np.random.default_rng(7)
    one_d = [_tf(rng.gamma(2.0, 1.0, 5000)) for _ in range(2)]
    assert compute_threshold(one_d, 0.01, edge=1.0) == _reference_formula(one_d, 0.01, edge=1.0)
    np.testing.assert_allclose(compute_threshold(one_d, 0.01, alp=2.0, edge=1.0),
                               _reference_formula(one_d, 0.01, alp=2.0, edge=1.0), rtol=1e-12)
    with pytest.raises(ValueError):
        compute_threshold([], 0.01)
