"""Network-pixel selection (coherence_native/selection.py:125-243).

CPU: the oracle (oracle/select_oracle.c) on the known answers of the
reference's tests (coherence_native/tests/test_circular_cat2_semantics.py:
207-263: the shifted-veto live mask, the per-detector source bins [4, 6]) and
against a dense float64 numpy formulation from the definitions (whole-map
sums of circularly rolled energy maps, masks, neighbourhood sums by array
shifts) on the reference test's edge cases (zero lag, wrapped shift, shifted
veto, edge crop, frequency band, empty) and random maps.
GPU: every payload field bit-exact against the oracle, for many lags per call.
"""

from types import SimpleNamespace

import numpy as np
import pytest


def _tf(data, dt=1.0, df=1.0, f_low=0.0, f_high=None):
    data = np.asarray(data, dtype=float)
    return SimpleNamespace(data=data, dt=dt, df=df, start=0.0, stop=float(data.shape[1]) * dt,
                           f_low=f_low,
                           f_high=float(data.shape[0] - 1) * df if f_high is None else f_high)


def _band(n_freq, df, f_low, f_high):
    ib, ie = 1, n_freq
    for k in range(n_freq):
        if k * df <= f_high:
            ie = k
        if k * df <= f_low:
            ib = k + 1
    return max(ib, 1), min(ie, n_freq - 1)


def _dense(E, shift, veto, edge, ib, ie, eo):
    """Float64 numpy selection from the definitions."""
    n_ifo, n_freq, n_time = E.shape
    vs, ve = edge, n_time - edge
    nn = ve - vs
    win = E[:, :, vs:ve]
    rolled = np.stack([np.roll(win[d], -int(shift[d]), axis=1) for d in range(n_ifo)])
    raw = np.zeros((n_freq, n_time))
    raw[:, vs:ve] = rolled.sum(0)
    live = np.zeros(n_time, bool)
    live[vs:ve] = True
    if veto is not None:
        vw = np.asarray(veto)[vs:ve]
        ok = np.all([np.roll(vw, -int(shift[d])) != 0 for d in range(n_ifo)], axis=0)
        live[vs:ve] = ok
    c = np.where(raw < eo, 0.0, np.where(raw > 2 * eo, 2 * eo + 0.1, raw))
    c[:, ~live] = 0.0
    c[:ib] = 0.0
    em, eh = 2 * eo, 4 * eo * eo
    z = np.zeros((n_freq + 4, n_time + 4))
    z[2:-2, 2:-2] = c
    S = lambda df_, dt_: z[2 + df_:2 + df_ + n_freq, 2 + dt_:2 + dt_ + n_time]
    ct = S(1, 0) + S(0, 1) + S(1, 1)
    cb = S(-1, 0) + S(0, -1) + S(-1, -1)
    fidx = np.arange(n_freq)[:, None]
    ht = S(1, 2) + np.where(fidx < n_freq - 2, S(2, 2) + S(2, 1), 0.0)
    hb = S(-1, -2) + np.where(fidx >= 2, S(-2, -2) + S(-2, -1), 0.0)
    quiet = ((ct + cb) * c < eh) & ((ct + ht) * c < eh) & ((cb + hb) * c < eh) & (c < em)
    sel = (c >= eo) & ~quiet
    margin = max(edge, 2)
    win_m = np.zeros_like(sel)
    win_m[ib:max(min(max(ie, ib), n_freq - 1), ib), margin:n_time - margin] = True
    sel &= win_m
    f, t = np.nonzero(sel)
    src = np.stack([vs + (t - vs + int(shift[d])) % nn for d in range(n_ifo)], axis=1)
    de = np.stack([E[d, f, src[:, d]] for d in range(n_ifo)], axis=1)
    return {"mask": sel, "frequency": f, "time": t, "live_mask": live,
            "energy": raw[f, t], "pix_det_energy": np.maximum(de, 0.0),
            "pix_det_index": src * n_freq + f[:, None]}


def _cases():
    base0 = np.zeros((7, 12))
    base1 = np.zeros_like(base0)
    base0[1:6, 2:10] = 4.0
    base1[1:6, 2:10] = 3.0
    base0[3, 4] = 30.0
    base1[3, 7] = 25.0
    v = np.array([0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0], np.int16)
    return [("zero_lag", base0, base1, [0.0, 0.0], 0.0, None, 0.0, None),
            ("shifted_wrap", base0, base1, [0.0, 3.0], 0.0, None, 0.0, None),
            ("shifted_veto", base0, base1, [0.0, 3.0], 0.0, v, 0.0, None),
            ("edge_crop", base0, base1, [0.0, 0.0], 2.0, None, 0.0, None),
            ("frequency_band", base0, base1, [0.0, 0.0], 0.0, None, 2.0, 4.0),
            ("empty", 0 * base0, 0 * base1, [0.0, 0.0], 0.0, None, 0.0, None)]


def _run_oracle(oracle, maps, shifts_sec, edge, veto, f_low, f_high, eo):
    E = np.stack([np.asarray(m, float) for m in maps])
    n_freq = E.shape[1]
    ref = min(shifts_sec)
    shift = np.array([int((s - ref) * 1.0 + 0.001) for s in shifts_sec], np.int64)
    ib, ie = _band(n_freq, 1.0, f_low, (n_freq - 1) if f_high is None else f_high)
    return E, shift, ib, ie, oracle.select_pixels(E, shift, veto, int(edge), ib, ie, eo)


def _assert_same(a, b, exact=False):
    for k in ("mask", "frequency", "time", "live_mask", "pix_det_index"):
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
    for k in ("energy", "pix_det_energy"):
        if exact:
            np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
        else:
            np.testing.assert_allclose(np.asarray(a[k]), np.asarray(b[k]), rtol=1e-12, err_msg=k)


@pytest.mark.parametrize("case", range(6))
def test_oracle_matches_dense_definitions_on_reference_cases(oracle, case):
    name, b0, b1, sh, edge, veto, f_low, f_high = _cases()[case]
    E, shift, ib, ie, o = _run_oracle(oracle, [b0, b1], sh, edge, veto, f_low, f_high, 5.0)
    _assert_same(o, _dense(E, shift, veto, int(edge), ib, ie, 5.0))
    if name == "empty":
        assert o["time"].size == 0
    else:
        assert o["time"].size > 0


def test_oracle_known_answers_shifted_veto(oracle):
    """test_circular_cat2_semantics.py:207-263."""
    data = np.ones((5, 8)) * 10.0
    E = np.stack([data, data])
    o = oracle.select_pixels(E, [0, 2], None, 0, 1, 4, 1.0)
    assert 4 in set(o["time"])
    veto = np.zeros(8, np.int16)
    veto[4] = 1
    o = oracle.select_pixels(E, [0, 2], veto, 0, 1, 4, 1.0)
    assert 4 not in set(o["time"]) and not o["live_mask"][4]
    veto[6] = 1
    o = oracle.select_pixels(E, [0, 2], veto, 0, 1, 4, 1.0)
    assert 4 in set(o["time"]) and o["live_mask"][4]
    idx = np.where(o["time"] == 4)[0][0]
    assert list(o["pix_det_index"][idx] // 5) == [4, 6]
    veto = np.zeros(8, np.int16)
    veto[4] = 1
    o = oracle.select_pixels(E, [0, 0], veto, 0, 1, 4, 1.0)
    assert 4 in set(o["time"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_dense_definitions_random(oracle, seed):
    rng = np.random.default_rng(seed)
    n_ifo, n_freq, n_time = 3, 17, 90
    E = rng.exponential(1.0, size=(n_ifo, n_freq, n_time))
    E[:, 5:9, 30:40] += 6.0
    shift = rng.integers(0, 40, size=n_ifo)
    veto = (rng.random(n_time) > 0.15).astype(np.int16)
    for edge, v in ((0, None), (3, veto), (6, None)):
        o = oracle.select_pixels(E, shift, v, edge, 2, 14, 4.0)
        _assert_same(o, _dense(E, shift, v, edge, 2, 14, 4.0))
        assert o["time"].size > 0


def _gpu_cache(E, edge_bins, ib, ie):
    import torch

    n_ifo, n_freq, n_time = E.shape
    return {"arrays_stack": torch.from_numpy(np.ascontiguousarray(E)).to("cuda:0"),
            "n_ifo": n_ifo, "n_freq": n_freq, "n_time": n_time, "dt": 1.0, "rate": 1.0,
            "edge_bins": edge_bins, "valid_start": edge_bins, "valid_stop": n_time - edge_bins,
            "nn_valid": n_time - 2 * edge_bins, "ib": ib, "ie": ie, "start": 0.0, "stop": 0.0,
            "f_low": 0.0, "f_high": 0.0, "shift_bins_by_lag": None}


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n_ifo,n_freq,n_time,edge,use_veto",
                         [(0, 2, 9, 64, 0, False), (1, 3, 33, 500, 4, True),
                          (2, 5, 65, 1200, 10, True), (3, 1, 5, 40, 2, False)])
def test_gpu_selection_bit_exact_many_lags(oracle, seed, n_ifo, n_freq, n_time, edge, use_veto):
    from pycwb_amd.coherence_native import select_network_pixels_lags

    print(f"case select seed={seed} n_ifo={n_ifo} {n_freq}x{n_time}", flush=True)
    rng = np.random.default_rng(seed)
    E = rng.exponential(1.0, size=(n_ifo, n_freq, n_time))
    for _ in range(6):
        f0, t0 = rng.integers(1, n_freq - 3), rng.integers(0, n_time - 12)
        E[:, f0:f0 + 3, t0:t0 + 10] += rng.uniform(2, 12)
    veto = (rng.random(n_time) > 0.1).astype(np.int16) if use_veto else None
    nn = n_time - 2 * edge
    shifts = np.concatenate([np.zeros((1, n_ifo), np.int64),
                             rng.integers(0, nn, size=(11, n_ifo))])
    ib, ie = 1, n_freq - 2
    eo = 2.5 * n_ifo
    out = select_network_pixels_lags(_gpu_cache(E, edge, ib, ie), eo, shifts, veto)
    assert len(out) == len(shifts)
    n_total = 0
    for lag, g in enumerate(out):
        o = oracle.select_pixels(E, shifts[lag], veto, edge, ib, ie, eo)
        _assert_same(g, o, exact=True)
        n_total += o["time"].size
    assert n_total > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(6))
def test_gpu_reference_cases(oracle, case):
    from pycwb_amd.coherence_native import select_network_pixels

    print(f"case select reference {case}", flush=True)
    name, b0, b1, sh, edge, veto, f_low, f_high = _cases()[case]
    maps = [_tf(b0, f_low=f_low, f_high=f_high), _tf(b1, f_low=f_low, f_high=f_high)]
    g = select_network_pixels(maps, 0, 5.0, lag_shifts=sh, veto=veto, edge=edge)
    E, shift, ib, ie, o = _run_oracle(oracle, [b0, b1], sh, edge, veto, f_low, f_high, 5.0)
    _assert_same(g, o, exact=True)
    assert g["live_samples"] == int(np.sum(o["live_mask"]))


@pytest.mark.gpu
def test_gpu_known_answers_and_cached_lags():
    from pycwb_amd.coherence_native import _build_selection_cache, select_network_pixels

    print("case select known answers", flush=True)
    data = np.ones((5, 8)) * 10.0
    maps = [_tf(data), _tf(data)]
    veto = np.zeros(8, np.int16)
    veto[4] = 1
    r = select_network_pixels(maps, 0, 1.0, lag_shifts=[0.0, 2.0], veto=veto)
    assert 4 not in set(r["time"]) and not r["live_mask"][4]
    veto[6] = 1
    r = select_network_pixels(maps, 0, 1.0, lag_shifts=[0.0, 2.0], veto=veto)
    idx = np.where(r["time"] == 4)[0][0]
    assert list(r["pix_det_index"][idx] // 5) == [4, 6]
    cache = _build_selection_cache(maps, edge=0.0, lag_shifts_by_lag=[[0.0, 0.0], [0.0, 2.0]])
    a = select_network_pixels(maps, 1, 1.0, selection_cache=cache)
    b = select_network_pixels(maps, 0, 1.0, lag_shifts=[0.0, 2.0])
    _assert_same(a, b, exact=True)
