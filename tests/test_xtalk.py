"""Cross-talk catalog format and the oracle's CSR builders (CPU).

Format: the reference's header tests (pycwb/modules/xtalk/tests/test_monster.py)
re-run against this reader, plus a full write -> parse round trip and the
``.npz`` cache.  Oracle pins: the fast CSR builder (getXTalk_pixels_fast,
monster.py:240-440) holds, row by row, exactly the entries of the all-pairs
builder (getXTalk_pixels_numba, :195-237, every ordered pair through
getXTalk) -- the reference's own claim for the fast path -- and a hand-built
two-resolution catalog fixes the coefficient placement and the c2/c3 swap.
"""

import struct

import numpy as np
import pytest

from pycwb_amd import xtalk as X


def _pack_header(nRes, tag=0.0, beta_order=0.0, precision=0.0, kwdm=0.0, layers=None):
    if layers is None:
        layers = list(range(1, nRes + 1))
    if tag != 0.0 or beta_order != 0.0 or precision != 0.0 or kwdm != 0.0:
        return struct.pack("<f4f" + f"{nRes}f", float(-nRes), tag, beta_order, precision, kwdm,
                           *layers)
    return struct.pack("<f" + f"{nRes}f", float(nRes), *layers)


def test_header_plain_and_extended():
    r = X._read_header(_pack_header(3))
    assert r["nRes"] == 3 and r["tag"] == 0.0 and r["layers"].tolist() == [1, 2, 3]
    assert r["offset"] == struct.calcsize("<f3f")
    r = X._read_header(_pack_header(2, tag=1.5, beta_order=2.0, precision=3.0, kwdm=4.0))
    assert (r["nRes"], r["tag"], r["beta_order"], r["precision"], r["kwdm"]) == \
        (2, 1.5, 2.0, 3.0, 4.0)
    assert r["layers"].tolist() == [1, 2]
    r = X._read_header(_pack_header(10, layers=[2 * i for i in range(10)]))
    assert r["layers"].tolist() == [2 * i for i in range(10)]


def test_metadata_from_bin_and_npz(tmp_path):
    p = tmp_path / "cat.bin"
    p.write_bytes(_pack_header(4, layers=[2, 4, 8, 16]))
    m = X.read_catalog_metadata(p)
    assert m["nRes"] == 4 and m["tag"] == 0.0 and m["layers"].tolist() == [2, 4, 8, 16]
    q = tmp_path / "meta.npz"
    np.savez(q, nRes=2, tag=1.0, beta_order=2.0, precision=0.01, kwdm=4.0,
             layers=np.array([1, 2], np.int32))
    m = X.read_catalog_metadata(q)
    assert (m["nRes"], m["tag"], m["beta_order"], m["precision"], m["kwdm"]) == \
        (2, 1.0, 2.0, 0.01, 4.0)


def test_catalog_round_trip_and_cache(tmp_path):
    cat = X.synthetic_catalog([4, 8, 16], seed=3)
    p = tmp_path / "OverlapCatalog-test.bin"
    X.write_catalog(p, cat.coeff, cat.lookup_table, cat.layers, tag=1.0, beta_order=6.0,
                    precision=10.0, kwdm=4.0)
    coeff, table, layers, nres, meta = X.parse_catalog(p.read_bytes())
    np.testing.assert_array_equal(coeff, cat.coeff)
    np.testing.assert_array_equal(table, cat.lookup_table)
    assert layers.tolist() == [4, 8, 16] and nres == 3 and meta["beta_order"] == 6.0
    loaded = X.XTalk.load(p, dump=True)
    assert (tmp_path / "OverlapCatalog-test.npz").exists()
    np.testing.assert_array_equal(loaded.coeff, cat.coeff)
    again = X.XTalk.load(p)          # served from the .npz cache
    np.testing.assert_array_equal(again.lookup_table, cat.lookup_table)
    with pytest.raises(ValueError):
        X.XTalk(cat.coeff, cat.lookup_table[:, :, :3], cat.layers, 3)


def _cluster(rng, cat, n_per_res=(12, 10, 8), t_span=8):
    """Pixels in a compact time-frequency block at every resolution."""
    lay, tim = [], []
    for r, n in enumerate(n_per_res):
        L = int(cat.layers[r])
        span = t_span * (int(cat.layers[-1]) // L)
        ts = rng.integers(10, 10 + span, size=n)
        fs = rng.integers(0, L + 1, size=n)
        lay.append(np.full(n, L + 1))
        tim.append(ts * (L + 1) + fs)
    lay, tim = np.concatenate(lay), np.concatenate(tim)
    _, first = np.unique(lay.astype(np.int64) * (1 << 32) + tim, return_index=True)
    lay, tim = lay[np.sort(first)], tim[np.sort(first)]     # distinct pixels
    perm = rng.permutation(lay.size)
    return lay[perm].astype(np.int32), tim[perm].astype(np.int64)



@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fast_builder_matches_all_pairs(oracle, seed):
    rng = np.random.default_rng(seed)
    cat = X.synthetic_catalog([4, 8, 16], seed=seed)
    lay, tim = _cluster(rng, cat)
    lk_f, xt_f = oracle.xtalk_pixels_fast(lay, tim, cat.layers, cat.lookup_table, cat.coeff)
    lk_p, xt_p = oracle.xtalk_pixels_pairwise(lay, tim, cat.layers, cat.lookup_table,
                                              cat.coeff)
    assert xt_f.shape[0] == xt_p.shape[0] > lay.size          # more than the self entries
    for i, ((a, b), (c, d)) in enumerate(zip(lk_f, lk_p)):
        f = xt_f[a:b][np.lexsort(xt_f[a:b, ::-1].T)]
        p = xt_p[c:d][np.lexsort(xt_p[c:d, ::-1].T)]
        assert f.shape == p.shape, i
        cols = [0, 1, 2, 4, 5, 6, 7]
        np.testing.assert_array_equal(f[:, cols], p[:, cols])
        np.testing.assert_allclose(f[:, 3], p[:, 3], rtol=2e-7)   # sum order differs


def test_host_and_oracle_pair_lookup_agree(oracle):
    rng = np.random.default_rng(8)
    cat = X.synthetic_catalog([4, 8, 16], seed=8)
    lay, tim = _cluster(rng, cat)
    for _ in range(300):
        i, j = rng.integers(lay.size, size=2)
        a = X.get_xtalk(lay[i], tim[i], lay[j], tim[j], cat.layers, cat.coeff, cat.lookup_table)
        b = oracle.xtalk_get(lay[i], tim[i], lay[j], tim[j], cat.layers, cat.lookup_table,
                             cat.coeff)
        np.testing.assert_array_equal(a, b)


def test_two_resolution_placement_and_swap(oracle):
    """Layers (2, 4): one coarse-grid neighbour of the fine pixel at
    frequency 1, even time bin 4.  Fine pixel i: nLayer 5, t = 4*5 + 1 = 21;
    coarse pixel j: nLayer 3, relative index 7 from base (4 - 0) * 2 * 3 = 24,
    t = 31.  Row i gets (j, c1, c3, c2, c4), row j the swapped (i, c1, c2, c3, c4)."""
    layers = np.array([2, 4], np.int32)
    table = np.zeros((2, 2, 5, 2, 2), np.int32)
    coeff = np.array([[7, 0.5, -0.25, 0.125, 2.0]], np.float32)
    table[1, 0, 1, 0] = [0, 1]
    lay = np.array([5, 3], np.int32)
    tim = np.array([21, 31], np.int64)
    lk, xt = oracle.xtalk_pixels_fast(lay, tim, layers, table, coeff)
    c1, c2, c3, c4 = np.float32([0.5, -0.25, 0.125, 2.0])
    assert lk.tolist() == [[0, 1], [1, 2]]
    np.testing.assert_array_equal(xt[0], [1, c1 * c1 + c2 * c2, c3 * c3 + c4 * c4,
                                          c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4, c1, c3, c2, c4])
    np.testing.assert_array_equal(xt[1], [0, c1 * c1 + c3 * c3, c2 * c2 + c4 * c4,
                                          c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4, c1, c2, c3, c4])
    np.testing.assert_array_equal(oracle.xtalk_get(5, 21, 3, 31, layers, table, coeff),
                                  [c1, c2, c3, c4])
    np.testing.assert_array_equal(oracle.xtalk_get(3, 31, 5, 21, layers, table, coeff),
                                  [c1, c3, c2, c4])
    # the other time parity and a shifted neighbour find nothing
    assert oracle.xtalk_get(5, 26, 3, 31, layers, table, coeff)[0] == 3.0
    assert oracle.xtalk_get(5, 21, 3, 32, layers, table, coeff)[0] == 3.0
