"""WDM cross-talk catalog and the clusters' cross-talk CSR on the GPU.

Drop-in for pycwb/modules/xtalk (monster.py, type.py):
  read_catalog_metadata / load_catalog  (monster.py:13-153)  the catalog file
      formats: cWB's binary ``.bin`` / ``.xbin`` and the ``.npz`` cache;
  XTalk.load / get_xtalk_pixels / get_xtalk_pixels_from_arrays / get_xtalk
      (type.py:6-74), with the CSR built by HIP kernels
      (getXTalk_pixels_fast, monster.py:240-440, is the reference's builder);
  XTalk.build_batch: the CSRs of many clusters in one pass, left on the device
      in the layout the best-sky statistics and the sub-network cut take.

Binary format (little-endian, as written by cWB's monster::write): float nRes
(negative: four more floats tag, BetaOrder, precision, KWDM follow), nRes
floats of layer counts, then for every resolution pair i >= j, frequency index
k = 0..layers[i] and time parity l = 0, 1: a float count and that many
(int32 relative index, 4 float32 coefficients) records.
"""

from __future__ import annotations

import ctypes as C
import logging
import pathlib
import struct
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import lib, ptr, require_gpu, stream_ptr

logger = logging.getLogger(__name__)

RECORD = np.dtype([("index", "<i4"), ("CC", "<f4", (4,))])


def _read_header(data: bytes) -> dict:
    """monster.py:13-40."""
    off = 0
    (n,) = struct.unpack_from("<f", data, off)
    off += 4
    nres = int(n)
    tag = beta = precision = kwdm = 0.0
    if nres < 0:
        nres = -nres
        tag, beta, precision, kwdm = struct.unpack_from("<4f", data, off)
        off += 16
    layers = np.array(struct.unpack_from(f"<{nres}f", data, off), dtype=np.float32)
    off += 4 * nres
    return {"nRes": nres, "tag": tag, "beta_order": beta, "precision": precision,
            "kwdm": kwdm, "layers": layers.astype(np.int32), "offset": off}


def read_catalog_metadata(fn) -> dict:
    """monster.py:43-60."""
    fn = pathlib.Path(fn)
    if fn.suffix == ".npz":
        with np.load(fn, allow_pickle=False) as d:
            get = lambda k: float(d[k]) if k in d.files else 0.0
            return {"nRes": int(d["nRes"]), "tag": get("tag"), "beta_order": get("beta_order"),
                    "precision": get("precision"), "kwdm": get("kwdm"),
                    "layers": np.array(d["layers"], dtype=np.int32)}
    meta = _read_header(fn.read_bytes())
    meta.pop("offset")
    return meta


def parse_catalog(data: bytes):
    """(coeff (E, 5) float32, table (nRes, nRes, maxL+1, 2, 2) int32, layers,
    nRes, metadata) of a binary catalog (monster.py:121-145)."""
    meta = _read_header(data)
    nres, layers, off = meta["nRes"], meta["layers"], meta["offset"]
    maxl = int(layers.max()) if nres else 0
    table = np.zeros((nres, nres, maxl + 1, 2, 2), dtype=np.int32)
    parts = []
    entry = 0
    for i in range(nres):
        for j in range(i + 1):
            for k in range(int(layers[i]) + 1):
                for par in range(2):
                    (cnt,) = struct.unpack_from("<f", data, off)
                    off += 4
                    cnt = int(cnt)
                    rec = np.frombuffer(data, dtype=RECORD, count=cnt, offset=off)
                    off += cnt * RECORD.itemsize
                    table[i, j, k, par, 0] = entry
                    entry += cnt
                    table[i, j, k, par, 1] = entry
                    if cnt:
                        parts.append(np.concatenate(
                            [rec["index"].astype(np.float32)[:, None], rec["CC"]], axis=1))
    coeff = np.concatenate(parts).astype(np.float32) if parts else np.zeros((0, 5), np.float32)
    return coeff, table, layers, nres, meta


def write_catalog(fn, coeff, table, layers, tag=0.0, beta_order=0.0, precision=0.0,
                  kwdm=0.0) -> None:
    """Binary catalog in the format parse_catalog reads (for tests and for
    converting ``.npz`` caches back)."""
    layers = np.asarray(layers, dtype=np.int32)
    nres = layers.size
    ext = any(v != 0.0 for v in (tag, beta_order, precision, kwdm))
    out = [struct.pack("<f", float(-nres if ext else nres))]
    if ext:
        out.append(struct.pack("<4f", tag, beta_order, precision, kwdm))
    out.append(struct.pack(f"<{nres}f", *[float(x) for x in layers]))
    coeff = np.asarray(coeff, dtype=np.float32)
    for i in range(nres):
        for j in range(i + 1):
            for k in range(int(layers[i]) + 1):
                for par in range(2):
                    a, b = int(table[i, j, k, par, 0]), int(table[i, j, k, par, 1])
                    out.append(struct.pack("<f", float(b - a)))
                    rec = np.zeros(b - a, dtype=RECORD)
                    rec["index"] = coeff[a:b, 0].astype(np.int32)
                    rec["CC"] = coeff[a:b, 1:5]
                    out.append(rec.tobytes())
    pathlib.Path(fn).write_bytes(b"".join(out))


def load_catalog(fn, dump=True):
    """monster.py:63-153: (xtalk_coeff, xtalk_lookup_table, layers, nRes).  A
    ``.npz`` cache next to a ``.bin``/``.xbin`` file is used when present; with
    ``dump`` the parsed catalog is saved as that cache."""
    p = pathlib.Path(fn)
    if p.suffix == ".npz":
        logger.info("Loading %s", fn)
        with np.load(p, allow_pickle=False) as d:
            return (np.asarray(d["xtalk_coeff"], np.float32),
                    np.asarray(d["xtalk_lookup_table"], np.int32),
                    np.asarray(d["layers"], np.int32), int(d["nRes"]))
    if p.suffix in (".bin", ".xbin") and p.with_suffix(".npz").exists():
        logger.info(".npz file found: %s, loading the catalog from the .npz file.",
                    p.with_suffix(".npz"))
        return load_catalog(p.with_suffix(".npz"))
    coeff, table, layers, nres, meta = parse_catalog(p.read_bytes())
    if dump:
        np.savez(p.with_suffix(".npz"), xtalk_coeff=coeff, xtalk_lookup_table=table,
                 layers=layers, nRes=nres, tag=meta["tag"], beta_order=meta["beta_order"],
                 precision=meta["precision"], kwdm=meta["kwdm"])
    return coeff, table, layers, nres


def get_xtalk(nlayer1, indx1, nlayer2, indx2, layers, coeff, table) -> np.ndarray:
    """getXTalk (monster.py:157-192): (c1, c2, c3, c4) of a pixel pair, 3s when
    the catalog has no entry (c2 and c3 swapped when pixel 1 is the coarser)."""
    layers = np.asarray(layers)
    r1 = r2 = -1
    for i, L in enumerate(layers):
        if L == nlayer1 - 1:
            r1 = i
        if L == nlayer2 - 1:
            r2 = i
    if r1 == -1 or r2 == -1:
        raise ValueError(f"Resolution not found: {nlayer1} {nlayer2} in layers")
    swap = r1 < r2
    if swap:
        r1, r2, indx1, indx2 = r2, r1, indx2, indx1
    L1 = int(layers[r1])
    time1, freq1 = divmod(int(indx1), L1 + 1)
    odd = time1 & 1
    index = int(indx2) - (time1 - odd) * (L1 // int(layers[r2])) * (int(layers[r2]) + 1)
    ret = np.full(4, 3.0, dtype=np.float32)
    a, b = table[r1, r2, freq1, odd]
    for row in coeff[a:b]:
        if index == int(row[0]):
            ret[:] = row[1:5]
            if swap:
                ret[1], ret[2] = ret[2], ret[1]
            break
    return ret


@dataclass
class XTalkCSR:
    """Cross-talk CSRs of a batch of clusters on the device: cluster c's rows
    are lookup[pix_off[c]:pix_off[c+1]] ([start, end) relative to xt_off[c])
    into xt[xt_off[c]:xt_off[c+1]].  The buffers hold at least one row."""
    pix_off: np.ndarray     # int64 (C + 1,)
    xt_off: np.ndarray      # int64 (C + 1,)
    lookup: torch.Tensor    # int32 (max(P_total, 1), 2)
    xt: torch.Tensor        # float32 (max(total, 1), 8)

    def cluster(self, c):
        """(lookup (P, 2) int32, xt (m, 8) float32) of cluster c, on the host."""
        a, b = int(self.pix_off[c]), int(self.pix_off[c + 1])
        u, v = int(self.xt_off[c]), int(self.xt_off[c + 1])
        return self.lookup[a:b].cpu().numpy(), self.xt[u:v].cpu().numpy()


def self_xtalk_batch(n_pix, device=None) -> XTalkCSR:
    """Self-only cross-talk of a whole batch as device CSRs: each pixel lists
    itself with coefficients (1, 0, 0, 1) in columns 4-7, the CSRs of
    pycwb_amd.skystats.self_xtalk for every cluster laid end to end."""
    device = device or require_gpu()
    n_pix = np.asarray(n_pix, dtype=np.int64).reshape(-1)
    if (n_pix < 0).any():
        raise ValueError("cluster sizes must be >= 0")
    pix_off = np.concatenate([[0], np.cumsum(n_pix)]).astype(np.int64)
    n = int(pix_off[-1])
    first = torch.from_numpy(np.repeat(pix_off[:-1], n_pix)).to(device)
    j = torch.arange(n, dtype=torch.int64, device=device) - first   # index inside the cluster
    lookup = torch.zeros((max(n, 1), 2), dtype=torch.int32, device=device)
    lookup[:n, 0] = j.to(torch.int32)
    lookup[:n, 1] = (j + 1).to(torch.int32)
    xt = torch.zeros((max(n, 1), 8), dtype=torch.float32, device=device)
    xt[:n, 0] = j.to(torch.float32)
    xt[:n, 4] = 1.0
    xt[:n, 7] = 1.0
    return XTalkCSR(pix_off=pix_off, xt_off=pix_off.copy(), lookup=lookup, xt=xt)


@dataclass
class XTalk:
    """The WDM cross-talk catalog (type.py:6-74)."""
    coeff: np.ndarray          # (E, 5) float32: relative index, c1..c4
    lookup_table: np.ndarray   # (nRes, nRes, maxL+1, 2, 2) int32
    layers: np.ndarray         # (nRes,) int32
    nRes: int

    @classmethod
    def load(cls, fn, dump=True) -> "XTalk":
        coeff, table, layers, nres = load_catalog(fn, dump)
        return cls(np.asarray(coeff, np.float32).reshape(-1, 5),
                   np.asarray(table, np.int32), np.asarray(layers, np.int32), int(nres))

    def __post_init__(self):
        self.coeff = np.ascontiguousarray(self.coeff, np.float32).reshape(-1, 5)
        self.lookup_table = np.ascontiguousarray(self.lookup_table, np.int32)
        self.layers = np.ascontiguousarray(self.layers, np.int32)
        if self.layers.size != int(self.nRes) or (self.layers <= 0).any():
            raise ValueError("catalog layers must be nRes positive counts")
        maxl = int(self.layers.max())
        if self.lookup_table.shape != (self.nRes, self.nRes, maxl + 1, 2, 2):
            raise ValueError(f"lookup table shape {self.lookup_table.shape} does not match "
                             f"nRes={self.nRes}, max layers={maxl}")
        t = self.lookup_table
        if t.size and (t.min() < 0 or t.max() > self.coeff.shape[0] or
                       (t[..., 1] < t[..., 0]).any()):
            raise ValueError("lookup table ranges outside the coefficient table")
        self._dev = None

    def _device(self, device):
        if self._dev is None or self._dev[0] != device:
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
            self._dev = (device, t(self.layers), t(self.lookup_table),
                         t(np.concatenate([self.coeff.ravel(), np.zeros(5, np.float32)])))
        return self._dev[1:]

    def build_batch(self, clusters, stream=None, device=None) -> XTalkCSR:
        """CSRs of many clusters; clusters[c] = (layers, times) of its pixels
        (PixelArrays.layers, PixelArrays.time)."""
        lay, tim, sizes = [], [], []
        for layers, times in clusters:
            layers = np.asarray(layers).reshape(-1)
            times = np.asarray(times).reshape(-1)
            if layers.shape != times.shape:
                raise ValueError("layers and times must have the same length")
            lay.append(layers)
            tim.append(times)
            sizes.append(layers.size)
        if not sizes:
            raise ValueError("no clusters")
        return self.build_batch_flat(np.concatenate(lay), np.concatenate(tim), sizes, stream,
                                     device)

    def build_batch_flat(self, layers, times, sizes, stream=None, device=None) -> XTalkCSR:
        """build_batch on the batch's pixels laid end to end: ``layers`` and
        ``times`` (host arrays or device tensors) hold cluster 0's pixels, then
        cluster 1's, ...; ``sizes[c]`` is cluster c's pixel count."""
        device = device or require_gpu()
        d = device
        sizes = np.asarray(sizes, dtype=np.int64).reshape(-1)
        n_cl = sizes.size
        if n_cl == 0:
            raise ValueError("no clusters")
        if (sizes < 0).any():
            raise ValueError("cluster sizes must be >= 0")
        n = int(sizes.sum())

        def dev(a, dtype):
            if isinstance(a, torch.Tensor):
                return a.reshape(-1).to(device=d, dtype=dtype).contiguous()
            return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1),
                                                         dtype=np.dtype(str(dtype)[6:]))).to(d)

        nl_d, tm_d = dev(layers, torch.int64), dev(times, torch.int64)
        if nl_d.numel() != n or tm_d.numel() != n:
            raise ValueError("layers and times must hold sum(sizes) pixels")
        if n:
            lo_l, hi_l = torch.aminmax(nl_d)
            lo_t, hi_t = torch.aminmax(tm_d)
            lim = torch.stack([lo_l, hi_l, lo_t, hi_t]).cpu().numpy()
            if lim[2] < 0 or lim[0] < 1 or lim[1] >= 1 << 20 or lim[3] >= 1 << 32:
                raise ValueError("pixel layers must be in [1, 2^20) and times in [0, 2^32)")
        nl_d = torch.cat([nl_d.to(torch.int32), torch.zeros(1, dtype=torch.int32, device=d)])
        tm_d = torch.cat([tm_d, torch.zeros(1, dtype=torch.int64, device=d)])
        pix_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        # open-addressing table of each cluster: the power of two >= 2 P - 1 (at least 2)
        _, e = np.frexp(np.maximum(2 * sizes - 1, 1).astype(np.float64))
        slots = np.left_shift(1, np.maximum(1, e.astype(np.int64)))
        hash_off = np.concatenate([[0], np.cumsum(slots)]).astype(np.int64)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(d)
        layers_d, table_d, coeff_d = self._device(d)
        pix_off_d, hash_off_d = t(pix_off), t(hash_off)
        n_slots = int(hash_off[-1])
        hkey = torch.empty(n_slots, dtype=torch.int64, device=d)
        hval = torch.empty(n_slots, dtype=torch.int32, device=d)
        counts = torch.empty(max(n, 1), dtype=torch.int32, device=d)
        row_off = torch.empty(n + 1, dtype=torch.int64, device=d)
        block_sums = torch.empty(int(lib().pycwb_xtalk_scan_blocks(n)), dtype=torch.int64,
                                 device=d)
        xt_off = torch.empty(n_cl + 1, dtype=torch.int64, device=d)
        lookup = torch.empty((max(n, 1), 2), dtype=torch.int32, device=d)
        desc = _lib.XtalkDesc()
        desc.n_clusters, desc.n_pixels, desc.n_slots = n_cl, n, n_slots
        desc.pix_off, desc.nlayer, desc.time, desc.hash_off = (ptr(pix_off_d), ptr(nl_d),
                                                               ptr(tm_d), ptr(hash_off_d))
        desc.n_res, desc.max_layer = int(self.nRes), int(self.layers.max())
        desc.layers, desc.table, desc.coeff = ptr(layers_d), ptr(table_d), ptr(coeff_d)
        desc.n_coeff = int(self.coeff.shape[0])
        desc.hash_keys, desc.hash_vals, desc.counts = ptr(hkey), ptr(hval), ptr(counts)
        desc.row_off, desc.block_sums, desc.xt_off = ptr(row_off), ptr(block_sums), ptr(xt_off)
        desc.lookup = ptr(lookup)
        stream = stream or torch.cuda.current_stream()
        st = stream_ptr(stream)
        # the uploads above ran on torch's current stream
        stream.wait_stream(torch.cuda.current_stream(d))
        _lib.check(lib().pycwb_xtalk_count(C.byref(desc), st), "pycwb_xtalk_count")
        stream.synchronize()
        xt_off_h = xt_off.cpu().numpy()
        total = int(xt_off_h[-1])
        xt = torch.empty((max(total, 1), 8), dtype=torch.float32, device=d)
        keys = torch.empty(max(total, 1), dtype=torch.int64, device=d)
        desc.sort_keys, desc.xt, desc.xt_capacity = ptr(keys), ptr(xt), total
        _lib.check(lib().pycwb_xtalk_fill(C.byref(desc), st), "pycwb_xtalk_fill")
        stream.synchronize()
        return XTalkCSR(pix_off=pix_off, xt_off=xt_off_h, lookup=lookup, xt=xt)

    def get_xtalk_pixels_from_arrays(self, layers, times, check=True):
        """type.py:51-60: (lookup (P, 2) int32, xtalk (m, 8) float32)."""
        return self.build_batch([(layers, times)]).cluster(0)

    def get_xtalk_pixels(self, pixels, check=True):
        """type.py:34-49 for PixelArrays (``layers``/``time`` arrays) or a list of
        objects with ``layers`` and ``time`` attributes."""
        if hasattr(pixels, "layers") and hasattr(pixels, "time") and \
                not isinstance(pixels, (list, tuple)):
            return self.get_xtalk_pixels_from_arrays(pixels.layers, pixels.time, check)
        lay = np.array([p.layers for p in pixels], dtype=np.int64)
        tim = np.array([p.time for p in pixels], dtype=np.int64)
        return self.get_xtalk_pixels_from_arrays(lay, tim, check)

    def get_xtalk(self, pix1, pix2):
        """type.py:62-74: coefficients of one pixel pair (host)."""
        l1, t1 = pix1 if isinstance(pix1, tuple) else (pix1.layers, pix1.time)
        l2, t2 = pix2 if isinstance(pix2, tuple) else (pix2.layers, pix2.time)
        return get_xtalk(l1, t1, l2, t2, self.layers, self.coeff, self.lookup_table)


def synthetic_catalog(layers, seed=0, per_bucket=(2, 6), dt=2):
    """A catalog with the real one's structure and random coefficients: each
    (i >= j, k, parity) bucket lists neighbours within +-dt time bins of
    resolution j around the pixel's position.  For tests and benchmarks (the
    real catalog comes from cWB's wdmXTalk directory)."""
    rng = np.random.default_rng(seed)
    layers = np.asarray(layers, dtype=np.int32)
    nres = layers.size
    maxl = int(layers.max())
    table = np.zeros((nres, nres, maxl + 1, 2, 2), dtype=np.int32)
    rows = []
    e = 0
    for i in range(nres):
        for j in range(i + 1):
            Li, Lj = int(layers[i]), int(layers[j])
            ratio = Li // Lj
            for k in range(Li + 1):
                for par in range(2):
                    # neighbour time bins (in resolution j) and frequencies near k's band
                    fj = int(round(k * Lj / Li))
                    cand = [(tj, f) for tj in range(-dt, ratio + dt + 1)
                            for f in range(max(0, fj - 1), min(Lj, fj + 1) + 1)]
                    m = int(rng.integers(per_bucket[0], per_bucket[1] + 1))
                    pick = rng.choice(len(cand), size=min(m, len(cand)), replace=False)
                    if i == j and k <= Li:
                        self_pos = (par, k)
                        if self_pos not in [cand[q] for q in pick] and self_pos in cand:
                            pick = np.concatenate([[cand.index(self_pos)], pick[:-1]])
                    table[i, j, k, par, 0] = e
                    for q in pick:
                        tj, f = cand[q]
                        idx = tj * (Lj + 1) + f
                        cc = rng.normal(scale=0.3, size=4)
                        if i == j and (tj, f) == (par, k):
                            cc = np.array([1.0, 0.0, 0.0, 1.0])
                        rows.append([idx, *cc])
                        e += 1
                    table[i, j, k, par, 1] = e
    coeff = np.asarray(rows, dtype=np.float32).reshape(-1, 5)
    return XTalk(coeff, table, layers, nres)
