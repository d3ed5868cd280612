"""HEALPix pixelisation (RING and NESTED) without healpy, host side, float64.

The reference calls ``healpy.pix2ang`` for its sky grid
(pycwb/types/detector.py:1220-1240) and ``healpy.ang2pix`` / ``read_map`` for
Custom sky masks (pycwb/modules/likelihoodWP/sky_mask.py:143-175).  The
pixelisation is fixed by its definition (Gorski et al. 2005, ApJ 622, 759):
12 base faces, ``nside**2`` equal-area pixels each, pixel centres on
iso-latitude rings; the polar caps (|z| > 2/3) hold 4i pixels on ring i, the
equatorial belt 4 nside per ring.  The functions below restate it from those
equations (the near-pole branches use the sin(theta) form for accuracy).

Also here: a reader of the FITS binary-table maps that ``healpy.write_map``
writes (the first column, any ORDERING), and ``.npy`` maps.
"""

from __future__ import annotations

import math
import os

import numpy as np

TWO_PI = 2.0 * math.pi
HALF_PI = 0.5 * math.pi


def nside2npix(nside: int) -> int:
    return 12 * int(nside) * int(nside)


def npix2nside(npix: int) -> int:
    nside = int(round(math.sqrt(int(npix) / 12.0)))
    if nside2npix(nside) != int(npix):
        raise ValueError(f"{npix} is not a valid HEALPix pixel count")
    return nside


def _isqrt(v: np.ndarray) -> np.ndarray:
    r = np.floor(np.sqrt(v.astype(np.float64))).astype(np.int64)
    r = np.where(r * r > v, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= v, r + 1, r)
    return r


def pix2ang_ring(nside: int, ipix) -> tuple[np.ndarray, np.ndarray]:
    """RING pixel -> (theta, phi) radians (Gorski et al. 2005, eqs. 4-9)."""
    ipix = np.asarray(ipix, dtype=np.int64)
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    theta = np.empty(ipix.shape, dtype=np.float64)
    phi = np.empty(ipix.shape, dtype=np.float64)
    fact2 = 4.0 / npix
    fact1 = (nside << 1) * fact2

    north = ipix < ncap
    south = ipix >= npix - ncap
    equ = ~(north | south)

    if np.any(north):
        p = ipix[north]
        iring = (1 + _isqrt(1 + 2 * p)) >> 1
        iphi = p + 1 - 2 * iring * (iring - 1)
        tmp = (iring * iring).astype(np.float64) * fact2
        z = 1.0 - tmp
        sth = np.sqrt(tmp * (2.0 - tmp))
        theta[north] = np.arctan2(sth, z)
        phi[north] = (iphi - 0.5) * (0.5 * np.pi / iring)
    if np.any(equ):
        p = ipix[equ] - ncap
        iring = p // (4 * nside) + nside
        iphi = p % (4 * nside) + 1
        fodd = np.where(((iring + nside) & 1) != 0, 1.0, 0.5)
        z = (2 * nside - iring) * fact1
        theta[equ] = np.arccos(z)
        phi[equ] = (iphi - fodd) * (np.pi * 0.75 * fact1)
    if np.any(south):
        p = npix - ipix[south]
        iring = (1 + _isqrt(2 * p - 1)) >> 1
        iphi = 4 * iring + 1 - (p - 2 * iring * (iring - 1))
        tmp = (iring * iring).astype(np.float64) * fact2
        z = tmp - 1.0
        sth = np.sqrt(tmp * (2.0 - tmp))
        theta[south] = np.arctan2(sth, z)
        phi[south] = (iphi - 0.5) * (0.5 * np.pi / iring)
    return theta, phi


def _loc(theta, phi):
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    theta, phi = np.broadcast_arrays(theta, phi)
    if np.any((theta < 0.0) | (theta > np.pi)) or not np.all(np.isfinite(theta)):
        raise ValueError("THETA is out of range [0,pi]")
    z = np.cos(theta)
    sth = np.sin(theta)
    tt = np.mod(phi, TWO_PI) * (1.0 / HALF_PI)
    tt = np.where(tt >= 4.0, tt - 4.0, tt)
    return z, sth, tt


def _polar_tmp(nside, z, sth):
    """nside * sqrt(3 (1 - |z|)), from sin(theta) near the poles."""
    za = np.abs(z)
    near = sth < 0.01  # sin(theta)^2 / (1 + |z|) avoids the 1 - |z| cancellation
    safe = np.where(near, sth * sth / (1.0 + za), 1.0 - za)
    return nside * np.sqrt(3.0 * safe)


def ang2pix_ring(nside: int, theta, phi) -> np.ndarray:
    """(theta, phi) radians -> RING pixel (Gorski et al. 2005, sect. 4)."""
    nside = int(nside)
    z, sth, tt = _loc(theta, phi)
    za = np.abs(z)
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    out = np.empty(z.shape, dtype=np.int64)

    eq = za <= 2.0 / 3.0
    if np.any(eq):
        t1 = nside * (0.5 + tt[eq])
        t2 = nside * (z[eq] * 0.75)
        jp = (t1 - t2).astype(np.int64)      # ascending edge line index
        jm = (t1 + t2).astype(np.int64)      # descending edge line index
        ir = nside + 1 + jp - jm             # ring number counted from z = 2/3
        kshift = 1 - (ir & 1)
        ip = (jp + jm - nside + kshift + 1) // 2
        ip = np.mod(ip, 4 * nside)
        out[eq] = ncap + (ir - 1) * 4 * nside + ip
    pol = ~eq
    if np.any(pol):
        tp = tt[pol] - np.floor(tt[pol])
        tmp = _polar_tmp(nside, z[pol], sth[pol])
        jp = (tp * tmp).astype(np.int64)
        jm = ((1.0 - tp) * tmp).astype(np.int64)
        ir = jp + jm + 1                     # ring number counted from the closest pole
        ip = (tt[pol] * ir).astype(np.int64)
        ip = np.mod(ip, 4 * ir)
        out[pol] = np.where(z[pol] > 0, 2 * ir * (ir - 1) + ip, npix - 2 * ir * (ir + 1) + ip)
    return out


def _spread_bits(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.int64) & 0xFFFFFFFF
    v = (v | (v << 16)) & 0x0000FFFF0000FFFF
    v = (v | (v << 8)) & 0x00FF00FF00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v << 2)) & 0x3333333333333333
    v = (v | (v << 1)) & 0x5555555555555555
    return v


def _compress_bits(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.int64) & 0x5555555555555555
    v = (v | (v >> 1)) & 0x3333333333333333
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FF
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFF
    v = (v | (v >> 16)) & 0x00000000FFFFFFFF
    return v


def ang2pix_nest(nside: int, theta, phi) -> np.ndarray:
    """(theta, phi) radians -> NESTED pixel: base face, then the (x, y)
    position inside it with the bits of x and y interleaved."""
    nside = int(nside)
    if nside & (nside - 1):
        raise ValueError("NESTED ordering needs a power-of-two nside")
    z, sth, tt = _loc(theta, phi)
    za = np.abs(z)
    face = np.empty(z.shape, dtype=np.int64)
    ix = np.empty(z.shape, dtype=np.int64)
    iy = np.empty(z.shape, dtype=np.int64)

    eq = za <= 2.0 / 3.0
    if np.any(eq):
        t1 = nside * (0.5 + tt[eq])
        t2 = nside * (z[eq] * 0.75)
        jp = (t1 - t2).astype(np.int64)
        jm = (t1 + t2).astype(np.int64)
        ifp = jp // nside
        ifm = jm // nside
        face[eq] = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))
        ix[eq] = jm & (nside - 1)
        iy[eq] = nside - (jp & (nside - 1)) - 1
    pol = ~eq
    if np.any(pol):
        ntt = np.minimum(3, tt[pol].astype(np.int64))
        tp = tt[pol] - ntt
        tmp = _polar_tmp(nside, z[pol], sth[pol])
        jp = np.minimum((tp * tmp).astype(np.int64), nside - 1)
        jm = np.minimum(((1.0 - tp) * tmp).astype(np.int64), nside - 1)
        north = z[pol] >= 0
        face[pol] = np.where(north, ntt, ntt + 8)
        ix[pol] = np.where(north, nside - jm - 1, jp)
        iy[pol] = np.where(north, nside - jp - 1, jm)
    return face * nside * nside + _spread_bits(ix) + (_spread_bits(iy) << 1)


# ring of the south-west corner of each base face (in units of nside) and its
# longitude offset (in units of pi/4)
_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4], dtype=np.int64)
_JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7], dtype=np.int64)


def pix2ang_nest(nside: int, ipix) -> tuple[np.ndarray, np.ndarray]:
    """NESTED pixel -> (theta, phi) radians."""
    nside = int(nside)
    ipix = np.asarray(ipix, dtype=np.int64)
    npface = nside * nside
    face = ipix // npface
    sub = ipix & (npface - 1)
    ix = _compress_bits(sub)
    iy = _compress_bits(sub >> 1)
    jr = _JRLL[face] * nside - ix - iy - 1          # ring index 1 .. 4 nside - 1
    nl4 = 4 * nside
    theta = np.empty(ipix.shape, dtype=np.float64)
    nr = np.empty(ipix.shape, dtype=np.int64)
    kshift = np.empty(ipix.shape, dtype=np.int64)
    fact = 1.0 / (3.0 * npface)

    north = jr < nside
    south = jr > 3 * nside
    equ = ~(north | south)
    if np.any(north):
        r = jr[north]
        tmp = (r * r) * fact
        theta[north] = np.arctan2(np.sqrt(tmp * (2.0 - tmp)), 1.0 - tmp)
        nr[north] = r
        kshift[north] = 0
    if np.any(south):
        r = nl4 - jr[south]
        tmp = (r * r) * fact
        theta[south] = np.arctan2(np.sqrt(tmp * (2.0 - tmp)), tmp - 1.0)
        nr[south] = r
        kshift[south] = 0
    if np.any(equ):
        theta[equ] = np.arccos((2 * nside - jr[equ]) * (2.0 / (3.0 * nside)))
        nr[equ] = nside
        kshift[equ] = (jr[equ] - nside) & 1
    jp = (_JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > nl4, jp - nl4, jp)
    jp = np.where(jp < 1, jp + nl4, jp)
    phi = (jp - (kshift + 1) * 0.5) * (HALF_PI / nr)
    return theta, phi


def nest2ring(nside: int, ipix) -> np.ndarray:
    th, ph = pix2ang_nest(nside, ipix)
    return ang2pix_ring(nside, th, ph)


def ring2nest(nside: int, ipix) -> np.ndarray:
    th, ph = pix2ang_ring(nside, ipix)
    return ang2pix_nest(nside, th, ph)


def ang2pix(nside: int, theta, phi, nest: bool = False) -> np.ndarray:
    return ang2pix_nest(nside, theta, phi) if nest else ang2pix_ring(nside, theta, phi)


def pix2ang(nside: int, ipix, nest: bool = False):
    return pix2ang_nest(nside, ipix) if nest else pix2ang_ring(nside, ipix)


# ---- map files -------------------------------------------------------------

_FITS_BLOCK = 2880
_TFORM = {"E": ">f4", "D": ">f8", "J": ">i4", "K": ">i8", "I": ">i2", "B": "u1"}


def _fits_header(f) -> dict:
    cards = {}
    while True:
        block = f.read(_FITS_BLOCK)
        if len(block) < _FITS_BLOCK:
            raise ValueError("truncated FITS header")
        for k in range(0, _FITS_BLOCK, 80):
            card = block[k:k + 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                return cards
            if card[8:10] == "= ":
                rest = card[10:].strip()
                if rest.startswith("'"):
                    cards[key] = rest[1:].split("'", 1)[0].strip()
                else:
                    cards[key] = rest.split("/", 1)[0].strip()


def read_fits_map(path: str) -> tuple[np.ndarray, str]:
    """First column of the first binary table of a HEALPix FITS file, with the
    file's ORDERING ('RING' or 'NESTED')."""
    with open(path, "rb") as f:
        primary = _fits_header(f)
        naxis = int(primary.get("NAXIS", 0))
        size = 0
        if naxis > 0:
            size = abs(int(primary.get("BITPIX", 8))) // 8
            for a in range(1, naxis + 1):
                size *= int(primary[f"NAXIS{a}"])
        f.seek(((size + _FITS_BLOCK - 1) // _FITS_BLOCK) * _FITS_BLOCK, os.SEEK_CUR)
        ext = _fits_header(f)
        if ext.get("XTENSION") != "BINTABLE":
            raise ValueError(f"{path}: no binary table extension")
        row_bytes, rows = int(ext["NAXIS1"]), int(ext["NAXIS2"])
        form = ext["TFORM1"].strip()
        rep = int(form[:-1]) if len(form) > 1 else 1
        code = form[-1]
        if code not in _TFORM:
            raise ValueError(f"{path}: unsupported column format {form!r}")
        dt = np.dtype(_TFORM[code])
        raw = np.frombuffer(f.read(row_bytes * rows), dtype=np.uint8).reshape(rows, row_bytes)
        col = raw[:, :rep * dt.itemsize].copy().view(dt).reshape(-1)
        ordering = str(ext.get("ORDERING", "RING")).upper()
    return col.astype(np.float64), ("NESTED" if ordering.startswith("NEST") else "RING")


def read_map(path: str, nest: bool = False) -> np.ndarray:
    """A HEALPix map in the requested ordering (``healpy.read_map`` field 0).

    FITS files are read with their ORDERING keyword and reordered if needed;
    ``.npy`` files are taken to be in the requested ordering already.
    """
    if str(path).endswith(".npy"):
        return np.asarray(np.load(path, allow_pickle=False), dtype=np.float64).reshape(-1)
    m, ordering = read_fits_map(path)
    nside = npix2nside(m.size)
    if ordering == "NESTED" and not nest:
        out = np.empty_like(m)
        out[nest2ring(nside, np.arange(m.size))] = m
        return out
    if ordering == "RING" and nest:
        out = np.empty_like(m)
        out[ring2nest(nside, np.arange(m.size))] = m
        return out
    return m
