"""Sky grid, detector geometry and per-direction tables (host side, float64).

Restates, without healpy or astropy:

* ``_build_sky_directions`` (pycwb/types/detector.py:1220-1240): the HEALPix RING
  grid of order ``k`` (nside = 2**k) in cWB's Earth-fixed frame, returned as
  ``(phi_geo, latitude)`` in radians (``healpix.pix2ang_ring``).
* ``earth_centered_vectors`` (pycwb/types/detector.py:1061-1123): WGS84
  geodetic -> ECEF vertex (the ellipsoid formula behind astropy's
  ``EarthLocation.from_geodetic``) and the arm response tensors built with
  astropy's ``rotation_matrix`` convention.
* ``compute_sky_delay_and_patterns`` (pycwb/types/detector.py:1243-1329): delay
  taps ``ml = clip(rint((R_i - R_ref).n/c * rate), +-td_size)`` and F+/Fx at
  psi = 0 with gha = -phi (GMST cancels in the reference).

These tables are computed once per segment; the GPU scan reads them from HBM.
"""

from __future__ import annotations

import math

import numpy as np

from .detectors import DETECTORS
from .healpix import pix2ang_ring  # noqa: F401  (the grid of build_sky_directions)

C_LIGHT = 299792458.0          # m/s (astropy.constants.c)
WGS84_A = 6378137.0            # m
WGS84_F = 1.0 / 298.257223563


def nside_of(order: int) -> int:
    return 1 << int(order)


def npix_of(order: int) -> int:
    n = nside_of(order)
    return 12 * n * n


def build_sky_directions(n_sky: int, healpix_order: int | None = None):
    """cWB Earth-fixed ``(phi_geo, latitude)`` arrays (detector.py:1220-1240).

    With a positive HEALPix order the RING grid is returned.  Otherwise the
    reference's Fibonacci-sphere fallback of ``n_sky`` points is used.
    """
    if healpix_order is not None and int(healpix_order) > 0:
        nside = nside_of(healpix_order)
        theta, phi = pix2ang_ring(nside, np.arange(12 * nside * nside, dtype=np.int64))
        return phi.astype(np.float64), (np.pi / 2.0) - theta
    n_sky = int(max(1, n_sky))
    idx = np.arange(n_sky, dtype=np.float64)
    z = 1.0 - 2.0 * (idx + 0.5) / float(n_sky)
    phi = (np.pi * (3.0 - np.sqrt(5.0)) * idx) % (2.0 * np.pi)
    return phi.astype(np.float64), np.arcsin(np.clip(z, -1.0, 1.0)).astype(np.float64)


def _rot(angle: float, axis: str) -> np.ndarray:
    """Coordinate-frame rotation with astropy's ``rotation_matrix`` convention."""
    c, s = math.cos(angle), math.sin(angle)
    i = "xyz".index(axis)
    a1, a2 = (i + 1) % 3, (i + 2) % 3
    r = np.zeros((3, 3))
    r[i, i] = 1.0
    r[a1, a1] = c
    r[a1, a2] = s
    r[a2, a1] = -s
    r[a2, a2] = c
    return r


def geodetic_to_ecef(lon: float, lat: float, height: float) -> np.ndarray:
    e2 = WGS84_F * (2.0 - WGS84_F)
    sl = math.sin(lat)
    n = WGS84_A / math.sqrt(1.0 - e2 * sl * sl)
    return np.array([
        (n + height) * math.cos(lat) * math.cos(lon),
        (n + height) * math.cos(lat) * math.sin(lon),
        (n * (1.0 - e2) + height) * sl,
    ])


class Detector:
    """Vertex position and response tensor of an interferometer."""

    def __init__(self, name: str):
        if name not in DETECTORS:
            raise ValueError(f"unknown detector {name!r}")
        d = DETECTORS[name]
        self.name = name
        self.latitude = d["lat"]
        self.longitude = d["lon"]
        self.altitude = d["elevation"]
        yangle, xangle = d["y"]["az"], d["x"]["az"]
        yalt, xalt = d["y"]["alt"], d["x"]["alt"]
        base = np.array([[-1.0, 0, 0], [0, 0, 0], [0, 0, 0]])
        rm2 = _rot(-self.longitude, "z")
        rm1 = _rot(-1.0 * (np.pi / 2.0 - self.latitude), "y")
        resps, vecs = [], []
        for angle, azi in ((yangle, yalt), (xangle, xalt)):
            rm = rm2 @ rm1 @ _rot(angle, "z") @ _rot(-azi, "y")
            resps.append(rm @ base @ rm.T / 2.0)
            vecs.append(rm @ np.array([-1.0, 0, 0]))
        self.response = resps[0] - resps[1]
        self.x_vec = vecs[1]
        self.y_vec = vecs[0]
        self.vertex = geodetic_to_ecef(self.longitude, self.latitude, self.altitude)

    def antenna_pattern_gha(self, gha: np.ndarray, dec: np.ndarray, psi: float = 0.0):
        """F+, Fx for Greenwich hour angle ``gha`` (detector.py:272-378, tensor)."""
        cg, sg = np.cos(gha), np.sin(gha)
        cd, sd = np.cos(dec), np.sin(dec)
        cp, sp = math.cos(psi), math.sin(psi)
        x = np.stack([-cp * sg - sp * cg * sd, -cp * cg + sp * sg * sd, sp * cd])
        y = np.stack([sp * sg - cp * cg * sd, sp * cg + cp * sg * sd, cp * cd])
        dx = self.response @ x
        dy = self.response @ y
        fplus = (x * dx - y * dy).sum(axis=0)
        fcross = (x * dy + y * dx).sum(axis=0)
        return fplus, fcross


def compute_sky_delay_and_patterns(ifos, ref_ifo, sample_rate, td_size, healpix_order=None,
                                   n_sky=None):
    """``(ml, FP, FX)`` for a detector network (detector.py:1243-1329).

    Returns ``ml`` int32 (nIFO, nSky), ``FP``/``FX`` float64 (nIFO, nSky).
    """
    dets = [Detector(i) if isinstance(i, str) else i for i in ifos]
    if not dets:
        raise ValueError("No detectors provided")
    rate = float(sample_rate)
    td_size = int(max(1, td_size))
    if n_sky is None:
        n_sky = npix_of(healpix_order) if healpix_order and int(healpix_order) > 0 else 3072
    phi, lat = build_sky_directions(n_sky, healpix_order)
    ref_idx = 0
    for i, d in enumerate(dets):
        if d.name == ref_ifo:
            ref_idx = i
            break
    cl, sl = np.cos(lat), np.sin(lat)
    n_hat = np.vstack((cl * np.cos(phi), cl * np.sin(phi), sl)).T
    ref = dets[ref_idx].vertex
    n = len(dets)
    ml = np.zeros((n, phi.size), dtype=np.int32)
    FP = np.zeros((n, phi.size))
    FX = np.zeros((n, phi.size))
    for i, d in enumerate(dets):
        shift = n_hat @ (d.vertex - ref) / C_LIGHT
        ml[i] = np.clip(np.rint(shift * rate).astype(np.int32), -td_size, td_size)
        FP[i], FX[i] = d.antenna_pattern_gha(-phi, lat)
    return ml, FP, FX


def max_network_delay(ifos, ref_ifo) -> float:
    """Largest |(R_i - R_ref)| / c over the network, seconds."""
    dets = [Detector(i) for i in ifos]
    ref = next((d for d in dets if d.name == ref_ifo), dets[0]).vertex
    return max(float(np.linalg.norm(d.vertex - ref)) / C_LIGHT for d in dets)


def mirror_side(ifos, phi: np.ndarray, lat: np.ndarray) -> np.ndarray | None:
    """Side of each direction relative to the plane through the first three
    detector vertices (None for fewer than three detectors).  Mirror images
    through that plane share the delays of those three detectors."""
    if len(ifos) < 3:
        return None
    v = [Detector(i).vertex if isinstance(i, str) else i.vertex for i in ifos[:3]]
    nrm = np.cross(v[1] - v[0], v[2] - v[0])
    cl = np.cos(lat)
    nh = np.stack([cl * np.cos(phi), cl * np.sin(phi), np.sin(lat)])
    return (nrm @ nh) > 0


def delay_order(ml: np.ndarray, valid: np.ndarray, side: np.ndarray | None = None) -> np.ndarray:
    """Valid-list positions sorted by delay tuple ``ml[:, valid[k]]`` (then by k).

    Directions that share every delay tap read the same delayed amplitudes, so
    placing them side by side in a wavefront keeps the per-pixel energy test of
    the sky scan uniform across its lanes and the taps a wavefront reads within
    a narrow range.  With four or more detectors the two mirror sheets of the
    first three detectors' plane (``side``, per sky direction) are sorted
    apart: they share the leading delays but not the others, and interleaving
    them would widen the later detectors' tap ranges.
    """
    ml = np.asarray(ml)
    valid = np.asarray(valid, dtype=np.int64)
    keys = [np.arange(valid.size)] + [ml[i, valid] for i in reversed(range(ml.shape[0]))]
    if side is not None and ml.shape[0] >= 4:
        keys.append(np.asarray(side)[valid])
    return np.lexsort(keys).astype(np.int32)
