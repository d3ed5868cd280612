"""Sky masks, coordinate parsing and the HEALPix pixelisation (no GPU).

Cases follow the reference's sky-mask tests
(pycwb/modules/likelihoodWP/tests/test_compute_sky_valid_indices.py): all-sky
and empty configs, Patch caps in radians and degrees, zero and full radius,
Fixed nearest directions, ICRS masks placed with cWB's sidereal time and
recomputed at the cluster time, the cWB co-latitude frame, Custom HEALPix maps.
HEALPix itself is pinned by its definition (pixel centres on the published
rings, round trips, equal areas) and healpy's value for pixel 0 of nside 1.
"""

import math
from types import SimpleNamespace

import numpy as np
import pytest

from pycwb_amd import healpix, skymap_coord
from pycwb_amd.likelihoodWP.sky_mask import (compute_sky_valid_indices,
                                             sky_mask_requires_event_time,
                                             sky_valid_indices_for_cluster)

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")


def fib_sky(n):
    k = np.arange(n, dtype=np.float64)
    lat = np.arcsin(np.clip(1.0 - 2.0 * (k + 0.5) / n, -1.0, 1.0))
    lon = (np.pi * (3.0 - np.sqrt(5.0)) * k) % (2.0 * np.pi)
    return lon, lat


def sep(lon1, lat1, lon2, lat2):
    c = np.sin(lat1) * np.sin(lat2) + np.cos(lat1) * np.cos(lat2) * np.cos(lon1 - lon2)
    return np.arccos(np.clip(c, -1.0, 1.0))


# ---- all-sky and empty ----

@pytest.mark.parametrize("cfg", [None, {"type": "UniformAllSky"}])
def test_all_sky(cfg):
    lon, lat = fib_sky(77)
    out = compute_sky_valid_indices(lon, lat, cfg)
    assert out.dtype == np.int64 and np.array_equal(out, np.arange(77))
    assert len(compute_sky_valid_indices(np.array([]), np.array([]), cfg)) == 0


def test_shape_mismatch_raises():
    with pytest.raises(ValueError):
        compute_sky_valid_indices(np.zeros(3), np.zeros(4), None)


def test_unknown_type_raises():
    lon, lat = fib_sky(8)
    with pytest.raises(ValueError):
        compute_sky_valid_indices(lon, lat, {"type": "Galaxy", "coordsys": "geo"})


# ---- Patch ----

def _patch(phi, theta, radius, unit):
    return {"type": "Patch", "patch": {"center": {"phi": phi, "theta": theta},
                                       "radius": radius, "unit": unit}}


def test_patch_is_the_cap_and_legacy_theta_is_latitude():
    """Legacy phi/theta in the default geo frame: phi = longitude, theta =
    latitude; the mask is exactly the directions within the radius."""
    lon, lat = fib_sky(300)
    out = compute_sky_valid_indices(lon, lat, _patch(1.0, 0.3, 0.5, "rad"))
    inside = np.where(sep(lon, lat, 1.0, 0.3) <= 0.5)[0]
    assert np.array_equal(out, inside) and len(out) > 5
    assert np.all(np.diff(out) > 0)


def test_patch_degrees_equal_radians():
    lon, lat = fib_sky(300)
    a = compute_sky_valid_indices(lon, lat, _patch(1.0, 0.3, 0.5, "rad"))
    b = compute_sky_valid_indices(lon, lat, _patch(np.degrees(1.0), np.degrees(0.3),
                                                   np.degrees(0.5), "deg"))
    assert np.array_equal(a, b)


def test_patch_unit_strings():
    lon, lat = fib_sky(300)
    a = compute_sky_valid_indices(lon, lat, _patch(1.0, 0.3, 0.5, "rad"))
    cfg = {"type": "Patch", "coordsys": "geo",
           "patch": {"center": {"longitude": f"{np.degrees(1.0)} deg",
                                "latitude": f"{0.3} rad"},
                     "radius": f"{np.degrees(0.5) * 60} arcmin"}}
    assert np.array_equal(compute_sky_valid_indices(lon, lat, cfg), a)


def test_patch_zero_radius_is_the_exact_direction():
    lon, lat = fib_sky(300)
    out = compute_sky_valid_indices(lon, lat, _patch(float(lon[5]), float(lat[5]), 0.0, "rad"))
    assert out.tolist() == [5]


def test_patch_full_radius_and_bounds():
    lon, lat = fib_sky(300)
    assert len(compute_sky_valid_indices(lon, lat, _patch(1.0, 0.3, np.pi, "rad"))) == 300
    with pytest.raises(ValueError):
        compute_sky_valid_indices(lon, lat, _patch(1.0, 0.3, 4.0, "rad"))
    assert len(compute_sky_valid_indices(np.array([]), np.array([]),
                                         _patch(0.0, 0.0, 1.0, "rad"))) == 0


def test_bare_numbers_need_a_unit():
    lon, lat = fib_sky(30)
    cfg = {"type": "Patch", "coordsys": "geo",
           "patch": {"center": {"longitude": 1.0, "latitude": 0.3}, "radius": "10 deg"}}
    with pytest.raises(ValueError, match="angular unit"):
        compute_sky_valid_indices(lon, lat, cfg)
    cfg["patch"]["center"] = {"longitude": "1 m", "latitude": "0 deg"}
    with pytest.raises(ValueError, match="angular units"):
        compute_sky_valid_indices(lon, lat, cfg)


def test_frame_keys_are_cross_checked():
    lon, lat = fib_sky(30)
    cfg = {"type": "Fixed", "coordsys": "geo", "coordinates": {"ra": "1 deg", "dec": "0 deg"}}
    with pytest.raises(ValueError, match="does not match"):
        compute_sky_valid_indices(lon, lat, cfg)
    cfg["coordinates"] = {"longitude": "1 deg", "latitude": "0 deg", "phi": 1.0}
    with pytest.raises(ValueError, match="mixes"):
        compute_sky_valid_indices(lon, lat, cfg)
    with pytest.raises(ValueError, match="not recognized"):
        compute_sky_valid_indices(lon, lat, {"type": "Fixed", "coordsys": "galactic",
                                             "coordinates": {}})


# ---- Fixed ----

def _fixed(phi, theta, unit="rad"):
    return {"type": "Fixed", "coordinates": {"sky_loc": {"phi": phi, "theta": theta},
                                             "unit": unit}}


def test_fixed_is_the_nearest_direction():
    lon, lat = fib_sky(200)
    out = compute_sky_valid_indices(lon, lat, _fixed(1.0, 0.2))
    assert out.dtype == np.int64 and out.tolist() == [int(np.argmin(sep(lon, lat, 1.0, 0.2)))]
    deg = compute_sky_valid_indices(lon, lat, _fixed(np.degrees(2.0), np.degrees(-0.4), "deg"))
    rad = compute_sky_valid_indices(lon, lat, _fixed(2.0, -0.4))
    assert deg.tolist() == rad.tolist()


def test_fixed_exact_match():
    out = compute_sky_valid_indices(np.array([0.0, 1.0, 2.0]), np.array([0.0, 0.5, -0.5]),
                                    _fixed(1.0, 0.5))
    assert out.tolist() == [1]


def test_cwb_frame_converts_colatitude():
    cfg = {"type": "Fixed", "coordsys": "cwb",
           "coordinates": {"phi_geo": "0 deg", "theta_cwb": "60 deg"}}
    out = compute_sky_valid_indices(np.zeros(2), np.deg2rad([30.0, -30.0]), cfg)
    assert out.tolist() == [0]


# ---- ICRS ----

def test_gmst_matches_the_published_degree_form():
    """cWB's polynomial in seconds equals the IAU 1982 GMST in Meeus's degree
    form, 280.46061837 + 360.98564736629 d + 0.000387933 T^2 - T^3/38710000."""
    for gps in (630763213.0, 1126259162.4, 1261873618.0, 1e9, 1.4e9):
        d = (gps - 630763213.0) / 86400.0
        T = d / 36525.0
        deg = 280.46061837 + 360.98564736629 * d + 0.000387933 * T * T - T ** 3 / 38710000.0
        want = math.radians(deg % 360.0)
        got = skymap_coord.gmst_cwb(gps)
        assert abs((got - want + math.pi) % (2 * math.pi) - math.pi) < 2e-9, gps
    # at J2000.0 the sidereal time is 18h41m50.54841s
    assert skymap_coord.gmst_cwb(630763213.0) == pytest.approx(
        math.radians(67310.54841 / 240.0), abs=1e-12)


def test_icrs_fixed_uses_ra_minus_gmst():
    gps = 1261873618.0
    g = skymap_coord.gmst_cwb(gps)
    cfg = {"type": "Fixed", "coordsys": "icrs",
           "coordinates": {"ra": f"{np.degrees(g)} deg", "dec": "0 deg"}}
    out = compute_sky_valid_indices(np.array([0.0, g]), np.zeros(2), cfg, t_ref=gps)
    assert out.tolist() == [0]
    assert sky_mask_requires_event_time(cfg)
    with pytest.raises(ValueError, match="requires a GPS reference time"):
        compute_sky_valid_indices(np.array([0.0]), np.array([0.0]), cfg)


def test_icrs_mask_is_recomputed_at_cluster_time():
    start, offset = 1261873618.0, 600.0
    ra = skymap_coord.gmst_cwb(start + offset)
    cfg = {"type": "Fixed", "coordsys": "icrs",
           "coordinates": {"ra": f"{np.degrees(ra)} deg", "dec": "0 deg"}}
    phi_at_start = float((ra - skymap_coord.gmst_cwb(start)) % (2 * np.pi))
    setup = {"sky_mask_config": cfg, "segment_start_gps": start,
             "phi_geo_arr": np.array([0.0, phi_at_start]), "latitude_arr": np.zeros(2),
             "sky_valid_indices": np.array([1], dtype=np.int64)}
    out = sky_valid_indices_for_cluster(setup, SimpleNamespace(cluster_time=offset))
    assert out.tolist() == [0]
    # without a time offset the cached set is the answer at segment start
    meta = SimpleNamespace(c_time=0.0)
    out0 = sky_valid_indices_for_cluster(setup, SimpleNamespace(cluster_time=0.0,
                                                                cluster_meta=meta))
    assert out0.tolist() == [1]
    # geo masks keep the cached set
    geo = dict(setup, sky_mask_config=_fixed(0.0, 0.0))
    assert sky_valid_indices_for_cluster(geo, SimpleNamespace(cluster_time=5.0)).tolist() == [1]


# ---- Custom ----

def _write_healpix_fits(path, m, ordering="RING", rows_of=1024):
    """A minimal healpy-style FITS file: empty primary HDU, one BINTABLE with a
    float32 vector column of ``rows_of`` values per row."""
    def header(cards):
        text = "".join(f"{c:<80}" for c in cards + ["END"])
        text += " " * (-len(text) % 2880)
        return text.encode("ascii")

    m = np.asarray(m, dtype=">f4")
    rep = min(rows_of, m.size)
    rows = m.size // rep
    nside = healpix.npix2nside(m.size)
    prim = header(["SIMPLE  =                    T", "BITPIX  =                    8",
                   "NAXIS   =                    0", "EXTEND  =                    T"])
    ext = header(["XTENSION= 'BINTABLE'", "BITPIX  =                    8",
                  "NAXIS   =                    2", f"NAXIS1  = {4 * rep:>20d}",
                  f"NAXIS2  = {rows:>20d}", "PCOUNT  =                    0",
                  "GCOUNT  =                    1", "TFIELDS =                    1",
                  "TTYPE1  = 'TEMPERATURE'", f"TFORM1  = '{rep}E'",
                  "PIXTYPE = 'HEALPIX '", f"ORDERING= '{ordering}'",
                  f"NSIDE   = {nside:>20d}"])
    data = m.tobytes()
    data += b"\0" * (-len(data) % 2880)
    with open(path, "wb") as f:
        f.write(prim + ext + data)


def test_custom_icrs_map(tmp_path):
    gps, nside = 1261873618.0, 8
    lon, lat = np.array([0.0]), np.array([0.0])
    ra = skymap_coord.convert_phi_to_ra(lon, gps)
    good = int(healpix.ang2pix_ring(nside, np.array([np.pi / 2]), ra)[0])
    raw = int(healpix.ang2pix_ring(nside, np.array([np.pi / 2]), lon)[0])
    assert good != raw
    m = np.zeros(healpix.nside2npix(nside))
    m[good] = 1.0
    path = tmp_path / "mask.fits"
    _write_healpix_fits(path, m)
    cfg = {"type": "Custom", "coordsys": "icrs",
           "custom": {"healpix_map": str(path), "nside": nside, "ordering": "ring",
                      "threshold": 0.0}}
    assert compute_sky_valid_indices(lon, lat, cfg, t_ref=gps).tolist() == [0]
    with pytest.raises(ValueError, match="requires a GPS reference time"):
        compute_sky_valid_indices(lon, lat, cfg)
    bad = dict(cfg, custom=dict(cfg["custom"], nside=4))
    with pytest.raises(ValueError, match="nside"):
        compute_sky_valid_indices(lon, lat, bad, t_ref=gps)


def test_custom_map_orderings(tmp_path):
    """A NESTED file read for a RING mask (and the reverse) selects the same
    grid directions as the map written in the requested ordering."""
    nside = 4
    rng = np.random.default_rng(1)
    ring_map = (rng.uniform(size=healpix.nside2npix(nside)) > 0.6).astype(np.float64)
    nest_map = np.empty_like(ring_map)
    nest_map[healpix.ring2nest(nside, np.arange(ring_map.size))] = ring_map
    lon, lat = fib_sky(500)
    want = np.where(ring_map[healpix.ang2pix_ring(nside, np.pi / 2 - lat, lon)] > 0.5)[0]
    for fname, data, file_order in (("r.fits", ring_map, "RING"), ("n.fits", nest_map, "NESTED")):
        p = tmp_path / fname
        _write_healpix_fits(p, data, file_order, rows_of=64)
        for ordering in ("ring", "nested"):
            cfg = {"type": "Custom", "coordsys": "geo",
                   "custom": {"healpix_map": str(p), "ordering": ordering, "threshold": 0.5}}
            assert np.array_equal(compute_sky_valid_indices(lon, lat, cfg), want)
    np.save(tmp_path / "m.npy", ring_map)
    cfg = {"type": "Custom", "coordsys": "geo",
           "custom": {"healpix_map": str(tmp_path / "m.npy"), "threshold": 0.5}}
    assert np.array_equal(compute_sky_valid_indices(lon, lat, cfg), want)


# ---- HEALPix ----

def test_healpix_nside1_known_centres():
    """nside 1: pixels 0-3 on z = 2/3 at phi = pi/4 + k pi/2 (healpy gives
    (0.8410686705679303, 0.7853981633974483) for pixel 0), 4-7 on the equator
    at k pi/2, 8-11 on z = -2/3."""
    th, ph = healpix.pix2ang_ring(1, np.arange(12))
    assert th[0] == pytest.approx(0.8410686705679303, abs=1e-15)
    assert ph[0] == pytest.approx(0.7853981633974483, abs=1e-15)
    np.testing.assert_allclose(np.cos(th), [2 / 3] * 4 + [0] * 4 + [-2 / 3] * 4, atol=1e-15)
    np.testing.assert_allclose(ph[4:8], np.arange(4) * np.pi / 2, atol=1e-15)
    np.testing.assert_allclose(ph[8:], np.pi / 4 + np.arange(4) * np.pi / 2, atol=1e-15)


@pytest.mark.parametrize("nside", [2, 4, 16, 128])
def test_healpix_rings_from_the_definition(nside):
    """Pixel centres built ring by ring from the published equations: polar
    ring i (1..nside-1) has 4i pixels at z = 1 - i^2/(3 nside^2),
    phi = (j - 1/2) pi/(2i); equatorial ring i (nside..3nside) has 4 nside
    pixels at z = 4/3 - 2i/(3 nside), phi = (j - s/2) pi/(2 nside) with
    s = (i - nside + 1) mod 2 -- healpy numbers the s = 0 rings from the pixel
    at phi = 0 (the paper's j = 4 nside), so j runs 0..4nside-1 there."""
    z, phi = [], []
    for i in range(1, nside):
        z += [1 - i * i / (3 * nside * nside)] * (4 * i)
        phi += [(j - 0.5) * np.pi / (2 * i) for j in range(1, 4 * i + 1)]
    for i in range(nside, 3 * nside + 1):
        s = (i - nside + 1) % 2
        z += [4 / 3 - 2 * i / (3 * nside)] * (4 * nside)
        first = 1 if s else 0
        phi += [(j - s / 2) * np.pi / (2 * nside) for j in range(first, first + 4 * nside)]
    for i in range(nside - 1, 0, -1):
        z += [-(1 - i * i / (3 * nside * nside))] * (4 * i)
        phi += [(j - 0.5) * np.pi / (2 * i) for j in range(1, 4 * i + 1)]
    th, ph = healpix.pix2ang_ring(nside, np.arange(12 * nside * nside))
    np.testing.assert_allclose(np.cos(th), z, atol=1e-13)
    np.testing.assert_allclose(ph, phi, atol=1e-12)


@pytest.mark.parametrize("nside", [1, 2, 8, 64])
def test_healpix_round_trips_and_equal_area(nside):
    p = np.arange(healpix.nside2npix(nside))
    for nest in (False, True):
        th, ph = healpix.pix2ang(nside, p, nest=nest)
        assert np.array_equal(healpix.ang2pix(nside, th, ph, nest=nest), p)
    assert np.array_equal(np.sort(healpix.nest2ring(nside, p)), p)
    rng = np.random.default_rng(nside)
    n = 400 * p.size
    z, ph = rng.uniform(-1, 1, n), rng.uniform(0, 2 * np.pi, n)
    for nest in (False, True):
        c = np.bincount(healpix.ang2pix(nside, np.arccos(z), ph, nest=nest), minlength=p.size)
        # equal areas: counts are Poisson around one mean (chi-square within 6 sigma)
        chi2 = float(np.sum((c - c.mean()) ** 2) / c.mean())
        dof = p.size - 1
        assert abs(chi2 - dof) < 6 * math.sqrt(2 * dof) + 6, (nest, chi2, dof)
