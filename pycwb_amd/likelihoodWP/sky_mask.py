"""Sky masks on the Earth-fixed grid (pycwb/modules/likelihoodWP/sky_mask.py:71-207).

Mask types ``UniformAllSky`` (or no mask), ``Patch`` (a cap around a centre),
``Fixed`` (the one grid direction nearest a position) and ``Custom`` (grid
directions whose HEALPix-map value exceeds a threshold), in the frames
``geo`` (default, with a deprecation warning when ``coordsys`` is absent),
``cwb`` and ``icrs``.  Coordinates are parsed with their units
(``pycwb_amd.skymap_coord``).  ICRS masks rotate with the Earth: the centre's
right ascension becomes an Earth-fixed longitude through cWB's sidereal time
at the reference GPS time, so ``sky_valid_indices_for_cluster`` recomputes
them at each cluster's time (segment start + cluster time), as cWB's sky loop
does.  Custom maps are read from FITS (healpy's binary-table layout, any
ORDERING) or ``.npy`` files and indexed with ``healpix.ang2pix``.
"""

from __future__ import annotations

import logging
import warnings
from functools import lru_cache

import numpy as np

from .. import healpix
from ..skymap_coord import (convert_phi_to_ra, convert_ra_to_phi, convert_theta_to_dec,
                            normalize_coordinate_system, parse_angle_quantity,
                            parse_sky_coordinates)

logger = logging.getLogger(__name__)


def sky_mask_requires_event_time(sky_mask_config) -> bool:
    """True for masks that move on the Earth-fixed grid (ICRS)."""
    if not sky_mask_config or sky_mask_config.get("type", "UniformAllSky") == "UniformAllSky":
        return False
    return normalize_coordinate_system(sky_mask_config.get("coordsys", "geo")) == "icrs"


def _mask_coordsys(cfg) -> str:
    if "coordsys" not in cfg:
        warnings.warn("Restricted sky masks without explicit coordsys are deprecated; "
                      "legacy behavior is coordsys='geo'", DeprecationWarning, stacklevel=3)
    return normalize_coordinate_system(cfg.get("coordsys", "geo"))


def _earth_fixed(first, second, coordsys, t_ref, *, context):
    """(longitude, latitude) on the Earth-fixed grid of a parsed pair."""
    if coordsys == "icrs":
        if t_ref is None:
            raise ValueError(f"{context} with coordsys='icrs' requires a GPS reference time")
        return float(convert_ra_to_phi(first, t_ref)), second
    if coordsys == "geo":
        return first % (2.0 * np.pi), second
    return first % (2.0 * np.pi), float(convert_theta_to_dec(second))


def _unwrap_sky_loc(container):
    if "sky_loc" in container:
        warnings.warn("coordinates.sky_loc is deprecated; put semantic coordinate keys directly "
                      "under coordinates", DeprecationWarning, stacklevel=3)
        return container["sky_loc"]
    return container


@lru_cache(maxsize=16)
def _read_healpix_map(path: str, nest: bool):
    return healpix.read_map(path, nest=nest)


def _cos_distance(phi, lat, c_phi, c_lat):
    return np.sin(lat) * np.sin(c_lat) + np.cos(lat) * np.cos(c_lat) * np.cos(phi - c_phi)


def compute_sky_valid_indices(phi_geo_arr, latitude_arr, sky_mask_config, t_ref=None):
    """Valid grid indices (ascending, int64) of a configured mask.

    ``phi_geo_arr`` is cWB's Earth-fixed longitude and ``latitude_arr`` the
    geographic latitude, in radians; ``t_ref`` is the GPS time at which an ICRS
    mask is placed on the grid.
    """
    phi = np.asarray(phi_geo_arr, dtype=np.float64)
    lat = np.asarray(latitude_arr, dtype=np.float64)
    if phi.shape != lat.shape:
        raise ValueError("Sky longitude and latitude arrays must have identical shapes")
    n_sky = len(phi)
    if sky_mask_config is None:
        return np.arange(n_sky, dtype=np.int64)
    kind = sky_mask_config.get("type", "UniformAllSky")
    if kind == "UniformAllSky":
        return np.arange(n_sky, dtype=np.int64)
    coordsys = _mask_coordsys(sky_mask_config)
    logger.info("Applying sky mask type=%s coordsys=%s at GPS=%s", kind, coordsys, t_ref)

    if kind == "Patch":
        patch = sky_mask_config["patch"]
        unit = patch.get("unit", "rad")
        first, second = parse_sky_coordinates(patch["center"], coordsys,
                                              context="sky_mask.patch.center", legacy_unit=unit)
        legacy = "phi" in patch["center"] and "theta" in patch["center"]
        radius = parse_angle_quantity(patch["radius"], name="sky_mask.patch.radius",
                                      legacy_unit=unit, allow_legacy_numeric=legacy)
        if not 0.0 <= radius <= np.pi:
            raise ValueError("sky_mask.patch.radius must be within [0 deg, 180 deg]")
        c_phi, c_lat = _earth_fixed(first, second, coordsys, t_ref, context="sky_mask.patch")
        d = np.arccos(np.clip(_cos_distance(phi, lat, c_phi, c_lat), -1.0, 1.0))
        return np.where(d <= radius)[0].astype(np.int64)

    if kind == "Fixed":
        coords = sky_mask_config["coordinates"]
        first, second = parse_sky_coordinates(_unwrap_sky_loc(coords), coordsys,
                                              context="sky_mask.coordinates",
                                              legacy_unit=coords.get("unit", "rad"))
        f_phi, f_lat = _earth_fixed(first, second, coordsys, t_ref,
                                    context="sky_mask.coordinates")
        if n_sky == 0:
            return np.array([], dtype=np.int64)
        c = np.clip(_cos_distance(phi, lat, f_phi, f_lat), -1.0, 1.0)
        return np.array([int(np.argmax(c))], dtype=np.int64)

    if kind == "Custom":
        if coordsys == "icrs" and t_ref is None:
            raise ValueError("sky_mask Custom with coordsys='icrs' requires a GPS reference time")
        custom = sky_mask_config["custom"]
        ordering = str(custom.get("ordering", "ring")).lower()
        if ordering not in {"ring", "nested"}:
            raise ValueError("sky_mask.custom.ordering must be 'ring' or 'nested'")
        nest = ordering == "nested"
        skymap = _read_healpix_map(str(custom["healpix_map"]), nest)
        map_nside = healpix.npix2nside(len(skymap))
        nside = int(custom.get("nside", map_nside))
        if nside != map_nside or len(skymap) != healpix.nside2npix(nside):
            raise ValueError("sky_mask.custom.nside does not match the loaded HEALPix map")
        lon = convert_phi_to_ra(phi, t_ref) if coordsys == "icrs" else phi
        pix = healpix.ang2pix(nside, np.pi / 2.0 - lat, lon, nest=nest)
        threshold = float(custom.get("threshold", 0.0))
        return np.where(skymap[pix] > threshold)[0].astype(np.int64)

    raise ValueError(f"Unknown sky_mask type: {kind!r}; supported types are UniformAllSky, "
                     "Patch, Fixed, Custom")


def cluster_gps_time(setup, cluster) -> float:
    """Segment start + the cluster's time offset (``cluster_time``, else the
    metadata's ``c_time``)."""
    start = setup.get("segment_start_gps")
    if start is None:
        raise ValueError("An ICRS sky mask requires segment_start_gps so it can be evaluated "
                         "at the cluster time")
    offset = float(getattr(cluster, "cluster_time", 0.0) or 0.0)
    if offset == 0.0:
        meta = getattr(cluster, "cluster_meta", None)
        offset = float(getattr(meta, "c_time", 0.0) or 0.0)
    return float(start) + offset


def sky_valid_indices_for_cluster(setup, cluster, *, use_big_grid: bool = False):
    """Valid indices at the cluster's GPS time (sky_mask.py:179-207): the
    setup's cached set unless the mask is ICRS, which is recomputed."""
    cfg = setup.get("sky_mask_config")
    cached = setup.get("sky_valid_indices_big" if use_big_grid else "sky_valid_indices")
    if not sky_mask_requires_event_time(cfg):
        return cached
    gps = cluster_gps_time(setup, cluster)
    if use_big_grid:
        phi, lat = setup.get("phi_geo_arr_big_cluster"), setup.get("latitude_arr_big_cluster")
    else:
        phi = setup.get("phi_geo_arr", setup.get("ra_arr"))
        lat = setup.get("latitude_arr", setup.get("dec_arr"))
    return compute_sky_valid_indices(phi, lat, cfg, t_ref=gps)


__all__ = ["compute_sky_valid_indices", "sky_mask_requires_event_time",
           "sky_valid_indices_for_cluster"]
