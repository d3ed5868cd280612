"""Sky masks on the Earth-fixed grid (pycwb/modules/likelihoodWP/sky_mask.py:71-207).

Supported: no mask / ``UniformAllSky``, ``Patch`` and ``Fixed`` in Earth-fixed
(``geo``) coordinates, angles in radians.  ICRS masks need the event-time GMST
and ``Custom`` masks need a HEALPix map reader; both raise NotImplementedError.
"""

from __future__ import annotations

import numpy as np


def compute_sky_valid_indices(phi_geo_arr, latitude_arr, sky_mask_config, t_ref=None):
    phi = np.asarray(phi_geo_arr, dtype=np.float64)
    lat = np.asarray(latitude_arr, dtype=np.float64)
    if phi.shape != lat.shape:
        raise ValueError("Sky longitude and latitude arrays must have identical shapes")
    n_sky = len(phi)
    if not sky_mask_config or sky_mask_config.get("type", "UniformAllSky") == "UniformAllSky":
        return np.arange(n_sky, dtype=np.int64)
    kind = sky_mask_config["type"]
    coordsys = str(sky_mask_config.get("coordsys", "geo")).lower()
    if coordsys not in ("geo", "earth", "earth_fixed"):
        raise NotImplementedError(f"sky mask coordsys {coordsys!r} is not supported here")
    if kind == "Patch":
        patch = sky_mask_config["patch"]
        c = patch["center"]
        c_phi, c_lat = float(c["phi"]) % (2 * np.pi), float(c.get("latitude", c.get("dec", 0.0)))
        radius = float(patch["radius"])
        if not 0.0 <= radius <= np.pi:
            raise ValueError("sky_mask.patch.radius must be within [0 deg, 180 deg]")
        cos_d = np.sin(lat) * np.sin(c_lat) + np.cos(lat) * np.cos(c_lat) * np.cos(phi - c_phi)
        return np.where(np.arccos(np.clip(cos_d, -1.0, 1.0)) <= radius)[0].astype(np.int64)
    if kind == "Fixed":
        c = sky_mask_config["coordinates"]
        f_phi, f_lat = float(c["phi"]) % (2 * np.pi), float(c.get("latitude", c.get("dec", 0.0)))
        if n_sky == 0:
            return np.array([], dtype=np.int64)
        cos_d = np.sin(lat) * np.sin(f_lat) + np.cos(lat) * np.cos(f_lat) * np.cos(phi - f_phi)
        return np.array([int(np.argmax(np.clip(cos_d, -1.0, 1.0)))], dtype=np.int64)
    if kind == "Custom":
        raise NotImplementedError("Custom HEALPix-map sky masks are not supported here")
    raise ValueError(f"Unknown sky_mask type: {kind!r}; supported types are UniformAllSky, "
                     "Patch, Fixed, Custom")


def sky_valid_indices_for_cluster(setup, cluster, *, use_big_grid=False):
    key = "sky_valid_indices_big" if use_big_grid else "sky_valid_indices"
    return setup.get(key)
