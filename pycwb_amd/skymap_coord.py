"""Sky-coordinate parsing and frame conversion for sky masks, without astropy.

Same names, argument meanings and errors as pycwb/utils/skymap_coord.py:
angles are radians internally; a coordinate pair is named by its frame
(``ra``/``dec`` for ICRS, ``longitude``/``latitude`` for geographic,
``phi_geo``/``theta_cwb`` for cWB's Earth-fixed frame with co-latitude,
skymap_coord.py:28-32); user values carry their unit (``"120 deg"``); bare
numbers are accepted only on the legacy ``phi``/``theta`` path with a detached
``unit`` (skymap_coord.py:50-160).

Quantity strings are parsed here (a number followed by an angular unit)
instead of by ``astropy.units.Quantity``.  The units recognised are those
astropy treats as angles in sky configurations: rad, mrad, urad, deg, arcmin,
arcsec, mas, uas, hourangle and cycle (with their long names).  A value whose
unit is not an angle raises ``ValueError`` as the reference does.

The sidereal time of ICRS masks is cWB's, ``gmst_cwb`` (skymap_coord.py:281-298,
``cwb-core/skymap.hh::phiRA``): the IAU 1982 mean sidereal time polynomial
(Aoki et al. 1982) in seconds of the Julian centuries since J2000.0
(GPS 630763213), without a UT1 correction.
"""

from __future__ import annotations

import math
import re
import warnings
from collections.abc import Mapping

import numpy as np

TWO_PI = 2.0 * np.pi
EPOCH_J2000_0_GPS = 630763213.0

COORDINATE_KEYS = {
    "icrs": ("ra", "dec"),
    "geo": ("longitude", "latitude"),
    "cwb": ("phi_geo", "theta_cwb"),
}
_ALL_SEMANTIC_KEYS = frozenset(k for pair in COORDINATE_KEYS.values() for k in pair)

_DEG = math.pi / 180.0
_ANGLE_UNITS = {
    "rad": 1.0, "radian": 1.0, "radians": 1.0,
    "mrad": 1e-3, "milliradian": 1e-3, "urad": 1e-6, "microradian": 1e-6,
    "deg": _DEG, "degree": _DEG, "degrees": _DEG, "°": _DEG,
    "arcmin": _DEG / 60.0, "arcminute": _DEG / 60.0, "arcminutes": _DEG / 60.0,
    "arcsec": _DEG / 3600.0, "arcsecond": _DEG / 3600.0, "arcseconds": _DEG / 3600.0,
    "mas": _DEG / 3.6e6, "uas": _DEG / 3.6e9,
    "hourangle": 15.0 * _DEG,
    "cycle": TWO_PI, "cy": TWO_PI,
}
_QUANTITY = re.compile(
    r"^\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)\s*([A-Za-z°_]*)\s*$")


def normalize_coordinate_system(coordinate_system, *, default: str = "icrs") -> str:
    """Validated lower-case frame name (skymap_coord.py:35-44)."""
    value = default if coordinate_system is None else str(coordinate_system).lower()
    if value not in COORDINATE_KEYS:
        supported = ", ".join(sorted(COORDINATE_KEYS))
        raise ValueError(f"Coordinate system {coordinate_system!r} is not recognized; "
                         f"supported values are: {supported}")
    return value


def _unit_factor(unit, name: str) -> float:
    key = str(unit).strip()
    if key in _ANGLE_UNITS:
        return _ANGLE_UNITS[key]
    if key.lower() in _ANGLE_UNITS:
        return _ANGLE_UNITS[key.lower()]
    raise ValueError(f"{name} must have angular units, got {unit}")


def parse_angle_quantity(value, *, name: str, legacy_unit=None,
                         allow_legacy_numeric: bool = False) -> float:
    """One angular value in radians (skymap_coord.py:47-81)."""
    if isinstance(value, str):
        m = _QUANTITY.match(value)
        if m is None:
            raise ValueError(f"{name} must be an angular quantity, got {value!r}")
        if not m.group(2):
            raise ValueError(f"{name} must have angular units, got dimensionless")
        return float(m.group(1)) * _unit_factor(m.group(2), name)
    if allow_legacy_numeric and legacy_unit is not None and np.isscalar(value) \
            and not isinstance(value, (str, bytes)):
        try:
            factor = _unit_factor(legacy_unit, name)
        except ValueError as exc:
            raise ValueError(f"Invalid legacy angular unit {legacy_unit!r} for {name}") from exc
        return float(value) * factor
    raise ValueError(f"{name} must include an angular unit, for example '120 deg' or "
                     "'2.094 rad'; bare numeric angles are ambiguous")


def parse_sky_coordinates(values, coordinate_system: str, *, context: str = "sky coordinates",
                          legacy_unit=None, allow_legacy: bool = True) -> tuple[float, float]:
    """A validated coordinate pair in radians (skymap_coord.py:84-160).

    The key pair must match the frame; the legacy ``phi``/``theta`` pair is
    read in the frame's own order (first = longitude-like, second as the frame
    defines it) with the detached unit, and warns.
    """
    if not isinstance(values, Mapping):
        raise ValueError(f"{context} must be a mapping")
    coordsys = normalize_coordinate_system(coordinate_system)
    expected = COORDINATE_KEYS[coordsys]
    present_semantic = _ALL_SEMANTIC_KEYS.intersection(values)
    present_legacy = {"phi", "theta"}.intersection(values)
    used_legacy = False
    if all(k in values for k in expected):
        unexpected = present_semantic.difference(expected)
        if unexpected or present_legacy:
            raise ValueError(f"{context} mixes coordinate frames: coordsys={coordsys!r} requires "
                             f"{expected}, but also found {sorted(unexpected | present_legacy)}")
        first = parse_angle_quantity(values[expected[0]], name=f"{context}.{expected[0]}")
        second = parse_angle_quantity(values[expected[1]], name=f"{context}.{expected[1]}")
    elif present_semantic:
        raise ValueError(f"{context} does not match coordsys={coordsys!r}: expected keys "
                         f"{expected}, found {sorted(present_semantic)}")
    elif allow_legacy and "phi" in values and "theta" in values:
        used_legacy = True
        if legacy_unit is None:
            raise ValueError(f"Legacy {context} phi/theta values require an explicit angular unit")
        warnings.warn(f"Legacy {context} keys 'phi'/'theta' and detached 'unit' are deprecated; "
                      f"use {expected[0]!r}/{expected[1]!r} with unit-bearing values",
                      DeprecationWarning, stacklevel=2)
        first = parse_angle_quantity(values["phi"], name=f"{context}.phi",
                                     legacy_unit=legacy_unit, allow_legacy_numeric=True)
        second = parse_angle_quantity(values["theta"], name=f"{context}.theta",
                                      legacy_unit=legacy_unit, allow_legacy_numeric=True)
    else:
        raise ValueError(f"{context} with coordsys={coordsys!r} requires keys "
                         f"{expected[0]!r} and {expected[1]!r}")

    if coordsys == "icrs":
        if not used_legacy and not 0.0 <= first <= TWO_PI:
            raise ValueError(f"{context}.ra must be within [0 deg, 360 deg]")
        first %= TWO_PI
        if not -np.pi / 2.0 <= second <= np.pi / 2.0:
            raise ValueError(f"{context}.dec must be within [-90 deg, 90 deg]")
    elif coordsys == "geo":
        if not used_legacy and not -np.pi <= first <= np.pi:
            raise ValueError(f"{context}.longitude must be within [-180 deg, 180 deg]")
        first = (first + np.pi) % TWO_PI - np.pi
        if not -np.pi / 2.0 <= second <= np.pi / 2.0:
            raise ValueError(f"{context}.latitude must be within [-90 deg, 90 deg]")
    else:
        if not used_legacy and not 0.0 <= first <= TWO_PI:
            raise ValueError(f"{context}.phi_geo must be within [0 deg, 360 deg]")
        first %= TWO_PI
        if not 0.0 <= second <= np.pi:
            raise ValueError(f"{context}.theta_cwb must be within [0 deg, 180 deg]")
    return first, second


def gmst_cwb(gps_time):
    """cWB's mean sidereal time in radians (skymap_coord.py:281-298): the IAU
    1982 GMST polynomial in Julian centuries T since J2000.0,
    67310.54841 s + (876600 h + 8640184.812866 s) T + 0.093104 s T^2
    - 6.2e-6 s T^3, reduced modulo one sidereal day."""
    gps = np.asarray(gps_time, dtype=np.float64)
    t = (gps - EPOCH_J2000_0_GPS) / 86400.0 / 36525.0
    seconds = ((-6.2e-6 * t + 0.093104) * t * t + 67310.54841 + 8640184.812866 * t
               + 3155760000.0 * t)
    out = np.deg2rad((360.0 * seconds / 86400.0) % 360.0)
    return float(out) if out.ndim == 0 else out


def convert_phi_to_ra(phi, gps_time, *, gmst_model: str = "cwb"):
    """Earth-fixed longitude -> ICRS right ascension (radians)."""
    return (np.asarray(phi) + _gmst(gps_time, gmst_model)) % TWO_PI


def convert_ra_to_phi(ra, gps_time, *, gmst_model: str = "cwb"):
    """ICRS right ascension -> Earth-fixed longitude (radians)."""
    return (np.asarray(ra) - _gmst(gps_time, gmst_model)) % TWO_PI


def _gmst(gps_time, model: str):
    if str(model).lower() != "cwb":
        raise ValueError("only gmst_model='cwb' is available (the sky masks' model); "
                         "the astropy model needs astropy")
    return gmst_cwb(gps_time)


def convert_theta_to_dec(theta):
    return np.pi / 2.0 - np.asarray(theta)


def convert_dec_to_theta(dec):
    return np.pi / 2.0 - np.asarray(dec)


__all__ = ["COORDINATE_KEYS", "normalize_coordinate_system", "parse_angle_quantity",
           "parse_sky_coordinates", "gmst_cwb", "convert_phi_to_ra", "convert_ra_to_phi",
           "convert_theta_to_dec", "convert_dec_to_theta"]
