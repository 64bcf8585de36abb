"""Minimal time / location / sky-position types and the per-time rotation R_t.

matvis's API takes astropy ``Time``, ``EarthLocation`` and ``SkyCoord`` objects.
This module accepts those by duck typing (``.jd``; ``.lat``/``.lon`` with
``.rad``; ``.ra``/``.dec`` with ``.rad``) and provides small stand-ins with the
same attributes for environments without astropy.

:func:`eci_to_enu_rotation` builds the 3x3 matrix that takes ICRS (J2000)
unit vectors to topocentric ENU at one time: mean precession (IAU 1976),
the leading nutation terms (IAU 1980, four largest), apparent sidereal time
(GMST IAU 1982 plus the equation of the equinoxes) and the ENU rotation of
coordinates.py:91-128.  It neglects annual aberration (up to 20.5"), light
deflection, polar motion and UT1-UTC, so it agrees with astropy's AltAz to
about 20-30 arcsec, not to the 10 mas of the reference's ERFA path.  The same
R_t feeds the device kernels and the host oracle, so parity of the per-time
computation does not depend on this accuracy.
"""

from __future__ import annotations

import math

import numpy as np

ARCSEC = math.pi / (180.0 * 3600.0)
J2000 = 2451545.0
TT_MINUS_UTC_S = 69.184  # TAI-UTC (37 s) + 32.184 s, valid from 2017


class Angle:
    """Array of angles with astropy-like ``.rad`` / ``.deg`` accessors."""

    def __init__(self, value, unit: str = "rad"):
        v = np.asarray(value, dtype=np.float64)
        self.rad = v if unit == "rad" else np.deg2rad(v)

    @property
    def deg(self):
        return np.rad2deg(self.rad)

    def __len__(self):
        return self.rad.size

    def __getitem__(self, i):
        return Angle(self.rad[i])


class Time:
    """Julian dates (UTC), with the parts of astropy's ``Time`` matvis uses."""

    def __init__(self, jd, format: str = "jd"):
        if format != "jd":
            raise ValueError("only format='jd' is supported")
        self.jd = np.atleast_1d(np.asarray(jd, dtype=np.float64))

    @property
    def ndim(self):
        return self.jd.ndim

    @property
    def size(self):
        return self.jd.size

    def __len__(self):
        return self.jd.size

    def __getitem__(self, i):
        return Time(self.jd[i])


class EarthLocation:
    """Geodetic location; ``from_geodetic(lon_deg, lat_deg, height_m)``."""

    def __init__(self, lat_rad: float, lon_rad: float, height: float = 0.0):
        self.lat = Angle(lat_rad)
        self.lon = Angle(lon_rad)
        self.height = float(height)

    @classmethod
    def from_geodetic(cls, lon, lat, height=0.0):
        return cls(np.deg2rad(lat), np.deg2rad(lon), height)


class SkyCoord:
    """Source positions (ICRS ra/dec in radians)."""

    def __init__(self, ra, dec, frame: str = "icrs"):
        self.ra = ra if hasattr(ra, "rad") else Angle(ra)
        self.dec = dec if hasattr(dec, "rad") else Angle(dec)
        self.frame = frame

    def __len__(self):
        return len(self.ra)


def _as_rad(x) -> np.ndarray:
    """Radians from an astropy Quantity/Angle, our Angle, or a bare float."""
    if hasattr(x, "rad"):
        return np.asarray(x.rad, dtype=np.float64)
    if hasattr(x, "to_value"):
        return np.asarray(x.to_value("rad"), dtype=np.float64)
    return np.asarray(x, dtype=np.float64)


def times_jd(times) -> np.ndarray:
    """UTC Julian dates of a Time-like object or a plain array."""
    if hasattr(times, "jd"):
        return np.atleast_1d(np.asarray(times.jd, dtype=np.float64))
    return np.atleast_1d(np.asarray(times, dtype=np.float64))


def location_lat_lon(loc) -> tuple[float, float]:
    """(latitude, east longitude) in radians of an EarthLocation-like object."""
    return float(_as_rad(loc.lat)), float(_as_rad(loc.lon))


def skycoord_ra_dec(sc) -> tuple[np.ndarray, np.ndarray]:
    """(ra, dec) in radians of a SkyCoord-like object."""
    return (
        np.atleast_1d(_as_rad(sc.ra)).astype(np.float64),
        np.atleast_1d(_as_rad(sc.dec)).astype(np.float64),
    )


# ------------------------------------------------------------------ rotation


def _r1(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def _r2(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])


def _r3(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


def gmst_iau1982(jd_ut1: float) -> float:
    """Greenwich mean sidereal time (rad), IAU 1982 expression."""
    t = (jd_ut1 - J2000) / 36525.0
    # seconds of time at the instant: 876600 h per Julian century of solar
    # rotation plus the secular terms; 67310.54841 s = GMST at J2000 (12h UT1)
    gmst_s = (
        67310.54841
        + (876600.0 * 3600.0 + 8640184.812866) * t
        + 0.093104 * t * t
        - 6.2e-6 * t * t * t
    )
    return (gmst_s % 86400.0) / 86400.0 * 2.0 * math.pi


def precession_iau1976(jd_tt: float) -> np.ndarray:
    """Mean-equator precession matrix J2000 -> date (Lieske et al. 1977)."""
    t = (jd_tt - J2000) / 36525.0
    zeta = (2306.2181 * t + 0.30188 * t**2 + 0.017998 * t**3) * ARCSEC
    z = (2306.2181 * t + 1.09468 * t**2 + 0.018203 * t**3) * ARCSEC
    theta = (2004.3109 * t - 0.42665 * t**2 - 0.041833 * t**3) * ARCSEC
    return _r3(-z) @ _r2(theta) @ _r3(-zeta)


def nutation_iau1980_leading(jd_tt: float) -> tuple[float, float, float]:
    """(dpsi, deps, eps0) in rad from the four largest IAU 1980 terms."""
    t = (jd_tt - J2000) / 36525.0
    omega = math.radians(125.04452 - 1934.136261 * t)
    l_sun = math.radians(280.4665 + 36000.7698 * t)
    l_moon = math.radians(218.3165 + 481267.8813 * t)
    dpsi = (
        -17.20 * math.sin(omega)
        - 1.32 * math.sin(2 * l_sun)
        - 0.23 * math.sin(2 * l_moon)
        + 0.21 * math.sin(2 * omega)
    ) * ARCSEC
    deps = (
        9.20 * math.cos(omega)
        + 0.57 * math.cos(2 * l_sun)
        + 0.10 * math.cos(2 * l_moon)
        - 0.09 * math.cos(2 * omega)
    ) * ARCSEC
    eps0 = (
        84381.448 - 46.8150 * t - 0.00059 * t**2 + 0.001813 * t**3
    ) * ARCSEC
    return dpsi, deps, eps0


def eci_to_enu_matrix(ha, lat):
    """3x3 rotation of the reference's coordinates.py:91-128."""
    return np.array(
        [
            [-math.sin(ha), math.cos(ha), 0.0],
            [-math.sin(lat) * math.cos(ha), -math.sin(lat) * math.sin(ha),
             math.cos(lat)],
            [math.cos(lat) * math.cos(ha), math.cos(lat) * math.sin(ha),
             math.sin(lat)],
        ]
    )


def eci_to_enu_rotation(jd_utc: float, lat: float, lon: float) -> np.ndarray:
    """R_t: ICRS unit vector -> topocentric ENU at one UTC Julian date."""
    jd_tt = jd_utc + TT_MINUS_UTC_S / 86400.0
    prec = precession_iau1976(jd_tt)
    dpsi, deps, eps0 = nutation_iau1980_leading(jd_tt)
    nut = _r1(-(eps0 + deps)) @ _r3(-dpsi) @ _r1(eps0)
    last = gmst_iau1982(jd_utc) + dpsi * math.cos(eps0 + deps) + lon
    return eci_to_enu_matrix(last, lat) @ nut @ prec


def rotation_matrices(times, telescope_loc) -> np.ndarray:
    """R_t for every time, shape (ntimes, 3, 3)."""
    lat, lon = location_lat_lon(telescope_loc)
    return np.stack(
        [eci_to_enu_rotation(jd, lat, lon) for jd in times_jd(times)]
    )
