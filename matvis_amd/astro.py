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

:func:`erfa_context` gives the per-time astrometry context of the ERFA-style
rotation (the fields of ERFA's ``eraApco13`` that matvis's ERFA path uses).
From 2015 to 2045 the date-dependent inputs -- the CIP coordinates x, y and
CIO locator s of the IAU 2006/2000A bias-precession-nutation, the Earth's
heliocentric position and barycentric velocity -- are interpolated (cubic,
daily nodes) from ``data/astrometry_2015_2045.npz``, written with ERFA by
``data/make_astrometry_table.py``; the observer's geocentric position and
velocity (WGS84, IAU 2000 Earth rotation angle, UT1 = UTC, no polar motion)
and the GCRS -> CIRS matrix from (x, y, s) are formed here.  Outside that
range the context falls back to the low-precision solar series of the
Astronomical Almanac and the R_t model above (0.05-0.09 arcsec from ERFA),
marked ``approx``.
"""

from __future__ import annotations

import math

import numpy as np

ARCSEC = math.pi / (180.0 * 3600.0)

#: measured distance of the low-precision series context's apparent
#: directions from ERFA's ``eraAtco13`` (arcsec; 48 sources x 5 dates at the
#: HERA site); the table-based context (2015-2045) is within 0.0031-0.0044
BUILTIN_CONTEXT_ERROR_ARCSEC = (0.05, 0.09)


class AstrometryAccuracyWarning(UserWarning):
    """The low-precision series context is in use (a date outside the
    astrometry table's 2015-2045): apparent directions are 0.05-0.09 arcsec
    from ERFA's, not within the reference's 10 mas."""
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


# ------------------------------------------------------------------ ERFA-style context

#: speed of light in au per day (IAU 2012 au, c = 299792458 m/s)
C_AU_PER_DAY = 299792458.0 * 86400.0 / 149597870700.0
#: Earth rotation angle rate, rad per UT1 day (IAU 2000)
ERA_RATE = 2.0 * math.pi * 1.00273781191135448
WGS84_A, WGS84_F = 6378137.0, 1.0 / 298.257223563


def earth_rotation_angle(jd_ut1: float) -> float:
    """IAU 2000 Earth rotation angle (rad)."""
    d = jd_ut1 - J2000
    f = (jd_ut1 % 1.0) + (J2000 % 1.0)  # fraction of day, split for precision
    return (2.0 * math.pi * (f + 0.7790572732640 + 0.00273781191135448 * d)) \
        % (2.0 * math.pi)


def sun_geocentric(jd_tt: float):
    """Geocentric position (au) and velocity (au/day) of the Sun on the mean
    equator and equinox of date, from the Astronomical Almanac's
    low-precision series (about 0.01 deg in longitude, 1950-2050):
    L = 280.460 + 0.9856474 n, g = 357.528 + 0.9856003 n (deg),
    lambda = L + 1.915 sin g + 0.020 sin 2g (+20.5" to undo the annual
    aberration folded into L), R = 1.00014 - 0.01671 cos g - 0.00014 cos 2g,
    obliquity 23.439 - 4e-7 n; n = days from J2000."""
    n = jd_tt - J2000
    deg = math.pi / 180.0
    L = (280.460 + 0.9856474 * n) * deg
    g = (357.528 + 0.9856003 * n) * deg
    dg = 0.9856003 * deg  # rad/day
    lam = L + (1.915 * math.sin(g) + 0.020 * math.sin(2 * g)) * deg \
        + 20.496 * ARCSEC
    dlam = 0.9856474 * deg + (1.915 * math.cos(g) + 0.040 * math.cos(2 * g)) \
        * deg * dg
    r = 1.00014 - 0.01671 * math.cos(g) - 0.00014 * math.cos(2 * g)
    dr = (0.01671 * math.sin(g) + 0.00028 * math.sin(2 * g)) * dg
    eps = (23.439 - 0.0000004 * n) * deg
    ce, se = math.cos(eps), math.sin(eps)
    u = np.array([math.cos(lam), ce * math.sin(lam), se * math.sin(lam)])
    du = np.array([-math.sin(lam), ce * math.cos(lam), se * math.cos(lam)])
    return r * u, dr * u + r * dlam * du


# ------------------------------------------------------------ ERFA-table path

_TABLE = None
_TABLE_PATH = None


def _table():
    """The daily astrometry table (loaded once), or None if absent."""
    global _TABLE, _TABLE_PATH
    if _TABLE is None:
        from pathlib import Path

        path = Path(__file__).with_name("data") / "astrometry_2015_2045.npz"
        if not path.exists():
            _TABLE = {}
            return None
        with np.load(path) as d:
            _TABLE = {k: np.asarray(d[k]) for k in ("x", "y", "s", "eh",
                                                     "vb")}
            _TABLE["jd0"] = float(d["jd_utc0"])
        _TABLE_PATH = path
    return _TABLE or None


def _interp_daily(values: np.ndarray, jd0: float, jd: float):
    """4-point Lagrange interpolation of daily samples at jd (needs a node
    on each side of the two around jd); None outside the table."""
    t = jd - jd0
    i = int(math.floor(t))
    if i < 1 or i + 2 >= len(values):
        return None
    u = t - i  # in [0, 1): nodes i - 1, i, i + 1, i + 2 at -1, 0, 1, 2
    w = np.array([-u * (u - 1.0) * (u - 2.0) / 6.0,
                  (u + 1.0) * (u - 1.0) * (u - 2.0) / 2.0,
                  -(u + 1.0) * u * (u - 2.0) / 2.0,
                  (u + 1.0) * u * (u - 1.0) / 6.0])
    return np.tensordot(w, values[i - 1:i + 3], axes=(0, 0))


def c2i_from_xys(x: float, y: float, s: float) -> np.ndarray:
    """GCRS -> CIRS matrix from the CIP coordinates and the CIO locator
    (IERS Conventions 2010, eq. 5.10: R3(-(E + s)) R2(d) R3(E))."""
    r2 = x * x + y * y
    e = math.atan2(y, x) if r2 > 0.0 else 0.0
    d = math.atan(math.sqrt(r2 / (1.0 - r2)))
    return _r3(-(e + s)) @ _r2(d) @ _r3(e)


def observer_geocentric(lat: float, lon: float, height: float,
                        era: float):
    """Position (m) and velocity (m/s) of a WGS84 site in the frame that the
    Earth rotation angle era rotates (no polar motion)."""
    e2 = WGS84_F * (2.0 - WGS84_F)
    nrad = WGS84_A / math.sqrt(1.0 - e2 * math.sin(lat) ** 2)
    rxy = (nrad + height) * math.cos(lat)
    z = (nrad * (1.0 - e2) + height) * math.sin(lat)
    x0, y0 = rxy * math.cos(lon), rxy * math.sin(lon)
    c, s = math.cos(era), math.sin(era)
    pos = np.array([c * x0 - s * y0, s * x0 + c * y0, z])
    om = ERA_RATE / 86400.0
    vel = om * np.array([-s * x0 - c * y0, c * x0 - s * y0, 0.0])
    return pos, vel


AU_M = 149597870700.0


def _table_context(jd_utc, lat, lon, height):
    tab = _table()
    if tab is None:
        return None
    vals = [_interp_daily(tab[k], tab["jd0"], jd_utc)
            for k in ("x", "y", "s", "eh", "vb")]
    if any(v is None for v in vals):
        return None
    x, y, s_cio, eh_earth, vb_earth = vals
    bpn = c2i_from_xys(float(x), float(y), float(s_cio))
    era = earth_rotation_angle(jd_utc)
    pos, vel = observer_geocentric(lat, lon, height, era)
    # observer in GCRS (bpn takes GCRS to CIRS; the ERA frame is CIRS)
    pos_g, vel_g = bpn.T @ pos, bpn.T @ vel
    ph = eh_earth + pos_g / AU_M
    em = float(np.linalg.norm(ph))
    v = (vb_earth + vel_g * 86400.0 / AU_M) / C_AU_PER_DAY
    return {
        "eh": ph / em, "em": em, "v": v, "bm1": math.sqrt(1.0 - float(v @ v)),
        "bpn": bpn, "eral": (era + lon) % (2.0 * math.pi), "xpl": 0.0,
        "ypl": 0.0, "sphi": math.sin(lat), "cphi": math.cos(lat),
        "approx": False,
    }


def erfa_context(jd_utc: float, lat: float, lon: float,
                 height: float = 0.0) -> dict:
    """Astrometry context at one UTC Julian date for a geodetic site
    (radians, metres): ``eh`` (unit vector Sun -> observer, GCRS), ``em``
    (au), ``v`` (observer barycentric velocity / c, GCRS), ``bm1``,
    ``bpn`` (GCRS -> CIRS), ``eral`` (local Earth rotation angle), ``xpl =
    ypl = 0`` (no polar motion), ``sphi``, ``cphi``, and ``approx`` (True
    outside the table's 2015-2045, where the low-precision series gives the
    context).  UT1 = UTC."""
    ctx = _table_context(jd_utc, lat, lon, height)
    if ctx is not None:
        return ctx
    ctx = erfa_context_series(jd_utc, lat, lon, height)
    ctx["approx"] = True
    return ctx


def erfa_context_series(jd_utc: float, lat: float, lon: float,
                        height: float = 0.0) -> dict:
    """The context of :func:`erfa_context` from the low-precision series
    (Almanac solar series, IAU 1976 precession, four IAU 1980 nutation
    terms): 0.05-0.09 arcsec from ERFA, at any date; its ``bpn`` takes GCRS
    to the frame ``eral`` rotates, so that with light deflection and
    aberration left out it reproduces R_t."""
    jd_tt = jd_utc + TT_MINUS_UTC_S / 86400.0
    prec = precession_iau1976(jd_tt)
    dpsi, deps, eps0 = nutation_iau1980_leading(jd_tt)
    nut = _r1(-(eps0 + deps)) @ _r3(-dpsi) @ _r1(eps0)
    gast = gmst_iau1982(jd_utc) + dpsi * math.cos(eps0 + deps)
    era = earth_rotation_angle(jd_utc)
    bpn = _r3(gast - era) @ nut @ prec  # R3(era + lon) bpn = R3(last) N P
    eral = (era + lon) % (2.0 * math.pi)
    # Earth about the Sun (mean equator of date -> GCRS by the precession)
    s_pos, s_vel = sun_geocentric(jd_tt)
    s_pos, s_vel = prec.T @ s_pos, prec.T @ s_vel
    em = float(np.linalg.norm(s_pos))
    eh = -s_pos / em
    v = -s_vel / C_AU_PER_DAY
    # the site's rotation about the Earth's axis, in the rotating frame of
    # eral, back to GCRS
    e2 = WGS84_F * (2.0 - WGS84_F)
    nrad = WGS84_A / math.sqrt(1.0 - e2 * math.sin(lat) ** 2)
    rxy = (nrad + height) * math.cos(lat)
    omega = ERA_RATE / 86400.0  # rad/s
    v_site = np.array([-rxy * math.sin(eral), rxy * math.cos(eral), 0.0]) \
        * omega / 299792458.0
    v = v + bpn.T @ v_site
    return {
        "eh": eh, "em": em, "v": v, "bm1": math.sqrt(1.0 - float(v @ v)),
        "bpn": bpn, "eral": eral, "xpl": 0.0, "ypl": 0.0,
        "sphi": math.sin(lat), "cphi": math.cos(lat),
    }
