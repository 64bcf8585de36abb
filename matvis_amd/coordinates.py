"""Host-side coordinate helpers (numpy), as matvis's ``coordinates.py`` has.

These run on the host for set-up only (source unit vectors once, az/za for
host-side checks); the per-time work runs in the HIP kernels.
"""

from __future__ import annotations

import numpy as np


def point_source_crd_eq(ra, dec):
    """ECI unit vectors, shape (3, nsrc) (coordinates.py:174-200)."""
    ra = np.asarray(ra, dtype=np.float64)
    dec = np.asarray(dec, dtype=np.float64)
    return np.asarray(
        [np.cos(ra) * np.cos(dec), np.cos(dec) * np.sin(ra), np.sin(dec)]
    )


def enu_to_az_za(enu_e, enu_n, orientation="astropy", periodic_azimuth=True):
    """Direction cosines -> (az, za) (coordinates.py:31-88)."""
    assert orientation in ["astropy", "uvbeam"], (
        "orientation must be either 'astropy' or 'uvbeam'"
    )
    enu_e = np.asarray(enu_e, dtype=np.float64)
    enu_n = np.asarray(enu_n, dtype=np.float64)
    lsqr = enu_n**2.0 + enu_e**2.0
    mask = lsqr < 1
    zeta = np.zeros_like(lsqr)
    zeta[mask] = np.sqrt(1 - lsqr[mask])
    az = np.arctan2(enu_e, enu_n)
    za = 0.5 * np.pi - np.arcsin(zeta)
    if orientation == "uvbeam":
        az = 0.5 * np.pi - az
    if periodic_azimuth:
        az = np.mod(az, 2 * np.pi)
    return az, za


def eci_to_enu_matrix(ha, lat):
    """3x3 ECI -> ENU rotation (coordinates.py:91-128)."""
    return np.array(
        [
            [-np.sin(ha), np.cos(ha), 0.0 * ha],
            [-np.sin(lat) * np.cos(ha), -np.sin(lat) * np.sin(ha), np.cos(lat)],
            [np.cos(lat) * np.cos(ha), np.cos(lat) * np.sin(ha), np.sin(lat)],
        ]
    )


def enu_to_eci_matrix(ha, lat):
    """3x3 ENU -> ECI rotation, the inverse of :func:`eci_to_enu_matrix`
    (coordinates.py:131-160)."""
    return np.array(
        [
            [-np.sin(ha), -np.cos(ha) * np.sin(lat), np.cos(ha) * np.cos(lat)],
            [np.cos(ha), -np.sin(ha) * np.sin(lat), np.sin(ha) * np.cos(lat)],
            [0.0 * ha, np.cos(lat), np.sin(lat)],
        ]
    )


def altaz_to_enu(el, az):
    """(elevation, azimuth east of north) -> ENU (coordinates.py:163-171)."""
    el = np.asarray(el, dtype=np.float64)
    az = np.asarray(az, dtype=np.float64)
    return np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)])
