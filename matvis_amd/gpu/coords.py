"""Registered coordinate-rotation methods.

Names follow matvis's registry (core/coords.py:32-38): ``CoordinateRotationERFA``
and ``GPUCoordinateRotationERFA`` (cpu/coords.py:41-230, gpu/coords.py:38-53)
and ``CoordinateRotationAstropy`` (cpu/coords.py:15-38).  Added here:
``CoordinateRotationMatrix`` for caller-supplied per-time matrices.

The ERFA-named methods compute R_t on the host with the model of
:mod:`matvis_amd.astro` (precession, leading nutation, apparent sidereal time);
the device applies it fused with the horizon cull.  They do not yet apply the
per-source BCRS corrections (light deflection, aberration) of the reference's
``_set_bcrs``; ``update_bcrs_every`` is accepted for signature compatibility.
"""

from __future__ import annotations

import numpy as np

from ..astro import rotation_matrices, skycoord_ra_dec, times_jd
from ..core.coords import CoordinateRotation


class CoordinateRotationERFA(CoordinateRotation):
    """Rigid per-time rotation R_t from the host astrometry model."""

    def __init__(self, update_bcrs_every=0.0, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.update_bcrs_every = update_bcrs_every
        self._rots = rotation_matrices(self.times, self.telescope_loc)

    def rotation_matrix(self, t: int) -> np.ndarray:
        return self._rots[t]


class GPUCoordinateRotationERFA(CoordinateRotationERFA):
    """Same rotation; the name the reference's GPU path registers."""

    requires_gpu: bool = True


class CoordinateRotationMatrix(CoordinateRotation):
    """Caller-supplied rotations: ``rotations`` is (ntimes, 3, 3), ECI -> ENU."""

    def __init__(self, rotations, *args, **kwargs):
        super().__init__(*args, **kwargs)
        rot = np.asarray(rotations, dtype=np.float64)
        if rot.shape != (len(times_jd(self.times)), 3, 3):
            raise ValueError("rotations must have shape (ntimes, 3, 3)")
        self._rots = rot

    def rotation_matrix(self, t: int) -> np.ndarray:
        return self._rots[t]


class CoordinateRotationAstropy(CoordinateRotation):
    """Host astropy AltAz per time (cpu/coords.py:15-38); needs astropy."""

    def __init__(self, *args, **kwargs):
        try:
            from astropy.coordinates import AltAz  # noqa: F401
        except ImportError as e:  # pragma: no cover - depends on environment
            raise ImportError(
                "CoordinateRotationAstropy needs astropy; use "
                "coord_method='CoordinateRotationERFA' without it"
            ) from e
        super().__init__(*args, **kwargs)

    def setup(self):
        super().setup()
        from astropy.coordinates import AltAz

        frame = AltAz(obstime=self.times[0], location=self.telescope_loc)
        self.skycoords[0].transform_to(frame)

    def rotate(self, t: int) -> None:
        from astropy.coordinates import AltAz

        frame = AltAz(obstime=self.times[t], location=self.telescope_loc)
        altaz = self.skycoords.transform_to(frame)
        el, az = np.asarray(altaz.alt.rad), np.asarray(altaz.az.rad)
        self.set_topocentric(
            np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az),
                      np.sin(el)])
        )


def source_vectors(skycoords) -> np.ndarray:
    """ECI unit vectors (3, nsrc) of a SkyCoord-like object."""
    from ..coordinates import point_source_crd_eq

    ra, dec = skycoord_ra_dec(skycoords)
    return point_source_crd_eq(ra, dec)
