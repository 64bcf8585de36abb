"""Beam models accepted by matvis_amd and their preparation for the device.

The reference takes pyuvdata ``UVBeam`` / ``AnalyticBeam`` / ``BeamInterface``
objects.  Those are accepted here by duck typing:

* a *grid beam* exposes ``data_array`` shaped
  ``(Naxes_vec, Nfeeds | Npols, Nfreqs, Naxes2 = n_za, Naxes1 = n_az)``,
  ``axis1_array`` (azimuth, rad), ``axis2_array`` (zenith angle, rad),
  ``freq_array``, ``beam_type`` ("efield" or "power") and
  ``pixel_coordinate_system`` ("az_za") -- the attributes of a pyuvdata UVBeam;
* an *analytic beam* is :class:`GaussianBeam` (this module);
* a ``BeamInterface``-like wrapper exposes ``.beam`` (and ``.beam_type``).

:class:`GridBeam` is a small UVBeam stand-in with those attributes.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

C_LIGHT = 299792458.0


class GridBeam:
    """Regular az/za grid beam with the UVBeam attributes matvis reads.

    ``data_array``: (Naxes_vec, Nfeeds, Nfreqs, Nza, Naz) complex for an
    E-field beam; (1, Npols, Nfreqs, Nza, Naz) real for a power beam.
    """

    pixel_coordinate_system = "az_za"

    def __init__(self, data_array, az_array, za_array, freq_array=None,
                 beam_type: str = "efield", feed_array=("x", "y")):
        self.data_array = np.asarray(data_array)
        self.axis1_array = np.asarray(az_array, dtype=np.float64)
        self.axis2_array = np.asarray(za_array, dtype=np.float64)
        nfreq = self.data_array.shape[2]
        self.freq_array = (np.zeros(nfreq) if freq_array is None
                           else np.atleast_1d(np.asarray(freq_array, float)))
        if beam_type not in ("efield", "power"):
            raise ValueError("beam_type must be 'efield' or 'power'")
        self.beam_type = beam_type
        self.feed_array = list(feed_array)
        if self.data_array.ndim != 5:
            raise ValueError("data_array must have 5 axes "
                             "(Naxes_vec, Nfeeds|Npols, Nfreqs, Nza, Naz)")
        if self.data_array.shape[3:] != (self.axis2_array.size,
                                         self.axis1_array.size):
            raise ValueError("data_array grid does not match az/za arrays")

    @property
    def Naxes_vec(self):
        return self.data_array.shape[0]

    @property
    def Nfeeds(self):
        return self.data_array.shape[1]


@dataclass
class GaussianBeam:
    """Circular zenith-pointed Gaussian beam.

    E-field amplitude ``exp(-za^2 / (2 sigma^2))`` (power = its square), with
    ``sigma`` in radians scaled by ``(freq / reference_frequency) **
    spectral_index`` when both are set.  ``diameter`` (m) sets sigma from an
    Airy-disk match, ``arcsin(2.2150894 * lambda / (pi D)) * 2 / 2.355``:
    the Gaussian E-field's half maximum where the Airy E-field 2 J1(x) / x
    has its own (x = 2.2150894).  Restates pyuvdata's GaussianBeam, which is
    absent here.
    """

    sigma: float | None = None
    diameter: float | None = None
    spectral_index: float = 0.0
    reference_frequency: float | None = None
    beam_type: str = "efield"

    def __post_init__(self):
        if (self.sigma is None) == (self.diameter is None):
            raise ValueError("set exactly one of sigma and diameter")

    def get_sigma(self, freq: float) -> float:
        if self.diameter is not None:
            lam = C_LIGHT / freq
            return float(np.arcsin(2.2150894 * lam / (np.pi * self.diameter))
                         * 2.0 / 2.355)
        s = float(self.sigma)
        if self.reference_frequency and self.spectral_index:
            s *= (freq / self.reference_frequency) ** self.spectral_index
        return s


class BeamInterface:
    """Wrapper carrying a beam and the type it is to be used as."""

    def __init__(self, beam, beam_type: str | None = None):
        self.beam = beam.beam if isinstance(beam, BeamInterface) else beam
        self.beam_type = beam_type or getattr(self.beam, "beam_type", "efield")

    @property
    def _isuvbeam(self) -> bool:
        return is_grid_beam(self.beam)


def unwrap(beam):
    """Return ``(inner beam, beam_type)`` for a beam or a BeamInterface-like."""
    if hasattr(beam, "beam") and not hasattr(beam, "data_array"):
        return beam.beam, getattr(beam, "beam_type", None)
    return beam, getattr(beam, "beam_type", None)


def is_grid_beam(beam) -> bool:
    inner, _ = unwrap(beam)
    return hasattr(inner, "data_array")


def is_analytic_beam(beam) -> bool:
    inner, _ = unwrap(beam)
    return hasattr(inner, "get_sigma")


#: Distance (Hz) within which a requested frequency takes the nearest beam
#: node's data instead of an interpolation (UVBeam.interp's freq_interp_tol
#: default, reached from core/beams.py:82-93).
FREQ_INTERP_TOL = 1.0


def _interp_freq(data: np.ndarray, freqs: np.ndarray, freq: float):
    """data[:, :, f] at one frequency, as ``UVBeam.interp(freq_array=[freq])``
    with its defaults (core/beams.py:82-93):

    * within ``FREQ_INTERP_TOL`` of a node: that node's data;
    * otherwise a cubic spline through all nodes along the frequency axis
      (``freq_interp_kind="cubic"``: scipy's ``interp1d``, not-a-knot ends,
      real and imaginary parts separately), which needs at least 4 nodes and
      the frequency inside their range; both conditions raise ValueError as
      there.

    A beam with a single frequency is used unchanged at every frequency (a
    frequency-independent grid beam); the reference raises for it instead
    ("Only one frequency in UVBeam so cannot interpolate")."""
    if data.shape[2] == 1:
        return data[:, :, 0]
    freqs = np.asarray(freqs, dtype=np.float64)
    dist = np.abs(freqs - freq)
    near = int(np.argmin(dist))
    if dist[near] < FREQ_INTERP_TOL:
        return data[:, :, near]
    if not freqs.min() <= freq <= freqs.max():
        raise ValueError(f"frequency {freq} Hz outside the beam's range "
                         f"[{freqs.min()}, {freqs.max()}]")
    if freqs.size < 4:
        raise ValueError(
            f"frequency {freq} Hz is between beam frequencies and the cubic "
            f"frequency interpolation needs at least 4 of them "
            f"(the beam has {freqs.size})")
    from scipy.interpolate import interp1d

    def spline(x):
        return interp1d(freqs, x, kind="cubic", axis=2)(freq)

    if np.iscomplexobj(data):
        return spline(data.real) + 1j * spline(data.imag)
    return spline(data)


def grid_planes(beam, freq: float, polarized: bool, beam_type=None):
    """Planes of a grid beam as the device kernels take them.

    Returns ``(planes, daz, dza, azmin, is_power)`` with ``planes`` shaped
    ``(nax, nfeed, nza, naz')`` -- the ``(nbeam, nax, nfeed, nza, naz)`` layout
    of gpu/beams.py:143-145 for one beam.

    * unpolarized: an E-field beam becomes the xx power beam
      ``|E[0, x]|^2 + |E[1, x]|^2`` (core/beams.py:15-32); a power beam keeps
      its first polarization;
    * frequency, after the power conversion as there (core/beams.py:65-66
      before :82-93): :func:`_interp_freq`;
    * azimuth wrap, as pyuvdata's ``UVBeam._prepare_coordinate_data`` that
      gpu/beams.py:15-19 calls: when span + spacing is within one spacing of
      2 pi, 3 columns of wrapped data are added on each side and ``azmin``
      moves 3 spacings below the first azimuth;
    * the map coordinates are [za / dza, (az - azmin) / daz]
      (gpu/beams.py:185), so the za grid must start at 0.
    """
    inner, bt = unwrap(beam)
    bt = beam_type or bt or getattr(inner, "beam_type", "efield")
    if getattr(inner, "pixel_coordinate_system", "az_za") != "az_za":
        raise ValueError('pixel coordinate system must be "az_za"')
    az = np.asarray(inner.axis1_array, dtype=np.float64)
    za = np.asarray(inner.axis2_array, dtype=np.float64)
    if az.size < 2 or za.size < 2:
        raise ValueError("a grid beam needs at least 2 points on each axis")
    if abs(za[0]) > 1e-12:
        raise ValueError("the zenith-angle grid must start at 0")
    data = np.asarray(inner.data_array)
    freqs = np.asarray(inner.freq_array, float)
    # data: (Naxes_vec, Nfeeds|Npols, Nfreqs, nza, naz); its meaning follows
    # the stored beam's own type, the requested use follows `bt`
    stored = getattr(inner, "beam_type", "efield")
    if polarized:
        if bt != "efield" or stored != "efield":
            raise ValueError("beam type must be efield if using polarized=True")
        planes = _interp_freq(data[:2, :2], freqs, freq).astype(np.complex128)
        is_power = False
    else:
        if stored == "efield":
            # feed x, every vector axis: (nvec, Nfreqs, ...); the power in
            # FP64 whatever the stored precision
            e = data[:, 0].astype(np.complex128)
            power = (e.real ** 2 + e.imag ** 2).sum(axis=0)
        else:
            power = np.real(data[0, 0])
        planes = _interp_freq(power[None, None], freqs, freq).astype(
            np.float64)
        is_power = True
    step = az[1] - az[0]
    span = abs(az[0] - az[-1]) + step
    if np.isclose(span, 2 * np.pi, atol=step):
        ext = 3
        az = np.concatenate([az[0] + np.arange(-ext, 0) * step, az,
                             az[-1] + np.arange(1, ext + 1) * step])
        planes = np.concatenate([planes[..., -ext:], planes,
                                 planes[..., :ext]], axis=-1)
    # spacings as the reference takes them: first differences of the
    # (padded) axes (gpu/beams.py:19)
    daz = float(np.diff(az)[0])
    dza = float(np.diff(za)[0])
    return planes, daz, dza, float(az.min()), is_power
