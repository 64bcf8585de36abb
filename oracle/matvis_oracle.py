"""CPU oracle for matvis's per-time visibility computation.

TEST INFRASTRUCTURE ONLY.  The product package (``matvis_amd``) never imports
this module.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it: as the reference result the HIP
path is compared against, and as the CPU baseline timed beside the GPU figures.

It is a numpy (float64 / complex128) restatement of matvis's CPU path at
``precision=2``.  Every function cites the reference file:line it follows
(paths relative to the matvis source tree).  The reference itself cannot be
imported here (it hard-imports pyuvdata, which is absent), so the restatement is
pinned instead by:

* the reference's own known-answer tests, ported with the same inputs
  (``tests/test_matprod.py``, ``tests/test_cublas.py``,
  ``tests/test_beam_interp_gpu.py``, ``tests/test_coordinates.py``,
  ``tests/test_matvis_gpu.py::test_antizenith``);
* ``scipy.ndimage.map_coordinates`` (the routine the reference's CPU beam path
  reaches through pyuvdata, and that cupyx mirrors on its GPU path), which the
  beam interpolation here calls directly;
* closed forms and brute force on tiny inputs (zenith phase = 1, Hermitian
  symmetry, an explicit triple loop for ``conj(Z) Z^T`` with the feed swap).

Parity status per function (see ``tests/test_oracle.py``):

==========================  =====================================================
``point_source_crd_eq``     pinned: cardinal points (test_coordinates.py:21-68)
``eci_to_enu_matrix``       pinned: zenith at HA=0, orthonormal (:71-106)
``rotate``                  pinned: brute force vs an explicit matrix product
``select_chunk``            pinned: ascending-order compaction, zero padding
``enu_to_az_za``            pinned: E/N/U conventions (test_coordinates.py:108-160)
``prepare_grid``            pinned: continuity across az = 0 / 2 pi and
                            accuracy on both sides of the seam for a
                            trigonometric-polynomial beam (a missing wrap or
                            a wrong azmin fails); power conversion before
                            the frequency step; **parity unpinned**: the
                            3-column padding width of pyuvdata's
                            _prepare_coordinate_data
``beam_at_frequency``       pinned: scipy interp1d(kind="cubic"), the
                            routine UVBeam.interp calls; beams cubic in
                            frequency reproduced exactly off the nodes (a
                            linear restatement fails); node within 1 Hz;
                            **parity unpinned**: that "cubic" is pyuvdata's
                            default kind (its documentation; absent here)
``interp_grid_beam``        pinned: scipy map_coordinates; node identity and
                            linear exactness (test_beam_interp_gpu.py:12-64)
``gaussian_beam``           pinned in part: with the diameter rule's sigma
                            the E-field ``exp(-za^2 / (2 sigma^2))`` is 1/2
                            at the Airy E-field's half maximum (scipy j1;
                            a power-beam sigma or the intensity half
                            maximum fails); **parity unpinned**: the
                            ``/sqrt(2)`` per feed/axis when polarized and
                            pyuvdata's 2.355 (pyuvdata is absent)
``exptau``                  pinned: zenith source gives exactly 1; closed form
                            ``exp(+2 pi i nu b.s / c)`` on a 1-antenna case
``zmatrix``                 pinned: per-antenna beam loop == broadcast
                            (test_matvis_cpu.py:61-74 bit-exact invariance)
``vis_from_z``              pinned: test_matprod.py:14-71 (arange + 1j),
                            test_cublas.py:13-25, explicit triple loop
``simulate``                pinned: anti-zenith source gives V == 0
                            (test_matvis_gpu.py:15-25); chunk-count invariance
``erfa_ld``, ``erfa_ab``    pinned: pyerfa's eraLd / eraAb on the golden
                            sources and contexts (tests/golden/erfa_vectors.npz)
``erfa_bcrs``,              pinned: pyerfa's eraAtco13 observed directions; the
``erfa_atioq_matrix``       residual equals the diurnal aberration the
                            reference omits (checked by applying it)
``vecs2rot``,               pinned: carries r1 onto r2, orthonormal with
``axis_angle_rotation``     determinant 1, quarter turns about the axes
``coherency_rotation``      pinned: for a rigid Earth the angle X is minus
                            the textbook parallactic angle
                            atan2(sin H, tan(lat) cos(dec) - sin(dec) cos H)
``rotate_coherency``        pinned: Stokes I, V and Q^2 + U^2 invariant;
                            Q' = Q cos 2X - U sin 2X, U' = Q sin 2X + U cos 2X
``select_chunk_polarized``  pinned: the two above plus the compaction of
                            ``select_chunk``
==========================  =====================================================
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage

#: Speed of light in m/s, the value of ``astropy.constants.c`` used by the
#: reference (core/tau.py:4,31).
C_LIGHT = 299792458.0


def get_dtypes(precision: int):
    """(real, complex) dtypes for a precision level (_utils.py:291-319)."""
    if precision == 1:
        return np.float32, np.complex64
    return np.float64, np.complex128


def point_source_crd_eq(ra, dec):
    """ECI unit vectors ``(cos ra cos dec, cos dec sin ra, sin dec)``.

    Follows coordinates.py:174-200.  Shape ``(3, nsrc)``.
    """
    ra = np.asarray(ra, dtype=np.float64)
    dec = np.asarray(dec, dtype=np.float64)
    return np.asarray(
        [np.cos(ra) * np.cos(dec), np.cos(dec) * np.sin(ra), np.sin(dec)]
    )


def eci_to_enu_matrix(ha, lat):
    """3x3 ECI -> ENU rotation at hour angle ``ha`` and latitude ``lat``.

    Follows coordinates.py:91-128.
    """
    return np.array(
        [
            [-np.sin(ha), np.cos(ha), 0.0 * ha],
            [-np.sin(lat) * np.cos(ha), -np.sin(lat) * np.sin(ha), np.cos(lat)],
            [np.cos(lat) * np.cos(ha), np.cos(lat) * np.sin(ha), np.sin(lat)],
        ]
    )


def rotate(rot, eci):
    """Topocentric ENU vectors ``R_t . X_eq``.

    Step 2 of docs/understanding_the_algorithm.rst:118-121 (the ERFA path's
    ``_atioq`` matmul, cpu/coords.py:89-116).  Each output row is evaluated as
    ``(R[i,0]*x + R[i,1]*y) + R[i,2]*z`` in float64, one rounding per
    operation and no fused multiply-add, so that the horizon test ``z > 0``
    is reproducible bit for bit by any implementation that keeps this order.
    """
    rot = np.asarray(rot, dtype=np.float64)
    x, y, z = np.asarray(eci, dtype=np.float64)
    out = np.empty((3, x.size), dtype=np.float64)
    for i in range(3):
        out[i] = rot[i, 0] * x + rot[i, 1] * y + rot[i, 2] * z
    return out


def select_chunk(topo, flux, chunk: int, chunk_size: int, nsrc_alloc: int):
    """Horizon cull and stable compaction of one source chunk.

    Follows core/coords.py:99-150: ``where(topo[2] > 0)`` keeps ascending
    source order; coordinates and fluxes are gathered into buffers of length
    ``nsrc_alloc``; fluxes past ``n`` are zero.  Coordinates past ``n`` are
    zero here (the reference leaves stale values there, which its zero flux
    makes irrelevant).  Returns ``(crd, flux, n, index)`` where ``index`` holds
    the global source indices kept.
    """
    slc = slice(chunk * chunk_size, (chunk + 1) * chunk_size)
    t = topo[:, slc]
    f = flux[slc]
    above = np.where(t[2] > 0)[0]
    n = len(above)
    if n > nsrc_alloc:
        raise ValueError(
            f"nsrc_alloc ({nsrc_alloc}) is too small for the number of "
            f"sources above horizon ({n}). Try increasing source_buffer."
        )
    crd = np.zeros((3, nsrc_alloc), dtype=np.float64)
    fl = np.zeros(nsrc_alloc, dtype=f.dtype)
    crd[:, :n] = t[:, above]
    fl[:n] = f[above]
    return crd, fl, n, above + chunk * chunk_size


def enu_to_az_za(enu_e, enu_n, orientation="astropy", periodic_azimuth=True):
    """Direction cosines -> (azimuth, zenith angle).

    Follows coordinates.py:31-88, including the uvbeam orientation flip
    ``az -> pi/2 - az`` and the ``mod 2 pi``.
    """
    assert orientation in ("astropy", "uvbeam")
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


def theta_hat(theta, phi):
    """Unit vector of increasing colatitude at (theta, phi)
    (coordinates.py:537-556)."""
    theta, phi = np.asarray(theta, np.float64), np.asarray(phi, np.float64)
    return np.array([np.cos(phi) * np.cos(theta), np.sin(phi) * np.cos(theta),
                     -np.sin(theta)])


def phi_hat(theta, phi):
    """Unit vector of increasing longitude at (theta, phi)
    (coordinates.py:559-576)."""
    phi = np.asarray(phi, np.float64)
    return np.array([-np.sin(phi), np.cos(phi), np.zeros_like(phi)])


def axis_angle_rotation(axis, angle):
    """Rodrigues: I + sin(a) K + (1 - cos(a)) K^2 with K the cross-product
    matrix of each unit axis; (3, 3, N) (coordinates.py:414-452)."""
    axis = np.asarray(axis, np.float64)
    z = np.zeros(axis.shape[1])
    k = np.array([[z, -axis[2], axis[1]],
                  [axis[2], z, -axis[0]],
                  [-axis[1], axis[0], z]])
    k2 = np.einsum("ijn,jkn->ikn", k, k)
    return np.eye(3)[..., None] + np.sin(angle) * k + (1.0 - np.cos(angle)) * k2


def vecs2rot(r1, r2):
    """Rotation about r1 x r2 by the angle between them, carrying each r1 to
    r2 (coordinates.py:387-411).  Where r1 x r2 vanishes the reference's
    axis is 0/0 (NaN); here it is the identity, the limit as the angle goes
    to zero."""
    cross = np.cross(r1, r2, axis=0)
    sin_psi = np.linalg.norm(cross, axis=0)
    cos_psi = np.sum(r1 * r2, axis=0)
    safe = np.where(sin_psi > 0, sin_psi, 1.0)
    axis = cross / safe
    psi = np.where(sin_psi > 0, np.arctan2(sin_psi, cos_psi), 0.0)
    return axis_angle_rotation(axis, psi)


def coherency_rotation(ra, dec, alt, az, r_avg):
    """2x2 rotation (2, 2, N) carrying a coherency matrix from the RA/Dec
    basis to the alt/az basis (coordinates.py:288-384, 455-490): the average
    ICRS -> (N, E, U) rotation ``r_avg`` (3x3) composed with the per-source
    rotation that lands it on the source's (alt, az) direction; X is the
    angle between the equatorial theta_hat / phi_hat and the alt/az
    theta_hat carried back to ICRS."""
    frame = point_source_crd_eq(ra, dec)
    altaz = point_source_crd_eq(az, alt)
    r_avg = np.asarray(r_avg, np.float64)
    perturb = vecs2rot(r_avg @ frame, altaz)
    exact = np.einsum("abn,bc->acn", perturb, r_avg)
    theta, phi = 0.5 * np.pi - np.asarray(dec), np.asarray(ra)
    bh = np.einsum("ban,bn->an", exact,
                   theta_hat(0.5 * np.pi - np.asarray(alt), az))
    cos_x = np.einsum("an,an->n", bh, theta_hat(theta, phi))
    sin_x = np.einsum("an,an->n", bh, phi_hat(theta, phi))
    return np.array([[cos_x, sin_x], [-sin_x, cos_x]])


def rotate_coherency(coh, rot):
    """rot^T C rot per source and frequency: coh (N, nfreq, 2, 2), rot
    (2, 2, N) (core/coords.py:169-175)."""
    return np.einsum("ban,nfbc,cdn->nfad", rot, coh, rot)


def select_chunk_polarized(topo, coh, ra, dec, r_avg, chunk: int,
                           chunk_size: int, nsrc_alloc: int):
    """Polarized branch of CoordinateRotation.select_chunk
    (core/coords.py:99-150): cull and compact as :func:`select_chunk`, then
    rotate the kept sources' coherency matrices into the alt/az basis.
    Returns (coords (3, nsrc_alloc), coherency (nsrc_alloc, nfreq, 2, 2),
    n, global indices kept)."""
    topo = np.asarray(topo)
    crd, _, n, index = select_chunk(topo, np.zeros(topo.shape[1]), chunk,
                                    chunk_size, nsrc_alloc)
    coh = np.asarray(coh)
    out = np.zeros((nsrc_alloc,) + coh.shape[1:],
                   np.complex128 if np.iscomplexobj(coh) else np.float64)
    if n:
        az, za = enu_to_az_za(crd[0, :n], crd[1, :n])
        rot = coherency_rotation(np.asarray(ra)[index],
                                 np.asarray(dec)[index], 0.5 * np.pi - za,
                                 az, r_avg)
        out[:n] = rotate_coherency(coh[index], rot)
    return crd, out, n, index


#: UVBeam.interp's freq_interp_tol default (Hz).
FREQ_INTERP_TOL = 1.0


def beam_at_frequency(d, freq_array, freq: float):
    """``d[:, :, f]`` at one frequency, as ``UVBeam.interp(freq_array=[freq])``
    with its defaults, reached from core/beams.py:82-93 (pyuvdata 3.2.4,
    ``uv.lock:2932-2968``; absent here): the nearest node's data when it is
    within ``freq_interp_tol`` = 1 Hz, otherwise ``freq_interp_kind="cubic"``,
    i.e. scipy's ``interp1d(kind="cubic")`` along the frequency axis, real
    and imaginary parts separately (at least 4 nodes; frequencies outside
    the nodes' range raise).  A single-frequency beam is returned unchanged
    (the reference raises there; here it is a frequency-independent beam).
    """
    from scipy.interpolate import interp1d

    fa = np.atleast_1d(np.asarray(freq_array, dtype=np.float64))
    if fa.size == 1:
        return d[:, :, 0]
    dist = np.abs(fa - freq)
    if dist.min() < FREQ_INTERP_TOL:
        return d[:, :, int(np.argmin(dist))]
    if not fa.min() <= freq <= fa.max():
        raise ValueError("frequency outside the beam's range")
    if fa.size < 4:
        raise ValueError("cubic frequency interpolation needs 4 frequencies")
    out = interp1d(fa, np.real(d), kind="cubic", axis=2)(freq)
    if np.iscomplexobj(d):
        out = out + 1j * interp1d(fa, np.imag(d), kind="cubic", axis=2)(freq)
    return out


def prepare_grid(data_array, az, za, freq_array, freq: float,
                 polarized: bool, beam_type: str = "efield"):
    """Interpolation grid of one UVBeam-shaped az/za beam at one frequency.

    The preparation the reference applies before ``gpu_beam_interpolation``:

    * unpolarized: ``prepare_beam_unpolarized`` (core/beams.py:15-32), run
      first (core/beams.py:65-66): an E-field beam becomes the xx power beam
      ``sum_vec |E[vec, x]|^2`` (the first feed) at every frequency node, a
      power beam keeps its first polarization;
    * frequency: ``UVBeam.interp(freq_array=[freq])`` (core/beams.py:82-93),
      :func:`beam_at_frequency`;
    * ``prepare_for_map_coords`` (gpu/beams.py:15-19), which takes pyuvdata's
      ``UVBeam._prepare_coordinate_data``: when the azimuth axis covers the
      circle (span + spacing within one spacing of 2 pi) the grid is padded
      with 3 columns of wrapped data on each side; ``daz`` is the first
      spacing, ``azmin`` the padded axis's minimum.

    Returns ``(planes (nax, nfeed, nza, naz'), daz, dza, azmin)``.
    Pinned: the wrap and ``azmin`` by continuity across az = 0 / 2 pi and
    by accuracy against a trigonometric-polynomial beam on both sides of the
    seam; the frequency step by exact reproduction of beams cubic in
    frequency.  PARITY UNPINNED: the padding width of 3 columns (it sets the
    cubic prefilter's boundary near the seam) and pyuvdata's default kind
    "cubic" itself (from its documentation; pyuvdata is absent here).
    """
    data = np.asarray(data_array)
    az = np.asarray(az, dtype=np.float64)
    za = np.asarray(za, dtype=np.float64)
    if polarized:
        d = data[:2, :2].astype(np.complex128)
    elif beam_type == "efield":
        e = data[:, 0].astype(np.complex128)
        d = (e.real ** 2 + e.imag ** 2).sum(axis=0)[None, None]
    else:
        d = np.real(data[0, 0])[None, None].astype(np.float64)
    planes = beam_at_frequency(d, freq_array, freq)
    step = az[1] - az[0]
    span = abs(az[0] - az[-1]) + step
    if np.isclose(span, 2 * np.pi, atol=step):
        ext = 3
        az = np.concatenate([az[0] + np.arange(-ext, 0) * step, az,
                             az[-1] + np.arange(1, ext + 1) * step])
        planes = np.concatenate([planes[..., -ext:], planes,
                                 planes[..., :ext]], axis=-1)
    return planes, float(np.diff(az)[0]), float(np.diff(za)[0]), float(az.min())


def interp_grid_beam(planes, daz, dza, azmin, az, za, order: int = 1):
    """Interpolate az/za grid beams at source positions.

    Follows gpu/beams.py:127-200 (``gpu_beam_interpolation``): ``planes`` has
    shape ``(nbeam, nax, nfeed, nza, naz)`` and is real (power beam) or complex
    (E-field beam); the map coordinates are ``[za / dza, (az - azmin) / daz]``
    (gpu/beams.py:185); ``map_coordinates`` runs with its defaults
    (``mode='constant'``, ``cval=0``, spline prefilter for ``order > 1``); a
    power beam ends with a square root taken in the complex output buffer
    (gpu/beams.py:195-197 with ``out`` complex, gpu/beams.py:122), so a
    negative interpolated power ``p`` becomes ``1j * sqrt(-p)``.

    Returns ``(nbeam, nfeed, nax, nsrc)`` complex128.
    """
    planes = np.asarray(planes)
    nbeam, nax, nfeed = planes.shape[:3]
    az = np.asarray(az, dtype=np.float64)
    za = np.asarray(za, dtype=np.float64)
    out = np.zeros((nbeam, nfeed, nax, az.size), dtype=np.complex128)
    for bm in range(nbeam):
        coords = np.asarray([za / dza[bm], (az - azmin[bm]) / daz[bm]])
        for fd in range(nfeed):
            for ax in range(nax):
                out[bm, fd, ax] = ndimage.map_coordinates(
                    planes[bm, ax, fd], coords, order=order
                )
    if not np.iscomplexobj(planes):
        out = np.sqrt(out)
    return out


def gaussian_beam(sigma, za, polarized: bool):
    """Analytic Gaussian beam response, shape ``(nbeam, nfeed, nax, nsrc)``.

    Restates pyuvdata's ``GaussianBeam`` as the reference consumes it
    (core/beams.py:15-32, cpu/beams.py:41-46): the E-field amplitude is
    ``exp(-za^2 / (2 sigma^2))``; unpolarized simulations take the square
    root of the xx power beam, which is that amplitude; polarized ones see
    the amplitude divided by sqrt(2) on every (feed, axis).
    Pinned in part (the E-field convention against the Airy half maximum of
    the diameter rule); the polarized normalization is parity unpinned.
    """
    sigma = np.atleast_1d(np.asarray(sigma, dtype=np.float64))
    za = np.asarray(za, dtype=np.float64)
    nf = na = 2 if polarized else 1
    out = np.zeros((sigma.size, nf, na, za.size), dtype=np.complex128)
    for b, sg in enumerate(sigma):
        e = np.exp(-(za**2) / (2.0 * sg**2))
        out[b] = e / np.sqrt(2.0) if polarized else e
    return out


def exptau(antpos, freq: float, crd):
    """``exp(i 2 pi nu (b . s) / c)``, shape ``(nant, nsrc)``.

    Follows core/tau.py:30-31 (``antpos * 2 pi nu * 1j / c``, positive sign)
    and core/tau.py:65-66 (matmul then exp).
    """
    ang_freq = np.float64(2.0 * np.pi * freq)
    antpos_c = np.asarray(antpos, dtype=np.float64) * ang_freq * 1j / C_LIGHT
    return np.exp(antpos_c @ np.asarray(crd, dtype=np.float64))


def zmatrix(sqrt_flux, beam, exptau_, beam_idx, nfeed: int, nax: int):
    """``Z = A sqrt(I) exp(tau)``, shape ``(nant * nfeed, nax * nsrc)``.

    Follows core/getz.py:80-100: row = ``ant * nfeed + feed``, column =
    ``ax * nsrc + src``; ``beam`` is ``(nbeam, nfeed, nax, nsrc)`` broadcast
    over antennas, or indexed per antenna through ``beam_idx``.
    """
    nant, nsrc = exptau_.shape
    e = exptau_ * sqrt_flux
    z = np.empty((nant, nfeed, nax, nsrc), dtype=np.complex128)
    z[:] = e[:, None, None, :]
    if beam_idx is None:
        z *= beam
    else:
        for ant, b in enumerate(beam_idx):
            z[ant] *= beam[b]
    return z.reshape(nant * nfeed, nax * nsrc)


def vis_from_z(z, nant: int, nfeed: int, antpairs=None):
    """``V = conj(Z) Z^T`` in the API layout ``(npairs, nfeed, nfeed)``.

    Follows cpu/matprod.py:21-30: the product is reshaped to
    ``(nant, nfeed, nant, nfeed)`` and transposed ``(0, 2, 3, 1)``, so that
    ``out[i * nant + j, p, q] = sum conj(Z[i*nf + q]) Z[j*nf + p]`` (the feed
    swap that tests/test_matprod.py:56-65 pins).
    """
    v = z.conj().dot(z.T)
    v = v.reshape(nant, nfeed, nant, nfeed).transpose(0, 2, 3, 1)
    if antpairs is None:
        return v.reshape(nant * nant, nfeed, nfeed)
    antpairs = np.asarray(antpairs)
    return v[antpairs[:, 0], antpairs[:, 1]]


def desired_chunks(nsrc: int, nchunks: int):
    """``(nchunks, sources per chunk)`` as get_desired_chunks returns them.

    Follows _utils.py:259-288: ``nchunks = min(nchunks, nsrc)`` and
    ``ceil(nsrc / nchunks)`` sources per chunk.
    """
    nchunks = max(1, min(nchunks, nsrc))
    return nchunks, int(np.ceil(nsrc / nchunks))


def beam_response(beam: dict, crd, polarized: bool):
    """Evaluate a beam description at compacted topocentric coordinates.

    ``beam`` is ``{"kind": "grid", "planes", "daz", "dza", "azmin",
    "order"}`` or ``{"kind": "gaussian", "sigma"}``.  Grid beams go through
    ``enu_to_az_za(..., orientation="uvbeam")`` first (gpu/beams.py:110).
    """
    az, za = enu_to_az_za(crd[0], crd[1], orientation="uvbeam")
    if beam["kind"] == "grid":
        return interp_grid_beam(
            beam["planes"], beam["daz"], beam["dza"], beam["azmin"], az, za,
            order=beam.get("order", 1),
        )
    if beam["kind"] == "gaussian":
        return gaussian_beam(beam["sigma"], za, polarized)
    raise ValueError(f"unknown beam kind {beam['kind']!r}")


def simulate(
    *,
    antpos,
    freq: float,
    eci,
    rotations,
    I_sky,
    beam: dict,
    polarized: bool = False,
    antpairs=None,
    beam_idx=None,
    nchunks: int = 1,
    source_buffer: float = 1.0,
):
    """The per-time loop of ``cpu.simulate`` at precision 2.

    Follows cpu/cpu.py:151-277: flux ``sqrt(0.5 I)`` (cpu/cpu.py:188 via
    gpu/gpu.py:127), per time a rotation, per chunk a cull/compaction, beam,
    exp(tau), Z and ``conj(Z) Z^T``; the chunk partials are summed
    (core/matprod.py:103-115).  A chunk with no source above the horizon
    contributes exact zeros (the CPU path has no ``nsrcs_up < 1`` skip).
    ``rotations`` is ``(ntimes, 3, 3)``: the host-computed ``R_t`` per time
    that the device path applies too.  ``nsrc_alloc`` follows the GPU rule
    ``int(chunk * source_buffer)`` (gpu/gpu.py:112).

    Returns ``(ntimes, npairs, nfeed, nfeed)`` if polarized, else
    ``(ntimes, npairs)``, complex128.
    """
    I_sky = np.asarray(I_sky, dtype=np.float64)
    assert I_sky.ndim == 1, "I_sky must have shape (NSRCS,)."
    nax = nfeed = 2 if polarized else 1
    antpos = np.asarray(antpos, dtype=np.float64)
    nant = antpos.shape[0]
    nsrc = I_sky.size
    nchunks, chunk_size = desired_chunks(nsrc, nchunks)
    nsrc_alloc = int(chunk_size * source_buffer)
    flux = np.sqrt(0.5 * I_sky)
    rotations = np.asarray(rotations, dtype=np.float64)
    ntimes = rotations.shape[0]
    npairs = nant * nant if antpairs is None else len(antpairs)
    vis = np.zeros((ntimes, npairs, nfeed, nfeed), dtype=np.complex128)
    for t in range(ntimes):
        topo = rotate(rotations[t], eci)
        for c in range(nchunks):
            crd, fl, _, _ = select_chunk(topo, flux, c, chunk_size, nsrc_alloc)
            a = beam_response(beam, crd, polarized)
            et = exptau(antpos, freq, crd)
            z = zmatrix(fl, a, et, beam_idx, nfeed, nax)
            vis[t] += vis_from_z(z, nant, nfeed, antpairs)
    return vis if polarized else vis[:, :, 0, 0]


#: Schwarzschild radius of the Sun in au (cpu/coords.py:10).
ERFA_SRS = 1.97412574336e-8


def erfa_ld(p, e, em: float, dlim: float):
    """Light deflection by the Sun (cpu/coords.py:110-137): p (3, n) unit
    directions, e (3,) Sun -> observer unit vector, em its distance in au;
    p + w p x (e x p) with w = SRS / em / max(p . (p + e), dlim)."""
    p = np.asarray(p, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    qdqpe = np.maximum(np.sum(p * (p + e[:, None]), axis=0), dlim)
    w = ERFA_SRS / em / qdqpe
    eq = np.cross(e[None, :], p.T)
    peq = np.cross(p.T, eq).T
    return p + w[None] * peq


def erfa_ab(pnat, v, s: float, bm1: float):
    """Aberration (cpu/coords.py:139-163): natural direction (3, n) to
    proper direction for an observer at barycentric velocity v (units of c),
    distance s (au) from the Sun, bm1 = sqrt(1 - |v|^2); renormalized."""
    pnat = np.asarray(pnat, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    pdv = v @ pnat
    w1 = 1.0 + pdv / (1.0 + bm1)
    w2 = ERFA_SRS / s
    p = bm1 * pnat + np.outer(v, w1) + w2 * (v[:, None] - pdv[None] * pnat * bm1)
    return p / np.linalg.norm(p, axis=0)


def erfa_bcrs(eci, ctx: dict):
    """CIRS proper directions of ICRS unit vectors (cpu/coords.py:176-215):
    light deflection (limiter 1e-6), aberration, bias-precession-nutation."""
    p = erfa_ld(eci, ctx["eh"], ctx["em"], 1e-6)
    p = erfa_ab(p, ctx["v"], ctx["em"], ctx["bm1"])
    return np.asarray(ctx["bpn"], dtype=np.float64) @ p


def erfa_atioq_matrix(ctx: dict):
    """CIRS -> (east, north, up) (cpu/coords.py:78-108): rotation by the
    local Earth rotation angle, polar motion, then HA/Dec -> ENU."""
    ce, se = np.cos(ctx["eral"]), np.sin(ctx["eral"])
    c2h = np.array([[ce, se, 0.0], [-se, ce, 0.0], [0.0, 0.0, 1.0]])
    sx, cx = np.sin(ctx["xpl"]), np.cos(ctx["xpl"])
    sy, cy = np.sin(ctx["ypl"]), np.cos(ctx["ypl"])
    pm = np.array([[cx, 0.0, sx], [sx * sy, cy, -cx * sy],
                   [-sx * cy, sy, cx * cy]])
    sphi, cphi = ctx["sphi"], ctx["cphi"]
    enu = np.array([[0.0, 1.0, 0.0], [-sphi, 0.0, cphi], [cphi, 0.0, sphi]])
    return enu @ pm @ c2h


def normalized_error(v_test, v_ref, nant: int, nfeed: int):
    """Max and median of ``|V - V_ref| / sqrt(|V_ref_ii| |V_ref_jj|)``.

    ``v_*`` are ``(..., nant*nant, nfeed, nfeed)`` in the API layout.  The
    normaliser of element ``(i, j, p, q)`` is the geometric mean of the two
    autocorrelations it pairs, ``V[i,i,q,q]`` and ``V[j,j,p,p]``.  Returns
    ``(max, median, per-element max relative, per-element median relative)``.
    """
    vt = np.asarray(v_test).reshape(-1, nant, nant, nfeed, nfeed)
    vr = np.asarray(v_ref).reshape(-1, nant, nant, nfeed, nfeed)
    idx = np.arange(nant)
    auto = np.abs(vr[:, idx, idx])  # (T, nant, nfeed, nfeed)
    autod = np.stack([auto[:, :, f, f] for f in range(nfeed)], axis=-1)
    # element (i, j, p, q) pairs conj(Z[i,q]) with Z[j,p]
    norm = np.sqrt(autod[:, :, None, None, :] * autod[:, None, :, :, None])
    diff = np.abs(vt - vr)
    with np.errstate(divide="ignore", invalid="ignore"):
        nerr = np.where(norm > 0, diff / norm, 0.0)
        rel = np.where(np.abs(vr) > 0, diff / np.abs(vr), 0.0)
    return (
        float(nerr.max()),
        float(np.median(nerr)),
        float(rel.max()),
        float(np.median(rel)),
    )
