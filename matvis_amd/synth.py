"""Seeded synthetic inputs for the benchmark workloads.

Arrays, skies and beams of the documented configurations:

* ``disk_array``: antennas uniform in a disk (z = 0);
* ``hex_array``: the n antennas nearest the centre of a hexagonal grid, with
  uniform position jitter;
* ``random_sky``: RA uniform, sin(dec) uniform on [sin dec_min, sin dec_max];
* ``powerlaw_fluxes`` (dN/dS ~ S^-2 on [smin, smax]) and lognormal fluxes;
* ``healpix_centres``: pixel centres of a HEALPix RING grid;
* ``harmonic_efield_beam``: a complex E-field beam on a 1-degree az/za grid,
  a seeded sum of cos(m az) P_l(cos za) terms (l <= 8) tapered to zero at the
  horizon.
"""

from __future__ import annotations

import numpy as np
from scipy.special import eval_legendre

from .astro import EarthLocation, Time
from .beams import GridBeam

HERA_LAT_DEG = -30.72
HERA_LON_DEG = 21.43
START_JD = 2459863.0


def hera_location():
    return EarthLocation.from_geodetic(HERA_LON_DEG, HERA_LAT_DEG, 1073.0)


def times_spaced(ntimes: int, minutes: float, start_jd: float = START_JD):
    return Time(start_jd + np.arange(ntimes) * minutes / (24.0 * 60.0))


def disk_array(nant: int, radius: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.uniform(size=nant))
    phi = rng.uniform(0, 2 * np.pi, size=nant)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(nant)], 1)


def hex_array(nant: int, spacing: float = 14.6, jitter: float = 0.02,
              seed: int = 0) -> np.ndarray:
    rings = int(np.ceil(np.sqrt(nant))) + 2
    pts = []
    for i in range(-rings, rings + 1):
        for j in range(-rings, rings + 1):
            pts.append((spacing * (i + 0.5 * j), spacing * j * np.sqrt(3) / 2))
    pts = np.asarray(pts)
    order = np.lexsort((pts[:, 1], pts[:, 0], np.hypot(pts[:, 0], pts[:, 1])))
    pos = pts[order[:nant]]
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    return np.concatenate([pos, np.zeros((nant, 1))], 1)


def random_sky(nsrc: int, seed: int, dec_min_deg: float = -90.0,
               dec_max_deg: float = 30.0):
    rng = np.random.default_rng(seed)
    ra = rng.uniform(0, 2 * np.pi, size=nsrc)
    s = rng.uniform(np.sin(np.radians(dec_min_deg)),
                    np.sin(np.radians(dec_max_deg)), size=nsrc)
    return ra, np.arcsin(s)


def powerlaw_fluxes(nsrc: int, seed: int, smin: float = 0.1,
                    smax: float = 10.0) -> np.ndarray:
    """dN/dS ~ S^-2 on [smin, smax] by inverse transform."""
    u = np.random.default_rng(seed).uniform(size=nsrc)
    return 1.0 / (1.0 / smin - u * (1.0 / smin - 1.0 / smax))


def lognormal_fluxes(nsrc: int, seed: int, mu: float = 0.0,
                     sigma: float = 1.0) -> np.ndarray:
    return np.random.default_rng(seed).lognormal(mu, sigma, size=nsrc)


def _isqrt(x):
    r = np.floor(np.sqrt(x.astype(np.float64))).astype(np.int64)
    r -= (r * r > x)
    r += ((r + 1) * (r + 1) <= x)
    return r


def healpix_centres(nside: int):
    """(ra, dec) of HEALPix RING pixel centres (Gorski et al. 2005)."""
    npix = 12 * nside * nside
    p = np.arange(npix, dtype=np.int64)
    ncap = 2 * nside * (nside - 1)
    z = np.empty(npix)
    phi = np.empty(npix)
    half_pi = 0.5 * np.pi
    # north polar cap: rings 1 .. nside-1 with 4 i pixels
    n = p < ncap
    ir = (1 + _isqrt(1 + 2 * p[n])) >> 1
    ip = p[n] + 1 - 2 * ir * (ir - 1)
    z[n] = 1.0 - ir * ir / (3.0 * nside * nside)
    phi[n] = (ip - 0.5) * half_pi / ir
    # equatorial belt: rings nside .. 3 nside with 4 nside pixels
    e = (p >= ncap) & (p < npix - ncap)
    q = p[e] - ncap
    ir = q // (4 * nside) + nside
    ip = q % (4 * nside) + 1
    fodd = np.where((ir + nside) & 1, 1.0, 0.5)
    z[e] = (2 * nside - ir) / (1.5 * nside)
    phi[e] = (ip - fodd) * half_pi / nside
    # south polar cap
    sc = p >= npix - ncap
    q = npix - p[sc]
    ir = (1 + _isqrt(2 * q - 1)) >> 1
    ip = 4 * ir + 1 - (q - 2 * ir * (ir - 1))
    z[sc] = -1.0 + ir * ir / (3.0 * nside * nside)
    phi[sc] = (ip - 0.5) * half_pi / ir
    return phi, np.arcsin(z)


def harmonic_efield_beam(seed: int, naz: int = 360, nza: int = 91,
                         lmax: int = 8, mmax: int = 4, polarized: bool = True,
                         perturb: float = 0.0, perturb_seed: int | None = None,
                         dtype=np.complex64) -> GridBeam:
    """Seeded smooth E-field beam on a 1-degree grid, zero at the horizon."""
    rng = np.random.default_rng(seed)
    az = np.radians(np.arange(naz) * 360.0 / naz)
    za = np.radians(np.linspace(0.0, 90.0, nza))
    AZ, ZA = np.meshgrid(az, za)  # (nza, naz)
    nax = nfd = 2 if polarized else 1
    data = np.zeros((nax, nfd, 1, nza, naz), dtype=np.complex128)
    taper = np.cos(ZA)
    for a in range(nax):
        for f in range(nfd):
            acc = np.zeros_like(AZ, dtype=np.complex128)
            for l in range(lmax + 1):
                pl = eval_legendre(l, np.cos(ZA))
                for m in range(min(l, mmax) + 1):
                    c = (rng.normal() + 1j * rng.normal()) / (1.0 + l + m)
                    ph = rng.uniform(0, 2 * np.pi)
                    acc += c * np.cos(m * AZ + ph) * pl
            data[a, f, 0] = acc * taper
    # main response on the co-polar terms
    data[0, 0, 0] += 2.0 * taper**2
    if nax > 1:
        data[1, 1, 0] += 2.0 * taper**2
    data /= np.abs(data).max()
    if perturb:
        prng = np.random.default_rng(perturb_seed)
        data *= 1.0 + perturb * prng.uniform(-1, 1, size=data.shape)
    return GridBeam(data.astype(dtype), az, za, freq_array=np.array([150e6]),
                    beam_type="efield")


#: Antenna counts of the serving workload's requests.
SERVE_NANTS = (32, 128, 350, 512)


def serving_draw(rng) -> dict:
    """Sizes of one request of the Poisson serving workload: nant from
    {32, 128, 350, 512}, nsrc log-uniform on [1e3, 2e5], nfreq uniform on
    [1, 64], ntime uniform on [1, 16], polarized with p = 0.5, plain FP64 or
    split-BF16 with p = 0.5 (drawn in that order from ``rng``)."""
    return dict(nant=int(rng.choice(SERVE_NANTS)),
                nsrc=int(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))),
                nfreq=int(rng.integers(1, 65)),
                ntime=int(rng.integers(1, 17)),
                polarized=bool(rng.random() < 0.5),
                fp64=bool(rng.random() < 0.5))


class ServingSky:
    """Shared inputs of the serving workload: one harmonic E-field beam on
    the 1-degree 360 x 91 grid (bicubic), 16 times 2 min apart at the HERA
    site with their rotation matrices, hex arrays with 14.6 m spacing."""

    def __init__(self, beam=None):
        from . import astro

        self.loc = hera_location()
        self.beam = beam or harmonic_efield_beam(seed=4, naz=360, nza=91)
        self.times = times_spaced(16, 2.0)
        self.rots = astro.rotation_matrices(self.times, self.loc)
        self._ants: dict[int, np.ndarray] = {}

    def antpos(self, nant: int) -> np.ndarray:
        if nant not in self._ants:
            self._ants[nant] = hex_array(nant, 14.6, 0.02, seed=nant)
        return self._ants[nant]

    def request(self, nant: int, nsrc: int, nfreq: int, ntime: int,
                polarized: bool, fp64: bool, seed: int,
                flux_scale: float = 1.0) -> dict:
        """``simulate_multi`` keywords of one request: RA uniform, sin(dec)
        uniform on [-90, +30] deg, power-law S^-2 fluxes on [0.1, 10] Jy
        (times ``flux_scale``), nfreq frequencies evenly over 100-200 MHz,
        the first ntime times."""
        from . import astro

        ra, dec = random_sky(nsrc, seed, -90.0, 30.0)
        flux = powerlaw_fluxes(nsrc, seed + 1) * flux_scale
        freqs = np.linspace(100e6, 200e6, nfreq)
        return dict(antpos=self.antpos(nant), freqs=freqs,
                    times=self.times[:ntime],
                    skycoords=astro.SkyCoord(ra, dec), telescope_loc=self.loc,
                    fluxes=np.repeat(flux[:, None], nfreq, 1),
                    beam_list=[self.beam], polarized=polarized, precision=2,
                    coord_method="CoordinateRotationMatrix",
                    coord_method_params={"rotations": self.rots[:ntime]},
                    beam_spline_opts={"order": 3},
                    matprod_method="GPUMatMul" if fp64
                    else "GPUMatMulSplitBF16")
