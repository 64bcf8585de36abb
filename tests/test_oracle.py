"""Pins the CPU oracle against the reference's known answers and mathematics.

Each test names the reference test or closed form it ports.  None of them
restates the oracle's own formula: the checks are cardinal points, inverse
matrices, scipy, explicit loops over a different algorithm, or invariances.
"""

import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from _problems import small_problem

# ------------------------------------------------------------------ (a)


def test_point_source_crd_eq_cardinal_points():
    """tests/test_coordinates.py:21-68."""
    ra = np.radians([0.0, 45.0, 90.0, 135.0, 180.0, 270.0, 360.0])
    pole = np.full(ra.size, 0.5 * np.pi)
    c = O.point_source_crd_eq(ra, pole)
    assert np.allclose(c[0], 0) and np.allclose(c[1], 0) and np.allclose(c[2], 1)
    c = O.point_source_crd_eq(ra, -pole)
    assert np.allclose(c[2], -1)
    c = O.point_source_crd_eq(ra, 0 * pole)
    assert np.isclose(c[0, 0], 1) and np.isclose(c[1, 0], 0)
    assert np.isclose(c[0, 2], 0) and np.isclose(c[1, 2], 1)
    assert np.allclose(c[2], 0)


def test_eci_to_enu_zenith_and_conventions():
    """tests/test_coordinates.py:71-160: RA = LST at dec = lat is overhead;
    E/N/U map to the documented az/za in both orientations."""
    lat = np.radians(-30.7215)
    lsts = np.linspace(0, 2 * np.pi, 24)
    crd = O.point_source_crd_eq(lsts, np.full(lsts.size, lat))
    for i, lst in enumerate(lsts):
        m = O.eci_to_enu_matrix(lst, lat)
        assert np.allclose(m @ m.T, np.eye(3))
        top = O.rotate(m, crd)
        assert np.allclose(top[:, i], [0, 0, 1])
        az, za = O.enu_to_az_za(top[0], top[1], "astropy")
        assert np.isclose(za[i], 0.0)
    az, za = O.enu_to_az_za(1.0, 0.0, "astropy")
    assert np.isclose(az, 0.5 * np.pi) and np.isclose(za, 0.5 * np.pi)
    az, za = O.enu_to_az_za(0.0, 1.0, "astropy")
    assert np.isclose(az, 0.0) and np.isclose(za, 0.5 * np.pi)
    az, za = O.enu_to_az_za(1.0, 0.0, "uvbeam")
    assert np.isclose(az, 0.0) and np.isclose(za, 0.5 * np.pi)
    az, za = O.enu_to_az_za(0.0, 1.0, "uvbeam")
    assert np.isclose(az, 0.5 * np.pi)
    az, za = O.enu_to_az_za(0.0, -1.0, "uvbeam", periodic_azimuth=False)
    assert np.isclose(az, -0.5 * np.pi)
    _, za = O.enu_to_az_za(0.0, 0.0, "uvbeam")
    assert np.isclose(za, 0.0)


def test_rotate_agrees_with_blas_product(rng):
    r = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    x = rng.normal(size=(3, 1000))
    np.testing.assert_allclose(O.rotate(r, x), r @ x, rtol=0, atol=1e-14)


def test_select_chunk_order_padding_and_overflow():
    """core/coords.py:109-150: ascending order, zero flux past n, error text."""
    topo = np.array([[1, 2, 3, 4, 5, 6.0], [0] * 6, [0.5, -1, 0.2, 0.0, 1, 1]])
    flux = np.arange(6.0) + 1
    crd, fl, n, idx = O.select_chunk(topo, flux, 0, 6, 5)
    assert n == 4 and list(idx) == [0, 2, 4, 5]
    assert list(crd[0, :4]) == [1, 3, 5, 6] and list(fl) == [1, 3, 5, 6, 0]
    crd, fl, n, idx = O.select_chunk(topo, flux, 1, 4, 4)
    assert n == 2 and list(idx) == [4, 5]
    with pytest.raises(ValueError, match="is too small for the number of"):
        O.select_chunk(topo, flux, 0, 6, 3)


# ------------------------------------------------------------------ (b)


def test_beam_interp_identity_nodes():
    """tests/test_beam_interp_gpu.py:12-38: grid nodes reproduce sqrt(power)."""
    za = np.linspace(0, 1, 3)
    az = np.linspace(0, 2 * np.pi, 6)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA**2, 2 - ZA**2 + np.sin(AZ)])[:, None, None]
    out = O.interp_grid_beam(beam, [az[1] - az[0]] * 2, [za[1] - za[0]] * 2,
                             [0.0] * 2, AZ.ravel(), ZA.ravel())
    for b in range(2):
        np.testing.assert_allclose(out[b, 0, 0], np.sqrt(beam[b, 0, 0].ravel()))
    assert np.isclose(out[0, 0, 0, 0], 1) and np.isclose(out[0, 0, 0, -1], 0)
    assert np.isclose(out[1, 0, 0, -1], 1)


def test_beam_interp_linear_exact():
    """tests/test_beam_interp_gpu.py:41-64: bilinear is exact for a beam
    linear in za at the midpoints of a decimated grid."""
    za = np.linspace(0, 1, 5)
    az = np.linspace(0, 1, 5)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA])[:, None, None]
    dec = beam[..., ::2, ::2]
    out = O.interp_grid_beam(dec, [2 * (az[1] - az[0])], [2 * (za[1] - za[0])],
                             [0.0], AZ.ravel(), ZA.ravel())
    np.testing.assert_allclose(out[0, 0, 0], np.sqrt(beam[0, 0, 0].ravel()))


def _bspline3_reference(grid, x, y):
    """Explicit cubic B-spline with the mirror prefilter of scipy's
    mode='constant' and zero outside [0, n-1]: the semantics the HIP kernel
    implements, written independently of the oracle."""
    z = math.sqrt(3) - 2

    def filt(c):
        c = c.astype(complex) * 6.0
        n = len(c)
        zn1 = z ** (n - 1)
        c0 = c[0] + zn1 * c[n - 1]
        zi = z
        for i in range(1, n - 1):
            c0 += zi * (c[i] + zn1 * c[n - 1 - i])
            zi *= z
        c[0] = c0 / (1 - zn1 * zn1)
        for i in range(1, n):
            c[i] += z * c[i - 1]
        c[n - 1] = (z * c[n - 2] + c[n - 1]) * z / (z * z - 1)
        for i in range(n - 2, -1, -1):
            c[i] = z * (c[i + 1] - c[i])
        return c

    coef = np.apply_along_axis(filt, 1, grid)
    coef = np.apply_along_axis(filt, 0, coef)

    def w(t):
        t = abs(t)
        return 2 / 3 - t * t + t**3 / 2 if t < 1 else (
            (2 - t) ** 3 / 6 if t < 2 else 0.0)

    def fold(i, n):
        p = 2 * (n - 1)
        i = abs(i) % p
        return p - i if i >= n else i

    nx, ny = grid.shape
    out = np.zeros(len(x), complex)
    for k, (xx, yy) in enumerate(zip(x, y)):
        if not (0 <= xx <= nx - 1 and 0 <= yy <= ny - 1):
            continue
        x0, y0 = int(np.floor(xx)) - 1, int(np.floor(yy)) - 1
        for i in range(4):
            for j in range(4):
                out[k] += (coef[fold(x0 + i, nx), fold(y0 + j, ny)]
                           * w(xx - x0 - i) * w(yy - y0 - j))
    return out


def test_bicubic_semantics_match_scipy(rng):
    g = rng.normal(size=(7, 11)) + 1j * rng.normal(size=(7, 11))
    x = rng.uniform(-0.5, 6.5, 300)
    y = rng.uniform(-0.5, 10.5, 300)
    x[:4] = [0, 6, 0, 6]
    y[:4] = [0, 0, 10, 10]
    ref = ndimage.map_coordinates(g, [x, y], order=3)
    np.testing.assert_allclose(_bspline3_reference(g, x, y), ref, rtol=0,
                               atol=1e-12)


def test_power_beam_negative_takes_imaginary_root():
    """gpu/beams.py:195-197 square-roots in the complex output buffer."""
    out = O.interp_grid_beam(np.full((1, 1, 1, 2, 2), -4.0), [1.0], [1.0],
                             [0.0], [0.5], [0.5])
    assert out[0, 0, 0, 0] == 2j


# ------------------------------------------------------------------ (c)


def test_exptau_zenith_is_one_and_closed_form():
    antpos = np.array([[10.0, -3.0, 0.0], [123.0, 45.0, 0.0]])
    crd = np.array([[0.0], [0.0], [1.0]])
    assert np.all(O.exptau(antpos, 150e6, crd) == 1.0)
    s = np.array([[0.3], [-0.2], [math.sqrt(1 - 0.13)]])
    got = O.exptau(antpos, 150e6, s)
    for a in range(2):
        ph = 2 * math.pi * 150e6 * (antpos[a] @ s[:, 0]) / 299792458.0
        assert abs(got[a, 0] - complex(math.cos(ph), math.sin(ph))) < 1e-12


def test_zmatrix_beam_idx_equals_broadcast(rng):
    """tests/test_matvis_cpu.py:61-74: per-antenna identical beams through
    beam_idx equal one shared beam, bit for bit."""
    nant, nsrc = 4, 9
    et = np.exp(1j * rng.uniform(0, 6, (nant, nsrc)))
    fl = rng.uniform(size=nsrc)
    beam = rng.normal(size=(1, 2, 2, nsrc)) + 1j * rng.normal(size=(1, 2, 2, nsrc))
    z1 = O.zmatrix(fl, beam, et, None, 2, 2)
    z2 = O.zmatrix(fl, np.repeat(beam, 3, 0), et, np.array([0, 2, 1, 0]), 2, 2)
    assert np.array_equal(z1, z2)


# ------------------------------------------------------------------ (d)


@pytest.mark.parametrize("nfeed", [1, 2])
@pytest.mark.parametrize("antpairs", [True, False])
def test_vis_from_z_known_answer(nfeed, antpairs):
    """tests/test_matprod.py:14-71 with z = arange + 1j, checked against an
    explicit loop over (i, j, p, q, k)."""
    nant, nsrc = 5, 15
    z = np.arange(nfeed * nant * nsrc, dtype=complex).reshape(nfeed * nant, nsrc)
    z = z + 1j
    pairs = (np.array([(i, j) for i in range(nant) for j in range(nant)])
             if antpairs else None)
    got = O.vis_from_z(z, nant, nfeed, pairs)
    want = np.zeros((nant * nant, nfeed, nfeed), complex)
    for i in range(nant):
        for j in range(nant):
            for p in range(nfeed):
                for q in range(nfeed):
                    acc = 0j
                    for k in range(nsrc):
                        acc += np.conj(z[i * nfeed + q, k]) * z[j * nfeed + p, k]
                    want[i * nant + j, p, q] = acc
    np.testing.assert_allclose(got, want, rtol=1e-15)


def test_zdotz_known_answer():
    """tests/test_cublas.py:13-25."""
    a = np.arange(6).reshape((2, 3)).astype(complex) + 1j
    v = O.vis_from_z(a, 2, 1)[:, 0, 0].reshape(2, 2)
    np.testing.assert_allclose(v, np.dot(a.conj(), a.T))


# ------------------------------------------------------------------ loop


def test_simulate_antizenith_is_zero():
    """tests/test_matvis_gpu.py:15-25: one source below the horizon at every
    time gives exact zeros."""
    p = small_problem(nant=3, nsrc=1, ntimes=2)
    lat = np.radians(-30.72)
    # the south celestial pole region never rises at dec = -lat (north) ...
    eci = O.point_source_crd_eq([0.0], [np.pi / 2 + lat - 0.2])
    v = O.simulate(antpos=p["antpos"], freq=p["freq"], eci=eci,
                   rotations=p["rotations"], I_sky=np.array([2.0]),
                   beam=p["oracle_beam"])
    assert np.all(v == 0)


def test_simulate_chunking_invariance():
    """tests/test_coordrot.py:147-167 in spirit: chunk count does not change
    the result beyond rounding."""
    p = small_problem(nant=5, nsrc=257, ntimes=2)
    kw = dict(antpos=p["antpos"], freq=p["freq"], eci=p["eci"],
              rotations=p["rotations"], I_sky=p["I_sky"], beam=p["oracle_beam"])
    v1 = O.simulate(nchunks=1, **kw)
    v3 = O.simulate(nchunks=3, **kw)
    np.testing.assert_allclose(v3, v1, rtol=0, atol=1e-12 * np.abs(v1).max())


def test_simulate_brute_force_tiny():
    """Two antennas, polarized Gaussian beam, explicit per-source sum with
    complex exponentials from the math module."""
    p = small_problem(nant=2, nsrc=40, ntimes=1, polarized=True)
    v = O.simulate(antpos=p["antpos"], freq=p["freq"], eci=p["eci"],
                   rotations=p["rotations"], I_sky=p["I_sky"],
                   beam=p["oracle_beam"], polarized=True)
    r = p["rotations"][0]
    sig = p["oracle_beam"]["sigma"][0]
    want = np.zeros((4, 2, 2), complex)
    for s in range(40):
        top = r @ p["eci"][:, s]
        if not top[2] > 0:
            continue
        za = math.pi / 2 - math.asin(math.sqrt(max(0.0, 1 - top[0] ** 2 - top[1] ** 2)))
        amp = math.exp(-za * za / (2 * sig * sig)) / math.sqrt(2)
        f = math.sqrt(0.5 * p["I_sky"][s])
        zz = []
        for a in range(2):
            ph = 2 * math.pi * p["freq"] * (p["antpos"][a] @ top) / 299792458.0
            zz.append(amp * f * complex(math.cos(ph), math.sin(ph)))
        for i in range(2):
            for j in range(2):
                # every (feed, axis) carries the same amplitude: V sums nax=2
                want[i * 2 + j] += 2 * (zz[i].conjugate() * zz[j])
    np.testing.assert_allclose(v[0], want, rtol=1e-12, atol=1e-14)


def test_normalized_error_detects_a_planted_mistake(rng):
    # a grid beam whose feeds differ, so a feed transpose changes V
    p = small_problem(nant=4, nsrc=200, ntimes=1, polarized=True, beam="grid")
    v = O.simulate(antpos=p["antpos"], freq=p["freq"], eci=p["eci"],
                   rotations=p["rotations"], I_sky=p["I_sky"],
                   beam=p["oracle_beam"], polarized=True)
    e0 = O.normalized_error(v, v, 4, 2)
    assert e0[0] == 0.0
    bad = v.copy()
    bad[0, 5, 1, 0] *= 1 + 1e-6
    e1 = O.normalized_error(bad, v, 4, 2)
    assert 1e-9 < e1[0] < 1e-5
    # the transposed feed layout is a real mistake the metric must flag
    e2 = O.normalized_error(v.transpose(0, 1, 3, 2), v, 4, 2)
    assert e2[0] > 1e-3


# ---------------------------------------------------------------- beam grid
# preparation: azimuth seam and frequency step

def _trig_beam(naz=72, nza=19):
    """A beam that is an exact trigonometric polynomial in azimuth, smooth
    in zenith angle, on a grid covering the circle (5-degree spacing)."""
    az = np.radians(np.arange(naz) * 360.0 / naz)
    za = np.radians(np.linspace(0, 90, nza))

    def f(a, z):
        return ((1 + 0.5 * np.cos(a) + 0.3 * np.sin(2 * a)
                 + 0.2 * np.cos(3 * a)) * (1 + 0.3 * np.cos(2 * z)))

    AZ, ZA = np.meshgrid(az, za)
    data = np.broadcast_to(f(AZ, ZA).astype(complex), (2, 2, 1, za.size,
                                                       az.size)).copy()
    # bilinear error bound h^2/8 (max|f_az az| + max|f_za za|), h = 5 deg
    h = np.radians(5.0)
    bound = h * h / 8 * ((0.5 + 0.3 * 4 + 0.2 * 9) * 1.3 + 1.2 * 2.0)
    # max |df/daz|
    slope = (0.5 + 0.6 + 0.6) * 1.3
    return az, za, data, f, bound, slope


def _seam_errors(planes, daz, dza, azmin, f, order):
    rng = np.random.default_rng(3)
    h = np.radians(5.0)
    pa = np.mod(2 * np.pi - 4 * h + np.linspace(0, 8 * h, 401), 2 * np.pi)
    pz = rng.uniform(np.radians(20), np.radians(70), pa.size)
    s = O.interp_grid_beam(planes[None], [daz], [dza], [azmin], pa, pz,
                           order)[0, 0, 0].real
    eps = 1e-9
    c = O.interp_grid_beam(planes[None], [daz], [dza], [azmin],
                           np.array([2 * np.pi - eps, eps]),
                           np.array([0.7, 0.7]), order)[0, 0, 0].real
    return np.abs(s - f(pa, pz)).max(), abs(c[0] - c[1]), eps


@pytest.mark.parametrize("order", [1, 3])
def test_prepare_grid_azimuth_seam(order):
    """Interpolation across az = 0 / 2 pi: continuous at the seam (within
    the function's own change over 2e-9 rad) and within the bilinear error
    bound of the true beam on both sides (the bicubic spline is inside it
    too).  A grid without the wrap, or with azmin left at the first
    azimuth, fails both checks by orders of magnitude."""
    az, za, data, f, bound, slope = _trig_beam()
    planes, daz, dza, azmin = O.prepare_grid(data, az, za, [150e6], 150e6,
                                             True)
    err, jump, eps = _seam_errors(planes, daz, dza, azmin, f, order)
    assert err < bound, err
    assert jump < 2 * eps * slope + 1e-12, jump
    # planted: no wrap (the unpadded grid at its own azimuth origin)
    e2, j2, _ = _seam_errors(planes[..., 3:-3], daz, dza, az.min(), f, order)
    assert e2 > 100 * bound and j2 > 0.1, (e2, j2)
    # planted: padded grid, azmin not moved below the first azimuth
    e3, _, _ = _seam_errors(planes, daz, dza, az.min(), f, order)
    assert e3 > 10 * bound, e3


def _freq_cubic_beam(nodes_mhz):
    """E-field data cubic in frequency: A + B t + C t^2 + D t^3 with
    t = (nu - 150 MHz) / 50 MHz, on a small grid."""
    rng = np.random.default_rng(11)
    shape = (2, 2, 5, 8)
    coef = [rng.normal(size=shape) + 1j * rng.normal(size=shape)
            for _ in range(4)]

    def at(nu):
        t = (nu - 150e6) / 50e6
        return coef[0] + coef[1] * t + coef[2] * t ** 2 + coef[3] * t ** 3

    nodes = np.asarray(nodes_mhz, float) * 1e6
    data = np.stack([at(nu) for nu in nodes], axis=2)
    return data, nodes, at


def test_beam_at_frequency_reproduces_cubics():
    """Off the nodes, a beam cubic in frequency comes back exactly (a
    not-a-knot cubic spline reproduces cubics); linear interpolation
    between the bracketing nodes would miss it by ~1e-2.  Within 1 Hz of a
    node: that node's data."""
    data, nodes, at = _freq_cubic_beam([100, 120, 140, 160, 180, 200])
    for nu in (103.7e6, 133.3e6, 151.0e6, 199.2e6):
        got = O.beam_at_frequency(data, nodes, nu)
        np.testing.assert_allclose(got, at(nu), rtol=1e-12,
                                   atol=1e-12 * np.abs(at(nu)).max())
        k = np.searchsorted(nodes, nu) - 1
        w = (nu - nodes[k]) / (nodes[k + 1] - nodes[k])
        lin = (1 - w) * data[:, :, k] + w * data[:, :, k + 1]
        assert np.abs(lin - at(nu)).max() > 1e-3
    np.testing.assert_array_equal(O.beam_at_frequency(data, nodes,
                                                      140e6 + 0.4),
                                  data[:, :, 2])
    with pytest.raises(ValueError):
        O.beam_at_frequency(data, nodes, 99e6)
    with pytest.raises(ValueError):
        O.beam_at_frequency(data[:, :, :3], nodes[:3], 110e6)


def test_prepare_grid_power_before_frequency_step():
    """Unpolarized: the E-field becomes power at every node first, then the
    frequency step (core/beams.py:65-66 before :82-93).  For an E-field
    cubic in frequency the two orders differ; the planes follow the
    reference's."""
    from scipy.interpolate import interp1d
    data, nodes, at = _freq_cubic_beam([100, 120, 140, 160, 180, 200])
    az = np.radians(np.arange(8) * 5.0)  # partial circle: no padding
    za = np.radians(np.linspace(0, 40, 5))
    nu = 133.3e6
    planes, *_ = O.prepare_grid(data, az, za, nodes, nu, False)
    power_nodes = (np.abs(data[:, 0]) ** 2).sum(0)  # (nfreq, nza, naz)
    want = interp1d(nodes, power_nodes, kind="cubic", axis=0)(nu)
    np.testing.assert_allclose(planes[0, 0], want, rtol=1e-13)
    other = (np.abs(at(nu)[:, 0]) ** 2).sum(0)
    assert np.abs(other - want).max() > 1e-6 * np.abs(want).max()


def test_gaussian_beam_matches_airy_half_maximum():
    """The diameter rule of GaussianBeam (sigma = arcsin(2.2150894 lambda /
    (pi D)) x 2 / 2.355) matches the Gaussian's E-field half maximum to the
    Airy pattern's: 2.2150894 is where the Airy E-field 2 J1(x) / x falls to
    1/2, and the oracle's E-field exp(-za^2 / (2 sigma^2)) is 1/2 at that
    angle (to the 1e-4 of 2.355 against 2 sqrt(2 ln 2)).  A sigma taken as
    the power beam's, the intensity half maximum (x = 1.6163) or a missing
    factor 2 fails."""
    from scipy.optimize import brentq
    from scipy.special import j1

    from matvis_amd.beams import GaussianBeam

    x_half = brentq(lambda x: 2.0 * j1(x) / x - 0.5, 1.0, 3.0, xtol=1e-14)
    assert abs(x_half - 2.2150894) < 1e-6
    for D, f in ((14.0, 150e6), (8.0, 100e6), (14.0, 200e6)):
        lam = 299792458.0 / f
        theta = np.arcsin(x_half * lam / (np.pi * D))
        sigma = GaussianBeam(diameter=D).get_sigma(f)
        e = O.gaussian_beam(sigma, np.array([theta]), False)[0, 0, 0, 0].real
        assert abs(e - 0.5) < 2e-4, (D, f, e)
        # the plausible mistakes
        assert abs(np.exp(-theta**2 / sigma**2) - 0.5) > 0.1
        s_int = np.arcsin(1.6163 * lam / (np.pi * D)) * 2 / 2.355
        assert abs(np.exp(-theta**2 / (2 * s_int**2)) - 0.5) > 0.05
        assert abs(np.exp(-theta**2 / (2 * (sigma / 2) ** 2)) - 0.5) > 0.1
