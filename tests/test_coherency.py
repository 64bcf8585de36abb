"""The oracle's polarized-sky coherency rotation, pinned by mathematics the
reference's formulas must satisfy (coordinates.py:288-576,
core/coords.py:123-175): Rodrigues rotations that carry one vector onto
another, the parallactic angle of spherical astronomy, and the rotation of
Stokes Q/U under a change of polarization basis."""

import numpy as np
import pytest

import oracle as O

LAT = np.deg2rad(-30.72)


def _unit(rng, n):
    v = rng.normal(size=(3, n))
    return v / np.linalg.norm(v, axis=0)


def test_vecs2rot_carries_r1_onto_r2():
    rng = np.random.default_rng(0)
    r1, r2 = _unit(rng, 200), _unit(rng, 200)
    rot = O.vecs2rot(r1, r2)
    np.testing.assert_allclose(np.einsum("abn,bn->an", rot, r1), r2,
                               atol=1e-13)
    for k in range(200):
        np.testing.assert_allclose(rot[..., k].T @ rot[..., k], np.eye(3),
                                   atol=1e-13)
        assert abs(np.linalg.det(rot[..., k]) - 1.0) < 1e-13


def test_quarter_turns_about_the_axes():
    x, y, z = np.eye(3)
    # right-handed: x -> y is +90 degrees about z, y -> z about x
    want_z = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float)
    want_x = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], float)
    np.testing.assert_allclose(O.vecs2rot(x[:, None], y[:, None])[..., 0],
                               want_z, atol=1e-15)
    np.testing.assert_allclose(O.vecs2rot(y[:, None], z[:, None])[..., 0],
                               want_x, atol=1e-15)
    np.testing.assert_allclose(
        O.axis_angle_rotation(z[:, None], np.array([np.pi / 2]))[..., 0],
        want_z, atol=1e-15)


def test_parallel_vectors_give_the_identity():
    v = _unit(np.random.default_rng(1), 5)
    rot = O.vecs2rot(v, v.copy())
    assert np.all(np.isfinite(rot))
    np.testing.assert_allclose(rot, np.repeat(np.eye(3)[..., None], 5, -1),
                               atol=1e-15)


def _rigid_sky(lst, n=4000, seed=2):
    """Sources above the horizon of a rigid Earth at local sidereal angle
    lst: (ra, dec, alt, az, R_avg ICRS -> (N, E, U))."""
    rng = np.random.default_rng(seed)
    ra = rng.uniform(0, 2 * np.pi, n)
    dec = np.arcsin(rng.uniform(-1, 1, n))
    enu = O.eci_to_enu_matrix(lst, LAT)
    topo = enu @ O.point_source_crd_eq(ra, dec)
    up = topo[2] > 0
    az, za = O.enu_to_az_za(topo[0, up], topo[1, up])
    return ra[up], dec[up], 0.5 * np.pi - za, az, enu[[1, 0, 2]]


@pytest.mark.parametrize("lst", [0.0, 1.3, 4.0])
def test_angle_is_minus_the_parallactic_angle(lst):
    """For a rigid Earth the alt/az theta_hat (away from the zenith) makes
    with the equatorial theta_hat (away from the north pole) the angle
    between the directions to the zenith and to the pole: the parallactic
    angle q, positive west of the meridian, tan q = sin H / (tan(lat)
    cos(dec) - sin(dec) cos H).  The zenith direction is cos q (north) +
    sin q (east) on the sky, while X is measured from theta_hat (south)
    towards phi_hat (east): X = -q."""
    ra, dec, alt, az, r_avg = _rigid_sky(lst)
    rot = O.coherency_rotation(ra, dec, alt, az, r_avg)
    h = lst - ra
    q = np.arctan2(np.sin(h), np.tan(LAT) * np.cos(dec)
                   - np.sin(dec) * np.cos(h))
    # away from the poles of both frames, where the angle is defined
    ok = (np.abs(alt) < np.deg2rad(89.9)) & (np.abs(dec) < np.deg2rad(89.9))
    np.testing.assert_allclose(rot[0, 0][ok], np.cos(q)[ok], atol=1e-9)
    np.testing.assert_allclose(rot[0, 1][ok], -np.sin(q)[ok], atol=1e-9)
    np.testing.assert_allclose(rot[1, 0], -rot[0, 1], atol=0)
    np.testing.assert_allclose(rot[0, 0] ** 2 + rot[0, 1] ** 2, 1.0,
                               atol=1e-12)


def test_alignment_rotation_absorbs_a_perturbed_direction():
    """A small per-source offset between R_avg . frame and the observed
    direction leaves the angle continuous: it changes by about the offset."""
    ra, dec, alt, az, r_avg = _rigid_sky(0.7, n=500)
    base = O.coherency_rotation(ra, dec, alt, az, r_avg)
    moved = O.coherency_rotation(ra, dec, alt + 1e-5, az + 1e-5, r_avg)
    x0 = np.arctan2(base[0, 1], base[0, 0])
    x1 = np.arctan2(moved[0, 1], moved[0, 0])
    ok = alt < np.deg2rad(80)
    assert np.max(np.abs(np.angle(np.exp(1j * (x1 - x0)))[ok])) < 1e-3


def test_coherency_rotation_is_a_stokes_rotation():
    """With C = [[I + Q, U - iV], [U + iV, I - Q]] / 2 and rot = [[c, s],
    [-s, c]], rot^T C rot keeps I and V and turns (Q, U) by 2X:
    Q' = Q cos 2X - U sin 2X, U' = Q sin 2X + U cos 2X."""
    rng = np.random.default_rng(3)
    n, nf = 50, 3
    i_, q_, u_, v_ = (rng.normal(size=(n, nf)) for _ in range(4))
    coh = 0.5 * np.array([[i_ + q_, u_ - 1j * v_], [u_ + 1j * v_, i_ - q_]])
    coh = coh.transpose(2, 3, 0, 1)  # (n, nf, 2, 2)
    x = rng.uniform(-np.pi, np.pi, n)
    rot = np.array([[np.cos(x), np.sin(x)], [-np.sin(x), np.cos(x)]])
    out = O.rotate_coherency(coh, rot)
    c2, s2 = np.cos(2 * x)[:, None], np.sin(2 * x)[:, None]
    i2 = (out[..., 0, 0] + out[..., 1, 1]).real
    q2 = (out[..., 0, 0] - out[..., 1, 1]).real
    u2 = (out[..., 0, 1] + out[..., 1, 0]).real
    v2 = (out[..., 1, 0] - out[..., 0, 1]).imag
    np.testing.assert_allclose(i2, i_, atol=1e-13)
    np.testing.assert_allclose(v2, v_, atol=1e-13)
    np.testing.assert_allclose(q2, q_ * c2 - u_ * s2, atol=1e-13)
    np.testing.assert_allclose(u2, q_ * s2 + u_ * c2, atol=1e-13)


def test_select_chunk_polarized_compacts_like_select_chunk():
    rng = np.random.default_rng(4)
    n, nf, chunk = 3000, 2, 1000
    ra = rng.uniform(0, 2 * np.pi, n)
    dec = np.arcsin(rng.uniform(-1, 1, n))
    enu = O.eci_to_enu_matrix(0.4, LAT)
    topo = enu @ O.point_source_crd_eq(ra, dec)
    flux = rng.uniform(1, 2, n)
    coh = np.zeros((n, nf, 2, 2))
    coh[:, :, 0, 0] = coh[:, :, 1, 1] = flux[:, None] / 2
    for c in range(3):
        crd, f, k, idx = O.select_chunk(topo, flux, c, chunk, 700)
        crd2, coh2, k2, idx2 = O.select_chunk_polarized(
            topo, coh, ra, dec, enu[[1, 0, 2]], c, chunk, 700)
        assert k == k2
        np.testing.assert_array_equal(idx, idx2)
        np.testing.assert_array_equal(crd, crd2)
        # unpolarized coherency is invariant under the basis rotation
        np.testing.assert_allclose(coh2[:, :, 0, 0],
                                   np.repeat(f[:, None] / 2, nf, 1),
                                   rtol=1e-14, atol=0)
        np.testing.assert_allclose(coh2[:, :, 0, 1], 0, atol=1e-14)
        assert not coh2[k:].any()
