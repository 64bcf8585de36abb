"""Golden vectors (tests/golden/golden_vectors.npz, written by
tests/golden/make_golden.py from closed forms, explicit loops, scipy and the
reference's own beam file): the oracle and the GPU path against them."""

from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLD = np.load(Path(__file__).parent / "golden" / "golden_vectors.npz")
FITS = Path("/root/reference/src/matvis/data/NF_HERA_Dipole_small.fits")
has_fits = pytest.mark.skipif("fits_data" not in GOLD.files,
                              reason="fixture written without the FITS file")


# ---------------------------------------------------------------- oracle


def test_oracle_cardinal_points():
    got = O.point_source_crd_eq(GOLD["crd_ra"], GOLD["crd_dec"])
    np.testing.assert_allclose(got, GOLD["crd_want"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("nf", [1, 2])
def test_oracle_matprod_known_answer(nf):
    z = GOLD[f"matprod_nf{nf}_z"]
    got = O.vis_from_z(z, 5, nf)
    np.testing.assert_array_equal(got, GOLD[f"matprod_nf{nf}_want"])


@pytest.mark.parametrize("case", ["identity", "linear"])
def test_oracle_beam_grids(case):
    g = lambda k: GOLD[f"beam_{case}_{k}"]  # noqa: E731
    planes = g("planes")
    got = O.interp_grid_beam(planes, g("daz"), g("dza"),
                             np.zeros(planes.shape[0]), g("az"), g("za"),
                             order=1)
    np.testing.assert_allclose(got[:, 0, 0], g("want"), rtol=1e-12,
                               atol=1e-14)


def test_oracle_tiny_simulation():
    topo = GOLD["tiny_topo"]
    flux = GOLD["tiny_flux"]
    et = O.exptau(GOLD["tiny_antpos"], float(GOLD["tiny_freq"]), topo)
    beam = np.ones((1, 1, 1, topo.shape[1]), np.complex128)
    z = O.zmatrix(np.sqrt(0.5 * flux), beam, et, None, 1, 1)
    got = O.vis_from_z(z, 3, 1).reshape(3, 3)
    want = GOLD["tiny_want"]
    np.testing.assert_allclose(got, want, rtol=0,
                               atol=1e-13 * np.abs(want).max())


@has_fits
@pytest.mark.parametrize("order", [1, 3])
def test_oracle_fits_beam_interpolation(order):
    planes, daz, dza, azmin = O.prepare_grid(
        GOLD["fits_data"], GOLD["fits_az"], GOLD["fits_za"],
        GOLD["fits_freq"], float(GOLD["fits_freq"][0]), True)
    got = O.interp_grid_beam(planes[None], [daz], [dza], [azmin],
                             GOLD["fits_pt_az"], GOLD["fits_pt_za"], order)
    np.testing.assert_allclose(got[0], GOLD[f"fits_want_order{order}"],
                               rtol=0, atol=1e-13)


@pytest.mark.skipif(not FITS.exists(), reason="reference beam file absent")
@has_fits
def test_beamfits_reader_reproduces_fixture():
    from matvis_amd.beamfits import read_beamfits

    b = read_beamfits(FITS)
    assert b.beam_type == "efield" and b.feed_array == ["x", "y"]
    assert b.data_array.shape == (2, 2, 2, 91, 180)
    np.testing.assert_array_equal(b.data_array[:, :, :1].astype(np.complex64),
                                  GOLD["fits_data"])
    np.testing.assert_array_equal(b.axis1_array, GOLD["fits_az"])
    np.testing.assert_array_equal(b.axis2_array, GOLD["fits_za"])
    np.testing.assert_array_equal(b.freq_array, [100e6, 101e6])


# ---------------------------------------------------------------- GPU


def _fits_gridbeam():
    from matvis_amd.beams import GridBeam

    return GridBeam(GOLD["fits_data"].astype(np.complex128), GOLD["fits_az"],
                    GOLD["fits_za"], freq_array=GOLD["fits_freq"])


@pytest.mark.gpu
@has_fits
@pytest.mark.parametrize("order", [1, 3])
def test_gpu_fits_beam_interpolation(order):
    """The product's grid preparation + prefilter + interpolation kernel on
    the reference's beam, at points spread over the sky."""
    import matvis_amd._device as D
    from matvis_amd._lib import call
    from matvis_amd.beams import grid_planes

    planes, daz, dza, azmin, is_power = grid_planes(
        _fits_gridbeam(), float(GOLD["fits_freq"][0]), True)
    nax, nfeed, nza, naz = planes.shape
    print(f"fits beam order={order} grid {nza}x{naz}", flush=True)
    pl = D.to_dev(planes[None].astype(np.complex128))
    if order == 3:
        call("mv_beam_prefilter", D.ptr(pl), 1, nax * nfeed, nza, naz,
             D.stream_ptr())
    n = GOLD["fits_pt_az"].size
    grid = D.to_dev(np.array([[daz, dza, azmin]]), np.float64)
    az = D.to_dev(GOLD["fits_pt_az"], np.float64)
    za = D.to_dev(GOLD["fits_pt_za"], np.float64)
    out = D.zeros((1, nfeed, nax, n), np.complex128)
    call("mv_beam_interp_azza", D.ptr(pl), 1, 1, nfeed, nax, nza, naz,
         D.ptr(grid), order, int(is_power), D.ptr(az), D.ptr(za), n,
         D.ptr(out), n, D.stream_ptr())
    got = D.to_host(out)[0]
    np.testing.assert_allclose(got, GOLD[f"fits_want_order{order}"], rtol=0,
                               atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("nf", [1, 2])
@pytest.mark.parametrize("method", ["GPUMatMul", "GPUVectorDot"])
def test_gpu_matprod_known_answer(nf, method):
    import matvis_amd._device as D
    from matvis_amd.gpu import matprod

    obj = getattr(matprod, method)(nchunks=1, nfeed=nf, nant=5, antpairs=None,
                                   precision=2)
    obj.setup()
    obj(D.to_dev(GOLD[f"matprod_nf{nf}_z"], np.complex128), chunk=0)
    out = np.zeros((25, nf, nf), np.complex128)
    obj.sum_chunks(out)
    np.testing.assert_array_equal(out, GOLD[f"matprod_nf{nf}_want"])


@pytest.mark.gpu
def test_gpu_tiny_simulation():
    """simulate() end to end with R = identity (so ECI = topocentric) and a
    constant power beam against the explicit-loop visibilities."""
    from matvis_amd import astro
    from matvis_amd.beams import GridBeam
    from matvis_amd.gpu.gpu import simulate

    topo = GOLD["tiny_topo"]
    ra = np.arctan2(topo[1], topo[0])
    dec = np.arcsin(np.clip(topo[2], -1, 1))
    az = np.radians(np.arange(36) * 10.0)
    za = np.radians(np.linspace(0, 90, 10))
    ones = np.ones((1, 1, 1, za.size, az.size))
    beam = GridBeam(ones, az, za, beam_type="power")
    v = simulate(antpos=GOLD["tiny_antpos"], freq=float(GOLD["tiny_freq"]),
                 times=astro.Time(np.array([2459863.0]), format="jd"),
                 skycoords=astro.SkyCoord(ra, dec),
                 telescope_loc=astro.EarthLocation.from_geodetic(21.4, -30.7,
                                                                 1073.0),
                 I_sky=GOLD["tiny_flux"], beam_list=[beam], polarized=False,
                 coord_method="CoordinateRotationMatrix",
                 coord_method_params={"rotations": np.eye(3)[None]},
                 precision=2)
    want = GOLD["tiny_want"]
    np.testing.assert_allclose(v[0].reshape(3, 3), want, rtol=0,
                               atol=1e-13 * np.abs(want).max())
