"""Host-side logic: chunk model, beam preparation, astrometry, synthetic
inputs, plugin argument validation (no GPU needed)."""

import math

import numpy as np
import pytest

from matvis_amd import astro, synth
from matvis_amd._utils import get_desired_chunks, get_required_chunks
from matvis_amd.beams import BeamInterface, GaussianBeam, GridBeam, grid_planes
from matvis_amd.core.beams import _wrangle_beams


def test_required_chunks_monotone_and_validated():
    """tests/test_utils.py:20-82."""
    args = (2, 2, 350, 100_000, 1, 360 * 91, 2)
    big = get_required_chunks(np.inf, *args)
    assert big == 1
    counts = [get_required_chunks(mem, *args) for mem in (2**34, 2**31, 2**29)]
    assert counts == sorted(counts)
    for bad in (1.5, 1.1, 0.0, -0.5):
        with pytest.raises(ValueError, match=r"memory_buffer must be in the range"):
            get_required_chunks(1e9, *args, memory_buffer=bad)
    # memory_buffer = 1 is valid; a smaller buffer never needs fewer chunks;
    # the count is clamped to 1 when everything fits
    small = (2, 2, 4, 100, 2, 0, 1)
    assert get_required_chunks(10 * 1024**3, *small, memory_buffer=1.0) >= 1
    assert (get_required_chunks(2**20, *args, memory_buffer=0.1)
            >= get_required_chunks(2**20, *args, memory_buffer=0.9))
    assert get_required_chunks(1e18, *small, memory_buffer=1.0) == 1
    n, per = get_desired_chunks(np.inf, 3, [], 1, 1, 8, 10, 2)
    assert (n, per) == (3, 4)
    n, per = get_desired_chunks(np.inf, 50, [], 1, 1, 8, 10, 2)
    assert (n, per) == (10, 1)


def _grid_beam(naz=36, nza=10, full_circle=True):
    az = np.radians(np.arange(naz) * (360.0 / naz if full_circle else 5.0))
    za = np.radians(np.linspace(0, 90, nza))
    rng = np.random.default_rng(0)
    data = rng.normal(size=(2, 2, 1, nza, naz)) + 1j * rng.normal(size=(2, 2, 1, nza, naz))
    return GridBeam(data, az, za)


def test_grid_planes_wraps_azimuth_and_converts_power():
    b = _grid_beam()
    planes, daz, dza, azmin, is_power = grid_planes(b, 150e6, polarized=True)
    # full circle: 3 wrapped columns on each side (pyuvdata's padding)
    assert planes.shape == (2, 2, 10, 42) and not is_power
    d = b.data_array[:2, :2, 0]
    np.testing.assert_array_equal(planes[..., :3], d[..., -3:])
    np.testing.assert_array_equal(planes[..., 3:39], d)
    np.testing.assert_array_equal(planes[..., -3:], d[..., :3])
    assert math.isclose(daz, math.radians(10))
    assert math.isclose(azmin, -3 * math.radians(10))
    planes, *_, is_power = grid_planes(b, 150e6, polarized=False)
    e = b.data_array[:, 0, 0]
    np.testing.assert_allclose(planes[0, 0, :, 3:39], (np.abs(e) ** 2).sum(0))
    assert is_power and planes.shape == (1, 1, 10, 42)


@pytest.mark.parametrize("polarized", [True, False])
@pytest.mark.parametrize("naz,full", [(36, True), (20, False), (180, True)])
def test_grid_planes_matches_oracle_preparation(polarized, naz, full):
    """The product's grid preparation equals the oracle's restatement of
    _wrangle_beams + prepare_for_map_coords, frequency interpolation
    included."""
    import oracle as O
    az = np.radians(np.arange(naz) * (360.0 / naz if full else 5.0))
    if naz == 180:  # the spacing of the reference's NF_HERA_Dipole_small.fits
        az = np.radians(np.arange(naz) * 2.005586592178771)
    za = np.radians(np.linspace(0, 90, 10))
    rng = np.random.default_rng(naz)
    freqs = np.array([100e6, 130e6, 150e6, 180e6, 200e6])
    shape = (2, 2, freqs.size, za.size, az.size)
    data = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    b = GridBeam(data, az, za, freq_array=freqs)
    for f in (100e6, 137.5e6, 150e6 + 0.5, 200e6):
        planes, daz, dza, azmin, _ = grid_planes(b, f, polarized)
        want, wdaz, wdza, wazmin = O.prepare_grid(data, az, za, freqs, f,
                                                  polarized)
        np.testing.assert_allclose(planes, want, rtol=1e-15, atol=0)
        assert (daz, dza, azmin) == (wdaz, wdza, wazmin)


def test_grid_planes_partial_azimuth_not_wrapped():
    b = _grid_beam(naz=20, full_circle=False)
    planes, *_ = grid_planes(b, 150e6, polarized=True)
    assert planes.shape[-1] == 20


def test_grid_planes_rejects_bad_grids():
    b = _grid_beam()
    b.pixel_coordinate_system = "healpix"
    with pytest.raises(ValueError, match='pixel coordinate system must be "az_za"'):
        grid_planes(b, 150e6, True)
    b = _grid_beam()
    b.axis2_array = b.axis2_array + 0.1
    with pytest.raises(ValueError, match="must start at 0"):
        grid_planes(b, 150e6, True)


def test_grid_planes_frequency_interpolation():
    """Off the nodes, the cubic frequency step reproduces a beam cubic in
    frequency; a node within 1 Hz gives its data; outside the range, or
    between fewer than 4 nodes, it raises; one node is used at every
    frequency."""
    az = np.radians(np.arange(8) * 45.0)
    za = np.radians([0.0, 45.0, 90.0])
    nodes = np.array([100e6, 125e6, 150e6, 175e6, 200e6])
    rng = np.random.default_rng(2)
    coef = rng.normal(size=(4, 2, 2, 3, 8))

    def at(nu):
        t = (nu - 150e6) / 50e6
        return coef[0] + coef[1] * t + coef[2] * t ** 2 + coef[3] * t ** 3

    data = np.stack([at(nu) for nu in nodes], axis=2).astype(complex)
    b = GridBeam(data, az, za, freq_array=nodes)
    planes, *_ = grid_planes(b, 140e6, True)
    np.testing.assert_allclose(planes[..., 3:-3], at(140e6), rtol=1e-12,
                               atol=1e-12)
    planes, *_ = grid_planes(b, 125e6 + 0.3, True)
    np.testing.assert_array_equal(planes[..., 3:-3], data[:, :, 1])
    with pytest.raises(ValueError, match="outside the beam's range"):
        grid_planes(b, 250e6, True)
    b3 = GridBeam(data[:, :, :3], az, za, freq_array=nodes[:3])
    with pytest.raises(ValueError, match="at least 4"):
        grid_planes(b3, 110e6, True)
    b1 = GridBeam(data[:, :, :1], az, za, freq_array=nodes[:1])
    planes, *_ = grid_planes(b1, 190e6, True)
    np.testing.assert_array_equal(planes[..., 3:-3], data[:, :, 0])


def test_wrangle_beams_errors():
    """tests/test_beams.py:58-88."""
    g = GaussianBeam(sigma=0.2)
    kw = dict(beam_idx=None, beam_list=[g], polarized=True, nant=3, freq=1e8)
    with pytest.raises(ValueError, match="beam_idx must be provided"):
        _wrangle_beams(**(kw | {"beam_list": [g, g]}))
    with pytest.raises(ValueError, match="beam_idx must be length nant"):
        _wrangle_beams(**(kw | {"beam_idx": np.zeros(2, int)}))
    with pytest.raises(ValueError, match="beam_idx contains indices greater"):
        _wrangle_beams(**(kw | {"beam_idx": np.array([0, 1, 1])}))
    with pytest.raises(ValueError, match="beam type must be efield"):
        _wrangle_beams(**(kw | {"beam_list": [BeamInterface(g, "power")]}))
    bl, nb, bi = _wrangle_beams(**(kw | {"polarized": False}))
    assert nb == 1 and bl[0].beam_type == "power"


def test_gaussian_beam_sigma():
    assert GaussianBeam(sigma=0.1).get_sigma(1e8) == 0.1
    b = GaussianBeam(sigma=0.1, spectral_index=-1.0, reference_frequency=1e8)
    assert math.isclose(b.get_sigma(2e8), 0.05)
    with pytest.raises(ValueError):
        GaussianBeam()


def test_gmst_at_j2000_and_rotation_properties():
    # GMST at 2000-01-01 12:00 UT1 is 18h41m50.54841s
    g = astro.gmst_iau1982(astro.J2000)
    assert abs(g - math.radians(280.46061837)) < 1e-9
    # one sidereal day later GMST is back to the same value
    g2 = astro.gmst_iau1982(astro.J2000 + 0.99726956633)
    assert abs(math.remainder(g2 - g, 2 * math.pi)) < 1e-6
    lat, lon = math.radians(-30.72), math.radians(21.43)
    r = astro.eci_to_enu_rotation(2459863.3, lat, lon)
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-15)
    assert abs(np.linalg.det(r) - 1) < 1e-14
    # at J2000 the precession matrix is the identity
    np.testing.assert_allclose(astro.precession_iau1976(astro.J2000), np.eye(3),
                               atol=1e-15)
    # the south celestial pole of date (J2000 vector: (N P)^T (0, 0, -1)) is
    # due south at altitude -lat: ENU (0, cos lat ... ) with z = sin(-lat)
    jd_tt = 2459863.3 + astro.TT_MINUS_UTC_S / 86400.0
    dpsi, deps, eps0 = astro.nutation_iau1980_leading(jd_tt)
    n_mat = astro._r1(-(eps0 + deps)) @ astro._r3(-dpsi) @ astro._r1(eps0)
    pole = (n_mat @ astro.precession_iau1976(jd_tt)).T @ np.array([0, 0, -1.0])
    enu = r @ pole
    np.testing.assert_allclose(enu, [0.0, -math.cos(lat), math.sin(-lat)],
                               atol=1e-12)
    # the J2000 pole is off by the precession since J2000 (0.127 deg here)
    off = math.degrees(math.acos(np.clip(pole @ [0, 0, -1.0], -1, 1)))
    assert 0.12 < off < 0.14


def test_duck_typed_inputs():
    t = astro.Time([2459863.0, 2459863.1])
    assert astro.times_jd(t).tolist() == [2459863.0, 2459863.1]
    loc = astro.EarthLocation.from_geodetic(21.43, -30.72, 1000.0)
    lat, lon = astro.location_lat_lon(loc)
    assert math.isclose(lat, math.radians(-30.72))
    sc = astro.SkyCoord(np.array([0.1, 0.2]), np.array([-0.3, 0.4]))
    ra, dec = astro.skycoord_ra_dec(sc)
    assert ra.tolist() == [0.1, 0.2] and len(sc) == 2


def test_synthetic_inputs():
    pos = synth.hex_array(350, 14.6, 0.02, seed=1)
    assert pos.shape == (350, 3) and np.all(pos[:, 2] == 0)
    d = np.hypot(*(pos[:, None, :2] - pos[None, :, :2]).transpose(2, 0, 1))
    d[np.arange(350), np.arange(350)] = np.inf
    assert abs(d.min(1).max() - 14.6) < 0.1
    s = synth.powerlaw_fluxes(10000, 1)
    assert s.min() >= 0.1 and s.max() <= 10.0
    ra, dec = synth.healpix_centres(8)
    assert ra.size == 768
    v = np.stack([np.cos(ra) * np.cos(dec), np.sin(ra) * np.cos(dec), np.sin(dec)])
    assert abs(v.sum(1)).max() < 1e-9  # centres balance on the sphere
    assert np.sum(np.sin(dec) > 1e-12) == np.sum(np.sin(dec) < -1e-12)
    ra, dec = synth.healpix_centres(1)
    assert math.isclose(math.sin(dec[0]), 2 / 3) and math.isclose(ra[0], math.pi / 4)
    beam = synth.harmonic_efield_beam(seed=3, naz=36, nza=10)
    assert beam.data_array.shape == (2, 2, 1, 10, 36)
    assert np.abs(beam.data_array[..., -1, :]).max() < 1e-6  # horizon taper


def test_simulate_vis_requires_gpu_flag_and_validates():
    import matvis_amd

    with pytest.raises(NotImplementedError, match="use_gpu=True"):
        matvis_amd.simulate_vis({0: (0, 0, 0)}, np.ones((1, 1)), np.zeros(1),
                                np.zeros(1), np.array([1e8]),
                                astro.Time([2459863.0]),
                                [GaussianBeam(sigma=0.1)],
                                synth.hera_location())
    from matvis_amd.gpu.gpu import simulate_multi

    for kw, msg in [({"source_buffer": 0.0}, "source_buffer must satisfy"),
                    ({"memory_buffer": 1.1}, "memory_buffer must satisfy")]:
        with pytest.raises(ValueError, match=msg):
            simulate_multi(antpos=np.zeros((1, 3)), freqs=[1e8],
                           times=astro.Time([2459863.0]),
                           skycoords=astro.SkyCoord([0.0], [0.0]),
                           telescope_loc=synth.hera_location(),
                           fluxes=np.ones((1, 1)),
                           beam_list=[GaussianBeam(sigma=0.1)], **kw)


def test_matprod_antpairs_validation():
    from matvis_amd.gpu.matprod import GPUMatMul

    with pytest.raises(ValueError, match="out of range"):
        GPUMatMul(1, 1, 3, np.array([[0, 3]]))
    mp = GPUMatMul(2, 2, 3, None, precision=2)
    assert mp.npairs == 9 and mp.antpairs[4].tolist() == [1, 1]


def test_split_products_default_follows_terms():
    """Given only the term count, a split form takes the products that its
    kernels implement for it: 2 terms -> 3, 3 terms -> 6; plain forms 1."""
    from matvis_amd._lib import FORM_C64, FORM_SPLIT_BF16, FORM_SPLIT_F16
    from matvis_amd.gpu.gpu import _resolve_matprod
    from matvis_amd.gpu.herk import SPLIT_VARIANTS
    from matvis_amd.gpu.matprod import GPUMatMul, GPUMatMulSplitBF16

    r = _resolve_matprod
    assert r("GPUMatMulSplitBF16", 2, None, None) == (FORM_SPLIT_BF16, 2, 3)
    assert r("GPUMatMulSplitBF16", 2, 3, None) == (FORM_SPLIT_BF16, 3, 6)
    assert r("GPUMatMulSplitBF16", 2, 2, 4) == (FORM_SPLIT_BF16, 2, 4)
    assert r("GPUMatMulSplitF16", 1, 3, None) == (FORM_SPLIT_F16, 3, 6)
    assert r("GPUMatMul", 1, None, None) == (FORM_C64, 1, 1)
    for t in (2, 3):
        assert (t, r("GPUMatMulSplitBF16", 2, t, None)[2]) in SPLIT_VARIANTS
    mp = GPUMatMulSplitBF16(1, 2, 3, None, precision=2, split_terms=3)
    assert (mp.nterms, mp.nprod) == (3, 6)
    assert (GPUMatMul(1, 2, 3, None).nterms, GPUMatMul(1, 2, 3, None).nprod) \
        == (1, 1)


@pytest.mark.parametrize("nf", [1, 2])
def test_product_error_metric_matches_oracle(nf):
    """The metric the bench reports equals the oracle's definition, and a
    planted error on one baseline shows at the scale the autos give it."""
    import oracle as O
    from matvis_amd.metrics import normalized_error

    rng = np.random.default_rng(3)
    nant = 7
    z = rng.normal(size=(nant * nf, 40)) + 1j * rng.normal(size=(nant * nf, 40))
    want = O.vis_from_z(z, nant, nf)[None]
    got = want.copy()
    got[0, 2 * nant + 5, 0, nf - 1] += 1e-3
    a = normalized_error(got, want, nant, nf)
    b = O.normalized_error(got, want, nant, nf)
    np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)
    auto = lambda i, p: abs(want[0, i * nant + i, p, p])  # noqa: E731
    # element (2, 5, 0, nf-1) pairs conj(Z[2, nf-1]) with Z[5, 0]
    np.testing.assert_allclose(a[0], 1e-3 / np.sqrt(auto(2, nf - 1) * auto(5, 0)),
                               rtol=1e-12)


def test_data_path_resolution(tmp_path, monkeypatch):
    """DATA_PATH: $MATVIS_DATA_PATH, else an installed matvis package's
    data directory (found without importing it), else matvis_amd/data."""
    import sys

    import matvis_amd

    monkeypatch.setenv("MATVIS_DATA_PATH", str(tmp_path))
    assert matvis_amd._data_path() == tmp_path
    monkeypatch.delenv("MATVIS_DATA_PATH")
    pkg = tmp_path / "site" / "matvis"
    (pkg / "data").mkdir(parents=True)
    (pkg / "__init__.py").write_text("raise RuntimeError('imported')\n")
    monkeypatch.syspath_prepend(str(tmp_path / "site"))
    assert matvis_amd._data_path() == pkg / "data"
    assert "matvis" not in sys.modules
    monkeypatch.setattr(sys, "path", [p for p in sys.path
                                      if p != str(tmp_path / "site")])
    import importlib

    importlib.invalidate_caches()
    got = matvis_amd._data_path()
    assert got.name == "data" and got.parent.name in ("matvis_amd", "matvis")



def test_chunks_fit_this_builds_footprint():
    """max_memory bounds this build's device memory: the reference's chunk
    model (which sizes its own buffers) is the starting point, raised while
    the footprint model (Z terms, split-K partials, accumulator, beam
    tables and values, outputs, the rotation plugin's arrays) exceeds the
    budget.  FP64 at 350 antennas x 100k sources under 1.5 GiB: the
    reference's model gives 2 chunks, the partials need a third."""
    from matvis_amd._lib import FORM_C128, FORM_SPLIT_BF16
    from matvis_amd._utils import get_desired_chunks
    from matvis_amd.gpu.engine import device_bytes
    from matvis_amd.gpu.gpu import fit_chunks

    beam = synth.harmonic_efield_beam(seed=4, naz=360, nza=91)
    budget = 1.5 * 2**30
    kw = dict(nant=350, polarized=True, nfreq=1, nsrc=100_000,
              source_buffer=1.0, beams=[beam], npairs=350 * 350,
              precision=2)
    ref, _ = get_desired_chunks(budget, 1, [beam], 2, 2, 350, 100_000, 2)
    assert ref == 2
    ch, npc = fit_chunks(0.9 * budget, ref, form=FORM_C128, nterms=1,
                         nprod=1, **kw)
    assert (ch, npc) == (3, 33334)

    def total(form, nt, npd, nch):
        n = -(-100_000 // nch)
        return device_bytes(form=form, nterms=nt, nprod=npd, nant=350,
                            polarized=True, nfreq=1, nsrc=100_000,
                            nchunks=nch, chunk_size=n, nsrc_alloc=n,
                            nbeam=1, beam_pixels=91 * 363, beam_sets=1,
                            npairs=350 * 350, precision=2)["total"]

    assert total(FORM_C128, 1, 1, 2) > 0.9 * budget >= total(FORM_C128, 1, 1,
                                                              3)
    # split-BF16 fits at the reference's count: unchanged
    assert fit_chunks(0.9 * budget, ref, form=FORM_SPLIT_BF16, nterms=2,
                      nprod=3, **kw)[0] == ref
    # a budget nothing fits: the cap, not an endless loop
    assert fit_chunks(1e6, 1, form=FORM_C128, nterms=1, nprod=1,
                      **kw)[0] == 1000


def test_human_readable_size_and_ceildiv():
    """The reference's expected strings (tests/test_utils.py:20-38)."""
    from matvis_amd._utils import ceildiv, human_readable_size as h

    assert h(0) == "0.00 B" and h(1) == "1.00 B" and h(1023) == "1023.00 B"
    assert h(1024) == "1.00 KiB" and h(1024**2) == "1.00 MiB"
    assert h(1024**3) == "1.00 GiB" and h(1024**4) == "1.00 TiB"
    assert h(1024**5) == "1.00 PiB" and h(1024**6) == "1024.00 PiB"
    assert h(1024**6, decimal_places=3) == "1024.000 PiB"
    assert h(1024**6, decimal_places=3, indicate_sign=True) == "+1024.000 PiB"
    assert h(-2048, indicate_sign=True) == "-2.00 KiB"
    assert [ceildiv(a, 4) for a in (0, 1, 4, 5, 8)] == [0, 1, 1, 2, 2]


def test_log_progress_reports_at_info(caplog):
    import logging

    from matvis_amd._utils import log_progress

    class Proc:
        def memory_info(self):
            class M:
                rss = 3 * 1024**3
            return M()

    with caplog.at_level(logging.WARNING, logger="matvis_amd._utils"):
        assert log_progress(0.0, 1.0, 2, 10, Proc(), 5) == (1.0, 5)
    assert not caplog.records
    with caplog.at_level(logging.INFO, logger="matvis_amd._utils"):
        t, rss = log_progress(0.0, 0.0, 2, 10, Proc(), 2 * 1024**3)
    assert rss == 3 * 1024**3 and t > 0
    msg = caplog.records[-1].getMessage()
    assert "2/10 times (20.0%)" in msg and "3.00 GiB (+1.00 GiB)" in msg


def test_autotune_classes_signature_and_cache(tmp_path, monkeypatch):
    """Problem classes (m to a power of two, K to a power of four), batch
    signatures independent of order, default configurations, the
    per-problem assignment and the file cache."""
    from matvis_amd._lib import FORM_C128, FORM_SPLIT_BF16
    from matvis_amd.gpu import autotune as A

    pc = A.problem_class
    assert pc({"m": 700, "nseg": 2, "seg_len": 59700}) == (1024, 4 ** 9)
    assert pc({"m": 64, "nseg": 1, "seg_len": 4}) == (64, 4)
    assert pc({"m": 65, "nseg": 1, "seg_len": 5}) == (128, 16)
    assert pc({"m": 1, "nseg": 1, "seg_len": 0}) == (1, 1)
    ps = [{"m": 700, "nseg": 2, "seg_len": 59700},
          {"m": 64, "nseg": 1, "seg_len": 4000}]
    sig = A.signature(FORM_SPLIT_BF16, 2, 3, ps)
    assert sig == A.signature(FORM_SPLIT_BF16, 2, 3, ps[::-1])
    assert sig != A.signature(FORM_SPLIT_BF16, 2, 4, ps)
    assert A.default_config(FORM_SPLIT_BF16, (64, 4096)) == (64, 192)
    assert A.default_config(FORM_SPLIT_BF16, (1024, 4 ** 9)) == (128, 192)
    assert A.default_config(FORM_C128, (1024, 16)) == (64, 64)
    cfg = {pc(ps[0]): (128, 96), pc(ps[1]): (64, 48)}
    assert A.assignment(cfg, ps) == ([128, 64], [96, 48])
    monkeypatch.setenv("MATVIS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    key = A.signature(FORM_SPLIT_BF16, 2, 4, ps)
    A._store_file(key, cfg)
    assert A._cached(key) == cfg
