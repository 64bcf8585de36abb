"""The ERFA-style rotation: the oracle against pyerfa's own values
(tests/golden/erfa_vectors.npz, written by tests/golden/make_erfa_golden.py
with pyerfa), the host context model against them, and the device path
against the oracle."""

from pathlib import Path

import numpy as np
import pytest

import oracle as O

G = np.load(Path(__file__).parent / "golden" / "erfa_vectors.npz")
NT = G["utc"].size
MAS = np.pi / (180 * 3600e3)


def ctx(i):
    return {k[4:]: G[k][i] for k in G.files if k.startswith("ctx_")}


def eci():
    return O.point_source_crd_eq(G["ra"], G["dec"])


def ang(a, b):
    """Angles (rad) between unit-vector columns."""
    return np.arccos(np.clip(np.sum(a * b, axis=0), -1.0, 1.0))


# ---------------------------------------------------------------- oracle


@pytest.mark.parametrize("i", range(NT))
def test_oracle_deflection_and_aberration_are_erfa(i):
    c = ctx(i)
    ld = O.erfa_ld(eci(), c["eh"], float(c["em"]), 1e-6)
    np.testing.assert_allclose(ld, G["ld"][i].T, rtol=0, atol=2e-16)
    ab = O.erfa_ab(G["ld"][i].T, c["v"], float(c["em"]), float(c["bm1"]))
    np.testing.assert_allclose(ab, G["ab"][i].T, rtol=0, atol=2e-16)


@pytest.mark.parametrize("i", range(NT))
def test_oracle_chain_matches_erfa_observed(i):
    """BCRS corrections + CIRS -> ENU against eraAtco13 (no refraction):
    within 10 mas, the reference's own budget for this path."""
    c = ctx(i)
    topo = O.erfa_atioq_matrix(c) @ O.erfa_bcrs(eci(), c)
    assert ang(topo, G["obs_enu"][i]).max() < 10 * MAS


def test_oracle_chain_planted_errors_fail():
    """Dropping aberration or transposing the Earth rotation is caught."""
    c = ctx(0)
    p = O.erfa_ld(eci(), c["eh"], float(c["em"]), 1e-6)
    no_ab = O.erfa_atioq_matrix(c) @ (np.asarray(c["bpn"]) @ p)
    assert ang(no_ab, G["obs_enu"][0]).max() > 5000 * MAS
    bad = dict(c, eral=-c["eral"])
    wrong = O.erfa_atioq_matrix(bad) @ O.erfa_bcrs(eci(), c)
    assert ang(wrong, G["obs_enu"][0]).max() > 1e-3


# ---------------------------------------------------------------- host model


@pytest.mark.parametrize("i", range(NT))
def test_host_context_matches_erfa(i):
    """The host context from the ERFA-generated daily tables (2015-2045)
    against eraApco13's own at the golden dates (two of them a fraction of
    a day past a node): the full chain within the reference's 10 mas of
    eraAtco13 (tests/test_coordrot.py:124-125)."""
    from matvis_amd import astro

    lon, lat, hm = G["site"]
    c = astro.erfa_context(float(G["utc"][i]), lat, lon, hm)
    g = ctx(i)
    assert not c["approx"]
    d_eral = (c["eral"] - g["eral"] + np.pi) % (2 * np.pi) - np.pi
    assert abs(d_eral) < 1e-9
    np.testing.assert_allclose(c["bpn"], g["bpn"], rtol=0, atol=2e-9)
    np.testing.assert_allclose(c["eh"], g["eh"], rtol=0, atol=5e-9)
    np.testing.assert_allclose(c["v"], g["v"], rtol=0, atol=1e-10)
    assert abs(c["em"] - g["em"]) < 5e-9
    topo = O.erfa_atioq_matrix(c) @ O.erfa_bcrs(eci(), c)
    assert ang(topo, G["obs_enu"][i]).max() < 10 * MAS


@pytest.mark.parametrize("i", range(NT))
def test_series_context_close_to_erfa(i):
    """The low-precision series context (dates outside the table): within
    150 mas of eraAtco13."""
    from matvis_amd import astro

    lon, lat, hm = G["site"]
    c = astro.erfa_context_series(float(G["utc"][i]), lat, lon, hm)
    g = ctx(i)
    d_eral = (c["eral"] - g["eral"] + np.pi) % (2 * np.pi) - np.pi
    assert abs(d_eral) < 1e-9
    assert np.linalg.norm(c["v"] - g["v"]) < 2e-3 * np.linalg.norm(g["v"])
    assert abs(c["em"] - g["em"]) < 2e-4
    topo = O.erfa_atioq_matrix(c) @ O.erfa_bcrs(eci(), c)
    assert ang(topo, G["obs_enu"][i]).max() < 150 * MAS


def test_table_interpolation_and_range():
    """The daily table: nodes interpolate to themselves, the cubic is exact
    for cubic data, and dates outside 2015-2045 fall back to the series
    (marked approx)."""
    from matvis_amd import astro

    assert astro._table() is not None
    vals = np.arange(10.0) ** 3 - 2 * np.arange(10.0)
    for jd in (3.0, 4.25, 6.999):
        got = astro._interp_daily(vals, 0.0, jd)
        assert abs(got - (jd ** 3 - 2 * jd)) < 1e-9
    assert astro._interp_daily(vals, 0.0, 0.5) is None
    assert astro._interp_daily(vals, 0.0, 8.0) is None
    lon, lat, hm = G["site"]
    assert astro.erfa_context(2469000.5, lat, lon, hm)["approx"]  # 2048


def test_series_context_without_corrections_is_the_rotation_matrix():
    """With light deflection and aberration left out the series context
    reproduces R_t of astro.eci_to_enu_rotation."""
    from matvis_amd import astro
    from matvis_amd.gpu.coords import atioq_matrix

    lon, lat, hm = G["site"]
    for jd in G["utc"]:
        c = astro.erfa_context_series(float(jd), lat, lon, hm)
        np.testing.assert_allclose(atioq_matrix(c) @ c["bpn"],
                                   astro.eci_to_enu_rotation(jd, lat, lon),
                                   rtol=0, atol=1e-12)
        np.testing.assert_allclose(atioq_matrix(c), O.erfa_atioq_matrix(c),
                                   rtol=0, atol=0)


def test_context_array_layout():
    from matvis_amd.gpu.coords import bcrs_context_array

    c = ctx(1)
    a = bcrs_context_array(c)
    assert a.shape == (17,)
    np.testing.assert_array_equal(a[:3], c["eh"])
    assert a[3] == c["em"] and a[7] == c["bm1"]
    np.testing.assert_array_equal(a[4:7], c["v"])
    np.testing.assert_array_equal(a[8:].reshape(3, 3), c["bpn"])


# ---------------------------------------------------------------- device


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(NT))
def test_gpu_bcrs_update_matches_oracle(i):
    import matvis_amd._device as D
    from matvis_amd._lib import call
    from matvis_amd.gpu.coords import bcrs_context_array

    print(f"bcrs update time {i}", flush=True)
    e = eci()
    n = e.shape[1]
    src = D.to_dev(e, np.float64)
    out = D.zeros((3, n), np.float64)
    ctx_arr = bcrs_context_array(ctx(i))  # alive for the whole call
    call("mv_bcrs_update", D.ptr(src), n, n, ctx_arr.ctypes.data, 1e-6,
         D.ptr(out), n, D.stream_ptr())
    np.testing.assert_allclose(D.to_host(out), O.erfa_bcrs(e, ctx(i)),
                               rtol=0, atol=4e-16)


@pytest.mark.gpu
@pytest.mark.parametrize("every", [0.0, 1e9])
def test_gpu_erfa_rotation_per_time(every):
    """GPUCoordinateRotationERFA with the golden contexts: the per-time
    topocentric vectors equal the oracle chain (BCRS from the last
    evaluation, ENU matrix of the current time) and lie within 10 mas of
    eraAtco13 when re-evaluated every time."""
    from matvis_amd import astro
    from matvis_amd.gpu.coords import GPUCoordinateRotationERFA

    print(f"erfa rotation update_bcrs_every={every}", flush=True)
    lon, lat, hm = G["site"]
    loc = astro.EarthLocation.from_geodetic(np.degrees(lon), np.degrees(lat),
                                            hm)
    co = GPUCoordinateRotationERFA(
        every, flux=np.ones(G["ra"].size), times=G["utc"], telescope_loc=loc,
        skycoords=astro.SkyCoord(G["ra"], G["dec"]), precision=2,
        astrom=[ctx(i) for i in range(NT)])
    co.setup()
    import matvis_amd._device as D
    for t in range(NT):
        co.rotate(t)
        got = D.to_host(co.all_coords_topo)
        c_b = ctx(0 if every > 0 else t)
        want = O.erfa_atioq_matrix(ctx(t)) @ O.erfa_bcrs(eci(), c_b)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-15)
        if every == 0.0:
            assert ang(got, G["obs_enu"][t]).max() < 10 * MAS


@pytest.mark.gpu
def test_gpu_simulate_with_erfa_rotation():
    """simulate() through GPUCoordinateRotationERFA (golden contexts, BCRS
    evaluated once) against the oracle fed the same CIRS directions and
    per-time ENU matrices: FP64 within 1e-12."""
    from matvis_amd import astro
    from matvis_amd.beams import GaussianBeam
    from matvis_amd.gpu.gpu import simulate
    from matvis_amd import synth

    print("simulate with GPUCoordinateRotationERFA", flush=True)
    lon, lat, hm = G["site"]
    loc = astro.EarthLocation.from_geodetic(np.degrees(lon), np.degrees(lat),
                                            hm)
    rng = np.random.default_rng(5)
    nsrc, nant = 3000, 10
    ra, dec = synth.random_sky(nsrc, 9)
    flux = rng.uniform(0.1, 5.0, nsrc)
    antpos = synth.disk_array(nant, 80.0, 4)
    contexts = [ctx(i) for i in range(NT)]
    v = simulate(antpos=antpos, freq=150e6, times=G["utc"],
                 skycoords=astro.SkyCoord(ra, dec), telescope_loc=loc,
                 I_sky=flux, beam_list=[GaussianBeam(sigma=0.3)],
                 polarized=False, precision=2,
                 coord_method="GPUCoordinateRotationERFA",
                 coord_method_params={"update_bcrs_every": 1e12,
                                      "astrom": contexts})
    e0 = O.point_source_crd_eq(ra, dec)
    ref = O.simulate(antpos=antpos, freq=150e6, eci=O.erfa_bcrs(e0, contexts[0]),
                     rotations=np.stack([O.erfa_atioq_matrix(c)
                                         for c in contexts]),
                     I_sky=flux, beam={"kind": "gaussian", "sigma": [0.3]},
                     polarized=False)
    e = O.normalized_error(v[..., None, None], ref[..., None, None], nant, 1)
    print(f"  normalised max {e[0]:.2e}", flush=True)
    assert e[0] < 1e-12


@pytest.mark.gpu
def test_gpu_erfa_precision1_and_chunk_sizes():
    """tests/test_coordrot.py:100-127 at precision 1 (within 150 arcsec of
    the observed directions; single-precision output) and :147-168 (chunk
    sizes 16, 24 and all sources give the same culled coordinates)."""
    from matvis_amd import astro
    from matvis_amd.gpu.coords import GPUCoordinateRotationERFA
    import matvis_amd._device as D

    print("erfa precision 1 and chunk sizes", flush=True)
    lon, lat, hm = G["site"]
    loc = astro.EarthLocation.from_geodetic(np.degrees(lon), np.degrees(lat),
                                            hm)
    arcsec = np.pi / (180 * 3600)

    def make(chunk):
        co = GPUCoordinateRotationERFA(
            0.0, flux=np.ones(G["ra"].size), times=G["utc"],
            telescope_loc=loc, skycoords=astro.SkyCoord(G["ra"], G["dec"]),
            precision=1, chunk_size=chunk, source_buffer=1.0,
            astrom=[ctx(i) for i in range(NT)])
        co.setup()
        return co

    co = make(None)
    for t in range(NT):
        co.rotate(t)
        got = D.to_host(co.all_coords_topo)
        assert got.dtype == np.float32
        got = got.astype(np.float64)
        got /= np.linalg.norm(got, axis=0)
        assert ang(got, G["obs_enu"][t]).max() < 150 * arcsec
    small, mid = make(16), make(24)
    for t in (0, NT - 1):
        for c in (co, small, mid):
            c.rotate(t)
        whole, _, n_all = co.select_chunk(0, t)
        whole = D.to_host(whole)[:, :n_all].copy()
        parts = []
        for c, size in ((small, 16), (mid, 24)):
            cols = []
            for k in range(-(-G["ra"].size // size)):
                crd, _, n = c.select_chunk(k, t)
                cols.append(D.to_host(crd)[:, :n].copy())
            parts.append(np.concatenate(cols, axis=1))
        for p in parts:
            np.testing.assert_array_equal(p, whole)


def test_builtin_context_warns_once():
    """The series context (outside 2015-2045) is 0.05-0.09 arcsec from ERFA:
    using it warns (once per rotation object); the table context and
    caller-supplied contexts do not."""
    import warnings

    from matvis_amd import astro
    from matvis_amd.gpu.coords import CoordinateRotationERFA

    lon, lat, hm = G["site"]
    loc = astro.EarthLocation.from_geodetic(np.degrees(lon), np.degrees(lat),
                                            hm)
    late = np.array([2469000.5, 2469000.6])  # 2048, past the table
    kw = dict(flux=np.ones(G["ra"].size), telescope_loc=loc,
              skycoords=astro.SkyCoord(G["ra"], G["dec"]), precision=2)
    co = CoordinateRotationERFA(0.0, times=late, **kw)
    with pytest.warns(astro.AstrometryAccuracyWarning, match="mas"):
        c = co.context(0)
    assert c["approx"]
    np.testing.assert_array_equal(
        c["bpn"], astro.erfa_context_series(float(late[0]), lat, lon,
                                            hm)["bpn"])
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        co.context(1)  # once per object
        CoordinateRotationERFA(0.0, times=G["utc"], **kw).context(0)
        CoordinateRotationERFA(
            0.0, times=G["utc"], astrom=[ctx(i) for i in range(NT)],
            **kw).context(0)
