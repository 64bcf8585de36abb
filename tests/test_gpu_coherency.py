"""Polarized sky models through CoordinateRotation.select_chunk: the device
rotation of coherency matrices into the alt/az basis (mv_coherency_rotate)
against the oracle (core/coords.py:99-175, coordinates.py:288-576)."""

import numpy as np
import pytest

import matvis_amd._device as D
import oracle as O
from matvis_amd import astro, synth

pytestmark = pytest.mark.gpu


def _coherency(nsrc, nf, cplx, seed):
    """0.5 [[I + Q, U - iV], [U + iV, I - Q]] with random Stokes per source
    and frequency; real (V = 0) unless cplx."""
    rng = np.random.default_rng(seed)
    i_ = rng.uniform(1, 3, (nsrc, nf))
    q_, u_ = rng.uniform(-0.5, 0.5, (2, nsrc, nf))
    v_ = rng.uniform(-0.3, 0.3, (nsrc, nf)) if cplx else np.zeros((nsrc, nf))
    coh = 0.5 * np.array([[i_ + q_, u_ - 1j * v_], [u_ + 1j * v_, i_ - q_]])
    coh = np.ascontiguousarray(coh.transpose(2, 3, 0, 1))
    return coh if cplx else coh.real.copy()


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("cplx", [False, True])
def test_polarized_select_chunk_matches_oracle(precision, cplx):
    from matvis_amd.gpu.coords import CoordinateRotationMatrix

    print(f"coherency rotation precision={precision} complex={cplx}",
          flush=True)
    nsrc, nf, chunk, ntimes = 5000, 3, 2000, 2
    ra, dec = synth.random_sky(nsrc, 11, -90.0, 90.0)
    coh = _coherency(nsrc, nf, cplx, 12)
    times = synth.times_spaced(ntimes, 60.0)
    loc = synth.hera_location()
    rots = astro.rotation_matrices(times, loc)
    co = CoordinateRotationMatrix(
        rots, flux=coh, times=times, telescope_loc=loc,
        skycoords=astro.SkyCoord(ra, dec), chunk_size=chunk,
        precision=precision, source_buffer=1.0)
    co.setup()
    want_dtype = {(1, False): np.float32, (1, True): np.complex64,
                  (2, False): np.float64, (2, True): np.complex128}
    eci = O.point_source_crd_eq(ra, dec)
    for t in range(ntimes):
        co.rotate(t)
        topo = O.rotate(rots[t], eci)
        if precision == 1:  # the device's single-precision coordinates
            topo = topo.astype(np.float32).astype(np.float64)
        for c in range(3):
            crd, flux, n = co.select_chunk(c, t)
            got = D.to_host(flux)
            assert got.dtype == want_dtype[(precision, cplx)]
            assert got.shape == (co.nsrc_alloc, nf, 2, 2)
            _, want, k, _ = O.select_chunk_polarized(
                topo, coh, ra, dec, rots[t][[1, 0, 2]], c, chunk,
                co.nsrc_alloc)
            assert n == k
            err = np.abs(got - want).max() / np.abs(coh).max()
            print(f"  t={t} chunk={c} n={n} max err {err:.2e}", flush=True)
            assert err < (1e-12 if precision == 2 else 1e-6)
            assert not got[n:].any()
            # trace (Stokes I) is unchanged by the basis rotation
            idx = O.select_chunk(topo, np.zeros(nsrc), c, chunk,
                                 co.nsrc_alloc)[3]
            tr = np.trace(got[:n], axis1=2, axis2=3)
            np.testing.assert_allclose(
                tr, np.trace(coh[idx], axis1=2, axis2=3),
                rtol=1e-12 if precision == 2 else 1e-6)


def test_polarized_select_chunk_with_erfa_rotation():
    """GPUCoordinateRotationERFA (golden contexts): the alignment rotation
    absorbs the per-source deflection and aberration; against the oracle
    fed the same CIRS -> ENU directions and average rotation."""
    from test_erfa import G, NT, ctx

    from matvis_amd.gpu.coords import GPUCoordinateRotationERFA

    print("coherency rotation with GPUCoordinateRotationERFA", flush=True)
    lon, lat, hm = G["site"]
    loc = astro.EarthLocation.from_geodetic(np.degrees(lon), np.degrees(lat),
                                            hm)
    ra, dec = G["ra"], G["dec"]
    coh = _coherency(ra.size, 2, True, 13)
    co = GPUCoordinateRotationERFA(
        0.0, flux=coh, times=G["utc"], telescope_loc=loc,
        skycoords=astro.SkyCoord(ra, dec), precision=2,
        astrom=[ctx(i) for i in range(NT)])
    co.setup()
    eci = O.point_source_crd_eq(ra, dec)
    for t in range(NT):
        co.rotate(t)
        crd, flux, n = co.select_chunk(0, t)
        c = ctx(t)
        topo = O.erfa_atioq_matrix(c) @ O.erfa_bcrs(eci, c)
        r_avg = (O.erfa_atioq_matrix(c) @ np.asarray(c["bpn"]))[[1, 0, 2]]
        _, want, k, _ = O.select_chunk_polarized(topo, coh, ra, dec, r_avg,
                                                 0, ra.size, co.nsrc_alloc)
        got = D.to_host(flux)
        assert n == k
        err = np.abs(got - want).max() / np.abs(coh).max()
        print(f"  t={t} n={n} max err {err:.2e}", flush=True)
        assert err < 1e-12


def test_simulate_rejects_a_polarized_sky():
    """gpu.simulate keeps the reference's I_sky check
    (core/__init__.py:30)."""
    from matvis_amd.beams import GaussianBeam
    from matvis_amd.gpu.gpu import simulate

    ra, dec = synth.random_sky(10, 1)
    with pytest.raises(AssertionError, match="I_sky must have shape"):
        simulate(antpos=np.zeros((2, 3)), freq=1e8,
                 times=synth.times_spaced(1, 1.0),
                 skycoords=astro.SkyCoord(ra, dec),
                 telescope_loc=synth.hera_location(),
                 I_sky=_coherency(10, 1, False, 1),
                 beam_list=[GaussianBeam(sigma=0.3)])
