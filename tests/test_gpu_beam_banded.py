"""Beam interpolation with the grid staged in LDS by za band
(mv_beam_interp_banded) against the tap-gather kernel (mv_beam_interp), bit
for bit; and a 0.25 deg beam, 33 MB of planes, through the drop-in path
against the oracle."""

import numpy as np
import pytest
import torch

import oracle as O
from _problems import small_problem
from matvis_amd import _device as D
from matvis_amd import astro
from matvis_amd._lib import call, lib
from matvis_amd.gpu.beams import DeviceBeamSet
from matvis_amd.gpu.gpu import simulate

pytestmark = pytest.mark.gpu


def _positions(count, seed):
    """Topocentric unit vectors (3, count): random directions over the
    upper hemisphere (za up to 90 deg, past the grid's 80 deg edge), the
    zenith and a few horizon points."""
    rng = np.random.default_rng(seed)
    za = np.arccos(rng.uniform(0.0, 1.0, count))
    az = rng.uniform(0, 2 * np.pi, count)
    za[:3] = [0.0, np.pi / 2, np.radians(80.0) - 1e-9]
    e, n = np.sin(za) * np.sin(az), np.sin(za) * np.cos(az)
    return np.stack([e, n, np.cos(za)])


@pytest.mark.parametrize("naz,nza", [(363, 91), (1443, 181), (4, 2)])
@pytest.mark.parametrize("is_f64", [0, 1])
@pytest.mark.parametrize("order", [1, 3])
@pytest.mark.parametrize("polarized,is_power", [(True, 0), (False, 0),
                                                (False, 1)])
def test_banded_equals_gather(naz, nza, is_f64, order, polarized, is_power):
    rng = np.random.default_rng(naz + 7 * is_f64 + order)
    nbeam, nf = 2, (2 if polarized else 1)
    dt = np.complex128 if is_f64 else np.complex64
    shape = (nbeam, nf, nf, nza, naz)
    planes = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dt)
    if is_power:
        planes = (np.abs(planes) - 0.3).astype(dt)  # some negative powers
    dza = np.radians(80.0) / (nza - 1)
    daz = 2 * np.pi / (naz - 3)
    grid = np.array([[daz, dza, 0.0], [daz, dza, -2e-3]])
    count, n = (5000, 4711) if nza > 2 else (7, 5)
    crd = _positions(count, naz)
    bufs = dict(planes=D.to_dev(planes), grid=D.to_dev(grid, np.float64),
                crd=D.to_dev(crd, np.float64),
                crd32=D.to_dev(crd, np.float32),
                n=D.to_dev(np.array([n], np.int32)))
    nplanes = nbeam * nf * nf
    for crd_key, crd_f64 in (("crd", 1), ("crd32", 0)):
        outs = [D.zeros((nplanes, count), dt) + 9.0 for _ in range(2)]
        args = (D.ptr(bufs["planes"]), is_f64, nbeam, nf, nf, nza, naz,
                D.ptr(bufs["grid"]), order, is_power, D.ptr(bufs[crd_key]),
                crd_f64, count, count, D.ptr(bufs["n"]))
        call("mv_beam_interp", *args, D.ptr(outs[0]), count, D.stream_ptr())
        need = lib().mv_beam_interp_banded_workspace_bytes(is_f64, nza, naz,
                                                           count)
        ws = D.empty((max(need, 1),), np.uint8)
        call("mv_beam_interp_banded", *args, D.ptr(outs[1]), count,
             D.ptr(ws), ws.numel(), D.stream_ptr())
        a, b = D.to_host(outs[0]), D.to_host(outs[1])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), crd_key
        assert np.all(a[:, n:] == 0)
        if nza > 2:
            assert np.count_nonzero(a[:, :n]) > 0.7 * nplanes * n


def test_large_beam_matches_oracle_and_bands_match_gathers():
    """A 0.25 deg polarized E-field beam (1440 x 361, 33 MB of complex128
    planes) through the drop-in path agrees with the oracle at FP64; the
    same beam set interpolated through the za bands (two rows a band at
    complex128) gives the gathers' values bit for bit."""
    from matvis_amd import synth

    p = small_problem(nant=6, nsrc=1500, ntimes=2, polarized=True,
                      beam="grid", order=3)
    beam = synth.harmonic_efield_beam(seed=31, naz=1440, nza=361,
                                      polarized=True, dtype=np.complex128)
    assert lib().mv_beam_interp_banded_workspace_bytes(1, 361, 1443, 1500) > 0
    planes, daz, dza, azmin = O.prepare_grid(
        beam.data_array, beam.axis1_array, beam.axis2_array, beam.freq_array,
        p["freq"], True)
    obeam = {"kind": "grid", "planes": planes[None], "daz": [daz],
             "dza": [dza], "azmin": [azmin], "order": 3}
    v = simulate(antpos=p["antpos"], freq=p["freq"], times=p["times"],
                 skycoords=astro.SkyCoord(p["ra"], p["dec"]),
                 telescope_loc=p["loc"], I_sky=p["I_sky"], beam_list=[beam],
                 polarized=True, coord_method="CoordinateRotationMatrix",
                 coord_method_params={"rotations": p["rotations"]},
                 beam_spline_opts={"order": 3}, precision=2)
    ref = O.simulate(antpos=p["antpos"], freq=p["freq"], eci=p["eci"],
                     rotations=p["rotations"], I_sky=p["I_sky"], beam=obeam,
                     polarized=True)
    e = O.normalized_error(v, ref, 6, 2)
    print(f"0.25 deg beam: normalised max {e[0]:.2e}", flush=True)
    assert e[0] < 1e-12
    crd = D.to_dev(_positions(3000, 5), np.float64)
    outs = [D.zeros((4, 3000), np.complex128) for _ in range(2)]
    for out, banded in zip(outs, (False, True)):
        dbs = DeviceBeamSet([beam], p["freq"], True, np.complex128, order=3,
                            banded=banded)
        dbs.eval(crd, 3000, 3000, out, 3000)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
