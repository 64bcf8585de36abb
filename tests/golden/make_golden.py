"""Write tests/golden/golden_vectors.npz: inputs and expected outputs that pin
the oracle and the GPU path independently of the oracle's own code.

Every expected value here comes from a closed form, from explicit Python
loops over tiny inputs, from scipy's ``map_coordinates`` (the routine the
reference's beam interpolation calls), or from the reference's own data file;
none is computed with the oracle.

Sources of the inputs:

* ``crd_*``: cardinal points of ``point_source_crd_eq``
  (tests/test_coordinates.py:21-68 in the reference).
* ``matprod_nf{1,2}_*``: the ``arange + 1j`` Z of tests/test_matprod.py:45-46
  with nant = 5, nsrc = 15; expected output in the API layout
  out[i nant + j, a, b] = V[i nf + b, j nf + a] (test_matprod.py:56-65).
* ``beam_identity_*`` / ``beam_linear_*``: the analytic grids of
  tests/test_beam_interp_gpu.py:12-64 (power beams: the output is the
  square root of the interpolated power).
* ``tiny_*``: a 3-antenna, 4-source, one-time unpolarized simulation with a
  constant power beam, whose visibilities are
  V_ij = sum_s I_s / 2 exp(2 pi i nu (b_j - b_i) . s / c) (core/tau.py:30-31,
  core/getz.py:80-100, cpu/matprod.py:21-30), summed in explicit loops.
* ``fits_*``: one frequency (100 MHz) of the reference's beam fixture
  src/matvis/data/NF_HERA_Dipole_small.fits (read with matvis_amd.beamfits),
  with map_coordinates results at 64 seeded (az, za) points on the grid
  prepared by pyuvdata's rule (3 wrapped columns on each side of a
  full-circle azimuth axis).  Needs the reference's data file; the other
  groups do not.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""

from __future__ import annotations

import cmath
import math
import sys
from pathlib import Path

import numpy as np
from scipy import ndimage

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
FITS = Path("/root/reference/src/matvis/data/NF_HERA_Dipole_small.fits")
C_LIGHT = 299792458.0


def cardinal_points(out):
    ra = np.array([0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 0.0, 0.0])
    dec = np.array([0.0, 0.0, 0.0, 0.0, np.pi / 2, -np.pi / 2])
    want = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0],
                     [0, 0, 1], [0, 0, -1]], dtype=np.float64).T
    out.update(crd_ra=ra, crd_dec=dec, crd_want=want)


def matprod(out):
    nant, nsrc = 5, 15
    for nf in (1, 2):
        z = [[complex(r * nsrc + s, 1.0) for s in range(nsrc)]
             for r in range(nf * nant)]
        m = nf * nant
        v = [[sum(z[a][s].conjugate() * z[b][s] for s in range(nsrc))
              for b in range(m)] for a in range(m)]
        api = np.zeros((nant * nant, nf, nf), np.complex128)
        for i in range(nant):
            for j in range(nant):
                for a in range(nf):
                    for b in range(nf):
                        api[i * nant + j, a, b] = v[i * nf + b][j * nf + a]
        out[f"matprod_nf{nf}_z"] = np.array(z)
        out[f"matprod_nf{nf}_want"] = api


def beam_grids(out):
    za = np.linspace(0, 1, 3)
    az = np.linspace(0, 2 * np.pi, 6)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA**2, 2 - ZA**2 + np.sin(AZ)])
    out.update(beam_identity_planes=beam[:, None, None],
               beam_identity_daz=np.array([az[1] - az[0]] * 2),
               beam_identity_dza=np.array([za[1] - za[0]] * 2),
               beam_identity_az=AZ.ravel(), beam_identity_za=ZA.ravel(),
               beam_identity_want=np.sqrt(beam.reshape(2, -1)))
    za = np.linspace(0, 1, 5)
    az = np.linspace(0, 1, 5)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA])
    dec = beam[:, None, None][..., ::2, ::2]
    out.update(beam_linear_planes=dec,
               beam_linear_daz=np.array([2 * (az[1] - az[0])]),
               beam_linear_dza=np.array([2 * (za[1] - za[0])]),
               beam_linear_az=AZ.ravel(), beam_linear_za=ZA.ravel(),
               beam_linear_want=np.sqrt(beam.reshape(1, -1)))


def tiny_sim(out):
    antpos = np.array([[0.0, 0.0, 0.0], [14.6, 0.0, 0.0], [7.3, 12.64, 0.0]])
    # topocentric unit vectors (east, north, up), all above the horizon
    s = np.array([[0.0, 0.0, 1.0], [0.3, 0.1, 0.0], [-0.2, 0.5, 0.0],
                  [0.05, -0.4, 0.0]])
    s[1:, 2] = np.sqrt(1 - s[1:, 0] ** 2 - s[1:, 1] ** 2)
    flux = np.array([2.0, 1.0, 0.5, 3.0])
    freq = 150e6
    v = np.zeros((3, 3), np.complex128)
    for i in range(3):
        for j in range(3):
            acc = 0j
            for k in range(4):
                b = [antpos[j][d] - antpos[i][d] for d in range(3)]
                tau = sum(b[d] * s[k][d] for d in range(3)) / C_LIGHT
                acc += flux[k] / 2 * cmath.exp(2j * math.pi * freq * tau)
            v[i, j] = acc
    out.update(tiny_antpos=antpos, tiny_topo=s.T.copy(), tiny_flux=flux,
               tiny_freq=np.array(freq), tiny_want=v)


def fits_beam(out):
    sys.path.insert(0, str(ROOT))
    from matvis_amd.beamfits import read_beamfits

    b = read_beamfits(FITS)
    # 100 MHz only, rounded to complex64 (as stored) before anything is
    # computed from it
    data = b.data_array[:, :, :1].astype(np.complex64).astype(np.complex128)
    az, za = b.axis1_array, b.axis2_array
    # pyuvdata's padding of a full-circle azimuth axis (gpu/beams.py:15-19)
    step = az[1] - az[0]
    azp = np.concatenate([az[0] + np.arange(-3, 0) * step, az,
                          az[-1] + np.arange(1, 4) * step])
    grid = np.concatenate([data[..., -3:], data, data[..., :3]], axis=-1)
    daz, azmin = np.diff(azp)[0], azp.min()
    dza = np.diff(za)[0]
    rng = np.random.default_rng(2024)
    paz = rng.uniform(0, 2 * np.pi, 64)
    pza = rng.uniform(0, np.pi / 2, 64)
    coords = np.array([pza / dza, (paz - azmin) / daz])
    for order in (1, 3):
        res = np.zeros((2, 2, 64), np.complex128)  # (feed, vec, point)
        for fd in range(2):
            for vec in range(2):
                plane = grid[vec, fd, 0]
                res[fd, vec] = (ndimage.map_coordinates(plane.real, coords,
                                                        order=order)
                                + 1j * ndimage.map_coordinates(
                                    plane.imag, coords, order=order))
        out[f"fits_want_order{order}"] = res
    out.update(fits_data=data.astype(np.complex64), fits_az=az, fits_za=za,
               fits_freq=b.freq_array[:1], fits_pt_az=paz, fits_pt_za=pza)


def main():
    out = {}
    cardinal_points(out)
    matprod(out)
    beam_grids(out)
    tiny_sim(out)
    if FITS.exists():
        fits_beam(out)
    else:
        print(f"{FITS} not found: fits_* vectors skipped")
    np.savez_compressed(HERE / "golden_vectors.npz", **out)
    print(f"wrote {HERE / 'golden_vectors.npz'} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
