"""Seeded small problems shared by the oracle and GPU tests."""

import numpy as np


def small_problem(nant=6, nsrc=300, ntimes=3, polarized=False, seed=1,
                  beam="gaussian", order=1, nbeam=1):
    """A small seeded simulation description shared by oracle and GPU tests.

    Returns a dict with antpos, eci, rotations, I_sky, freq and beam
    descriptions for both the oracle (``oracle_beam``) and the product
    (``beam_list``).
    """
    from matvis_amd import astro, synth
    from matvis_amd.beams import GaussianBeam

    antpos = synth.disk_array(nant, 100.0, seed)
    ra, dec = synth.random_sky(nsrc, seed + 1)
    from oracle import point_source_crd_eq

    eci = point_source_crd_eq(ra, dec)
    times = synth.times_spaced(ntimes, 6.0)
    loc = synth.hera_location()
    rotations = astro.rotation_matrices(times, loc)
    I_sky = synth.powerlaw_fluxes(nsrc, seed + 2)
    freq = 150e6
    if beam == "gaussian":
        sig = np.radians([10.0, 14.0, 7.0, 20.0][:nbeam])
        beam_list = [GaussianBeam(sigma=s) for s in sig]
        oracle_beam = {"kind": "gaussian", "sigma": sig}
    else:
        beam_list = [synth.harmonic_efield_beam(seed=10 + b, naz=72, nza=19,
                                                polarized=True,
                                                dtype=np.complex128)
                     for b in range(nbeam)]
        from oracle import prepare_grid

        preps = [prepare_grid(b.data_array, b.axis1_array, b.axis2_array,
                              b.freq_array, freq, polarized, b.beam_type)
                 for b in beam_list]
        oracle_beam = {
            "kind": "grid",
            "planes": np.stack([p[0] for p in preps]),
            "daz": [p[1] for p in preps],
            "dza": [p[2] for p in preps],
            "azmin": [p[3] for p in preps],
            "order": order,
        }
    return dict(antpos=antpos, ra=ra, dec=dec, eci=eci, times=times, loc=loc,
                rotations=rotations, I_sky=I_sky, freq=freq,
                beam_list=beam_list, oracle_beam=oracle_beam,
                polarized=polarized, order=order)
