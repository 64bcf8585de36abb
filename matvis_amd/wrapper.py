"""``simulate_vis``: matvis's user-facing wrapper (wrapper.py:25-200)."""

from __future__ import annotations

import logging

import numpy as np

from .astro import SkyCoord

logger = logging.getLogger(__name__)


def simulate_vis(
    ants: dict,
    fluxes: np.ndarray,
    ra: np.ndarray,
    dec: np.ndarray,
    freqs: np.ndarray,
    times,
    beams: list,
    telescope_loc,
    polarized: bool = False,
    precision: int = 1,
    use_feed: str = "x",
    use_gpu: bool = False,
    beam_spline_opts: dict | None = None,
    beam_idx: np.ndarray | None = None,
    antpairs=None,
    source_buffer: float = 1.0,
    coord_method: str = "CoordinateRotationAstropy",
    coord_method_params: dict | None = None,
    matprod_method: str = "MatMul",
    **backend_kwargs,
):
    """Simulate visibilities for every frequency and time.

    Returns ``(NFREQS, NTIMES, NBLS, NFEED, NFEED)`` if polarized, else
    ``(NFREQS, NTIMES, NBLS)``.  Antenna positions come from ``ants`` in its
    key order (wrapper.py:163).  ``use_feed`` is accepted and unused, as in
    the reference (wrapper.py:36).

    matvis_amd implements the GPU backend only: ``use_gpu=False`` raises.
    """
    if not use_gpu:
        raise NotImplementedError(
            "matvis_amd implements matvis's GPU path (use_gpu=True); the CPU "
            "backend is matvis's own"
        )
    from .gpu.gpu import simulate_multi

    freqs = np.atleast_1d(np.asarray(freqs))
    ra = np.atleast_1d(np.asarray(ra))
    fluxes = np.asarray(fluxes)
    assert fluxes.shape == (ra.size, freqs.size), (
        "The `fluxes` array must have shape (NSRCS, NFREQS)."
    )
    antpos = np.array([ants[k] for k in ants.keys()], dtype=np.float64)
    if coord_method == "CoordinateRotationAstropy":
        import astropy.units as un
        from astropy.coordinates import SkyCoord as AstropySkyCoord

        skycoords = AstropySkyCoord(ra=ra * un.rad, dec=dec * un.rad,
                                    frame="icrs")
    else:
        skycoords = SkyCoord(ra, dec)
    if matprod_method in ("MatMul", "VectorLoop"):
        matprod_method = f"GPU{matprod_method}"
    vis = simulate_multi(
        antpos=antpos, freqs=freqs, times=times, skycoords=skycoords,
        telescope_loc=telescope_loc, fluxes=fluxes, beam_list=beams,
        polarized=polarized, antpairs=antpairs, beam_idx=beam_idx,
        precision=precision, beam_spline_opts=beam_spline_opts,
        coord_method=coord_method, matprod_method=matprod_method,
        source_buffer=source_buffer, coord_method_params=coord_method_params,
        **backend_kwargs,
    )
    return vis if polarized else vis[..., 0, 0]
