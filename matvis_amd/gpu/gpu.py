"""``simulate``: the GPU visibility simulator (matvis gpu/gpu.py:57-256).

Same keyword signature as the reference, plus ``split_terms`` /
``split_products`` for the mixed-precision HERK forms.  The per-time loop
runs as the fused :class:`~matvis_amd.gpu.engine.VisibilityEngine`.
"""

from __future__ import annotations

import logging

import numpy as np

from .. import _device as D
from .._utils import get_desired_chunks, get_dtypes
from ..beams import unwrap
from ..core import _validate_inputs
from ..core.beams import _wrangle_beams
from ..core.coords import CoordinateRotation
from . import coords as _coords  # noqa: F401  (registers rotation methods)
from . import matprod as mp
from .beams import spline_order
from .engine import VisibilityEngine
from .herk import products_for

logger = logging.getLogger(__name__)

HAVE_CUDA = True


def _resolve_matprod(matprod_method: str, precision: int, split_terms,
                     split_products):
    cls = getattr(mp, matprod_method)
    form = cls.split_form
    if form is None:
        from .._lib import FORM_C64, FORM_C128

        form = FORM_C128 if precision == 2 else FORM_C64
        return form, 1, 1
    nterms = split_terms or cls.default_terms
    return form, nterms, products_for(nterms, split_products)


def build_engine(
    *,
    antpos,
    freqs,
    times,
    skycoords,
    telescope_loc,
    fluxes,
    beam_list,
    polarized: bool = False,
    antpairs=None,
    beam_idx=None,
    max_memory=np.inf,
    min_chunks: int = 1,
    precision: int = 1,
    beam_spline_opts: dict | None = None,
    coord_method: str = "CoordinateRotationAstropy",
    matprod_method: str = "GPUMatMul",
    source_buffer: float = 1.0,
    coord_method_params: dict | None = None,
    memory_buffer: float = 0.9,
    split_terms: int | None = None,
    split_products: int | None = None,
    beam_cache: dict | None = None,
):
    """The set-up :class:`VisibilityEngine` of a simulation and its number
    of times; arguments as :func:`simulate_multi`.  ``beam_cache``: a dict
    shared between engines so that device beam tables of the same beam
    objects are built once (see ``VisibilityEngine``)."""
    if not 0 < source_buffer <= 1:
        raise ValueError("source_buffer must satisfy 0 < source_buffer <= 1")
    if not 0 < memory_buffer <= 1:
        raise ValueError("memory_buffer must satisfy 0 < memory_buffer <= 1")
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    fluxes = np.asarray(fluxes, dtype=np.float64)
    if fluxes.ndim == 1:
        fluxes = fluxes[:, None]
    nax, nfeed, nant, ntimes = _validate_inputs(
        precision, polarized, np.asarray(antpos), times, fluxes[:, 0]
    )
    rtype, ctype = get_dtypes(precision)
    nsrc = fluxes.shape[0]
    form, nterms, nprod = _resolve_matprod(matprod_method, precision,
                                           split_terms, split_products)
    freemem = min(max_memory, D.free_memory())
    inner_beams = [unwrap(b)[0] for b in beam_list]
    nchunks, npixc = get_desired_chunks(
        freemem, min_chunks, inner_beams, nax, nfeed, nant, nsrc, precision,
        source_buffer=source_buffer, memory_buffer=memory_buffer,
    )
    beams, _, beam_idx = _wrangle_beams(beam_idx, beam_list, polarized, nant,
                                        float(freqs[0]))
    nchunks, npixc = fit_chunks(
        freemem * memory_buffer, nchunks, form=form, nterms=nterms,
        nprod=nprod, nant=nant, polarized=polarized, nfreq=freqs.size,
        nsrc=nsrc, source_buffer=source_buffer, beams=beams,
        npairs=nant * nant if antpairs is None else len(antpairs),
        precision=precision)
    nsrc_alloc = int(npixc * source_buffer)
    coord_cls = CoordinateRotation._methods[coord_method]
    coords = coord_cls(
        flux=np.sqrt(0.5 * fluxes[:, 0]), times=times,
        telescope_loc=telescope_loc, skycoords=skycoords, chunk_size=npixc,
        precision=precision, source_buffer=source_buffer, gpu=True,
        **(coord_method_params or {}),
    )
    coords.setup()
    engine = VisibilityEngine(
        antpos=antpos, freqs=freqs, coords=coords, fluxes=fluxes,
        beam_list=beams, beam_idx=beam_idx, polarized=polarized,
        precision=precision, form=form, nterms=nterms, nprod=nprod,
        order=spline_order(beam_spline_opts), nchunks=nchunks,
        chunk_size=npixc, nsrc_alloc=nsrc_alloc, antpairs=antpairs,
        beam_cache=beam_cache, beam_src=beam_list,
    )
    logger.info(f"Running with {nchunks} chunks, form {form}")
    engine.setup()
    return engine, ntimes


def fit_chunks(budget: float, nchunks: int, *, form, nterms, nprod, nant,
               polarized, nfreq, nsrc, source_buffer, beams, npairs,
               precision) -> tuple[int, int]:
    """(nchunks, sources per chunk): the reference's chunk count
    (get_desired_chunks, which sizes the reference's own buffers), raised
    until this build's device footprint (engine.device_bytes: Z terms,
    split-K workspace, FP64 accumulator, beam tables and values, output
    buffers, the rotation plugin's arrays) fits the budget
    (min(max_memory, free memory) x memory_buffer).  If even min(nsrc, 1000)
    chunks do not fit, that count is used and a warning logged."""
    from ..beams import is_grid_beam, unwrap
    from .engine import device_bytes

    grid = [unwrap(b)[0] for b in beams if is_grid_beam(b)]
    pixels = max((int(np.prod(g.data_array.shape[-2:])) + 3
                  * int(g.data_array.shape[-2]) for g in grid), default=0)
    per_freq = nfreq > 1 and any(np.asarray(g.freq_array).size > 1
                                 for g in grid)

    def need(ch):
        npc = int(np.ceil(nsrc / ch))
        return device_bytes(
            form=form, nterms=nterms, nprod=nprod, nant=nant,
            polarized=polarized, nfreq=nfreq, nsrc=nsrc, nchunks=ch,
            chunk_size=npc, nsrc_alloc=int(npc * source_buffer),
            nbeam=len(beams), beam_pixels=pixels,
            beam_sets=nfreq if per_freq else 1, npairs=npairs,
            precision=precision)["total"]

    limit = max(nchunks, min(nsrc, 1000))
    ch = nchunks
    while ch < limit and need(ch) > budget:
        ch = min(limit, ch + max(1, ch // 8))
    if need(ch) > budget:
        logger.warning(
            f"device footprint {need(ch) / 2**30:.2f} GiB with {ch} chunks "
            f"exceeds the budget of {budget / 2**30:.2f} GiB")
    elif ch != nchunks:
        logger.info(f"{ch} chunks (not {nchunks}) to fit this build's "
                    f"footprint in {budget / 2**30:.2f} GiB")
    return ch, int(np.ceil(nsrc / ch))


def simulate_multi(**kwargs):
    """All frequencies at once: returns (nfreq, ntimes, npairs, nfeed, nfeed).

    Keywords as :func:`simulate` with ``freqs`` for ``freq`` and ``fluxes``
    (NSRCS, NFREQS) for ``I_sky``.  Equivalent to calling :func:`simulate`
    once per frequency (wrapper.py:180-199), with the rotation shared.
    """
    engine, ntimes = build_engine(**kwargs)
    return engine.run(ntimes)


def simulate(
    *,
    antpos,
    freq: float,
    times,
    skycoords,
    telescope_loc,
    I_sky,
    beam_list,
    polarized: bool = False,
    antpairs=None,
    beam_idx=None,
    max_memory=np.inf,
    min_chunks: int = 1,
    precision: int = 1,
    beam_spline_opts: dict | None = None,
    coord_method: str = "CoordinateRotationAstropy",
    matprod_method: str = "GPUMatMul",
    source_buffer: float = 1.0,
    coord_method_params: dict | None = None,
    memory_buffer: float = 0.9,
    split_terms: int | None = None,
    split_products: int | None = None,
) -> np.ndarray:
    """GPU implementation of the visibility simulator at one frequency.

    Returns ``(NTIMES, NBLS, NFEED, NFEED)`` if polarized, else
    ``(NTIMES, NBLS)`` (gpu/gpu.py:256), complex64 or complex128 by precision.
    """
    vis = simulate_multi(
        antpos=antpos, freqs=[freq], times=times, skycoords=skycoords,
        telescope_loc=telescope_loc, fluxes=np.asarray(I_sky)[:, None],
        beam_list=beam_list, polarized=polarized, antpairs=antpairs,
        beam_idx=beam_idx, max_memory=max_memory, min_chunks=min_chunks,
        precision=precision, beam_spline_opts=beam_spline_opts,
        coord_method=coord_method, matprod_method=matprod_method,
        source_buffer=source_buffer, coord_method_params=coord_method_params,
        memory_buffer=memory_buffer, split_terms=split_terms,
        split_products=split_products,
    )[0]
    return vis if polarized else vis[:, :, 0, 0]
