"""Host-side helpers: dtypes and the source-chunk memory model.

Restated from matvis's ``_utils.py`` (get_required_chunks :124-209,
get_desired_chunks :212-288, get_dtypes :291-319).  The byte model is kept
term for term so that a given ``max_memory`` / ``min_chunks`` produces the same
chunking as the reference.
"""

from __future__ import annotations

import logging

import numpy as np

logger = logging.getLogger(__name__)


def get_dtypes(precision: int):
    """Return ``(real_dtype, complex_dtype)`` for precision 1 or 2."""
    if precision == 1:
        return np.float32, np.complex64
    return np.float64, np.complex128


def get_required_chunks(
    freemem: int,
    nax: int,
    nfeed: int,
    nant: int,
    nsrc: int,
    nbeam: int,
    nbeampix: int,
    precision: int,
    source_buffer: float = 1.0,
    memory_buffer: float = 0.9,
) -> int:
    """Number of source chunks needed to fit in ``freemem * memory_buffer``.

    Same byte model as _utils.py:186-199 (antpos, flux, beam, coordinates,
    per-chunk exptau/beam/Z, and one visibility matrix per chunk).
    """
    if not (0 < memory_buffer <= 1.0):
        raise ValueError("memory_buffer must be in the range (0, 1]")

    rsize = 4 * precision
    csize = 2 * rsize

    total = freemem
    ch = 0
    while total >= freemem * memory_buffer and ch < 100:
        ch += 1
        nchunk = int(nsrc // ch * source_buffer)
        sizes = {
            "antpos": nant * 3 * rsize,
            "flux": nsrc * rsize,
            "beam": nbeampix * nfeed * nax * csize,
            "crd_eq": 3 * nsrc * rsize,
            "eq2top": 3 * 3 * rsize,
            "crd_top": 3 * nsrc * rsize,
            "crd_chunk": 3 * nchunk * rsize,
            "flux_chunk": nchunk * rsize,
            "exptau": nant * nchunk * csize,
            "beam_interp": nbeam * nfeed * nax * nchunk * csize,
            "zmat": nchunk * nfeed * nant * nax * csize,
            "vis": ch * nfeed * nant * nfeed * nant * csize,
        }
        total = sum(sizes.values())
    ch = max(ch, 1)
    logger.info(
        f"Total free mem: {freemem / 1024**3:.2f} GB. Requires {ch} chunks "
        f"(estimate {total / 1024**3:.2f} GB)"
    )
    return ch


def get_desired_chunks(
    freemem,
    min_chunks: int,
    beam_list,
    nax: int,
    nfeed: int,
    nant: int,
    nsrc: int,
    precision: int,
    source_buffer: float = 1.0,
    memory_buffer: float = 0.9,
) -> tuple[int, int]:
    """``(nchunks, sources per chunk)`` (_utils.py:259-288)."""
    nbeampix = sum(
        beam.data_array.shape[-2] * beam.data_array.shape[-1]
        for beam in beam_list
        if hasattr(beam, "data_array")
    )
    nchunks = min(
        max(
            min_chunks,
            get_required_chunks(
                freemem, nax, nfeed, nant, nsrc, len(beam_list), nbeampix,
                precision, source_buffer, memory_buffer,
            ),
        ),
        nsrc,
    )
    return nchunks, int(np.ceil(nsrc / nchunks))
