"""Host-side helpers: dtypes, the source-chunk memory model and progress
logging.

Restated from matvis's ``_utils.py`` (get_required_chunks :124-209,
get_desired_chunks :212-288, get_dtypes :291-319; ceildiv, human_readable_size
and log_progress :30-123).  The byte model is kept term for term so that a
given ``max_memory`` / ``min_chunks`` produces the same chunking as the
reference.
"""

from __future__ import annotations

import datetime
import logging
import time

import numpy as np

logger = logging.getLogger(__name__)

_SIZE_UNITS = ("B", "KiB", "MiB", "GiB", "TiB", "PiB")


def ceildiv(a: int, b: int) -> int:
    """Integer a / b rounded up (src/matvis/_utils.py:30-35)."""
    return -(a // -b)


def human_readable_size(size, decimal_places: int = 2,
                        indicate_sign: bool = False) -> str:
    """A byte count in binary units, B to PiB (beyond 1024 PiB it stays in
    PiB), with a leading sign when ``indicate_sign`` (src/matvis/_utils.py:38-52)."""
    value, k = float(size), 0
    while abs(value) >= 1024.0 and k < len(_SIZE_UNITS) - 1:
        value /= 1024.0
        k += 1
    sign = "+" if indicate_sign else ""
    return f"{value:{sign}.{decimal_places}f} {_SIZE_UNITS[k]}"


def log_progress(start_time: float, prev_time: float, iters: int, niters: int,
                 pr, last_mem: float) -> tuple[float, int]:
    """At INFO level, log iterations done of ``niters``, the time since the
    last report and since the start, the projected total and the process's
    resident memory with its change since ``last_mem`` (``pr`` a
    ``psutil.Process``); return the new (report time, resident bytes)
    (src/matvis/_utils.py:92-123)."""
    if not logger.isEnabledFor(logging.INFO):
        return prev_time, last_mem
    now = time.time()
    total = datetime.timedelta(seconds=now - start_time)
    per_iter = total / max(iters, 1)
    expected = per_iter * niters
    rss = pr.memory_info().rss
    logger.info(
        f"Progress: {iters}/{niters} times ({100 * iters / niters:.1f}%); "
        f"since last report {datetime.timedelta(seconds=now - prev_time)}; "
        f"total {total} ({per_iter} a time), expected {expected} "
        f"({expected - total} left); resident memory "
        f"{human_readable_size(rss)} "
        f"({human_readable_size(rss - last_mem, indicate_sign=True)})")
    return now, rss


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
