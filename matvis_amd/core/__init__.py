"""Plugin contracts of the per-time visibility computation.

Same class names, constructor keywords and ``setup()`` / ``__call__``
contracts as matvis's ``core`` package; the implementations call the HIP
kernels of ``libmatvis_hip.so`` and hold their buffers in device memory.
"""

import numpy as np


def _validate_inputs(precision, polarized, antpos, times, I_sky):
    """Validate input shapes and types (core/__init__.py:11-32)."""
    assert precision in {1, 2}
    if polarized:
        nax = nfeed = 2
    else:
        nax = nfeed = 1
    nant, ncrd = np.shape(antpos)
    assert ncrd == 3, "antpos must have shape (NANTS, 3)."
    ntimes = len(times)
    assert np.ndim(I_sky) == 1, "I_sky must have shape (NSRCS,)."
    return nax, nfeed, nant, ntimes
