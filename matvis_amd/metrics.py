"""Error metrics of visibilities against a reference.

The normalised error of element (i, j, p, q) of the API layout
``V[i*nant + j, p, q]`` is ``|V - V_ref| / sqrt(|A_iq| |A_jp|)`` with
``A_ip = V_ref[i*nant + i, p, p]`` the reference autocorrelations: the
element pairs ``conj(Z[i, q])`` with ``Z[j, p]``, so the two autocorrelations
bound it by Cauchy-Schwarz.  Its maximum over elements is the figure quoted
against the 1e-6 bound; the per-element relative error ``|V - V_ref| /
|V_ref|`` is reported beside it.
"""

from __future__ import annotations

import numpy as np


def normalized_error(v_test, v_ref, nant: int, nfeed: int):
    """``(max, median)`` normalised and ``(max, median)`` per-element relative
    error; ``v_*``: (..., nant*nant, nfeed, nfeed)."""
    t = np.asarray(v_test).reshape(-1, nant, nant, nfeed, nfeed)
    r = np.asarray(v_ref).reshape(-1, nant, nant, nfeed, nfeed)
    a = np.abs(r[:, np.arange(nant), np.arange(nant)])  # (T, nant, nf, nf)
    auto = np.diagonal(a, axis1=2, axis2=3)  # (T, nant, nf): A[i, p]
    # scale[T, i, j, p, q] = sqrt(A[i, q] A[j, p])
    scale = np.sqrt(auto[:, :, None, None, :] * auto[:, None, :, :, None])
    d = np.abs(t - r)
    ar = np.abs(r)
    with np.errstate(divide="ignore", invalid="ignore"):
        ne = np.where(scale > 0, d / scale, 0.0)
        re = np.where(ar > 0, d / ar, 0.0)
    return (float(ne.max()), float(np.median(ne)), float(re.max()),
            float(np.median(re)))
