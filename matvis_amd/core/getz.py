"""Z matrix plugin (matvis core/getz.py:13-105) on the device."""

from __future__ import annotations

import numpy as np

from .. import _device as D
from .._lib import call


class ZMatrixCalc:
    """``Z = A sqrt(I) exp(tau)`` of shape ``(nfeed * nant, nax * nsrc)``.

    Rows are ``ant * nfeed + feed`` and columns ``ax * nsrc + src``
    (core/getz.py:82-100).  As in the reference, ``exptau`` is multiplied
    by ``sqrt_flux`` in place (core/getz.py:80).
    """

    def __init__(self, nant: int, nfeed: int, nax: int, nsrc: int, ctype,
                 gpu: bool = True):
        if not gpu:
            raise ValueError("matvis_amd's ZMatrixCalc runs on the GPU")
        self.nant, self.nfeed, self.nax, self.nsrc = nant, nfeed, nax, nsrc
        self.ctype = np.dtype(ctype).type
        self.gpu = True
        self._bidx_key = None
        self._bidx_dev = None

    def setup(self):
        self.z = D.zeros((self.nfeed * self.nant, self.nax * self.nsrc),
                         self.ctype)

    def _beam_idx(self, beam_idx):
        if beam_idx is None:
            return None
        key = tuple(np.asarray(beam_idx).tolist()) if not hasattr(
            beam_idx, "data_ptr") else ("dev", beam_idx.data_ptr())
        if key != self._bidx_key:
            self._bidx_dev = D.as_dev(beam_idx, np.int32)
            self._bidx_key = key
        return self._bidx_dev

    def __call__(self, sqrt_flux, beam, exptau, beam_idx):
        is_f64 = self.ctype == np.complex128
        rtype = np.float64 if is_f64 else np.float32
        flux = D.as_dev(sqrt_flux, rtype)
        beam_t = D.as_dev(beam, self.ctype)
        if exptau.dtype != D.tdtype(self.ctype) or not exptau.is_contiguous():
            raise ValueError("exptau must be a contiguous device array of the "
                             "calculator's complex dtype")
        nbeam = beam_t.shape[0]
        if tuple(beam_t.shape) != (nbeam, self.nfeed, self.nax, self.nsrc):
            raise ValueError(f"beam has shape {tuple(beam_t.shape)}, expected "
                             f"(nbeam, {self.nfeed}, {self.nax}, {self.nsrc})")
        call("mv_zmat_parts", D.ptr(flux), D.ptr(beam_t),
             D.ptr(self._beam_idx(beam_idx)), nbeam, D.ptr(exptau), self.nant,
             self.nfeed, self.nax, self.nsrc, D.ptr(self.z), int(is_f64),
             D.stream_ptr())
        return self.z
