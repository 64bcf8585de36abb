"""Host driver of the Hermitian rank-k update kernels (``mv_herk``).

A :class:`Herk` fixes the numeric form and the row count m = nfeed * nant
and owns the split-K workspace.  The kernels choose the number of active
K-slices on the device (the column count after the horizon cull is only known
there); the host sizes the workspace for the largest count they may choose.
"""

from __future__ import annotations

import numpy as np

from .. import _device as D
from .._lib import (FORM_C64, FORM_C128, FORM_SPLIT_BF16, FORM_SPLIT_F16, call,
                    lib)

FORMS = {
    "c128": FORM_C128,
    "c64": FORM_C64,
    "bf16": FORM_SPLIT_BF16,
    "f16": FORM_SPLIT_F16,
}

#: Complex columns per K tile in both kernel families.
K_TILE = 16
#: (nterms, nprod) pairs the split kernels implement.
SPLIT_VARIANTS = {(2, 3), (2, 4), (3, 6)}
#: products taken when only the term count is given: 2 terms drop lo*lo,
#: 3 terms drop the three products below FP32 resolution (m*l, l*m, l*l)
DEFAULT_PRODUCTS = {2: 3, 3: 6}


def products_for(nterms: int, nprod: int | None) -> int:
    """The product count of a split form: as given, else the default for
    the term count."""
    return nprod or DEFAULT_PRODUCTS.get(nterms, 1)


def cdiv(a: int, b: int) -> int:
    return -(-a // b)


def split_terms_alloc(nterms: int, m: int, ncols: int) -> list:
    """Zeroed device term arrays of a split form for an m-row Z of `ncols`
    columns (a multiple of 16, segments included), K-tile-major as
    ``mv_zmat_fused`` and ``mv_split_z`` write them and the split HERK reads
    them: K tile q (16 columns) is m rows of 16 int32 words, so the arrays
    are (ncols / 16, m, 16) and the ``ldz`` / ``ldt`` to pass is m."""
    if ncols % K_TILE:
        raise ValueError(f"split terms need whole {K_TILE}-column K tiles "
                         f"(ncols={ncols})")
    return [D.zeros((ncols // K_TILE, m, K_TILE), np.int32)
            for _ in range(nterms)]


def is_split(form: int) -> bool:
    return form in (FORM_SPLIT_BF16, FORM_SPLIT_F16)


class Herk:
    """V (+)= scale * conj(Z) Z^T over the upper tiles, into an FP64 buffer."""

    def __init__(self, form: int, m: int, nterms: int = 1, nprod: int = 1):
        if is_split(form):
            if (nterms, nprod) not in SPLIT_VARIANTS:
                raise ValueError(
                    f"unsupported split (nterms={nterms}, nprod={nprod}); "
                    f"supported: {sorted(SPLIT_VARIANTS)}"
                )
        elif (nterms, nprod) != (1, 1):
            raise ValueError("plain forms take nterms=1, nprod=1")
        self.form, self.m, self.nterms, self.nprod = form, m, nterms, nprod
        self.tile = lib().mv_herk_tile(form)
        self.nb = cdiv(m, self.tile)
        self.ntiles = self.nb * (self.nb + 1) // 2
        self.ldv = self.nb * self.tile
        self._ws = None

    def nslices(self, kmax: int) -> int:
        """Slice capacity covering every count the kernels may choose for up
        to kmax complex columns."""
        return max(1, int(lib().mv_herk_max_slices(self.form, self.m,
                                                   int(kmax))))

    def active_slices(self, nseg: int, seg_len: int, nslices: int) -> int:
        """The slice count the kernels use for seg_len columns per segment."""
        n = lib().mv_herk_slices(self.form, self.nterms, self.nprod, self.m,
                                 nseg, int(seg_len), nslices)
        if n < 1:
            call("mv_herk_slices", self.form, self.nterms, self.nprod, self.m,
                 nseg, int(seg_len), nslices)
        return int(n)

    def workspace(self, nslices: int):
        need = lib().mv_herk_workspace_bytes(self.form, self.m, nslices)
        if self._ws is None or self._ws.numel() < need:
            self._ws = D.empty((max(need, 256),), np.uint8)
        return self._ws

    def new_vacc(self):
        return D.zeros((self.ldv, self.ldv), np.complex128)

    def run(self, z_terms, ldz: int, nseg: int, seg_stride: int, seg_len: int,
            vacc, accumulate: bool, seg_len_dev=None, scale: float = 1.0,
            nslices: int | None = None, stream=None, between=None):
        """Enqueue the HERK; ``between(stream)``, if given, runs after the
        compute kernel is enqueued and before the reduce."""
        kmax = nseg * seg_len
        ns = nslices or self.nslices(kmax)
        ws = self.workspace(ns)
        zp = [D.ptr(t) for t in z_terms] + [None] * (3 - len(z_terms))
        args = (self.form, self.nterms, self.nprod, zp[0], zp[1], zp[2],
                self.m, ldz, nseg, seg_stride, seg_len, D.ptr(seg_len_dev), ns,
                float(scale), D.ptr(vacc), self.ldv, int(bool(accumulate)),
                D.ptr(ws), ws.numel(), D.stream_ptr(stream))
        if between is None:
            call("mv_herk", *args)
        else:
            call("mv_herk_compute", *args)
            between(stream)
            call("mv_herk_reduce", *args)
        return vacc


def expand(vacc, ldv: int, tile: int, nant: int, nfeed: int, pairs_dev,
           npairs: int, out, stream=None):
    """API layout (npairs, nfeed, nfeed) from the FP64 accumulator."""
    call("mv_vis_expand", D.ptr(vacc), ldv, tile, nant, nfeed,
         D.ptr(pairs_dev), npairs, D.ptr(out),
         int(out.dtype == D.tdtype(np.complex128)), D.stream_ptr(stream))
    return out
