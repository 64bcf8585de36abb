"""Host driver of the Hermitian rank-k update kernels (``mv_herk``).

A :class:`Herk` fixes the numeric form and the row count m = nfeed * nant,
owns the split-K workspace, and picks the number of K-slices.
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


def cdiv(a: int, b: int) -> int:
    return -(-a // b)


def is_split(form: int) -> bool:
    return form in (FORM_SPLIT_BF16, FORM_SPLIT_F16)


class Herk:
    """V (+)= scale * conj(Z) Z^T over the upper tiles, into an FP64 buffer."""

    def __init__(self, form: int, m: int, nterms: int = 1, nprod: int = 1,
                 target_workgroups: int = 1024, min_ktiles: int = 4):
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
        self.target_workgroups = target_workgroups
        self.min_ktiles = min_ktiles
        self._ws = None

    def nslices(self, kmax: int) -> int:
        """K-slices for at most kmax complex columns: enough workgroups to
        fill the chip, each with at least `min_ktiles` K tiles."""
        ktiles = max(1, cdiv(kmax, K_TILE))
        ns = cdiv(self.target_workgroups, self.ntiles)
        return max(1, min(ns, ktiles // self.min_ktiles))

    def workspace(self, nslices: int):
        need = lib().mv_herk_workspace_bytes(self.form, self.m, nslices)
        if self._ws is None or self._ws.numel() < need:
            self._ws = D.empty((max(need, 256),), np.uint8)
        return self._ws

    def new_vacc(self):
        return D.zeros((self.ldv, self.ldv), np.complex128)

    def run(self, z_terms, ldz: int, nseg: int, seg_stride: int, seg_len: int,
            vacc, accumulate: bool, seg_len_dev=None, scale: float = 1.0,
            nslices: int | None = None, stream=None):
        kmax = nseg * seg_len
        ns = nslices or self.nslices(kmax)
        ws = self.workspace(ns)
        zp = [D.ptr(t) for t in z_terms] + [None] * (3 - len(z_terms))
        call("mv_herk", self.form, self.nterms, self.nprod, zp[0], zp[1],
             zp[2], self.m, ldz, nseg, seg_stride, seg_len,
             D.ptr(seg_len_dev), ns, float(scale), D.ptr(vacc), self.ldv,
             int(bool(accumulate)), D.ptr(ws), ws.numel(),
             D.stream_ptr(stream))
        return vacc


def expand(vacc, ldv: int, tile: int, nant: int, nfeed: int, pairs_dev,
           npairs: int, out, stream=None):
    """API layout (npairs, nfeed, nfeed) from the FP64 accumulator."""
    call("mv_vis_expand", D.ptr(vacc), ldv, tile, nant, nfeed,
         D.ptr(pairs_dev), npairs, D.ptr(out),
         int(out.dtype == D.tdtype(np.complex128)), D.stream_ptr(stream))
    return out
