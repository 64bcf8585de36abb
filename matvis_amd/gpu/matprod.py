"""GPU source-summing plugins: HERK on MFMA tiles.

``GPUMatMul`` replaces matvis's cuBLAS ZGEMM/CGEMM (gpu/matprod.py:10-63,
gpu/_cublas.py:12-65) with an upper-triangle HERK accumulated in FP64.
``GPUVectorDot`` (gpu/matprod.py:66-98) computes one conj(z_i) z_j^T per
listed pair; here the HERK already produces every pair in one pass, so it
shares GPUMatMul's kernel and differs only in which pairs it returns.
The split classes are the mixed-precision forms: Z is split into 16-bit terms
and the significant term products run on BF16 / FP16 MFMA with FP32
accumulation, summed across K-slices in FP64.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .. import _device as D
from .._lib import FORM_C64, FORM_C128, FORM_SPLIT_BF16, FORM_SPLIT_F16, call
from ..core.matprod import MatProd
from .herk import Herk, expand, products_for, split_terms_alloc


class GPUMatMul(MatProd):
    """Plain form: complex128 on FP64 MFMA (precision 2) or complex64 on
    FP32 MFMA (precision 1)."""

    split_form: int | None = None
    default_terms = 1
    default_products = 1

    def __init__(self, nchunks: int, nfeed: int, nant: int, antpairs,
                 precision=1, split_terms: int | None = None,
                 split_products: int | None = None):
        super().__init__(nchunks, nfeed, nant, antpairs, precision)
        self.nterms = split_terms or self.default_terms
        self.nprod = (products_for(self.nterms, split_products)
                      if self.split_form is not None
                      else split_products or self.default_products)
        if self.split_form is None:
            self.form = FORM_C128 if precision == 2 else FORM_C64
        else:
            self.form = self.split_form

    def allocate_vis(self):
        self.m = self.nfeed * self.nant
        self.herk = Herk(self.form, self.m, self.nterms, self.nprod)
        ldv = self.herk.ldv
        self._vis_all = D.zeros((self.nchunks, ldv, ldv), np.complex128)
        self.vis = [self._vis_all[i] for i in range(self.nchunks)]
        self._vsum = (D.zeros((ldv, ldv), np.complex128)
                      if self.nchunks > 1 else None)
        self._pairs_dev = (None if self.all_pairs
                           else D.to_dev(self.antpairs, np.int32))
        self._out_dev = D.empty((self.npairs, self.nfeed, self.nfeed),
                                self.ctype)

    def _check_z(self, z):
        if not isinstance(z, torch.Tensor):
            z = D.to_dev(np.asarray(z))
        if z.dim() != 2 or z.shape[0] != self.m:
            raise ValueError(f"z must have shape ({self.m}, K); got "
                             f"{tuple(z.shape)}")
        if z.dtype not in (torch.complex128, torch.complex64):
            raise TypeError(f"invalid dtype: {z.dtype}")
        if z.stride(1) != 1:
            z = z.contiguous()
        return z

    def compute(self, z, out):
        z = self._check_z(z)
        want = torch.complex128 if self.form == FORM_C128 else torch.complex64
        if z.dtype != want:
            z = z.to(want)
        k = z.shape[1]
        self.herk.run([z], ldz=z.stride(0), nseg=1, seg_stride=k, seg_len=k,
                      vacc=out, accumulate=False)
        return out

    def summed_device(self):
        """The FP64 chunk sum on the device (ldv x ldv, upper tiles)."""
        if self.nchunks == 1:
            return self.vis[0]
        ldv = self.herk.ldv
        call("mv_vacc_sum", D.ptr(self._vis_all), self.nchunks, 2 * ldv * ldv,
             D.ptr(self._vsum), 2 * ldv * ldv, D.stream_ptr())
        return self._vsum

    def sum_chunks(self, out):
        """Chunk sum, expansion to (npairs, nfeed, nfeed) and copy to host
        (gpu/matprod.py:38-63)."""
        v = self.summed_device()
        expand(v, self.herk.ldv, self.herk.tile, self.nant, self.nfeed,
               self._pairs_dev, self.npairs, self._out_dev)
        out[:] = D.to_host(self._out_dev)
        return out


class GPUVectorDot(GPUMatMul):
    """Per-pair source sums; the HERK covers all pairs, the expansion keeps
    the listed ones."""


#: matvis's wrapper accepts "VectorLoop" (wrapper.py:48-55) although its
#: backends name the class VectorDot; both names resolve here.
GPUVectorLoop = GPUVectorDot


class GPUMatMulSplitBF16(GPUMatMul):
    """Mixed precision: Z = hi + lo in bfloat16, products hi*hi + hi*lo +
    lo*hi on BF16 MFMA (``split_products=4`` adds lo*lo), FP32 accumulation
    per K-slice, FP64 across slices, the diagonal summed exactly."""

    split_form = FORM_SPLIT_BF16
    default_terms = 2
    default_products = 3

    def _scale(self, z) -> float:
        return 1.0

    def compute(self, z, out):
        z = self._check_z(z)
        k = z.shape[1]
        # the HERK reads whole 16-column K tiles: columns [k, ncols) of the
        # last one must be zero
        ncols = max(16, -(-k // 16) * 16)
        if (getattr(self, "_terms", None) is None
                or self._terms[0].shape[0] * 16 < ncols):
            self._terms = split_terms_alloc(self.nterms, self.m, ncols)
        elif k % 16:
            for term in self._terms:
                term[k // 16].zero_()
        t = self._terms + [None] * (3 - self.nterms)
        s = self._scale(z)
        call("mv_split_z", D.ptr(z), int(z.dtype == torch.complex128), self.m,
             1, k, z.stride(0), k, self.form, self.nterms, s, D.ptr(t[0]),
             D.ptr(t[1]), D.ptr(t[2]), self.m, ncols, D.stream_ptr())
        self.herk.run(self._terms, ldz=self.m, nseg=1, seg_stride=ncols,
                      seg_len=k, vacc=out, accumulate=False,
                      scale=1.0 / (s * s))
        return out


class GPUMatMulSplitF16(GPUMatMulSplitBF16):
    """Mixed precision with float16 terms (11-bit significands) and a
    power-of-two scale per chunk that puts max|Z| near 2^14; default
    products hi*hi + hi*lo + lo*hi."""

    split_form = FORM_SPLIT_F16
    default_terms = 2
    default_products = 3

    def _scale(self, z) -> float:
        zmax = float(torch.view_as_real(z).abs().max())
        return f16_scale(zmax)


def f16_scale(bound: float) -> float:
    """Power of two that maps magnitudes up to `bound` to at most 2^14."""
    if not bound > 0 or not math.isfinite(bound):
        return 1.0
    return 2.0 ** (14 - math.ceil(math.log2(bound)))
