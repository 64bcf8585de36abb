"""exp(tau) plugin (matvis core/tau.py:16-71) on the device."""

from __future__ import annotations

import numpy as np
import torch

from .. import _device as D
from .._lib import call
from .._utils import get_dtypes

C_LIGHT = 299792458.0


class TauCalculator:
    """``exptau[a, s] = exp(+i 2 pi nu (b_a . s) / c)`` (core/tau.py:30-31)."""

    def __init__(self, antpos, freq: float, nsrc: int, precision: int = 1,
                 gpu: bool = True):
        if not gpu:
            raise ValueError("matvis_amd's TauCalculator runs on the GPU")
        self.rtype, self.ctype = get_dtypes(precision)
        self.precision = precision
        self.antpos_m = np.asarray(antpos, dtype=np.float64)
        self.nant = len(self.antpos_m)
        self.nsrc = nsrc
        self.freq = float(freq)
        self.gpu = True

    def setup(self):
        self.antpos = D.to_dev(self.antpos_m, np.float64)
        self.exptau = D.zeros((self.nant, self.nsrc), self.ctype)

    def __call__(self, crdtop):
        """Fill and return ``exptau`` for topocentric vectors (3, nsrc)."""
        crd = crdtop if isinstance(crdtop, torch.Tensor) else D.to_dev(
            np.asarray(crdtop))
        if crd.dtype not in (torch.float64, torch.float32):
            crd = crd.to(torch.float64)
        if crd.device.type != "cuda":
            crd = crd.to(D.require_gpu())
        if crd.stride(1) != 1:
            crd = crd.contiguous()
        n = crd.shape[1]
        if n > self.nsrc:
            raise ValueError(f"crdtop has {n} sources, exptau holds {self.nsrc}")
        call("mv_exptau", D.ptr(self.antpos), self.nant, self.freq, D.ptr(crd),
             int(crd.dtype == torch.float64), crd.stride(0), n,
             D.ptr(self.exptau), int(self.precision == 2), self.nsrc,
             D.stream_ptr())
        return self.exptau
