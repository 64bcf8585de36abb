"""Source-summing plugin contract (matvis core/matprod.py:11-115)."""

from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np

from .._utils import get_dtypes


class MatProd(ABC):
    """Per-chunk ``V = conj(Z) Z^T`` followed by a chunk sum.

    ``antpairs`` None means every ordered pair ``(i, j)``, i-major, which is
    the order of the API output (core/matprod.py:37-43).
    """

    def __init__(self, nchunks: int, nfeed: int, nant: int, antpairs,
                 precision=1):
        if antpairs is None:
            self.all_pairs = True
            self.antpairs = np.array(
                [(i, j) for i in range(nant) for j in range(nant)]
            ).reshape(-1, 2)
        else:
            self.all_pairs = False
            self.antpairs = np.asarray(antpairs).reshape(-1, 2)
            if self.antpairs.size and (
                self.antpairs.min() < 0 or self.antpairs.max() >= nant
            ):
                raise ValueError("antpairs contains an antenna index out of range")
        self.nchunks = nchunks
        self.nfeed = nfeed
        self.nant = nant
        self.npairs = len(self.antpairs)
        self.precision = precision
        self.ctype = get_dtypes(precision)[1]
        self.ant1_idx = self.antpairs[:, 0]
        self.ant2_idx = self.antpairs[:, 1]

    @abstractmethod
    def allocate_vis(self):
        """Allocate ``self.vis``, indexable by chunk."""

    def setup(self):
        self.allocate_vis()

    @abstractmethod
    def compute(self, z, out):
        """Source sum of one chunk's Z into ``out``."""

    def __call__(self, z, chunk: int):
        self.compute(z, out=self.vis[chunk])
        return self.vis[chunk]

    @abstractmethod
    def sum_chunks(self, out):
        """Sum the chunk partials into the host array ``out``
        (npairs, nfeed, nfeed)."""
