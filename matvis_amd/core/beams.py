"""Beam plugin contract (matvis core/beams.py:15-208)."""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Literal

import numpy as np
import torch

from .. import _device as D
from ..beams import BeamInterface, unwrap


def prepare_beam_unpolarized(
    beam,
    use_feed: Literal["x", "y"] = "x",
    allow_beam_mutation: bool = False,
):
    """Mark a beam for unpolarized use (core/beams.py:15-32).

    The E-field -> xx power conversion itself happens when the grid is
    prepared for the device (:func:`matvis_amd.beams.grid_planes`), as the
    reference's ``as_power_beam`` + ``with_feeds(["x"])`` produce it.
    """
    inner, bt = unwrap(beam)
    if bt == "power" and isinstance(beam, BeamInterface):
        return beam
    return BeamInterface(inner, beam_type="power")


def _wrangle_beams(beam_idx, beam_list, polarized: bool, nant: int, freq: float):
    """Checks on the beam list and beam indices (core/beams.py:35-99)."""
    nbeam = len(beam_list)
    beam_list = [
        b if isinstance(b, BeamInterface) else BeamInterface(*unwrap(b))
        for b in beam_list
    ]
    if not polarized:
        beam_list = [prepare_beam_unpolarized(b) for b in beam_list]
    if beam_idx is None and nbeam not in (1, nant):
        raise ValueError(
            "If number of beams provided is not 1 or nant, beam_idx must be "
            "provided."
        )
    if beam_idx is not None:
        beam_idx = np.asarray(beam_idx)
        if beam_idx.shape != (nant,):
            raise ValueError("beam_idx must be length nant")
        if not all(0 <= i < nbeam for i in beam_idx):
            raise ValueError(
                "beam_idx contains indices greater than the number of beams"
            )
    if polarized and any(b.beam_type != "efield" for b in beam_list):
        raise ValueError("beam type must be efield if using polarized=True")
    return beam_list, nbeam, beam_idx


class BeamInterpolator(ABC):
    """Base class of beam evaluation (core/beams.py:102-208)."""

    def __init__(
        self,
        beam_list,
        beam_idx,
        polarized: bool,
        nant: int,
        freq: float,
        nsrc: int,
        spline_opts: dict | None = None,
        precision: int = 1,
    ):
        self.beam_list, self.nbeam, self.beam_idx = _wrangle_beams(
            beam_idx=beam_idx, beam_list=beam_list, polarized=polarized,
            nant=nant, freq=freq,
        )
        self.polarized = polarized
        self.nant = nant
        self.freq = freq
        self.spline_opts = spline_opts or {}
        self.nfeed = self.nax = 2 if polarized else 1
        if precision == 1:
            self.complex_dtype, self.real_dtype = np.complex64, np.float32
        else:
            self.complex_dtype, self.real_dtype = np.complex128, np.float64
        self.precision = precision
        self.nsrc = nsrc

    def setup(self):
        self.interpolated_beam = D.zeros(
            (self.nbeam, self.nfeed, self.nax, self.nsrc), self.complex_dtype
        )

    @abstractmethod
    def interp(self, tx, ty, out):
        """Fill ``out`` (nbeam, nfeed, nax, nsrc) at positions (tx, ty)."""

    def __call__(self, tx, ty, check: bool = True):
        self.interp(tx, ty, self.interpolated_beam)
        if check:
            if not bool(torch.isfinite(self.interpolated_beam).all()):
                raise ValueError("Beam interpolation resulted in an invalid value")
        return self.interpolated_beam
