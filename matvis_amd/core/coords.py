"""Coordinate rotation plugin: ECI -> topocentric ENU, horizon cull, compaction.

Contract of matvis's ``CoordinateRotation`` (core/coords.py:20-179): a
registry of subclasses by class name, ``setup()`` allocating the buffers,
``rotate(t)`` for one time index, and ``select_chunk(chunk, t)`` returning the
above-horizon coordinates, their fluxes (zero past ``n``) and ``n``.

Subclasses provide the per-time rotation either as a 3x3 matrix R_t
(:meth:`rotation_matrix`; the device applies it fused with the cull) or by
filling topocentric vectors themselves (:meth:`set_topocentric`).

A polarized sky model is a flux of shape (nsrc, nfreq, 2, 2), real or
complex: the coherency matrix of each source per frequency in the RA/Dec
basis.  ``select_chunk`` then returns the kept sources' coherency matrices
rotated into the alt/az basis (``mv_coherency_rotate``), which needs the
time's average ICRS -> (north, east, up) rotation (:meth:`average_rotation`).
"""

from __future__ import annotations

from abc import ABC

import numpy as np

from .. import _device as D
from .._lib import call, lib
from .._utils import get_dtypes
from ..astro import skycoord_ra_dec
from ..coordinates import point_source_crd_eq

_EYE = np.eye(3, dtype=np.float64)


class CoordinateRotation(ABC):
    """Base class; every subclass registers itself in ``_methods``."""

    _methods: dict = {}
    requires_gpu: bool = False

    def __init_subclass__(cls) -> None:
        CoordinateRotation._methods[cls.__name__] = cls
        return super().__init_subclass__()

    def __init__(
        self,
        flux: np.ndarray,
        times,
        telescope_loc,
        skycoords,
        chunk_size: int | None = None,
        source_buffer: float = 0.55,
        precision: int = 1,
        gpu: bool = True,
    ):
        if not gpu:
            raise ValueError(
                "matvis_amd coordinate rotations run on the GPU (gpu=True)"
            )
        self.gpu = True
        self.precision = precision
        self.rtype, self.ctype = get_dtypes(precision)
        flux = np.asarray(flux)
        self._polarized = flux.ndim == 4
        if self._polarized:
            if flux.shape[2:] != (2, 2):
                raise ValueError(
                    "a polarized flux must have shape (NSRCS, NFREQ, 2, 2)")
        elif flux.ndim != 1 or np.iscomplexobj(flux):
            raise NotImplementedError(
                "flux must be real with shape (NSRCS,), or a coherency "
                "matrix per source and frequency, (NSRCS, NFREQ, 2, 2)"
            )
        # dtype rule of core/coords.py:59-63
        self.sky_model_dtype = (self.ctype if np.iscomplexobj(flux)
                                else self.rtype)
        self.flux = (flux.astype(self.sky_model_dtype) if self._polarized
                     else flux.astype(np.float64))
        self.nsrc = len(flux)
        self.times = times
        self.telescope_loc = telescope_loc
        self.skycoords = skycoords
        if hasattr(times, "ndim"):
            assert times.ndim == 1
        assert len(skycoords) == self.nsrc
        self.chunk_size = chunk_size or self.nsrc
        self.source_buffer = source_buffer
        # allocation rule of core/coords.py:79-82
        if self.chunk_size > 1000:
            self.nsrc_alloc = int(self.chunk_size * self.source_buffer)
        else:
            self.nsrc_alloc = self.chunk_size
        self._rot = None
        # "eci" (R_t applied to the ICRS vectors), "topo" (host-filled ENU
        # vectors, identity rotation) or "bcrs" (the CIRS directions of the
        # ERFA-style path, with its per-time CIRS -> ENU matrix)
        self._mode = None
        self._topo_valid = False

    # ----------------------------------------------------------- buffers

    def setup(self):
        """Allocate device buffers and upload source vectors and fluxes."""
        ra, dec = skycoord_ra_dec(self.skycoords)
        self.eci = D.to_dev(point_source_crd_eq(ra, dec), np.float64)
        self._topo = D.zeros((3, self.nsrc), self.rtype)
        self._topo64 = None
        self.coords_above_horizon = D.zeros((3, self.nsrc_alloc), self.rtype)
        if self._polarized:
            self.flux_dev = D.to_dev(self.flux)
            self.flux_above_horizon = D.zeros(
                (self.nsrc_alloc,) + self.flux.shape[1:], self.sky_model_dtype)
            self._ra = D.to_dev(np.asarray(ra, np.float64))
            self._dec = D.to_dev(np.asarray(dec, np.float64))
            self._idx = D.zeros((self.nsrc_alloc,), np.int32)
        else:
            self.flux_dev = D.to_dev(self.flux, np.float64)
            self.flux_above_horizon = D.zeros((self.nsrc_alloc,), self.rtype)
        self.n_up_dev = D.zeros((1,), np.int32)
        nbytes = lib().mv_rotate_cull_workspace_bytes(self.chunk_size)
        self.workspace = D.empty((nbytes,), np.uint8)

    # ----------------------------------------------------------- rotation

    def rotation_matrix(self, t: int) -> np.ndarray | None:
        """R_t (3x3, ECI -> ENU) for time index t, or None if the subclass
        fills topocentric vectors itself."""
        return None

    def rotate(self, t: int) -> None:
        """Prepare the rotation for time index t."""
        rot = self.rotation_matrix(t)
        if rot is None:
            raise NotImplementedError(
                f"{type(self).__name__} must implement rotation_matrix() or "
                "override rotate() and call set_topocentric()"
            )
        self._rot = np.ascontiguousarray(rot, dtype=np.float64)
        self._mode = "eci"
        self._topo_valid = False

    def average_rotation(self, t: int) -> np.ndarray:
        """3x3 rotation carrying ICRS unit vectors to the (north, east, up)
        frame at time index t: the role of ``_calc_average_rotation_matrix``
        (coordinates.py:493-534).  Here R_t with its first two rows swapped;
        subclasses without an R_t override it."""
        rot = self.rotation_matrix(t)
        if rot is None:
            raise NotImplementedError(
                f"{type(self).__name__} must implement average_rotation() "
                "for polarized sky models"
            )
        return np.ascontiguousarray(np.asarray(rot, np.float64)[[1, 0, 2]])

    def set_topocentric(self, enu: np.ndarray) -> None:
        """Use host-computed ENU unit vectors (3, nsrc) for the current time."""
        if self._topo64 is None:
            self._topo64 = D.zeros((3, self.nsrc), np.float64)
        self._topo64.copy_(D.to_dev(np.asarray(enu, dtype=np.float64)))
        self._rot = _EYE
        self._mode = "topo"
        self._topo_valid = False

    def rotation_source(self):
        """Device (3, nsrc) f64 vectors the current rotation applies to."""
        if self._mode == "topo":
            return self._topo64
        if self._mode == "bcrs":
            return self._bcrs
        return self.eci

    @property
    def all_coords_topo(self):
        """Topocentric unit vectors of every source, (3, nsrc), rtype."""
        if self._mode is None:
            raise RuntimeError("call rotate(t) first")
        if not self._topo_valid:
            src = self.rotation_source()
            call("mv_rotate", D.ptr(src), self.nsrc, self._rot.ctypes.data,
                 self.nsrc, D.ptr(self._topo), int(self.precision == 2),
                 self.nsrc, D.stream_ptr())
            self._topo_valid = True
        return self._topo

    # ----------------------------------------------------------- cull

    def select_chunk_async(self, chunk: int, t: int):
        """Enqueue the fused rotate/cull/compaction of one chunk; the count
        stays on the device in ``n_up_dev``."""
        if self._mode is None:
            raise RuntimeError("call rotate(t) first")
        src = self.rotation_source()
        chunk0 = chunk * self.chunk_size
        chunk_len = max(0, min(self.chunk_size, self.nsrc - chunk0))
        chunk0 = min(chunk0, self.nsrc)
        pol = self._polarized
        f64 = int(self.precision == 2)
        call(
            "mv_rotate_cull", D.ptr(src), self.nsrc, self._rot.ctypes.data,
            None if pol else D.ptr(self.flux_dev), chunk0, chunk_len,
            self.nsrc_alloc, D.ptr(self.coords_above_horizon), self.nsrc_alloc,
            None if pol else D.ptr(self.flux_above_horizon), f64,
            D.ptr(self._idx) if pol else None, D.ptr(self.n_up_dev),
            D.ptr(self.workspace), self.workspace.numel(), D.stream_ptr(),
        )
        if pol:
            r_avg = np.ascontiguousarray(self.average_rotation(t), np.float64)
            if r_avg.shape != (3, 3):
                raise ValueError("average_rotation() must return a 3x3 matrix")
            call(
                "mv_coherency_rotate", D.ptr(self._ra), D.ptr(self._dec),
                self.nsrc, D.ptr(self._idx), D.ptr(self.coords_above_horizon),
                self.nsrc_alloc, f64, D.ptr(self.n_up_dev), self.nsrc_alloc,
                r_avg.ctypes.data, D.ptr(self.flux_dev), self.flux.shape[1],
                f64, int(np.iscomplexobj(self.flux)),
                D.ptr(self.flux_above_horizon), D.stream_ptr(),
            )
        return self.coords_above_horizon, self.flux_above_horizon, self.n_up_dev

    def select_chunk(self, chunk: int, t: int):
        """Cull and compact one chunk; returns ``(crd, flux, n)`` with the
        reference's error if ``n`` exceeds ``nsrc_alloc``
        (core/coords.py:117-121)."""
        crd, flux, n_dev = self.select_chunk_async(chunk, t)
        n = int(n_dev.item())
        if n > self.nsrc_alloc:
            raise ValueError(
                f"nsrc_alloc ({self.nsrc_alloc}) is too small for the number of "
                f"sources above horizon ({n}). Try increasing source_buffer."
            )
        return crd, flux, n
