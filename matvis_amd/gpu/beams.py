"""GPU beam evaluation (matvis gpu/beams.py:15-200) on HIP kernels.

Grid beams are interpolated by ``mv_beam_interp`` (bilinear, or bicubic
B-spline on planes prefiltered once by ``mv_beam_prefilter``), whose taps
gather from L2 and the Infinity Cache; ``mv_beam_interp_banded`` stages the
grid in LDS by za band instead (the same values, opt-in: it measured slower
on every grid it fits).  Gaussian analytic beams are evaluated on the device
by ``mv_beam_gaussian`` instead of the reference's host round trip
(gpu/beams.py:95-101).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _device as D
from .._lib import call, lib
from ..beams import grid_planes, is_analytic_beam, is_grid_beam, unwrap
from ..core.beams import BeamInterpolator

_MIXED = ("GPUBeamInterpolator only supports beam_lists with either all UVBeam "
          "or all AnalyticBeam objects.")


def spline_order(spline_opts: dict | None) -> int:
    """Interpolation order from ``beam_spline_opts`` (default 1, as the
    reference's gpu_beam_interpolation, gpu/beams.py:135)."""
    opts = spline_opts or {}
    order = int(opts.get("order", 1))
    if order not in (1, 3):
        raise ValueError(f"beam interpolation order must be 1 or 3, got {order}")
    return order


class DeviceBeamSet:
    """Device-resident tables of a beam list at one frequency."""

    def __init__(self, beam_list, freq: float, polarized: bool, complex_dtype,
                 order: int = 1, banded: bool = False):
        """``banded``: interpolate with the grid staged in LDS by za band
        (the same values as the default tap gathers, which measured faster
        on MI355X for every grid whose rows fit in LDS); the bands need
        every beam to share beam 0's za spacing."""
        self.nbeam = len(beam_list)
        self.nfeed = self.nax = 2 if polarized else 1
        self.polarized = polarized
        self.complex_dtype = np.dtype(complex_dtype).type
        self.is_f64 = self.complex_dtype == np.complex128
        self.order = order
        self.use_interp = is_grid_beam(beam_list[0])
        if self.use_interp and not all(is_grid_beam(b) for b in beam_list):
            raise ValueError(_MIXED)
        if not self.use_interp:
            if any(is_grid_beam(b) for b in beam_list):
                raise ValueError(_MIXED)
            if not all(is_analytic_beam(b) for b in beam_list):
                raise NotImplementedError(
                    "only Gaussian analytic beams are evaluated on the device"
                )
        if self.use_interp:
            preps = [grid_planes(b, freq, polarized) for b in beam_list]
            shapes = {p[0].shape for p in preps}
            if len(shapes) != 1:
                raise ValueError("all grid beams must share one az/za grid shape")
            planes = np.stack([p[0] for p in preps])
            self.is_power = bool(preps[0][4])
            self.nza, self.naz = planes.shape[-2:]
            grid = np.array([[p[1], p[2], p[3]] for p in preps], np.float64)
            self.grid = D.to_dev(grid, np.float64)
            self.planes = D.to_dev(planes.astype(self.complex_dtype))
            amax = float(np.abs(planes).max()) if planes.size else 1.0
            self.max_abs = np.sqrt(amax) if self.is_power else amax
            if order == 3:
                call("mv_beam_prefilter", D.ptr(self.planes), int(self.is_f64),
                     self.nbeam * self.nax * self.nfeed, self.nza, self.naz,
                     D.stream_ptr())
            if banded and not np.all(grid[:, 1] == grid[0, 1]):
                raise ValueError("the za-band path needs every beam to share "
                                 "one za spacing")
            self.banded = bool(banded)
            self._band_ws = None
        else:
            sig = [unwrap(b)[0].get_sigma(freq) for b in beam_list]
            self.sigma = D.to_dev(np.asarray(sig, np.float64))
            self.is_power = False
            self.max_abs = 1.0
            self.banded = False

    def eval(self, crd, ld_crd: int, count: int, out, ld_out: int,
             n_dev=None, stream=None):
        """A (nbeam, nfeed, nax, ld_out) at topocentric vectors crd (3, ld)."""
        crd_is_f64 = int(crd.dtype == torch.float64)
        if self.use_interp and self.banded:
            need = lib().mv_beam_interp_banded_workspace_bytes(
                int(self.is_f64), self.nza, self.naz, count)
            if self._band_ws is None or self._band_ws.numel() < need:
                self._band_ws = D.empty((max(need, 1),), np.uint8)
            call("mv_beam_interp_banded", D.ptr(self.planes),
                 int(self.is_f64), self.nbeam, self.nfeed, self.nax, self.nza,
                 self.naz, D.ptr(self.grid), self.order, int(self.is_power),
                 D.ptr(crd), crd_is_f64, ld_crd, count, D.ptr(n_dev),
                 D.ptr(out), ld_out, D.ptr(self._band_ws),
                 self._band_ws.numel(), D.stream_ptr(stream))
        elif self.use_interp:
            call("mv_beam_interp", D.ptr(self.planes), int(self.is_f64),
                 self.nbeam, self.nfeed, self.nax, self.nza, self.naz,
                 D.ptr(self.grid), self.order, int(self.is_power), D.ptr(crd),
                 crd_is_f64, ld_crd, count, D.ptr(n_dev), D.ptr(out), ld_out,
                 D.stream_ptr(stream))
        else:
            call("mv_beam_gaussian", D.ptr(self.sigma), self.nbeam,
                 int(self.polarized), D.ptr(crd), crd_is_f64, ld_crd, count,
                 D.ptr(n_dev), D.ptr(out), int(self.is_f64), ld_out,
                 D.stream_ptr(stream))
        return out


def _crd_rows(tx, ty):
    """(pointer source tensor, leading dimension) covering rows tx, ty."""
    if (isinstance(tx, torch.Tensor) and isinstance(ty, torch.Tensor)
            and tx.is_cuda and ty.is_cuda and tx.dtype == ty.dtype
            and tx.dtype in (torch.float64, torch.float32) and tx.dim() == 1
            and tx.stride(0) == 1 and ty.stride(0) == 1):
        item = tx.element_size()
        gap = ty.data_ptr() - tx.data_ptr()
        if gap % item == 0 and gap // item >= tx.numel():
            return tx, gap // item
    tx = torch.as_tensor(tx, dtype=torch.float64).to(D.require_gpu())
    ty = torch.as_tensor(ty, dtype=torch.float64).to(D.require_gpu())
    both = torch.stack([tx, ty]).contiguous()
    return both, both.stride(0)


class GPUBeamInterpolator(BeamInterpolator):
    """Beam evaluation on the device for grid (UVBeam-like) or Gaussian
    beams; a beam list must be all one kind (gpu/beams.py:34-38)."""

    def setup(self):
        super().setup()
        self.use_interp = is_grid_beam(self.beam_list[0])
        if self.beam_idx is not None:
            self.beam_idx = D.as_dev(self.beam_idx, np.int32)
        self.tables = DeviceBeamSet(
            self.beam_list, self.freq, self.polarized, self.complex_dtype,
            spline_order(self.spline_opts),
        )

    def interp(self, tx, ty, out):
        src, ld = _crd_rows(tx, ty)
        count = int(tx.shape[0]) if hasattr(tx, "shape") else len(tx)
        ld_out = out.shape[-1]
        if count > ld_out:
            raise ValueError(f"{count} positions do not fit the beam buffer "
                             f"({ld_out})")
        self.tables.eval(src, ld, count, out, ld_out)
        if count < ld_out:
            out[..., count:] = 0
        return out


def gpu_beam_interpolation(beam, daz, dza, azmin, az, za, beam_at_src=None,
                           order: int = 1):
    """Interpolate (nbeam, nax, nfeed, nza, naz) grids at (az, za).

    The reference utility of gpu/beams.py:127-200: real grids are power beams
    (square root after interpolation), complex grids E-field beams; the
    output is ``(nbeam, nfeed, nax, nsrc)`` complex on the device.
    """
    b = beam.detach().cpu().numpy() if isinstance(beam, torch.Tensor) else \
        np.asarray(beam)
    if b.dtype in (np.dtype("float32"), np.dtype("complex64")):
        ctype = np.complex64
    elif b.dtype in (np.dtype("float64"), np.dtype("complex128")):
        ctype = np.complex128
    else:
        raise ValueError(
            f"Got {b.dtype} as the dtype for beam, which is unrecognized"
        )
    is_power = not np.iscomplexobj(b)
    nbeam, nax, nfeed, nza, naz = b.shape
    az_d = D.as_dev(az, np.float64).reshape(-1)
    za_d = D.as_dev(za, np.float64).reshape(-1)
    nsrc = az_d.numel()
    grid = np.stack([np.broadcast_to(np.asarray(daz, float), (nbeam,)),
                     np.broadcast_to(np.asarray(dza, float), (nbeam,)),
                     np.broadcast_to(np.asarray(azmin, float), (nbeam,))], 1)
    planes = D.to_dev(b.astype(ctype))
    is_f64 = int(ctype == np.complex128)
    if order == 3:
        call("mv_beam_prefilter", D.ptr(planes), is_f64, nbeam * nax * nfeed,
             nza, naz, D.stream_ptr())
    elif order != 1:
        raise ValueError(f"beam interpolation order must be 1 or 3, got {order}")
    if beam_at_src is None:
        beam_at_src = D.zeros((nbeam, nfeed, nax, nsrc), ctype)
    else:
        assert tuple(beam_at_src.shape) == (nbeam, nfeed, nax, nsrc)
    g = D.to_dev(grid, np.float64)
    call("mv_beam_interp_azza", D.ptr(planes), is_f64, nbeam, nfeed, nax, nza,
         naz, D.ptr(g), order, int(is_power), D.ptr(az_d), D.ptr(za_d), nsrc,
         D.ptr(beam_at_src), nsrc, D.stream_ptr())
    torch.cuda.current_stream().synchronize()
    return beam_at_src
