"""Grouped HERK: many instances of one form and different sizes in one
persistent launch (``mv_herk_grouped``), optionally captured once in a HIP
graph and replayed.  Split forms use 64- or 128-row tiles, the plain forms
(complex128, complex64) 64-row tiles.

A :class:`HerkGroup` holds one tile class (64 or 128 rows per workgroup
tile) of a batch: the packed problem descriptors on the host and the device,
and the workspace.  :class:`GroupedBatch` splits a batch into tile classes
(the tile per problem comes from :mod:`matvis_amd.gpu.autotune`) and runs
them back to back on one stream.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _device as D
from .._lib import call, lib
from .herk import Herk, is_split


def plain_tile() -> int:
    """Tile side of the plain forms' grouped launch."""
    from .._lib import FORM_C128

    return int(lib().mv_herk_tile(FORM_C128))


def tile_for(form: int, m: int) -> int:
    """Default tile of one problem: 64 rows for small split problems, the
    form's own tile otherwise."""
    if not is_split(form):
        return plain_tile()
    return 64 if m <= 128 else 128


#: Layout of ``mv_herk_problem`` (include/matvis_hip.h), 104 bytes.
PROBLEM_DTYPE = np.dtype([
    ("z", "<u8", (3,)), ("ldz", "<i8"), ("seg_stride", "<i8"),
    ("seg_len", "<i8"), ("seg_len_dev", "<u8"), ("vacc", "<u8"),
    ("ldv", "<i8"), ("scale", "<f8"), ("m", "<i4"), ("nseg", "<i4"),
    ("accumulate", "<i4"), ("nslices", "<i4"), ("kchain", "<i4"),
    ("reserved", "<i4"),
])

#: K-tile width in complex columns of the grouped kernels, split and plain
#: alike (kSK and kPK in csrc/herk.hip)
K_TILE = 16


def slice_capacity(form: int, m: int, nterms: int, nprod: int, kmax: int,
                   kchain: int = 0) -> int:
    """Slices a problem of up to kmax complex columns may run: the kernels'
    own capacity (mv_herk_max_slices), raised to cover ceil(K tiles /
    kchain) when a shorter chain is asked for."""
    cap = Herk(form, m, nterms, nprod).nslices(kmax)
    if kchain > 0:
        ktiles = -(-int(kmax) // K_TILE)
        cap = max(cap, -(-ktiles // int(kchain)))
    return max(1, cap)


def _addr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


class HerkGroup:
    """One grouped launch over problems sharing a form and a tile.

    ``problems``: list of dicts with ``terms`` (list of int32 term tensors
    for split forms, one complex tensor for the plain forms),
    ``m``, ``ldz``, ``nseg``, ``seg_stride``, ``seg_len`` (host bound),
    optional ``seg_len_dev`` (device int32 tensor), ``vacc`` (complex128
    tensor, at least ldv x ldv), ``ldv``, ``scale``, ``accumulate``.
    The tensors must stay alive and in place while the group is used.
    """

    def __init__(self, form: int, nterms: int, nprod: int, tile: int,
                 problems: list):
        if lib().mv_herk_problem_bytes() != PROBLEM_DTYPE.itemsize:
            raise RuntimeError("mv_herk_problem layout mismatch")
        self.form, self.nterms, self.nprod, self.tile = form, nterms, nprod, tile
        self.problems = problems
        host = self._pack(problems)
        self.host = host
        self.dev = D.to_dev(host.view(np.uint8))
        need = lib().mv_herk_grouped_workspace_bytes(
            form, tile, host.ctypes.data, len(problems))
        if need == 0:
            raise ValueError("invalid grouped HERK problems")
        self.ws = D.empty((need,), np.uint8)

    def _pack(self, problems):
        host = np.zeros(len(problems), PROBLEM_DTYPE)
        for i, p in enumerate(problems):
            terms = list(p["terms"]) + [None] * (3 - len(p["terms"]))
            host["z"][i] = [_addr(t) for t in terms]
            host["ldz"][i] = p["ldz"]
            host["seg_stride"][i] = p["seg_stride"]
            host["seg_len"][i] = p["seg_len"]
            host["seg_len_dev"][i] = _addr(p.get("seg_len_dev"))
            host["vacc"][i] = _addr(p["vacc"])
            host["ldv"][i] = p["ldv"]
            host["scale"][i] = p.get("scale", 1.0)
            host["m"][i] = p["m"]
            host["nseg"][i] = p["nseg"]
            host["accumulate"][i] = int(bool(p.get("accumulate", False)))
            kchain = int(p.get("kchain", 0))
            host["kchain"][i] = kchain
            host["nslices"][i] = slice_capacity(
                self.form, p["m"], self.nterms, self.nprod,
                p["nseg"] * p["seg_len"], kchain)
        return host

    def update(self, problems):
        """New contents for the same problem shapes (addresses, scales,
        accumulate flags); the device copy is rewritten in place, so a
        captured graph replays with it.  The copy is asynchronous on the
        current stream, from pinned memory that torch's host allocator keeps
        until the copy has run: the host does not wait for the work queued
        before it."""
        host = self._pack(problems)
        for f in ("m", "nseg", "ldz", "seg_stride", "seg_len", "ldv",
                  "nslices", "kchain"):
            if not np.array_equal(host[f], self.host[f]):
                raise ValueError(f"grouped problems changed shape ({f})")
        self.problems, self.host = problems, host
        staging = torch.from_numpy(host.view(np.uint8).copy()).pin_memory()
        self.dev.copy_(staging, non_blocking=True)

    def run(self, stream=None):
        call("mv_herk_grouped", self.form, self.nterms, self.nprod, self.tile,
             self.host.ctypes.data, D.ptr(self.dev), len(self.problems),
             D.ptr(self.ws), self.ws.numel(), D.stream_ptr(stream))


class GroupedBatch:
    """A batch of HERK problems of one form, one grouped launch per tile class;
    ``capture()`` records the launches in a HIP graph for ``replay()``."""

    def __init__(self, form: int, nterms: int, nprod: int, problems: list,
                 tiles: list[int], kchains: list[int] | None = None):
        self.tiles = [int(t) for t in tiles]
        self.kchains = ([int(k) for k in kchains] if kchains is not None
                        else None)
        classes: dict[int, list] = {}
        for p, t in zip(self._with_chains(problems), self.tiles):
            classes.setdefault(t, []).append(p)
        self.groups = [HerkGroup(form, nterms, nprod, t, ps)
                       for t, ps in sorted(classes.items())]
        self.graph = None

    def _with_chains(self, problems):
        if self.kchains is None:
            return problems
        return [dict(p, kchain=k) for p, k in zip(problems, self.kchains)]

    def update(self, problems):
        """Same problems in the same order with new contents (see
        :meth:`HerkGroup.update`)."""
        classes: dict[int, list] = {}
        for p, t in zip(self._with_chains(problems), self.tiles):
            classes.setdefault(t, []).append(p)
        for g in self.groups:
            g.update(classes[g.tile])

    def run(self, stream=None):
        for g in self.groups:
            g.run(stream)

    def capture(self):
        """Record run() in a graph.  The launches first run once eagerly, so
        the kernels' occupancy queries are cached outside the capture: that
        run is a real execution of the batch (problems with ``accumulate``
        add into their accumulators), and a caller that captures in place of
        a step must not also replay in that step."""
        self.run()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(self.graph, stream=s):
                self.run(torch.cuda.current_stream())
        torch.cuda.current_stream().wait_stream(s)
        return self

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()
