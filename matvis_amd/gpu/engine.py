"""Fused per-time visibility pipeline on one GPU.

For each time: one rotation + cull + compaction per source chunk (shared by
every frequency); then for each frequency and chunk the beam evaluation, the
fused phase/Z kernel (Z written once) and the HERK into an FP64 accumulator;
then the expansion into the API layout.  Everything is enqueued on the current
stream without host synchronization; results go to the host through pinned
buffers on a second stream, double-buffered so the copy of time t overlaps the
compute of time t+1.

This is the loop body of matvis's ``gpu.simulate`` (gpu/gpu.py:197-254) for
all frequencies of ``simulate_vis`` at once (wrapper.py:180-199): the rotation
depends only on time, so it is computed once per time instead of once per
(time, frequency).
"""

from __future__ import annotations

import logging
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import psutil
import torch

from .. import _device as D
from .._lib import FORM_C128, FORM_SPLIT_F16, call, lib
from .._utils import log_progress
from .._utils import logger as _progress_log
from .beams import DeviceBeamSet
from .herk import Herk, cdiv, expand, is_split, split_terms_alloc
from .matprod import f16_scale

Z_PAD = 32  # mv_zmat_fused pads Z columns to a multiple of this


def device_bytes(*, form: int, nterms: int, nprod: int, nant: int,
                 polarized: bool, nfreq: int, nsrc: int, nchunks: int,
                 chunk_size: int, nsrc_alloc: int, nbeam: int,
                 beam_pixels: int, beam_sets: int, npairs: int,
                 precision: int) -> dict:
    """Device bytes an engine and its rotation plugin hold, term by term, as
    their ``setup`` and ``run`` allocate them (beam_pixels: grid points of
    one beam plane, 0 for analytic beams; beam_sets: 1, or nfreq when the
    beam tables differ by frequency).  ``total`` is their sum."""
    nf = 2 if polarized else 1
    m = nf * nant
    seg = max(Z_PAD, cdiv(nsrc_alloc, Z_PAD) * Z_PAD)
    rsz = 8 if precision == 2 else 4
    esz = 16 if form == FORM_C128 else 8  # beam values, plain Z elements
    herk = Herk(form, m, nterms, nprod)
    cull_ws = int(lib().mv_rotate_cull_workspace_bytes(chunk_size))
    terms = {
        # rotation plugin: ECI, topocentric (and FP64 / BCRS copies), the
        # kept coordinates, fluxes, indices and its cull workspace
        "coords": nsrc * (24 + 3 * rsz + 24 + 24 + 16 + 8)
        + nsrc_alloc * (3 * rsz + 4 * 16 + 4) + cull_ws,
        "antpos_flux": nant * 24 + nfreq * nsrc * 8,
        "chunk_coords": nchunks * seg * (24 + 4) + cull_ws,
        "beam_tables": beam_sets * nbeam * nf * nf * beam_pixels * esz,
        "beam_values": (nchunks if beam_sets == 1 else 1) * nbeam * nf * nf
        * seg * esz,
        "z": m * nf * seg * (4 * nterms if is_split(form) else esz),
        "vacc": herk.ldv * herk.ldv * 16,
        "herk_workspace": int(lib().mv_herk_workspace_bytes(
            form, m, herk.nslices(nf * nsrc_alloc))),
        "output": 2 * nfreq * npairs * nf * nf * 2 * rsz,
    }
    terms["total"] = sum(terms.values())
    return terms


class VisibilityEngine:
    """Device-resident state of one simulation and its per-time pipeline.

    Parameters mirror ``gpu.simulate``: ``coords`` is a set-up
    :class:`~matvis_amd.core.coords.CoordinateRotation`; ``fluxes`` is
    (nsrc, nfreq) Stokes I; ``form`` / ``nterms`` / ``nprod`` select the HERK.
    """

    def __init__(self, *, antpos, freqs, coords, fluxes, beam_list, beam_idx,
                 polarized: bool, precision: int, form: int, nterms: int = 1,
                 nprod: int = 1, order: int = 1, nchunks: int, chunk_size: int,
                 nsrc_alloc: int, antpairs=None, beam_cache: dict | None = None,
                 beam_src=None):
        self.antpos_h = np.asarray(antpos, dtype=np.float64)
        self.nant = self.antpos_h.shape[0]
        self.freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        self.nfreq = self.freqs.size
        self.coords = coords
        self.nsrc = coords.nsrc
        fl = np.asarray(fluxes, dtype=np.float64).reshape(self.nsrc, -1)
        if fl.shape[1] != self.nfreq:
            raise ValueError("fluxes must have shape (NSRCS, NFREQS)")
        self.sqrt_flux_h = np.sqrt(0.5 * fl.T)  # (nfreq, nsrc)
        self.beam_list = beam_list
        self.beam_idx = beam_idx
        self.polarized = polarized
        self.nfeed = self.nax = 2 if polarized else 1
        self.m = self.nfeed * self.nant
        self.precision = precision
        self.ctype = np.complex128 if precision == 2 else np.complex64
        self.form = form
        self.nterms, self.nprod = nterms, nprod
        self.order = order
        self.nchunks, self.chunk_size = nchunks, chunk_size
        self.nsrc_alloc = nsrc_alloc
        self.seg = max(Z_PAD, cdiv(nsrc_alloc, Z_PAD) * Z_PAD)
        if antpairs is None:
            self.pairs_h = None
            self.npairs = self.nant * self.nant
        else:
            self.pairs_h = np.asarray(antpairs, dtype=np.int32).reshape(-1, 2)
            if self.pairs_h.size and (self.pairs_h.min() < 0
                                      or self.pairs_h.max() >= self.nant):
                raise ValueError("antpairs contains an antenna index out of range")
            self.npairs = len(self.pairs_h)
        #: device beam tables shared between engines (a resident server's
        #: cache), keyed by the caller's beam objects ``beam_src``
        self.beam_cache = beam_cache
        self.beam_src = list(beam_src) if beam_src is not None else None
        self._ready = False
        #: optional (before, after) callables taking the stream, run around
        #: every HERK compute kernel (event timing of the dominant kernel;
        #: the FP64 reduce that follows it is outside the pair)
        self.herk_hooks = None

    # ---------------------------------------------------------------- setup

    def setup(self):
        seg, nalloc = self.seg, self.nsrc_alloc
        self.antpos = D.to_dev(self.antpos_h, np.float64)
        self.sqrt_flux = D.to_dev(self.sqrt_flux_h, np.float64)
        self.beam_idx_dev = (None if self.beam_idx is None
                             else D.as_dev(self.beam_idx, np.int32))
        self.pairs_dev = (None if self.pairs_h is None
                          else D.to_dev(self.pairs_h, np.int32))
        nb = self.nchunks
        self.crd = [D.zeros((3, seg), np.float64) for _ in range(nb)]
        self.idx = [D.zeros((seg,), np.int32) for _ in range(nb)]
        self.n_up = D.zeros((2, nb), np.int32)
        ws = lib().mv_rotate_cull_workspace_bytes(self.chunk_size)
        self.rot_ws = D.empty((ws,), np.uint8)
        # beams: one table set per distinct frequency response
        self.beam_ctype = np.complex128 if self.form == FORM_C128 else np.complex64
        self.beam_sets = self._make_beam_sets()
        self.nbeam = self.beam_sets[0].nbeam
        shared = len(self.beam_sets) == 1
        nA = nb if shared else 1
        self.A = [D.zeros((self.nbeam, self.nfeed, self.nax, seg),
                          self.beam_ctype) for _ in range(nA)]
        # Z: (m, nax, seg) complex (ldz the row stride), or nterms arrays of
        # 16-bit terms, K-tile-major (ldz the rows of a 16-column K tile;
        # 32 bits a complex element, 16-byte chunks of 4 columns: re x 4,
        # im x 4)
        if is_split(self.form):
            self.ldz = self.m
            self.z = split_terms_alloc(self.nterms, self.m, self.nax * seg)
        else:
            self.ldz = self.nax * seg
            self.z = [D.zeros((self.m, self.ldz), self.beam_ctype)]
        self.herk = Herk(self.form, self.m, self.nterms, self.nprod)
        self.vacc = self.herk.new_vacc()
        self.nslices = self.herk.nslices(self.nax * nalloc)
        self.herk.workspace(self.nslices)
        # per-(chunk, freq) power-of-two scales for the fp16 split
        self.scales = np.ones((nb, self.nfreq))
        if self.form == FORM_SPLIT_F16:
            bmax = max(b.max_abs for b in self.beam_sets) * 1.5
            for c in range(nb):
                sl = slice(c * self.chunk_size, (c + 1) * self.chunk_size)
                part = self.sqrt_flux_h[:, sl]
                for f in range(self.nfreq):
                    mx = float(part[f].max()) if part.shape[1] else 0.0
                    self.scales[c, f] = f16_scale(mx * bmax)
        self._ready = True

    def _make_beam_sets(self):
        keys = [self._beam_key(f) for f in self.freqs]
        if all(k == keys[0] for k in keys):
            return [self._beam_set(self.freqs[0])]
        return [self._beam_set(f) for f in self.freqs]

    def _beam_set(self, f):
        """Tables at frequency f, from the shared cache when one is given
        (keyed by the caller's beam objects, which the cache keeps alive so
        that their ids stay unique, and by everything the tables depend on)."""
        if self.beam_cache is None or self.beam_src is None:
            return DeviceBeamSet(self.beam_list, f, self.polarized,
                                 self.beam_ctype, self.order)
        fkey = tuple(k[2:] if k[0] == "grid" else k for k in self._beam_key(f))
        key = (tuple(id(b) for b in self.beam_src), self.polarized,
               np.dtype(self.beam_ctype).str, self.order, fkey)
        hit = self.beam_cache.get(key)
        if hit is None:
            hit = (DeviceBeamSet(self.beam_list, f, self.polarized,
                                 self.beam_ctype, self.order), self.beam_src)
            self.beam_cache[key] = hit
        return hit[0]

    def _beam_key(self, f):
        """Frequency-response key: equal keys give identical beam tables."""
        from ..beams import is_grid_beam, unwrap

        key = []
        for b in self.beam_list:
            inner, _ = unwrap(b)
            if is_grid_beam(b):
                key.append(("grid", id(inner)) if np.asarray(
                    inner.freq_array).size <= 1 else ("grid", id(inner), f))
            else:
                key.append(("gauss", inner.get_sigma(f)))
        return tuple(key)

    # ---------------------------------------------------------------- compute

    def stage_time(self, t: int, parity: int = 0, stream=None):
        """Rotation, cull and compaction of time t for every chunk (and the
        beam at the kept sources when one beam table serves all
        frequencies); the chunk counts land in n_up[parity]."""
        if not self._ready:
            self.setup()
        co = self.coords
        co.rotate(t)
        src = co.rotation_source()
        rot = co._rot
        sp = D.stream_ptr(stream)
        nup = self.n_up[parity]
        shared = len(self.beam_sets) == 1
        for c in range(self.nchunks):
            chunk0 = min(c * self.chunk_size, self.nsrc)
            clen = max(0, min(self.chunk_size, self.nsrc - chunk0))
            call("mv_rotate_cull", D.ptr(src), self.nsrc, rot.ctypes.data,
                 None, chunk0, clen, self.nsrc_alloc, D.ptr(self.crd[c]),
                 self.seg, None, 1, D.ptr(self.idx[c]), D.ptr(nup[c:c + 1]),
                 D.ptr(self.rot_ws), self.rot_ws.numel(), sp)
            if shared:
                self.beam_sets[0].eval(self.crd[c], self.seg, self.nsrc_alloc,
                                       self.A[c], self.seg, nup[c:c + 1],
                                       stream)
        self._parity = parity

    def stage_z(self, fi: int, c: int, stream=None):
        """Z of frequency index fi, chunk c (after stage_time) into the Z
        buffers; returns the HERK problem it feeds (as ``HerkGroup`` and
        ``Herk.run`` take it)."""
        n_dev = self.n_up[self._parity][c:c + 1]
        if len(self.beam_sets) == 1:
            a = self.A[c]
        else:
            a = self.A[0]
            self.beam_sets[fi].eval(self.crd[c], self.seg, self.nsrc_alloc, a,
                                    self.seg, n_dev, stream)
        s = float(self.scales[c, fi])
        zp = [D.ptr(z) for z in self.z] + [None] * (3 - len(self.z))
        call("mv_zmat_fused", D.ptr(self.antpos), self.nant,
             float(self.freqs[fi]), D.ptr(self.crd[c]), self.seg,
             D.ptr(self.sqrt_flux[fi]), D.ptr(self.idx[c]), D.ptr(a),
             int(self.beam_ctype == np.complex128), self.seg, self.nbeam,
             D.ptr(self.beam_idx_dev), self.nfeed, self.nax,
             self.nsrc_alloc, D.ptr(n_dev), self.form, self.nterms, s,
             zp[0], zp[1], zp[2], self.ldz, self.seg, D.stream_ptr(stream))
        return dict(terms=self.z, m=self.m, ldz=self.ldz, nseg=self.nax,
                    seg_stride=self.seg, seg_len=self.nsrc_alloc,
                    seg_len_dev=n_dev, vacc=self.vacc, ldv=self.herk.ldv,
                    scale=1.0 / (s * s), accumulate=c > 0)

    def run_herk(self, prob: dict, stream=None):
        """The HERK of one staged (frequency, chunk) on its own."""
        self.herk.run(prob["terms"], ldz=prob["ldz"], nseg=prob["nseg"],
                      seg_stride=prob["seg_stride"], seg_len=prob["seg_len"],
                      seg_len_dev=prob["seg_len_dev"], vacc=prob["vacc"],
                      accumulate=prob["accumulate"], scale=prob["scale"],
                      nslices=self.nslices, stream=stream,
                      between=(self.herk_hooks[1] if self.herk_hooks
                               else None))

    def expand_freq(self, fi: int, out_dev, stream=None):
        """The accumulated V of frequency fi into out_dev[fi] (API layout)."""
        expand(self.vacc, self.herk.ldv, self.herk.tile, self.nant,
               self.nfeed, self.pairs_dev, self.npairs, out_dev[fi], stream)

    def compute_time(self, t: int, out_dev, parity: int = 0, stream=None):
        """Enqueue the whole of time index t; results land in out_dev
        (nfreq, npairs, nfeed, nfeed) and the chunk counts in n_up[parity]."""
        self.stage_time(t, parity, stream)
        for fi in range(self.nfreq):
            for c in range(self.nchunks):
                prob = self.stage_z(fi, c, stream)
                if self.herk_hooks:
                    self.herk_hooks[0](stream)
                self.run_herk(prob, stream)
            self.expand_freq(fi, out_dev, stream)
        return out_dev

    def new_output(self):
        return D.empty((self.nfreq, self.npairs, self.nfeed, self.nfeed),
                       self.ctype)

    def check_counts(self, counts: np.ndarray):
        worst = int(np.max(counts)) if counts.size else 0
        if worst > self.nsrc_alloc:
            raise ValueError(
                f"nsrc_alloc ({self.nsrc_alloc}) is too small for the number "
                f"of sources above horizon ({worst}). Try increasing "
                "source_buffer."
            )

    def run(self, ntimes: int, out: np.ndarray | None = None) -> np.ndarray:
        """All times; returns host (nfreq, ntimes, npairs, nfeed, nfeed)."""
        if not self._ready:
            self.setup()
        if out is None:
            out = np.zeros((self.nfreq, ntimes, self.npairs, self.nfeed,
                            self.nfeed), dtype=self.ctype)
        comp = torch.cuda.current_stream()
        copy = torch.cuda.Stream()
        dev = [self.new_output(), self.new_output()]
        pin = [torch.empty(tuple(dev[0].shape), dtype=dev[0].dtype,
                           pin_memory=True) for _ in range(2)]
        nup_pin = torch.empty((2, self.nchunks), dtype=torch.int32,
                              pin_memory=True)
        done_comp = [torch.cuda.Event(), torch.cuda.Event()]
        done_copy = [torch.cuda.Event(), torch.cuda.Event()]
        pending = None

        # the pinned -> caller copy of a time (nfreq x npairs x nfeed^2
        # elements, 157 MB at 350 antennas x 20 frequencies) in frequency
        # blocks on several host threads: one thread copies ~6 GB/s, slower
        # than the GPU produces a time
        nthr = max(1, min(self.nfreq, 8, os.cpu_count() or 1))
        blocks = np.array_split(np.arange(self.nfreq), nthr)
        pool = ThreadPoolExecutor(nthr) if nthr > 1 else None

        def copy_block(dst, src, b):
            if b.size:
                dst[b[0]:b[-1] + 1] = src[b[0]:b[-1] + 1]

        def finish(item):
            tt, kk = item
            done_copy[kk].synchronize()
            self.check_counts(nup_pin[kk].numpy())
            src, dst = pin[kk].numpy(), out[:, tt]
            if pool is None:
                dst[...] = src
            else:
                list(pool.map(lambda b: copy_block(dst, src, b), blocks))

        # progress at INFO about every 1% of the times, as the reference's
        # time loop reports (gpu/gpu.py)
        report = ntimes // 100 + 1
        proc = (psutil.Process() if _progress_log.isEnabledFor(logging.INFO)
                else None)
        t_start = t_last = time.time()
        mem_last = proc.memory_info().rss if proc is not None else 0
        try:
            for t in range(ntimes):
                k = t % 2
                if t >= 2:
                    comp.wait_event(done_copy[k])
                self.compute_time(t, dev[k], parity=k, stream=comp)
                done_comp[k].record(comp)
                copy.wait_event(done_comp[k])
                with torch.cuda.stream(copy):
                    pin[k].copy_(dev[k], non_blocking=True)
                    nup_pin[k].copy_(self.n_up[k], non_blocking=True)
                done_copy[k].record(copy)
                if pending is not None:
                    finish(pending)
                pending = (t, k)
                if proc is not None and (t % report == 0 or t == ntimes - 1):
                    t_last, mem_last = log_progress(t_start, t_last, t + 1,
                                                    ntimes, proc, mem_last)
            if pending is not None:
                finish(pending)
        finally:
            if pool is not None:
                pool.shutdown()
        comp.wait_stream(copy)
        return out
