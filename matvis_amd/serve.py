"""A resident visibility server on one GPU: continuous batching of
simulation requests of uneven sizes.

Requests (antenna layout, sky, frequencies, times, beam, polarization,
precision form) are admitted while their projected device footprint fits the
HBM budget; admitted requests stay resident (their engine keeps antenna
positions, sky, beam tables and buffers on the device) until their last
(time, frequency) instance is done.  Each scheduler step takes the next
instance of every admitted request: the per-time rotation and cull, the beam
and the fused Z kernel run per request; the HERKs of all instances of the
step then run as grouped persistent launches, one per HERK form present
(split-BF16, split-FP16, complex128, complex64), each captured in a HIP graph
the first time a set of requests steps together and replayed while that set
lasts.  Requests that arrive while others run are admitted at the next step,
as memory frees up.

Every instance's HERK depends only on its own data (the slice count comes
from its own column count), so a request's result is the same alone and in
any mix.

Results of each time go to the host through pinned buffers, double-buffered
per request, so the device runs ahead of the host by up to two times of every
resident request; device beam tables are shared between requests that use the
same beam objects.

Latency is submit -> result on the host; :meth:`VisibilityServer.stats`
reports p50 / p99 and visibilities/s, overall and per request label.
"""

from __future__ import annotations

import itertools
import threading
import time
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
import torch

from . import _device as D
from ._lib import FORM_C64, FORM_C128
from .gpu.herk import is_split

_ids = itertools.count()


@dataclass
class SimRequest:
    """One simulation; keywords as ``matvis_amd.gpu.gpu.simulate_multi``
    (``freqs``, ``fluxes`` of shape (nsrc, nfreq), ``times``, ``skycoords``,
    ``telescope_loc``, ``beam_list``, ``polarized``, ``precision``,
    ``matprod_method``, ``coord_method`` ...)."""

    kwargs: dict
    #: keep the whole (nfreq, ntimes, npairs, nfeed, nfeed) result in
    #: ``result``; with False only ``on_time`` sees the data
    keep: bool = True
    #: called as on_time(request, t, values) with the (nfreq, npairs, nfeed,
    #: nfeed) result of time t (a buffer reused afterwards: copy to keep)
    on_time: Callable | None = None
    #: free-form class name for per-class statistics
    label: str = ""
    rid: int = field(default_factory=lambda: next(_ids))
    t_submit: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    visibilities: int = 0
    #: projected device work (set when the shortest-first policy needs it)
    work: float | None = None
    result: np.ndarray | None = None
    error: Exception | None = None
    done: threading.Event = field(default_factory=threading.Event)


def projected_work(engine_kwargs: dict) -> float:
    """Relative device work of a request: instances x m^2 x sources (the
    HERK term that dominates), FP64 forms weighted 5x."""
    nant = len(engine_kwargs["antpos"])
    nf = 2 if engine_kwargs.get("polarized") else 1
    nsrc = np.asarray(engine_kwargs["fluxes"]).shape[0]
    nfreq = np.atleast_1d(engine_kwargs["freqs"]).size
    times = engine_kwargs["times"]
    ntimes = len(np.atleast_1d(getattr(times, "jd", times)))
    split = "Split" in engine_kwargs.get("matprod_method", "GPUMatMul")
    fp64 = not split and engine_kwargs.get("precision", 1) == 2
    return float(nfreq * ntimes * (nf * nant) ** 2 * nf * nsrc
                 * (5.0 if fp64 else 1.0))


def projected_bytes(engine_kwargs: dict) -> int:
    """Device bytes a request will hold while resident: Z, accumulator,
    HERK workspace, coordinates, beam values and the output of one time."""
    from .gpu.herk import Herk

    nant = len(engine_kwargs["antpos"])
    nf = 2 if engine_kwargs.get("polarized") else 1
    m = nant * nf
    nsrc = np.asarray(engine_kwargs["fluxes"]).shape[0]
    nfreq = np.atleast_1d(engine_kwargs["freqs"]).size
    seg = -(-nsrc // 32) * 32
    method = engine_kwargs.get("matprod_method", "GPUMatMul")
    split = "Split" in method
    zbytes = m * nf * seg * (8 if split else
                             (16 if engine_kwargs.get("precision", 1) == 2
                              else 8))
    ldv = -(-m // 128) * 128
    work = ldv * ldv * 16 * 64  # accumulator and a bound on the partials
    out = nfreq * nant * nant * nf * nf * 16
    return int(zbytes + work + 3 * seg * 12 + 4 * nf * nf * seg * 16 + out
               + nsrc * nfreq * 8)


class _Active:
    """An admitted request, its position in (time, frequency, chunk) and its
    double-buffered device -> host path."""

    def __init__(self, req: SimRequest, engine, ntimes: int, nbytes: int):
        self.req, self.engine, self.ntimes, self.nbytes = req, engine, ntimes, nbytes
        self.t = self.fi = self.c = 0
        e = engine
        self.out_dev = [e.new_output(), e.new_output()]
        shape = tuple(self.out_dev[0].shape)
        self.pin = [torch.empty(shape, dtype=self.out_dev[0].dtype,
                                pin_memory=True) for _ in range(2)]
        self.nup_pin = torch.empty((2, e.nchunks), dtype=torch.int32,
                                   pin_memory=True)
        self.ready = [torch.cuda.Event(), torch.cuda.Event()]
        self.pending: list[tuple[int, int]] = []  # (time, buffer)
        self.host = (np.zeros((e.nfreq, ntimes, e.npairs, e.nfeed, e.nfeed),
                              dtype=e.ctype) if req.keep else None)

    @property
    def split(self) -> bool:
        return is_split(self.engine.form)

    def advance(self) -> bool:
        """Move past the current instance; True when the request is done."""
        e = self.engine
        self.c += 1
        if self.c < e.nchunks:
            return False
        self.c = 0
        self.fi += 1
        if self.fi < e.nfreq:
            return False
        self.fi = 0
        self.t += 1
        return self.t >= self.ntimes

    def ship(self, t: int, copy_stream):
        """Enqueue the copy of time t's result to the host on
        `copy_stream`, after the compute stream's work so far."""
        k = t % 2
        if len(self.pending) == 2:
            self.deliver()
        e = self.engine
        computed = torch.cuda.Event()
        computed.record()
        copy_stream.wait_event(computed)
        with torch.cuda.stream(copy_stream):
            self.pin[k].copy_(self.out_dev[k], non_blocking=True)
            self.nup_pin[k].copy_(e.n_up[k], non_blocking=True)
            self.ready[k].record(copy_stream)
        self.pending.append((t, k))

    def stage_time(self):
        """Rotation and cull of the current time, once the copy out of the
        buffers it reuses (those of time t - 2) has finished."""
        if self.t >= 2:
            torch.cuda.current_stream().wait_event(self.ready[self.t % 2])
        self.engine.stage_time(self.t, parity=self.t % 2)

    def deliver(self):
        """Hand the oldest shipped time to the request (waits for its copy)."""
        t, k = self.pending.pop(0)
        self.ready[k].synchronize()
        self.engine.check_counts(self.nup_pin[k].numpy())
        vals = self.pin[k].numpy()
        if self.host is not None:
            self.host[:, t] = vals
        if self.req.on_time is not None:
            self.req.on_time(self.req, t, vals)


class VisibilityServer:
    """See the module docstring.  ``hbm_budget``: bytes of device memory the
    resident requests may hold (default: 80% of free memory); ``max_active``:
    at most this many requests resident at once; ``graphs``: capture the
    grouped HERK launches; ``policy``: admission order, ``"fifo"`` (arrival)
    or ``"shortest"`` (least projected work first, aged: a request's work
    counts as work / (1 + wait / ``age_s``), so long waits are admitted)."""

    def __init__(self, hbm_budget: int | None = None, max_active: int = 8,
                 graphs: bool = True, policy: str = "fifo",
                 age_s: float = 10.0):
        D.require_gpu()
        if policy not in ("fifo", "shortest"):
            raise ValueError("policy must be 'fifo' or 'shortest'")
        self.budget = int(hbm_budget or 0.8 * D.free_memory())
        self.max_active = max_active
        self.graphs = graphs
        self.policy = policy
        self.age_s = float(age_s)
        self.queue: list[SimRequest] = []
        self.active: list[_Active] = []
        self.finished: list[SimRequest] = []
        self.used = 0
        self.vis_done = 0
        self._lock = threading.Lock()
        self._batches: dict[tuple, object] = {}
        self._beams: dict = {}
        self._copy = torch.cuda.Stream()
        self._thread = None
        self._stop = threading.Event()

    # ---------------------------------------------------------------- API

    def submit(self, req: SimRequest) -> SimRequest:
        req.t_submit = time.perf_counter()
        with self._lock:
            self.queue.append(req)
        return req

    def step(self) -> int:
        """Admit what fits, then run one instance of every active request.
        Returns the number of instances run."""
        self._admit()
        if not self.active:
            return 0
        split_probs, plain, failed = [], [], []
        for a in self.active:
            try:
                e = a.engine
                if a.fi == 0 and a.c == 0:
                    a.stage_time()
                prob = e.stage_z(a.fi, a.c)
                (split_probs if a.split else plain).append((a, prob))
            except Exception as exc:  # reported on the request
                failed.append((a, exc))
        if split_probs or plain:
            self._grouped(split_probs + plain)
        done = []
        for a in self.active:
            if any(a is f for f, _ in failed):
                continue
            e = a.engine
            try:
                if a.c == e.nchunks - 1:
                    e.expand_freq(a.fi, a.out_dev[a.t % 2])
                    if a.fi == e.nfreq - 1:
                        a.ship(a.t, self._copy)
                        n = e.nfreq * e.npairs * e.nfeed * e.nfeed
                        a.req.visibilities += n
                        self.vis_done += n
                if a.advance():
                    while a.pending:
                        a.deliver()
                    done.append(a)
            except Exception as exc:
                failed.append((a, exc))
        for a, exc in failed:
            a.req.error = exc
            done.append(a)
        for a in done:
            self._finish(a)
        return len(split_probs) + len(plain)

    def run_until_idle(self, timeout: float = 600.0):
        t0 = time.perf_counter()
        while (self.queue or self.active) and time.perf_counter() - t0 < timeout:
            self.step()

    def start(self):
        """Serve from a background thread until :meth:`stop`."""
        def loop():
            while not self._stop.is_set():
                if not self.step():
                    time.sleep(0.001)
        self._stop.clear()
        self._thread = threading.Thread(target=loop, daemon=True)
        self._thread.start()

    def stop(self):
        self._stop.set()
        if self._thread is not None:
            self._thread.join()

    def stats(self) -> dict:
        """p50 / p99 latency (s) and visibilities of the finished requests,
        overall and per label."""
        def summary(reqs):
            lat = np.array([r.t_done - r.t_submit for r in reqs
                            if r.error is None])
            return {
                "requests": len(reqs),
                "failed": sum(r.error is not None for r in reqs),
                "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
                "visibilities": int(sum(r.visibilities for r in reqs)),
            }

        out = summary(self.finished)
        out["visibilities_all"] = int(self.vis_done)
        labels = sorted({r.label for r in self.finished if r.label})
        if labels:
            out["by_label"] = {lb: summary([r for r in self.finished
                                            if r.label == lb])
                               for lb in labels}
        return out

    # ---------------------------------------------------------------- internals

    def _admit(self):
        from .gpu.gpu import build_engine

        while len(self.active) < self.max_active:
            with self._lock:
                if not self.queue:
                    return
                i = self._next_index()
                need = projected_bytes(self.queue[i].kwargs)
                if self.active and self.used + need > self.budget:
                    return
                req = self.queue.pop(i)
            try:
                engine, ntimes = build_engine(**req.kwargs,
                                              beam_cache=self._beams)
            except Exception as exc:  # reported on the request
                req.error = exc
                req.t_done = time.perf_counter()
                with self._lock:
                    self.finished.append(req)
                req.done.set()
                continue
            req.t_admit = time.perf_counter()
            self.used += need
            self.active.append(_Active(req, engine, ntimes, need))

    def _next_index(self) -> int:
        """Queue position of the request to admit next (lock held)."""
        if self.policy == "fifo" or len(self.queue) == 1:
            return 0
        now = time.perf_counter()
        for r in self.queue:
            if r.work is None:
                r.work = projected_work(r.kwargs)
        return min(range(len(self.queue)), key=lambda i: self.queue[i].work
                   / (1.0 + (now - self.queue[i].t_submit) / self.age_s))

    def _grouped(self, items):
        from .gpu.herk_group import GroupedBatch, tile_for

        # one grouped launch per form variant present in the step
        variants: dict[tuple, list] = {}
        for a, prob in items:
            e = a.engine
            variants.setdefault((e.form, e.nterms, e.nprod), []).append((a, prob))
        for (form, nterms, nprod), its in variants.items():
            key = (form, nterms, nprod) + tuple(a.req.rid for a, _ in its)
            probs = [p for _, p in its]
            batch = self._batches.get(key)
            if batch is None:
                tiles = [tile_for(form, p["m"]) for p in probs]
                batch = GroupedBatch(form, nterms, nprod, probs, tiles)
                self._batches[key] = batch
                if self.graphs:
                    # capture() executes the launches once before recording
                    # them: that execution is this step's, so no replay here
                    # (a replay would add accumulating chunks twice)
                    batch.capture()
                    continue
            else:
                batch.update(probs)
            if self.graphs:
                batch.replay()
            else:
                batch.run()

    def _finish(self, a: _Active):
        req = a.req
        req.result = a.host if req.error is None else None
        req.t_done = time.perf_counter()
        with self._lock:
            self.active.remove(a)
            self.used -= a.nbytes
            self.finished.append(req)
            # graphs of sets that included this request cannot recur
            for k in [k for k in self._batches if req.rid in k[3:]]:
                del self._batches[k]
        req.done.set()
