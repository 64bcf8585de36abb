"""Launch-shape autotuner for grouped HERKs, by measurement.

The knobs, chosen per class of problems (m rounded up to a power of two,
K = nseg x seg_len rounded up to a power of four):

* tile side (split forms): 64 rows (one wave per workgroup tile, up to
  eight workgroups per CU) or 128 rows (four waves of 64 x 64, two
  workgroups per CU).  Waves per tile go with the side.  Small problems
  waste most of a 128-row tile (m = 64 keeps one of its four waves busy);
  large ones lose operand reuse on 64-row tiles.  A 256-row tile would need
  256 x 256 complex FP32 accumulators (512 KB, a CU's whole register file),
  so it is not a candidate.  The plain forms have one tile (64 rows).
* kchain, the longest K-slice in K tiles (split forms 192, 96 or 48; plain
  forms 64, 32 or 16): shorter slices give a batch that does not fill the
  GPU more work items, at the cost of more partial tiles to write and
  reduce.  The split forms' 192 is also their accuracy cap on an FP32
  chain, so the tuner only shortens chains.

Search: coordinate descent from the default (64-row tiles for m <= 128,
128 otherwise; the form's chain cap).  Classes are visited in order of
descending work.  Each candidate changes one knob of one class and is timed
as the whole batch's grouped launches (HIP events, the mean of `reps` runs
after one warm run).  A change is kept when it is at least 1% faster.  The
search ends after a pass without a kept change or when the time budget is
spent.  Results are cached per batch signature (form, terms, products and
the multiset of classes) in memory and, if ``MATVIS_AMD_TUNE_CACHE`` names a
file, in that JSON file.  The problems' accumulators are overwritten while
tuning.
"""

from __future__ import annotations

import json
import math
import os
import threading
import time
from dataclasses import dataclass, field

import torch

from .herk import is_split
from .herk_group import GroupedBatch, tile_for

#: K-slice lengths (K tiles) the search tries, longest (the default) first
CHAINS = {True: (192, 96, 48), False: (64, 32, 16)}
#: tile sides of the grouped kernels
TILES = {True: (64, 128), False: (64,)}
#: a candidate must beat the incumbent by this fraction to be kept
GAIN = 0.01

_cache: dict[str, dict] = {}
_lock = threading.Lock()


def problem_class(p: dict) -> tuple[int, int]:
    """(m rounded up to a power of two, K rounded up to a power of four)."""
    k = max(1, int(p["nseg"]) * int(p["seg_len"]))
    return (1 << max(0, math.ceil(math.log2(max(1, int(p["m"]))))),
            4 ** max(0, math.ceil(math.log(k, 4) - 1e-12)))


def signature(form: int, nterms: int, nprod: int, problems: list) -> str:
    counts: dict[tuple, int] = {}
    for p in problems:
        c = problem_class(p)
        counts[c] = counts.get(c, 0) + 1
    return json.dumps([form, nterms, nprod,
                       sorted([list(c), n] for c, n in counts.items())])


def default_config(form: int, cls: tuple[int, int]) -> tuple[int, int]:
    split = is_split(form)
    return (tile_for(form, cls[0]), CHAINS[split][0])


@dataclass
class Plan:
    """The tuned assignment: per-problem tile sides and chain lengths, the
    per-class configuration they come from, and the timings behind it."""
    tiles: list[int]
    kchains: list[int]
    config: dict = field(default_factory=dict)
    default_ms: float | None = None
    best_ms: float | None = None
    trials: int = 0
    cached: bool = False

    def describe(self) -> dict:
        return {"config": {f"m<={c[0]},K<={c[1]}": {"tile": t, "kchain": k}
                           for c, (t, k) in sorted(self.config.items())},
                "default_ms": self.default_ms, "best_ms": self.best_ms,
                "trials": self.trials, "cached": self.cached}


def assignment(config: dict, problems: list) -> tuple[list, list]:
    cls = [problem_class(p) for p in problems]
    return ([config[c][0] for c in cls], [config[c][1] for c in cls])


def time_assignment(form, nterms, nprod, problems, tiles, kchains,
                    reps: int = 3) -> float:
    """Mean ms of one grouped run of the batch under this assignment."""
    b = GroupedBatch(form, nterms, nprod, problems, tiles, kchains)
    b.run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _load_file() -> dict:
    path = os.environ.get("MATVIS_AMD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_file(key: str, config: dict):
    path = os.environ.get("MATVIS_AMD_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data[key] = [[list(c), list(v)] for c, v in config.items()]
    with open(path, "w") as fh:
        json.dump(data, fh)


def _cached(key: str):
    with _lock:
        hit = _cache.get(key)
    if hit is None:
        raw = _load_file().get(key)
        if raw is not None:
            hit = {tuple(c): tuple(v) for c, v in raw}
    return hit


def tune(form: int, nterms: int, nprod: int, problems: list,
         budget_s: float = 5.0, reps: int = 3) -> Plan:
    """Per-problem (tile, kchain) for the batch, from cache or by search."""
    key = signature(form, nterms, nprod, problems)
    hit = _cached(key)
    if hit is not None:
        tiles, chains = assignment(hit, problems)
        return Plan(tiles, chains, dict(hit), cached=True)
    split = is_split(form)
    classes = sorted({problem_class(p) for p in problems})
    work = {c: 0.0 for c in classes}
    for p in problems:
        work[problem_class(p)] += float(p["m"]) ** 2 * p["nseg"] * p["seg_len"]
    order = sorted(classes, key=lambda c: -work[c])
    config = {c: default_config(form, c) for c in classes}

    def measure(cfg):
        return time_assignment(form, nterms, nprod, problems,
                               *assignment(cfg, problems), reps=reps)

    t_end = time.perf_counter() + budget_s
    best = default = measure(config)
    trials = 1
    improved = True
    while improved and time.perf_counter() < t_end:
        improved = False
        for c in order:
            tile, chain = config[c]
            cands = [(t, chain) for t in TILES[split] if t != tile]
            cands += [(tile, k) for k in CHAINS[split] if k != chain]
            for cand in cands:
                if time.perf_counter() >= t_end:
                    break
                trial = dict(config)
                trial[c] = cand
                ms = measure(trial)
                trials += 1
                if ms < best * (1.0 - GAIN):
                    best, config, improved = ms, trial, True
    with _lock:
        _cache[key] = dict(config)
    _store_file(key, config)
    tiles, chains = assignment(config, problems)
    return Plan(tiles, chains, config, default_ms=default, best_ms=best,
                trials=trials)
