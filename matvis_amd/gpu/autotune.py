"""Tile choice for grouped split HERKs, by measurement.

A grouped launch runs one tile side per launch (64 or 128 rows per
workgroup tile).  Small problems waste most of a 128-row tile (m = 64 keeps
one of its four waves busy) while large ones lose operand reuse on 64-row
tiles, and where the crossover lies depends on the batch's mix of sizes.
:func:`choose_tiles` times a few tile-assignment policies on the batch
itself (a grouped launch per tile class, HIP events) and returns the fastest
assignment; results are cached per batch signature (split form and the
multiset of (m, K rounded to a power of two)) in memory and, if
``MATVIS_AMD_TUNE_CACHE`` names a file, in that JSON file.
"""

from __future__ import annotations

import json
import math
import os
import threading

import torch

#: policy name -> largest m assigned to 64-row tiles (0: all 128)
POLICIES = {"all128": 0, "m<=128:64": 128, "m<=256:64": 256,
            "m<=384:64": 384}

_cache: dict[str, str] = {}
_lock = threading.Lock()


def signature(form: int, nterms: int, nprod: int, problems: list) -> str:
    sizes = sorted((p["m"], 1 << max(0, math.ceil(math.log2(max(
        1, p["nseg"] * p["seg_len"]))))) for p in problems)
    return json.dumps([form, nterms, nprod, sizes])


def tiles_for(policy: str, problems: list) -> list[int]:
    limit = POLICIES[policy]
    return [64 if p["m"] <= limit else 128 for p in problems]


def _load_file():
    path = os.environ.get("MATVIS_AMD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_file(key: str, policy: str):
    path = os.environ.get("MATVIS_AMD_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data[key] = policy
    with open(path, "w") as fh:
        json.dump(data, fh)


def time_policy(form, nterms, nprod, problems, policy, reps=3):
    """Mean ms of one grouped run of the batch under `policy`."""
    from .herk_group import GroupedBatch

    b = GroupedBatch(form, nterms, nprod, problems, tiles_for(policy, problems))
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


def choose_tiles(form: int, nterms: int, nprod: int, problems: list,
                 policies=None, report: dict | None = None) -> list[int]:
    """Per-problem tile sides for the batch, from cache or by timing each
    policy.  The problems' accumulators are overwritten while tuning.
    ``report``, if given, receives {policy: ms} of the timed candidates."""
    key = signature(form, nterms, nprod, problems)
    with _lock:
        hit = _cache.get(key) or _load_file().get(key)
    if hit in POLICIES:
        return tiles_for(hit, problems)
    times = {}
    for pol in policies or POLICIES:
        tiles = tiles_for(pol, problems)
        # policies that give the same assignment need one timing
        same = [q for q in times if tiles_for(q, problems) == tiles]
        times[pol] = times[same[0]] if same else time_policy(
            form, nterms, nprod, problems, pol)
    best = min(times, key=times.get)
    if report is not None:
        report.update(times)
    with _lock:
        _cache[key] = best
    _store_file(key, best)
    return tiles_for(best, problems)
