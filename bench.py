"""Benchmark: visibilities/s of the per-time computation on one MI355X.

Workload (the headline configuration): 350 antennas on a 14.6 m hexagonal grid with
+-2 cm jitter, 100,000 point sources (RA uniform, sin(dec) uniform on
[-90, +30] deg, log-normal(0, 1) Jy fluxes), polarized (nfeed = nax = 2),
20 frequencies over 100-200 MHz, times 2 min apart, one shared complex64
E-field beam on a 360 x 91 (1 deg) az/za grid with bicubic interpolation.
Synthetic, seeded, generated here; inputs are resident in HBM before timing.

A step is one time sample: rotation + cull once, then for each of the 20
frequencies the beam, the fused phase/Z kernel, the HERK and the expansion
into the API layout, results left in device memory.  Forms timed: split-BF16
(the headline `value`; Z = hi + lo bfloat16, three term products on BF16
MFMA with FP32 accumulation per K-slice, FP64 cross-slice sums and an exact
FP64 diagonal), its four-product and three-term variants, split-FP16,
plain complex64 and plain FP64.  The split-BF16 forms are also timed through
the drop-in path into host memory over the workload's 30 times.

Prints one JSON line.  `--gpus N` with N > 1 exits: the library runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

NANT = 350
NSRC = 100_000
NFREQ = 20
NTIMES_POOL = 30
MFMA_BF16_PEAK = 2516.6e12
FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12
#: dense MFMA peaks of MI355X (TFLOP/s) for each form's arithmetic
PEAK_TFLOPS = {"FORM_SPLIT_BF16": 2516.6, "FORM_SPLIT_F16": 2516.6,
               "FORM_C128": 78.6, "FORM_C64": 157.3}


def build_workload(nsrc=NSRC, nant=NANT, nfreq=NFREQ, seed=1):
    from matvis_amd import astro, synth

    antpos = synth.hex_array(nant, 14.6, 0.02, seed=seed)
    ra, dec = synth.random_sky(nsrc, seed + 1, -90.0, 30.0)
    flux0 = synth.lognormal_fluxes(nsrc, seed + 2, 0.0, 1.0)
    freqs = np.linspace(100e6, 200e6, nfreq)
    fluxes = np.repeat(flux0[:, None], nfreq, 1)
    times = synth.times_spaced(NTIMES_POOL, 2.0)
    loc = synth.hera_location()
    beam = synth.harmonic_efield_beam(seed=seed + 3, naz=360, nza=91)
    rot = astro.rotation_matrices(times, loc)
    return dict(antpos=antpos, ra=ra, dec=dec, fluxes=fluxes, freqs=freqs,
                times=times, loc=loc, beam=beam, rotations=rot)


FORMS = {
    # name: (form, nterms, nprod)
    "bf16": ("FORM_SPLIT_BF16", 2, 3),
    "bf16x4": ("FORM_SPLIT_BF16", 2, 4),
    "bf16x3": ("FORM_SPLIT_BF16", 3, 6),
    "f16": ("FORM_SPLIT_F16", 2, 3),
    "fp64": ("FORM_C128", 1, 1),
    "fp32": ("FORM_C64", 1, 1),
}
FORM_LABEL = {
    "bf16": "split-BF16 (2 terms, 3 products)",
    "bf16x4": "split-BF16 (2 terms, 4 products)",
    "bf16x3": "split-BF16 (3 terms, 6 products)",
    "f16": "split-FP16 (2 terms, 3 products)",
    "fp64": "plain FP64",
    "fp32": "plain complex64 (precision=1, the reference's default)",
}


def make_engine(w, form, nterms=1, nprod=1, precision=2):
    from matvis_amd import astro
    from matvis_amd.core.beams import _wrangle_beams
    from matvis_amd.gpu import coords as gcoords
    from matvis_amd.gpu.engine import VisibilityEngine

    nsrc = w["ra"].size
    co = gcoords.CoordinateRotationMatrix(
        rotations=w["rotations"], flux=np.sqrt(0.5 * w["fluxes"][:, 0]),
        times=w["times"], telescope_loc=w["loc"],
        skycoords=astro.SkyCoord(w["ra"], w["dec"]), chunk_size=nsrc,
        precision=precision, source_buffer=1.0)
    co.setup()
    beams, _, _ = _wrangle_beams(None, [w["beam"]], True, len(w["antpos"]),
                                 float(w["freqs"][0]))
    eng = VisibilityEngine(
        antpos=w["antpos"], freqs=w["freqs"], coords=co, fluxes=w["fluxes"],
        beam_list=beams, beam_idx=None, polarized=True, precision=precision,
        form=form, nterms=nterms, nprod=nprod, order=3, nchunks=1,
        chunk_size=nsrc, nsrc_alloc=nsrc)
    eng.setup()
    return eng


class HerkTimer:
    """HIP events around every HERK compute kernel, on the launching
    stream (the split-K reduce that follows is not inside the pair)."""

    def __init__(self):
        import torch

        self.torch = torch
        self.pairs = []
        self.enabled = False

    def before(self, stream):
        if self.enabled:
            e = self.torch.cuda.Event(enable_timing=True)
            e.record(stream)
            self.pairs.append([e, None])

    def after(self, stream):
        if self.enabled:
            e = self.torch.cuda.Event(enable_timing=True)
            e.record(stream)
            self.pairs[-1][1] = e

    def mean_ms(self):
        return float(np.mean([a.elapsed_time(b) for a, b in self.pairs]))


def time_form(eng, steps, warmup, timer=None):
    """Time `steps` time samples; returns (seconds, per-step ms, n_up list)."""
    import torch

    stream = torch.cuda.current_stream()
    out = eng.new_output()
    for i in range(warmup):
        eng.compute_time(i % NTIMES_POOL, out, stream=stream)
    torch.cuda.synchronize()
    if timer is not None:
        eng.herk_hooks = (timer.before, timer.after)
        timer.enabled = True
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    nups = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(steps):
        t = (warmup + i) % NTIMES_POOL
        eng.compute_time(t, out, stream=stream)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if timer is not None:
        timer.enabled = False
        eng.herk_hooks = None
    per_step = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
    for i in range(steps):
        t = (warmup + i) % NTIMES_POOL
        rot = eng.coords._rots[t]
        nups.append(int(((rot[2] @ eng.coords_eci_host) > 0).sum()))
    return elapsed, per_step, nups


def batched_bench(ninst: int = 256, seed: int = 7, reps: int = 5):
    """Many instances in one grouped launch: nant from {64, 128, 350},
    unpolarized / polarized 50/50, n_up log-uniform on [5k, 80k] above the
    horizon, synthetic complex64 Z split to bf16 (2 terms, 3 products);
    tiles from the autotuner; the grouped launches captured in a HIP graph
    and replayed.  Compared with one plain complex64 HERK launch per
    instance, one split launch per instance, and the plain complex64 HERKs
    of the same instances in one grouped, captured launch."""
    import torch

    import matvis_amd._device as D
    import matvis_amd._lib as L
    from matvis_amd.gpu import autotune
    from matvis_amd.gpu.herk import Herk, split_terms_alloc
    from matvis_amd.gpu.herk_group import GroupedBatch

    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    form, nt, npd = L.FORM_SPLIT_BF16, 2, 3
    probs, zs, vis, flops = [], [], 0, 0.0
    for i in range(ninst):
        nant = int(rng.choice([64, 128, 350]))
        nf = 2 if rng.random() < 0.5 else 1
        nup = int(np.exp(rng.uniform(np.log(5000), np.log(80000))))
        m, seg = nant * nf, -(-nup // 16) * 16
        z = torch.randn((m, nf * seg), dtype=torch.complex64, device="cuda",
                        generator=gen)
        z.view(m, nf, seg)[:, :, nup:] = 0
        terms = split_terms_alloc(nt, m, nf * seg)  # m rows a K tile
        L.call("mv_split_z", D.ptr(z), 0, m, nf, nup, nf * seg, seg, form, nt,
               1.0, D.ptr(terms[0]), D.ptr(terms[1]), None, m, seg,
               D.stream_ptr())
        ldv = -(-m // 128) * 128
        probs.append(dict(terms=terms, m=m, ldz=m, nseg=nf,
                          seg_stride=seg, seg_len=nup,
                          vacc=D.zeros((ldv, ldv), np.complex128), ldv=ldv,
                          scale=1.0, accumulate=False))
        zs.append(z)
        vis += nant * nant * nf * nf
        flops += 8.0 * nf * nup * m * (m + 1) / 2
    torch.cuda.synchronize()
    log(f"batched: {ninst} instances built ({flops / 1e12:.1f} Tflop useful)")
    # the tuner on the whole batch, and on its 16 smallest instances (a
    # batch that does not fill the GPU): the picks differ with the mix
    work = [p["m"] ** 2 * p["nseg"] * p["seg_len"] for p in probs]
    small = [probs[i] for i in np.argsort(work)[:16]]
    t0 = time.perf_counter()
    plan_small = autotune.tune(form, nt, npd, small, budget_s=5.0)
    tune_small_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = autotune.tune(form, nt, npd, probs, budget_s=10.0)
    tune_s = time.perf_counter() - t0
    log(f"batched: tuned in {tune_s:.1f} s: {plan.describe()}")
    tiles = plan.tiles
    batch = GroupedBatch(form, nt, npd, probs, plan.tiles,
                         plan.kchains).capture()
    batch.replay()
    torch.cuda.synchronize()
    e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    e0.record()
    for _ in range(reps):
        batch.replay()
    e1.record()
    torch.cuda.synchronize()
    t_group = e0.elapsed_time(e1) / reps

    def sequential(run_one):
        run_one()
        torch.cuda.synchronize()
        a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
        a.record()
        run_one()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    herks = {}

    def herk_for(f, m, n1, n2):
        return herks.setdefault((f, m), Herk(f, m, n1, n2))

    def zld(p):  # row stride of the complex Z (the terms' ldz is m)
        return p["nseg"] * p["seg_stride"]

    def plain_all():
        for p, z in zip(probs, zs):
            h = herk_for(L.FORM_C64, p["m"], 1, 1)
            h.run([z], ldz=zld(p), nseg=p["nseg"], seg_stride=p["seg_stride"],
                  seg_len=p["seg_len"], vacc=p["vacc"], accumulate=False,
                  nslices=h.nslices(p["nseg"] * p["seg_len"]))

    def split_all():
        for p in probs:
            h = herk_for(form, p["m"], nt, npd)
            h.run(p["terms"], ldz=p["ldz"], nseg=p["nseg"],
                  seg_stride=p["seg_stride"], seg_len=p["seg_len"],
                  vacc=p["vacc"], accumulate=False,
                  nslices=h.nslices(p["nseg"] * p["seg_len"]))

    t_plain = sequential(plain_all)
    t_split = sequential(split_all)
    # the plain complex64 HERKs of the same instances as one grouped launch
    from matvis_amd.gpu.herk_group import plain_tile
    probs64 = [dict(p, terms=[z], ldz=zld(p)) for p, z in zip(probs, zs)]
    batch64 = GroupedBatch(L.FORM_C64, 1, 1, probs64,
                           [plain_tile()] * ninst).capture()
    t_group64 = sequential(batch64.replay)
    del batch64, probs64
    # errors of the grouped result against the FP64 HERK on a sample
    batch.replay()
    torch.cuda.synchronize()
    errs, meds = [], []
    for i in rng.choice(ninst, 8, replace=False):
        p, z = probs[i], zs[i]
        h = Herk(L.FORM_C128, p["m"])
        v = h.new_vacc()
        h.run([z.to(torch.complex128)], ldz=zld(p), nseg=p["nseg"],
              seg_stride=p["seg_stride"], seg_len=p["seg_len"], vacc=v,
              accumulate=False, nslices=h.nslices(p["nseg"] * p["seg_len"]))
        m = p["m"]
        ref = torch.triu(v[:m, :m])
        got = torch.triu(p["vacc"][:m, :m])
        d = torch.sqrt(torch.diagonal(ref).real.abs())
        rel = torch.triu(torch.abs(got - ref) / (d[:, None] * d[None]))
        errs.append(float(rel.max()))
        iu = torch.triu_indices(m, m, device=rel.device)
        meds.append(float(torch.median(rel[iu[0], iu[1]])))
    del zs, probs, batch, herks
    torch.cuda.empty_cache()
    return {
        "instances": ninst, "visibilities": vis, "useful_tflop": flops / 1e12,
        "grouped_graph_ms": t_group, "vis_per_s": vis / t_group * 1e3,
        "useful_tflops": flops / t_group / 1e9,
        "sequential_plain_c64_ms": t_plain,
        "sequential_split_bf16_ms": t_split,
        "speedup_vs_sequential_plain": t_plain / t_group,
        "speedup_vs_sequential_split": t_split / t_group,
        "grouped_graph_plain_c64_ms": t_group64,
        "plain_c64_speedup_grouped_vs_sequential": t_plain / t_group64,
        "tiles": {str(k): tiles.count(k) for k in sorted(set(tiles))},
        "autotune": dict(plan.describe(), seconds=tune_s),
        "autotune_16_smallest": dict(plan_small.describe(),
                                     seconds=tune_small_s),
        "max_err_vs_fp64_8_instances": max(errs),
        "median_err_vs_fp64_8_instances": float(np.median(meds)),
    }


def diffuse_bench(ntimes: int = 2, warmup: int = 1):
    """Diffuse-sky scale: HEALPix nside 128 (196,608 pixel centres),
    log-normal(-2, 1.5) brightness, 350 hex antennas, polarized, four +-5%
    perturbations of the 360 x 91 E-field beam assigned round-robin
    (beam_idx = ant % 4), 8 frequencies over 120-180 MHz, source chunks of
    16,384 (12 chunks, FP64 cross-chunk sums).  Times `ntimes` time samples
    of the plain FP64 and the split-BF16 forms; errors of split-BF16 against
    FP64 on the first time, and of an FP32-only cross-chunk sum (each
    chunk's FP64 visibilities rounded to FP32 and summed in FP32) at the
    first frequency."""
    import torch

    from matvis_amd import astro, synth
    from matvis_amd.gpu.gpu import build_engine, simulate
    from matvis_amd.metrics import normalized_error

    ra, dec = synth.healpix_centres(128)
    flux0 = synth.lognormal_fluxes(ra.size, 3, -2.0, 1.5)
    nant, nfreq = 350, 8
    antpos = synth.hex_array(nant, 14.6, 0.02, seed=3)
    beams = [synth.harmonic_efield_beam(seed=4, naz=360, nza=91, perturb=0.05,
                                        perturb_seed=100 + b)
             for b in range(4)]
    freqs = np.linspace(120e6, 180e6, nfreq)
    times = synth.times_spaced(16, 10.0)
    loc = synth.hera_location()
    rot = astro.rotation_matrices(times, loc)
    common = dict(antpos=antpos, times=times, skycoords=astro.SkyCoord(ra, dec),
                  telescope_loc=loc, beam_list=beams,
                  beam_idx=np.arange(nant) % 4, polarized=True, precision=2,
                  coord_method="CoordinateRotationMatrix",
                  coord_method_params={"rotations": rot},
                  beam_spline_opts={"order": 3}, min_chunks=12)
    fluxes = np.repeat(flux0[:, None], nfreq, 1)
    out = {"nsrc": int(ra.size), "nant": nant, "nfreq": nfreq, "chunks": 12,
           "chunk_sources": 16384, "times_timed": ntimes}
    first = {}
    vis_per_time = nfreq * nant * nant * 4
    for name, method in (("fp64", "GPUMatMul"),
                         ("bf16", "GPUMatMulSplitBF16")):
        log(f"diffuse: {name} engine")
        eng, _ = build_engine(freqs=freqs, fluxes=fluxes,
                              matprod_method=method, **common)
        dev = eng.new_output()
        stream = torch.cuda.current_stream()
        for t in range(warmup):
            eng.compute_time(t, dev, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(ntimes):
            eng.compute_time(warmup + t, dev, stream=stream)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        eng.compute_time(0, dev, stream=stream)
        first[name] = dev.cpu().numpy()
        out[name] = {"ms_per_time": 1e3 * sec / ntimes,
                     "vis_per_s": vis_per_time * ntimes / sec,
                     "nchunks": eng.nchunks}
        del eng, dev
        torch.cuda.empty_cache()
    e = [normalized_error(first["bf16"][f], first["fp64"][f], nant, 2)
         for f in range(nfreq)]
    out["bf16_vs_fp64_max_median"] = [max(x[0] for x in e),
                                      float(np.median([x[1] for x in e]))]
    # FP32-only cross-chunk sums at the first frequency and time
    log("diffuse: FP32-only chunk sums")
    one = dict(common, times=times[:1], coord_method_params={"rotations": rot[:1]},
               min_chunks=1)
    acc32 = None
    n = ra.size
    for c in range(12):
        sl = slice(c * n // 12, (c + 1) * n // 12)
        v = simulate(freq=float(freqs[0]), I_sky=flux0[sl],
                     **dict(one, skycoords=astro.SkyCoord(ra[sl], dec[sl])))
        v = v.astype(np.complex64)
        acc32 = v if acc32 is None else acc32 + v
    ref = first["fp64"][0]
    out["fp32_chunk_sums_vs_fp64_max_median"] = list(normalized_error(
        acc32[0].astype(np.complex128), ref, nant, 2)[:2])
    out["speedup_bf16_over_fp64"] = (out["bf16"]["vis_per_s"]
                                     / out["fp64"]["vis_per_s"])
    return out


def serving_requests(seconds: float, rate: float, seed: int = 11):
    """A Poisson request stream: arrival offsets and request keywords.

    Each request: nant from {32, 128, 350, 512} on a 14.6 m hex grid; nsrc
    log-uniform on [1e3, 2e5] (RA uniform, sin(dec) uniform on [-90, +30]
    deg, power-law S^-2 fluxes on [0.1, 10] Jy); nfreq uniform on [1, 64]
    over 100-200 MHz; ntime uniform on [1, 16], 2 min apart; polarized with
    p = 0.5; plain FP64 or split-BF16 with p = 0.5; one shared harmonic
    E-field beam (bicubic).  Sizes from ``synth.serving_draw``, keywords
    from ``synth.ServingSky.request``."""
    from matvis_amd import synth

    rng = np.random.default_rng(seed)
    n = max(1, int(rate * seconds * 2) + 16)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals = arrivals[arrivals < seconds]
    sky = synth.ServingSky()
    reqs = []
    for i in range(arrivals.size):
        d = synth.serving_draw(rng)
        kw = sky.request(seed=seed * 1000 + i, **d)
        label = ("fp64" if d["fp64"] else "bf16") + (
            "/pol" if d["polarized"] else "/unpol")
        work = (d["ntime"] * d["nfreq"]
                * (d["nant"] * (2 if d["polarized"] else 1)) ** 2 * d["nsrc"])
        reqs.append((float(arrivals[i]), kw, label, work))
    return reqs


def serving_bench(seconds: float = 60.0, rate: float = 20.0,
                  drain: float = 30.0, seed: int = 11, max_active: int = 16,
                  policy: str = "fifo", errors: bool = True):
    """The resident server under a Poisson stream (arrivals for `seconds`,
    then up to `drain` s to finish).  Reports latency of the finished
    requests, lower bounds that count unfinished ones at their age, the
    visibilities delivered per second and, for the two smallest finished
    requests of each class, the error against the same request run alone on
    the plain FP64 GPU path (which the GPU tests hold within 1e-12 of the
    oracle).  Latency trend: the p99 (unfinished requests at their age) of
    the requests arriving in the first and in the last third of the window;
    a backlog that grows shows as a last-third p99 well above the first."""
    import threading

    import torch

    from matvis_amd.gpu.gpu import simulate_multi
    from matvis_amd.metrics import normalized_error
    from matvis_amd.serve import SimRequest, VisibilityServer

    stream = serving_requests(seconds, rate, seed)
    log(f"serving: {len(stream)} requests over {seconds:.0f} s "
        f"(Poisson {rate:g}/s)")
    # per class, the two smallest of its first ten arrivals keep their whole
    # result (early, so that they finish under saturation)
    keep = set()
    for lb in {s[2] for s in stream}:
        idx = [i for i, s in enumerate(stream) if s[2] == lb][:10]
        keep.update(sorted(idx, key=lambda i: stream[i][3])[:2])
    srv = VisibilityServer(max_active=max_active, policy=policy)
    reqs = [SimRequest(kw, keep=i in keep, label=lb)
            for i, (_, kw, lb, _) in enumerate(stream)]
    srv.start()
    t0 = time.perf_counter()
    sent = 0

    def feeder():
        nonlocal sent
        for (at, _, _, _), r in zip(stream, reqs):
            wait = t0 + at - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            srv.submit(r)
            sent += 1

    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    last = t0
    backlog = []  # (s since start, queued + active) every 0.5 s
    while True:
        now = time.perf_counter()
        if not backlog or now - t0 - backlog[-1][0] >= 0.5:
            backlog.append((now - t0, len(srv.queue) + len(srv.active)))
        if th.is_alive() is False and all(r.done.is_set() for r in reqs):
            break
        if now - t0 > seconds + drain:
            break
        if now - last > 5:
            st = srv.stats()
            log(f"serving: t={now - t0:.0f}s sent {sent} finished "
                f"{st['requests']} active {len(srv.active)} queued "
                f"{len(srv.queue)}")
            last = now
        time.sleep(0.05)
    t_end = time.perf_counter()
    th.join()
    srv.stop()
    torch.cuda.synchronize()
    elapsed = t_end - t0
    st = srv.stats()
    # unfinished requests count at their age when the run ended
    ages = np.array([(r.t_done if r.done.is_set() else t_end) - r.t_submit
                     for r in reqs])
    at = np.array([a for a, _, _, _ in stream])

    def p99_of(mask):
        return float(np.percentile(ages[mask], 99)) if mask.any() else None
    out = {
        "arrival_window_s": seconds, "rate_per_s": rate, "drain_s": drain,
        "elapsed_s": elapsed, "submitted": len(reqs),
        "finished": st["requests"], "failed": st["failed"],
        "unfinished": sum(not r.done.is_set() for r in reqs),
        "p50_s": st["p50_s"], "p99_s": st["p99_s"],
        "p50_s_all_lower_bound": float(np.percentile(ages, 50)),
        "p99_s_all_lower_bound": float(np.percentile(ages, 99)),
        "vis_per_s": st["visibilities_all"] / elapsed,
        "finished_per_s": st["requests"] / elapsed,
        "p99_s_first_third": p99_of(at < seconds / 3),
        "p99_s_last_third": p99_of(at >= 2 * seconds / 3),
        "backlog_slope_per_s": backlog_slope(backlog, seconds),
        "backlog_max": max(b for _, b in backlog),
        "drain_needed_s": max(0.0, max(
            (r.t_done - t0 for r in reqs if r.done.is_set()),
            default=0.0) - seconds),
        "by_class": st.get("by_label", {}),
        "max_active": max_active,
        "policy": policy,
    }
    del srv
    torch.cuda.empty_cache()
    if not errors:
        return out

    def nerr(got, ref, nant, nf):
        a = got.reshape(-1, nant * nant, nf, nf)
        b = ref.reshape(-1, nant * nant, nf, nf)
        e = [normalized_error(a[j], b[j], nant, nf) for j in range(a.shape[0])]
        return max(x[0] for x in e), [x[1] for x in e]

    # served result vs the request alone on the plain FP64 path; for split
    # requests also the request alone through simulate_multi in its own
    # form, which separates the form's rounding from anything the mix adds
    errs = {}
    for i in sorted(keep):
        r = reqs[i]
        if r.result is None:
            continue
        nant = len(r.kwargs["antpos"])
        nf = 2 if r.kwargs["polarized"] else 1
        ref = simulate_multi(**dict(r.kwargs, matprod_method="GPUMatMul"))
        emax, emed = nerr(r.result, ref, nant, nf)
        cls = errs.setdefault(r.label, {"max": 0.0, "median": [],
                                        "requests": []})
        cls["max"] = max(cls["max"], emax)
        cls["median"] += emed
        desc = {"nant": nant, "nsrc": int(r.kwargs["fluxes"].shape[0]),
                "instances": int(r.result.shape[0] * r.result.shape[1])}
        if r.kwargs["matprod_method"] != "GPUMatMul":
            alone = simulate_multi(**r.kwargs)
            desc["alone_same_form_max"] = nerr(alone, ref, nant, nf)[0]
        cls["requests"].append(desc)
    for cls in errs.values():
        cls["median"] = float(np.median(cls["median"]))
    out["error_vs_fp64_alone_normalised"] = errs
    return out


def backlog_slope(samples, seconds: float) -> float:
    """Least-squares slope (requests/s) of the backlog (queued + active)
    over the last two thirds of the arrival window: about zero while the
    server keeps up, the arrival rate minus the service rate once it does
    not."""
    pts = np.array([(t, b) for t, b in samples
                    if seconds / 3 <= t <= seconds], dtype=np.float64)
    if len(pts) < 2:
        return 0.0
    return float(np.polyfit(pts[:, 0], pts[:, 1], 1)[0])


def serving_sweep(rates, seconds: float = 40.0, drain: float = 20.0,
                  seed: int = 11, policy: str = "fifo") -> dict:
    """The serving workload at a ladder of Poisson rates under `policy`.  A rate is
    sustainable when the backlog does not grow (its slope over the last two
    thirds of the window is under 10% of the arrival rate) and every request
    finished within the drain; the sweep stops at the first rate that is
    not, and the sustainable rate is the largest one before it."""
    runs = []
    for rate in rates:
        log(f"serving sweep: {rate:g} req/s for {seconds:g} s")
        r = serving_bench(seconds, rate, drain, seed=seed, policy=policy,
                          errors=False)
        r["sustainable"] = bool(
            r["unfinished"] == 0 and r["failed"] == 0
            and r["backlog_slope_per_s"] < 0.1 * rate)
        runs.append(r)
        log(f"serving sweep: {rate:g}/s finished {r['finished']} of "
            f"{r['submitted']}, backlog slope "
            f"{r['backlog_slope_per_s']:.2f}/s, p50 {r['p50_s']} p99 "
            f"{r['p99_s']} s, sustainable {r['sustainable']}")
        if not r["sustainable"]:
            break
    ok = [r["rate_per_s"] for r in runs if r["sustainable"]]
    return {"window_s": seconds, "drain_s": drain, "policy": policy,
            "sustainable_rate_per_s": max(ok) if ok else None,
            "rule": "backlog slope over the last 2/3 of the window < 0.1 x "
                    "rate and every request finished within the drain",
            "runs": runs}


def host_output_rate(eng, ntimes: int, vis_per_time: int) -> dict:
    """The drop-in API's rate: ``engine.run`` returning host arrays (each
    time's output copied on a second stream into pinned buffers, then into
    the caller's array), PCIe and host copies included; wall clock."""
    import torch

    out = np.empty((eng.nfreq, ntimes, eng.npairs, eng.nfeed, eng.nfeed),
                   dtype=eng.ctype)
    out.fill(0)  # first touch outside the timed region
    eng.run(1, out[:, :1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(ntimes, out)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    log(f"host output: {ntimes} times in {sec:.3f} s")
    return dict(times=ntimes, seconds=sec,
                vis_per_s=ntimes * vis_per_time / sec,
                output_bytes=int(out.nbytes))


def beam_interp_bench(count: int = 59700, reps: int = 50) -> list:
    """Bicubic interpolation of a polarized beam (4 planes) at `count`
    random directions over the upper hemisphere: the tap-gather kernel
    (mv_beam_interp) against the za bands staged in LDS
    (mv_beam_interp_banded), at 1, 0.5 and 0.25 deg grids in complex64 and
    complex128; HIP-event means over `reps` launches after 3 warm ones."""
    import torch

    import matvis_amd._device as D
    import matvis_amd._lib as L

    rng = np.random.default_rng(3)
    za = np.arccos(rng.uniform(0.0, 1.0, count))
    az = rng.uniform(0, 2 * np.pi, count)
    crd = D.to_dev(np.stack([np.sin(za) * np.sin(az),
                             np.sin(za) * np.cos(az), np.cos(za)]), np.float64)
    rows = []
    for naz, nza in ((363, 91), (723, 181), (1443, 361)):
        for is_f64 in (0, 1):
            dt = np.complex128 if is_f64 else np.complex64
            planes = D.to_dev((rng.normal(size=(1, 2, 2, nza, naz))
                               + 1j * rng.normal(size=(1, 2, 2, nza, naz)))
                              .astype(dt))
            grid = D.to_dev(np.array([[2 * np.pi / (naz - 3),
                                       (np.pi / 2) / (nza - 1), 0.0]]),
                            np.float64)
            out = D.zeros((4, count), dt)
            need = L.lib().mv_beam_interp_banded_workspace_bytes(
                is_f64, nza, naz, count)
            ws = D.empty((max(need, 1),), np.uint8)
            args = (D.ptr(planes), is_f64, 1, 2, 2, nza, naz, D.ptr(grid), 3,
                    0, D.ptr(crd), 1, count, count, None, D.ptr(out), count)
            row = {"grid": f"{naz}x{nza}", "dtype": "c128" if is_f64 else "c64",
                   "planes_mb": planes.numel() * planes.element_size() / 1e6}
            for name in ("gather", "banded"):
                def run():
                    if name == "gather":
                        L.call("mv_beam_interp", *args, D.stream_ptr())
                    else:
                        L.call("mv_beam_interp_banded", *args, D.ptr(ws),
                               ws.numel(), D.stream_ptr())
                for _ in range(3):
                    run()
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    run()
                e1.record()
                torch.cuda.synchronize()
                row[f"{name}_us"] = 1e3 * e0.elapsed_time(e1) / reps
            log(f"beam interpolation {row}")
            rows.append(row)
    return rows


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def blas_threads() -> int:
    """Threads numpy's BLAS runs with (the CPU baseline's core count)."""
    try:
        from threadpoolctl import threadpool_info

        n = [i["num_threads"] for i in threadpool_info()
             if i.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return os.cpu_count() or 1


TRAFFIC = ROOT / "profiles" / "hbm_traffic_350ant_100ksrc.json"

#: the kernels of one headline step (one time, every frequency) in the
#: committed counter summary: (stage, kernel key, launches per step)
STAGE_KERNELS = [
    ("rotate+cull: count", "mv::cull_count_kernel", 1),
    ("rotate+cull: scan", "mv::cull_scan_kernel", 1),
    ("rotate+cull: scatter", "mv::cull_scatter_kernel<double>", 1),
    ("beam interpolation", "mv::interp_kernel<float, double, 3, false>", 1),
    ("fused Z (split-BF16 terms)", "mv::zmat_fused_kernel<2, 2, true>",
     "nfreq"),
    ("HERK (split-BF16)", "mv::herk_split_kernel<2, 1, false, 128>",
     "nfreq"),
    ("split-K reduce", "mv::herk_reduce_kernel<float>", "nfreq"),
    ("expansion to the API layout", "mv::expand_kernel<double>", "nfreq"),
]


def _traffic(nant: int, nsrc: int):
    """The committed counter summary (FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by what the counter reads for a 1 GiB copy) when
    it was measured on this workload; None otherwise."""
    try:
        rec = json.loads(TRAFFIC.read_text())
    except (OSError, ValueError):
        return None
    wl = rec.get("_workload", {})
    if wl.get("nant") != nant or wl.get("nsrc") != nsrc:
        return None
    return rec


def measured_traffic(kernel: str, nant: int, nsrc: int):
    """HBM bytes per launch of `kernel` from the counter summary, or None."""
    rec = _traffic(nant, nsrc)
    if rec is None or kernel not in rec:
        return None
    return float(rec[kernel]["bytes_per_launch"])


def stage_traffic(nant: int, nsrc: int, nfreq: int):
    """Counter-measured HBM bytes and GB/s of every kernel in one step
    (durations: the same dispatches' own timestamps in the counter runs)."""
    rec = _traffic(nant, nsrc)
    if rec is None:
        return None
    out = {}
    for stage, k, per in STAGE_KERNELS:
        if k not in rec:
            continue
        r = rec[k]
        out[stage] = {
            "kernel": k, "launches_per_step": nfreq if per == "nfreq" else per,
            "read_bytes": r["read_bytes"], "write_bytes": r["write_bytes"],
            "us": r["mean_us_counter_runs"], "gb_per_s": r["gb_per_s"],
            "frac_of_hbm_peak": r["gb_per_s"] * 1e9 / HBM_PEAK}
    return out


def _kernel_key(name: str) -> str:
    """`ns::kernel<args>` from a demangled kernel name (parameters dropped)."""
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            name = name[:i]
            break
    name = name.strip()
    return name[5:] if name.startswith("void ") else name


def summarize_counters(fetch_dir: str, write_dir: str) -> dict:
    """Per-kernel HBM traffic from two rocprofv3 counter runs of the same
    command (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch): means over the
    dispatches of each kernel, FETCH_SIZE and WRITE_SIZE each scaled by what
    the counter reads for the 1 GiB device copy of --copy-probe, and the
    mean dispatch duration of the counter runs (End - Start timestamps)."""
    import csv
    import glob

    def load(d, counter):
        per = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"),
                           recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] != counter:
                    continue
                k = _kernel_key(r["Kernel_Name"])
                dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                per.setdefault(k, []).append(
                    (float(r["Counter_Value"]) * 1024.0, dur))
        return per

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    gib = float(1 << 30)
    probe = [k for k in write if "copyBuffer" in k]
    if not probe:
        raise RuntimeError("no copy-probe dispatch in the counter runs")
    # the probe is the copy whose writes are largest (1 GiB)
    pw = max(v for k in probe for v, _ in write[k])
    pf = max(v for k in probe for v, _ in fetch.get(k, [(0.0, 0)]))
    fs, ws = gib / pf, gib / pw
    out = {"_calibration": {
        "fetch_scale": fs, "write_scale": ws,
        "probe": "one 1 GiB device-to-device copy (bench.py --copy-probe)"}}
    for k in sorted(set(fetch) & set(write)):
        if "copyBuffer" in k:
            continue
        rb = fs * float(np.mean([v for v, _ in fetch[k]]))
        wb = ws * float(np.mean([v for v, _ in write[k]]))
        us = float(np.mean([d for _, d in fetch[k] + write[k]])) / 1e3
        out[k] = {"dispatches": len(fetch[k]), "read_bytes": rb,
                  "write_bytes": wb, "bytes_per_launch": rb + wb,
                  "mean_us_counter_runs": us,
                  "gb_per_s": (rb + wb) / (us * 1e3)}
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(w, ninst_freqs=2):
    """The oracle on one time x `ninst_freqs` frequencies, all BLAS threads."""
    import oracle as O

    cores = blas_threads()
    eci = O.point_source_crd_eq(w["ra"], w["dec"])
    b = w["beam"]
    planes, daz, dza, azmin = O.prepare_grid(
        b.data_array, b.axis1_array, b.axis2_array, b.freq_array,
        float(w["freqs"][0]), True, b.beam_type)
    beam = {"kind": "grid", "planes": planes[None], "daz": [daz],
            "dza": [dza], "azmin": [azmin], "order": 3}
    t0 = time.perf_counter()
    outs = []
    for f in range(ninst_freqs):
        log(f"cpu baseline instance {f + 1}/{ninst_freqs} ({cores} threads)")
        outs.append(O.simulate(antpos=w["antpos"], freq=w["freqs"][f], eci=eci,
                               rotations=w["rotations"][:1],
                               I_sky=w["fluxes"][:, f], beam=beam,
                               polarized=True))
    dt = time.perf_counter() - t0
    nant = len(w["antpos"])
    vis = ninst_freqs * nant * nant * 4
    return vis / dt, cores, outs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nsrc", type=int, default=NSRC)
    ap.add_argument("--nant", type=int, default=NANT)
    ap.add_argument("--nfreq", type=int, default=NFREQ)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline (for profiling runs)")
    ap.add_argument("--forms", default="bf16,bf16x4,bf16x3,f16,fp32,fp64",
                    help="comma list of " + ",".join(FORMS))
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the grouped-HERK batch (256 mixed instances)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving run (Poisson request stream)")
    ap.add_argument("--serve-sweep", default=None,
                    help="comma-separated Poisson rates (req/s): run only "
                         "the serving sweep at these rates and print it")
    ap.add_argument("--serve-policy", default="fifo",
                    choices=("fifo", "shortest"),
                    help="admission policy of the serving sweep")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="arrival window of the serving run")
    ap.add_argument("--serve-rate", type=float, default=20.0,
                    help="mean requests per second of the serving run")
    ap.add_argument("--serve-drain", type=float, default=20.0,
                    help="seconds after the window for requests to finish")
    ap.add_argument("--no-diffuse", action="store_true",
                    help="skip the diffuse-sky (HEALPix nside 128) run")
    ap.add_argument("--summarize-counters", nargs=2,
                    metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print per-kernel HBM traffic (JSON) from two "
                    "rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of a "
                    "command that included --copy-probe, and exit")
    ap.add_argument("--beam-interp", action="store_true",
                    help="time beam interpolation, tap gathers against za "
                         "bands staged in LDS, print JSON and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device-to-device copy first (the "
                         "known-size kernel that calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.summarize_counters:
        print(json.dumps(summarize_counters(*args.summarize_counters),
                         indent=1))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: matvis_amd has no multi-GPU path "
              "(instances shard by (time, frequency) across GPUs)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import matvis_amd._lib as L
    from matvis_amd.coordinates import point_source_crd_eq
    from matvis_amd.metrics import normalized_error

    torch.cuda.set_device(0)
    if args.copy_probe:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.beam_interp:
        print(json.dumps({"beam_interp": beam_interp_bench()}))
        return
    if args.serve_sweep:
        rates = [float(x) for x in args.serve_sweep.split(",")]
        print(json.dumps({"serving_sweep": serving_sweep(
            rates, args.serve_seconds, args.serve_drain,
            policy=args.serve_policy)}))
        return
    nant, nfreq = args.nant, args.nfreq
    log(f"building workload: {nant} antennas x {args.nsrc} sources")
    w = build_workload(nsrc=args.nsrc, nant=nant, nfreq=nfreq)
    eci_h = point_source_crd_eq(w["ra"], w["dec"])
    forms = args.forms.split(",")
    results = {}
    outs = {}
    vis_per_inst = nant * nant * 4
    m = 2 * nant
    for name in forms:
        fname, nt, npd = FORMS[name]
        form = getattr(L, fname)
        log(f"form {name}: setting up")
        eng = make_engine(w, form, nt, npd,
                          precision=1 if name == "fp32" else 2)
        eng.coords_eci_host = eci_h
        log(f"form {name}: timing {args.steps} steps after {args.warmup} "
            "warmup")
        timer = HerkTimer()
        steps = args.steps
        elapsed, per_step, nups = time_form(eng, steps, args.warmup, timer)
        herk_ms = timer.mean_ms()
        k_mean = 2 * float(np.mean(nups))
        alg_flops = 8.0 * k_mean * m * (m + 1) / 2
        inst = steps * nfreq
        log(f"form {name}: {1e3 * elapsed / steps:.1f} ms/step, HERK "
            f"{herk_ms:.3f} ms/launch")
        results[name] = dict(
            seconds=elapsed, ms_per_step=1e3 * elapsed / steps,
            vis_per_s=inst * vis_per_inst / elapsed,
            src_bl_per_s=nfreq * sum(n * nant * (nant + 1) / 2 for n in nups)
            / elapsed,
            p50_ms=float(np.percentile(per_step, 50)),
            p99_ms=float(np.percentile(per_step, 99)),
            herk_ms=herk_ms, herk_alg_flops=alg_flops,
            herk_exec_flops=npd * alg_flops,
            herk_tflops=alg_flops / herk_ms / 1e9,
            herk_exec_tflops=npd * alg_flops / herk_ms / 1e9,
            herk_peak_tflops=PEAK_TFLOPS[FORMS[name][0]],
            herk_frac_of_peak=npd * alg_flops / herk_ms / 1e9
            / PEAK_TFLOPS[FORMS[name][0]],
            n_up_mean=float(np.mean(nups)),
        )
        if FORMS[name][0] == "FORM_SPLIT_BF16" or name == forms[0]:
            results[name]["host_output"] = host_output_rate(
                eng, NTIMES_POOL, nfreq * vis_per_inst)
        # keep one instance (time 0, first frequencies) for the error figures
        out = eng.new_output()
        eng.compute_time(0, out)
        torch.cuda.synchronize()
        outs[name] = out[:2].cpu().numpy().astype(np.complex128)
        del eng
        torch.cuda.empty_cache()

    def nerr(a, b):
        return normalized_error(a, b, nant, 2)

    err = {}
    if "fp64" in outs:
        for name in outs:
            if name != "fp64":
                err[f"{name}_vs_fp64_gpu"] = nerr(outs[name], outs["fp64"])
    cpu = None
    if not args.no_cpu:
        cpu_vis, cores, ref = cpu_baseline(w, 2)
        ref = np.stack([r[0] for r in ref])
        for name in outs:
            err[f"{name}_vs_oracle"] = nerr(outs[name], ref)
        cpu = {"value": cpu_vis, "unit": "vis/s", "cores": cores,
               "host_cpus": os.cpu_count(),
               "affinity_cpus": len(os.sched_getaffinity(0)),
               "kind": "port",
               "sample": "time 0 x first 2 frequencies of the same workload: "
                         "2 instances through the numpy oracle (BLAS ZGEMM "
                         "over the padded source axis) on " + cpu_model()}

    # headline: the fastest split-BF16 form within 1e-6 (normalised max) of
    # the FP64 GPU path; else the most accurate split-BF16 form measured;
    # FP64 itself if no BF16 form ran
    bf = [n for n in results if FORMS[n][0] == "FORM_SPLIT_BF16"]
    ok = [n for n in bf if err.get(f"{n}_vs_fp64_gpu", (1.0,))[0] <= 1e-6]
    if ok:
        head_name = max(ok, key=lambda n: results[n]["vis_per_s"])
    elif bf:
        head_name = min(bf, key=lambda n: err.get(f"{n}_vs_fp64_gpu",
                                                  (1.0,))[0])
    else:
        head_name = "fp64" if "fp64" in results else forms[0]
    head = results[head_name]
    is_split = head_name != "fp64"
    peak = MFMA_BF16_PEAK if is_split else FP64_PEAK
    # split forms: the executed term-product flops on the 16-bit MFMA
    # (nprod x the useful 8 k m(m+1)/2); FP64: the useful flops
    achieved = head["herk_exec_tflops"]
    line = {
        "metric": "vis/s at 350 ant\u00d7100k src pol, split-BF16 vs FP64 GPU "
                  "@\u22641e-6 rel err; p50/p99",
        "value": head["vis_per_s"],
        "unit": "vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16": "f16", "fp64": "f64"}.get(head_name, "bf16"),
        "data": "synthetic (seeded hex array, random sky, harmonic E-field "
                "beam)",
        "config": {"workload": f"{nant} antennas x {args.nsrc} sources, "
                               f"polarized, {nfreq} freqs, bicubic c64 beam; "
                               f"step = 1 time x {nfreq} freqs",
                   "nant": nant, "nsrc": args.nsrc, "nfreq": nfreq,
                   "form": FORM_LABEL[head_name]},
        "meets_1e-6": bool(head_name in ok) if head_name != "fp64" else None,
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        # the drop-in API's rate with the visibilities copied back into the
        # caller's host arrays (PCIe and host copies included)
        "value_host_delivered": head.get("host_output", {}).get("vis_per_s"),
        "roofline": {"bound": "mfma", "achieved": achieved,
                     "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / (peak / 1e12),
                     "traffic": measured_traffic(
                         "mv::herk_split_kernel<2, 1, false, 128>"
                         if head_name == "bf16"
                         else "mv::herk_plain_kernel<double>"
                         if head_name == "fp64" else "", nant, args.nsrc),
                     "kernel": "herk_split_kernel" if is_split
                     else "herk_plain_kernel<double>",
                     "useful_tflops": head["herk_tflops"]},
        "cpu_baseline": cpu,
        # the counter runs profiled the two-term, three-product step
        "stages": (stage_traffic(nant, args.nsrc, nfreq)
                   if head_name == "bf16" else None),
        "forms": results,
        "errors_normalised_max_median_elemmax_elemmedian": err,
    }
    if not args.no_batched:
        log("batched: grouped HERK over 256 mixed instances")
        line["batched"] = batched_bench()
    if not args.no_diffuse:
        line["diffuse"] = diffuse_bench()
    if not args.no_serving:
        line["serving"] = serving_bench(args.serve_seconds, args.serve_rate,
                                        args.serve_drain, policy="fifo")
        line["serving_shortest_first"] = serving_bench(
            args.serve_seconds, args.serve_rate, args.serve_drain,
            policy="shortest")
    if "fp64" in results:
        line["speedup_over_fp64"] = {
            n: results[n]["vis_per_s"] / results["fp64"]["vis_per_s"]
            for n in results if n != "fp64"}
    if "fp32" in results:
        # against the plain complex64 form (precision=1, the reference's
        # default), whose accuracy the three-term split matches or betters
        line["speedup_over_fp32"] = {
            n: results[n]["vis_per_s"] / results["fp32"]["vis_per_s"]
            for n in results if n not in ("fp32", "fp64")}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
