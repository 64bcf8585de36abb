"""The resident server: results identical alone and in a mix (with and
without graphs), equal to simulate_multi, admission within the budget."""

import numpy as np
import pytest

from _problems import small_problem
from matvis_amd import astro
from matvis_amd.gpu.gpu import simulate_multi
from matvis_amd.serve import SimRequest, VisibilityServer

pytestmark = pytest.mark.gpu


def _kwargs(nant, nsrc, nfreq, ntimes, polarized, method, seed):
    p = small_problem(nant=nant, nsrc=nsrc, ntimes=ntimes,
                      polarized=polarized, seed=seed, beam="grid", order=3)
    freqs = np.linspace(110e6, 190e6, nfreq)
    fluxes = np.repeat(p["I_sky"][:, None], nfreq, 1) * np.linspace(
        1.0, 0.7, nfreq)
    return dict(antpos=p["antpos"], freqs=freqs, times=p["times"],
                skycoords=astro.SkyCoord(p["ra"], p["dec"]),
                telescope_loc=p["loc"], fluxes=fluxes,
                beam_list=p["beam_list"], polarized=polarized, precision=2,
                coord_method="CoordinateRotationMatrix",
                coord_method_params={"rotations": p["rotations"]},
                beam_spline_opts={"order": 3}, matprod_method=method)


MIX = [
    (16, 1500, 3, 2, True, "GPUMatMulSplitBF16", 1),
    (8, 800, 2, 3, False, "GPUMatMul", 2),
    (40, 3000, 2, 1, True, "GPUMatMulSplitBF16", 3),
    (70, 2500, 1, 2, False, "GPUMatMulSplitF16", 4),
    (24, 1200, 2, 2, True, "GPUMatMul", 5),
]


def _serve(specs, graphs=True, budget=None):
    srv = VisibilityServer(hbm_budget=budget, graphs=graphs)
    reqs = [srv.submit(SimRequest(_kwargs(*s))) for s in specs]
    srv.run_until_idle(timeout=300)
    for r in reqs:
        assert r.error is None, r.error
        assert r.done.is_set()
    return srv, [r.result for r in reqs]


@pytest.mark.parametrize("graphs", [True, False])
def test_results_same_alone_and_in_mix(graphs):
    print(f"serve mix graphs={graphs}", flush=True)
    alone = [_serve([s], graphs)[1][0] for s in MIX]
    srv, mixed = _serve(MIX, graphs)
    for s, a, b in zip(MIX, alone, mixed):
        assert np.array_equal(a, b), s
    st = srv.stats()
    print(f"  {st}", flush=True)
    assert st["requests"] == len(MIX) and st["failed"] == 0


def test_server_matches_simulate_multi():
    print("serve vs simulate_multi", flush=True)
    _, got = _serve(MIX)
    for s, g in zip(MIX, got):
        want = simulate_multi(**_kwargs(*s))
        # grouped vs single-problem launches slice K differently: FP64
        # agrees to rounding, the split forms to their accumulation error
        import oracle as O
        nf = 2 if s[4] else 1
        bound = 1e-12 if s[5] == "GPUMatMul" else 1e-6
        for fi in range(g.shape[0]):
            a, b = g[fi], want[fi]
            if not s[4]:
                a, b = a[..., None, None], b[..., None, None]
            e = O.normalized_error(a, b, s[0], nf)[0]
            assert e < bound, (s, fi, e)


def test_admission_within_budget():
    print("serve admission", flush=True)
    from matvis_amd.serve import projected_bytes
    need = [projected_bytes(_kwargs(*s)) for s in MIX]
    budget = max(need) + min(need) // 2  # at most one large request at a time
    srv, got = _serve(MIX, budget=budget)
    assert srv.stats()["requests"] == len(MIX)
    assert srv.used == 0 and not srv.active


def test_streamed_results_and_labels():
    """keep=False with an on_time consumer sees the same values as a kept
    result; statistics per label; device beam tables shared."""
    print("serve streamed", flush=True)
    _, kept = _serve(MIX)
    srv = VisibilityServer()
    got = {}

    def consume(req, t, vals):
        got.setdefault(req.rid, {})[t] = vals.copy()

    reqs = [srv.submit(SimRequest(_kwargs(*s), keep=False, on_time=consume,
                                  label=s[5])) for s in MIX]
    srv.run_until_idle(timeout=300)
    for r, want in zip(reqs, kept):
        assert r.error is None and r.result is None
        times = got[r.rid]
        assert sorted(times) == list(range(want.shape[1]))
        for t, v in times.items():
            assert np.array_equal(v, want[:, t])
    st = srv.stats()
    assert set(st["by_label"]) == {s[5] for s in MIX}
    assert sum(v["requests"] for v in st["by_label"].values()) == len(MIX)
    # one grid beam object per request: tables keyed by it, not rebuilt for
    # a second request on the same beam
    n_tables = len(srv._beams)
    again = srv.submit(SimRequest(reqs[0].kwargs))
    srv.run_until_idle(timeout=300)
    assert again.error is None and len(srv._beams) == n_tables


def test_requests_admitted_while_others_run():
    """Continuous batching from the background thread: requests submitted
    while a long one is resident join the running set before it is done,
    and every result equals the one computed alone."""
    import time

    print("serve background", flush=True)
    long_spec = (40, 3000, 4, 8, True, "GPUMatMulSplitBF16", 9)
    specs = [long_spec] + MIX
    kws = [_kwargs(*s) for s in specs]
    alone = [_serve([s])[1][0] for s in specs]
    srv = VisibilityServer()
    srv.start()
    try:
        first = srv.submit(SimRequest(kws[0]))
        t0 = time.perf_counter()
        while not any(a.req is first for a in list(srv.active)):
            assert time.perf_counter() - t0 < 60
            time.sleep(0.0005)
        rest = [srv.submit(SimRequest(k)) for k in kws[1:]]
        for r in [first] + rest:
            assert r.done.wait(300), "request did not finish"
    finally:
        srv.stop()
    for r, a, s in zip([first] + rest, alone, specs):
        assert r.error is None, r.error
        assert np.array_equal(r.result, a), s
    # the later requests ran while the long one was resident
    assert all(r.t_admit < first.t_done for r in rest)


def test_shortest_first_policy():
    """policy='shortest' with one resident slot admits the request with the
    least projected work first; results still equal the alone results."""
    print("serve shortest-first", flush=True)
    specs = [MIX[2], MIX[4], MIX[1]]
    from matvis_amd.serve import projected_work
    works = [projected_work(_kwargs(*s)) for s in specs]
    assert len(set(works)) == 3
    alone = [_serve([s])[1][0] for s in specs]
    srv = VisibilityServer(max_active=1, policy="shortest")
    reqs = [srv.submit(SimRequest(_kwargs(*s))) for s in specs]
    srv.run_until_idle(timeout=300)
    order = sorted(range(3), key=lambda i: reqs[i].t_admit)
    assert order == sorted(range(3), key=lambda i: works[i]), (order, works)
    assert order != [0, 1, 2]  # not the arrival order
    for r, a in zip(reqs, alone):
        assert r.error is None and np.array_equal(r.result, a)


@pytest.mark.parametrize("graphs", [True, False])
def test_multichunk_request_matches_simulate_multi(graphs):
    """A request cut into source chunks (min_chunks=3): the chunks after the
    first accumulate into the same FP64 accumulator inside the grouped
    launches, captured or not, alone and beside other requests; the result
    equals simulate_multi's."""
    import oracle as O
    print(f"serve multichunk graphs={graphs}", flush=True)
    specs = [(20, 2400, 2, 2, True, "GPUMatMul", 31),
             (20, 2400, 2, 2, True, "GPUMatMulSplitBF16", 32),
             (12, 900, 1, 2, False, "GPUMatMul", 33)]
    kws = [dict(_kwargs(*s), min_chunks=3) for s in specs]
    srv = VisibilityServer(graphs=graphs)
    reqs = [srv.submit(SimRequest(k)) for k in kws]
    srv.run_until_idle(timeout=300)
    for s, k, r in zip(specs, kws, reqs):
        assert r.error is None, r.error
        want = simulate_multi(**k)
        nf = 2 if s[4] else 1
        bound = 1e-12 if s[5] == "GPUMatMul" else 1e-6
        for fi in range(want.shape[0]):
            a, b = r.result[fi], want[fi]
            if not s[4]:
                a, b = a[..., None, None], b[..., None, None]
            e = O.normalized_error(a, b, s[0], nf)[0]
            assert e < bound, (s, fi, e)
