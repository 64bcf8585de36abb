"""The batched-instance and serving workloads at their documented sizes,
against references that do not share the code under test.

Batched instances: 256 (time, frequency) HERK instances in grouped
persistent launches, tiles and K-slice lengths chosen by the autotuner,
captured in a HIP graph and replayed: nant from {64, 128, 350},
unpolarized / polarized 50/50, n_up
log-uniform on [5k, 80k] columns per E-field axis, complex64 Z split into two
bf16 terms (3 products).  Every instance is checked against one plain FP64
HERK launch per instance on the same Z promoted to complex128 (the 256
sequential plain launches), and a seeded sample of 8 instances, the largest
and the smallest included, against numpy's FP64 ``conj(Z) Z^T``.  Bound:
1e-6 normalised (|V - V_ref| / sqrt(V_ii V_jj)).

Serving: requests drawn from the Poisson workload's distribution (nant from
{32, 128, 350, 512}, nsrc log-uniform on [1e3, 2e5], polarized with p = 0.5,
plain FP64 or split-BF16) plus its largest shape (512 antennas, 2e5 sources,
polarized) in both forms.  The number of frequencies is capped at 3 and of
times at 2 (the workload draws up to 64 and 16; instances are independent,
so the cap shortens the run without changing what an instance computes).
Checked: every result bit-identical alone and in the mix; Hermitian output
(exactly, off the self-correlations), real positive autocorrelations; 4x the
fluxes gives exactly 4x the visibilities at the largest shape; the small
requests against the CPU oracle (FP64 <= 1e-12; split-BF16 <= 1.6e-5, twice
the 2^-17 relative representation bound of a two-term bf16 split, which a
sky of a few thousand power-law sources does not average down).
"""

import numpy as np
import pytest
import torch

import matvis_amd._device as D
import oracle as O
from matvis_amd import synth
from matvis_amd._lib import FORM_C128, FORM_SPLIT_BF16, call
from matvis_amd.gpu.herk import split_terms_alloc

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------ batched

def _batch(ninst=256, seed=7):
    """Instances of the batched workload: complex64 Z with per-column
    amplitudes log-uniform over two decades and uniform phases, columns past
    n_up zero in every segment, split into two bf16 terms."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    probs, zs = [], []
    for _ in range(ninst):
        nant = int(rng.choice([64, 128, 350]))
        nf = 2 if rng.random() < 0.5 else 1
        nup = int(np.exp(rng.uniform(np.log(5000), np.log(80000))))
        m, seg = nant * nf, -(-nup // 16) * 16
        amp = 10.0 ** (-2.0 * torch.rand((1, nf * seg), generator=gen,
                                         device="cuda", dtype=torch.float64))
        ph = 2 * np.pi * torch.rand((m, nf * seg), generator=gen,
                                    device="cuda", dtype=torch.float64)
        z = torch.polar(amp.expand(m, -1), ph).to(torch.complex64)
        z.view(m, nf, seg)[:, :, nup:] = 0
        terms = split_terms_alloc(2, m, nf * seg)  # m rows a K tile
        call("mv_split_z", D.ptr(z), 0, m, nf, nup, nf * seg, seg,
             FORM_SPLIT_BF16, 2, 1.0, D.ptr(terms[0]), D.ptr(terms[1]), None,
             m, seg, D.stream_ptr())
        ldv = -(-m // 128) * 128
        probs.append(dict(terms=terms, m=m, ldz=m, nseg=nf,
                          seg_stride=seg, seg_len=nup,
                          vacc=D.zeros((ldv, ldv), np.complex128), ldv=ldv,
                          scale=1.0, accumulate=False))
        zs.append(z)
    torch.cuda.synchronize()
    return probs, zs


def _upper_nerr(got, ref, m):
    """Normalised max error over the upper triangle (r <= c < m)."""
    ref = torch.triu(ref[:m, :m])
    got = torch.triu(got[:m, :m])
    d = torch.sqrt(torch.diagonal(ref).real.abs())
    rel = torch.triu(torch.abs(got - ref) / (d[:, None] * d[None]))
    return float(rel.max())


def test_batched_workload_against_fp64():
    from matvis_amd.gpu import autotune
    from matvis_amd.gpu.herk import Herk
    from matvis_amd.gpu.herk_group import GroupedBatch

    print("batched workload: 256 instances", flush=True)
    probs, zs = _batch()
    plan = autotune.tune(FORM_SPLIT_BF16, 2, 3, probs, budget_s=5.0)
    print(f"  tuned: {plan.describe()}", flush=True)
    batch = GroupedBatch(FORM_SPLIT_BF16, 2, 3, probs, plan.tiles,
                         plan.kchains)
    # the autotuner's timing runs wrote other tile assignments into the
    # accumulators (past m too): start from zeros, as the replays do
    for p in probs:
        p["vacc"].zero_()
    batch.capture()  # runs the batch once, then records it
    torch.cuda.synchronize()
    first = [p["vacc"].clone() for p in probs]
    for rep in range(2):
        for p in probs:
            p["vacc"].zero_()
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        for p, w in zip(probs, first):
            assert torch.equal(p["vacc"], w), f"replay {rep} differs"
    # every instance against one plain FP64 launch on the same Z
    worst = 0.0
    for i, (p, z) in enumerate(zip(probs, zs)):
        h = Herk(FORM_C128, p["m"])
        v = h.new_vacc()
        h.run([z.to(torch.complex128)], ldz=p["nseg"] * p["seg_stride"],
              nseg=p["nseg"],
              seg_stride=p["seg_stride"], seg_len=p["seg_len"], vacc=v,
              accumulate=False, nslices=h.nslices(p["nseg"] * p["seg_len"]))
        e = _upper_nerr(p["vacc"], v, p["m"])
        worst = max(worst, e)
        assert e < 1e-6, (i, p["m"], p["seg_len"], e)
    print(f"  max over 256 instances vs plain FP64 HERK: {worst:.2e}",
          flush=True)
    # a sample against numpy FP64, the largest and smallest included
    work = [p["m"] ** 2 * p["nseg"] * p["seg_len"] for p in probs]
    rng = np.random.default_rng(70)
    rest = [i for i in range(len(probs))
            if i not in (int(np.argmax(work)), int(np.argmin(work)))]
    sample = [int(np.argmax(work)), int(np.argmin(work))] + [
        int(i) for i in rng.choice(rest, 6, replace=False)]
    for i in sample:
        p, z = probs[i], zs[i]
        m = p["m"]
        zh = D.to_host(z).astype(np.complex128)
        ref = np.conj(zh) @ zh.T
        got = D.to_host(p["vacc"])[:m, :m]
        d = np.sqrt(np.abs(np.diag(ref)))
        iu = np.triu_indices(m)
        e = float((np.abs(got - ref) / (d[:, None] * d[None]))[iu].max())
        print(f"  instance {i}: m={m} K={p['nseg'] * p['seg_len']} "
              f"vs numpy FP64 {e:.2e}", flush=True)
        assert e < 1e-6, (i, e)


# ------------------------------------------------------------ serving

#: (nant, nsrc, nfreq, ntime, polarized, fp64, sky seed)
LARGEST = [(512, 200_000, 2, 1, True, True, 501),
           (512, 200_000, 2, 1, True, False, 502)]
#: small enough for the oracle
SMALL = [(32, 2500, 2, 2, True, True, 601), (32, 1800, 3, 1, False, True, 602),
         (32, 3000, 2, 2, True, False, 603), (128, 1200, 1, 2, False, False,
                                              604)]


def _drawn(n=5, seed=23):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        d = synth.serving_draw(rng)
        out.append((d["nant"], d["nsrc"], min(d["nfreq"], 3),
                    min(d["ntime"], 2), d["polarized"], d["fp64"], 700 + i))
    return out


@pytest.fixture(scope="module")
def serving_runs():
    from matvis_amd.serve import SimRequest, VisibilityServer

    sky = synth.ServingSky()
    specs = LARGEST + _drawn() + SMALL
    kws = [sky.request(*s[:6], seed=s[6]) for s in specs]
    # the largest shape again at 4x the fluxes, both forms
    kws4 = [sky.request(*s[:6], seed=s[6], flux_scale=4.0) for s in LARGEST]

    def serve(kwargs_list):
        srv = VisibilityServer()
        reqs = [srv.submit(SimRequest(k)) for k in kwargs_list]
        srv.run_until_idle(timeout=600)
        for r in reqs:
            assert r.error is None, r.error
        return [r.result for r in reqs]

    for s in specs:
        print(f"  request {s}", flush=True)
    print("serving workload: mix", flush=True)
    mixed = serve(kws + kws4)
    print("serving workload: alone", flush=True)
    alone = [serve([k])[0] for k in kws]
    return sky, specs, kws, mixed[:len(kws)], mixed[len(kws):], alone


def test_serving_alone_equals_mix(serving_runs):
    _, specs, _, mixed, _, alone = serving_runs
    for s, a, b in zip(specs, alone, mixed):
        assert np.array_equal(a, b), s


def test_serving_hermitian_and_real_autos(serving_runs):
    _, specs, _, mixed, _, _ = serving_runs
    for s, v in zip(specs, mixed):
        nant, nfreq, ntime, pol, fp64 = s[0], s[2], s[3], s[4], s[5]
        nf = 2 if pol else 1
        w = v.reshape(nfreq, ntime, nant, nant, nf, nf)
        herm = np.conj(w.transpose(0, 1, 3, 2, 5, 4))
        off = np.ones(w.shape, bool)
        idx = np.arange(nant)
        for p in range(nf):
            off[:, :, idx, idx, p, p] = False
        np.testing.assert_array_equal(w[off], herm[off], err_msg=str(s))
        autos = np.diagonal(w[:, :, idx, idx], axis1=-2, axis2=-1)
        assert np.all(autos.real > 0), s
        if fp64:
            assert np.all(np.abs(autos.imag) <= 1e-13 * autos.real), s
        else:
            assert np.all(autos.imag == 0), s


def test_serving_flux_linear_at_largest_shape(serving_runs):
    _, specs, _, mixed, mixed4, _ = serving_runs
    for i, s in enumerate(LARGEST):
        j = specs.index(s)
        # sqrt(4 I / 2) = 2 sqrt(I / 2) exactly, and the 16-bit split of 2Z
        # is twice that of Z: V scales by exactly 4
        np.testing.assert_array_equal(mixed4[i], 4.0 * mixed[j],
                                      err_msg=str(s))


def test_serving_small_requests_match_oracle(serving_runs):
    sky, specs, kws, mixed, _, _ = serving_runs
    for s in SMALL:
        j = specs.index(s)
        nant, nsrc, nfreq, ntime, pol, fp64, seed = s
        nf = 2 if pol else 1
        kw = kws[j]
        ra, dec = synth.random_sky(nsrc, seed, -90.0, 30.0)
        eci = O.point_source_crd_eq(ra, dec)
        b = sky.beam
        worst = 0.0
        for fi, freq in enumerate(kw["freqs"]):
            planes, daz, dza, azmin = O.prepare_grid(
                b.data_array, b.axis1_array, b.axis2_array, b.freq_array,
                freq, pol, b.beam_type)
            beam = {"kind": "grid", "planes": planes[None], "daz": [daz],
                    "dza": [dza], "azmin": [azmin], "order": 3}
            ref = O.simulate(antpos=kw["antpos"], freq=freq, eci=eci,
                             rotations=kw["coord_method_params"]["rotations"],
                             I_sky=kw["fluxes"][:, fi], beam=beam,
                             polarized=pol)
            if not pol:
                ref = ref[..., None, None]
            worst = max(worst, O.normalized_error(mixed[j][fi], ref, nant,
                                                  nf)[0])
        bound = 1e-12 if fp64 else 1.6e-5
        print(f"  {s}: vs oracle {worst:.2e} (bound {bound:g})", flush=True)
        assert worst < bound, (s, worst)
