"""Grouped HERK (mv_herk_grouped), split and plain forms, against
single-problem launches, and its HIP-graph capture replayed several times."""

import numpy as np
import pytest
import torch

import matvis_amd._device as D
from matvis_amd._lib import FORM_SPLIT_BF16, FORM_SPLIT_F16, call

pytestmark = pytest.mark.gpu

# (m, columns per segment, segments, device column count or None)
CASES = [(64, 1000, 1, None), (128, 3000, 2, 2500), (200, 500, 1, None),
         (350, 4000, 2, None), (700, 2000, 2, 1900), (1, 16, 1, None),
         (65, 17, 1, 9), (130, 70, 2, None), (90, 48, 2, 0)]


def _problem(m, seg_len, nseg, n_dev, form, nterms, seed):
    rng = np.random.default_rng(seed)
    seg = -(-seg_len // 16) * 16
    amp = np.exp(rng.uniform(-2, 0, (1, nseg * seg)) * np.log(10))
    z = (amp * np.exp(2j * np.pi * rng.uniform(size=(m, nseg * seg))))
    z = z.astype(np.complex64)
    n = seg_len if n_dev is None else n_dev
    for s in range(nseg):  # zero beyond the column count of every segment
        z[:, s * seg + n:(s + 1) * seg] = 0
    from matvis_amd.gpu.matprod import f16_scale
    scale = f16_scale(float(np.abs(z).max())) if form == FORM_SPLIT_F16 else 1.0
    zd = D.to_dev(z)
    from matvis_amd.gpu.herk import split_terms_alloc
    terms = split_terms_alloc(nterms, m, nseg * seg)  # m rows a K tile
    t = terms + [None] * (3 - nterms)
    call("mv_split_z", D.ptr(zd), 0, m, nseg, seg_len, nseg * seg, seg, form,
         nterms, scale, D.ptr(t[0]), D.ptr(t[1]), D.ptr(t[2]), m,
         seg, D.stream_ptr())
    ldv = -(-m // 128) * 128
    return dict(terms=terms, m=m, ldz=m, nseg=nseg, seg_stride=seg,
                seg_len=seg_len,
                seg_len_dev=None if n_dev is None else D.to_dev(
                    np.array([n_dev], np.int32)),
                vacc=D.zeros((ldv, ldv), np.complex128), ldv=ldv,
                scale=1.0 / (scale * scale), accumulate=False, z=zd)


def _single(p, form, nterms, nprod):
    from matvis_amd.gpu.herk import Herk
    h = Herk(form, p["m"], nterms, nprod)
    v = h.new_vacc()
    h.run(p["terms"], ldz=p["ldz"], nseg=p["nseg"], seg_stride=p["seg_stride"],
          seg_len=p["seg_len"], seg_len_dev=p["seg_len_dev"], vacc=v,
          accumulate=False, scale=p["scale"],
          nslices=h.nslices(p["nseg"] * p["seg_len"]))
    return h, v


def _full(vacc, ldv, m, tile):
    out = D.zeros((m * m, 1, 1), np.complex128)
    call("mv_vis_expand", D.ptr(vacc), ldv, tile, m, 1, None, m * m,
         D.ptr(out), 1, D.stream_ptr())
    return D.to_host(out).reshape(m, m)


def _nerr(got, want):
    d = np.sqrt(np.abs(np.diag(want)))
    d[d == 0] = 1
    return (np.abs(got - want) / (d[:, None] * d[None])).max()


@pytest.mark.parametrize("tile", [64, 128])
@pytest.mark.parametrize("form,nterms,nprod", [(FORM_SPLIT_BF16, 2, 3),
                                               (FORM_SPLIT_F16, 2, 3),
                                               (FORM_SPLIT_BF16, 3, 6)])
def test_grouped_matches_single_launches(tile, form, nterms, nprod):
    from matvis_amd.gpu.herk_group import HerkGroup
    print(f"grouped tile={tile} form={form} ({nterms},{nprod})", flush=True)
    probs = [_problem(*c, form, nterms, seed=i) for i, c in enumerate(CASES)]
    g = HerkGroup(form, nterms, nprod, tile, probs)
    g.run()
    for p, c in zip(probs, CASES):
        h, v = _single(p, form, nterms, nprod)
        want = _full(v, h.ldv, p["m"], h.tile)
        got = _full(p["vacc"], p["ldv"], p["m"], tile)
        e = _nerr(got, want)
        print(f"  case {c}: grouped vs single {e:.2e}", flush=True)
        # same terms, FP32 accumulation over different K-slices
        assert e < 1e-6, (c, e)
        if c[3] == 0:  # no columns on the device: an exact zero
            assert not got.any() and not want.any(), c
        # the diagonal is summed exactly in both: bit-identical up to the
        # order of the slice sums
        np.testing.assert_allclose(np.diag(got).real, np.diag(want).real,
                                   rtol=1e-13, atol=0)


def _plain_problems(ct, seed):
    """Plain-form problems over CASES: complex Z of dtype ct, zero beyond
    each segment's column count."""
    from matvis_amd.gpu.herk import Herk
    from matvis_amd._lib import FORM_C64, FORM_C128
    f = FORM_C128 if ct == np.complex128 else FORM_C64
    probs = []
    for i, (m, seg_len, nseg, n_dev) in enumerate(CASES):
        rng = np.random.default_rng(seed + i)
        seg = -(-seg_len // 16) * 16
        z = (rng.normal(size=(m, nseg * seg))
             + 1j * rng.normal(size=(m, nseg * seg))).astype(ct)
        n = seg_len if n_dev is None else n_dev
        for s_ in range(nseg):
            z[:, s_ * seg + n:(s_ + 1) * seg] = 0
        h = Herk(f, m)
        probs.append(dict(terms=[D.to_dev(z)], m=m, ldz=nseg * seg,
                          nseg=nseg, seg_stride=seg, seg_len=seg_len,
                          seg_len_dev=None if n_dev is None else D.to_dev(
                              np.array([n_dev], np.int32)),
                          vacc=h.new_vacc(), ldv=h.ldv, scale=1.0,
                          accumulate=False))
    return f, probs


@pytest.mark.parametrize("kind", ["split_bf16", "c128", "c64"])
def test_grouped_graph_replays(kind):
    from matvis_amd.gpu.herk_group import GroupedBatch, plain_tile
    print(f"grouped graph capture {kind}", flush=True)
    if kind == "split_bf16":
        form, nterms, nprod = FORM_SPLIT_BF16, 2, 3
        probs = [_problem(*c, form, nterms, seed=10 + i)
                 for i, c in enumerate(CASES)]
        tiles = [64 if p["m"] <= 128 else 128 for p in probs]
    else:
        nterms = nprod = 1
        form, probs = _plain_problems(
            np.complex128 if kind == "c128" else np.complex64, 50)
        tiles = [plain_tile()] * len(probs)
    b = GroupedBatch(form, nterms, nprod, probs, tiles)
    b.run()
    torch.cuda.synchronize()
    want = [p["vacc"].clone() for p in probs]
    b.capture()
    for rep in range(3):
        for p in probs:  # eager work between replays
            p["vacc"].zero_()
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for p, w in zip(probs, want):
            assert torch.equal(p["vacc"], w), f"replay {rep} differs"


@pytest.mark.parametrize("form", ["c128", "c64"])
def test_grouped_plain_matches_single_launches(form):
    """Plain forms in one grouped launch (64-row tiles) against one plain
    launch per problem: FP64 to rounding, complex64 to FP32 accumulation."""
    from matvis_amd.gpu.herk import Herk
    from matvis_amd.gpu.herk_group import HerkGroup, plain_tile

    print(f"grouped plain {form}", flush=True)
    f, probs = _plain_problems(
        np.complex128 if form == "c128" else np.complex64, 40)
    singles = []
    for p in probs:
        h = Herk(f, p["m"])
        v = h.new_vacc()
        h.run(p["terms"], ldz=p["ldz"], nseg=p["nseg"],
              seg_stride=p["seg_stride"], seg_len=p["seg_len"],
              seg_len_dev=p["seg_len_dev"], vacc=v, accumulate=False,
              nslices=h.nslices(p["nseg"] * p["seg_len"]))
        singles.append((h, v))
    g = HerkGroup(f, 1, 1, plain_tile(), probs)
    g.run()
    for p, (h, v), c in zip(probs, singles, CASES):
        want = _full(v, h.ldv, p["m"], h.tile)
        got = _full(p["vacc"], p["ldv"], p["m"], plain_tile())
        e = _nerr(got, want)
        print(f"  case {c}: grouped vs single {e:.2e}", flush=True)
        assert e < (1e-13 if form == "c128" else 2e-6), (c, e)
        if c[3] == 0:  # no columns on the device: an exact zero
            assert not got.any() and not want.any(), c


def _numpy_upper_nerr(got, z, m):
    """Normalised max error of the upper triangle of `got` against numpy's
    FP64 conj(Z) Z^T of the complex Z."""
    zh = D.to_host(z).astype(np.complex128)
    ref = np.conj(zh) @ zh.T
    d = np.sqrt(np.abs(np.diag(ref)))
    d[d == 0] = 1
    iu = np.triu_indices(m)
    return float((np.abs(got[:m, :m] - ref) / (d[:, None] * d[None]))[iu].max())


@pytest.mark.parametrize("kind", ["split_bf16", "c128"])
def test_grouped_accumulate_eager_and_replayed(kind):
    """accumulate=True adds into a pre-filled accumulator: the eager grouped
    launch, the graph capture's own run and every replay (the accumulator
    reset to the pre-filled values between them) give the pre-filled values
    plus the accumulate=False result, which matches numpy FP64."""
    from matvis_amd.gpu.herk_group import GroupedBatch, HerkGroup, plain_tile
    print(f"grouped accumulate {kind}", flush=True)
    if kind == "split_bf16":
        form, nterms, nprod = FORM_SPLIT_BF16, 2, 3
        probs = [_problem(*c, form, nterms, seed=80 + i)
                 for i, c in enumerate(CASES)]
        zs = [p["z"] for p in probs]
        tiles = [64 if p["m"] <= 128 else 128 for p in probs]
        # against numpy on the unsplit Z: the two-term bf16 representation
        # (2^-17 relative per element) is not averaged down over the few
        # hundred columns of these problems (the bound of the split-HERK
        # kernel tests)
        bound = 3e-5
    else:
        nterms = nprod = 1
        form, probs = _plain_problems(np.complex128, 90)
        zs = [p["terms"][0] for p in probs]
        tiles = [plain_tile()] * len(probs)
        bound = 1e-12
    # the accumulate=False result, per tile class as GroupedBatch runs it
    GroupedBatch(form, nterms, nprod, probs, tiles).run()
    torch.cuda.synchronize()
    base = [p["vacc"].clone() for p in probs]
    for p, b, z, c in zip(probs, base, zs, CASES):
        m, n = p["m"], (c[1] if c[3] is None else c[3])
        # numpy sees only the columns the device counted
        zc = z.view(m, p["nseg"], p["seg_stride"])[:, :, :n].reshape(m, -1)
        e = _numpy_upper_nerr(D.to_host(b), zc, m)
        assert e < bound, (c, e)
    rng = np.random.default_rng(5)
    pre = [D.to_dev(rng.normal(size=b.shape) + 1j * rng.normal(size=b.shape))
           for b in base]

    def reset():
        for p, q in zip(probs, pre):
            p["vacc"].copy_(q)
        torch.cuda.synchronize()

    def check(what):
        torch.cuda.synchronize()
        for p, b, q, c in zip(probs, base, pre, CASES):
            m = p["m"]
            got = D.to_host(p["vacc"])[:m, :m]
            want = (D.to_host(q) + D.to_host(b))[:m, :m]
            iu = np.triu_indices(m)
            # one rounding of the add (fused or not) per element
            np.testing.assert_allclose(got[iu], want[iu], rtol=1e-15,
                                       atol=1e-15 * np.abs(want).max(),
                                       err_msg=f"{what} {c}")

    for p in probs:
        p["accumulate"] = True
    reset()
    GroupedBatch(form, nterms, nprod, probs, tiles).run()
    check("eager")
    reset()
    batch = GroupedBatch(form, nterms, nprod, probs, tiles)
    batch.capture()  # one eager execution, then the recording
    check("capture")
    for rep in range(3):
        reset()
        batch.replay()
        check(f"replay {rep}")
    # a single HerkGroup launch of the same problems accumulates the same way
    if kind == "c128":
        reset()
        HerkGroup(form, 1, 1, plain_tile(), probs).run()
        check("group")


@pytest.mark.parametrize("kind", ["split_bf16", "c128"])
def test_grouped_kchain_matches_single_launches(kind):
    """Per-problem K-slice lengths (kchain): shorter slices change only the
    FP32 (split) or FP64 (plain) partial sums.  kchain 1 on the 700-row
    case runs 250 slices, more than mv_herk_max_slices gives (2 + 64): the
    capacity is raised for it."""
    from matvis_amd.gpu.herk_group import GroupedBatch, slice_capacity
    print(f"grouped kchain {kind}", flush=True)
    if kind == "split_bf16":
        form, nt, npd = FORM_SPLIT_BF16, 2, 3
        probs = [_problem(*c, form, nt, seed=40 + i)
                 for i, c in enumerate(CASES)]
        tiles, bound = [64 if p["m"] <= 128 else 128 for p in probs], 1e-6
    else:
        form, probs = _plain_problems(np.complex128, 60)
        nt = npd = 1
        tiles, bound = [64] * len(probs), 1e-12
    chains = [(0, 48, 1, 7)[i % 4] for i in range(len(probs))]
    big = CASES.index((700, 2000, 2, 1900))
    chains[big] = 1
    assert slice_capacity(form, 700, nt, npd, 2 * 2000, 1) == 250
    b = GroupedBatch(form, nt, npd, probs, tiles, chains)
    b.run()
    for p, c, k in zip(probs, CASES, chains):
        h, v = _single(p, form, nt, npd)
        want = _full(v, h.ldv, p["m"], h.tile)
        got = _full(p["vacc"], p["ldv"], p["m"], h.tile)
        e = _nerr(got, want)
        print(f"  case {c} kchain {k}: grouped vs single {e:.2e}", flush=True)
        assert e < bound, (c, k, e)


def test_autotuner_shortens_slices_of_an_underfilled_batch():
    """Two 128-row problems of 80k columns fill few of the GPU's workgroup
    slots at the default K-slice length (192 K tiles: 27 slices each); the
    tuner takes shorter slices for them, is no slower than the default,
    keeps the result within the split form's bound of a single launch, and
    answers a second call from its cache."""
    from matvis_amd.gpu import autotune
    from matvis_amd.gpu.herk_group import GroupedBatch
    print("autotuner: under-filled batch", flush=True)
    probs = [_problem(128, 80000, 1, None, FORM_SPLIT_BF16, 2, seed=90 + i)
             for i in range(2)]
    plan = autotune.tune(FORM_SPLIT_BF16, 2, 3, probs, budget_s=5.0)
    print(f"  {plan.describe()}", flush=True)
    assert not plan.cached
    assert all(k < 192 for k in plan.kchains), plan.describe()
    assert plan.best_ms <= plan.default_ms
    for p in probs:
        p["vacc"].zero_()
    GroupedBatch(FORM_SPLIT_BF16, 2, 3, probs, plan.tiles,
                 plan.kchains).run()
    for p, t in zip(probs, plan.tiles):
        h, v = _single(p, FORM_SPLIT_BF16, 2, 3)
        e = _nerr(_full(p["vacc"], p["ldv"], 128, t),
                  _full(v, h.ldv, 128, h.tile))
        assert e < 1e-6, e
    again = autotune.tune(FORM_SPLIT_BF16, 2, 3, probs, budget_s=5.0)
    assert again.cached and again.config == plan.config
