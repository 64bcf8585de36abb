"""mv_herk_noalloc (the HERK with its split-K partials in the library's own
device scratch array) against mv_herk with a caller workspace, through the
C ABI: two calls in a row, the second with a smaller problem, for the
split-BF16 and complex128 forms; sizes beyond the scratch are rejected.

Bounds: complex128 1e-12 normalised (both FP64; the slice counts, and so
the summation order, may differ); split-BF16 1e-6 normalised (FP32 chains
over different K-slices)."""

import numpy as np
import pytest
import torch

import matvis_amd._device as D
from matvis_amd._lib import FORM_C128, FORM_SPLIT_BF16, HipError, call, lib
from matvis_amd.gpu.herk import Herk, split_terms_alloc

pytestmark = pytest.mark.gpu


def _z(m, nseg, nup, nvalid, seed, dtype):
    """Z with nup columns a segment, zero past nvalid in each."""
    rng = np.random.default_rng(seed)
    seg = -(-nup // 32) * 32
    amp = 10.0 ** rng.uniform(-2, 0, (1, nseg * seg))
    z = amp * np.exp(2j * np.pi * rng.uniform(size=(m, nseg * seg)))
    for s in range(nseg):
        z[:, s * seg + nvalid:(s + 1) * seg] = 0
    return D.to_dev(z.astype(dtype)), seg


def _terms(z, m, nseg, nup, seg):
    ts = split_terms_alloc(2, m, nseg * seg)  # K-tile-major, m rows a tile
    call("mv_split_z", D.ptr(z), 0, m, nseg, nup, nseg * seg, seg,
         FORM_SPLIT_BF16, 2, 1.0, D.ptr(ts[0]), D.ptr(ts[1]), None, m,
         seg, D.stream_ptr())
    return ts


def _nerr(got, ref, m):
    ref = torch.triu(ref[:m, :m])
    got = torch.triu(got[:m, :m])
    d = torch.sqrt(torch.diagonal(ref).real.abs())
    d[d == 0] = 1
    return float(torch.triu(torch.abs(got - ref) / (d[:, None] * d[None])).max())


def _both(form, terms, m, nseg, seg, nup, ndev=None):
    """(mv_herk result, mv_herk_noalloc result) on the same terms."""
    nt, npd = (2, 3) if form == FORM_SPLIT_BF16 else (1, 1)
    h = Herk(form, m, nt, npd)
    dev = D.to_dev(np.array([ndev], np.int32)) if ndev is not None else None
    p = [D.ptr(t) for t in terms] + [None] * (3 - len(terms))
    ldz = m if form == FORM_SPLIT_BF16 else nseg * seg
    want = h.run(terms, ldz=ldz, nseg=nseg, seg_stride=seg,
                 seg_len=nup, seg_len_dev=dev, vacc=h.new_vacc(),
                 accumulate=False)
    got = h.new_vacc()
    call("mv_herk_noalloc", form, nt, npd, p[0], p[1], p[2], m, ldz,
         nseg, seg, nup, D.ptr(dev), 1.0, D.ptr(got), h.ldv, 0,
         D.stream_ptr())
    torch.cuda.synchronize()
    return want, got


@pytest.mark.parametrize("form", [FORM_SPLIT_BF16, FORM_C128])
def test_noalloc_matches_workspace_entry(form):
    bound = 1e-6 if form == FORM_SPLIT_BF16 else 1e-12
    # a large call, then a smaller one with a device column count
    for (m, nseg, nup, ndev, seed) in [(700, 2, 30000, None, 1),
                                       (130, 2, 3000, 2500, 2)]:
        print(f"noalloc form={form} m={m} nseg={nseg} nup={nup} ndev={ndev}",
              flush=True)
        ct = np.complex64 if form == FORM_SPLIT_BF16 else np.complex128
        z, seg = _z(m, nseg, nup, nup if ndev is None else ndev, seed, ct)
        terms = _terms(z, m, nseg, nup, seg) if form == FORM_SPLIT_BF16 else [z]
        want, got = _both(form, terms, m, nseg, seg, nup, ndev)
        e = _nerr(got, want, m)
        print(f"  noalloc vs workspace entry: {e:.2e}", flush=True)
        assert torch.isfinite(torch.triu(got[:m, :m]).abs()).all()
        assert e < bound, (m, e)


def test_noalloc_rejects_sizes_beyond_scratch():
    """Split partials of one slice at m = 5760 exceed the 128 MiB array:
    a clear error, before any device work."""
    z = D.zeros((1, 32), np.int32)
    v = D.zeros((1, 1), np.complex128)
    with pytest.raises(HipError, match="scratch"):
        call("mv_herk_noalloc", FORM_SPLIT_BF16, 2, 3, D.ptr(z), D.ptr(z),
             None, 5760, 32, 1, 32, 16, None, 1.0, D.ptr(v), 5760, 0,
             D.stream_ptr())
    assert lib().mv_herk_noalloc is not None
