"""Kernel-level parity: each C-ABI entry point against the CPU oracle.

Tolerances: integer/index work bit-exact; FP64 (complex128) forms <= 1e-12
relative (normalised for V); FP32 forms <= 1e-5; split forms as stated per
test.
"""

import numpy as np
import pytest
import torch

import oracle as O
from matvis_amd import _device as D
from matvis_amd import astro, synth
from matvis_amd._lib import (FORM_C64, FORM_C128, FORM_SPLIT_BF16,
                             FORM_SPLIT_F16, call, lib)

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    return D.to_dev(np.asarray(a) if dtype is None else np.asarray(a, dtype))


def _rot(jd=2459863.2):
    return astro.eci_to_enu_rotation(jd, np.radians(-30.72), np.radians(21.43))


def _sky(n, seed):
    ra, dec = synth.random_sky(n, seed, -90, 90)
    return O.point_source_crd_eq(ra, dec)


# ------------------------------------------------------------------ (a)


@pytest.mark.parametrize("nsrc,chunk0,chunk_len", [
    (5000, 0, 5000), (5000, 1234, 2049), (70000, 0, 70000),
    (70000, 4097, 30000), (100, 100, 0), (3, 0, 3)])
def test_rotate_cull_bit_exact(nsrc, chunk0, chunk_len):
    print(f"rotate_cull nsrc={nsrc} chunk0={chunk0} len={chunk_len}", flush=True)
    eci = _sky(nsrc, 3)
    rot = _rot()
    flux = np.random.default_rng(1).uniform(0.1, 5, nsrc)
    topo = O.rotate(rot, eci)
    idx_o = chunk0 + np.where(topo[2, chunk0:chunk0 + chunk_len] > 0)[0]
    n_o = idx_o.size
    nalloc = max(chunk_len, 1) + 5
    eci_d = _dev(eci, np.float64)
    flux_d = _dev(flux, np.float64)
    crd = D.zeros((3, nalloc), np.float64) - 7.0
    fl = D.zeros((nalloc,), np.float64) - 7.0
    idx = D.zeros((nalloc,), np.int32) - 7
    n_up = D.zeros((1,), np.int32)
    ws = D.empty((lib().mv_rotate_cull_workspace_bytes(chunk_len),), np.uint8)
    call("mv_rotate_cull", D.ptr(eci_d), nsrc, rot.ctypes.data, D.ptr(flux_d),
         chunk0, chunk_len, nalloc, D.ptr(crd), nalloc, D.ptr(fl), 1,
         D.ptr(idx), D.ptr(n_up), D.ptr(ws), ws.numel(), D.stream_ptr())
    n = int(n_up.item())
    assert n == n_o
    c = D.to_host(crd)
    assert np.array_equal(c[:, :n], topo[:, idx_o])  # bit-exact rotation
    assert np.array_equal(D.to_host(idx)[:n], idx_o)
    assert np.array_equal(D.to_host(fl)[:n], flux[idx_o])
    assert np.all(c[:, n:] == 0) and np.all(D.to_host(fl)[n:] == 0)
    assert np.all(D.to_host(idx)[n:] == -1)


def test_rotate_cull_overflow_reports_true_count():
    nsrc = 4000
    eci = _sky(nsrc, 5)
    rot = _rot()
    n_true = int((O.rotate(rot, eci)[2] > 0).sum())
    nalloc = n_true // 2
    crd = D.zeros((3, nalloc), np.float32)
    n_up = D.zeros((1,), np.int32)
    ws = D.empty((lib().mv_rotate_cull_workspace_bytes(nsrc),), np.uint8)
    eci_d = _dev(eci, np.float64)
    call("mv_rotate_cull", D.ptr(eci_d), nsrc, rot.ctypes.data,
         None, 0, nsrc, nalloc, D.ptr(crd), nalloc, None, 0, None,
         D.ptr(n_up), D.ptr(ws), ws.numel(), D.stream_ptr())
    assert int(n_up.item()) == n_true


def test_rotate_all_bit_exact():
    eci = _sky(3000, 7)
    rot = _rot(2459863.7)
    out = D.zeros((3, 3000), np.float64)
    eci_d = _dev(eci, np.float64)
    call("mv_rotate", D.ptr(eci_d), 3000, rot.ctypes.data,
         3000, D.ptr(out), 1, 3000, D.stream_ptr())
    assert np.array_equal(D.to_host(out), O.rotate(rot, eci))


# ------------------------------------------------------------------ (b)


def _beam_case(nbeam, polarized, order, is_f64, seed=0):
    beams = [synth.harmonic_efield_beam(seed=seed + b, naz=72, nza=19,
                                        polarized=True, dtype=np.complex128)
             for b in range(nbeam)]
    from matvis_amd.beams import grid_planes
    preps = [grid_planes(b, 150e6, polarized) for b in beams]
    planes = np.stack([p[0] for p in preps])
    grid = np.array([[p[1], p[2], p[3]] for p in preps])
    # give beam 1 a shifted azimuth origin to exercise per-beam grids
    if nbeam > 1:
        grid[1, 2] = 0.01
    return planes, grid, bool(preps[0][4])


@pytest.mark.parametrize("order", [1, 3])
@pytest.mark.parametrize("polarized", [False, True])
@pytest.mark.parametrize("is_f64", [True, False])
def test_beam_interp_vs_oracle(order, polarized, is_f64):
    print(f"beam_interp order={order} pol={polarized} f64={is_f64}", flush=True)
    planes, grid, is_power = _beam_case(2, polarized, order, is_f64)
    nbeam, nax, nfeed, nza, naz = planes.shape
    eci = _sky(4000, 11)
    topo = O.rotate(_rot(), eci)
    nalloc = 2400
    crd, _, n, _ = O.select_chunk(topo, np.ones(4000), 0, 4000, nalloc)
    ctype = np.complex128 if is_f64 else np.complex64
    pl = _dev(planes.astype(ctype))
    if order == 3:
        call("mv_beam_prefilter", D.ptr(pl), int(is_f64), nbeam * nax * nfeed,
             nza, naz, D.stream_ptr())
    out = D.zeros((nbeam, nfeed, nax, nalloc), ctype) + 3.0
    n_dev = _dev(np.array([n], np.int32))  # positions n.. get zeros
    grid_d, crd_d = _dev(grid, np.float64), _dev(crd, np.float64)
    call("mv_beam_interp", D.ptr(pl), int(is_f64), nbeam, nfeed, nax, nza, naz,
         D.ptr(grid_d), order, int(is_power),
         D.ptr(crd_d), 1, nalloc, nalloc, D.ptr(n_dev),
         D.ptr(out), nalloc, D.stream_ptr())
    got = D.to_host(out)
    want = O.interp_grid_beam(planes, grid[:, 0], grid[:, 1], grid[:, 2],
                              *O.enu_to_az_za(crd[0, :n], crd[1, :n], "uvbeam"),
                              order=order)
    tol = 1e-12 if is_f64 else 2e-5
    np.testing.assert_allclose(got[..., :n], want, rtol=0,
                               atol=tol * np.abs(want).max())
    assert np.all(got[..., n:] == 0)


def test_gpu_beam_interpolation_reference_cases():
    """tests/test_beam_interp_gpu.py:12-87 on the device."""
    from matvis_amd.gpu.beams import gpu_beam_interpolation

    za = np.linspace(0, 1, 3)
    az = np.linspace(0, 2 * np.pi, 6)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA**2, 2 - ZA**2 + np.sin(AZ)])[:, None, None]
    new = D.to_host(gpu_beam_interpolation(
        beam, [az[1] - az[0]] * 2, [za[1] - za[0]] * 2, [0.0] * 2,
        AZ.flatten(), ZA.flatten()))
    for b in range(2):
        np.testing.assert_allclose(new[b, 0, 0], np.sqrt(beam[b, 0, 0].ravel()))
    za = np.linspace(0, 1, 5)
    az = np.linspace(0, 1, 5)
    AZ, ZA = np.meshgrid(az, za, indexing="xy")
    beam = np.array([1 - ZA])[:, None, None]
    new = D.to_host(gpu_beam_interpolation(
        beam[..., ::2, ::2], [2 * (az[1] - az[0])], [2 * (za[1] - za[0])],
        [0.0], AZ.flatten(), ZA.flatten()))
    assert np.allclose(np.sqrt(beam[0, 0, 0].ravel()), new[0, 0, 0])
    with pytest.raises(ValueError, match="as the dtype for beam, which is unrecognized"):
        gpu_beam_interpolation(beam.astype(int), 1.0, 1.0, 0.0, AZ.flatten(),
                               ZA.flatten())


@pytest.mark.parametrize("polarized", [False, True])
def test_gaussian_beam_vs_oracle(polarized):
    eci = _sky(3000, 13)
    crd, _, n, _ = O.select_chunk(O.rotate(_rot(), eci), np.ones(3000), 0,
                                  3000, 1800)
    sig = np.radians([10.0, 25.0])
    nf = 2 if polarized else 1
    out = D.zeros((2, nf, nf, 1800), np.complex128)
    sig_d, crd_d = _dev(sig), _dev(crd, np.float64)
    call("mv_beam_gaussian", D.ptr(sig_d), 2, int(polarized),
         D.ptr(crd_d), 1, 1800, 1800, None, D.ptr(out), 1,
         1800, D.stream_ptr())
    want = O.gaussian_beam(sig, O.enu_to_az_za(crd[0], crd[1], "uvbeam")[1],
                           polarized)
    np.testing.assert_allclose(D.to_host(out), want, rtol=1e-13, atol=1e-15)


# ------------------------------------------------------------------ (c)


@pytest.mark.parametrize("is_f64", [True, False])
def test_exptau_vs_oracle(is_f64):
    antpos = synth.hex_array(37, 14.6, 0.02, 1) * 8  # up to ~600 m baselines
    eci = _sky(2000, 17)
    crd, _, n, _ = O.select_chunk(O.rotate(_rot(), eci), np.ones(2000), 0,
                                  2000, 1300)
    ctype = np.complex128 if is_f64 else np.complex64
    out = D.zeros((37, 1300), ctype)
    ant_d, crd_d = _dev(antpos, np.float64), _dev(crd, np.float64)
    call("mv_exptau", D.ptr(ant_d), 37, 200e6,
         D.ptr(crd_d), 1, 1300, 1300, D.ptr(out), int(is_f64),
         1300, D.stream_ptr())
    want = O.exptau(antpos, 200e6, crd)
    err = np.abs(D.to_host(out) - want).max()
    assert err < (1e-12 if is_f64 else 2e-6), err


def test_zmat_parts_plugin_vs_oracle():
    from matvis_amd.core.getz import ZMatrixCalc

    rng = np.random.default_rng(3)
    nant, nsrc = 7, 333
    et = np.exp(1j * rng.uniform(0, 6, (nant, nsrc)))
    fl = rng.uniform(0.1, 2, nsrc)
    beam = rng.normal(size=(3, 2, 2, nsrc)) + 1j * rng.normal(size=(3, 2, 2, nsrc))
    bidx = np.array([0, 2, 1, 0, 1, 2, 2])
    zc = ZMatrixCalc(nant=nant, nfeed=2, nax=2, nsrc=nsrc, ctype=np.complex128)
    zc.setup()
    et_d = _dev(et)
    z = zc(_dev(fl), _dev(beam), et_d, bidx)
    want = O.zmatrix(fl, beam, et, bidx, 2, 2)
    np.testing.assert_allclose(D.to_host(z), want, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(D.to_host(et_d), et * fl, rtol=1e-15)


@pytest.mark.parametrize("polarized", [False, True])
def test_zmat_fused_c128_vs_oracle(polarized):
    nant, nsrc = 9, 1500
    antpos = synth.disk_array(nant, 200.0, 4)
    eci = _sky(nsrc, 19)
    flux = np.random.default_rng(2).uniform(0.1, 3, nsrc)
    crd, fl, n, idx = O.select_chunk(O.rotate(_rot(), eci), np.sqrt(0.5 * flux),
                                     0, nsrc, 900)
    assert n < 900
    nf = 2 if polarized else 1
    rng = np.random.default_rng(4)
    beam = rng.normal(size=(2, nf, nf, 900)) + 1j * rng.normal(size=(2, nf, nf, 900))
    bidx = (np.arange(nant) % 2).astype(np.int32)
    seg = 928
    z = D.zeros((nant * nf, nf * seg), np.complex128) + 5.0
    # every buffer stays referenced until the call returns
    bufs = dict(
        ant=_dev(antpos, np.float64),
        crd=_dev(np.pad(crd, ((0, 0), (0, seg - 900))), np.float64),
        flux=_dev(np.sqrt(0.5 * flux), np.float64),
        idx=_dev(np.pad(idx, (0, seg - n)).astype(np.int32)),
        beam=_dev(np.pad(beam, ((0, 0),) * 3 + ((0, seg - 900),))),
        bidx=_dev(bidx), n=_dev(np.array([n], np.int32)))
    call("mv_zmat_fused", D.ptr(bufs["ant"]), nant, 150e6,
         D.ptr(bufs["crd"]), seg, D.ptr(bufs["flux"]), D.ptr(bufs["idx"]),
         D.ptr(bufs["beam"]), 1, seg,
         2, D.ptr(bufs["bidx"]), nf, nf, 900, D.ptr(bufs["n"]),
         FORM_C128, 1, 1.0, D.ptr(z), None, None, nf * seg, seg,
         D.stream_ptr())
    got = D.to_host(z).reshape(nant * nf, nf, seg)
    want = O.zmatrix(fl, beam, O.exptau(antpos, 150e6, crd), bidx, nf, nf)
    want = want.reshape(nant * nf, nf, 900)
    np.testing.assert_allclose(got[..., :n], want[..., :n], rtol=0, atol=1e-12)
    npad = -(-n // 32) * 32
    assert np.all(got[..., n:npad] == 0)
    assert np.all(got[..., npad:] == 5.0)  # untouched past the pad


# ------------------------------------------------------------------ (d)


def _herk_case(m, k, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))).astype(dtype)


def _run_herk(form, zs, m, ldz, nseg, seg_stride, seg_len, nterms=1, nprod=1,
              seg_len_dev=None, nslices=3, scale=1.0):
    from matvis_amd.gpu.herk import Herk
    h = Herk(form, m, nterms, nprod)
    v = h.new_vacc()
    h.run(zs, ldz=ldz, nseg=nseg, seg_stride=seg_stride, seg_len=seg_len,
          vacc=v, accumulate=False, seg_len_dev=seg_len_dev, scale=scale,
          nslices=nslices)
    return h, v


def _vis_full(h, v, m):
    out = D.zeros((m * m, 1, 1), np.complex128)
    call("mv_vis_expand", D.ptr(v), h.ldv, h.tile, m, 1, None, m * m,
         D.ptr(out), 1, D.stream_ptr())
    return D.to_host(out).reshape(m, m)


@pytest.mark.parametrize("m,k", [(1, 5), (5, 0), (63, 17), (64, 16), (65, 33),
                                 (130, 1000), (700, 300)])
@pytest.mark.parametrize("form", [FORM_C128, FORM_C64])
def test_herk_plain_vs_numpy(m, k, form):
    print(f"herk plain form={form} m={m} k={k}", flush=True)
    dtype = np.complex128 if form == FORM_C128 else np.complex64
    z = _herk_case(m, k, dtype, seed=m + k)
    zd = _dev(z) if k else D.zeros((m, 1), dtype)
    h, v = _run_herk(form, [zd], m, max(k, 1), 1, max(k, 1), k)
    got = _vis_full(h, v, m)
    want = z.astype(np.complex128).conj() @ z.astype(np.complex128).T
    scale = np.sqrt(np.abs(np.diag(want))[:, None] * np.abs(np.diag(want))[None])
    scale[scale == 0] = 1
    err = (np.abs(got - want) / scale).max() if k else np.abs(got).max()
    assert err < (1e-13 if form == FORM_C128 else 2e-6), err


def test_herk_segments_and_device_length():
    m, seg, n = 70, 48, 37
    z = _herk_case(m, 2 * seg, np.complex128, seed=9)
    z_valid = np.concatenate([z[:, :n], z[:, seg:seg + n]], 1)
    h, v = _run_herk(FORM_C128, [_dev(z)], m, 2 * seg, 2, seg, seg,
                     seg_len_dev=_dev(np.array([n], np.int32)), nslices=5)
    want = z_valid.conj() @ z_valid.T
    np.testing.assert_allclose(_vis_full(h, v, m), want, rtol=1e-13,
                               atol=1e-13 * np.abs(want).max())
    # a device count above the host bound is clamped to it
    h, v = _run_herk(FORM_C128, [_dev(z)], m, 2 * seg, 2, seg, n,
                     seg_len_dev=_dev(np.array([10 * seg], np.int32)))
    np.testing.assert_allclose(_vis_full(h, v, m), want, rtol=1e-13,
                               atol=1e-13 * np.abs(want).max())


def _split_terms(z, form, nterms, scale=1.0):
    """Split terms of z (K-tile-major, ldt = m rows a K tile) and their
    column count, zero-padded to whole 16-column K tiles."""
    from matvis_amd.gpu.herk import split_terms_alloc
    m, k = z.shape
    ncols = max(16, -(-k // 16) * 16)
    ts = split_terms_alloc(nterms, m, ncols)
    t = ts + [None] * (3 - nterms)
    zd = _dev(z)
    call("mv_split_z", D.ptr(zd), int(z.dtype == np.complex128), m, 1, k,
         k, k, form, nterms, scale, D.ptr(t[0]), D.ptr(t[1]), D.ptr(t[2]), m,
         ncols, D.stream_ptr())
    return ts, ncols


def _row_major(t):
    """Words of a K-tile-major term array (ntiles, rows, 16) as the
    row-major (rows, ntiles * 16) matrix they store."""
    w = D.to_host(t).view(np.uint32)
    return w.transpose(1, 0, 2).reshape(w.shape[1], -1)


def _np_term_words(re, im, f16, nterms):
    """The term arrays the HERK reads, restated in numpy from float32 real
    and imaginary parts (rows x columns): each term the round-to-nearest-even
    16-bit value of the remainder the previous terms left; per 4-column
    chunk the 16-bit real parts of the 4 columns (two a word, the lower
    column in the low half), then their imaginary parts; columns past the
    last one up to a multiple of 4 are zeros."""
    def split(x):
        out, r = [], x.astype(np.float32)
        for _ in range(nterms):
            if f16:
                h = r.astype(np.float16)
                b = h.view(np.uint16).astype(np.uint32)
                v = h.astype(np.float32)
            else:
                u = r.view(np.uint32).astype(np.uint64)
                b = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(
                    np.uint32)
                v = (b << np.uint32(16)).view(np.float32)
            out.append(b)
            r = (r - v).astype(np.float32)
        return out

    m, k = re.shape
    k4 = -(-k // 4) * 4
    words = []
    for tr, ti in zip(split(re), split(im)):
        R = np.pad(tr, ((0, 0), (0, k4 - k))).reshape(m, k4 // 4, 4)
        I = np.pad(ti, ((0, 0), (0, k4 - k))).reshape(m, k4 // 4, 4)
        W = np.stack([R[..., 0] | R[..., 1] << 16, R[..., 2] | R[..., 3] << 16,
                      I[..., 0] | I[..., 1] << 16, I[..., 2] | I[..., 3] << 16],
                     -1)
        words.append(W.reshape(m, k4))
    return words


@pytest.mark.parametrize("form,nterms", [(FORM_SPLIT_BF16, 2),
                                         (FORM_SPLIT_BF16, 3),
                                         (FORM_SPLIT_F16, 2),
                                         (FORM_SPLIT_F16, 3)])
@pytest.mark.parametrize("zdtype", [np.complex64, np.complex128])
def test_split_z_terms_bit_exact(form, nterms, zdtype):
    """mv_split_z's term words equal the numpy restatement bit for bit, on
    values spanning 8 decades with exact zeros and both signs, 77 columns
    (a partial last chunk) and a scale."""
    rng = np.random.default_rng(nterms + (form == FORM_SPLIT_F16))
    m, k = 5, 77
    mag = np.exp(rng.uniform(-8, 0, (m, k)) * np.log(10))
    z = (mag * np.exp(2j * np.pi * rng.uniform(size=(m, k)))).astype(zdtype)
    z[1, 3] = 0
    z[2, 10] = -0.75 + 0.0j
    scale = 0.5
    ts, ncols = _split_terms(z, form, nterms, scale)
    re = z.real.astype(np.float32) * np.float32(scale)
    im = z.imag.astype(np.float32) * np.float32(scale)
    want = _np_term_words(re, im, form == FORM_SPLIT_F16, nterms)
    k4 = want[0].shape[1]
    for t, w in zip(ts, want):
        got = _row_major(t)
        assert np.array_equal(got[:, :k4], w)
        assert np.all(got[:, k4:] == 0)


@pytest.mark.parametrize("polarized", [False, True])
def test_zmat_fused_split_is_the_split_of_its_complex64(polarized):
    """The fused kernel's split forms write exactly the terms of the
    complex64 Z its complex64 form writes for the same inputs (the same
    FP32 arithmetic up to the store), in the restated layout; padding
    columns are zero terms and nothing past them is written."""
    nant, nsrc, seg = 9, 1500, 928
    antpos = synth.disk_array(nant, 200.0, 4)
    eci = _sky(nsrc, 19)
    flux = np.random.default_rng(2).uniform(0.1, 3, nsrc)
    crd, fl, n, idx = O.select_chunk(O.rotate(_rot(), eci), np.sqrt(0.5 * flux),
                                     0, nsrc, 900)
    nf = 2 if polarized else 1
    rng = np.random.default_rng(4)
    beam = (rng.normal(size=(2, nf, nf, 900))
            + 1j * rng.normal(size=(2, nf, nf, 900))).astype(np.complex64)
    bufs = dict(
        ant=_dev(antpos, np.float64),
        crd=_dev(np.pad(crd, ((0, 0), (0, seg - 900))), np.float64),
        flux=_dev(np.sqrt(0.5 * flux), np.float64),
        idx=_dev(np.pad(idx, (0, seg - n)).astype(np.int32)),
        beam=_dev(np.pad(beam, ((0, 0),) * 3 + ((0, seg - 900),))),
        bidx=_dev((np.arange(nant) % 2).astype(np.int32)),
        n=_dev(np.array([n], np.int32)))

    def fused(form, nterms, outs, ldz):
        zp = [D.ptr(o) for o in outs] + [None] * (3 - len(outs))
        call("mv_zmat_fused", D.ptr(bufs["ant"]), nant, 150e6,
             D.ptr(bufs["crd"]), seg, D.ptr(bufs["flux"]), D.ptr(bufs["idx"]),
             D.ptr(bufs["beam"]), 0, seg, 2, D.ptr(bufs["bidx"]), nf, nf, 900,
             D.ptr(bufs["n"]), form, nterms, 1.0, zp[0], zp[1], zp[2],
             ldz, seg, D.stream_ptr())

    z64 = D.zeros((nant * nf, nf * seg), np.complex64)
    fused(FORM_C64, 1, [z64], nf * seg)
    z = D.to_host(z64)
    npad = -(-n // 32) * 32
    for form, nterms in [(FORM_SPLIT_BF16, 2), (FORM_SPLIT_BF16, 3),
                         (FORM_SPLIT_F16, 2)]:
        # K-tile-major terms, nant * nf rows a K tile
        outs = [D.zeros((nf * seg // 16, nant * nf, 16), np.int32)
                + 0x5A5A5A5A for _ in range(nterms)]
        fused(form, nterms, outs, nant * nf)
        for ax in range(nf):
            cols = slice(ax * seg, ax * seg + npad)
            want = _np_term_words(z[:, cols].real, z[:, cols].imag,
                                  form == FORM_SPLIT_F16, nterms)
            for o, w in zip(outs, want):
                got = _row_major(o)
                assert np.array_equal(got[:, cols], w), (form, nterms, ax)
                assert np.all(got[:, ax * seg + npad:(ax + 1) * seg]
                              == 0x5A5A5A5A)


# Bounds: the split's representation error with the exact diagonal (bf16
# two terms up to 3.4e-6 at k = 77, 3 or 4 products alike; 3 terms ~5e-9;
# fp16 ~6e-8 on these inputs) plus FP32 accumulation within one K-slice.
@pytest.mark.parametrize("form,nterms,nprod,bound", [
    (FORM_SPLIT_BF16, 2, 3, 6e-6), (FORM_SPLIT_BF16, 2, 4, 6e-6),
    (FORM_SPLIT_BF16, 3, 6, 1e-6), (FORM_SPLIT_F16, 2, 3, 1e-6),
    (FORM_SPLIT_F16, 2, 4, 1e-6)])
@pytest.mark.parametrize("m,k", [(60, 200), (300, 4000), (129, 77)])
def test_herk_split_accuracy(form, nterms, nprod, bound, m, k):
    print(f"herk split form={form} nt={nterms} np={nprod} m={m} k={k}",
          flush=True)
    rng = np.random.default_rng(m * k)
    # beam-like magnitudes spanning 3 decades, unit phases
    amp = np.exp(rng.uniform(-3, 0, (1, k)) * np.log(10))
    z = (amp * np.exp(2j * np.pi * rng.uniform(size=(m, k)))).astype(np.complex64)
    from matvis_amd.gpu.matprod import f16_scale
    s = f16_scale(float(np.abs(z).max())) if form == FORM_SPLIT_F16 else 1.0
    ts, ncols = _split_terms(z, form, nterms, s)
    # FP32 accumulation spans one K-slice; the slice sums are FP64
    from matvis_amd.gpu.herk import Herk
    cap = Herk(form, m, nterms, nprod).nslices(k)
    h, v = _run_herk(form, ts, m, m, 1, ncols, k, nterms, nprod,
                     nslices=cap, scale=1.0 / (s * s))
    got = _vis_full(h, v, m)
    z64 = z.astype(np.complex128)
    want = z64.conj() @ z64.T
    d = np.sqrt(np.abs(np.diag(want)))
    err = (np.abs(got - want) / (d[:, None] * d[None])).max()
    print(f"  normalised max error {err:.3e} (bound {bound:.0e})", flush=True)
    assert err < bound


@pytest.mark.parametrize("nfeed", [1, 2])
@pytest.mark.parametrize("antpairs", [True, False])
@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("method", ["GPUMatMul", "GPUVectorDot",
                                    "GPUMatMulSplitBF16", "GPUMatMulSplitF16"])
@pytest.mark.parametrize("nchunks", [1, 2])
def test_matprod_known_answer(nfeed, antpairs, precision, method, nchunks):
    """tests/test_matprod.py:14-71 with the GPU classes."""
    from matvis_amd.gpu import matprod as module

    cls = getattr(module, method)
    nant, nsrc = 5, 15
    pairs = (np.array([(i, j) for i in range(nant) for j in range(nant)])
             if antpairs else None)
    obj = cls(nchunks=nchunks, nfeed=nfeed, nant=nant, antpairs=pairs,
              precision=precision)
    obj.setup()
    ctype = np.complex64 if precision == 1 else np.complex128
    z = np.arange(nfeed * nant * nsrc, dtype=ctype).reshape(nfeed * nant, nsrc)
    z = z + 1j
    out = np.zeros((obj.npairs, nfeed, nfeed), dtype=ctype)
    obj(_dev(z), chunk=0)
    obj.sum_chunks(out)
    true = (z.astype(np.complex128).conj().dot(z.astype(np.complex128).T)
            .reshape((nant, nfeed, nant, nfeed)).transpose((0, 2, 3, 1))
            .reshape((-1, nfeed, nfeed)))
    assert out.dtype.type == ctype and out.shape == (obj.npairs, nfeed, nfeed)
    rtol = 1e-7 if (precision == 2 and "Split" not in method) else 1e-5
    np.testing.assert_allclose(out, true, rtol=rtol)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_zdotz(dtype):
    """tests/test_cublas.py:13-25: a.conj() @ a.T for a = arange(6) + 1j."""
    from matvis_amd.gpu.matprod import GPUMatMul

    a = np.arange(6).reshape((2, 3)).astype(dtype) + 1j
    mp = GPUMatMul(1, 1, 2, None, precision=2 if dtype == np.complex128 else 1)
    mp.setup()
    out = np.zeros((4, 1, 1), dtype)
    mp(_dev(a), 0)
    mp.sum_chunks(out)
    np.testing.assert_allclose(out.reshape(2, 2), np.dot(a.conj(), a.T),
                               rtol=1e-5 if dtype == np.complex64 else 1e-10)


def test_vacc_sum_and_pair_expand():
    m, nant, nf = 6, 3, 2
    za = _herk_case(m, 40, np.complex128, 1)
    zb = _herk_case(m, 25, np.complex128, 2)
    from matvis_amd.gpu.herk import Herk
    h = Herk(FORM_C128, m)
    va, vb = h.new_vacc(), h.new_vacc()
    h.run([_dev(za)], 40, 1, 40, 40, va, False)
    h.run([_dev(zb)], 25, 1, 25, 25, vb, False)
    both = torch.stack([va, vb])
    dst = h.new_vacc()
    call("mv_vacc_sum", D.ptr(both), 2, 2 * h.ldv * h.ldv, D.ptr(dst),
         2 * h.ldv * h.ldv, D.stream_ptr())
    pairs = np.array([[2, 0], [1, 1], [0, 2]], np.int32)
    out = D.zeros((3, nf, nf), np.complex64)
    pairs_d = _dev(pairs)
    call("mv_vis_expand", D.ptr(dst), h.ldv, h.tile, nant, nf,
         D.ptr(pairs_d), 3, D.ptr(out), 0, D.stream_ptr())
    z = np.concatenate([za, zb], 1)
    want = O.vis_from_z(z, nant, nf, pairs)
    np.testing.assert_allclose(D.to_host(out), want, rtol=1e-6)


@pytest.mark.parametrize("form,nterms,nprod", [(FORM_C128, 1, 1),
                                               (FORM_SPLIT_BF16, 2, 3)])
def test_herk_phases_match_single_call(form, nterms, nprod):
    """mv_herk_compute then mv_herk_reduce equals mv_herk bit for bit."""
    from matvis_amd.gpu.herk import Herk
    print(f"herk phases form={form}", flush=True)
    m, k = 150, 700
    z = _herk_case(m, k, np.complex64, seed=5)
    if form == FORM_C128:
        zs, ld = [_dev(z.astype(np.complex128))], k
        ldz = ld
    else:
        zs, ld = _split_terms(z, form, nterms)
        ldz = m  # rows of a K tile
    h = Herk(form, m, nterms, nprod)
    v1, v2 = h.new_vacc(), h.new_vacc()
    cap = h.nslices(k)
    h.run(zs, ldz=ldz, nseg=1, seg_stride=ld, seg_len=k, vacc=v1,
          accumulate=False, nslices=cap)
    marks = []
    h.run(zs, ldz=ldz, nseg=1, seg_stride=ld, seg_len=k, vacc=v2,
          accumulate=False, nslices=cap, between=lambda s: marks.append(s))
    assert len(marks) == 1
    np.testing.assert_array_equal(_vis_full(h, v1, m), _vis_full(h, v2, m))
