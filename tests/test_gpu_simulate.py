"""End-to-end parity of the drop-in entry points against the CPU oracle.

The oracle and the device share the per-time rotations R_t (passed through
``CoordinateRotationMatrix``), so every difference is the per-time pipeline's.
Error metric: |V - V_ref| / sqrt(|V_ref_ii| |V_ref_jj|), max over baselines,
feeds and times.  Bounds: FP64 plain 1e-12, complex64 1e-5, split forms as
stated.
"""

import numpy as np
import pytest

import oracle as O
from _problems import small_problem
from matvis_amd import astro
from matvis_amd.beams import GaussianBeam
from matvis_amd.gpu.gpu import simulate, simulate_multi

pytestmark = pytest.mark.gpu


def run_gpu(p, **kw):
    opts = dict(antpos=p["antpos"], freq=p["freq"], times=p["times"],
                skycoords=astro.SkyCoord(p["ra"], p["dec"]),
                telescope_loc=p["loc"], I_sky=p["I_sky"],
                beam_list=p["beam_list"], polarized=p["polarized"],
                coord_method="CoordinateRotationMatrix",
                coord_method_params={"rotations": p["rotations"]},
                beam_spline_opts={"order": p["order"]}, precision=2)
    opts.update(kw)
    return simulate(**opts)


def run_oracle(p, **kw):
    opts = dict(antpos=p["antpos"], freq=p["freq"], eci=p["eci"],
                rotations=p["rotations"], I_sky=p["I_sky"],
                beam=p["oracle_beam"], polarized=p["polarized"])
    opts.update(kw)
    return O.simulate(**opts)


def nerr(v, ref, nant, polarized):
    nf = 2 if polarized else 1
    if not polarized:
        v, ref = v[..., None, None], ref[..., None, None]
    return O.normalized_error(v, ref, nant, nf)


@pytest.mark.parametrize("polarized", [False, True])
def test_gaussian_fp64_matches_oracle(polarized):
    """Config 0 scaled to 3 times: 16 antennas, 2000 sources."""
    p = small_problem(nant=16, nsrc=2000, ntimes=3, polarized=polarized)
    v = run_gpu(p)
    ref = run_oracle(p)
    assert v.dtype == np.complex128 and v.shape == ref.shape
    e = nerr(v, ref, 16, polarized)
    print(f"gaussian fp64 pol={polarized}: normalised max {e[0]:.2e}, "
          f"per-element max {e[2]:.2e}", flush=True)
    assert e[0] < 1e-12


@pytest.mark.parametrize("order", [1, 3])
def test_grid_beam_fp64_matches_oracle(order):
    p = small_problem(nant=12, nsrc=1500, ntimes=2, polarized=True,
                      beam="grid", order=order)
    e = nerr(run_gpu(p), run_oracle(p), 12, True)
    print(f"grid order={order}: normalised max {e[0]:.2e}", flush=True)
    assert e[0] < 1e-12


@pytest.mark.parametrize("polarized", [True, False])
def test_grid_beam_off_node_frequency_matches_oracle(polarized):
    """A grid beam with 5 frequency nodes whose patterns change with
    frequency (a cubic in frequency of three harmonic patterns), simulated
    at 137.3 MHz, between nodes: the cubic frequency step (after the power
    conversion when unpolarized) on the host, then the device pipeline,
    against the oracle at FP64."""
    from matvis_amd import synth
    from matvis_amd.beams import GridBeam

    p = small_problem(nant=10, nsrc=1200, ntimes=2, polarized=polarized,
                      beam="grid", order=3)
    pats = [synth.harmonic_efield_beam(seed=20 + k, naz=72, nza=19,
                                       dtype=np.complex128).data_array[:, :, 0]
            for k in range(3)]
    nodes = np.array([100e6, 125e6, 150e6, 175e6, 200e6])
    t = (nodes - 150e6) / 50e6
    data = np.stack([pats[0] + ti * pats[1] + ti ** 3 * pats[2] for ti in t],
                    axis=2)
    b0 = p["beam_list"][0]
    beam = GridBeam(data, b0.axis1_array, b0.axis2_array, freq_array=nodes)
    freq = 137.3e6
    planes, daz, dza, azmin = O.prepare_grid(
        data, beam.axis1_array, beam.axis2_array, nodes, freq, polarized)
    obeam = {"kind": "grid", "planes": planes[None], "daz": [daz],
             "dza": [dza], "azmin": [azmin], "order": 3}
    v = run_gpu(p, beam_list=[beam], freq=freq)
    ref = run_oracle(p, beam=obeam, freq=freq)
    e = nerr(v, ref, 10, polarized)
    print(f"off-node frequency pol={polarized}: normalised max {e[0]:.2e}",
          flush=True)
    assert e[0] < 1e-12


def test_multichunk_source_buffer_and_antpairs():
    p = small_problem(nant=10, nsrc=3001, ntimes=2, polarized=True)
    pairs = np.array([(0, 0), (3, 7), (7, 3), (9, 1), (2, 2)])
    v = run_gpu(p, min_chunks=3, source_buffer=0.8, antpairs=pairs)
    ref = run_oracle(p, nchunks=3, source_buffer=0.8, antpairs=pairs)
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    v1 = run_gpu(p, min_chunks=1, antpairs=pairs)
    np.testing.assert_allclose(v1, v, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("method,terms,bound", [
    ("GPUMatMulSplitBF16", 2, 2e-6), ("GPUMatMulSplitBF16", 3, 1e-6),
    ("GPUMatMulSplitF16", 2, 1e-6)])
def test_split_forms_multichunk_beams_and_pairs(method, terms, bound):
    """The split forms through the drop-in path with 3 source chunks (each
    chunk's HERK adds into the FP64 accumulator; split-FP16 scales each
    chunk by its own power of two), a source buffer and two grid beams
    assigned per antenna, against the oracle; a pair subset equals the same
    pairs of the full output bit for bit."""
    nant = 12
    p = small_problem(nant=nant, nsrc=3001, ntimes=2, polarized=True,
                      beam="grid", order=3, nbeam=2)
    bidx = np.arange(nant) % 2
    kw = dict(matprod_method=method, split_terms=terms, min_chunks=3,
              source_buffer=0.8, beam_idx=bidx)
    v = run_gpu(p, **kw)
    ref = run_oracle(p, nchunks=3, source_buffer=0.8, beam_idx=bidx)
    e = nerr(v, ref, nant, True)
    print(f"{method} ({terms} terms), 3 chunks: normalised max {e[0]:.2e}",
          flush=True)
    assert e[0] < bound
    pairs = np.array([(0, 0), (3, 7), (7, 3), (11, 1), (2, 2)])
    vp = run_gpu(p, antpairs=pairs, **kw)
    full = v.reshape(v.shape[0], nant, nant, 2, 2)
    assert np.array_equal(vp, full[:, pairs[:, 0], pairs[:, 1]])


def test_source_buffer_overflow_raises():
    p = small_problem(nant=4, nsrc=3000, ntimes=1)
    with pytest.raises(ValueError, match="is too small for the number of"):
        run_gpu(p, source_buffer=0.3)


def test_antizenith_is_exact_zero():
    """tests/test_matvis_gpu.py:15-25."""
    p = small_problem(nant=3, nsrc=1, ntimes=2, polarized=True)
    lat = np.radians(-30.72)
    p["ra"], p["dec"] = np.array([0.0]), np.array([np.pi / 2 + lat - 0.2])
    p["I_sky"] = np.array([2.0])
    assert np.all(run_gpu(p) == 0)


def test_multibeam_invariances():
    """tests/test_matvis_gpu.py:28-110: per-antenna copies == one beam;
    permuted beam list with remapped beam_idx is the same; a different
    assignment differs."""
    p = small_problem(nant=3, nsrc=400, ntimes=1)
    ga, gb = GaussianBeam(sigma=0.2), GaussianBeam(sigma=0.12)
    v1 = run_gpu(p, beam_list=[ga])
    v2 = run_gpu(p, beam_list=[ga] * 3, beam_idx=np.zeros(3, int))
    assert np.array_equal(v1, v2)
    vab = run_gpu(p, beam_list=[ga, gb], beam_idx=np.array([0, 1, 0]))
    vba = run_gpu(p, beam_list=[gb, ga], beam_idx=np.array([1, 0, 1]))
    assert np.array_equal(vab, vba)
    v101 = run_gpu(p, beam_list=[ga, gb], beam_idx=np.array([1, 0, 1]))
    assert not np.allclose(vab, v101, rtol=0, atol=1e-12)
    ref = run_oracle(p, beam={"kind": "gaussian", "sigma": [0.2, 0.12]},
                     beam_idx=np.array([0, 1, 0]))
    assert nerr(vab, ref, 3, False)[0] < 1e-12


def test_mixed_beams_raise():
    from matvis_amd import synth

    p = small_problem(nant=3, nsrc=10, ntimes=1)
    grid = synth.harmonic_efield_beam(1, naz=36, nza=10)
    with pytest.raises(ValueError, match="GPUBeamInterpolator only supports"):
        run_gpu(p, beam_list=[grid, GaussianBeam(sigma=0.1), grid])


def test_single_precision_matches_oracle():
    """tests/test_matvis_gpu.py:127-139 style: precision 1 within 1e-5."""
    p = small_problem(nant=12, nsrc=1500, ntimes=2, polarized=True, beam="grid")
    v = run_gpu(p, precision=1)
    assert v.dtype == np.complex64
    e = nerr(v, run_oracle(p), 12, True)
    print(f"precision 1: normalised max {e[0]:.2e}", flush=True)
    assert e[0] < 1e-5


def test_three_term_split_at_precision_1_betters_complex64():
    """Three bf16 terms hold a complex64 Z exactly (3 x 8 significand bits);
    the six kept products drop only terms below FP32 resolution, and the
    diagonal is summed exactly.  At precision 1 the form stays within the
    complex64 bound and is no less accurate than the plain complex64 HERK on
    the same problem (normalised max and median)."""
    p = small_problem(nant=20, nsrc=4000, ntimes=2, polarized=True,
                      beam="grid", order=3)
    ref = run_oracle(p)
    v64 = run_gpu(p, precision=1)
    v3 = run_gpu(p, precision=1, matprod_method="GPUMatMulSplitBF16",
                 split_terms=3)
    assert v3.dtype == np.complex64
    e64, e3 = nerr(v64, ref, 20, True), nerr(v3, ref, 20, True)
    print(f"precision 1: complex64 {e64[0]:.2e} (median {e64[1]:.2e}), "
          f"3-term bf16 {e3[0]:.2e} (median {e3[1]:.2e})", flush=True)
    assert e3[0] < 1e-5
    assert e3[0] <= 1.1 * e64[0] and e3[1] <= 1.1 * e64[1]


# Bounds on this 20-antenna, 4000-source bicubic problem: the two-term bf16
# split leaves a 2^-18 representation residual per element, about 1e-6 here
# with 3 or 4 products (the diagonal is summed exactly, so lo*lo only
# matters off the diagonal, where it averages out); 3 bf16 terms and the
# fp16 split are limited by FP32 accumulation, a few 1e-7.
@pytest.mark.parametrize("method,terms,prods,bound", [
    ("GPUMatMulSplitBF16", 2, 4, 2e-6), ("GPUMatMulSplitBF16", 2, 3, 2e-6),
    ("GPUMatMulSplitBF16", 3, 6, 1e-6), ("GPUMatMulSplitF16", 2, 3, 1e-6)])
def test_split_forms_match_oracle(method, terms, prods, bound):
    p = small_problem(nant=20, nsrc=4000, ntimes=2, polarized=True, beam="grid",
                      order=3)
    v = run_gpu(p, matprod_method=method, split_terms=terms,
                split_products=prods)
    e = nerr(v, run_oracle(p), 20, True)
    print(f"{method} ({terms},{prods}): normalised max {e[0]:.2e} median "
          f"{e[1]:.2e}", flush=True)
    assert e[0] < bound


@pytest.mark.parametrize("precision,method,bound", [
    (2, "GPUMatMul", 1e-12), (1, "GPUMatMul", 1e-5),
    (2, "GPUMatMulSplitBF16", 1e-5)])
def test_cpu_runnable_configuration(precision, method, bound):
    """The CPU-runnable configuration through simulate_vis: 16 antennas in a
    100 m disk, 2,000 sources (RA uniform, sin(dec) uniform on [-90, +30]
    deg, power-law S^-2 fluxes on [0.1, 10] Jy), 150 MHz, 10 times 6 min
    apart, Gaussian beam sigma = 10 deg, unpolarized; output
    (1, ntimes, nant^2) against the oracle.  The split bound is wider than
    the 100k-source sky's: a power-law sky's few bright sources do not
    average the bf16 representation residual down."""
    import matvis_amd

    p = small_problem(nant=16, nsrc=2000, ntimes=10, polarized=False,
                      beam="gaussian")
    ants = {i: tuple(p["antpos"][i]) for i in range(16)}
    vis = matvis_amd.simulate_vis(
        ants, p["I_sky"][:, None], p["ra"], p["dec"], np.array([p["freq"]]),
        p["times"], p["beam_list"], p["loc"], polarized=False,
        precision=precision, use_gpu=True, matprod_method=method,
        coord_method="CoordinateRotationMatrix",
        coord_method_params={"rotations": p["rotations"]})
    assert vis.shape == (1, 10, 256)
    assert vis.dtype == (np.complex128 if precision == 2 else np.complex64)
    e = nerr(vis[0], run_oracle(p), 16, False)
    print(f"16 ant x 2000 src x 10 times, {method} precision {precision}: "
          f"normalised max {e[0]:.2e} median {e[1]:.2e}", flush=True)
    assert e[0] < bound


def test_simulate_vis_frequencies_match_single_calls():
    """wrapper.py:180-199: the frequency loop; here rotations are shared."""
    import matvis_amd

    p = small_problem(nant=6, nsrc=500, ntimes=2, polarized=True)
    freqs = np.array([120e6, 150e6, 180e6])
    fluxes = p["I_sky"][:, None] * (freqs[None] / 150e6) ** -0.8
    ants = {i: tuple(p["antpos"][i]) for i in range(6)}
    kw = dict(coord_method="CoordinateRotationMatrix",
              coord_method_params={"rotations": p["rotations"]})
    vis = matvis_amd.simulate_vis(ants, fluxes, p["ra"], p["dec"], freqs,
                                  p["times"], p["beam_list"], p["loc"],
                                  polarized=True, precision=2, use_gpu=True,
                                  **kw)
    assert vis.shape == (3, 2, 36, 2, 2)
    for i, f in enumerate(freqs):
        single = run_gpu(p, freq=f, I_sky=fluxes[:, i])
        assert np.array_equal(vis[i], single)
        ref = run_oracle(p, freq=f, I_sky=fluxes[:, i])
        assert nerr(vis[i], ref, 6, True)[0] < 1e-12


def test_plugin_loop_matches_oracle():
    """Drive the plugin classes as gpu.simulate's loop does
    (gpu/gpu.py:197-254) and compare with the oracle."""
    from matvis_amd.core.coords import CoordinateRotation
    from matvis_amd.core.getz import ZMatrixCalc
    from matvis_amd.core.tau import TauCalculator
    from matvis_amd.gpu import coords as _  # noqa: F401
    from matvis_amd.gpu.beams import GPUBeamInterpolator
    from matvis_amd.gpu.matprod import GPUMatMul

    p = small_problem(nant=8, nsrc=1200, ntimes=2, polarized=True, beam="grid")
    nchunks, chunk = 2, 600
    co = CoordinateRotation._methods["CoordinateRotationMatrix"](
        rotations=p["rotations"], flux=np.sqrt(0.5 * p["I_sky"]),
        times=p["times"], telescope_loc=p["loc"],
        skycoords=astro.SkyCoord(p["ra"], p["dec"]), chunk_size=chunk,
        precision=2, source_buffer=1.0)
    bm = GPUBeamInterpolator(beam_list=p["beam_list"], beam_idx=None,
                             polarized=True, nant=8, freq=p["freq"],
                             nsrc=co.nsrc_alloc, precision=2)
    tau = TauCalculator(antpos=p["antpos"], freq=p["freq"], nsrc=co.nsrc_alloc,
                        precision=2)
    zc = ZMatrixCalc(nant=8, nfeed=2, nax=2, nsrc=co.nsrc_alloc,
                     ctype=np.complex128)
    mp = GPUMatMul(nchunks, 2, 8, None, precision=2)
    for obj in (tau, bm, co, zc, mp):
        obj.setup()
    vis = np.zeros((2, 64, 2, 2), np.complex128)
    for t in range(2):
        co.rotate(t)
        for c in range(nchunks):
            crd, isq, n = co.select_chunk(c, t)
            a = bm(crd[0], crd[1], check=t == 0)
            et = tau(crd)
            z = zc(isq, a, et, bm.beam_idx)
            mp(z, c)
        mp.sum_chunks(vis[t])
    ref = run_oracle(p, nchunks=2)
    assert nerr(vis, ref, 8, True)[0] < 1e-12
    # all_coords_topo is the full rotation of every source
    np.testing.assert_array_equal(
        co.all_coords_topo.cpu().numpy(), O.rotate(p["rotations"][1], p["eci"]))


def test_nan_beam_raises():
    """tests/test_beams.py:124-146: a NaN in the beam is reported."""
    from matvis_amd.gpu.beams import GPUBeamInterpolator
    from matvis_amd import synth

    b = synth.harmonic_efield_beam(2, naz=36, nza=10, dtype=np.complex128)
    b.data_array[1] = np.nan
    tx = np.linspace(-0.7, 0.7, 100)
    bm = GPUBeamInterpolator(beam_list=[b], beam_idx=np.zeros(1, int),
                             polarized=True, nant=1, freq=150e6, nsrc=100,
                             precision=2)
    bm.setup()
    with pytest.raises(ValueError, match="invalid value"):
        bm(tx, tx, check=True)


def test_simulate_multi_shapes_and_dtype():
    p = small_problem(nant=5, nsrc=300, ntimes=2)
    v = simulate_multi(antpos=p["antpos"], freqs=[1e8, 1.2e8], times=p["times"],
                       skycoords=astro.SkyCoord(p["ra"], p["dec"]),
                       telescope_loc=p["loc"],
                       fluxes=np.stack([p["I_sky"]] * 2, 1),
                       beam_list=p["beam_list"], precision=1,
                       coord_method="CoordinateRotationERFA")
    assert v.shape == (2, 2, 25, 1, 1) and v.dtype == np.complex64
    assert np.isfinite(v).all()


# ------------------------------------------------------------------ diffuse
# Diffuse sky with per-antenna beams and source chunks: HEALPix pixel
# centres, log-normal(-2, 1.5) brightness, four +-5% perturbations of one
# E-field beam on the 360 az x 91 za grid at 1 deg, assigned round-robin
# (beam_idx = ant % 4), source chunks.


def _diffuse(nside, nant, nfreq, ntimes, seed=3):
    from matvis_amd import synth
    ra, dec = synth.healpix_centres(nside)
    flux = synth.lognormal_fluxes(ra.size, seed, -2.0, 1.5)
    antpos = synth.hex_array(nant, 14.6, 0.02, seed=seed)
    beams = [synth.harmonic_efield_beam(seed=5, naz=360, nza=91,
                                        perturb=0.05, perturb_seed=100 + b,
                                        dtype=np.complex128)
             for b in range(4)]
    freqs = np.linspace(120e6, 180e6, nfreq)
    times = synth.times_spaced(ntimes, 10.0)
    loc = synth.hera_location()
    rot = astro.rotation_matrices(times, loc)
    return dict(ra=ra, dec=dec, flux=flux, antpos=antpos, beams=beams,
                freqs=freqs, times=times, loc=loc, rot=rot,
                beam_idx=np.arange(nant) % 4)


def _diffuse_gpu(d, freq, **kw):
    opts = dict(antpos=d["antpos"], freq=freq, times=d["times"],
                skycoords=astro.SkyCoord(d["ra"], d["dec"]),
                telescope_loc=d["loc"], I_sky=d["flux"],
                beam_list=d["beams"], beam_idx=d["beam_idx"], polarized=True,
                coord_method="CoordinateRotationMatrix",
                coord_method_params={"rotations": d["rot"]},
                beam_spline_opts={"order": 3}, precision=2)
    opts.update(kw)
    return simulate(**opts)


@pytest.mark.parametrize("method,bound", [("GPUMatMul", 1e-12),
                                          ("GPUMatMulSplitBF16", 2e-6)])
def test_diffuse_multibeam_chunked_matches_oracle(method, bound):
    """nside 16 (3072 pixels), 16 antennas, 2 frequencies, 2 times, chunks of
    512 sources, against the oracle with the same chunking."""
    d = _diffuse(16, 16, 2, 2)
    eci = O.point_source_crd_eq(d["ra"], d["dec"])
    preps = [O.prepare_grid(b.data_array, b.axis1_array, b.axis2_array,
                            b.freq_array, 150e6, True) for b in d["beams"]]
    beam = {"kind": "grid", "planes": np.stack([p[0] for p in preps]),
            "daz": [p[1] for p in preps], "dza": [p[2] for p in preps],
            "azmin": [p[3] for p in preps], "order": 3}
    for f in d["freqs"]:
        print(f"diffuse {method} freq={f / 1e6:.0f} MHz", flush=True)
        v = _diffuse_gpu(d, f, min_chunks=6, matprod_method=method)
        ref = O.simulate(antpos=d["antpos"], freq=f, eci=eci,
                         rotations=d["rot"], I_sky=d["flux"], beam=beam,
                         polarized=True, beam_idx=d["beam_idx"], nchunks=6)
        e = nerr(v, ref, 16, True)
        print(f"  normalised max {e[0]:.2e}", flush=True)
        assert e[0] < bound


def test_diffuse_full_size_chunk_sums():
    """HEALPix nside 128 (196,608 sources), 350 antennas, one frequency and
    time, chunks of 16,384 sources: the FP64 result does not depend on the
    chunking; split-BF16 stays within 1e-6; summing the per-chunk results in
    FP32 instead of FP64 (the FP32-only variant) is stated beside them."""
    d = _diffuse(128, 350, 1, 1)
    f = float(d["freqs"][0])
    print("diffuse full: fp64 12 chunks", flush=True)
    v12 = _diffuse_gpu(d, f, min_chunks=12)
    v1 = _diffuse_gpu(d, f, min_chunks=1)
    e_chunk = nerr(v12, v1, 350, True)[0]
    print("diffuse full: split-BF16 12 chunks", flush=True)
    vb = _diffuse_gpu(d, f, min_chunks=12, matprod_method="GPUMatMulSplitBF16")
    e_bf16 = nerr(vb, v12, 350, True)[0]
    # FP32-only cross-chunk sum: each chunk's visibilities rounded to FP32
    # and accumulated in FP32
    acc32 = np.zeros(v12.shape, np.complex64)
    n = d["ra"].size
    for c in range(12):
        sl = slice(c * n // 12, (c + 1) * n // 12)
        dc = dict(d, ra=d["ra"][sl], dec=d["dec"][sl], flux=d["flux"][sl])
        acc32 += _diffuse_gpu(dc, f, min_chunks=1).astype(np.complex64)
    e_fp32 = nerr(acc32.astype(np.complex128), v12, 350, True)[0]
    print(f"  chunking invariance (fp64) {e_chunk:.2e}; split-BF16 vs fp64 "
          f"{e_bf16:.2e}; FP32-only chunk sums vs FP64 {e_fp32:.2e}",
          flush=True)
    assert e_chunk < 1e-12
    assert e_bf16 < 1e-6
    assert e_fp32 > e_chunk


# ------------------------------------------------------------ full size
# The headline size (350 hex antennas x 100,000 sources, polarized, bicubic
# complex64 E-field beam), where the oracle is too slow: properties that hold
# at any size, and the split form against the FP64 form.


@pytest.fixture(scope="module")
def full_size():
    from matvis_amd import synth

    nant, nsrc = 350, 100_000
    antpos = synth.hex_array(nant, 14.6, 0.02, seed=1)
    ra, dec = synth.random_sky(nsrc, 2, -90.0, 30.0)
    flux = synth.lognormal_fluxes(nsrc, 3, 0.0, 1.0)
    freqs = np.array([100e6, 200e6])
    times = synth.times_spaced(1, 2.0)
    loc = synth.hera_location()
    rot = astro.rotation_matrices(times, loc)
    beam = synth.harmonic_efield_beam(seed=4, naz=360, nza=91)
    kw = dict(antpos=antpos, freqs=freqs, times=times,
              skycoords=astro.SkyCoord(ra, dec), telescope_loc=loc,
              beam_list=[beam], polarized=True, precision=2,
              coord_method="CoordinateRotationMatrix",
              coord_method_params={"rotations": rot},
              beam_spline_opts={"order": 3})
    out = {}
    for method in ("GPUMatMul", "GPUMatMulSplitBF16"):
        print(f"full size {method}", flush=True)
        for s in (1.0, 4.0):
            f = np.repeat(flux[:, None] * s, 2, 1)
            out[method, s] = simulate_multi(fluxes=f, matprod_method=method,
                                            **kw)
    return nant, out


@pytest.mark.parametrize("method", ["GPUMatMul", "GPUMatMulSplitBF16"])
def test_full_size_hermitian_and_flux_linear(full_size, method):
    nant, out = full_size
    v = out[method, 1.0]  # (nfreq, ntimes, nant^2, 2, 2)
    w = v.reshape(2, 1, nant, nant, 2, 2)
    # V[i, j, p, q] == conj(V[j, i, q, p]) exactly (both come from one
    # upper-triangle element) except on the self-correlations (i = j,
    # p = q); those are real (the split forms write the exact FP64 row norm;
    # FP64 cancels re*im - im*re to rounding) and positive
    herm = np.conj(w.transpose(0, 1, 3, 2, 5, 4))
    idx = np.arange(nant)
    off = np.ones(w.shape, bool)
    for p in range(2):
        off[:, :, idx, idx, p, p] = False
    np.testing.assert_array_equal(w[off], herm[off])
    autos = np.diagonal(w[:, :, idx, idx], axis1=-2, axis2=-1)
    assert np.all(autos.real > 0)
    assert np.all(np.abs(autos.imag) <= 1e-13 * autos.real)
    if method != "GPUMatMul":
        assert np.all(autos.imag == 0)
    # sqrt(4 I / 2) = 2 sqrt(I / 2): Z and its 16-bit split scale by 2
    # exactly, V by 4
    np.testing.assert_array_equal(out[method, 4.0], 4.0 * v)


@pytest.mark.parametrize("freq", [100e6, 200e6])
def test_config1_full_size_forms_match_oracle(freq):
    """The plain-vs-mixed-precision configuration at its size: 350 hex
    antennas (14.6 m, 2 cm jitter), 50,000 sources (RA uniform, sin(dec)
    uniform on [-90, +30] deg, log-normal(0, 1) fluxes), polarized, the
    360 az x 91 za complex64 E-field beam, bicubic; the first of its 30
    times, at the ends of its 100-200 MHz band.  Each form against the
    oracle: FP64 <= 1e-12, complex64 (precision 1) <= 1e-5, split-BF16 <=
    1e-6 (normalised max); the per-element max is printed beside them."""
    from matvis_amd import synth

    nant, nsrc = 350, 50_000
    antpos = synth.hex_array(nant, 14.6, 0.02, seed=1)
    ra, dec = synth.random_sky(nsrc, 2, -90.0, 30.0)
    flux = synth.lognormal_fluxes(nsrc, 3, 0.0, 1.0)
    times = synth.times_spaced(1, 2.0)
    loc = synth.hera_location()
    rot = astro.rotation_matrices(times, loc)
    beam = synth.harmonic_efield_beam(seed=4, naz=360, nza=91)
    kw = dict(antpos=antpos, freqs=np.array([freq]), times=times,
              skycoords=astro.SkyCoord(ra, dec), telescope_loc=loc,
              fluxes=flux[:, None], beam_list=[beam], polarized=True,
              coord_method="CoordinateRotationMatrix",
              coord_method_params={"rotations": rot},
              beam_spline_opts={"order": 3})
    print(f"config 1 full size, {freq / 1e6:.0f} MHz: oracle", flush=True)
    planes, daz, dza, azmin = O.prepare_grid(
        beam.data_array, beam.axis1_array, beam.axis2_array, beam.freq_array,
        freq, True, beam.beam_type)
    ref = O.simulate(antpos=antpos, freq=freq,
                     eci=O.point_source_crd_eq(ra, dec), rotations=rot,
                     I_sky=flux, beam={"kind": "grid", "planes": planes[None],
                                       "daz": [daz], "dza": [dza],
                                       "azmin": [azmin], "order": 3},
                     polarized=True)
    for method, precision, bound in (("GPUMatMul", 2, 1e-12),
                                     ("GPUMatMul", 1, 1e-5),
                                     ("GPUMatMulSplitBF16", 2, 1e-6)):
        v = simulate_multi(matprod_method=method, precision=precision, **kw)
        e = nerr(v[0], ref, nant, True)
        print(f"  {method} precision {precision}: normalised max {e[0]:.2e}"
              f" median {e[1]:.2e}, per-element max {e[2]:.2e}", flush=True)
        assert e[0] < bound, (method, precision, e[0])


def test_full_size_split_within_1e6_of_fp64(full_size):
    nant, out = full_size
    for fi in range(2):
        e = O.normalized_error(out["GPUMatMulSplitBF16", 1.0][fi],
                               out["GPUMatMul", 1.0][fi], nant, 2)
        print(f"full size freq {fi}: split-BF16 vs FP64 normalised max "
              f"{e[0]:.2e} median {e[1]:.2e}", flush=True)
        assert e[0] < 1e-6


def test_registered_rotations_are_repeatable():
    """tests/test_coordrot.py:83-97: every registered rotation method gives
    the same topocentric vectors when a time is rotated again, also after
    other times (methods whose dependencies are absent are skipped)."""
    from matvis_amd.core.coords import CoordinateRotation
    from matvis_amd.gpu import coords as _  # noqa: F401
    from matvis_amd import synth

    print("rotation repeatability", flush=True)
    ra, dec = synth.random_sky(500, 3)
    times = synth.times_spaced(3, 30.0)
    loc = synth.hera_location()
    extra = {"CoordinateRotationMatrix": {
        "rotations": astro.rotation_matrices(times, loc)}}
    checked = 0
    for name, cls in sorted(CoordinateRotation._methods.items()):
        try:
            co = cls(flux=np.ones(500), times=times, telescope_loc=loc,
                     skycoords=astro.SkyCoord(ra, dec), precision=2,
                     **extra.get(name, {}))
        except ImportError:
            continue
        co.setup()
        got = []
        for t in (0, 1, 2, 0, 2):
            co.rotate(t)
            got.append(co.all_coords_topo.cpu().numpy().copy())
        assert np.array_equal(got[0], got[3]), name
        assert np.array_equal(got[2], got[4]), name
        assert not np.array_equal(got[0], got[1]), name
        checked += 1
    assert checked >= 3


def test_max_memory_bounds_device_allocation():
    """max_memory is a bound on what this build allocates: FP64 at 350
    antennas x 100k sources under 1.5 GiB runs in 3 chunks (the reference's
    chunk model alone would give 2, which do not fit the split-K partials),
    the peak of the device allocations stays within the budget and within
    the footprint model, which overestimates it by at most 25%, and the
    visibilities equal the single-chunk run's within 1e-12."""
    import torch

    from matvis_amd import synth
    from matvis_amd.gpu.engine import device_bytes
    from matvis_amd.gpu.gpu import build_engine

    nant, nsrc, budget = 350, 100_000, int(1.5 * 2**30)
    antpos = synth.hex_array(nant, 14.6, 0.02, seed=1)
    ra, dec = synth.random_sky(nsrc, 2, -90.0, 30.0)
    flux = synth.lognormal_fluxes(nsrc, 3, 0.0, 1.0)
    times = synth.times_spaced(1, 2.0)
    loc = synth.hera_location()
    beam = synth.harmonic_efield_beam(seed=4, naz=360, nza=91)
    kw = dict(antpos=antpos, freqs=np.array([150e6]), times=times,
              skycoords=astro.SkyCoord(ra, dec), telescope_loc=loc,
              fluxes=flux[:, None], beam_list=[beam], polarized=True,
              precision=2, coord_method="CoordinateRotationMatrix",
              coord_method_params={"rotations": astro.rotation_matrices(
                  times, loc)},
              beam_spline_opts={"order": 3})
    print("max_memory 1.5 GiB, FP64 350 x 100k", flush=True)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    eng, nt = build_engine(max_memory=budget, **kw)
    v = eng.run(nt)
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    model = device_bytes(
        form=eng.form, nterms=eng.nterms, nprod=eng.nprod, nant=nant,
        polarized=True, nfreq=1, nsrc=nsrc, nchunks=eng.nchunks,
        chunk_size=eng.chunk_size, nsrc_alloc=eng.nsrc_alloc, nbeam=1,
        beam_pixels=91 * 363, beam_sets=1, npairs=nant * nant,
        precision=2)["total"]
    print(f"  {eng.nchunks} chunks; peak {peak / 2**30:.3f} GiB, model "
          f"{model / 2**30:.3f} GiB, budget {budget / 2**30:.2f} GiB",
          flush=True)
    assert eng.nchunks == 3
    assert peak <= budget
    assert peak <= model <= 1.25 * peak
    del eng
    torch.cuda.empty_cache()
    one = simulate_multi(**kw)
    assert nerr(v[0], one[0], nant, True)[0] < 1e-12
