/*
 * matvis_hip.h -- C ABI of libmatvis_hip.so, the MI355X (gfx950) kernels behind
 * matvis's per-time visibility computation.
 *
 * The reference boundary is Python only: matvis.gpu.simulate
 * (src/matvis/gpu/gpu.py:57-256) drives five plugin classes whose device work
 * is CuPy/cuBLAS.  Each entry point below replaces one of those device calls;
 * the comment on each names the reference code it stands in for.  The Python
 * package matvis_amd binds these through ctypes (matvis_amd/_lib.py).
 *
 * Conventions
 *  - Every entry point returns an int status: MV_OK (0) or a negative code.
 *    mv_last_error() returns the message of the last failure on this thread.
 *  - All array arguments are device pointers owned by the caller, except where
 *    a parameter is documented as host memory.  Complex arrays are interleaved
 *    (re, im) pairs of double (c128) or float (c64).
 *  - `stream` is a hipStream_t passed as an opaque pointer (NULL = the null
 *    stream).  Work is enqueued on it; nothing synchronizes unless stated.
 *  - Counts that are only known on the device (sources above the horizon) are
 *    passed as a device pointer `n_dev`; when it is NULL the host count is
 *    used, and otherwise the host count bounds it (min of the two), so a
 *    device count larger than the buffers never indexes past them.
 *  - Nothing here allocates device memory: scratch space is a caller-provided
 *    workspace whose size a *_workspace_bytes() query returns.
 */
#ifndef MATVIS_HIP_H
#define MATVIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *mv_stream_t;

#define MV_OK 0
#define MV_ERR_ARG (-1)      /* invalid argument (shape, pointer, enum)     */
#define MV_ERR_HIP (-2)      /* a HIP runtime call or launch failed          */
#define MV_ERR_CAPACITY (-3) /* a size exceeds what a buffer can hold        */

/* Z / HERK numeric forms. */
#define MV_FORM_C128 0       /* plain complex128, FP64 MFMA                  */
#define MV_FORM_C64 1        /* plain complex64, FP32 MFMA                   */
#define MV_FORM_SPLIT_BF16 2 /* Z split into bf16 terms, BF16 MFMA + FP32 acc */
#define MV_FORM_SPLIT_F16 3  /* Z split into fp16 terms, FP16 MFMA + FP32 acc */

/* Copies the last error message of the calling thread into buf (truncated,
 * NUL-terminated) and returns the last status code (0 if none).  The Python
 * layer raises the reference's exception types with these messages (the
 * reference raises directly: src/matvis/core/coords.py:117-121,
 * src/matvis/core/beams.py:203-207). */
int mv_last_error(char *buf, size_t len);

/* Library ABI version (incremented on any change of a signature or of an
 * argument's meaning; 3: split terms are K-tile-major).  No reference
 * counterpart: lets a binding check it loaded a matching library. */
int mv_abi_version(void);

/* ---------------------------------------------------------------- (a) ---- */

/* Rotate ECI unit vectors into the topocentric ENU frame, all sources.
 * Replaces the per-time rotation that fills CoordinateRotation.all_coords_topo
 * (src/matvis/cpu/coords.py:26-38 and the _atioq matmul, :89-116).
 * eci: (3, ld_eci) f64 rows x,y,z.  rot9: host, row-major 3x3 R_t.
 * topo: (3, ld_topo) f64 (topo_is_f64=1) or f32.  Rows are evaluated as
 * (R[i,0]*x + R[i,1]*y) + R[i,2]*z in FP64 without contraction. */
int mv_rotate(const double *eci, int64_t ld_eci, const double *rot9,
              int64_t nsrc, void *topo, int topo_is_f64, int64_t ld_topo,
              mv_stream_t stream);

/* Bytes of workspace mv_rotate_cull needs for a chunk of chunk_len sources
 * (the scan's block sums).  No reference counterpart: xp.where allocates
 * its own (src/matvis/core/coords.py:115). */
size_t mv_rotate_cull_workspace_bytes(int64_t chunk_len);

/* Fused rotation + horizon cull (z > 0) + stable stream compaction of sources
 * [chunk0, chunk0 + chunk_len).  Replaces CoordinateRotation.select_chunk
 * (src/matvis/core/coords.py:99-150): ascending source order is kept, flux is
 * gathered, and positions n_up..nalloc-1 are zero-filled (coords and flux).
 * rot9 (host) may be the identity when `eci` already holds ENU vectors.
 * crd_out: (3, ld_crd), flux_out: (nalloc) [nullable], both f64 if
 * out_is_f64 else f32.  flux_in: f64 (nsrc) indexed by global source id
 * [nullable].  idx_out: int32 (nalloc) global source ids [nullable].
 * n_up_dev: int32 on device, receives the count; if it exceeds nalloc only
 * nalloc sources are stored and the count still reports the true number
 * (the caller raises the reference's "nsrc_alloc ... too small" error). */
int mv_rotate_cull(const double *eci, int64_t ld_eci, const double *rot9,
                   const double *flux_in, int64_t chunk0, int64_t chunk_len,
                   int64_t nalloc, void *crd_out, int64_t ld_crd,
                   void *flux_out, int out_is_f64, int32_t *idx_out,
                   int32_t *n_up_dev, void *workspace, size_t ws_bytes,
                   mv_stream_t stream);

/* BCRS corrections of ICRS unit vectors, the per-source part of the ERFA
 * rotation: light deflection by the Sun (limiter dlim, 1e-6 in the
 * reference), aberration and the bias-precession-nutation matrix, giving CIRS
 * proper directions.  Replaces CoordinateRotationERFA._set_bcrs with _ld,
 * _ab and _bpn (src/matvis/cpu/coords.py:110-215) and the light-deflection
 * kernel of GPUCoordinateRotationERFA (src/matvis/gpu/coords.py:7-53).
 * eci, out: (3, ld) f64 SoA, nsrc sources (out may alias eci).
 * ctx17 (host): eh[3] (Sun -> observer unit vector), em (au), v[3]
 * (barycentric velocity / c), bm1 (sqrt(1 - |v|^2)), bpn[9] row-major.
 * mv_rotate_cull then applies the per-time CIRS -> ENU matrix to `out`. */
int mv_bcrs_update(const double *eci, int64_t ld_eci, int64_t nsrc,
                   const double *ctx17, double dlim, double *out,
                   int64_t ld_out, mv_stream_t stream);

/* Rotation of polarized sources' coherency matrices from the equatorial
 * (RA/Dec) basis into the alt/az basis, for the sources mv_rotate_cull kept.
 * Replaces the polarized branch of CoordinateRotation.select_chunk with
 * _rotate_frame_coherency (src/matvis/core/coords.py:123-175) and
 * calc_coherency_rotation (src/matvis/coordinates.py:288-491).
 * ra, dec: f64 (nsrc) radians, by global source id.  idx, crd, n_up_dev: the
 * idx_out (int32, nalloc), crd_out ((3, ld_crd), f64 if crd_is_f64 else f32)
 * and count mv_rotate_cull wrote.  r_avg9 (host): row-major 3x3 carrying ICRS
 * unit vectors to the (north, east, up) frame at the time.
 * coh: (nsrc, nfreq, 2, 2), out: (nalloc, nfreq, 2, 2), both f64 if
 * coh_is_f64 else f32, complex (interleaved re, im) if coh_is_complex.
 * out[s] = rot^T coh[idx[s]] rot with rot = [[cos X, sin X], [-sin X,
 * cos X]] per source; positions [n_up, nalloc) are zero-filled.  Computed
 * in FP64 throughout; where the per-source alignment rotation's axis is
 * undefined (the two directions coincide) it is the identity, the limit the
 * reference's 0/0 leaves as NaN. */
int mv_coherency_rotate(const double *ra, const double *dec, int64_t nsrc,
                        const int32_t *idx, const void *crd, int64_t ld_crd,
                        int crd_is_f64, const int32_t *n_up_dev,
                        int64_t nalloc, const double *r_avg9, const void *coh,
                        int64_t nfreq, int coh_is_f64, int coh_is_complex,
                        void *out, mv_stream_t stream);

/* ---------------------------------------------------------------- (b) ---- */

/* In-place cubic B-spline prefilter of beam planes (mirror boundary), the
 * prefilter map_coordinates(order=3) applies (src/matvis/gpu/beams.py:187).
 * planes: (nplanes, nza, naz) complex (is_f64 ? c128 : c64). */
int mv_beam_prefilter(void *planes, int is_f64, int64_t nplanes, int nza,
                      int naz, mv_stream_t stream);

/* Interpolate az/za grid beams at topocentric positions.  Replaces
 * GPUBeamInterpolator._interp + gpu_beam_interpolation
 * (src/matvis/gpu/beams.py:103-200): (tx, ty) -> (az, za) in the uvbeam
 * convention, coordinates [za/dza, (az-azmin)/daz], map_coordinates with
 * mode='constant' (cval 0), order 1 (bilinear) or 3 (bicubic B-spline on
 * prefiltered planes), then a complex sqrt for power beams.
 * planes: (nbeam, nax, nfeed, nza, naz) complex, is_f64 selects c128/c64.
 * grid: device (nbeam, 3) doubles {daz, dza, azmin}.
 * crd: (3, ld_crd) f64 or f32 (crd_is_f64).  out: (nbeam, nfeed, nax, ld_out)
 * complex in the planes' precision.  Positions p < n (n = *n_dev if n_dev else
 * count) are evaluated; positions n..count-1 are set to zero. */
int mv_beam_interp(const void *planes, int is_f64, int nbeam, int nfeed,
                   int nax, int nza, int naz, const double *grid, int order,
                   int is_power, const void *crd, int crd_is_f64,
                   int64_t ld_crd, int64_t count, const int32_t *n_dev,
                   void *out, int64_t ld_out, mv_stream_t stream);

/* mv_beam_interp with the grid staged in LDS by za band instead of tap
 * gathers.  On MI355X it measured slower than mv_beam_interp at every grid
 * whose rows fit (the planes of a 0.25 deg beam sit in the 256 MB Infinity
 * Cache), so the library uses it only when asked.
 * The positions are binned by za band (B grid rows; B + 3 rows of one plane,
 * 128 KB at most, staged per workgroup), and each result is written to its
 * own position's slot, so the output equals mv_beam_interp's when every
 * beam shares beam 0's za spacing grid[1] (the bands use it).  Replaces the
 * same reference call as mv_beam_interp.  workspace: device memory of the
 * size the next function returns for (is_f64, nza, naz, count); that is 0
 * when a band's rows would not fit in LDS, and then the gather kernel runs
 * and no workspace is needed.  count <= INT32_MAX.  Other arguments as
 * mv_beam_interp. */
size_t mv_beam_interp_banded_workspace_bytes(int is_f64, int nza, int naz,
                                             int64_t count);
int mv_beam_interp_banded(const void *planes, int is_f64, int nbeam,
                          int nfeed, int nax, int nza, int naz,
                          const double *grid, int order, int is_power,
                          const void *crd, int crd_is_f64, int64_t ld_crd,
                          int64_t count, const int32_t *n_dev, void *out,
                          int64_t ld_out, void *workspace, size_t ws_bytes,
                          mv_stream_t stream);

/* As mv_beam_interp, but at given (az, za) arrays (f64, count) instead of
 * topocentric vectors: the reference's gpu_beam_interpolation(beam, daz, dza,
 * azmin, az, za) utility (src/matvis/gpu/beams.py:127-200). */
int mv_beam_interp_azza(const void *planes, int is_f64, int nbeam, int nfeed,
                        int nax, int nza, int naz, const double *grid,
                        int order, int is_power, const double *az,
                        const double *za, int64_t count, void *out,
                        int64_t ld_out, mv_stream_t stream);

/* Analytic Gaussian beam: amplitude exp(-za^2 / (2 sigma_b^2)), divided by
 * sqrt(2) on every (feed, axis) when polarized.  Replaces the host-side
 * analytic-beam evaluation of GPUBeamInterpolator.interp
 * (src/matvis/gpu/beams.py:95-101).  sigma: device (nbeam) radians.
 * out as for mv_beam_interp, precision out_is_f64. */
int mv_beam_gaussian(const double *sigma, int nbeam, int polarized,
                     const void *crd, int crd_is_f64, int64_t ld_crd,
                     int64_t count, const int32_t *n_dev, void *out,
                     int out_is_f64, int64_t ld_out, mv_stream_t stream);

/* ---------------------------------------------------------------- (c) ---- */

/* exptau[a, s] = exp(i 2 pi freq (b_a . s) / c).  Replaces TauCalculator
 * (src/matvis/core/tau.py:48-71).  antpos: device (nant, 3) f64 metres.
 * crd: (3, ld_crd) f64/f32.  out: (nant, ld_out) complex, out_is_f64.
 * The phase is formed in FP64 and reduced modulo one turn before sincos. */
int mv_exptau(const double *antpos, int nant, double freq, const void *crd,
              int crd_is_f64, int64_t ld_crd, int64_t nsrc, void *out,
              int out_is_f64, int64_t ld_out, mv_stream_t stream);

/* Z = A * sqrt_flux * exptau in the reference layout (nant*nfeed, nax*nsrc)
 * with row ant*nfeed+feed and column ax*nsrc+src.  Replaces ZMatrixCalc
 * (src/matvis/core/getz.py:52-105), including its in-place update
 * exptau *= sqrt_flux (core/getz.py:80).  beam: (nbeam, nfeed, nax, nsrc);
 * beam_idx: int32 (nant) [nullable: broadcast beam 0 or per-antenna when
 * nbeam == nant]. All complex arrays share precision is_f64. */
int mv_zmat_parts(const void *sqrt_flux, const void *beam,
                  const int32_t *beam_idx, int nbeam, void *exptau, int nant,
                  int nfeed, int nax, int64_t nsrc, void *z, int is_f64,
                  mv_stream_t stream);

/* Fused phase + Z: Z[a*nfeed+f, ax, s] = A[bidx(a), f, ax, s] * flux[s] *
 * exp(i 2 pi freq (b_a . crd_s) / c), written once, coalesced along s.
 * The fusion of TauCalculator and ZMatrixCalc
 * (src/matvis/core/tau.py:65-66, core/getz.py:80-100).
 * antpos: device (nant, 3) f64 metres.  crd: (3, ld_crd) f64.
 * flux: f64 sqrt(I/2); read as flux[idx[s]] when idx (int32, count, the
 * compaction's global source ids) is non-NULL, else as flux[s].
 * beam: (nbeam, nfeed, nax, ld_beam) complex (beam_is_f64).
 * form: MV_FORM_*; z0..z2 are the output arrays: z0 the complex matrix for
 * plain forms, row r at r*ldz (complex elements), axis segment ax at
 * ax*seg_stride.  For split forms the hi/mid/lo term arrays (nterms of
 * them), 32 bits per complex element, K-tile-major: the columns of all
 * segments (column s of segment ax is column ax*seg_stride + s) are cut
 * into K tiles of 16; tile q holds ldz rows (ldz >= nant*nfeed) of 64
 * bytes, row r at byte (q*ldz + r)*64; in a row, each 16-byte chunk of 4
 * consecutive columns holds the 4 real parts (16 bits each) and then the 4
 * imaginary parts.  Split forms need seg_stride a multiple of 16, so that
 * segments start on K tiles.  Columns s in [n, ceil(n/pad)*pad) are
 * zero-filled (pad = 32).
 * scale: multiplies Z before the split (a power of two; 1 for plain). */
int mv_zmat_fused(const double *antpos, int nant, double freq,
                  const double *crd, int64_t ld_crd, const double *flux,
                  const int32_t *idx, const void *beam, int beam_is_f64,
                  int64_t ld_beam,
                  int nbeam, const int32_t *beam_idx, int nfeed, int nax,
                  int64_t count, const int32_t *n_dev, int form, int nterms,
                  double scale, void *z0, void *z1, void *z2, int64_t ldz,
                  int64_t seg_stride, mv_stream_t stream);

/* Split a complex matrix (c128 or c64, layout as mv_zmat_fused output) into
 * nterms 16-bit terms of `form` (SPLIT_BF16 / SPLIT_F16), after multiplying
 * by `scale`, in the K-tile-major term layout of mv_zmat_fused (ldt rows a
 * K tile, ldt >= rows; with nseg > 1, tseg_stride a multiple of 16).  Rows
 * [0, rows), columns
 * [0, cols) of each of nseg segments; the slots of a segment's last chunk
 * past cols are written as zeros.
 * For Z built outside mv_zmat_fused, e.g. by the plugin chain's ZMatrixCalc
 * (src/matvis/core/getz.py:80-100) ahead of a split MatProd. */
int mv_split_z(const void *z, int z_is_f64, int64_t rows, int nseg,
               int64_t cols, int64_t ldz, int64_t seg_stride, int form,
               int nterms, double scale, void *t0, void *t1, void *t2,
               int64_t ldt, int64_t tseg_stride, mv_stream_t stream);

/* ---------------------------------------------------------------- (d) ---- */

/* Tile side (rows of V per workgroup tile) used by a HERK form: the
 * padding of the accumulator (ldv >= tile-rounded m).  No reference
 * counterpart (cuBLAS writes the full m x m product). */
int mv_herk_tile(int form);

/* Workspace bytes for mv_herk with room for `nslices` K-slices at row
 * count m (partial tiles of the split-K).  No reference counterpart. */
size_t mv_herk_workspace_bytes(int form, int m, int nslices);

/* Slice capacity that covers every slice count mv_herk may pick for up to
 * kmax columns (nseg * seg_len).  No reference counterpart: the reference's
 * cuBLAS call has no split-K workspace. */
int mv_herk_max_slices(int form, int m, int64_t kmax);

/* The slice count mv_herk uses for seg_len columns per segment given a
 * capacity of nslices (the kernels choose it on the device from the device
 * column count by the same rule): enough slices that no FP32 accumulation
 * chain exceeds the form's limit, then the one that best fills the GPU's
 * resident workgroup slots.  Returns the count, or a negative status.  No
 * reference counterpart (reporting and workspace planning). */
int mv_herk_slices(int form, int nterms, int nprod, int m, int nseg,
                   int64_t seg_len, int nslices);

/* Hermitian rank-k update over the upper triangle of tiles:
 * Vacc[r, c] (+)= scale * sum_k conj(Z[r, k]) Z[c, k]   for tile(r) <= tile(c)
 * Replaces GPUMatMul.compute -> zdotz -> cuBLAS zgemm/cgemm
 * (src/matvis/gpu/matprod.py:32-36, src/matvis/gpu/_cublas.py:12-65), which
 * computes the full m x m product.  K is split into slices (at most
 * `nslices`, the workspace capacity; see mv_herk_slices) whose partial tiles
 * land in the workspace and are summed in FP64 in a fixed order.
 * z0..z2: Z (plain) or its split terms (nterms, nprod significant products:
 * 3 = hi*hi + hi*lo + lo*hi, 4 adds lo*lo, 6 = the 3-term set up to order 2).
 * Z columns: nseg segments, each seg_len long (*seg_len_dev if non-NULL),
 * starting at seg*seg_stride (complex units).  Plain forms: ldz is the row
 * stride (complex units).  Split forms: terms K-tile-major as mv_zmat_fused
 * writes them, ldz the rows of a K tile (>= m), seg_stride a multiple of 16.
 * Split forms read whole tiles of 16 columns: every segment must have room
 * for round_up(seg_len, 16) columns and hold zeros from its column count
 * (*seg_len_dev when given) up to the next multiple of 16, as
 * mv_zmat_fused leaves it.
 * Split forms sum the diagonal V[r, r] exactly (FP64 over |t0 + t1 (+ t2)|^2)
 * rather than from the MFMA accumulators: the MFMA adder keeps a window of
 * about 2^-25 below its largest operand, so inside a long FP32 accumulation
 * the small term products vanish and every diagonal element would sit low
 * by their share (about 2^-18 for bf16); off the diagonal those products
 * average out.
 * Vacc: complex128 (ldv x ldv, row-major), ldv >= ntile_side*tile; holds
 * V[r, c] for r <= c (other elements are scratch).
 * accumulate: 0 overwrites the upper tiles, 1 adds to them. */
int mv_herk(int form, int nterms, int nprod, const void *z0, const void *z1,
            const void *z2, int m, int64_t ldz, int nseg, int64_t seg_stride,
            int64_t seg_len, const int32_t *seg_len_dev, int nslices,
            double scale, double *vacc, int64_t ldv, int accumulate,
            void *workspace, size_t ws_bytes, mv_stream_t stream);

/* mv_herk with no workspace argument and no allocation: the split-K
 * partials go to a 128 MiB device array that the library holds from load
 * until the process ends (one copy on each device it runs on).  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Uses as many K-slices as the array holds, at most
 * mv_herk's count; m whose partials for one slice exceed the array (split
 * forms m > 5632, complex128 m > 4032, complex64 m > 5760) is rejected with
 * MV_ERR_CAPACITY.  Calls that overlap on one device share the array and
 * race: order them (one stream, or events).  Other arguments as mv_herk. */
int mv_herk_noalloc(int form, int nterms, int nprod, const void *z0,
                    const void *z1, const void *z2, int m, int64_t ldz,
                    int nseg, int64_t seg_stride, int64_t seg_len,
                    const int32_t *seg_len_dev, double scale, double *vacc,
                    int64_t ldv, int accumulate, mv_stream_t stream);

/* mv_herk in its two launches, same arguments: the MFMA kernel that writes
 * the slice partials to the workspace, and the FP64 reduce into Vacc.  A
 * caller can record events between them (the compute kernel's own duration)
 * or overlap other work.  No reference counterpart (one cuBLAS call there). */
int mv_herk_compute(int form, int nterms, int nprod, const void *z0,
                    const void *z1, const void *z2, int m, int64_t ldz,
                    int nseg, int64_t seg_stride, int64_t seg_len,
                    const int32_t *seg_len_dev, int nslices, double scale,
                    double *vacc, int64_t ldv, int accumulate, void *workspace,
                    size_t ws_bytes, mv_stream_t stream);
int mv_herk_reduce(int form, int nterms, int nprod, const void *z0,
                   const void *z1, const void *z2, int m, int64_t ldz,
                   int nseg, int64_t seg_stride, int64_t seg_len,
                   const int32_t *seg_len_dev, int nslices, double scale,
                   double *vacc, int64_t ldv, int accumulate, void *workspace,
                   size_t ws_bytes, mv_stream_t stream);

/* One instance of a grouped split HERK (mv_herk_grouped; no reference
 * counterpart, the reference runs one matprod per (time, frequency, chunk),
 * src/matvis/gpu/gpu.py:240-250): plain integers
 * and device addresses so that a binding can pack an array of them
 * (104 bytes each; mv_herk_problem_bytes() reports the size).  Fields as in
 * mv_herk: z = term arrays (z[2] = 0 unless 3 terms), ldz / seg_stride /
 * seg_len as there (seg_len the host bound), seg_len_dev = device
 * int32 column count or 0, vacc = device complex128 (ldv x ldv), scale,
 * accumulate; nslices = slice capacity given to this problem's share of the
 * workspace (mv_herk_max_slices, or more); kchain = the longest K-slice in
 * K tiles (0: the form's cap, 192 tiles for the split forms and 64 for the
 * plain ones; larger values are clamped to it): the problem runs
 * ceil(K tiles / kchain) slices, at most nslices (shorter slices give a
 * small batch more work items to spread over the GPU, at the cost of more
 * partials to reduce); reserved = 0. */
typedef struct mv_herk_problem {
  uint64_t z[3];
  int64_t ldz, seg_stride, seg_len;
  uint64_t seg_len_dev, vacc;
  int64_t ldv;
  double scale;
  int32_t m, nseg, accumulate, nslices;
  int32_t kchain, reserved;
} mv_herk_problem;

size_t mv_herk_problem_bytes(void);

/* Workspace bytes for mv_herk_grouped over these problems (host copy).
 * No reference counterpart. */
size_t mv_herk_grouped_workspace_bytes(int form, int tile,
                                       const mv_herk_problem *problems,
                                       int nprob);

/* Many HERKs of one form and different sizes in one persistent launch (plus
 * a plan and a reduce launch): every problem's (tile, K-slice) items,
 * problem-major, are swept by one workgroup per resident slot, so small
 * problems share the GPU instead of each launch idling most of it.  Split
 * forms: `tile` = 128 or 64 (rows per workgroup tile; 64 suits m <= 128);
 * plain forms (MV_FORM_C128, MV_FORM_C64, nterms = 1, z[0] the complex Z):
 * `tile` = 64.  A problem's slice count depends on that problem alone, so
 * its result is the same in any group.  problems_host is read
 * for validation and launch geometry, problems_dev (the same array in device
 * memory) by the kernels, so a captured graph replays with whatever the
 * device array holds.  No reference counterpart: the reference runs one
 * cuBLAS call per (time, frequency, chunk). */
int mv_herk_grouped(int form, int nterms, int nprod, int tile,
                    const mv_herk_problem *problems_host,
                    const mv_herk_problem *problems_dev, int nprob,
                    void *workspace, size_t ws_bytes, mv_stream_t stream);

/* Sum nsrc_v accumulators elementwise into dst (all ldv x ldv c128):
 * dst = sum_i src_i.  The chunk sum of MatProd.sum_chunks
 * (src/matvis/core/matprod.py:103-115, src/matvis/gpu/matprod.py:49-51). */
int mv_vacc_sum(const double *src, int nsrc_v, int64_t stride, double *dst,
                int64_t count, mv_stream_t stream);

/* Expand the accumulator (V[r, c] for r <= c, V[c, r] = conj(V[r, c])) into
 * the API layout (npairs, nf, nf):
 * out[p, a, b] = V[i*nf + b, j*nf + a] for pair p = (i, j), the feed-swapped
 * layout of GPUMatMul.sum_chunks (src/matvis/gpu/matprod.py:38-63) and
 * CPUMatMul (src/matvis/cpu/matprod.py:24-30).  pairs: int32 (npairs, 2)
 * [nullable: all nant*nant ordered pairs, i-major].  out: c128 or c64. */
int mv_vis_expand(const double *vacc, int64_t ldv, int tile, int nant,
                  int nfeed, const int32_t *pairs, int64_t npairs, void *out,
                  int out_is_f64, mv_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MATVIS_HIP_H */
