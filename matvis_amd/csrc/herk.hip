// (d) V = conj(Z) Z^T as a Hermitian rank-k update on MFMA tiles.
//
// Only tiles (bi, bj) with bi <= bj of the upper triangle are computed, and
// inside them only the 64x64 (split) / 32x32 (plain) wave sub-tiles that hold
// elements (r, c) with r <= c < m; the accumulator is defined on r <= c.
//
// K (the source axis, times nax) is cut into slices.  A work item is one
// (tile, slice) pair; it writes its partial tile to the workspace, and a
// reduce kernel sums the slices in FP64 in a fixed order, so results are
// deterministic.  The number of active slices is chosen on the device from
// the device-side column count (sources above the horizon are only known
// there): enough slices that no FP32 accumulation chain exceeds `cmax` K
// tiles, then the count among the next 64 that best fills the resident
// workgroup slots.  The grid is at most one workgroup per slot; workgroups
// loop over items, each XCD owning a contiguous run of slice-major items so
// that workgroups sharing an L2 read the same K range of Z.
//
// Complex arithmetic on real MFMA: with X_r = (re_0, im_0, re_1, im_1, ...)
// the row of Z read as reals, and Y_r = (im_0, -re_0, im_1, -re_1, ...),
//   Re V[r, c] = X_r . X_c        Im V[r, c] = X_r . Y_c
// so each output tile needs two real products over 2K reals: 8 flop per
// complex multiply-add, with no extra passes over Z.  (The split forms'
// terms order the reals of each 4 columns as re_0..re_3, im_0..im_3: the dot
// products are the same sums, and Y is a register reorder of X.)
#include <map>
#include <mutex>
#include <utility>

#include "common.h"

namespace mv {
namespace {

constexpr int kXcd = 8;

__host__ __device__ inline int icdiv(int a, int b) { return (a + b - 1) / b; }

// Upper-triangle tile index -> (bi, bj) with bi <= bj, row-major over bi.
__device__ __forceinline__ void tile_coords(int t, int nb, int &bi, int &bj) {
  int i = 0;
  while (t >= nb - i) {
    t -= nb - i;
    ++i;
  }
  bi = i;
  bj = i + t;
}

// A 64 x 64 wave sub-tile (wr, wc) of tile (bi, bj), tile side ts, that the
// split kernels leave idle: below the diagonal of a diagonal tile, or
// starting at or past row or column m.  Idle sub-tiles get no MFMAs and no
// partials, and the reduce skips them: their accumulator elements lie below
// the diagonal or past m, which the accumulator's contract leaves as scratch.
// (Plain tiles are one 64-wide sub-tile that always starts below m.)
__host__ __device__ inline bool subtile_idle(int bi, int bj, int wr, int wc,
                                             int ts, int m) {
  return (bi == bj && wr > wc) || bi * ts + wr * 64 >= m ||
         bj * ts + wc * 64 >= m;
}

struct SliceCfg {
  int max_slices;  // partial slots per tile in the workspace
  int slots;       // workgroups resident at once (multiple of kXcd)
  int cmax;        // at most this many K tiles per slice
  int ovh;         // per-item cost beyond its K tiles, in K-tile units
};

// Active slice count for `ktiles` K tiles over `ntiles` tiles: enough
// slices that no chain exceeds cmax K tiles, then among the next 63 counts
// (at most the capacity and ktiles) the one with the least cost = rounds of
// resident slots x (K tiles per slice + per-item overhead); the smallest
// count on ties.  Pure integer function of its arguments: the compute and
// reduce kernels and the host agree on it.
__host__ __device__ inline void slice_range(int ktiles, const SliceCfg &c,
                                            int &lo, int &hi) {
  lo = icdiv(ktiles, c.cmax);
  lo = lo < 1 ? 1 : (lo > c.max_slices ? c.max_slices : lo);
  hi = lo + 63;
  if (hi > c.max_slices) hi = c.max_slices;
  if (hi > ktiles) hi = ktiles > lo ? ktiles : lo;
}
__host__ __device__ inline int64_t slice_cost(int ktiles, int ntiles, int ns,
                                              const SliceCfg &c) {
  const int per_x = c.slots / kXcd > 0 ? c.slots / kXcd : 1;
  const int rounds = icdiv(icdiv(ntiles * ns, kXcd), per_x);
  return (int64_t)rounds * (icdiv(ktiles, ns) + c.ovh);
}
__host__ __device__ inline int active_slices(int ktiles, int ntiles,
                                             const SliceCfg &c) {
  if (ktiles <= 0) return 1;
  int lo, hi;
  slice_range(ktiles, c, lo, hi);
  int best = lo;
  int64_t best_cost = INT64_MAX;
  for (int ns = lo; ns <= hi; ++ns) {
    const int64_t cost = slice_cost(ktiles, ntiles, ns, c);
    if (cost < best_cost) {
      best_cost = cost;
      best = ns;
    }
  }
  return best;
}
// The same value with one candidate per lane of a full wave (wave-uniform
// arguments): the least (cost, count) over the lanes.
__device__ inline int active_slices_wave(int ktiles, int ntiles,
                                         const SliceCfg &c) {
  if (ktiles <= 0) return 1;
  int lo, hi;
  slice_range(ktiles, c, lo, hi);
  const int lane = (int)(threadIdx.x & 63);
  const int ns = lo + lane;
  // cost < 2^48 (rounds and K tiles per slice are below 2^24); the lane in
  // the low bits breaks ties toward the smaller count
  uint64_t key = ns <= hi
                     ? ((uint64_t)slice_cost(ktiles, ntiles, ns, c) << 6) | lane
                     : UINT64_MAX;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(key, off, 64);
    key = o < key ? o : key;
  }
  return lo + (int)(key & 63);
}

// Columns of each segment: the device count never exceeds the host bound the
// buffers were sized for.
__device__ __forceinline__ int64_t seg_columns(int64_t host,
                                               const int32_t *dev) {
  if (!dev) return host;
  const int64_t d = *dev;
  return d < 0 ? 0 : (d < host ? d : host);
}

// Item loop: workgroup b runs on XCD x = b % 8, which holds w_x =
// ceil((g - x) / 8) of the g workgroups.  Each XCD owns a contiguous item
// range proportional to w_x (empty when w_x = 0) and its workgroups sweep it.
template <class F>
__device__ __forceinline__ void for_items(int64_t items, F &&body) {
  const int g = gridDim.x, b = blockIdx.x;
  const int x = b % kXcd, j = b / kXcd;
  int before = 0;
  for (int y = 0; y < x; ++y) before += (g - y + kXcd - 1) / kXcd;
  const int workers = (g - x + kXcd - 1) / kXcd;
  const int64_t lo = items * before / g;
  const int64_t hi = items * (before + workers) / g;
  for (int64_t it = lo + j; it < hi; it += workers) body(it);
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads() is a
// release/acquire fence for all address spaces and so also waits for every
// global load in flight; the K loops keep the next tiles' loads in flight
// across their barriers, and wait for them where the registers are used.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------ plain forms

template <typename T> struct PlainMfma;
// Both take A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k; they
// differ in where D lands: register r of lane l holds D[row(l, r)][l % 16].
template <> struct PlainMfma<double> {
  typedef v4d acc;
  __device__ static acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return g + 4 * r; }
};
template <> struct PlainMfma<float> {
  typedef v4f acc;
  __device__ static acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return 4 * g + r; }
};

constexpr int kPT = 64;  // tile side, complex rows
constexpr int kPK = 16;  // complex columns per LDS stage (one K tile)
constexpr int kPRow = kPK + 1;  // LDS row length (one element of padding)

// 256 threads = 4 waves in a 2x2 arrangement, each wave a 32x32 sub-tile made
// of 2x2 blocks of 16x16.  One 16x16x4 MFMA consumes two complex columns:
// lane l supplies k-slot g = l / 16 -> column g / 2, part g % 2 (re / im).
// One (slice, tile) item of a plain HERK: the slice's K tiles staged through
// two LDS stages (As for the tile's rows, Bs for its columns; separate
// __shared__ arrays of the calling kernel, which the compiler aligns for
// 16-byte LDS reads), the partial tile written to `part`.
template <typename T>
__device__ __forceinline__ void plain_item(
    const cplx<T> *__restrict__ z, int m, int64_t ldz, int64_t seg_stride,
    int64_t seg_len, int64_t kt_per_seg, int nb, int ntiles, int ktiles,
    int ns, int slice, int tile, cplx<T> *__restrict__ part,
    cplx<T> (&As)[2][kPT * kPRow], cplx<T> (&Bs)[2][kPT * kPRow]) {
  typedef typename PlainMfma<T>::acc acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lrow = tid >> 2, lcol = (tid & 3) * 4;
  const int g = lane >> 4, rl = lane & 15;
  const bool im_part = g & 1;
  int bi, bj;
  tile_coords(tile, nb, bi, bj);
  const bool diag = bi == bj;
  // sub-tiles below the diagonal or wholly in the row padding do no MFMA
  const bool active = !(diag && wr > wc) && bi * kPT + wr * 32 < m &&
                      bj * kPT + wc * 32 < m;
  const int64_t kt0 = (int64_t)slice * ktiles / ns;
  const int64_t kt1 = (int64_t)(slice + 1) * ktiles / ns;
  const int64_t grow_a = (int64_t)bi * kPT + lrow;
  const int64_t grow_b = (int64_t)bj * kPT + lrow;

  acc_t acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = acc_t{};
      acc_im[i][j] = acc_t{};
    }

  cplx<T> ra[4], rb[4];
  auto gload = [&](int64_t kt, int64_t grow, cplx<T> *r) {
    const int64_t sg = kt / kt_per_seg;
    const int64_t c0 = (kt % kt_per_seg) * kPK + lcol;
    const cplx<T> *src = z + grow * ldz + sg * seg_stride + c0;
    const bool row_ok = grow < m;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      r[q] = (row_ok && c0 + q < seg_len) ? src[q] : cplx<T>{(T)0, (T)0};
  };
  auto lstore = [&](cplx<T> *dst, const cplx<T> *r) {
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[lrow * kPRow + lcol + q] = r[q];
  };

  if (kt0 < kt1) {
    gload(kt0, grow_a, ra);
    lstore(As[0], ra);
    if (!diag) {
      gload(kt0, grow_b, rb);
      lstore(Bs[0], rb);
    }
  }
  lds_barrier();

  for (int64_t kt = kt0; kt < kt1; ++kt) {
    const int buf = (int)((kt - kt0) & 1);
    const bool has_next = kt + 1 < kt1;
    if (has_next) {
      gload(kt + 1, grow_a, ra);
      if (!diag) gload(kt + 1, grow_b, rb);
    }
    const cplx<T> *Ab = As[buf];
    const cplx<T> *Bb = diag ? As[buf] : Bs[buf];
    if (active) {
#pragma unroll
      for (int ks = 0; ks < kPK / 2; ++ks) {
        const int col = 2 * ks + (g >> 1);
        T a[2], bre[2], bim[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          cplx<T> v = Ab[(wr * 32 + i * 16 + rl) * kPRow + col];
          a[i] = im_part ? v.im : v.re;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cplx<T> v = Bb[(wc * 32 + j * 16 + rl) * kPRow + col];
          bre[j] = im_part ? v.im : v.re;
          bim[j] = im_part ? -v.re : v.im;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc_re[i][j] = PlainMfma<T>::op(a[i], bre[j], acc_re[i][j]);
            acc_im[i][j] = PlainMfma<T>::op(a[i], bim[j], acc_im[i][j]);
          }
      }
    }
    if (has_next) {
      lstore(As[buf ^ 1], ra);
      if (!diag) lstore(Bs[buf ^ 1], rb);
    }
    lds_barrier();
  }

  // 16x16 accumulator: lane l = 16 g + rl, register r -> row
  // PlainMfma<T>::row(g, r), column rl
  cplx<T> *out = part + ((int64_t)slice * ntiles + tile) * kPT * kPT;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr * 32 + i * 16 + PlainMfma<T>::row(g, r);
        const int col = wc * 32 + j * 16 + rl;
        out[row * kPT + col] = cplx<T>{acc_re[i][j][r], acc_im[i][j][r]};
      }
}

template <typename T>
__global__ __launch_bounds__(256) void herk_plain_kernel(
    const cplx<T> *__restrict__ z, int m, int64_t ldz, int nseg,
    int64_t seg_stride, int64_t seg_len_host,
    const int32_t *__restrict__ seg_len_dev, int nb, int ntiles,
    SliceCfg cfg, cplx<T> *__restrict__ part) {
  __shared__ cplx<T> As[2][kPT * kPRow];
  __shared__ cplx<T> Bs[2][kPT * kPRow];
  const int64_t seg_len = seg_columns(seg_len_host, seg_len_dev);
  const int64_t kt_per_seg = cdiv(seg_len, (int64_t)kPK);
  const int ktiles = (int)(kt_per_seg * nseg);
  const int ns = active_slices_wave(ktiles, ntiles, cfg);
  for_items((int64_t)ntiles * ns, [&](int64_t it) {
    plain_item<T>(z, m, ldz, seg_stride, seg_len, kt_per_seg, nb, ntiles,
                  ktiles, ns, (int)(it / ntiles), (int)(it % ntiles), part,
                  As, Bs);
  });
}

// ------------------------------------------------------------ split forms

constexpr int kST = 128;  // tile side, complex rows
constexpr int kSK = 16;   // complex columns per LDS stage (one K tile)

template <bool F16> struct SplitMfma;
template <> struct SplitMfma<false> {
  __device__ static v16f op(u32x4 a, u32x4 b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        __builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
  }
};
template <> struct SplitMfma<true> {
  __device__ static v16f op(u32x4 a, u32x4 b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(
        __builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0, 0);
  }
};

// The imaginary-part operand Y of a 16-byte chunk of 4 columns in the split
// terms' layout (the 4 real parts, then the 4 imaginary parts, 16 bits
// each): (re x 4 | im x 4) -> (im x 4 | -re x 4), a register reorder and two
// sign flips of packed pairs.
__device__ __forceinline__ u32x4 swap_negate4(u32x4 v) {
  return u32x4{v[2], v[3], v[0] ^ 0x80008000u, v[1] ^ 0x80008000u};
}



// The two 16-bit halves of a packed (re, im) word as floats.
template <bool F16> __device__ __forceinline__ float lo_half(uint32_t w) {
  if constexpr (F16)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
  else
    return __builtin_bit_cast(float, w << 16);
}
template <bool F16> __device__ __forceinline__ float hi_half(uint32_t w) {
  if constexpr (F16)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
  else
    return __builtin_bit_cast(float, w & 0xffff0000u);
}

// sum over the 8 complex values a thread stages of |t0 + t1 (+ t2)|^2, in
// FP64 (two bf16 terms add exactly in FP32; three round once at 2^-24): the
// sum of the squares of all 16 summed components, whichever column each
// 16-bit half of a word belongs to
template <int NT, bool F16>
__device__ __forceinline__ double row_norm2(const u32x4 (*r)[2]) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const uint32_t w = r[t][q >> 2][q & 3];
      re += lo_half<F16>(w);
      im += hi_half<F16>(w);
    }
    s += (double)re * re + (double)im * im;
  }
  return s;
}

// One HERK instance of a split form as the kernels read it.  `part` holds
// `nslices` x ntiles partial tiles of TILE x TILE complex<float>, `dpart`
// the FP64 row norms of the diagonal tiles (nslices x ntiles x TILE).
//
// Split terms are stored K-tile-major: K tile q (columns 16 q .. 16 q + 15
// of the concatenated segments, seg_stride a multiple of 16) is `ldz` rows
// of 16 words, row r at word (q ldz + r) 16.  The rows one copy
// instruction reads (16 consecutive rows of one K tile) are 1 KB of
// contiguous memory and a workgroup's panel of a K tile is one run of
// 128 x 64 bytes, where a row-major layout puts every row's 64 bytes on its
// own page.
struct SplitProblem {
  const uint32_t *t[3];
  int64_t ldz, seg_stride, seg_len_host;
  const int32_t *seg_len_dev;
  cplx<float> *part;
  double *dpart;
  cplx<double> *vacc;
  int64_t ldv;
  double scale;
  int m, nseg, nb, ntiles, accumulate;
};

// Block shape of the split kernels for a tile side: 128 -> 256 threads as
// 2x2 waves of 64x64; 64 -> one wave of 64x64.  Each thread stages PAIRS
// (row, half-row) pieces of 8 words of each operand panel per K tile.
template <int TILE> struct SplitShape {
  static constexpr int kThreads = TILE == 128 ? 256 : 64;
  static constexpr int kPairs = TILE * 2 / kThreads;
};

// LDS position of word w (a multiple of 4) of row r in a K-tile panel of
// 16 words a row: the row's 16-byte chunk c is stored at chunk c ^ swz(r),
// so 16 (or 8) consecutive rows accessed at one chunk, as the operand reads
// and the staging stores do, fall in 16 (or 8) distinct 16-byte bank groups.
__device__ __forceinline__ int swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
__device__ __forceinline__ int lds_word(int r, int w) {
  return r * kSK + (((w >> 2) ^ swz(r)) << 2);
}

// 16 bytes from global memory per lane straight into LDS: lane l's bytes
// land at byte 16 l of the wave-uniform destination `lds`.
__device__ __forceinline__ void copy16_to_lds(const uint32_t *src,
                                              uint32_t *lds) {
  __builtin_amdgcn_global_load_lds(src, lds, 16, 0, 0);
}
// Wait until this wave's global loads, LDS copies included, have landed.
__device__ __forceinline__ void wait_global_loads() {
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0), other counters unchecked
}

// One K-tile stage: the A rows (xa) and B rows (xb) of every term.  The two
// stages of a kernel are separate __shared__ objects, so that the compiler
// sees that loads in flight into one do not alias reads from the other.
template <int NT, int TILE> struct alignas(16) SplitStage {
  uint32_t xa[NT][TILE * kSK];
  uint32_t xb[NT][TILE * kSK];
};

// [x lanes 0-31 | y lanes 0-31] and [x lanes 32-63 | y lanes 32-63] of each
// register: the lane-half exchange of v_permlane32_swap.
__device__ __forceinline__ u32x4 halves_lo(u32x4 x, u32x4 y) {
  u32x4 r;
#pragma unroll
  for (int d = 0; d < 4; ++d)
    r[d] = __builtin_amdgcn_permlane32_swap(x[d], y[d], false, false)[0];
  return r;
}
__device__ __forceinline__ u32x4 halves_hi(u32x4 x, u32x4 y) {
  u32x4 r;
#pragma unroll
  for (int d = 0; d < 4; ++d)
    r[d] = __builtin_amdgcn_permlane32_swap(x[d], y[d], false, false)[1];
  return r;
}

// Waves compute 64x64 sub-tiles as 2x2 blocks of 32x32x16 MFMA, whose
// K = 16 reals are 8 from lanes 0-31 and 8 from lanes 32-63.
//
// Term products are grouped so that small ones share an instruction with a
// large one: a lane half h in {0, 1} supplies its term (a_h for A, b_h for B)
// over the same 4 complex columns, so one MFMA adds t_a0 t_b0 + t_a1 t_b1
// before its single FP32 rounding into the accumulator.
//   (2 terms, 3 products): [hi|lo] x [lo|hi] per 4 columns, and hi*hi over 8
//     columns, whose operands are the lane-half exchange of two [hi|lo]
//     (A) and [lo|hi] (B) fragments already in registers
//   (2 terms, 4 products): [hi|lo] x [hi|lo] and [hi|lo] x [lo|hi]
//   (3 terms, 6 products): [h|m] x [h|m], [h|m] x [m|h], [h|l] x [l|h]
// Each costs the same number of MFMAs as issuing the products separately.
// The imaginary-part operand Y = (im, -re) of a B fragment is formed in
// registers.
//
// K tiles go through two LDS stages.  Each wave copies its share of the next
// tile's rows straight from global memory into the other stage
// (global_load_lds_dwordx4: 16 bytes a lane, lane l to byte 16 l of a 1 KB
// run of 16 rows), then computes the current tile; one barrier per K tile
// publishes the copy and frees the stage just read.  Rows past m are read as
// row m - 1: they only reach output rows and columns past m, which nothing
// reads.  For a diagonal tile the threads also sum |t0 + t1 (+ t2)|^2 of the
// staged rows in FP64 (the exact diagonal).
template <int NT, int MAXORD, bool F16, int TILE>
__device__ __forceinline__ void split_item(const SplitProblem &P, int slice,
                                           int tile, int ns, int ktiles,
                                           int kps, int64_t seg_len,
                                           SplitStage<NT, TILE> &L0,
                                           SplitStage<NT, TILE> &L1) {
  typedef SplitShape<TILE> S;
  constexpr int kWaves = S::kThreads / 64;
  constexpr int kRuns = TILE / 16 / kWaves;  // 16-row runs a wave copies
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = TILE == 128 ? wave >> 1 : 0, wc = TILE == 128 ? wave & 1 : 0;
  const int lr = lane & 31, lh = lane >> 5;
  int bi, bj;
  tile_coords(tile, P.nb, bi, bj);
  const bool diag = bi == bj;
  const bool active = !subtile_idle(bi, bj, wr, wc, TILE, P.m);
  const int kt0 = (int)((int64_t)slice * ktiles / ns);
  const int kt1 = (int)((int64_t)(slice + 1) * ktiles / ns);
  // copy lanes: run g = wave * kRuns + u covers rows 16 g .. 16 g + 15;
  // lane l fills LDS chunk l % 4 of row 16 g + l / 4, i.e. logical chunk
  // (l % 4) ^ swz(row) of that row
  int64_t coff_a[kRuns], coff_b[kRuns];
#pragma unroll
  for (int u = 0; u < kRuns; ++u) {
    const int r = (wave * kRuns + u) * 16 + (lane >> 2);
    const int c = (lane & 3) ^ swz(r);
    const int ra = bi * TILE + r, rb = bj * TILE + r;
    const int ga = ra < P.m ? ra : P.m - 1, gb = rb < P.m ? rb : P.m - 1;
    coff_a[u] = (int64_t)ga * kSK + 4 * c;
    coff_b[u] = (int64_t)gb * kSK + 4 * c;
  }
  // diagonal sums: piece pi covers half lw of row lrow (8 words)
  int lrow[S::kPairs], lw[S::kPairs];
  double dsum[S::kPairs];
#pragma unroll
  for (int p = 0; p < S::kPairs; ++p) {
    const int pi = tid + p * S::kThreads, l = pi & 63;
    lrow[p] = (pi >> 6) * 32 + 16 * (l >> 5) + (l & 15);
    lw[p] = ((l >> 4) & 1) * 8;
    dsum[p] = 0.0;
  }

  v16f acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = v16f{};
      acc_im[i][j] = v16f{};
    }

  // position of the K tile to copy next: segment sg, tile kc within it;
  // whole K tiles: columns [seg_len, round_up(seg_len, 16)) of every
  // segment are zero by contract
  int sg = kt0 / (kps > 0 ? kps : 1), kc = kt0 - sg * kps;
  const int64_t seg_tiles = P.seg_stride / kSK;
  auto copy = [&](SplitStage<NT, TILE> &D) {
    const int64_t k0 = (sg * seg_tiles + kc) * P.ldz * kSK;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int u = 0; u < kRuns; ++u) {
        const int run = (wave * kRuns + u) * 16 * kSK;
        copy16_to_lds(P.t[t] + k0 + coff_a[u], &D.xa[t][run]);
        if (!diag) copy16_to_lds(P.t[t] + k0 + coff_b[u], &D.xb[t][run]);
      }
    if (++kc == kps) {
      kc = 0;
      ++sg;
    }
  };
  auto diag_sums = [&](const SplitStage<NT, TILE> &D) {
#pragma unroll
    for (int p = 0; p < S::kPairs; ++p) {
      u32x4 r[NT][2];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        r[t][0] = *(const u32x4 *)&D.xa[t][lds_word(lrow[p], lw[p])];
        r[t][1] = *(const u32x4 *)&D.xa[t][lds_word(lrow[p], lw[p] + 4)];
      }
      dsum[p] += row_norm2<NT, F16>(r);
    }
  };
  // the MFMAs of one K tile from stage D
  auto compute = [&](const SplitStage<NT, TILE> &D) {
    const uint32_t(*Ax)[TILE * kSK] = D.xa;
    const uint32_t(*Bx)[TILE * kSK] = diag ? D.xa : D.xb;
    // lane half lh reads term t_lh at word offset w_lh of its row
    auto ld_a = [&](int t0, int t1, int w0, int w1, u32x4 (&a)[2]) {
      const uint32_t *base = Ax[lh ? t1 : t0];
      const int w = lh ? w1 : w0;
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *(const u32x4 *)&base[lds_word(wr * 64 + i * 32 + lr, w)];
    };
    auto ld_b = [&](int t0, int t1, int w0, int w1, u32x4 (&b)[2]) {
      const uint32_t *base = Bx[lh ? t1 : t0];
      const int w = lh ? w1 : w0;
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *(const u32x4 *)&base[lds_word(wc * 64 + j * 32 + lr, w)];
    };
    auto mma = [&](const u32x4 (&a)[2], const u32x4 (&b)[2]) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const u32x4 y = swap_negate4(b[j]);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          acc_re[i][j] = SplitMfma<F16>::op(a[i], b[j], acc_re[i][j]);
          acc_im[i][j] = SplitMfma<F16>::op(a[i], y, acc_im[i][j]);
        }
      }
    };
    u32x4 a[2], b[2];
    if constexpr (NT == 2 && MAXORD == 1) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        u32x4 a1[2], b1[2];
        ld_a(0, 1, 8 * ks, 8 * ks, a);  // [hi|lo] of columns 8ks..8ks+3
        ld_b(1, 0, 8 * ks, 8 * ks, b);  // [lo|hi]
        mma(a, b);
        ld_a(0, 1, 8 * ks + 4, 8 * ks + 4, a1);
        ld_b(1, 0, 8 * ks + 4, 8 * ks + 4, b1);
        mma(a1, b1);
        // hi of columns 8ks..8ks+3 in lanes 0-31 and of 8ks+4..8ks+7 in
        // lanes 32-63, for A and B alike
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[i] = halves_lo(a[i], a1[i]);
          b[i] = halves_hi(b[i], b1[i]);
        }
        mma(a, b);
      }
    } else if constexpr (NT == 2) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ld_a(0, 1, 4 * q, 4 * q, a);
        ld_b(0, 1, 4 * q, 4 * q, b);
        mma(a, b);
        ld_b(1, 0, 4 * q, 4 * q, b);
        mma(a, b);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ld_a(0, 1, 4 * q, 4 * q, a);
        ld_b(0, 1, 4 * q, 4 * q, b);
        mma(a, b);
        ld_b(1, 0, 4 * q, 4 * q, b);
        mma(a, b);
        ld_a(0, 2, 4 * q, 4 * q, a);
        ld_b(2, 0, 4 * q, 4 * q, b);
        mma(a, b);
      }
    }
  };
  // one K tile computed from stage C while the next is copied into N
  auto step = [&](SplitStage<NT, TILE> &C, SplitStage<NT, TILE> &N, int kt) {
    if (kt + 1 < kt1) copy(N);
    if (diag) diag_sums(C);
    if (active) {
      // issue priority to waves in their MFMA phase over waves copying or
      // waiting at the barrier
      __builtin_amdgcn_s_setprio(1);
      compute(C);
      __builtin_amdgcn_s_setprio(0);
    }
    wait_global_loads();  // this wave's copies into N
    lds_barrier();
  };

  if (kt0 < kt1) copy(L0);
  wait_global_loads();
  lds_barrier();
  for (int kt = kt0; kt < kt1; kt += 2) {
    step(L0, L1, kt);
    if (kt + 1 < kt1) step(L1, L0, kt + 1);
  }
  // every wave is past the last barrier: LDS is free for the next item
  if (diag) {
    // two threads share each row: lanes l and l ^ 16 of one wave
#pragma unroll
    for (int p = 0; p < S::kPairs; ++p) {
      const double s = dsum[p] + __shfl_xor(dsum[p], 16, 64);
      if (lw[p] == 0)
        P.dpart[((int64_t)slice * P.ntiles + tile) * TILE + lrow[p]] = s;
    }
  }

  if (!active) return;  // an idle sub-tile has no partials
  // 32x32 accumulator layout: lane l, register r ->
  //   row 8*(r/4) + 4*(l/32) + r%4, col l%32.
  // The thread's base offset is opaque to the optimizer so that the store
  // addresses are formed here rather than kept live across the K loop.
  cplx<float> *out = P.part + ((int64_t)slice * P.ntiles + tile) * TILE * TILE;
  int obase = (wr * 64 + 4 * lh) * TILE + wc * 64 + lr;
  asm volatile("" : "+v"(obase));
  out += obase;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = i * 32 + 8 * (r >> 2) + (r & 3);
        out[row * TILE + j * 32] =
            cplx<float>{acc_re[i][j][r], acc_im[i][j][r]};
      }
}

// One split HERK: K tiles and the slice count from the device column count,
// items (tile, slice) swept per XCD.
template <int NT, int MAXORD, bool F16, int TILE>
__global__ __launch_bounds__(SplitShape<TILE>::kThreads,
                             TILE == 128 ? 2 : 8) void herk_split_kernel(
    SplitProblem P, SliceCfg cfg) {
  __shared__ SplitStage<NT, TILE> L0, L1;
  const int64_t seg_len = seg_columns(P.seg_len_host, P.seg_len_dev);
  const int kps = (int)cdiv(seg_len, (int64_t)kSK);  // K tiles per segment
  const int ktiles = kps * P.nseg;
  const int ns = active_slices_wave(ktiles, P.ntiles, cfg);
  for_items((int64_t)P.ntiles * ns, [&](int64_t it) {
    split_item<NT, MAXORD, F16, TILE>(P, (int)(it / P.ntiles),
                                      (int)(it % P.ntiles), ns, ktiles, kps,
                                      seg_len, L0, L1);
  });
}

// ------------------------------------------------------------ grouped

// Per-launch plan of a grouped HERK, written by herk_group_plan_kernel into
// the head of the workspace: slice counts, item and tile prefix offsets.
struct GroupPlan {
  int32_t *ns;        // [nprob]
  int64_t *item_off;  // [nprob + 1]
  int32_t *tile_off;  // [nprob + 1]
};

// Single block: every problem's internal descriptor (tile grid, partial
// areas carved from `parts` in problem order: cap x ntiles x ts^2 partial
// elements of part_bytes each, then, with row_norms, cap x ntiles x ts FP64
// row norms), its slice count (enough slices that no chain exceeds cmax K
// tiles, at most its capacity: a function of the problem alone) and the
// prefix offsets.
__global__ __launch_bounds__(256) void herk_group_plan_kernel(
    const mv_herk_problem *__restrict__ pub, int nprob, int ts, int kc,
    int cmax, int part_bytes, int row_norms, SplitProblem *__restrict__ probs,
    char *__restrict__ parts, GroupPlan plan) {
  for (int p = threadIdx.x; p < nprob; p += blockDim.x) {
    const mv_herk_problem &q = pub[p];
    SplitProblem P;
    for (int t = 0; t < 3; ++t) P.t[t] = (const uint32_t *)q.z[t];
    P.ldz = q.ldz;
    P.seg_stride = q.seg_stride;
    P.seg_len_host = q.seg_len;
    P.seg_len_dev = (const int32_t *)q.seg_len_dev;
    P.vacc = (cplx<double> *)q.vacc;
    P.ldv = q.ldv;
    P.scale = q.scale;
    P.m = q.m;
    P.nseg = q.nseg;
    P.accumulate = q.accumulate;
    P.nb = icdiv(q.m, ts);
    P.ntiles = P.nb * (P.nb + 1) / 2;
    P.part = nullptr;
    P.dpart = nullptr;
    probs[p] = P;
    const int64_t seg_len = seg_columns(q.seg_len, P.seg_len_dev);
    const int ktiles = (int)(cdiv(seg_len, (int64_t)kc) * q.nseg);
    const int chain = q.kchain > 0 && q.kchain < cmax ? q.kchain : cmax;
    int ns = ktiles > 0 ? icdiv(ktiles, chain) : 1;
    ns = ns < 1 ? 1 : (ns > q.nslices ? q.nslices : ns);
    plan.ns[p] = ns;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t items = 0, bytes = 0;
    int32_t tiles = 0;
    for (int p = 0; p < nprob; ++p) {
      SplitProblem &P = probs[p];
      const int64_t cap = pub[p].nslices;
      P.part = (cplx<float> *)(parts + bytes);
      bytes += cap * P.ntiles * (int64_t)ts * ts * part_bytes;
      P.dpart = row_norms ? (double *)(parts + bytes) : nullptr;
      if (row_norms) bytes += cap * P.ntiles * (int64_t)ts * 8;
      plan.item_off[p] = items;
      plan.tile_off[p] = tiles;
      items += (int64_t)P.ntiles * plan.ns[p];
      tiles += P.ntiles;
    }
    plan.item_off[nprob] = items;
    plan.tile_off[nprob] = tiles;
  }
}

// index p with off[p] <= x < off[p + 1]
template <typename T>
__device__ __forceinline__ int find_segment(const T *off, int n, T x) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// All problems' items in one persistent launch, problem-major then
// slice-major, swept per XCD as in the single-problem kernel.
template <int NT, int MAXORD, bool F16, int TILE>
__global__ __launch_bounds__(SplitShape<TILE>::kThreads,
                             TILE == 128 ? 2 : 8) void herk_split_grouped_kernel(
    const SplitProblem *__restrict__ probs, int nprob, GroupPlan plan) {
  __shared__ SplitStage<NT, TILE> L0, L1;
  const int64_t items = plan.item_off[nprob];
  for_items(items, [&](int64_t it) {
    const int p = find_segment(plan.item_off, nprob, it);
    const SplitProblem P = probs[p];
    const int ns = plan.ns[p];
    const int64_t local = it - plan.item_off[p];
    const int64_t seg_len = seg_columns(P.seg_len_host, P.seg_len_dev);
    const int kps = (int)cdiv(seg_len, (int64_t)kSK);
    split_item<NT, MAXORD, F16, TILE>(P, (int)(local / P.ntiles),
                                      (int)(local % P.ntiles), ns,
                                      kps * P.nseg, kps, seg_len, L0, L1);
  });
}

// Grouped plain HERK: the items of all problems, as the grouped split
// kernel; each problem's Z is its term pointer 0.
template <typename T>
__global__ __launch_bounds__(256) void herk_plain_grouped_kernel(
    const SplitProblem *__restrict__ probs, int nprob, GroupPlan plan) {
  __shared__ cplx<T> As[2][kPT * kPRow];
  __shared__ cplx<T> Bs[2][kPT * kPRow];
  const int64_t items = plan.item_off[nprob];
  for_items(items, [&](int64_t it) {
    const int p = find_segment(plan.item_off, nprob, it);
    const SplitProblem P = probs[p];
    const int ns = plan.ns[p];
    const int64_t local = it - plan.item_off[p];
    const int64_t seg_len = seg_columns(P.seg_len_host, P.seg_len_dev);
    const int64_t kt_per_seg = cdiv(seg_len, (int64_t)kPK);
    plain_item<T>((const cplx<T> *)P.t[0], P.m, P.ldz, P.seg_stride, seg_len,
                  kt_per_seg, P.nb, P.ntiles, (int)(kt_per_seg * P.nseg), ns,
                  (int)(local / P.ntiles), (int)(local % P.ntiles),
                  (cplx<T> *)P.part, As, Bs);
  });
}

// ------------------------------------------------------------ epilogues

// Elements of upper tile `tile`: the sum of their `ns` slice partials in
// FP64 (fixed order), scaled, stored or added into the accumulator.  With
// `dpart` (split forms) the diagonal elements of diagonal tiles take the row
// norms summed there instead of the MFMA sums.
//
// A wave takes 32 groups of kReduceEpt adjacent elements of one row; each
// group's sum over slices s = 8q + j is eight independent chains j, chains
// 0-3 in lanes 0-31 and chains 4-7 in lanes 32-63, so that each lane keeps
// four loads in flight and a launch has twice the workgroups of one lane per
// group.  The chains combine in a fixed order, ((c0 + c1) + (c2 + c3)) +
// ((c4 + c5) + (c6 + c7)); lane half h then stores element h of the group.
// (Four elements a lane halve the grid and measured slower.)
template <typename TP> constexpr int kReduceEpt = 2;

// The first element of the calling lane's group and its lane half.
template <typename TP>
__device__ __forceinline__ int64_t reduce_group(int &half) {
  const int lane = (int)(threadIdx.x & 63);
  half = lane >> 5;
  return kReduceEpt<TP> * ((int64_t)blockIdx.x * 128 +
                           (int64_t)(threadIdx.x >> 6) * 32 + (lane & 31));
}

template <typename TP, bool DTILE>
__device__ __forceinline__ void reduce_elems_t(
    const cplx<TP> *__restrict__ part, const double *__restrict__ dpart,
    int ntiles, int bi, int bj, int ts, int m, int tile, int64_t e, int h,
    int ns, double scale, cplx<double> *__restrict__ vacc, int64_t ldv,
    int accumulate) {
  constexpr int E = kReduceEpt<TP>;
  static_assert(E == 2, "lane half h stores element h of its group");
  const int64_t tsz = (int64_t)ts * ts;
  const int64_t r = e / ts, c = e % ts;
  // the group's E adjacent columns share a sub-tile (E divides 64); both
  // lane halves of a group leave together
  if (subtile_idle(bi, bj, (int)(r >> 6), (int)(c >> 6), ts, m)) return;
  const cplx<TP> *src = part + (int64_t)tile * tsz + e;
  const int64_t step = (int64_t)ntiles * tsz;
  // diagonal tiles (DTILE) also sum the row norms of the group's row, in the
  // same loop (every lane of the tile, so no lane runs extra rounds of
  // dependent loads after it)
  constexpr bool dtile = DTILE;
  const double *dp = dtile ? dpart + (int64_t)tile * ts + r : nullptr;
  const int64_t dstep = (int64_t)ntiles * ts;
  double re[E][4], im[E][4], dn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    dn[j] = 0.0;
#pragma unroll
    for (int k = 0; k < E; ++k) re[k][j] = im[k][j] = 0.0;
  }
  int s = 0;
  for (; s + 8 <= ns; s += 8) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t sl = s + 4 * h + j;  // chain 4 h + j
      const cplx<TP> *q = src + sl * step;
      cplx<TP> v[E];
#pragma unroll
      for (int k = 0; k < E; ++k) v[k] = q[k];
      if constexpr (dtile) dn[j] += dp[sl * dstep];
#pragma unroll
      for (int k = 0; k < E; ++k) {
        re[k][j] += (double)v[k].re;
        im[k][j] += (double)v[k].im;
      }
    }
  }
  if (h == 0) {
    // the last ns % 8 slices join chain 0
    for (; s < ns; ++s) {
      const cplx<TP> *q = src + (int64_t)s * step;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        re[k][0] += (double)q[k].re;
        im[k][0] += (double)q[k].im;
      }
      if constexpr (dtile) dn[0] += dp[(int64_t)s * dstep];
    }
  }
  // this half's four chains, then (chains 0-3) + (chains 4-7) in both
  // halves (the same sum: an IEEE add commutes)
  auto both = [&](const double (&x)[4]) {
    const double a = (x[0] + x[1]) + (x[2] + x[3]);
    const double b = __shfl_xor(a, 32, 64);
    return h == 0 ? a + b : b + a;
  };
  double sr[E], si[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    sr[k] = both(re[k]);
    si[k] = both(im[k]);
  }
  double vr = h == 0 ? sr[0] : sr[1];
  double vi = h == 0 ? si[0] : si[1];
  if constexpr (dtile) {
    const double d = both(dn);
    if (r == c + h) {
      // the diagonal: the FP64 sum of |terms|^2 the compute kernel kept
      vr = d;
      vi = 0.0;
    }
  }
  cplx<double> *dst =
      vacc + ((int64_t)bi * ts + r) * ldv + (int64_t)bj * ts + c + h;
  if (accumulate) {
    dst->re += vr * scale;
    dst->im += vi * scale;
  } else {
    *dst = cplx<double>{vr * scale, vi * scale};
  }
}

template <typename TP>
__device__ __forceinline__ void reduce_elems(
    const cplx<TP> *__restrict__ part, const double *__restrict__ dpart,
    int ntiles, int nb, int ts, int m, int tile, int64_t e, int h, int ns,
    double scale, cplx<double> *__restrict__ vacc, int64_t ldv,
    int accumulate) {
  int bi, bj;
  tile_coords(tile, nb, bi, bj);
  if (dpart && bi == bj)  // uniform over the block: one tile a block
    reduce_elems_t<TP, true>(part, dpart, ntiles, bi, bj, ts, m, tile, e, h,
                             ns, scale, vacc, ldv, accumulate);
  else
    reduce_elems_t<TP, false>(part, dpart, ntiles, bi, bj, ts, m, tile, e, h,
                              ns, scale, vacc, ldv, accumulate);
}

// Blocks of 256 threads (128 element groups), per tile.
template <typename TP> dim3 reduce_grid(int ts, int64_t ntiles) {
  return dim3((unsigned)cdiv((int64_t)ts * ts,
                             (int64_t)128 * kReduceEpt<TP>),
              (unsigned)ntiles);
}

// grid reduce_grid<TP>(ts, ntiles); the slice count is found once per block
template <typename TP>
__global__ __launch_bounds__(256) void herk_reduce_kernel(
    const cplx<TP> *__restrict__ part, const double *__restrict__ dpart,
    int ntiles, int nb, int ts, int m, int64_t seg_len_host,
    const int32_t *__restrict__ seg_len_dev, int nseg, int kc, SliceCfg cfg,
    double scale, cplx<double> *__restrict__ vacc, int64_t ldv,
    int accumulate) {
  __shared__ int ns_shared;
  if (threadIdx.x < 64) {
    const int64_t seg_len = seg_columns(seg_len_host, seg_len_dev);
    const int ktiles = (int)(cdiv(seg_len, (int64_t)kc) * nseg);
    const int ns = active_slices_wave(ktiles, ntiles, cfg);
    if (threadIdx.x == 0) ns_shared = ns;
  }
  __syncthreads();
  int h;
  const int64_t e = reduce_group<TP>(h);
  if (e >= (int64_t)ts * ts) return;
  reduce_elems(part, dpart, ntiles, nb, ts, m, (int)blockIdx.y, e, h,
               ns_shared, scale, vacc, ldv, accumulate);
}

// Grouped reduce: blockIdx.y runs over the tiles of all problems (prefix
// offsets from the plan); lanes and element groups as above.
template <typename TP>
__global__ __launch_bounds__(256) void herk_group_reduce_kernel(
    const SplitProblem *__restrict__ probs, int nprob, int ts,
    GroupPlan plan) {
  int h;
  const int64_t e = reduce_group<TP>(h);
  if (e >= (int64_t)ts * ts) return;
  // the grid covers the tiles of the problems' largest shapes (a graph
  // replayed with smaller ones keeps it); tiles past this plan's total idle
  if ((int32_t)blockIdx.y >= plan.tile_off[nprob]) return;
  const int p = find_segment(plan.tile_off, nprob, (int32_t)blockIdx.y);
  const SplitProblem &P = probs[p];
  reduce_elems((const cplx<TP> *)P.part, P.dpart, P.ntiles, P.nb, ts, P.m,
               (int)blockIdx.y - plan.tile_off[p], e, h, plan.ns[p], P.scale,
               P.vacc, P.ldv, P.accumulate);
}

__global__ __launch_bounds__(256) void vacc_sum_kernel(
    const double *__restrict__ src, int nsrc_v, int64_t stride,
    double *__restrict__ dst, int64_t count) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= count) return;
  double s = 0.0;
  for (int i = 0; i < nsrc_v; ++i) s += src[(int64_t)i * stride + e];
  dst[e] = s;
}

// out[p, a, b] = V[i nf + b, j nf + a] from the r <= c half of the
// accumulator, V[r, c] = conj(V[c, r]) for the other half.
template <typename TO>
__global__ __launch_bounds__(256) void expand_kernel(
    const cplx<double> *__restrict__ vacc, int64_t ldv, int nant, int nf,
    const int32_t *__restrict__ pairs, int64_t npairs,
    cplx<TO> *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int per = nf * nf;
  if (idx >= npairs * per) return;
  const int64_t p = idx / per;
  const int a = (int)((idx / nf) % nf), b = (int)(idx % nf);
  int i, j;
  if (pairs) {
    i = pairs[2 * p];
    j = pairs[2 * p + 1];
  } else {
    i = (int)(p / nant);
    j = (int)(p % nant);
  }
  if (i < 0 || i >= nant || j < 0 || j >= nant) {
    out[idx] = cplx<TO>{(TO)0, (TO)0};
    return;
  }
  const int64_t r = (int64_t)i * nf + b, c = (int64_t)j * nf + a;
  cplx<double> v;
  if (r <= c) {
    v = vacc[r * ldv + c];
  } else {
    v = vacc[c * ldv + r];
    v.im = -v.im;
  }
  out[idx] = cplx<TO>{(TO)v.re, (TO)v.im};
}

int split_maxord(int nterms, int nprod) {
  if (nterms == 2 && nprod == 3) return 1;
  if (nterms == 2 && nprod == 4) return 2;
  if (nterms == 3 && nprod == 6) return 2;
  return -1;
}

// Longest FP32 accumulation chain per slice, in K tiles of 16 columns.  The
// split forms' diagonal is summed exactly, and off the diagonal the FP32
// chain error stays near 1e-7 relative at 62 tiles and grows about as the
// square root of the length; 192 tiles keeps it well under 1e-6 while the
// partial tiles written per launch stay small.
int form_cmax(int form) {
  switch (form) {
    case MV_FORM_C128: return 1 << 30;  // FP64 accumulation: no limit
    case MV_FORM_C64: return 256;
    default: return 192;
  }
}

// Resident workgroups of `kernel` on the current device (CUs x blocks per
// CU), rounded down to a multiple of the XCD count; cached per device.
template <class K> int kernel_slots(K kernel, int block = 256) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  const std::pair<const void *, int> key((const void *)kernel, dev);
  std::lock_guard<std::mutex> lock(mu);
  auto f = cache.find(key);
  if (f != cache.end()) return f->second;
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) !=
          hipSuccess)
    return -1;
  int slots = cus * (per > 0 ? per : 1);
  slots = slots >= kXcd ? slots / kXcd * kXcd : kXcd;
  cache[key] = slots;
  return slots;
}

struct HerkPlan {
  int ts, nb, ntiles, kc;
  SliceCfg cfg;
  int grid;
};

}  // namespace
}  // namespace mv

using namespace mv;

extern "C" int mv_herk_tile(int form) {
  switch (form) {
    case MV_FORM_C128:
    case MV_FORM_C64: return kPT;
    case MV_FORM_SPLIT_BF16:
    case MV_FORM_SPLIT_F16: return kST;
    default: return 0;
  }
}

namespace {
// split forms: FP32 partial tiles, then FP64 row norms of diagonal tiles
size_t split_partial_bytes(int m, int nslices) {
  const int64_t nb = cdiv((int64_t)m, (int64_t)kST);
  return (size_t)nslices * (nb * (nb + 1) / 2) * kST * kST * 8;
}
}  // namespace

extern "C" size_t mv_herk_workspace_bytes(int form, int m, int nslices) {
  const int ts = mv_herk_tile(form);
  if (ts == 0 || m <= 0 || nslices <= 0) return 0;
  const int64_t nb = cdiv((int64_t)m, (int64_t)ts);
  const int64_t ntiles = nb * (nb + 1) / 2;
  if (ts == kST)
    return split_partial_bytes(m, nslices) + (size_t)nslices * ntiles * kST * 8;
  const size_t elem = form == MV_FORM_C128 ? 16 : 8;
  return (size_t)nslices * ntiles * ts * ts * elem;
}

extern "C" int mv_herk_max_slices(int form, int m, int64_t kmax) {
  const int ts = mv_herk_tile(form);
  if (ts == 0 || m <= 0 || kmax < 0) return 0;
  const int64_t ktiles = cdiv(kmax, (int64_t)16);
  if (ktiles == 0) return 1;
  const int64_t lo = cdiv(ktiles, (int64_t)form_cmax(form));
  const int64_t ns = lo + 64 < ktiles ? lo + 64 : ktiles;
  return (int)(ns < 1 ? 1 : ns);
}

namespace {
// The kernel instantiation for (form, nterms, maxord) as a type-erased
// pointer for occupancy queries.
const void *herk_kernel(int form, int nterms, int maxord) {
  switch (form) {
    case MV_FORM_C128: return (const void *)&herk_plain_kernel<double>;
    case MV_FORM_C64: return (const void *)&herk_plain_kernel<float>;
    default: break;
  }
  const bool f16 = form == MV_FORM_SPLIT_F16;
  if (nterms == 2 && maxord == 1)
    return f16 ? (const void *)&herk_split_kernel<2, 1, true, kST>
               : (const void *)&herk_split_kernel<2, 1, false, kST>;
  if (nterms == 2)
    return f16 ? (const void *)&herk_split_kernel<2, 2, true, kST>
               : (const void *)&herk_split_kernel<2, 2, false, kST>;
  return f16 ? (const void *)&herk_split_kernel<3, 2, true, kST>
             : (const void *)&herk_split_kernel<3, 2, false, kST>;
}

// The grouped kernel instantiation for (form, nterms, maxord, tile).
template <int TILE>
const void *grouped_kernel_t(bool f16, int nterms, int maxord) {
  if (nterms == 2 && maxord == 1)
    return f16 ? (const void *)&herk_split_grouped_kernel<2, 1, true, TILE>
               : (const void *)&herk_split_grouped_kernel<2, 1, false, TILE>;
  if (nterms == 2)
    return f16 ? (const void *)&herk_split_grouped_kernel<2, 2, true, TILE>
               : (const void *)&herk_split_grouped_kernel<2, 2, false, TILE>;
  return f16 ? (const void *)&herk_split_grouped_kernel<3, 2, true, TILE>
             : (const void *)&herk_split_grouped_kernel<3, 2, false, TILE>;
}

int plan_herk(int form, int nterms, int maxord, int m, int nslices,
              HerkPlan &p) {
  p.ts = mv_herk_tile(form);
  p.nb = (int)cdiv((int64_t)m, (int64_t)p.ts);
  p.ntiles = p.nb * (p.nb + 1) / 2;
  p.kc = p.ts == kPT ? kPK : kSK;
  const int slots = kernel_slots(herk_kernel(form, nterms, maxord));
  if (slots <= 0) return -1;
  p.cfg = SliceCfg{nslices, slots, form_cmax(form),
                   form == MV_FORM_C128 || form == MV_FORM_C64 ? 2 : 4};
  const int64_t items = (int64_t)p.ntiles * nslices;
  p.grid = (int)(items < slots ? items : slots);
  return 0;
}
}  // namespace

extern "C" int mv_herk_slices(int form, int nterms, int nprod, int m,
                              int nseg, int64_t seg_len, int nslices) {
  const int maxord = mv_herk_tile(form) == kST ? split_maxord(nterms, nprod)
                                               : 0;
  if (mv_herk_tile(form) == 0 || maxord < 0 || m < 1 || nseg < 1 ||
      seg_len < 0 || nslices < 1)
    return MV_ERR_ARG;
  HerkPlan p;
  if (plan_herk(form, nterms, maxord, m, nslices, p))
    return MV_ERR_HIP;
  const int ktiles = (int)(cdiv(seg_len, (int64_t)p.kc) * nseg);
  return active_slices(ktiles, p.ntiles, p.cfg);
}

namespace {
constexpr int kPhaseCompute = 1, kPhaseReduce = 2;

// Validation shared by every phase, then the compute kernel (partial tiles
// into the workspace) and/or the reduce kernel (workspace into Vacc).
int herk_run(int phases, int form, int nterms, int nprod, const void *z0,
             const void *z1, const void *z2, int m, int64_t ldz, int nseg,
             int64_t seg_stride, int64_t seg_len, const int32_t *seg_len_dev,
             int nslices, double scale, double *vacc, int64_t ldv,
             int accumulate, void *workspace, size_t ws_bytes,
             mv_stream_t stream) {
  const int ts = mv_herk_tile(form);
  MV_REQUIRE(ts > 0, "mv_herk: unknown form %d", form);
  MV_REQUIRE(z0 && vacc && workspace, "mv_herk: null pointer");
  MV_REQUIRE(m >= 1 && nseg >= 1 && seg_len >= 0 && nslices >= 1,
             "mv_herk: bad sizes m=%d nseg=%d seg_len=%lld nslices=%d", m,
             nseg, (long long)seg_len, nslices);
  MV_REQUIRE(nseg == 1 || seg_stride >= seg_len,
             "mv_herk: segments overlap (seg_stride %lld < seg_len %lld)",
             (long long)seg_stride, (long long)seg_len);
  MV_REQUIRE(ldz >= (form == MV_FORM_C128 || form == MV_FORM_C64
                         ? (int64_t)(nseg - 1) * seg_stride + seg_len
                         : (int64_t)m),
             "mv_herk: ldz %lld too small", (long long)ldz);
  MV_REQUIRE(cdiv(seg_len, (int64_t)16) * nseg < (int64_t)INT32_MAX / 2,
             "mv_herk: K too large");
  const int64_t nb = cdiv((int64_t)m, (int64_t)ts);
  const int64_t ntiles = nb * (nb + 1) / 2;
  MV_REQUIRE(ldv >= nb * ts, "mv_herk: ldv %lld < %lld (tile-padded m)",
             (long long)ldv, (long long)(nb * ts));
  MV_REQUIRE(ntiles * nslices < (int64_t)INT32_MAX / 2,
             "mv_herk: too many work items");
  MV_REQUIRE(ws_bytes >= mv_herk_workspace_bytes(form, m, nslices),
             "mv_herk: workspace too small (%zu < %zu)", ws_bytes,
             mv_herk_workspace_bytes(form, m, nslices));
  const bool plain = form == MV_FORM_C128 || form == MV_FORM_C64;
  const int maxord = plain ? 0 : split_maxord(nterms, nprod);
  if (plain)
    MV_REQUIRE(nterms == 1, "mv_herk: plain forms take nterms=1");
  MV_REQUIRE(maxord >= 0,
             "mv_herk: unsupported split (nterms=%d, nprod=%d); supported: "
             "(2,3) (2,4) (3,6)",
             nterms, nprod);
  HerkPlan p;
  if (plan_herk(form, nterms, maxord, m, nslices, p)) {
    set_error(MV_ERR_HIP, "mv_herk: occupancy query failed");
    return MV_ERR_HIP;
  }
  hipStream_t st = as_stream(stream);
  if (form == MV_FORM_C128) {
    if (phases & kPhaseCompute) {
      herk_plain_kernel<double><<<p.grid, 256, 0, st>>>(
          (const cplx<double> *)z0, m, ldz, nseg, seg_stride, seg_len,
          seg_len_dev, p.nb, p.ntiles, p.cfg, (cplx<double> *)workspace);
      if (int e = check_launch("mv_herk/plain64")) return e;
    }
    if (!(phases & kPhaseReduce)) return MV_OK;
    herk_reduce_kernel<double>
        <<<reduce_grid<double>(ts, ntiles), 256, 0, st>>>(
        (const cplx<double> *)workspace, nullptr, p.ntiles, p.nb, ts, m,
        seg_len, seg_len_dev, nseg, p.kc, p.cfg, scale, (cplx<double> *)vacc,
        ldv, accumulate);
    return check_launch("mv_herk/reduce");
  }
  if (form == MV_FORM_C64) {
    if (phases & kPhaseCompute) {
      herk_plain_kernel<float><<<p.grid, 256, 0, st>>>(
          (const cplx<float> *)z0, m, ldz, nseg, seg_stride, seg_len,
          seg_len_dev, p.nb, p.ntiles, p.cfg, (cplx<float> *)workspace);
      if (int e = check_launch("mv_herk/plain32")) return e;
    }
    if (!(phases & kPhaseReduce)) return MV_OK;
    herk_reduce_kernel<float>
        <<<reduce_grid<float>(ts, ntiles), 256, 0, st>>>(
        (const cplx<float> *)workspace, nullptr, p.ntiles, p.nb, ts, m,
        seg_len, seg_len_dev, nseg, p.kc, p.cfg, scale, (cplx<double> *)vacc,
        ldv, accumulate);
    return check_launch("mv_herk/reduce");
  }
  MV_REQUIRE(z1 && (nterms < 3 || z2), "mv_herk: missing term arrays");
  MV_REQUIRE(((uintptr_t)z0 | (uintptr_t)z1 | (uintptr_t)(z2 ? z2 : z0)) %
                     16 == 0,
             "mv_herk: split terms must be 16-byte aligned");
  const int64_t padded = cdiv(seg_len, (int64_t)kSK) * kSK;
  MV_REQUIRE(nseg == 1 || (seg_stride % kSK == 0 && seg_stride >= padded),
             "mv_herk: split terms need segments of whole %d-column K tiles "
             "(seg_stride %lld, seg_len %lld)",
             kSK, (long long)seg_stride, (long long)seg_len);
  const bool f16 = form == MV_FORM_SPLIT_F16;
  double *dpart = (double *)((char *)workspace + split_partial_bytes(m, nslices));
  SplitProblem P;
  P.t[0] = (const uint32_t *)z0;
  P.t[1] = (const uint32_t *)z1;
  P.t[2] = (const uint32_t *)z2;
  P.ldz = ldz;
  P.seg_stride = seg_stride;
  P.seg_len_host = seg_len;
  P.seg_len_dev = seg_len_dev;
  P.part = (cplx<float> *)workspace;
  P.dpart = dpart;
  P.vacc = (cplx<double> *)vacc;
  P.ldv = ldv;
  P.scale = scale;
  P.m = m;
  P.nseg = nseg;
  P.nb = p.nb;
  P.ntiles = p.ntiles;
  P.accumulate = accumulate;
#define MV_SPLIT_LAUNCH(NT, MO, F)                                           \
  herk_split_kernel<NT, MO, F, kST><<<p.grid, 256, 0, st>>>(P, p.cfg)
  if (phases & kPhaseCompute) {
    if (nterms == 2 && maxord == 1) {
      if (f16) MV_SPLIT_LAUNCH(2, 1, true);
      else MV_SPLIT_LAUNCH(2, 1, false);
    } else if (nterms == 2) {
      if (f16) MV_SPLIT_LAUNCH(2, 2, true);
      else MV_SPLIT_LAUNCH(2, 2, false);
    } else {
      if (f16) MV_SPLIT_LAUNCH(3, 2, true);
      else MV_SPLIT_LAUNCH(3, 2, false);
    }
    if (int e = check_launch("mv_herk/split")) return e;
  }
#undef MV_SPLIT_LAUNCH
  if (!(phases & kPhaseReduce)) return MV_OK;
  herk_reduce_kernel<float><<<reduce_grid<float>(ts, ntiles), 256, 0, st>>>(
      (const cplx<float> *)workspace, dpart, p.ntiles, p.nb, ts, m, seg_len,
      seg_len_dev, nseg, p.kc, p.cfg, scale, (cplx<double> *)vacc, ldv,
      accumulate);
  return check_launch("mv_herk/reduce");
}

}  // namespace

#define MV_HERK_ARGS                                                        \
  int form, int nterms, int nprod, const void *z0, const void *z1,          \
      const void *z2, int m, int64_t ldz, int nseg, int64_t seg_stride,     \
      int64_t seg_len, const int32_t *seg_len_dev, int nslices,             \
      double scale, double *vacc, int64_t ldv, int accumulate,              \
      void *workspace, size_t ws_bytes, mv_stream_t stream
#define MV_HERK_PASS                                                        \
  form, nterms, nprod, z0, z1, z2, m, ldz, nseg, seg_stride, seg_len,       \
      seg_len_dev, nslices, scale, vacc, ldv, accumulate, workspace,        \
      ws_bytes, stream

extern "C" int mv_herk(MV_HERK_ARGS) {
  return herk_run(kPhaseCompute | kPhaseReduce, MV_HERK_PASS);
}
extern "C" int mv_herk_compute(MV_HERK_ARGS) {
  return herk_run(kPhaseCompute, MV_HERK_PASS);
}
extern "C" int mv_herk_reduce(MV_HERK_ARGS) {
  return herk_run(kPhaseReduce, MV_HERK_PASS);
}
#undef MV_HERK_ARGS
#undef MV_HERK_PASS

// ------------------------------------------------------------ no-alloc entry

namespace {
// The split-K partials of mv_herk_noalloc: device memory the library holds
// from load until the process ends, one copy on each device it runs on
// (128 MiB: room for 48 slices of the 350-antenna polarized split form,
// more than its accuracy cap of 192 K tiles a slice needs at 100k sources).
// Host code uses the per-device address from hipGetSymbolAddress, never the
// array's name.  Every call on a device shares that device's copy and
// leaves its partials in it; nothing in it needs clearing (each call's
// kernels write every partial they later read).
constexpr size_t kScratchBytes = size_t(128) << 20;
}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;

// The address of this device's scratch copy (looked up once per device).
int scratch_address(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return set_error(MV_ERR_HIP, "mv_herk_noalloc: hipGetDevice failed");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto f = g_scratch_ptr.find(dev);
  if (f != g_scratch_ptr.end()) {
    *out = f->second;
    return MV_OK;
  }
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(MV_ERR_HIP,
                     "mv_herk_noalloc: hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
  g_scratch_ptr[dev] = ptr;
  *out = ptr;
  return MV_OK;
}
}  // namespace

extern "C" int mv_herk_noalloc(int form, int nterms, int nprod,
                               const void *z0, const void *z1, const void *z2,
                               int m, int64_t ldz, int nseg,
                               int64_t seg_stride, int64_t seg_len,
                               const int32_t *seg_len_dev, double scale,
                               double *vacc, int64_t ldv, int accumulate,
                               mv_stream_t stream) {
  const int ts = mv_herk_tile(form);
  MV_REQUIRE(ts > 0, "mv_herk_noalloc: unknown form %d", form);
  MV_REQUIRE(m >= 1 && nseg >= 1 && seg_len >= 0,
             "mv_herk_noalloc: bad sizes m=%d nseg=%d seg_len=%lld", m, nseg,
             (long long)seg_len);
  // the slices the scratch holds, at most what mv_herk would size for
  const size_t one = mv_herk_workspace_bytes(form, m, 1);
  if (one == 0 || one > kScratchBytes)
    return set_error(MV_ERR_CAPACITY,
                     "mv_herk_noalloc: m=%d needs %zu bytes of partials for "
                     "one K-slice, more than the %zu-byte scratch",
                     m, one, kScratchBytes);
  int ns = mv_herk_max_slices(form, m, (int64_t)nseg * seg_len);
  const int fit = (int)(kScratchBytes / one);
  ns = ns < 1 ? 1 : (ns > fit ? fit : ns);
  void *ws = nullptr;
  if (int e = scratch_address(&ws)) return e;
  return herk_run(kPhaseCompute | kPhaseReduce, form, nterms, nprod, z0, z1,
                  z2, m, ldz, nseg, seg_stride, seg_len, seg_len_dev, ns,
                  scale, vacc, ldv, accumulate, ws,
                  mv_herk_workspace_bytes(form, m, ns), stream);
}

// ------------------------------------------------------------ grouped API

namespace {
int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Head of a grouped workspace: internal descriptors, then the plan arrays;
// the partial areas follow at `parts`.
struct GroupLayout {
  int64_t probs, ns, item_off, tile_off, parts, total;
};

// Bytes per partial element and whether FP64 row norms are kept, per form.
bool is_split_form(int form) {
  return form == MV_FORM_SPLIT_BF16 || form == MV_FORM_SPLIT_F16;
}
int part_bytes(int form) { return form == MV_FORM_C128 ? 16 : 8; }

GroupLayout group_layout(int form, int ts, const mv_herk_problem *q,
                         int nprob) {
  GroupLayout g;
  g.probs = 0;
  g.ns = align256((int64_t)nprob * sizeof(SplitProblem));
  g.item_off = g.ns + align256((int64_t)nprob * 4);
  g.tile_off = g.item_off + align256((int64_t)(nprob + 1) * 8);
  g.parts = g.tile_off + align256((int64_t)(nprob + 1) * 4);
  int64_t bytes = 0;
  for (int i = 0; i < nprob; ++i) {
    const int64_t nb = cdiv((int64_t)q[i].m, (int64_t)ts);
    const int64_t ntiles = nb * (nb + 1) / 2;
    bytes += (int64_t)q[i].nslices * ntiles *
             ((int64_t)ts * ts * part_bytes(form) +
              (is_split_form(form) ? (int64_t)ts * 8 : 0));
  }
  g.total = g.parts + bytes;
  return g;
}

const void *grouped_kernel(int tile, bool f16, int nterms, int maxord) {
  return tile == 64 ? grouped_kernel_t<64>(f16, nterms, maxord)
                    : grouped_kernel_t<kST>(f16, nterms, maxord);
}
}  // namespace

extern "C" size_t mv_herk_problem_bytes(void) {
  return sizeof(mv_herk_problem);
}

namespace {
bool group_tile_ok(int form, int tile) {
  return is_split_form(form) ? (tile == 64 || tile == kST)
                             : ((form == MV_FORM_C128 || form == MV_FORM_C64) &&
                                tile == kPT);
}
}  // namespace

extern "C" size_t mv_herk_grouped_workspace_bytes(
    int form, int tile, const mv_herk_problem *problems, int nprob) {
  if (!group_tile_ok(form, tile) || !problems || nprob < 1) return 0;
  return (size_t)group_layout(form, tile, problems, nprob).total;
}

extern "C" int mv_herk_grouped(int form, int nterms, int nprod, int tile,
                               const mv_herk_problem *problems_host,
                               const mv_herk_problem *problems_dev, int nprob,
                               void *workspace, size_t ws_bytes,
                               mv_stream_t stream) {
  const bool split = is_split_form(form);
  MV_REQUIRE(split || form == MV_FORM_C128 || form == MV_FORM_C64,
             "mv_herk_grouped: unknown form %d", form);
  MV_REQUIRE(group_tile_ok(form, tile),
             "mv_herk_grouped: tile %d (split forms: 64 or %d; plain: %d)",
             tile, kST, kPT);
  const int maxord = split ? split_maxord(nterms, nprod) : 0;
  MV_REQUIRE(split ? maxord >= 0 : nterms == 1,
             "mv_herk_grouped: unsupported terms (nterms=%d, nprod=%d)",
             nterms, nprod);
  MV_REQUIRE(problems_host && problems_dev && workspace && nprob >= 1,
             "mv_herk_grouped: null pointer or no problems");
  int64_t total_tiles = 0, max_items = 0;
  for (int i = 0; i < nprob; ++i) {
    const mv_herk_problem &q = problems_host[i];
    MV_REQUIRE(q.m >= 1 && q.nseg >= 1 && q.seg_len >= 0 &&
                   q.nslices >= 1 && q.kchain >= 0,
               "mv_herk_grouped: problem %d: bad sizes m=%d nseg=%d "
               "seg_len=%lld nslices=%d kchain=%d", i, q.m, q.nseg,
               (long long)q.seg_len, q.nslices, q.kchain);
    MV_REQUIRE(q.z[0] && (!split || q.z[1]) && (nterms < 3 || q.z[2]) &&
                   q.vacc,
               "mv_herk_grouped: problem %d: null pointer", i);
    if (split) {
      const int64_t padded = cdiv(q.seg_len, (int64_t)kSK) * kSK;
      MV_REQUIRE(((q.z[0] | q.z[1] | (q.z[2] ? q.z[2] : q.z[0])) % 16) == 0,
                 "mv_herk_grouped: problem %d: terms must be 16-byte "
                 "aligned", i);
      MV_REQUIRE((q.nseg == 1 ||
                  (q.seg_stride % kSK == 0 && q.seg_stride >= padded)) &&
                     q.ldz >= q.m,
                 "mv_herk_grouped: problem %d: split terms need segments of "
                 "whole %d-column K tiles and ldz >= m", i, kSK);
    } else {
      MV_REQUIRE((q.nseg == 1 || q.seg_stride >= q.seg_len) &&
                     q.ldz >= (int64_t)(q.nseg - 1) * q.seg_stride + q.seg_len,
                 "mv_herk_grouped: problem %d: segments overlap or ldz too "
                 "small", i);
    }
    const int64_t nb = cdiv((int64_t)q.m, (int64_t)tile);
    MV_REQUIRE(q.ldv >= nb * tile,
               "mv_herk_grouped: problem %d: ldv %lld < %lld", i,
               (long long)q.ldv, (long long)(nb * tile));
    total_tiles += nb * (nb + 1) / 2;
    max_items += nb * (nb + 1) / 2 * q.nslices;
  }
  MV_REQUIRE(total_tiles < 65536 && max_items < (int64_t)INT32_MAX / 2,
             "mv_herk_grouped: too many tiles (%lld) or items",
             (long long)total_tiles);
  const GroupLayout g = group_layout(form, tile, problems_host, nprob);
  MV_REQUIRE(ws_bytes >= (size_t)g.total,
             "mv_herk_grouped: workspace too small (%zu < %lld)", ws_bytes,
             (long long)g.total);
  const bool f16 = form == MV_FORM_SPLIT_F16;
  const void *kernel =
      split ? grouped_kernel(tile, f16, nterms, maxord)
            : (form == MV_FORM_C128
                   ? (const void *)&herk_plain_grouped_kernel<double>
                   : (const void *)&herk_plain_grouped_kernel<float>);
  const int threads = !split ? 256
                      : tile == 64 ? SplitShape<64>::kThreads
                                   : SplitShape<kST>::kThreads;
  const int slots = kernel_slots(kernel, threads);
  if (slots <= 0) {
    set_error(MV_ERR_HIP, "mv_herk_grouped: occupancy query failed");
    return MV_ERR_HIP;
  }
  char *ws = (char *)workspace;
  SplitProblem *probs = (SplitProblem *)(ws + g.probs);
  GroupPlan plan{(int32_t *)(ws + g.ns), (int64_t *)(ws + g.item_off),
                 (int32_t *)(ws + g.tile_off)};
  hipStream_t st = as_stream(stream);
  // slices of a problem: split forms bound the FP32 chain (form_cmax);
  // plain forms take slices of 64 K tiles (1024 columns) for parallelism
  herk_group_plan_kernel<<<1, 256, 0, st>>>(
      problems_dev, nprob, tile, split ? kSK : kPK,
      split ? form_cmax(form) : 64, part_bytes(form), split ? 1 : 0, probs,
      ws + g.parts, plan);
  if (int e = check_launch("mv_herk_grouped/plan")) return e;
  const int grid = (int)(max_items < slots ? max_items : slots);
  if (!split) {
    if (form == MV_FORM_C128) {
      herk_plain_grouped_kernel<double><<<grid, 256, 0, st>>>(probs, nprob,
                                                              plan);
      if (int e = check_launch("mv_herk_grouped/plain")) return e;
      herk_group_reduce_kernel<double>
          <<<reduce_grid<double>(tile, total_tiles), 256, 0, st>>>(
              probs, nprob, tile, plan);
    } else {
      herk_plain_grouped_kernel<float><<<grid, 256, 0, st>>>(probs, nprob,
                                                             plan);
      if (int e = check_launch("mv_herk_grouped/plain")) return e;
      herk_group_reduce_kernel<float>
          <<<reduce_grid<float>(tile, total_tiles), 256, 0, st>>>(
              probs, nprob, tile, plan);
    }
    return check_launch("mv_herk_grouped/reduce");
  }
#define MV_GROUP_LAUNCH(NT, MO, F, T)                                        \
  herk_split_grouped_kernel<NT, MO, F, T>                                    \
      <<<grid, SplitShape<T>::kThreads, 0, st>>>(probs, nprob, plan)
#define MV_GROUP_TILE(T)                                                     \
  if (nterms == 2 && maxord == 1) {                                          \
    if (f16) MV_GROUP_LAUNCH(2, 1, true, T);                                 \
    else MV_GROUP_LAUNCH(2, 1, false, T);                                    \
  } else if (nterms == 2) {                                                  \
    if (f16) MV_GROUP_LAUNCH(2, 2, true, T);                                 \
    else MV_GROUP_LAUNCH(2, 2, false, T);                                    \
  } else {                                                                   \
    if (f16) MV_GROUP_LAUNCH(3, 2, true, T);                                 \
    else MV_GROUP_LAUNCH(3, 2, false, T);                                    \
  }
  if (tile == 64) {
    MV_GROUP_TILE(64)
  } else {
    MV_GROUP_TILE(kST)
  }
#undef MV_GROUP_TILE
#undef MV_GROUP_LAUNCH
  if (int e = check_launch("mv_herk_grouped/split")) return e;
  herk_group_reduce_kernel<float>
      <<<reduce_grid<float>(tile, total_tiles), 256, 0, st>>>(probs, nprob,
                                                             tile, plan);
  return check_launch("mv_herk_grouped/reduce");
}

extern "C" int mv_vacc_sum(const double *src, int nsrc_v, int64_t stride,
                           double *dst, int64_t count, mv_stream_t stream) {
  MV_REQUIRE(src && dst, "mv_vacc_sum: null pointer");
  MV_REQUIRE(nsrc_v >= 1 && count >= 0 && stride >= count,
             "mv_vacc_sum: bad sizes");
  if (count == 0) return MV_OK;
  vacc_sum_kernel<<<(unsigned)cdiv(count, (int64_t)256), 256, 0,
                    as_stream(stream)>>>(src, nsrc_v, stride, dst, count);
  return check_launch("mv_vacc_sum");
}

extern "C" int mv_vis_expand(const double *vacc, int64_t ldv, int tile,
                             int nant, int nfeed, const int32_t *pairs,
                             int64_t npairs, void *out, int out_is_f64,
                             mv_stream_t stream) {
  MV_REQUIRE(vacc && out, "mv_vis_expand: null pointer");
  MV_REQUIRE(tile > 0 && nant >= 1 && (nfeed == 1 || nfeed == 2) &&
                 npairs >= 0 && ldv >= (int64_t)nant * nfeed,
             "mv_vis_expand: bad sizes");
  const int64_t total = npairs * nfeed * nfeed;
  if (total == 0) return MV_OK;
  dim3 g((unsigned)cdiv(total, (int64_t)256));
  hipStream_t st = as_stream(stream);
  if (out_is_f64)
    expand_kernel<double><<<g, 256, 0, st>>>((const cplx<double> *)vacc, ldv,
                                             nant, nfeed, pairs, npairs,
                                             (cplx<double> *)out);
  else
    expand_kernel<float><<<g, 256, 0, st>>>((const cplx<double> *)vacc, ldv,
                                            nant, nfeed, pairs, npairs,
                                            (cplx<float> *)out);
  return check_launch("mv_vis_expand");
}
