// (c) Geometric phase and the Z matrix.
//
// Phase: cycles = (b . s) * (freq / c) in FP64, reduced to [-1/2, 1/2] turns
// before sincos, so the sincos argument is small whatever the baseline length
// (the reference forms exp(i * 2 pi freq (b . s) / c) directly,
// src/matvis/core/tau.py:30-31,65-66).  FP64 forms take an FP64 sincospi of
// the reduced phase; FP32 and split forms take the FP32 sincospi.
#include "common.h"

namespace mv {
namespace {

constexpr int kThreads = 256;
constexpr int kZPad = 32;  // Z columns are zero-padded to a multiple of this
// longest Z row, and most term columns (elements): a lane's column is a
// 32-bit index
constexpr int64_t kMaxRow = int64_t(1) << 30;

__device__ __forceinline__ double reduced_turns(const double *__restrict__ ap,
                                                double fc, double sx,
                                                double sy, double sz) {
  double cyc = (ap[0] * sx + ap[1] * sy + ap[2] * sz) * fc;
  return cyc - rint(cyc);
}

template <typename TO> struct Phasor;
template <> struct Phasor<double> {
  __device__ static cplx<double> eval(double turns) {
    double s, c;
    sincospi(2.0 * turns, &s, &c);
    return {c, s};
  }
};
template <> struct Phasor<float> {
  __device__ static cplx<float> eval(double turns) {
    float s, c;
    sincospif((float)(2.0 * turns), &s, &c);
    return {c, s};
  }
};

template <typename TC, typename TO>
__global__ __launch_bounds__(kThreads) void exptau_kernel(
    const double *__restrict__ antpos, double fc, const TC *__restrict__ crd,
    int64_t ld_crd, int64_t nsrc, cplx<TO> *__restrict__ out,
    int64_t ld_out) {
  int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  int a = blockIdx.y;
  if (s >= nsrc) return;
  double t = reduced_turns(antpos + 3 * a, fc, (double)crd[s],
                           (double)crd[ld_crd + s], (double)crd[2 * ld_crd + s]);
  out[(int64_t)a * ld_out + s] = Phasor<TO>::eval(t);
}

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T>
__global__ __launch_bounds__(kThreads) void zmat_parts_kernel(
    const T *__restrict__ sqrt_flux, const cplx<T> *__restrict__ beam,
    const int32_t *__restrict__ beam_idx, int nbeam, cplx<T> *exptau, int nf,
    int na, int64_t nsrc, cplx<T> *__restrict__ z) {
  int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  int a = blockIdx.y;
  if (s >= nsrc) return;
  cplx<T> e = exptau[(int64_t)a * nsrc + s];
  T f = sqrt_flux[s];
  e.re *= f;
  e.im *= f;
  exptau[(int64_t)a * nsrc + s] = e;  // core/getz.py:80 updates exptau in place
  const int b = beam_idx ? beam_idx[a] : (nbeam == 1 ? 0 : a);
  // an index outside the beam list is bad input: NaN, never an out-of-range read
  const bool valid = (unsigned)b < (unsigned)nbeam;
  const T nan = (T)NAN;
  for (int fd = 0; fd < nf; ++fd)
    for (int ax = 0; ax < na; ++ax) {
      cplx<T> A = valid ? beam[((int64_t)(b * nf + fd) * na + ax) * nsrc + s]
                        : cplx<T>{nan, nan};
      z[((int64_t)a * nf + fd) * (na * nsrc) + (int64_t)ax * nsrc + s] =
          cmul(e, A);
    }
}

// ---------------------------------------------------------------- splitting

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// x and y rounded to the nearest-even 16-bit value, x in the low half of the
// word and y in the high half (one v_cvt_pk_bf16_f32 for bf16)
template <bool F16>
__device__ __forceinline__ uint32_t pack16(float x, float y) {
  if constexpr (F16)
    return __builtin_bit_cast(uint32_t,
                              __builtin_convertvector((f32x2){x, y}, f16x2));
  else
    return __builtin_bit_cast(uint32_t,
                              __builtin_convertvector((f32x2){x, y}, bf16x2));
}
// the values of the low and high halves of a packed word
template <bool F16>
__device__ __forceinline__ float half_lo(uint32_t w) {
  if constexpr (F16)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
  else
    return __uint_as_float(w << 16);
}
template <bool F16>
__device__ __forceinline__ float half_hi(uint32_t w) {
  if constexpr (F16)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
  else
    return __uint_as_float(w & 0xFFFF0000u);
}

// Splits (re, im) into NT 16-bit term pairs, re ~ sum of the low halves and
// im ~ sum of the high halves of t[0..NT): each term the round-to-nearest of
// the remainder the previous terms left (the remainders are exact in FP32).
template <bool F16, int NT>
__device__ __forceinline__ void split_terms(float re, float im, uint32_t *t) {
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    t[k] = pack16<F16>(re, im);
    if (k + 1 < NT) {
      re -= half_lo<F16>(t[k]);
      im -= half_hi<F16>(t[k]);
    }
  }
}

// Writers of one Z element at (row offset, column): `row` is the offset of
// the row (and segment) start, the same in every lane of the wave; `col` the
// lane's column within it, below 2^30 (the entry points check the row length).
template <int FORM, int NT> struct ZWriter;
// put(zs, r, ld, col, v) stores element (r, col).  Plain forms: row-major,
// row stride ld elements.
template <> struct ZWriter<MV_FORM_C128, 1> {
  typedef double T;
  __device__ static void put(void *const *zs, int64_t r, int64_t ld,
                             uint32_t col, cplx<double> v) {
    ((cplx<double> *)zs[0] + r * ld)[col] = v;
  }
};
template <> struct ZWriter<MV_FORM_C64, 1> {
  typedef float T;
  __device__ static void put(void *const *zs, int64_t r, int64_t ld,
                             uint32_t col, cplx<float> v) {
    ((cplx<float> *)zs[0] + r * ld)[col] = v;
  }
};
// Split forms: per term, K tiles of 16 columns, each `ld` rows of 64 bytes
// (K-tile-major: row r of tile q at word (q ld + r) 16, the layout the HERK
// copies from).  A row of a tile is 4 16-byte chunks of 4 consecutive
// columns, the real parts of the 4 columns (16 bits each) in the chunk's
// first 8 bytes and their imaginary parts in its last 8.  Column c is slot
// c % 4 of chunk (c % 16) / 4 of tile c / 16.  Lanes 4q ..
// 4q + 3 hold the 4 columns of one chunk.  Each lane packs its (re, im) term
// into one word, takes its xor-1 partner's word (a DPP quad permutation
// within the quad) and stores one whole word: even slots j the real parts
// of columns j, j + 1 (word j / 2), odd slots the imaginary parts of
// columns j - 1, j (word 2 + j / 2), formed by one byte permutation of the
// two words.  Every lane of a quad must call put() together.
template <bool F16, int NT>
__device__ __forceinline__ void put_split(void *const *zs, int64_t r,
                                          int64_t ld, uint32_t col,
                                          cplx<float> v) {
  uint32_t t[NT];
  split_terms<F16, NT>(v.re, v.im, t);
  const uint32_t j = col & 3u;
  const bool odd = j & 1u;
  // word of the tile row, then the byte offset of the lane's word in it
  const int64_t row = ((int64_t)(col >> 4) * ld + r) * 16;
  const uint32_t wb = ((col & 15u) - j + (odd ? 2u : 0u) + (j >> 1)) * 4u;
  // bytes of {partner, mine}: even lanes take my low half then the
  // partner's low half; odd lanes the partner's high half then my high half
  const uint32_t sel = odd ? 0x03020706u : 0x05040100u;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    // quad_perm [1, 0, 3, 2]: the word of lane l ^ 1
    const uint32_t p =
        (uint32_t)__builtin_amdgcn_mov_dpp((int)t[k], 0xB1, 0xF, 0xF, false);
    char *base = (char *)((uint32_t *)zs[k] + row);
    *(uint32_t *)(base + wb) = __builtin_amdgcn_perm(p, t[k], sel);
  }
}
template <int NT> struct ZWriter<MV_FORM_SPLIT_BF16, NT> {
  typedef float T;
  __device__ static void put(void *const *zs, int64_t r, int64_t ld,
                             uint32_t col, cplx<float> v) {
    put_split<false, NT>(zs, r, ld, col, v);
  }
};
template <int NT> struct ZWriter<MV_FORM_SPLIT_F16, NT> {
  typedef float T;
  __device__ static void put(void *const *zs, int64_t r, int64_t ld,
                             uint32_t col, cplx<float> v) {
    put_split<true, NT>(zs, r, ld, col, v);
  }
};

struct ZOut {
  void *z[3];
};

// One thread per source, ants_per_block antennas per block (grid.y).
// POL: nfeed = nax = 2 (polarized) or 1.
template <int FORM, int NT, bool POL>
__global__ __launch_bounds__(kThreads) void zmat_fused_kernel(
    const double *__restrict__ antpos, int nant, double fc,
    const double *__restrict__ crd, int64_t ld_crd,
    const double *__restrict__ flux, const int32_t *__restrict__ idx,
    const cplx<typename ZWriter<FORM, NT>::T> *__restrict__ beam,
    int64_t ld_beam, int nbeam, const int32_t *__restrict__ beam_idx,
    int64_t count, const int32_t *__restrict__ n_dev, double scale,
    ZOut zo, int64_t ldz, int64_t seg_stride, int ants_per_block) {
  typedef typename ZWriter<FORM, NT>::T T;
  constexpr int nf = POL ? 2 : 1;
  constexpr int na = POL ? 2 : 1;
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t n = n_dev ? min((int64_t)*n_dev, count) : count;
  const int64_t n_pad = cdiv(n, kZPad) * kZPad;
  if (s >= n_pad || s >= seg_stride) return;
  const int a0 = blockIdx.y * ants_per_block;
  const int a1 = min(a0 + ants_per_block, nant);
  // lanes past the column count write zeros; every lane reaches the writes
  // (the split forms' writes are wave operations)
  const bool valid = s < n;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  T amp = (T)0;
  if (valid) {
    sx = crd[s];
    sy = crd[ld_crd + s];
    sz = crd[2 * ld_crd + s];
    amp = (T)(flux[idx ? (int64_t)idx[s] : s] * scale);
  }
  int cached = -1;
  cplx<T> A[nf * na];
  for (int a = a0; a < a1; ++a) {
    cplx<T> out[nf * na];
#pragma unroll
    for (int q = 0; q < nf * na; ++q) out[q] = cplx<T>{(T)0, (T)0};
    if (valid) {
      const int b = beam_idx ? beam_idx[a] : (nbeam == 1 ? 0 : a);
      if (b != cached) {
        // an index outside the beam list is bad input: NaN, never an
        // out-of-range read
        const bool ok = (unsigned)b < (unsigned)nbeam;
#pragma unroll
        for (int q = 0; q < nf * na; ++q)
          A[q] = ok ? beam[((int64_t)b * nf * na + q) * ld_beam + s]
                    : cplx<T>{(T)NAN, (T)NAN};
        cached = b;
      }
      cplx<T> e =
          Phasor<T>::eval(reduced_turns(antpos + 3 * a, fc, sx, sy, sz));
      e.re *= amp;
      e.im *= amp;
#pragma unroll
      for (int q = 0; q < nf * na; ++q) out[q] = cmul(A[q], e);
    }
#pragma unroll
    for (int fd = 0; fd < nf; ++fd)
#pragma unroll
      for (int ax = 0; ax < na; ++ax)
        ZWriter<FORM, NT>::put(zo.z, (int64_t)a * nf + fd, ldz,
                               (uint32_t)(ax * seg_stride + s),
                               out[fd * na + ax]);
  }
}

template <typename TI, int FORM, int NT>
__global__ __launch_bounds__(kThreads) void split_kernel(
    const cplx<TI> *__restrict__ z, int64_t rows, int nseg, int64_t cols,
    int64_t ldz, int64_t seg_stride, float scale, ZOut to, int64_t ldt,
    int64_t tseg_stride) {
  int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  int64_t rs = blockIdx.y;  // row * nseg + seg
  // whole chunks of 4 columns: the slots past `cols` get zeros
  if (c >= cdiv(cols, (int64_t)4) * 4) return;
  int64_t r = rs / nseg, sg = rs % nseg;
  cplx<float> w = {0.0f, 0.0f};
  if (c < cols) {
    const cplx<TI> v = z[r * ldz + sg * seg_stride + c];
    w = cplx<float>{(float)v.re * scale, (float)v.im * scale};
  }
  ZWriter<FORM, NT>::put(to.z, r, ldt, (uint32_t)(sg * tseg_stride + c), w);
}

template <int FORM, int NT>
int launch_fused(dim3 g, hipStream_t st, const double *antpos, int nant,
                 double fc, const double *crd, int64_t ld_crd,
                 const double *flux, const int32_t *idx, const void *beam,
                 int64_t ld_beam,
                 int nbeam, const int32_t *beam_idx, int nf, int na,
                 int64_t count, const int32_t *n_dev, double scale, ZOut zo,
                 int64_t ldz, int64_t seg_stride, int apb) {
  typedef typename ZWriter<FORM, NT>::T T;
  (void)na;
  if (nf == 2)
    zmat_fused_kernel<FORM, NT, true><<<g, kThreads, 0, st>>>(
        antpos, nant, fc, crd, ld_crd, flux, idx, (const cplx<T> *)beam,
        ld_beam, nbeam, beam_idx, count, n_dev, scale, zo, ldz, seg_stride,
        apb);
  else
    zmat_fused_kernel<FORM, NT, false><<<g, kThreads, 0, st>>>(
        antpos, nant, fc, crd, ld_crd, flux, idx, (const cplx<T> *)beam,
        ld_beam, nbeam, beam_idx, count, n_dev, scale, zo, ldz, seg_stride,
        apb);
  return check_launch("mv_zmat_fused");
}

}  // namespace
}  // namespace mv

using namespace mv;

extern "C" int mv_exptau(const double *antpos, int nant, double freq,
                         const void *crd, int crd_is_f64, int64_t ld_crd,
                         int64_t nsrc, void *out, int out_is_f64,
                         int64_t ld_out, mv_stream_t stream) {
  MV_REQUIRE(antpos && crd && out, "mv_exptau: null pointer");
  MV_REQUIRE(nant >= 1 && nant <= 65535 && nsrc >= 0 && ld_crd >= nsrc &&
                 ld_out >= nsrc,
             "mv_exptau: bad sizes");
  if (nsrc == 0) return MV_OK;
  dim3 g((unsigned)cdiv(nsrc, kThreads), (unsigned)nant);
  hipStream_t s = as_stream(stream);
  const double fc = freq / kC;
#define MV_TAU(TC, TO)                                                       \
  exptau_kernel<TC, TO><<<g, kThreads, 0, s>>>(                              \
      antpos, fc, (const TC *)crd, ld_crd, nsrc, (cplx<TO> *)out, ld_out)
  if (crd_is_f64) {
    if (out_is_f64) MV_TAU(double, double);
    else MV_TAU(double, float);
  } else {
    if (out_is_f64) MV_TAU(float, double);
    else MV_TAU(float, float);
  }
#undef MV_TAU
  return check_launch("mv_exptau");
}

extern "C" int mv_zmat_parts(const void *sqrt_flux, const void *beam,
                             const int32_t *beam_idx, int nbeam, void *exptau,
                             int nant, int nfeed, int nax, int64_t nsrc,
                             void *z, int is_f64, mv_stream_t stream) {
  MV_REQUIRE(sqrt_flux && beam && exptau && z, "mv_zmat_parts: null pointer");
  MV_REQUIRE(nant >= 1 && nant <= 65535 && nbeam >= 1 && nsrc >= 0 &&
                 (nfeed == 1 || nfeed == 2) && (nax == 1 || nax == 2),
             "mv_zmat_parts: bad sizes");
  MV_REQUIRE(beam_idx || nbeam == 1 || nbeam == nant,
             "mv_zmat_parts: nbeam must be 1 or nant without beam_idx");
  if (nsrc == 0) return MV_OK;
  dim3 g((unsigned)cdiv(nsrc, kThreads), (unsigned)nant);
  hipStream_t s = as_stream(stream);
  if (is_f64)
    zmat_parts_kernel<double><<<g, kThreads, 0, s>>>(
        (const double *)sqrt_flux, (const cplx<double> *)beam, beam_idx,
        nbeam, (cplx<double> *)exptau, nfeed, nax, nsrc, (cplx<double> *)z);
  else
    zmat_parts_kernel<float><<<g, kThreads, 0, s>>>(
        (const float *)sqrt_flux, (const cplx<float> *)beam, beam_idx, nbeam,
        (cplx<float> *)exptau, nfeed, nax, nsrc, (cplx<float> *)z);
  return check_launch("mv_zmat_parts");
}

extern "C" int mv_zmat_fused(const double *antpos, int nant, double freq,
                             const double *crd, int64_t ld_crd,
                             const double *flux, const int32_t *idx,
                             const void *beam, int beam_is_f64,
                             int64_t ld_beam, int nbeam,
                             const int32_t *beam_idx, int nfeed, int nax,
                             int64_t count, const int32_t *n_dev, int form,
                             int nterms, double scale, void *z0, void *z1,
                             void *z2, int64_t ldz, int64_t seg_stride,
                             mv_stream_t stream) {
  MV_REQUIRE(antpos && crd && flux && beam && z0,
             "mv_zmat_fused: null pointer");
  MV_REQUIRE(nant >= 1 && nbeam >= 1 && count >= 0 && ld_crd >= count &&
                 ld_beam >= count && (nfeed == 1 || nfeed == 2) &&
                 nax == nfeed,
             "mv_zmat_fused: bad sizes (nfeed and nax must both be 1 or 2)");
  MV_REQUIRE(beam_idx || nbeam == 1 || nbeam == nant,
             "mv_zmat_fused: nbeam must be 1 or nant without beam_idx");
  MV_REQUIRE(seg_stride >= cdiv(count, kZPad) * kZPad,
             "mv_zmat_fused: seg_stride (%lld) must cover count rounded up "
             "to %d",
             (long long)seg_stride, kZPad);
  const bool plain = form == MV_FORM_C128 || form == MV_FORM_C64;
  // plain: row-major rows of ldz elements; split: K tiles of ldz rows
  MV_REQUIRE(plain ? ldz >= (int64_t)nax * seg_stride
                   : ldz >= (int64_t)nant * nfeed && seg_stride % 16 == 0,
             plain ? "mv_zmat_fused: ldz must be >= nax * seg_stride"
                   : "mv_zmat_fused: split terms need ldz >= nant * nfeed "
                     "rows a K tile and seg_stride a multiple of 16");
  MV_REQUIRE(ldz <= kMaxRow && (int64_t)nax * seg_stride <= kMaxRow,
             "mv_zmat_fused: ldz and nax * seg_stride must be at most 2^30");
  MV_REQUIRE(plain || form == MV_FORM_SPLIT_BF16 || form == MV_FORM_SPLIT_F16,
             "mv_zmat_fused: unknown form %d", form);
  MV_REQUIRE(plain ? nterms == 1 : (nterms == 2 || nterms == 3),
             "mv_zmat_fused: bad nterms %d for form %d", nterms, form);
  MV_REQUIRE(!!beam_is_f64 == (form == MV_FORM_C128),
             "mv_zmat_fused: the beam must be c128 for the c128 form and c64 "
             "otherwise");
  MV_REQUIRE(nterms < 2 || z1, "mv_zmat_fused: missing term array 1");
  MV_REQUIRE(nterms < 3 || z2, "mv_zmat_fused: missing term array 2");
  if (count == 0) return MV_OK;
  const int apb = 16;
  // blocks go to the 8 XCDs round-robin by linear index x + y gx: with gx a
  // multiple of 8, the antenna blocks of one source range share an XCD and
  // its L2, which then serves their re-reads of the per-source inputs (the
  // blocks past the column count return at once)
  dim3 g((unsigned)(cdiv(cdiv(cdiv(count, kZPad) * kZPad, kThreads),
                         (int64_t)8) * 8),
         (unsigned)cdiv(nant, apb));
  hipStream_t st = as_stream(stream);
  const double fc = freq / kC;
  ZOut zo = {{z0, z1, z2}};
#define MV_FUSED(F, NT)                                                      \
  return launch_fused<F, NT>(g, st, antpos, nant, fc, crd, ld_crd, flux,     \
                             idx, beam, ld_beam, nbeam, beam_idx, nfeed, nax,\
                             count, n_dev, scale, zo, ldz, seg_stride, apb)
  switch (form) {
    case MV_FORM_C128: MV_FUSED(MV_FORM_C128, 1);
    case MV_FORM_C64: MV_FUSED(MV_FORM_C64, 1);
    case MV_FORM_SPLIT_BF16:
      if (nterms == 2) MV_FUSED(MV_FORM_SPLIT_BF16, 2);
      MV_FUSED(MV_FORM_SPLIT_BF16, 3);
    default:
      if (nterms == 2) MV_FUSED(MV_FORM_SPLIT_F16, 2);
      MV_FUSED(MV_FORM_SPLIT_F16, 3);
  }
#undef MV_FUSED
}

extern "C" int mv_split_z(const void *z, int z_is_f64, int64_t rows, int nseg,
                          int64_t cols, int64_t ldz, int64_t seg_stride,
                          int form, int nterms, double scale, void *t0,
                          void *t1, void *t2, int64_t ldt,
                          int64_t tseg_stride, mv_stream_t stream) {
  MV_REQUIRE(z && t0, "mv_split_z: null pointer");
  MV_REQUIRE(form == MV_FORM_SPLIT_BF16 || form == MV_FORM_SPLIT_F16,
             "mv_split_z: form must be a split form");
  MV_REQUIRE(nterms == 2 || nterms == 3, "mv_split_z: nterms must be 2 or 3");
  MV_REQUIRE(nterms < 3 || (t1 && t2), "mv_split_z: missing term arrays");
  MV_REQUIRE(t1, "mv_split_z: missing term array 1");
  // terms: K tiles of 16 columns, ldt rows each (ldt >= rows)
  MV_REQUIRE(rows >= 0 && nseg >= 1 && cols >= 0 && rows * nseg <= 65535 &&
                 seg_stride >= cols && ldz >= nseg * seg_stride &&
                 tseg_stride >= cols && ldt >= rows,
             "mv_split_z: bad sizes");
  MV_REQUIRE(ldt <= kMaxRow && nseg * tseg_stride <= kMaxRow,
             "mv_split_z: ldt and nseg * tseg_stride must be at most 2^30");
  MV_REQUIRE(nseg == 1 || tseg_stride % 16 == 0,
             "mv_split_z: term segments must start on 16-column K tiles "
             "(tseg_stride a multiple of 16)");
  if (rows == 0 || cols == 0) return MV_OK;
  dim3 g((unsigned)cdiv(cols, kThreads), (unsigned)(rows * nseg));
  hipStream_t st = as_stream(stream);
  ZOut to = {{t0, t1, t2}};
  float sc = (float)scale;
#define MV_SPLIT(TI, F, NT)                                                  \
  split_kernel<TI, F, NT><<<g, kThreads, 0, st>>>(                           \
      (const cplx<TI> *)z, rows, nseg, cols, ldz, seg_stride, sc, to, ldt,   \
      tseg_stride)
  if (form == MV_FORM_SPLIT_BF16) {
    if (z_is_f64) {
      if (nterms == 2) MV_SPLIT(double, MV_FORM_SPLIT_BF16, 2);
      else MV_SPLIT(double, MV_FORM_SPLIT_BF16, 3);
    } else {
      if (nterms == 2) MV_SPLIT(float, MV_FORM_SPLIT_BF16, 2);
      else MV_SPLIT(float, MV_FORM_SPLIT_BF16, 3);
    }
  } else {
    if (z_is_f64) {
      if (nterms == 2) MV_SPLIT(double, MV_FORM_SPLIT_F16, 2);
      else MV_SPLIT(double, MV_FORM_SPLIT_F16, 3);
    } else {
      if (nterms == 2) MV_SPLIT(float, MV_FORM_SPLIT_F16, 2);
      else MV_SPLIT(float, MV_FORM_SPLIT_F16, 3);
    }
  }
#undef MV_SPLIT
  return check_launch("mv_split_z");
}
