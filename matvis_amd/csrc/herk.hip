// (d) V = conj(Z) Z^T as a Hermitian rank-k update on MFMA tiles.
//
// Only tiles (bi, bj) with bi <= bj of the upper triangle are computed.  K (the
// source axis, times nax) is split into slices; each workgroup owns one
// (tile, slice) pair and writes its partial tile to a workspace, and a reduce
// kernel sums the slices in FP64 into the accumulator.  The order of that sum
// is fixed, so results are deterministic.
//
// Complex arithmetic on real MFMA: with X_r = (re_0, im_0, re_1, im_1, ...)
// the row of Z read as reals, and Y_r = (im_0, -re_0, im_1, -re_1, ...),
//   Re V[r, c] = X_r . X_c        Im V[r, c] = X_r . Y_c
// so each output tile needs two real products over 2K reals: 8 flop per
// complex multiply-add, with no extra passes over Z.
#include "common.h"

namespace mv {
namespace {

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

// The hardware hands workgroup w to XCD w % 8.  Remap so that each XCD owns a
// contiguous run of logical ids, ordered slice-major: workgroups that share an
// XCD (and its L2) then work on neighbouring K-slices of every tile.
__device__ __forceinline__ int xcd_remap(int w, int total) {
  constexpr int nx = 8;
  const int per = total / nx, rem = total % nx;
  const int x = w % nx, j = w / nx;
  return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

struct KRange {
  int64_t kt_per_seg, kt0, kt1;
};

// K tiles of width kc over nseg segments of seg_len columns, split evenly
// across nslices; slice `slice` gets [kt0, kt1).
__device__ __forceinline__ KRange k_range(int64_t seg_len, int nseg,
                                          int nslices, int slice, int kc) {
  KRange r;
  r.kt_per_seg = cdiv(seg_len, kc);
  const int64_t total = r.kt_per_seg * nseg;
  const int64_t per = cdiv(total, nslices);
  r.kt0 = min((int64_t)slice * per, total);
  r.kt1 = min(r.kt0 + per, total);
  return r;
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
constexpr int kPK = 16;  // complex columns per LDS stage
constexpr int kPRow = kPK + 1;  // LDS row length (one element of padding)

// 256 threads = 4 waves in a 2x2 arrangement, each wave a 32x32 sub-tile made
// of 2x2 blocks of 16x16.  One 16x16x4 MFMA consumes two complex columns:
// lane l supplies k-slot g = l / 16 -> column g / 2, part g % 2 (re / im).
template <typename T>
__global__ __launch_bounds__(256) void herk_plain_kernel(
    const cplx<T> *__restrict__ z, int m, int64_t ldz, int nseg,
    int64_t seg_stride, int64_t seg_len_host,
    const int32_t *__restrict__ seg_len_dev, int nb, int ntiles, int nslices,
    cplx<T> *__restrict__ part) {
  typedef typename PlainMfma<T>::acc acc_t;
  __shared__ cplx<T> As[2][kPT * kPRow];
  __shared__ cplx<T> Bs[2][kPT * kPRow];

  const int logical = xcd_remap(blockIdx.x, ntiles * nslices);
  const int slice = logical / ntiles, tile = logical % ntiles;
  int bi, bj;
  tile_coords(tile, nb, bi, bj);
  const bool diag = bi == bj;
  // the device count never exceeds the host bound the buffers were sized for
  const int64_t seg_len =
      seg_len_dev ? min((int64_t)*seg_len_dev, seg_len_host) : seg_len_host;
  const KRange kr = k_range(seg_len, nseg, nslices, slice, kPK);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lrow = tid >> 2, lcol = (tid & 3) * 4;
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
    const int64_t sg = kt / kr.kt_per_seg;
    const int64_t c0 = (kt % kr.kt_per_seg) * kPK + lcol;
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

  if (kr.kt0 < kr.kt1) {
    gload(kr.kt0, grow_a, ra);
    lstore(As[0], ra);
    if (!diag) {
      gload(kr.kt0, grow_b, rb);
      lstore(Bs[0], rb);
    }
  }
  __syncthreads();

  const int g = lane >> 4, rl = lane & 15;
  const bool im_part = g & 1;
  for (int64_t kt = kr.kt0; kt < kr.kt1; ++kt) {
    const int buf = (int)((kt - kr.kt0) & 1);
    const bool has_next = kt + 1 < kr.kt1;
    if (has_next) {
      gload(kt + 1, grow_a, ra);
      if (!diag) gload(kt + 1, grow_b, rb);
    }
    const cplx<T> *Ab = As[buf];
    const cplx<T> *Bb = diag ? As[buf] : Bs[buf];
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
    if (has_next) {
      lstore(As[buf ^ 1], ra);
      if (!diag) lstore(Bs[buf ^ 1], rb);
    }
    __syncthreads();
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

// ------------------------------------------------------------ split forms

constexpr int kST = 128;  // tile side, complex rows
constexpr int kSK = 16;   // complex columns per LDS stage (2 MFMA k-steps)
constexpr int kSRow = kSK + 4;  // LDS row stride in 32-bit words (80 bytes)

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

// (re, im) -> (im, -re) on a packed pair of 16-bit floats.
__device__ __forceinline__ uint32_t swap_negate(uint32_t w) {
  return ((w >> 16) | (w << 16)) ^ 0x80000000u;
}

struct TermPtrs {
  const uint32_t *t[3];
};

// 256 threads = 4 waves in a 2x2 arrangement over a 128x128 tile; each wave
// computes a 64x64 sub-tile as 2x2 blocks of 32x32x16 MFMA.  A lane's MFMA
// operand is 8 consecutive 16-bit reals = 4 complex values of one Z row.
// Term products (ta, tb) with ta + tb <= MAXORD are accumulated in FP32.
template <int NT, int MAXORD, bool F16>
__global__ __launch_bounds__(256) void herk_split_kernel(
    TermPtrs tp, int m, int64_t ldz, int nseg, int64_t seg_stride,
    int64_t seg_len_host, const int32_t *__restrict__ seg_len_dev, int nb,
    int ntiles, int nslices, cplx<float> *__restrict__ part) {
  __shared__ uint32_t As[2][NT][kST * kSRow];
  __shared__ uint32_t Bs[2][NT][kST * kSRow];

  const int logical = xcd_remap(blockIdx.x, ntiles * nslices);
  const int slice = logical / ntiles, tile = logical % ntiles;
  int bi, bj;
  tile_coords(tile, nb, bi, bj);
  const bool diag = bi == bj;
  // the device count never exceeds the host bound the buffers were sized for
  const int64_t seg_len =
      seg_len_dev ? min((int64_t)*seg_len_dev, seg_len_host) : seg_len_host;
  const KRange kr = k_range(seg_len, nseg, nslices, slice, kSK);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // global -> LDS: thread covers row tid/2, words [(tid%2)*8, +8)
  const int lrow = tid >> 1, lw = (tid & 1) * 8;
  const int64_t grow_a = (int64_t)bi * kST + lrow;
  const int64_t grow_b = (int64_t)bj * kST + lrow;

  v16f acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = v16f{};
      acc_im[i][j] = v16f{};
    }

  u32x4 ra[NT][2], rb[NT][2];
  auto gload = [&](int64_t kt, int64_t grow, u32x4 (*r)[2]) {
    const int64_t sg = kt / kr.kt_per_seg;
    const int64_t c0 = (kt % kr.kt_per_seg) * kSK + lw;
    const bool row_ok = grow < m;
    const bool full = row_ok && c0 + 8 <= seg_len;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const uint32_t *src = tp.t[t] + grow * ldz + sg * seg_stride + c0;
      if (full) {
        r[t][0] = *(const u32x4 *)src;
        r[t][1] = *(const u32x4 *)(src + 4);
      } else {
        uint32_t w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          w[q] = (row_ok && c0 + q < seg_len) ? src[q] : 0u;
        r[t][0] = u32x4{w[0], w[1], w[2], w[3]};
        r[t][1] = u32x4{w[4], w[5], w[6], w[7]};
      }
    }
  };
  auto lstore = [&](uint32_t (*dst)[kST * kSRow], u32x4 (*r)[2]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      *(u32x4 *)&dst[t][lrow * kSRow + lw] = r[t][0];
      *(u32x4 *)&dst[t][lrow * kSRow + lw + 4] = r[t][1];
    }
  };

  if (kr.kt0 < kr.kt1) {
    gload(kr.kt0, grow_a, ra);
    lstore(As[0], ra);
    if (!diag) {
      gload(kr.kt0, grow_b, rb);
      lstore(Bs[0], rb);
    }
  }
  __syncthreads();

  const int lr = lane & 31, lh = lane >> 5;
  for (int64_t kt = kr.kt0; kt < kr.kt1; ++kt) {
    const int buf = (int)((kt - kr.kt0) & 1);
    const bool has_next = kt + 1 < kr.kt1;
    if (has_next) {
      gload(kt + 1, grow_a, ra);
      if (!diag) gload(kt + 1, grow_b, rb);
    }
    uint32_t(*Ab)[kST * kSRow] = As[buf];
    uint32_t(*Bb)[kST * kSRow] = diag ? As[buf] : Bs[buf];
#pragma unroll
    for (int ks = 0; ks < kSK / 8; ++ks) {
      const int woff = ks * 8 + lh * 4;
      u32x4 a[2][NT], b[2][NT], y[2][NT];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[i][t] = *(const u32x4 *)&Ab[t][(wr * 64 + i * 32 + lr) * kSRow + woff];
          b[i][t] = *(const u32x4 *)&Bb[t][(wc * 64 + i * 32 + lr) * kSRow + woff];
          y[i][t] = u32x4{swap_negate(b[i][t][0]), swap_negate(b[i][t][1]),
                          swap_negate(b[i][t][2]), swap_negate(b[i][t][3])};
        }
#pragma unroll
      for (int ta = 0; ta < NT; ++ta)
#pragma unroll
        for (int tb = 0; tb < NT; ++tb) {
          if (ta + tb > MAXORD) continue;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              acc_re[i][j] = SplitMfma<F16>::op(a[i][ta], b[j][tb], acc_re[i][j]);
              acc_im[i][j] = SplitMfma<F16>::op(a[i][ta], y[j][tb], acc_im[i][j]);
            }
        }
    }
    if (has_next) {
      lstore(As[buf ^ 1], ra);
      if (!diag) lstore(Bs[buf ^ 1], rb);
    }
    __syncthreads();
  }

  // 32x32 accumulator layout: lane l, register r ->
  //   row 8*(r/4) + 4*(l/32) + r%4, col l%32
  cplx<float> *out = part + ((int64_t)slice * ntiles + tile) * kST * kST;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = wr * 64 + i * 32 + 8 * (r >> 2) + 4 * lh + (r & 3);
        const int col = wc * 64 + j * 32 + lr;
        out[row * kST + col] = cplx<float>{acc_re[i][j][r], acc_im[i][j][r]};
      }
}

// ------------------------------------------------------------ epilogues

// Sum the slice partials of every upper tile in FP64, scale, and store or add
// into the accumulator.
template <typename TP>
__global__ __launch_bounds__(256) void herk_reduce_kernel(
    const cplx<TP> *__restrict__ part, int nslices, int ntiles, int nb,
    int ts, double scale, cplx<double> *__restrict__ vacc, int64_t ldv,
    int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int tile = blockIdx.y;
  const int64_t tsz = (int64_t)ts * ts;
  if (e >= tsz) return;
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < nslices; ++s) {
    cplx<TP> v = part[((int64_t)s * ntiles + tile) * tsz + e];
    sr += (double)v.re;
    si += (double)v.im;
  }
  int bi, bj;
  tile_coords(tile, nb, bi, bj);
  const int64_t row = (int64_t)bi * ts + e / ts, col = (int64_t)bj * ts + e % ts;
  cplx<double> *dst = vacc + row * ldv + col;
  if (accumulate) {
    dst->re += sr * scale;
    dst->im += si * scale;
  } else {
    *dst = cplx<double>{sr * scale, si * scale};
  }
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

template <typename TO>
__global__ __launch_bounds__(256) void expand_kernel(
    const cplx<double> *__restrict__ vacc, int64_t ldv, int tile, int nant,
    int nf, const int32_t *__restrict__ pairs, int64_t npairs,
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
  if (r / tile <= c / tile) {
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

extern "C" size_t mv_herk_workspace_bytes(int form, int m, int nslices) {
  const int ts = mv_herk_tile(form);
  if (ts == 0 || m <= 0 || nslices <= 0) return 0;
  const int64_t nb = cdiv(m, ts);
  const int64_t ntiles = nb * (nb + 1) / 2;
  const size_t elem = form == MV_FORM_C128 ? 16 : 8;
  return (size_t)nslices * ntiles * ts * ts * elem;
}

extern "C" int mv_herk(int form, int nterms, int nprod, const void *z0,
                       const void *z1, const void *z2, int m, int64_t ldz,
                       int nseg, int64_t seg_stride, int64_t seg_len,
                       const int32_t *seg_len_dev, int nslices, double scale,
                       double *vacc, int64_t ldv, int accumulate,
                       void *workspace, size_t ws_bytes, mv_stream_t stream) {
  const int ts = mv_herk_tile(form);
  MV_REQUIRE(ts > 0, "mv_herk: unknown form %d", form);
  MV_REQUIRE(z0 && vacc && workspace, "mv_herk: null pointer");
  MV_REQUIRE(m >= 1 && nseg >= 1 && seg_len >= 0 && nslices >= 1,
             "mv_herk: bad sizes m=%d nseg=%d seg_len=%lld nslices=%d", m,
             nseg, (long long)seg_len, nslices);
  MV_REQUIRE(nseg == 1 || seg_stride >= seg_len,
             "mv_herk: segments overlap (seg_stride %lld < seg_len %lld)",
             (long long)seg_stride, (long long)seg_len);
  MV_REQUIRE(ldz >= (int64_t)(nseg - 1) * seg_stride + seg_len,
             "mv_herk: ldz %lld too small", (long long)ldz);
  const int64_t nb = cdiv(m, ts);
  const int64_t ntiles = nb * (nb + 1) / 2;
  MV_REQUIRE(ldv >= nb * ts, "mv_herk: ldv %lld < %lld (tile-padded m)",
             (long long)ldv, (long long)(nb * ts));
  MV_REQUIRE(ntiles * nslices < (int64_t)INT32_MAX, "mv_herk: grid too large");
  MV_REQUIRE(ws_bytes >= mv_herk_workspace_bytes(form, m, nslices),
             "mv_herk: workspace too small (%zu < %zu)", ws_bytes,
             mv_herk_workspace_bytes(form, m, nslices));
  hipStream_t st = as_stream(stream);
  const unsigned grid = (unsigned)(ntiles * nslices);
  if (form == MV_FORM_C128 || form == MV_FORM_C64) {
    MV_REQUIRE(nterms == 1, "mv_herk: plain forms take nterms=1");
    if (form == MV_FORM_C128) {
      herk_plain_kernel<double><<<grid, 256, 0, st>>>(
          (const cplx<double> *)z0, m, ldz, nseg, seg_stride, seg_len,
          seg_len_dev, (int)nb, (int)ntiles, nslices,
          (cplx<double> *)workspace);
      if (int e = check_launch("mv_herk/plain64")) return e;
      herk_reduce_kernel<double>
          <<<dim3((unsigned)cdiv(ts * ts, 256), (unsigned)ntiles), 256, 0,
             st>>>((const cplx<double> *)workspace, nslices, (int)ntiles,
                   (int)nb, ts, scale, (cplx<double> *)vacc, ldv, accumulate);
    } else {
      herk_plain_kernel<float><<<grid, 256, 0, st>>>(
          (const cplx<float> *)z0, m, ldz, nseg, seg_stride, seg_len,
          seg_len_dev, (int)nb, (int)ntiles, nslices,
          (cplx<float> *)workspace);
      if (int e = check_launch("mv_herk/plain32")) return e;
      herk_reduce_kernel<float>
          <<<dim3((unsigned)cdiv(ts * ts, 256), (unsigned)ntiles), 256, 0,
             st>>>((const cplx<float> *)workspace, nslices, (int)ntiles,
                   (int)nb, ts, scale, (cplx<double> *)vacc, ldv, accumulate);
    }
    return check_launch("mv_herk/reduce");
  }
  const int maxord = split_maxord(nterms, nprod);
  MV_REQUIRE(maxord >= 0,
             "mv_herk: unsupported split (nterms=%d, nprod=%d); supported: "
             "(2,3) (2,4) (3,6)",
             nterms, nprod);
  MV_REQUIRE(z1 && (nterms < 3 || z2), "mv_herk: missing term arrays");
  MV_REQUIRE(ldz % 4 == 0 && (nseg == 1 || seg_stride % 4 == 0) &&
                 ((uintptr_t)z0 | (uintptr_t)z1 | (uintptr_t)(z2 ? z2 : z0)) %
                         16 == 0,
             "mv_herk: split terms need 16-byte aligned rows (ldz and "
             "seg_stride multiples of 4)");
  TermPtrs tp = {{(const uint32_t *)z0, (const uint32_t *)z1,
                  (const uint32_t *)z2}};
  const bool f16 = form == MV_FORM_SPLIT_F16;
#define MV_SPLIT_LAUNCH(NT, MO, F)                                           \
  herk_split_kernel<NT, MO, F><<<grid, 256, 0, st>>>(                        \
      tp, m, ldz, nseg, seg_stride, seg_len, seg_len_dev, (int)nb,           \
      (int)ntiles, nslices, (cplx<float> *)workspace)
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
#undef MV_SPLIT_LAUNCH
  if (int e = check_launch("mv_herk/split")) return e;
  herk_reduce_kernel<float>
      <<<dim3((unsigned)cdiv(ts * ts, 256), (unsigned)ntiles), 256, 0, st>>>(
          (const cplx<float> *)workspace, nslices, (int)ntiles, (int)nb, ts,
          scale, (cplx<double> *)vacc, ldv, accumulate);
  return check_launch("mv_herk/reduce");
}

extern "C" int mv_vacc_sum(const double *src, int nsrc_v, int64_t stride,
                           double *dst, int64_t count, mv_stream_t stream) {
  MV_REQUIRE(src && dst, "mv_vacc_sum: null pointer");
  MV_REQUIRE(nsrc_v >= 1 && count >= 0 && stride >= count,
             "mv_vacc_sum: bad sizes");
  if (count == 0) return MV_OK;
  vacc_sum_kernel<<<(unsigned)cdiv(count, 256), 256, 0, as_stream(stream)>>>(
      src, nsrc_v, stride, dst, count);
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
  dim3 g((unsigned)cdiv(total, 256));
  hipStream_t st = as_stream(stream);
  if (out_is_f64)
    expand_kernel<double><<<g, 256, 0, st>>>((const cplx<double> *)vacc, ldv,
                                             tile, nant, nfeed, pairs, npairs,
                                             (cplx<double> *)out);
  else
    expand_kernel<float><<<g, 256, 0, st>>>((const cplx<double> *)vacc, ldv,
                                            tile, nant, nfeed, pairs, npairs,
                                            (cplx<float> *)out);
  return check_launch("mv_vis_expand");
}
