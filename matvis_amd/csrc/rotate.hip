// (a) Rotation of ECI source vectors into the topocentric ENU frame, fused
// with the horizon cull and a stable (ascending-index) stream compaction.
//
// Compaction runs as count -> scan -> scatter.  Each block owns a fixed run of
// kItemsPerBlock sources and re-derives the same horizon flags in both the
// count and the scatter pass, so the output order equals the reference's
// `xp.where(topo[2] > 0)` order (src/matvis/core/coords.py:115) for any grid.
#include "common.h"

namespace mv {
namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 8;
constexpr int kItemsPerBlock = kThreads * kRounds;  // 2048 sources per block

struct Mat3 {
  double m[9];
};

// Row i of R . (x, y, z), evaluated as (R[i,0]*x + R[i,1]*y) + R[i,2]*z with
// one rounding per operation, so the host oracle reproduces it bit for bit.
__device__ __forceinline__ double rot_row(const Mat3 &R, int i, double x,
                                          double y, double z) {
#pragma clang fp contract(off)
  double a = R.m[3 * i + 0] * x;
  double b = R.m[3 * i + 1] * y;
  double c = R.m[3 * i + 2] * z;
  return (a + b) + c;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void rotate_kernel(
    const double *__restrict__ eci, int64_t ld_eci, Mat3 R, int64_t nsrc,
    T *__restrict__ topo, int64_t ld_topo) {
  int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nsrc) return;
  double x = eci[p], y = eci[ld_eci + p], z = eci[2 * ld_eci + p];
  topo[p] = (T)rot_row(R, 0, x, y, z);
  topo[ld_topo + p] = (T)rot_row(R, 1, x, y, z);
  topo[2 * ld_topo + p] = (T)rot_row(R, 2, x, y, z);
}

__device__ __forceinline__ bool above_horizon(const double *__restrict__ eci,
                                              int64_t ld_eci, const Mat3 &R,
                                              int64_t g) {
  double x = eci[g], y = eci[ld_eci + g], z = eci[2 * ld_eci + g];
  return rot_row(R, 2, x, y, z) > 0.0;
}

// Pass 1: number of above-horizon sources in each block's run.
__global__ __launch_bounds__(kThreads) void cull_count_kernel(
    const double *__restrict__ eci, int64_t ld_eci, Mat3 R, int64_t chunk0,
    int64_t chunk_len, int32_t *__restrict__ counts) {
  __shared__ int wave_sum[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kItemsPerBlock;
  int local = 0;
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    int64_t p = base + r * kThreads + threadIdx.x;
    if (p < chunk_len && above_horizon(eci, ld_eci, R, chunk0 + p)) ++local;
  }
  // wave reduction
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
  if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kThreads / 64; ++w) s += wave_sum[w];
    counts[blockIdx.x] = s;
  }
}

// Pass 2: exclusive scan of block counts (single workgroup); total -> n_up.
__global__ __launch_bounds__(1024) void cull_scan_kernel(
    int32_t *__restrict__ counts, int64_t nblocks,
    int32_t *__restrict__ n_up) {
  __shared__ int buf[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nblocks; b0 += 1024) {
    int64_t b = b0 + threadIdx.x;
    int v = b < nblocks ? counts[b] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 entries
    for (int off = 1; off < 1024; off <<= 1) {
      int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    int incl = buf[threadIdx.x];
    if (b < nblocks) counts[b] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_up = carry;
}

// Pass 3: recompute flags, place each kept source at its block offset plus its
// rank inside the block, then zero-fill positions [n_up, nalloc).
template <typename T>
__global__ __launch_bounds__(kThreads) void cull_scatter_kernel(
    const double *__restrict__ eci, int64_t ld_eci, Mat3 R,
    const double *__restrict__ flux_in, int64_t chunk0, int64_t chunk_len,
    int64_t nalloc, const int32_t *__restrict__ offsets,
    const int32_t *__restrict__ n_up_dev, T *__restrict__ crd, int64_t ld_crd,
    T *__restrict__ flux_out, int32_t *__restrict__ idx_out,
    int64_t nblocks_real) {
  __shared__ int wave_cnt[kThreads / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if ((int64_t)blockIdx.x < nblocks_real) {
    const int64_t base = (int64_t)blockIdx.x * kItemsPerBlock;
    int64_t run = offsets[blockIdx.x];
    for (int r = 0; r < kRounds; ++r) {
      int64_t p = base + r * kThreads + threadIdx.x;
      int64_t g = chunk0 + p;
      bool keep = false;
      double x = 0, y = 0, z = 0;
      if (p < chunk_len) {
        x = eci[g];
        y = eci[ld_eci + g];
        z = eci[2 * ld_eci + g];
        keep = rot_row(R, 2, x, y, z) > 0.0;
      }
      uint64_t mask = __ballot(keep);
      int rank = __popcll(mask & ((1ull << lane) - 1ull));
      if (lane == 0) wave_cnt[wave] = __popcll(mask);
      __syncthreads();
      int before = 0, total = 0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) {
        before += (w < wave) ? wave_cnt[w] : 0;
        total += wave_cnt[w];
      }
      if (keep) {
        int64_t dst = run + before + rank;
        if (dst >= 0 && dst < nalloc) {
          crd[dst] = (T)rot_row(R, 0, x, y, z);
          crd[ld_crd + dst] = (T)rot_row(R, 1, x, y, z);
          crd[2 * ld_crd + dst] = (T)rot_row(R, 2, x, y, z);
          if (flux_out) flux_out[dst] = (T)flux_in[g];
          if (idx_out) idx_out[dst] = (int32_t)g;
        }
      }
      run += total;
      __syncthreads();
    }
  }
  // zero-fill the tail [n_up, nalloc)
  const int64_t n_up = *n_up_dev;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t q = n_up + (int64_t)blockIdx.x * kThreads + threadIdx.x;
       q < nalloc; q += stride) {
    crd[q] = (T)0;
    crd[ld_crd + q] = (T)0;
    crd[2 * ld_crd + q] = (T)0;
    if (flux_out) flux_out[q] = (T)0;
    if (idx_out) idx_out[q] = -1;
  }
}

Mat3 load_mat(const double *rot9) {
  Mat3 R;
  for (int i = 0; i < 9; ++i) R.m[i] = rot9[i];
  return R;
}

}  // namespace
}  // namespace mv

using namespace mv;

extern "C" int mv_rotate(const double *eci, int64_t ld_eci, const double *rot9,
                         int64_t nsrc, void *topo, int topo_is_f64,
                         int64_t ld_topo, mv_stream_t stream) {
  MV_REQUIRE(eci && rot9 && topo, "mv_rotate: null pointer");
  MV_REQUIRE(nsrc >= 0 && ld_eci >= nsrc && ld_topo >= nsrc,
             "mv_rotate: bad sizes nsrc=%lld ld_eci=%lld ld_topo=%lld",
             (long long)nsrc, (long long)ld_eci, (long long)ld_topo);
  if (nsrc == 0) return MV_OK;
  Mat3 R = load_mat(rot9);
  dim3 grid((unsigned)cdiv(nsrc, kThreads));
  if (topo_is_f64)
    rotate_kernel<double><<<grid, kThreads, 0, as_stream(stream)>>>(
        eci, ld_eci, R, nsrc, (double *)topo, ld_topo);
  else
    rotate_kernel<float><<<grid, kThreads, 0, as_stream(stream)>>>(
        eci, ld_eci, R, nsrc, (float *)topo, ld_topo);
  return check_launch("mv_rotate");
}

namespace mv {
namespace {

// Per-time astrometry context of the BCRS corrections: observer direction
// from the Sun and distance (au), barycentric velocity (units of c) and
// sqrt(1 - |v|^2), bias-precession-nutation matrix.
struct BcrsCtx {
  double eh[3], em, v[3], bm1;
  Mat3 bpn;
};

constexpr double kSunSchwarzschildAu = 1.97412574336e-8;

// Light deflection by the Sun (limiter dlim), aberration, then the
// bias-precession-nutation matrix, for one unit vector: ICRS -> CIRS proper
// direction.  One source per thread, SoA in and out.
__global__ __launch_bounds__(kThreads) void bcrs_kernel(
    const double *__restrict__ eci, int64_t ld_eci, int64_t n, BcrsCtx c,
    double dlim, double *__restrict__ out, int64_t ld_out) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  double p[3] = {eci[i], eci[ld_eci + i], eci[2 * ld_eci + i]};
  // deflection: p + w p x (e x p), w = SRS / em / max(p . (p + e), dlim)
  double qpe = p[0] * (p[0] + c.eh[0]) + p[1] * (p[1] + c.eh[1]) +
               p[2] * (p[2] + c.eh[2]);
  if (qpe < dlim) qpe = dlim;
  const double w = kSunSchwarzschildAu / c.em / qpe;
  const double ex[3] = {c.eh[1] * p[2] - c.eh[2] * p[1],
                        c.eh[2] * p[0] - c.eh[0] * p[2],
                        c.eh[0] * p[1] - c.eh[1] * p[0]};
  const double d[3] = {p[1] * ex[2] - p[2] * ex[1], p[2] * ex[0] - p[0] * ex[2],
                       p[0] * ex[1] - p[1] * ex[0]};
  for (int k = 0; k < 3; ++k) p[k] += w * d[k];
  // aberration: bm1 p + w1 v + w2 (v - (p.v) bm1 p), renormalized
  const double pdv = p[0] * c.v[0] + p[1] * c.v[1] + p[2] * c.v[2];
  const double w1 = 1.0 + pdv / (1.0 + c.bm1);
  const double w2 = kSunSchwarzschildAu / c.em;
  double q[3], r2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double pb = c.bm1 * p[k];
    q[k] = pb + w1 * c.v[k] + w2 * (c.v[k] - pdv * pb);
    r2 += q[k] * q[k];
  }
  const double inv = 1.0 / sqrt(r2);
  for (int k = 0; k < 3; ++k) q[k] *= inv;
  for (int r = 0; r < 3; ++r)
    out[r * ld_out + i] =
        fma(c.bpn.m[3 * r + 2], q[2],
            fma(c.bpn.m[3 * r + 1], q[1], c.bpn.m[3 * r] * q[0]));
}

}  // namespace
}  // namespace mv

extern "C" int mv_bcrs_update(const double *eci, int64_t ld_eci,
                              int64_t nsrc, const double *ctx17, double dlim,
                              double *out, int64_t ld_out,
                              mv_stream_t stream) {
  MV_REQUIRE(eci && ctx17 && out, "mv_bcrs_update: null pointer");
  MV_REQUIRE(nsrc >= 0 && ld_eci >= nsrc && ld_out >= nsrc,
             "mv_bcrs_update: bad sizes nsrc=%lld ld_eci=%lld ld_out=%lld",
             (long long)nsrc, (long long)ld_eci, (long long)ld_out);
  MV_REQUIRE(ctx17[3] > 0.0 && ctx17[7] > 0.0,
             "mv_bcrs_update: context needs em > 0 and bm1 > 0");
  if (nsrc == 0) return MV_OK;
  BcrsCtx c;
  for (int k = 0; k < 3; ++k) {
    c.eh[k] = ctx17[k];
    c.v[k] = ctx17[4 + k];
  }
  c.em = ctx17[3];
  c.bm1 = ctx17[7];
  c.bpn = load_mat(ctx17 + 8);
  dim3 grid((unsigned)cdiv(nsrc, kThreads));
  bcrs_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(eci, ld_eci, nsrc, c,
                                                        dlim, out, ld_out);
  return check_launch("mv_bcrs_update");
}

extern "C" size_t mv_rotate_cull_workspace_bytes(int64_t chunk_len) {
  int64_t nb = chunk_len > 0 ? cdiv(chunk_len, kItemsPerBlock) : 1;
  return (size_t)(cdiv(nb * 4, 256) * 256);
}

extern "C" int mv_rotate_cull(const double *eci, int64_t ld_eci,
                              const double *rot9, const double *flux_in,
                              int64_t chunk0, int64_t chunk_len,
                              int64_t nalloc, void *crd_out, int64_t ld_crd,
                              void *flux_out, int out_is_f64, int32_t *idx_out,
                              int32_t *n_up_dev, void *workspace,
                              size_t ws_bytes, mv_stream_t stream) {
  MV_REQUIRE(eci && rot9 && crd_out && n_up_dev && workspace,
             "mv_rotate_cull: null pointer");
  MV_REQUIRE(!flux_out || flux_in,
             "mv_rotate_cull: flux_out requires flux_in");
  MV_REQUIRE(chunk0 >= 0 && chunk_len >= 0 && nalloc >= 0 &&
                 ld_crd >= nalloc && chunk0 + chunk_len <= ld_eci,
             "mv_rotate_cull: bad sizes chunk0=%lld chunk_len=%lld "
             "nalloc=%lld ld_crd=%lld ld_eci=%lld",
             (long long)chunk0, (long long)chunk_len, (long long)nalloc,
             (long long)ld_crd, (long long)ld_eci);
  MV_REQUIRE(chunk0 + chunk_len < (int64_t)INT32_MAX,
             "mv_rotate_cull: source index exceeds int32");
  MV_REQUIRE(ws_bytes >= mv_rotate_cull_workspace_bytes(chunk_len),
             "mv_rotate_cull: workspace too small (%zu < %zu)", ws_bytes,
             mv_rotate_cull_workspace_bytes(chunk_len));
  hipStream_t s = as_stream(stream);
  Mat3 R = load_mat(rot9);
  int32_t *counts = (int32_t *)workspace;
  int64_t nb = cdiv(chunk_len, kItemsPerBlock);
  if (nb > 0) {
    cull_count_kernel<<<(unsigned)nb, kThreads, 0, s>>>(eci, ld_eci, R, chunk0,
                                                        chunk_len, counts);
    if (int e = check_launch("mv_rotate_cull/count")) return e;
  }
  cull_scan_kernel<<<1, 1024, 0, s>>>(counts, nb, n_up_dev);
  if (int e = check_launch("mv_rotate_cull/scan")) return e;
  int64_t grid = nb > 0 ? nb : 1;
  if (out_is_f64)
    cull_scatter_kernel<double><<<(unsigned)grid, kThreads, 0, s>>>(
        eci, ld_eci, R, flux_in, chunk0, chunk_len, nalloc, counts, n_up_dev,
        (double *)crd_out, ld_crd, (double *)flux_out, idx_out, nb);
  else
    cull_scatter_kernel<float><<<(unsigned)grid, kThreads, 0, s>>>(
        eci, ld_eci, R, flux_in, chunk0, chunk_len, nalloc, counts, n_up_dev,
        (float *)crd_out, ld_crd, (float *)flux_out, idx_out, nb);
  return check_launch("mv_rotate_cull/scatter");
}

// ------------------------------------------------- coherency (polarized sky)

namespace mv {
namespace {

// Angle X between the equatorial basis (theta_hat, phi_hat at colatitude
// pi/2 - dec, longitude ra) and the alt/az basis (at za, az) of one source,
// returned as (cos X, sin X).  R_avg carries ICRS unit vectors to the
// (north, east, up) frame; the small rotation taking R_avg . frame onto the
// source's alt/az direction (Rodrigues about their cross product) completes
// it for the source (src/matvis/coordinates.py:288-491).  Parallel vectors
// give the identity, the limit of the rotation as its angle goes to zero.
__device__ void basis_angle(double ra, double dec, double e, double n,
                            const Mat3 &R, double &cx, double &sx) {
  // az / za from the direction cosines as enu_to_az_za (coordinates.py:31-88)
  const double lsqr = n * n + e * e;
  const double zeta = lsqr < 1.0 ? sqrt(1.0 - lsqr) : 0.0;
  const double az = atan2(e, n);
  const double za = 0.5 * kPi - asin(zeta);
  const double alt = 0.5 * kPi - za;
  double saz, caz, salt, calt, sra, cra, sdec, cdec;
  sincos(az, &saz, &caz);
  sincos(alt, &salt, &calt);
  sincos(ra, &sra, &cra);
  sincos(dec, &sdec, &cdec);
  const double f[3] = {cra * cdec, cdec * sra, sdec};      // ICRS direction
  const double h[3] = {caz * calt, calt * saz, salt};      // (N, E, U)
  double p[3];
  for (int i = 0; i < 3; ++i)
    p[i] = R.m[3 * i] * f[0] + R.m[3 * i + 1] * f[1] + R.m[3 * i + 2] * f[2];
  const double c3[3] = {p[1] * h[2] - p[2] * h[1], p[2] * h[0] - p[0] * h[2],
                        p[0] * h[1] - p[1] * h[0]};
  const double sp = sqrt(c3[0] * c3[0] + c3[1] * c3[1] + c3[2] * c3[2]);
  const double cp = p[0] * h[0] + p[1] * h[1] + p[2] * h[2];
  // perturbation P = I + sin(psi) K + (1 - cos(psi)) K^2, K = [axis]_x
  double P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (sp > 0.0) {
    const double ax = c3[0] / sp, ay = c3[1] / sp, az3 = c3[2] / sp;
    const double psi = atan2(sp, cp);
    double s, c;
    sincos(psi, &s, &c);
    const double K[9] = {0, -az3, ay, az3, 0, -ax, -ay, ax, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double k2 = 0.0;
        for (int q = 0; q < 3; ++q) k2 += K[3 * i + q] * K[3 * q + j];
        P[3 * i + j] += s * K[3 * i + j] + (1.0 - c) * k2;
      }
  }
  // theta_hat of the alt/az basis at (colatitude za, longitude az), carried
  // back to ICRS by (P R_avg)^T
  double sbe, cbe;
  sincos(0.5 * kPi - alt, &sbe, &cbe);
  const double th[3] = {caz * cbe, saz * cbe, -sbe};
  double q[3];
  for (int j = 0; j < 3; ++j)
    q[j] = P[j] * th[0] + P[3 + j] * th[1] + P[6 + j] * th[2];  // P^T th
  double b[3];
  for (int j = 0; j < 3; ++j)
    b[j] = R.m[j] * q[0] + R.m[3 + j] * q[1] + R.m[6 + j] * q[2];  // R^T q
  double sth, cth;
  sincos(0.5 * kPi - dec, &sth, &cth);
  cx = b[0] * cra * cth + b[1] * sra * cth - b[2] * sth;   // . theta_hat
  sx = -b[0] * sra + b[1] * cra;                          // . phi_hat
}

// C' = rot^T C rot with rot = [[c, s], [-s, c]], for real or complex
// (re, im pairs) entries of one 2x2 coherency matrix.
template <typename T, int NC>
__device__ __forceinline__ void rotate_2x2(const T *in, T *out, double c,
                                           double s) {
  for (int k = 0; k < NC; ++k) {
    const double a = in[0 * NC + k], b = in[1 * NC + k];
    const double d = in[2 * NC + k], e = in[3 * NC + k];
    // M = rot^T C
    const double m00 = c * a - s * d, m01 = c * b - s * e;
    const double m10 = s * a + c * d, m11 = s * b + c * e;
    out[0 * NC + k] = (T)(m00 * c - m01 * s);
    out[1 * NC + k] = (T)(m00 * s + m01 * c);
    out[2 * NC + k] = (T)(m10 * c - m11 * s);
    out[3 * NC + k] = (T)(m10 * s + m11 * c);
  }
}

// One thread per output position s: the kept source idx[s] (s < n_up) has
// its (nfreq, 2, 2) coherency rotated; positions [n_up, nalloc) are zeroed,
// as is a position whose index lies outside [0, nsrc).
template <typename T, typename C, int NC>
__global__ __launch_bounds__(kThreads) void coherency_kernel(
    const double *__restrict__ ra, const double *__restrict__ dec,
    int64_t nsrc, const int32_t *__restrict__ idx, const C *__restrict__ crd,
    int64_t ld_crd, const int32_t *__restrict__ n_up_dev, int64_t nalloc,
    Mat3 R, const T *__restrict__ coh, int64_t nfreq, T *__restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (s >= nalloc) return;
  const int64_t per = nfreq * 4 * NC;  // scalars per source
  T *o = out + s * per;
  const int64_t n_up = *n_up_dev;
  const int64_t g = s < n_up ? idx[s] : -1;
  if (g < 0 || g >= nsrc) {
    for (int64_t q = 0; q < per; ++q) o[q] = T(0);
    return;
  }
  double cx, sx;
  basis_angle(ra[g], dec[g], (double)crd[s], (double)crd[ld_crd + s], R, cx,
              sx);
  const T *in = coh + g * per;
  for (int64_t f = 0; f < nfreq; ++f)
    rotate_2x2<T, NC>(in + f * 4 * NC, o + f * 4 * NC, cx, sx);
}

template <typename C>
int launch_coherency(const double *ra, const double *dec, int64_t nsrc,
                     const int32_t *idx, const C *crd, int64_t ld_crd, const int32_t *n_up_dev,
                     int64_t nalloc, const Mat3 &R, const void *coh,
                     int64_t nfreq, int is_f64, int is_complex, void *out,
                     hipStream_t s) {
  const dim3 grid((unsigned)cdiv(nalloc, kThreads));
#define MV_COH(T, NC)                                                        \
  coherency_kernel<T, C, NC><<<grid, kThreads, 0, s>>>(                      \
      ra, dec, nsrc, idx, crd, ld_crd, n_up_dev, nalloc, R, (const T *)coh,  \
      nfreq, (T *)out)
  if (is_f64 && is_complex) MV_COH(double, 2);
  else if (is_f64) MV_COH(double, 1);
  else if (is_complex) MV_COH(float, 2);
  else MV_COH(float, 1);
#undef MV_COH
  return check_launch("mv_coherency_rotate");
}

}  // namespace
}  // namespace mv

extern "C" int mv_coherency_rotate(const double *ra, const double *dec,
                                   int64_t nsrc, const int32_t *idx,
                                   const void *crd, int64_t ld_crd,
                                   int crd_is_f64,
                                   const int32_t *n_up_dev, int64_t nalloc,
                                   const double *r_avg9, const void *coh,
                                   int64_t nfreq, int coh_is_f64,
                                   int coh_is_complex, void *out,
                                   mv_stream_t stream) {
  using namespace mv;
  MV_REQUIRE(ra && dec && crd && n_up_dev && r_avg9 && out,
             "mv_coherency_rotate: null pointer");
  MV_REQUIRE(nsrc >= 0 && nalloc >= 0 && ld_crd >= nalloc && nfreq >= 1,
             "mv_coherency_rotate: bad sizes nsrc=%lld nalloc=%lld "
             "ld_crd=%lld nfreq=%lld",
             (long long)nsrc, (long long)nalloc, (long long)ld_crd,
             (long long)nfreq);
  MV_REQUIRE(nalloc == 0 || (idx && coh),
             "mv_coherency_rotate: null idx or coherency");
  if (nalloc == 0) return MV_OK;
  const Mat3 R = load_mat(r_avg9);
  hipStream_t s = as_stream(stream);
  return crd_is_f64
             ? launch_coherency(ra, dec, nsrc, idx, (const double *)crd,
                                ld_crd, n_up_dev, nalloc, R, coh, nfreq, coh_is_f64,
                                coh_is_complex, out, s)
             : launch_coherency(ra, dec, nsrc, idx, (const float *)crd,
                                ld_crd, n_up_dev, nalloc, R, coh, nfreq, coh_is_f64,
                                coh_is_complex, out, s);
}
