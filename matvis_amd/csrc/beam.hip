// (b) Beam evaluation at topocentric source positions.
//
// Grid beams follow map_coordinates(mode='constant', cval=0) as the reference
// calls it (src/matvis/gpu/beams.py:184-192): coordinates outside [0, n-1] on
// either axis give 0; inside, taps that fall past an edge are folded back by
// mirror symmetry; order 3 evaluates a cubic B-spline on planes that were
// prefiltered with the mirror boundary (mv_beam_prefilter).
//
// The beam grid (a few MB at most) stays L2-resident; each source reads its
// 4 (bilinear) or 16 (bicubic) taps per plane straight from it.
#include "common.h"

namespace mv {
namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- prefilter

// Cubic B-spline pole and gain.
constexpr double kPole = -0.26794919243112270;  // sqrt(3) - 2
constexpr double kGain = 6.0;

// In-place 1-D cubic B-spline coefficient filter with the mirror boundary
// (whole-sample symmetric extension), on n complex values spaced `stride`.
template <typename T>
__device__ void bspline_filter_line(cplx<T> *c, int n, int64_t stride) {
  if (n < 2) return;
  const double z = kPole;
  double zn1 = 1.0;
  for (int i = 0; i < n - 1; ++i) zn1 *= z;
  // gain + causal initialisation
  double c0r = kGain * (double)c[0].re, c0i = kGain * (double)c[0].im;
  double lr = kGain * (double)c[(int64_t)(n - 1) * stride].re;
  double li = kGain * (double)c[(int64_t)(n - 1) * stride].im;
  double sr = c0r + zn1 * lr, si = c0i + zn1 * li;
  double zi = z;
  for (int i = 1; i < n - 1; ++i) {
    double ar = kGain * (double)c[(int64_t)i * stride].re;
    double ai = kGain * (double)c[(int64_t)i * stride].im;
    double br = kGain * (double)c[(int64_t)(n - 1 - i) * stride].re;
    double bi = kGain * (double)c[(int64_t)(n - 1 - i) * stride].im;
    sr += zi * (ar + zn1 * br);
    si += zi * (ai + zn1 * bi);
    zi *= z;
  }
  double inv = 1.0 / (1.0 - zn1 * zn1);
  double pr = sr * inv, pi = si * inv;
  // causal pass; keep the causal coefficients in the array (in FP64 registers
  // only for the running value, the array holds T)
  c[0].re = (T)pr;
  c[0].im = (T)pi;
  double prev_r = pr, prev_i = pi, prev2_r = 0, prev2_i = 0;
  for (int i = 1; i < n; ++i) {
    double vr = kGain * (double)c[(int64_t)i * stride].re + z * prev_r;
    double vi = kGain * (double)c[(int64_t)i * stride].im + z * prev_i;
    prev2_r = prev_r;
    prev2_i = prev_i;
    prev_r = vr;
    prev_i = vi;
    c[(int64_t)i * stride].re = (T)vr;
    c[(int64_t)i * stride].im = (T)vi;
  }
  // anticausal initialisation from the last two causal values
  double nr = (z * prev2_r + prev_r) * z / (z * z - 1.0);
  double ni = (z * prev2_i + prev_i) * z / (z * z - 1.0);
  c[(int64_t)(n - 1) * stride].re = (T)nr;
  c[(int64_t)(n - 1) * stride].im = (T)ni;
  double nxt_r = nr, nxt_i = ni;
  for (int i = n - 2; i >= 0; --i) {
    double cr = (double)c[(int64_t)i * stride].re;
    double ci = (double)c[(int64_t)i * stride].im;
    double vr = z * (nxt_r - cr);
    double vi = z * (nxt_i - ci);
    c[(int64_t)i * stride].re = (T)vr;
    c[(int64_t)i * stride].im = (T)vi;
    nxt_r = vr;
    nxt_i = vi;
  }
}

// One thread per line.  axis 1: lines along az (length naz, stride 1);
// axis 0: lines along za (length nza, stride naz).
template <typename T>
__global__ __launch_bounds__(kThreads) void prefilter_kernel(
    cplx<T> *planes, int64_t nplanes, int nza, int naz, int axis) {
  int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  int64_t plane_sz = (int64_t)nza * naz;
  if (axis == 1) {
    if (t >= nplanes * nza) return;
    int64_t p = t / nza, r = t % nza;
    bspline_filter_line<T>(planes + p * plane_sz + r * naz, naz, 1);
  } else {
    if (t >= nplanes * naz) return;
    int64_t p = t / naz, col = t % naz;
    bspline_filter_line<T>(planes + p * plane_sz + col, nza, naz);
  }
}

// ---------------------------------------------------------------- geometry

// (tx, ty) -> (az, za), uvbeam orientation, az in [0, 2 pi)
// (src/matvis/coordinates.py:75-88 with orientation="uvbeam").
__device__ __forceinline__ void enu_to_az_za_uvbeam(double e, double n,
                                                    double &az, double &za) {
  double lsqr = n * n + e * e;
  double zeta = lsqr < 1.0 ? sqrt(1.0 - lsqr) : 0.0;
  double a = 0.5 * kPi - atan2(e, n);
  za = 0.5 * kPi - asin(zeta);
  // numpy mod: result takes the sign of the divisor
  double m = fmod(a, 2.0 * kPi);
  if (m != 0.0 && m < 0.0) m += 2.0 * kPi;
  az = m;
}

// Mirror-fold an index into [0, n-1] (whole-sample symmetric).
__device__ __forceinline__ int mirror(int i, int n) {
  if (n == 1) return 0;
  const int period = 2 * (n - 1);
  i = i < 0 ? -i : i;
  i = i % period;
  return i >= n ? period - i : i;
}

template <int ORDER> struct Taps;
template <> struct Taps<1> {
  static constexpr int N = 2;
  __device__ static int weights(double x, double *w) {
    double f = floor(x);
    double t = x - f;
    w[0] = 1.0 - t;
    w[1] = t;
    return (int)f;
  }
};
template <> struct Taps<3> {
  static constexpr int N = 4;
  __device__ static int weights(double x, double *w) {
    double f = floor(x);
    double t = x - f;
    double t2 = t * t, t3 = t2 * t;
    double u = 1.0 - t;
    w[0] = u * u * u / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
    return (int)f - 1;
  }
};

// The interpolated value at (az, za) of one plane: `row(x)` returns the
// plane's grid row for the unfolded za index x (the caller folds it, or
// reads it from rows staged in LDS), so both interpolation kernels run the
// same arithmetic in the same order.
template <typename T, int ORDER, class Row>
__device__ __forceinline__ cplx<T> interp_value(double az, double za,
                                                double daz, double dza,
                                                double azmin, int nza,
                                                int naz, int is_power,
                                                Row row) {
  const double cx = za / dza;            // along nza
  const double cy = (az - azmin) / daz;  // along naz
  const bool inside = cx >= 0.0 && cx <= (double)(nza - 1) && cy >= 0.0 &&
                      cy <= (double)(naz - 1);
  double vr = 0.0, vi = 0.0;
  if (inside) {
    double wx[Taps<ORDER>::N], wy[Taps<ORDER>::N];
    int ry[Taps<ORDER>::N];
    const int x0 = Taps<ORDER>::weights(cx, wx);
    const int y0 = Taps<ORDER>::weights(cy, wy);
#pragma unroll
    for (int k = 0; k < Taps<ORDER>::N; ++k) ry[k] = mirror(y0 + k, naz);
#pragma unroll
    for (int i = 0; i < Taps<ORDER>::N; ++i) {
      const cplx<T> *r = row(x0 + i);
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int j = 0; j < Taps<ORDER>::N; ++j) {
        const cplx<T> g = r[ry[j]];
        sr += wy[j] * (double)g.re;
        si += wy[j] * (double)g.im;
      }
      vr += wx[i] * sr;
      vi += wx[i] * si;
    }
  }
  cplx<T> v;
  if (is_power) {
    // complex sqrt of (vr, +0): the power beam's final cp.sqrt
    if (vr >= 0.0) {
      v.re = (T)sqrt(vr);
      v.im = (T)0;
    } else {
      v.re = (T)0;
      v.im = (T)sqrt(-vr);
    }
  } else {
    v.re = (T)vr;
    v.im = (T)vi;
  }
  return v;
}

// One thread per (source, plane): grid.y = nbeam x nfeed x nax planes, so a
// polarized beam runs 4x the waves of one thread per source (at 60k
// sources that is what keeps the tap gathers' latency covered; the planes,
// about 1 MB for one polarized 360 x 91 complex64 beam, stay in L2).
// AZZA: positions come as (az, za) arrays instead of topocentric vectors.
template <typename T, typename TC, int ORDER, bool AZZA = false>
__global__ __launch_bounds__(kThreads) void interp_kernel(
    const cplx<T> *__restrict__ planes, int nbeam, int nfeed, int nax, int nza,
    int naz, const double *__restrict__ grid, int is_power,
    const TC *__restrict__ crd, int64_t ld_crd, int64_t count,
    const int32_t *__restrict__ n_dev, cplx<T> *__restrict__ out,
    int64_t ld_out, const double *__restrict__ az_in,
    const double *__restrict__ za_in) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= count) return;
  // output plane (b, fd, ax) and its input plane (b, ax, fd)
  const int q = blockIdx.y;
  const int ax = q % nax, fd = (q / nax) % nfeed, b = q / (nax * nfeed);
  cplx<T> *dst = out + (int64_t)q * ld_out + p;
  const int64_t n = n_dev ? min((int64_t)*n_dev, count) : count;
  if (p >= n) {
    *dst = cplx<T>{(T)0, (T)0};
    return;
  }
  double az, za;
  if (AZZA) {
    az = az_in[p];
    za = za_in[p];
  } else {
    enu_to_az_za_uvbeam((double)crd[p], (double)crd[ld_crd + p], az, za);
  }
  const int64_t plane_sz = (int64_t)nza * naz;
  const cplx<T> *pl = planes + ((int64_t)(b * nax + ax) * nfeed + fd) * plane_sz;
  *dst = interp_value<T, ORDER>(
      az, za, grid[3 * b + 0], grid[3 * b + 1], grid[3 * b + 2], nza, naz,
      is_power, [&](int x) { return pl + (int64_t)mirror(x, nza) * naz; });
}

// ---------------------------------------------------- za bands staged in LDS
//
// The alternative to interp_kernel's tap gathers: the sources are binned
// by za band and each workgroup stages its band's rows of one plane in LDS,
// then interpolates up to kBandChunk of the band's sources from there (the
// rows are read once a workgroup instead of by every source's scattered
// taps; on MI355X the gathers, served by L2 and the Infinity Cache, measured
// faster at every grid size whose rows fit, so callers ask for this path
// explicitly).  Band b holds the sources whose first za tap row
// floor(cx) lies in [b B, (b + 1) B), B = band_rows(naz): their bicubic taps
// reach rows b B - 1 .. b B + B + 1 (bilinear b B .. b B + B), the B + 3
// rows staged (folded at the grid's ends as the gather kernel folds them).
// Results go to each source's own slot, so the order the binning's atomics
// give inside a band does not change them, and they equal interp_kernel's
// (same arithmetic, interp_value) when every beam shares beam 0's za spacing,
// which the bands use.

// staged bytes a workgroup: gfx950 gives a workgroup up to 160 KB of LDS;
// 128 KB fits five complex128 rows of a 0.25 deg grid (1443 columns)
constexpr int kBandLds = 128 * 1024;
constexpr int kBandChunk = 512;      // sources a workgroup (2 a thread)
constexpr int kMaxBands = 1024;      // bands a workgroup histograms in LDS

__host__ __device__ inline int band_rows(int naz, int elem_bytes) {
  return kBandLds / (naz * elem_bytes) - 3;  // below 1: rows do not fit
}

struct BandWs {
  int32_t *cnt;     // [nbands] sources per band
  int32_t *off;     // [nbands + 1] band starts in `sorted`
  int32_t *boff;    // [nbands + 1] first workgroup of each band
  double *az, *za;  // [count] the sources' (az, za)
  int32_t *band;    // [count] band, or -1 (past the count or off the grid)
  int32_t *pos;     // [count] rank inside the band
  int32_t *sorted;  // [count] source indices grouped by band
};

__global__ __launch_bounds__(kThreads) void band_zero_kernel(int32_t *cnt,
                                                             int nbands) {
  for (int i = threadIdx.x; i < nbands; i += kThreads) cnt[i] = 0;
}

// (az, za) of each source, its band and rank; sources past the count or off
// the za range of the grid get zero outputs here, as interp_kernel gives them.
// Ranks come from a per-workgroup LDS histogram and one global add per band
// the workgroup touches (per-source global adds on a few band counters
// serialize).
template <typename T, typename TC>
__global__ __launch_bounds__(kThreads) void band_assign_kernel(
    const TC *__restrict__ crd, int64_t ld_crd, int64_t count,
    const int32_t *__restrict__ n_dev, const double *__restrict__ grid,
    int nza, int B, int nbands, BandWs ws, cplx<T> *__restrict__ out,
    int64_t ld_out, int nplanes) {
  __shared__ int32_t hist[kMaxBands], base[kMaxBands];
  for (int i = threadIdx.x; i < nbands; i += kThreads) hist[i] = 0;
  __syncthreads();
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t n = n_dev ? min((int64_t)*n_dev, count) : count;
  int band = -1, local = 0;
  if (p < n) {
    double az, za;
    enu_to_az_za_uvbeam((double)crd[p], (double)crd[ld_crd + p], az, za);
    ws.az[p] = az;
    ws.za[p] = za;
    const double cx = za / grid[1];
    if (cx >= 0.0 && cx <= (double)(nza - 1)) {
      band = (int)floor(cx) / B;
      local = atomicAdd(&hist[band], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbands; i += kThreads)
    if (hist[i] > 0) base[i] = atomicAdd(&ws.cnt[i], hist[i]);
  __syncthreads();
  if (p >= count) return;
  ws.band[p] = band;
  if (band >= 0) {
    ws.pos[p] = base[band] + local;
  } else {
    for (int q = 0; q < nplanes; ++q)
      out[(int64_t)q * ld_out + p] = cplx<T>{(T)0, (T)0};
  }
}

// Band starts and workgroup starts (one thread: a few hundred bands at most).
__global__ __launch_bounds__(64) void band_scan_kernel(BandWs ws, int nbands) {
  if (threadIdx.x != 0) return;
  int32_t o = 0, bo = 0;
  for (int b = 0; b < nbands; ++b) {
    ws.off[b] = o;
    ws.boff[b] = bo;
    o += ws.cnt[b];
    bo += (ws.cnt[b] + kBandChunk - 1) / kBandChunk;
  }
  ws.off[nbands] = o;
  ws.boff[nbands] = bo;
}

__global__ __launch_bounds__(kThreads) void band_scatter_kernel(BandWs ws,
                                                                int64_t count) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= count) return;
  const int b = ws.band[p];
  if (b >= 0) ws.sorted[ws.off[b] + ws.pos[p]] = (int32_t)p;
}

// grid (cdiv(count, kBandChunk) + nbands, planes): workgroup x is chunk
// x - boff[b] of band b (boff[b] <= x < boff[b + 1]); those past the last
// band's chunks return at once.
template <typename T, int ORDER>
__global__ __launch_bounds__(kThreads) void interp_band_kernel(
    const cplx<T> *__restrict__ planes, int nfeed, int nax, int nza, int naz,
    const double *__restrict__ grid, int is_power, int B, int nbands,
    BandWs ws, cplx<T> *__restrict__ out, int64_t ld_out) {
  __shared__ cplx<T> rows[kBandLds / sizeof(cplx<T>)];
  const int x = blockIdx.x;
  if (x >= ws.boff[nbands]) return;
  // the band: the last b with boff[b] <= x (empty bands share a start)
  int lo = 0, hi = nbands - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ws.boff[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  const int band = lo;
  const int q = blockIdx.y;
  const int ax = q % nax, fd = (q / nax) % nfeed, b = q / (nax * nfeed);
  const int64_t plane_sz = (int64_t)nza * naz;
  const cplx<T> *pl = planes + ((int64_t)(b * nax + ax) * nfeed + fd) * plane_sz;
  const int base = band * B - 1;  // unfolded za index of staged row 0
  const int nstage = (B + 3) * naz;
  for (int i = threadIdx.x; i < nstage; i += kThreads) {
    const int r = i / naz, c = i - r * naz;
    rows[i] = pl[(int64_t)mirror(base + r, nza) * naz + c];
  }
  __syncthreads();
  const int64_t first = ws.off[band] + (int64_t)(x - ws.boff[band]) * kBandChunk;
  const int64_t last = min(first + kBandChunk, (int64_t)ws.off[band + 1]);
  const double daz = grid[3 * b + 0], azmin = grid[3 * b + 2];
  const double dza = grid[1];  // the bands' za spacing (beam 0's)
  for (int64_t s = first + threadIdx.x; s < last; s += kThreads) {
    const int32_t p = ws.sorted[s];
    out[(int64_t)q * ld_out + p] = interp_value<T, ORDER>(
        ws.az[p], ws.za[p], daz, dza, azmin, nza, naz, is_power,
        [&](int xr) {
          // staged rows cover the band's taps; any other index is a caller
          // whose beams do not share the za spacing: read row 0, in bounds
          const int r = xr - base;
          return rows + (r >= 0 && r < B + 3 ? r : 0) * naz;
        });
  }
}

template <typename T, typename TC>
__global__ __launch_bounds__(kThreads) void gaussian_kernel(
    const double *__restrict__ sigma, int nbeam, int polarized,
    const TC *__restrict__ crd, int64_t ld_crd, int64_t count,
    const int32_t *__restrict__ n_dev, cplx<T> *__restrict__ out,
    int64_t ld_out) {
  int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= count) return;
  const int64_t n = n_dev ? min((int64_t)*n_dev, count) : count;
  const int nplane = polarized ? 4 : 1;
  double az, za;
  bool live = p < n;
  if (live) enu_to_az_za_uvbeam((double)crd[p], (double)crd[ld_crd + p], az, za);
  for (int b = 0; b < nbeam; ++b) {
    double e = 0.0;
    if (live) {
      double s = sigma[b];
      e = exp(-(za * za) / (2.0 * s * s));
      if (polarized) e = e / sqrt(2.0);
    }
    for (int q = 0; q < nplane; ++q)
      out[((int64_t)b * nplane + q) * ld_out + p] = cplx<T>{(T)e, (T)0};
  }
}

}  // namespace
}  // namespace mv

using namespace mv;

extern "C" int mv_beam_prefilter(void *planes, int is_f64, int64_t nplanes,
                                 int nza, int naz, mv_stream_t stream) {
  MV_REQUIRE(planes, "mv_beam_prefilter: null pointer");
  MV_REQUIRE(nplanes >= 0 && nza >= 2 && naz >= 2,
             "mv_beam_prefilter: need a grid of at least 2x2 (got %dx%d)",
             nza, naz);
  hipStream_t s = as_stream(stream);
  for (int axis = 1; axis >= 0; --axis) {
    int64_t lines = nplanes * (axis == 1 ? nza : naz);
    if (lines == 0) continue;
    dim3 g((unsigned)cdiv(lines, kThreads));
    if (is_f64)
      prefilter_kernel<double><<<g, kThreads, 0, s>>>(
          (cplx<double> *)planes, nplanes, nza, naz, axis);
    else
      prefilter_kernel<float><<<g, kThreads, 0, s>>>(
          (cplx<float> *)planes, nplanes, nza, naz, axis);
    if (int e = check_launch("mv_beam_prefilter")) return e;
  }
  return MV_OK;
}

template <typename T, typename TC>
static void launch_interp(int order, dim3 g, hipStream_t s, const void *planes,
                          int nbeam, int nfeed, int nax, int nza, int naz,
                          const double *grid, int is_power, const void *crd,
                          int64_t ld_crd, int64_t count, const int32_t *n_dev,
                          void *out, int64_t ld_out) {
  if (order == 1)
    interp_kernel<T, TC, 1><<<g, kThreads, 0, s>>>(
        (const cplx<T> *)planes, nbeam, nfeed, nax, nza, naz, grid, is_power,
        (const TC *)crd, ld_crd, count, n_dev, (cplx<T> *)out, ld_out,
        nullptr, nullptr);
  else
    interp_kernel<T, TC, 3><<<g, kThreads, 0, s>>>(
        (const cplx<T> *)planes, nbeam, nfeed, nax, nza, naz, grid, is_power,
        (const TC *)crd, ld_crd, count, n_dev, (cplx<T> *)out, ld_out,
        nullptr, nullptr);
}

extern "C" int mv_beam_interp(const void *planes, int is_f64, int nbeam,
                              int nfeed, int nax, int nza, int naz,
                              const double *grid, int order, int is_power,
                              const void *crd, int crd_is_f64, int64_t ld_crd,
                              int64_t count, const int32_t *n_dev, void *out,
                              int64_t ld_out, mv_stream_t stream) {
  MV_REQUIRE(planes && grid && crd && out, "mv_beam_interp: null pointer");
  MV_REQUIRE(order == 1 || order == 3,
             "mv_beam_interp: order must be 1 or 3 (got %d)", order);
  MV_REQUIRE(nbeam >= 1 && nfeed >= 1 && nax >= 1 && nza >= 1 && naz >= 1 &&
                 count >= 0 && ld_crd >= count && ld_out >= count,
             "mv_beam_interp: bad sizes");
  MV_REQUIRE((int64_t)nbeam * nfeed * nax <= 65535,
             "mv_beam_interp: too many planes");
  if (count == 0) return MV_OK;
  dim3 g((unsigned)cdiv(count, kThreads), (unsigned)(nbeam * nfeed * nax));
  hipStream_t s = as_stream(stream);
  if (is_f64) {
    if (crd_is_f64)
      launch_interp<double, double>(order, g, s, planes, nbeam, nfeed, nax,
                                    nza, naz, grid, is_power, crd, ld_crd,
                                    count, n_dev, out, ld_out);
    else
      launch_interp<double, float>(order, g, s, planes, nbeam, nfeed, nax,
                                   nza, naz, grid, is_power, crd, ld_crd,
                                   count, n_dev, out, ld_out);
  } else {
    if (crd_is_f64)
      launch_interp<float, double>(order, g, s, planes, nbeam, nfeed, nax, nza,
                                   naz, grid, is_power, crd, ld_crd, count,
                                   n_dev, out, ld_out);
    else
      launch_interp<float, float>(order, g, s, planes, nbeam, nfeed, nax, nza,
                                  naz, grid, is_power, crd, ld_crd, count,
                                  n_dev, out, ld_out);
  }
  return check_launch("mv_beam_interp");
}

namespace mv {
namespace {
// Workspace layout of the za-band path: 256-byte aligned arrays.
struct BandLayout {
  size_t cnt, off, boff, az, za, band, pos, sorted, total;
};
BandLayout band_layout(int nbands, int64_t count) {
  BandLayout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return at;
  };
  L.cnt = take(4 * (size_t)nbands);
  L.off = take(4 * ((size_t)nbands + 1));
  L.boff = take(4 * ((size_t)nbands + 1));
  L.az = take(8 * (size_t)count);
  L.za = take(8 * (size_t)count);
  L.band = take(4 * (size_t)count);
  L.pos = take(4 * (size_t)count);
  L.sorted = take(4 * (size_t)count);
  L.total = o;
  return L;
}

template <typename T, typename TC>
int launch_banded(int order, hipStream_t s, const void *planes, int nbeam,
                  int nfeed, int nax, int nza, int naz, const double *grid,
                  int is_power, const void *crd, int64_t ld_crd, int64_t count,
                  const int32_t *n_dev, void *out, int64_t ld_out, int B,
                  void *workspace) {
  const int nbands = (int)cdiv((int64_t)nza, (int64_t)B);
  const BandLayout L = band_layout(nbands, count);
  char *w = (char *)workspace;
  BandWs ws = {(int32_t *)(w + L.cnt), (int32_t *)(w + L.off),
               (int32_t *)(w + L.boff), (double *)(w + L.az),
               (double *)(w + L.za), (int32_t *)(w + L.band),
               (int32_t *)(w + L.pos), (int32_t *)(w + L.sorted)};
  const int nplanes = nbeam * nfeed * nax;
  const unsigned gs = (unsigned)cdiv(count, (int64_t)kThreads);
  band_zero_kernel<<<1, kThreads, 0, s>>>(ws.cnt, nbands);
  band_assign_kernel<T, TC><<<gs, kThreads, 0, s>>>(
      (const TC *)crd, ld_crd, count, n_dev, grid, nza, B, nbands, ws,
      (cplx<T> *)out, ld_out, nplanes);
  band_scan_kernel<<<1, 64, 0, s>>>(ws, nbands);
  band_scatter_kernel<<<gs, kThreads, 0, s>>>(ws, count);
  dim3 g((unsigned)(cdiv(count, (int64_t)kBandChunk) + nbands),
         (unsigned)nplanes);
  if (order == 1)
    interp_band_kernel<T, 1><<<g, kThreads, 0, s>>>(
        (const cplx<T> *)planes, nfeed, nax, nza, naz, grid, is_power, B,
        nbands, ws, (cplx<T> *)out, ld_out);
  else
    interp_band_kernel<T, 3><<<g, kThreads, 0, s>>>(
        (const cplx<T> *)planes, nfeed, nax, nza, naz, grid, is_power, B,
        nbands, ws, (cplx<T> *)out, ld_out);
  return check_launch("mv_beam_interp_banded");
}
}  // namespace
}  // namespace mv

extern "C" size_t mv_beam_interp_banded_workspace_bytes(int is_f64, int nza,
                                                        int naz,
                                                        int64_t count) {
  if (nza < 1 || naz < 1 || count < 0) return 0;
  const int B = band_rows(naz, is_f64 ? 16 : 8);
  if (B < 1 || cdiv((int64_t)nza, (int64_t)B) > kMaxBands)
    return 0;  // the gather path runs; no workspace
  return band_layout((int)cdiv((int64_t)nza, (int64_t)B), count).total;
}

extern "C" int mv_beam_interp_banded(
    const void *planes, int is_f64, int nbeam, int nfeed, int nax, int nza,
    int naz, const double *grid, int order, int is_power, const void *crd,
    int crd_is_f64, int64_t ld_crd, int64_t count, const int32_t *n_dev,
    void *out, int64_t ld_out, void *workspace, size_t ws_bytes,
    mv_stream_t stream) {
  MV_REQUIRE(planes && grid && crd && out,
             "mv_beam_interp_banded: null pointer");
  MV_REQUIRE(order == 1 || order == 3,
             "mv_beam_interp_banded: order must be 1 or 3 (got %d)", order);
  MV_REQUIRE(nbeam >= 1 && nfeed >= 1 && nax >= 1 && nza >= 1 && naz >= 1 &&
                 count >= 0 && count <= INT32_MAX && ld_crd >= count &&
                 ld_out >= count,
             "mv_beam_interp_banded: bad sizes");
  MV_REQUIRE((int64_t)nbeam * nfeed * nax <= 65535,
             "mv_beam_interp_banded: too many planes");
  if (count == 0) return MV_OK;
  const int B = band_rows(naz, is_f64 ? 16 : 8);
  // a band's rows do not fit in LDS, or more bands than a workgroup's
  // histogram holds: the gather path
  if (B < 1 || cdiv((int64_t)nza, (int64_t)B) > kMaxBands)
    return mv_beam_interp(planes, is_f64, nbeam, nfeed, nax, nza, naz, grid,
                          order, is_power, crd, crd_is_f64, ld_crd, count,
                          n_dev, out, ld_out, stream);
  const size_t need = mv_beam_interp_banded_workspace_bytes(is_f64, nza, naz,
                                                            count);
  MV_REQUIRE(workspace && ws_bytes >= need,
             "mv_beam_interp_banded: workspace too small (%zu < %zu)",
             ws_bytes, need);
  hipStream_t s = as_stream(stream);
#define MV_BANDED(T, TC)                                                     \
  return launch_banded<T, TC>(order, s, planes, nbeam, nfeed, nax, nza, naz, \
                              grid, is_power, crd, ld_crd, count, n_dev, out, \
                              ld_out, B, workspace)
  if (is_f64) {
    if (crd_is_f64) MV_BANDED(double, double);
    MV_BANDED(double, float);
  }
  if (crd_is_f64) MV_BANDED(float, double);
  MV_BANDED(float, float);
#undef MV_BANDED
}

extern "C" int mv_beam_interp_azza(const void *planes, int is_f64, int nbeam,
                                   int nfeed, int nax, int nza, int naz,
                                   const double *grid, int order, int is_power,
                                   const double *az, const double *za,
                                   int64_t count, void *out, int64_t ld_out,
                                   mv_stream_t stream) {
  MV_REQUIRE(planes && grid && az && za && out,
             "mv_beam_interp_azza: null pointer");
  MV_REQUIRE(order == 1 || order == 3,
             "mv_beam_interp_azza: order must be 1 or 3 (got %d)", order);
  MV_REQUIRE(nbeam >= 1 && nfeed >= 1 && nax >= 1 && nza >= 1 && naz >= 1 &&
                 count >= 0 && ld_out >= count,
             "mv_beam_interp_azza: bad sizes");
  MV_REQUIRE((int64_t)nbeam * nfeed * nax <= 65535,
             "mv_beam_interp_azza: too many planes");
  if (count == 0) return MV_OK;
  dim3 g((unsigned)cdiv(count, kThreads), (unsigned)(nbeam * nfeed * nax));
  hipStream_t s = as_stream(stream);
#define MV_AZZA(T, O)                                                        \
  interp_kernel<T, double, O, true><<<g, kThreads, 0, s>>>(                  \
      (const cplx<T> *)planes, nbeam, nfeed, nax, nza, naz, grid, is_power,  \
      nullptr, 0, count, nullptr, (cplx<T> *)out, ld_out, az, za)
  if (is_f64) {
    if (order == 1) MV_AZZA(double, 1);
    else MV_AZZA(double, 3);
  } else {
    if (order == 1) MV_AZZA(float, 1);
    else MV_AZZA(float, 3);
  }
#undef MV_AZZA
  return check_launch("mv_beam_interp_azza");
}

extern "C" int mv_beam_gaussian(const double *sigma, int nbeam, int polarized,
                                const void *crd, int crd_is_f64,
                                int64_t ld_crd, int64_t count,
                                const int32_t *n_dev, void *out,
                                int out_is_f64, int64_t ld_out,
                                mv_stream_t stream) {
  MV_REQUIRE(sigma && crd && out, "mv_beam_gaussian: null pointer");
  MV_REQUIRE(nbeam >= 1 && count >= 0 && ld_crd >= count && ld_out >= count,
             "mv_beam_gaussian: bad sizes");
  if (count == 0) return MV_OK;
  dim3 g((unsigned)cdiv(count, kThreads));
  hipStream_t s = as_stream(stream);
#define MV_GAUSS(T, TC)                                                    \
  gaussian_kernel<T, TC><<<g, kThreads, 0, s>>>(                           \
      sigma, nbeam, polarized, (const TC *)crd, ld_crd, count, n_dev,      \
      (cplx<T> *)out, ld_out)
  if (out_is_f64) {
    if (crd_is_f64) MV_GAUSS(double, double);
    else MV_GAUSS(double, float);
  } else {
    if (crd_is_f64) MV_GAUSS(float, double);
    else MV_GAUSS(float, float);
  }
#undef MV_GAUSS
  return check_launch("mv_beam_gaussian");
}
