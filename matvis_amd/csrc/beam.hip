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
  const double daz = grid[3 * b + 0], dza = grid[3 * b + 1],
               azmin = grid[3 * b + 2];
  const double cx = za / dza;            // along nza
  const double cy = (az - azmin) / daz;  // along naz
  const bool inside = cx >= 0.0 && cx <= (double)(nza - 1) && cy >= 0.0 &&
                      cy <= (double)(naz - 1);
  double vr = 0.0, vi = 0.0;
  if (inside) {
    double wx[Taps<ORDER>::N], wy[Taps<ORDER>::N];
    int rx[Taps<ORDER>::N], ry[Taps<ORDER>::N];
    const int x0 = Taps<ORDER>::weights(cx, wx);
    const int y0 = Taps<ORDER>::weights(cy, wy);
#pragma unroll
    for (int k = 0; k < Taps<ORDER>::N; ++k) {
      rx[k] = mirror(x0 + k, nza);
      ry[k] = mirror(y0 + k, naz);
    }
    const cplx<T> *pl =
        planes + ((int64_t)(b * nax + ax) * nfeed + fd) * plane_sz;
#pragma unroll
    for (int i = 0; i < Taps<ORDER>::N; ++i) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int j = 0; j < Taps<ORDER>::N; ++j) {
        const cplx<T> g = pl[(int64_t)rx[i] * naz + ry[j]];
        sr += wy[j] * (double)g.re;
        si += wy[j] * (double)g.im;
      }
      vr += wx[i] * sr;
      vi += wx[i] * si;
    }
  }
  cplx<T> r;
  if (is_power) {
    // complex sqrt of (vr, +0): the power beam's final cp.sqrt
    if (vr >= 0.0) {
      r.re = (T)sqrt(vr);
      r.im = (T)0;
    } else {
      r.re = (T)0;
      r.im = (T)sqrt(-vr);
    }
  } else {
    r.re = (T)vr;
    r.im = (T)vi;
  }
  *dst = r;
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
