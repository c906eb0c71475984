// Physical-space closure of a reconstructed eigenfield on a Yee window: the cropped Yee
// curl-curl action, the scalar permittivity on the three edge families, the pointwise
// energy and mass densities, the column-normalised local Gaussian Rayleigh quotient (LWRQ)
// and the residual norms (exp3/compute_yee_lwrq_metrics_exp4.m:16-94,
// exp3/apply_cropped_A_exp3.m:8-56, exp3/sample_scalar_yee_material_exp4.m:4-21,
// exp3/compute_lwrq_exp4.m:12-56, exp3/separable_conv3_same_exp4.m:3-6).
//
// Layout: a plus-grid field is the column-major vectorisation of a (2n+3)^3 array (x
// fastest), the window interior is its (2n+1)^3 core; densities and rho are column-major
// (2n+1)^3 real arrays.  All work is HBM-bound stencil and reduction traffic:
//   pass 1   u (3 complex fields, 48 B/point) -> a, b (16 B/point) + sums of a, b
//   x, y, z  separable 7-tap convolution of a/den and b/den (den = conv(ones) is the
//            product of one 1-D profile per axis, evaluated in closed form), 16+16 B/point
//            for x and y; the z pass reduces E, M on the fly and writes nothing (rho only
//            when asked)
//   pass 3   u again (48 B/point) -> |Au - lambda Bu|^2 for lambda_3D and lambda_6D,
//            ||Au||^2, ||Bu||^2 (recomputed rather than stored: 48 B read < 96 B write+read)
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

constexpr int kYeeThreads = 256;
constexpr int kYeeMaxBlocks = 2048;
constexpr int kYeeMaxTaps = 17;   // LWRQ neighbour radius <= 8 cells
constexpr int kYeeMaxRows = 16;   // rows of the 3D-to-6D projection P

struct YeeGeom {
  long long q, m;  // interior and plus-grid extents
  double h2inv;
};

// Permittivity eps_c(x) = epsC exp(alpha/6 sum_l cos(P_l . x)) from per-axis tables:
// tab[((s*3 + d)*L + l)*q + i] = (cos, sin)(P(l,d) * x_d(i)), x_d(i) = (i - n) h + s h/2.
// cos(a + b + c) is expanded into the products of the three axis factors.
struct YeeMat {
  const double2* tab;
  int L;
  double epsC, a6;
};

__device__ __forceinline__ double yee_eps(const YeeMat& mt, long long q, int comp, long long i, long long j,
                                          long long k) {
  const long long axis_stride = (long long)mt.L * q;
  const double2* tx = mt.tab + ((comp == 0 ? 3 : 0) + 0) * axis_stride;
  const double2* ty = mt.tab + ((comp == 1 ? 3 : 0) + 1) * axis_stride;
  const double2* tz = mt.tab + ((comp == 2 ? 3 : 0) + 2) * axis_stride;
  double S = 0.0;
  for (int l = 0; l < mt.L; ++l) {
    const double2 X = tx[l * q + i], Y = ty[l * q + j], Z = tz[l * q + k];
    S += X.x * Y.x * Z.x - X.x * Y.y * Z.y - X.y * Y.x * Z.y - X.y * Y.y * Z.x;
  }
  return mt.epsC * exp(mt.a6 * S);
}

// Cropped curl-curl at interior point (i, j, k); c is its plus-grid offset.  Operator
// order and signs as exp3/apply_cropped_A_exp3.m:11-28 with the difference operators of
// :32-56 (each a combination of the point and its neighbours at offsets in {-1, 0, 1}).
__device__ __forceinline__ void yee_curlcurl(const cpx<double>* u1, const cpx<double>* u2, const cpx<double>* u3,
                                             long long c, long long sy, long long sz, double h2inv,
                                             cpx<double>& A1, cpx<double>& A2, cpx<double>& A3) {
  const long long sx = 1;
  // u1: Dyy, Dzz, DxT_DyP, DxT_DzP
  {
    const cpx<double> uc = u1[c];
    const cpx<double> dyy = (-u1[c - sy] + 2.0 * uc) - u1[c + sy];
    const cpx<double> dzz = (-u1[c - sz] + 2.0 * uc) - u1[c + sz];
    const cpx<double> u2c = u2[c];
    const cpx<double> dytdxp = ((u2[c + sx - sy] - u2[c - sy]) - u2[c + sx]) + u2c;
    const cpx<double> u3c = u3[c];
    const cpx<double> dztdxp = ((u3[c + sx - sz] - u3[c - sz]) - u3[c + sx]) + u3c;
    A1 = h2inv * (((dyy + dzz) - dytdxp) - dztdxp);
  }
  {
    const cpx<double> uc = u2[c];
    const cpx<double> dxx = (-u2[c - sx] + 2.0 * uc) - u2[c + sx];
    const cpx<double> dzz = (-u2[c - sz] + 2.0 * uc) - u2[c + sz];
    const cpx<double> u1c = u1[c];
    const cpx<double> dxtdyp = ((-u1[c - sx] + u1c) + u1[c - sx + sy]) - u1[c + sy];
    const cpx<double> u3c = u3[c];
    const cpx<double> dztdyp = ((-u3[c - sz] + u3[c + sy - sz]) + u3c) - u3[c + sy];
    A2 = h2inv * (((dxx + dzz) - dxtdyp) - dztdyp);
  }
  {
    const cpx<double> uc = u3[c];
    const cpx<double> dxx = (-u3[c - sx] + 2.0 * uc) - u3[c + sx];
    const cpx<double> dyy = (-u3[c - sy] + 2.0 * uc) - u3[c + sy];
    const cpx<double> u1c = u1[c];
    const cpx<double> dxtdzp = ((-u1[c - sx] + u1c) + u1[c - sx + sz]) - u1[c + sz];
    const cpx<double> u2c = u2[c];
    const cpx<double> dytdzp = ((-u2[c - sy] + u2c) + u2[c - sy + sz]) - u2[c + sz];
    A3 = h2inv * (((dxx + dyy) - dxtdzp) - dytdzp);
  }
}

__device__ __forceinline__ double abs2(cpx<double> z) { return z.re * z.re + z.im * z.im; }

// NV block sums at once; the totals end in v[] of thread 0.  red: NV * kYeeThreads.
template <int NV>
__device__ __forceinline__ void yee_block_sums(double (&v)[NV], double* red) {
  for (int k = 0; k < NV; ++k) red[k * kYeeThreads + threadIdx.x] = v[k];
  __syncthreads();
  if (threadIdx.x < 64) {
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = threadIdx.x; i < kYeeThreads; i += 64) s += red[k * kYeeThreads + i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      v[k] = s;
    }
  }
}

enum YeeMode : int { YEE_APPLY_A = 0, YEE_DENSITIES = 1, YEE_RESIDUALS = 2 };

// One pass over the window interior.
//   YEE_APPLY_A:   A1..A3 = cropped curl-curl of u
//   YEE_DENSITIES: a = Re(u' A u), b = Re(u' B u) per point; partials [sum a, sum b]
//   YEE_RESIDUALS: partials [|Au - l3 Bu|^2, |Au - l6 Bu|^2, |Au|^2, |Bu|^2]
template <int MODE>
__global__ __launch_bounds__(kYeeThreads) void k_yee_points(YeeGeom g, const cpx<double>* u1,
                                                            const cpx<double>* u2, const cpx<double>* u3,
                                                            YeeMat mt, cpx<double>* A1o, cpx<double>* A2o,
                                                            cpx<double>* A3o, double* a_out, double* b_out,
                                                            double lam3, double lam6, double* partials) {
  constexpr int NV = MODE == YEE_RESIDUALS ? 4 : 2;
  __shared__ double red[NV * kYeeThreads];
  const long long q = g.q, m = g.m, sy = m, sz = m * m, Q3 = q * q * q;
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (long long t = blockIdx.x * (long long)kYeeThreads + threadIdx.x; t < Q3;
       t += (long long)gridDim.x * kYeeThreads) {
    const long long i = t % q, r = t / q, j = r % q, k = r / q;
    const long long c = (i + 1) + m * ((j + 1) + m * (k + 1));
    cpx<double> A1, A2, A3;
    yee_curlcurl(u1, u2, u3, c, sy, sz, g.h2inv, A1, A2, A3);
    if constexpr (MODE == YEE_APPLY_A) {
      A1o[t] = A1;
      A2o[t] = A2;
      A3o[t] = A3;
    } else {
      const cpx<double> v1 = u1[c], v2 = u2[c], v3 = u3[c];
      const double e1 = yee_eps(mt, q, 0, i, j, k), e2 = yee_eps(mt, q, 1, i, j, k), e3 = yee_eps(mt, q, 2, i, j, k);
      const cpx<double> B1 = e1 * v1, B2 = e2 * v2, B3 = e3 * v3;
      if constexpr (MODE == YEE_DENSITIES) {
        const double a = (v1.re * A1.re + v1.im * A1.im) + (v2.re * A2.re + v2.im * A2.im) +
                         (v3.re * A3.re + v3.im * A3.im);
        const double b = (v1.re * B1.re + v1.im * B1.im) + (v2.re * B2.re + v2.im * B2.im) +
                         (v3.re * B3.re + v3.im * B3.im);
        a_out[t] = a;
        b_out[t] = b;
        acc[0] += a;
        acc[1] += b;
      } else {
        acc[0] += abs2(A1 - lam3 * B1) + abs2(A2 - lam3 * B2) + abs2(A3 - lam3 * B3);
        acc[1] += abs2(A1 - lam6 * B1) + abs2(A2 - lam6 * B2) + abs2(A3 - lam6 * B3);
        acc[2] += abs2(A1) + abs2(A2) + abs2(A3);
        acc[3] += abs2(B1) + abs2(B2) + abs2(B3);
      }
    }
  }
  if constexpr (MODE != YEE_APPLY_A) {
    yee_block_sums<NV>(acc, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = acc[k];
  }
}

// The three permittivity families on the interior (exp3/sample_scalar_yee_material_exp4.m:7-9).
__global__ __launch_bounds__(kYeeThreads) void k_yee_eps(long long q, YeeMat mt, double* e1, double* e2,
                                                         double* e3) {
  const long long Q3 = q * q * q;
  for (long long t = blockIdx.x * (long long)kYeeThreads + threadIdx.x; t < Q3;
       t += (long long)gridDim.x * kYeeThreads) {
    const long long i = t % q, r = t / q, j = r % q, k = r / q;
    e1[t] = yee_eps(mt, q, 0, i, j, k);
    e2[t] = yee_eps(mt, q, 1, i, j, k);
    e3[t] = yee_eps(mt, q, 2, i, j, k);
  }
}

struct Taps {
  double k[kYeeMaxTaps];
  int R;
};

// Zero-padded 'same' convolution along one axis (stride s, extent q) of two arrays.
// With `prof`: each source value is first divided by den = prof(i) prof(j) prof(k) (the
// column normalisation of exp3/compute_lwrq_exp4.m:22-26) and partials [sum in_a, sum in_b]
// of the raw inputs are written when `partials` is set.
__global__ __launch_bounds__(kYeeThreads) void k_lwrq_conv(long long q, long long s, Taps tp, const double* ina,
                                                           const double* inb, double* outa, double* outb,
                                                           const double* prof, double* partials) {
  __shared__ double red[2 * kYeeThreads];
  const long long Q3 = q * q * q;
  double acc[2] = {0.0, 0.0};
  for (long long t = blockIdx.x * (long long)kYeeThreads + threadIdx.x; t < Q3;
       t += (long long)gridDim.x * kYeeThreads) {
    const long long i = t % q, r = t / q, j = r % q, k = r / q;
    const long long pos = (s == 1) ? i : (s == q ? j : k);
    double sa = 0.0, sb = 0.0;
    for (int tau = -tp.R; tau <= tp.R; ++tau) {
      const long long src = pos - tau;
      if (src < 0 || src >= q) continue;
      const long long ts = t - tau * s;
      double va = ina[ts], vb = inb[ts];
      if (prof) {
        const double den = (s == 1 ? prof[src] * prof[j] * prof[k]
                                   : (s == q ? prof[i] * prof[src] * prof[k] : prof[i] * prof[j] * prof[src]));
        va = va / den;
        vb = vb / den;
      }
      sa += tp.k[tau + tp.R] * va;
      sb += tp.k[tau + tp.R] * vb;
    }
    outa[t] = sa;
    outb[t] = sb;
    if (partials) {
      acc[0] += ina[t];
      acc[1] += inb[t];
    }
  }
  if (partials) {
    yee_block_sums<2>(acc, red);
    if (threadIdx.x == 0) {
      partials[blockIdx.x] = acc[0];
      partials[gridDim.x + blockIdx.x] = acc[1];
    }
  }
}

// Last (z) convolution with the LWRQ reductions: E, M at every point, rho = E/M (written
// when asked) and partials [sum E, sum M, sum (E - lam_a M)^2 / M, sum (E - lam_a M),
// max M, min M, non-finite count], lam_a = sc[2].
__global__ __launch_bounds__(kYeeThreads) void k_lwrq_conv_z(long long q, Taps tp, const double* ina,
                                                             const double* inb, const double* sc, double* rho,
                                                             double* partials) {
  __shared__ double red[5 * kYeeThreads];
  __shared__ double redm[2 * kYeeThreads];
  const long long s = q * q, Q3 = q * q * q;
  const double lam = sc[2];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // E, M, dev2, dev1, bad
  double mx = -INFINITY, mn = INFINITY;
  for (long long t = blockIdx.x * (long long)kYeeThreads + threadIdx.x; t < Q3;
       t += (long long)gridDim.x * kYeeThreads) {
    const long long k = t / s;
    double E = 0.0, M = 0.0;
    for (int tau = -tp.R; tau <= tp.R; ++tau) {
      const long long src = k - tau;
      if (src < 0 || src >= q) continue;
      const long long ts = t - tau * s;
      E += tp.k[tau + tp.R] * ina[ts];
      M += tp.k[tau + tp.R] * inb[ts];
    }
    const double d = E - lam * M;
    acc[0] += E;
    acc[1] += M;
    acc[2] += d * d / M;
    acc[3] += d;
    if (!isfinite(M) || !isfinite(E)) acc[4] += 1.0;
    mx = fmax(mx, M);
    mn = fmin(mn, M);
    if (rho) rho[t] = E / M;
  }
  yee_block_sums<5>(acc, red);
  redm[threadIdx.x] = mx;
  redm[kYeeThreads + threadIdx.x] = mn;
  __syncthreads();
  if (threadIdx.x < 64) {
    for (int i = threadIdx.x + 64; i < kYeeThreads; i += 64) {
      mx = fmax(mx, redm[i]);
      mn = fmin(mn, redm[kYeeThreads + i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmax(mx, __shfl_down(mx, o, 64));
      mn = fmin(mn, __shfl_down(mn, o, 64));
    }
  }
  if (threadIdx.x == 0) {
    for (int v = 0; v < 5; ++v) partials[v * gridDim.x + blockIdx.x] = acc[v];
    partials[5 * gridDim.x + blockIdx.x] = mx;
    partials[6 * gridDim.x + blockIdx.x] = mn;
  }
}

struct Kinds {
  int k[8];
};

// out[v] = reduction of partials[v * nb ...]: kind 0 sum, 1 max, 2 min (one workgroup per value).
__global__ __launch_bounds__(kYeeThreads) void k_yee_finish(const double* partials, int nb, Kinds kinds,
                                                            double* out) {
  __shared__ double red[kYeeThreads];
  const int v = blockIdx.x, kind = kinds.k[v];
  double s = kind == 0 ? 0.0 : (kind == 1 ? -INFINITY : INFINITY);
  for (int i = threadIdx.x; i < nb; i += kYeeThreads) {
    const double x = partials[(long long)v * nb + i];
    s = kind == 0 ? s + x : (kind == 1 ? fmax(s, x) : fmin(s, x));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int i = 1; i < kYeeThreads; ++i) r = kind == 0 ? r + red[i] : (kind == 1 ? fmax(r, red[i]) : fmin(r, red[i]));
    out[v] = r;
  }
}

// sc[2] = sc[0] / sc[1] (lambda of the raw densities, the centre of the deviation sums)
__global__ void k_yee_ratio(double* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc[2] = sc[0] / sc[1];
}

// ----------------------------------------------------------------------------- host side
static int yee_blocks(long long Q3) {
  long long b = (Q3 + kYeeThreads - 1) / kYeeThreads;
  return (int)std::max<long long>(1, std::min<long long>(b, kYeeMaxBlocks));
}

// Workspace layout (every part 256-byte aligned): 4 real (2n+1)^3 arrays | partials
// (8 x kYeeMaxBlocks) | scalars (64) | convolution profile (2n+1) | material tables
// 6 x kYeeMaxRows x (2n+1) (cos, sin).
struct YeeWs {
  double *w0, *w1, *w2, *w3, *part, *sc, *prof;
  double2* tab;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t yee_ws_bytes(long long q) {
  const size_t Q3 = (size_t)q * q * q;
  return 4 * align256(sizeof(double) * Q3) + align256(sizeof(double) * 8 * kYeeMaxBlocks) +
         align256(sizeof(double) * 64) + align256(sizeof(double) * q) +
         align256(sizeof(double2) * 6 * kYeeMaxRows * q);
}

static YeeWs yee_ws_map(void* base, long long q) {
  const size_t Q3 = (size_t)q * q * q;
  char* p = static_cast<char*>(base);
  YeeWs w{};
  double** arr[4] = {&w.w0, &w.w1, &w.w2, &w.w3};
  for (int i = 0; i < 4; ++i) {
    *arr[i] = reinterpret_cast<double*>(p);
    p += align256(sizeof(double) * Q3);
  }
  w.part = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 8 * kYeeMaxBlocks);
  w.sc = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 64);
  w.prof = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * q);
  w.tab = reinterpret_cast<double2*>(p);
  return w;
}

// Gaussian taps k = exp(-(off/sigma)^2), off = (-R:R) h, sigma = sigma_factor h
// (exp3/compute_lwrq_exp4.m:12-15), and the 'same' convolution of ones along one axis,
// prof(i) = sum of the taps that land inside 0..q-1; den = prof(i) prof(j) prof(k).
static Taps lwrq_taps(double h, int R, double sigma_factor, long long q, std::vector<double>& prof) {
  if (R < 0 || 2 * R + 1 > kYeeMaxTaps) throw Error(FAME6D_EARG, "LWRQ neighbour radius must be in 0..8");
  if (!(h > 0) || !(sigma_factor > 0)) throw Error(FAME6D_EARG, "h and sigma factor must be positive");
  Taps tp{};
  tp.R = R;
  const double sigma = sigma_factor * h;
  for (int t = -R; t <= R; ++t) {
    const double off = t * h;
    tp.k[t + R] = std::exp(-(off * off) / (sigma * sigma));
  }
  prof.assign((size_t)q, 0.0);
  for (long long i = 0; i < q; ++i) {
    double s = 0.0;
    for (int t = -R; t <= R; ++t)
      if (i - t >= 0 && i - t < q) s += tp.k[t + R];
    prof[(size_t)i] = s;
    if (!(std::isfinite(s) && s > 0)) throw Error(FAME6D_ENOCONV, "Invalid LWRQ column-normalization denominator.");
  }
  return tp;
}

// Per-axis (cos, sin) tables of the permittivity (see YeeMat).  `tab` must stay alive
// until the stream has been synchronised.
static YeeMat upload_material(const YeeWs& w, int n, double h, long long q, const double* P, int L,
                              double epsC, double alpha, hipStream_t s, std::vector<double2>& tab) {
  if (L < 1 || L > kYeeMaxRows) throw Error(FAME6D_EARG, "the projection must have 1..16 rows");
  tab.assign((size_t)6 * L * q, double2{0.0, 0.0});
  for (int st = 0; st < 2; ++st)
    for (int d = 0; d < 3; ++d)
      for (int l = 0; l < L; ++l)
        for (long long i = 0; i < q; ++i) {
          const double x = (double)(i - n) * h + (st ? h / 2 : 0.0);
          const double ph = P[l * 3 + d] * x;  // P row-major, L x 3
          tab[(((size_t)st * 3 + d) * L + l) * q + i] = double2{std::cos(ph), std::sin(ph)};
        }
  F6_HIP(hipMemcpyAsync(w.tab, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice, s));
  return YeeMat{w.tab, L, epsC, alpha / 6.0};
}

struct LwrqSums {
  double sum_a, sum_b, sum_E, sum_M, dev2, dev1, bad, max_M, min_M;
};

// x and y passes of the separable convolution (the x pass divides by den) and the z pass
// with the LWRQ reductions.  a, b are read; w2/w3 and w0/w1 are ping-ponged.  With
// sums_from_x the x pass also sums a and b into sc[0], sc[1]; otherwise the caller has
// put them there.  One synchronisation.
static LwrqSums lwrq_device(const YeeWs& w, long long q, const Taps& tp, const double* a, const double* b,
                            double* rho, hipStream_t s, bool sums_from_x) {
  const long long Q3 = q * q * q;
  const int nb = yee_blocks(Q3);
  k_lwrq_conv<<<nb, kYeeThreads, 0, s>>>(q, 1, tp, a, b, w.w2, w.w3, w.prof, sums_from_x ? w.part : nullptr);
  F6_CHECK_LAUNCH();
  if (sums_from_x) k_yee_finish<<<2, kYeeThreads, 0, s>>>(w.part, nb, Kinds{{0, 0}}, w.sc);
  k_yee_ratio<<<1, 64, 0, s>>>(w.sc);
  k_lwrq_conv<<<nb, kYeeThreads, 0, s>>>(q, q, tp, w.w2, w.w3, w.w0, w.w1, nullptr, nullptr);
  k_lwrq_conv_z<<<nb, kYeeThreads, 0, s>>>(q, tp, w.w0, w.w1, w.sc, rho, w.part);
  k_yee_finish<<<7, kYeeThreads, 0, s>>>(w.part, nb, Kinds{{0, 0, 0, 0, 0, 1, 2}}, w.sc + 8);
  F6_CHECK_LAUNCH();
  double h[16];
  F6_HIP(hipMemcpyAsync(h, w.sc, sizeof(double) * 16, hipMemcpyDeviceToHost, s));
  F6_HIP(hipStreamSynchronize(s));
  return LwrqSums{h[0], h[1], h[8], h[9], h[10], h[11], h[12], h[13], h[14]};
}

// compute_lwrq_exp4 scalars from the reduced sums (exp3/compute_lwrq_exp4.m:28-48).  The
// deviation sums were taken about lam_a = sum a / sum b; sum M (rho - lam)^2 about the
// reference's lam = sum E / sum M follows exactly from them:
//   sum (E - lam M)^2 / M = dev2 + 2 (lam_a - lam) dev1 + (lam_a - lam)^2 sum M.
static fame6d_lwrq_t lwrq_finish(const LwrqSums& r, double mass_floor_rel) {
  const double eps = 2.220446049250313e-16;
  if (r.bad != 0.0 || !std::isfinite(r.max_M) || !(r.min_M > mass_floor_rel * r.max_M))
    throw Error(FAME6D_ENOCONV, "Nonpositive or numerically tiny local LWRQ mass encountered.");
  fame6d_lwrq_t o{};
  const double lam = r.sum_E / r.sum_M;
  const double lam_a = r.sum_a / r.sum_b;
  const double dl = lam_a - lam;
  const double vnum = r.dev2 + 2.0 * dl * r.dev1 + dl * dl * r.sum_M;
  o.lambda = lam;
  o.V = vnum / (r.sum_M * lam * lam);
  o.nu = std::sqrt(std::max(o.V, 0.0));
  o.energy_partition_error = std::fabs(r.sum_E - r.sum_a) / std::max(std::fabs(r.sum_a), eps);
  o.mass_partition_error = std::fabs(r.sum_M - r.sum_b) / std::max(std::fabs(r.sum_b), eps);
  o.partition_error = std::max(o.energy_partition_error, o.mass_partition_error);
  o.sum_a = r.sum_a;
  o.sum_b = r.sum_b;
  return o;
}

// Scratch for calls that bring no workspace (released after the call).
struct YeeScratch {
  DevBuf buf;
  void* get(void* work, size_t bytes) {
    if (work) return work;
    buf.ensure(bytes);
    return buf.ptr;
  }
  ~YeeScratch() { buf.release(); }
};

static void check_window(int n, double h) {
  if (n < 1 || n > 1000) throw Error(FAME6D_EARG, "Yee window n must be in 1..1000");
  if (!(h > 0) || !std::isfinite(h)) throw Error(FAME6D_EARG, "Yee step h must be positive");
}

}  // namespace fame6d

using namespace fame6d;

extern "C" {

size_t fame6d_yee_workspace_bytes(int n) { return n < 1 ? 0 : yee_ws_bytes(2LL * n + 1); }

int fame6d_yee_apply_cropped_A(int n, double h, const void* u1, const void* u2, const void* u3, void* A1,
                               void* A2, void* A3, void* stream) {
  try {
    check_window(n, h);
    if (!u1 || !u2 || !u3 || !A1 || !A2 || !A3) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const long long q = 2LL * n + 1;
    const YeeGeom g{q, q + 2, (1.0 / h) * (1.0 / h)};
    k_yee_points<YEE_APPLY_A><<<yee_blocks(q * q * q), kYeeThreads, 0, s>>>(
        g, (const cpx<double>*)u1, (const cpx<double>*)u2, (const cpx<double>*)u3, YeeMat{}, (cpx<double>*)A1,
        (cpx<double>*)A2, (cpx<double>*)A3, nullptr, nullptr, 0.0, 0.0, nullptr);
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(s));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_yee_sample_eps(int n, double h, const double* P_host, int nrows, double eps_c, double alpha, void* e1,
                          void* e2, void* e3, void* work, void* stream) {
  try {
    check_window(n, h);
    if (!P_host || !e1 || !e2 || !e3) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const long long q = 2LL * n + 1;
    YeeScratch scr;
    const YeeWs w = yee_ws_map(scr.get(work, yee_ws_bytes(q)), q);
    std::vector<double2> tab;
    const YeeMat mt = upload_material(w, n, h, q, P_host, nrows, eps_c, alpha, s, tab);
    k_yee_eps<<<yee_blocks(q * q * q), kYeeThreads, 0, s>>>(q, mt, (double*)e1, (double*)e2, (double*)e3);
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(s));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_lwrq(int n, double h, const double* a, const double* b, int radius, double sigma_factor,
                double mass_floor_rel, double* rho, void* work, fame6d_lwrq_t* out, void* stream) {
  try {
    check_window(n, h);
    if (!a || !b || !out) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const long long q = 2LL * n + 1;
    std::vector<double> prof;
    const Taps tp = lwrq_taps(h, radius, sigma_factor, q, prof);
    YeeScratch scr;
    const YeeWs w = yee_ws_map(scr.get(work, yee_ws_bytes(q)), q);
    F6_HIP(hipMemcpyAsync(w.prof, prof.data(), sizeof(double) * q, hipMemcpyHostToDevice, s));
    const LwrqSums r = lwrq_device(w, q, tp, a, b, rho, s, true);
    *out = lwrq_finish(r, mass_floor_rel);
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_yee_lwrq_metrics(int n, double h, const void* u1, const void* u2, const void* u3, const double* P_host,
                            int nrows, double eps_c, double alpha, double lambda6d, int radius,
                            double sigma_factor, double mass_floor_rel, double* rho, void* work,
                            fame6d_yee_metrics_t* out, void* stream) {
  try {
    check_window(n, h);
    if (!u1 || !u2 || !u3 || !P_host || !out) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const long long q = 2LL * n + 1, Q3 = q * q * q;
    std::vector<double> prof;
    const Taps tp = lwrq_taps(h, radius, sigma_factor, q, prof);
    YeeScratch scr;
    const YeeWs w = yee_ws_map(scr.get(work, yee_ws_bytes(q)), q);
    F6_HIP(hipMemcpyAsync(w.prof, prof.data(), sizeof(double) * q, hipMemcpyHostToDevice, s));
    std::vector<double2> tab;
    const YeeMat mt = upload_material(w, n, h, q, P_host, nrows, eps_c, alpha, s, tab);
    const YeeGeom g{q, q + 2, (1.0 / h) * (1.0 / h)};
    const auto* f1 = (const cpx<double>*)u1;
    const auto* f2 = (const cpx<double>*)u2;
    const auto* f3 = (const cpx<double>*)u3;
    const int nb = yee_blocks(Q3);
    // pass 1: a, b and their sums
    k_yee_points<YEE_DENSITIES><<<nb, kYeeThreads, 0, s>>>(g, f1, f2, f3, mt, nullptr, nullptr, nullptr, w.w0,
                                                           w.w1, 0.0, 0.0, w.part);
    F6_CHECK_LAUNCH();
    k_yee_finish<<<2, kYeeThreads, 0, s>>>(w.part, nb, Kinds{{0, 0}}, w.sc);
    F6_CHECK_LAUNCH();
    // LWRQ on (a, b); w0/w1 are free once the x pass has read them
    const LwrqSums r = lwrq_device(w, q, tp, w.w0, w.w1, rho, s, false);
    if (!(r.sum_b > 0)) throw Error(FAME6D_ENOCONV, "Cropped Yee mass is not positive.");
    const fame6d_lwrq_t L = lwrq_finish(r, mass_floor_rel);
    const double eps = 2.220446049250313e-16;
    fame6d_yee_metrics_t o{};
    o.lambda_crop = r.sum_a / r.sum_b;
    o.lambda_3d = L.lambda;
    o.delta_identity = std::fabs(o.lambda_3d - o.lambda_crop) / std::max(std::fabs(o.lambda_crop), eps);
    // pass 3: residual norms at lambda_3D and lambda_6D
    k_yee_points<YEE_RESIDUALS><<<nb, kYeeThreads, 0, s>>>(g, f1, f2, f3, mt, nullptr, nullptr, nullptr, nullptr,
                                                           nullptr, o.lambda_3d, lambda6d, w.part);
    F6_CHECK_LAUNCH();
    k_yee_finish<<<4, kYeeThreads, 0, s>>>(w.part, nb, Kinds{{0, 0, 0, 0}}, w.sc + 16);
    F6_CHECK_LAUNCH();
    double hr[4];
    F6_HIP(hipMemcpyAsync(hr, w.sc + 16, sizeof(hr), hipMemcpyDeviceToHost, s));
    F6_HIP(hipStreamSynchronize(s));
    const double nA = std::sqrt(hr[2]), nB = std::sqrt(hr[3]);
    o.eta_3d = std::sqrt(hr[0]) / (nA + std::fabs(o.lambda_3d) * nB);
    o.eta_6d = std::sqrt(hr[1]) / (nA + std::fabs(lambda6d) * nB);
    o.delta_lambda = std::fabs(o.lambda_3d - lambda6d) / std::fabs(lambda6d);
    o.V = L.V;
    o.nu = L.nu;
    o.delta_part = L.partition_error;
    o.energy_partition_error = L.energy_partition_error;
    o.mass_partition_error = L.mass_partition_error;
    o.cropped_mass = r.sum_b;
    *out = o;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

}  // extern "C"
