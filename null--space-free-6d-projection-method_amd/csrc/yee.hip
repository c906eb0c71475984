// Physical-space closure of a reconstructed eigenfield on a Yee window: the cropped Yee
// curl-curl action, the scalar permittivity on the three edge families, the pointwise
// energy and mass densities, the column-normalised local Gaussian Rayleigh quotient (LWRQ)
// and the residual norms (exp3/compute_yee_lwrq_metrics_exp4.m:16-94,
// exp3/apply_cropped_A_exp3.m:8-56, exp3/sample_scalar_yee_material_exp4.m:4-21,
// exp3/compute_lwrq_exp4.m:12-56, exp3/separable_conv3_same_exp4.m:3-6).
//
// Layout: a plus-grid field is the column-major vectorisation of a (2n+3)^3 array (x
// fastest), the window interior is its (2n+1)^3 core; densities and rho are column-major
// (2n+1)^3 real arrays.  The work is HBM-bound stencil and reduction traffic (per interior
// point):
//   pass 1   u (3 complex fields, 48 B) -> a/den, b/den (16 B); sums of a, b and of the
//            residual terms about the 6D eigenvalue (below)
//   LWRQ     the convolution of the two densities along x, y and z with the LWRQ
//            reductions in one z-march, nothing written (16 B; rho when asked).  The
//            standalone LWRQ entry point (which also normalises its raw inputs) runs it as
//            three passes: x (16 B in, 16 B out), y marched through registers (16 + 16 B),
//            z with the reductions (16 B)
// den = conv(ones) is the product of one 1-D profile per axis and is applied in closed form.
// The residuals ||Au - lam Bu||^2 for lam_3D (known only after the LWRQ) follow from sums
// taken in pass 1 about lam_0 = lam_6D:
//   ||Au - lam Bu||^2 = S0 - 2 d S1 + d^2 S2,  d = lam - lam_0,
//   S0 = ||Au - lam_0 Bu||^2, S1 = Re <Au - lam_0 Bu, Bu>, S2 = ||Bu||^2,
// which is exact algebra; when the three terms cancel to less than 1e-6 of their size the
// residual is recomputed by a second pass over u (48 B) instead.
//
// Launch shape: workgroups of 64 x 4 threads own (x, y) tiles of 64 x 4 points (one wave per
// row segment, coalesced) and loop over a chunk of the third index, so no grid index is ever
// divided.  The point passes keep the plus-grid planes k-1, k, k+1 of their tile (with a
// one-point halo) in LDS; the convolutions march a register window.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

constexpr int kYeeTX = 64, kYeeTY = 4, kYeeThreads = kYeeTX * kYeeTY;
constexpr int kYeeMaxBlocks = 2048;
constexpr int kYeeMaxR = 8;        // LWRQ neighbour radius <= 8 cells
constexpr int kYeeMaxRows = 16;    // rows of the 3D-to-6D projection P
constexpr int kYeeMinChunk = 16;   // planes per workgroup of the point passes (2 more are loaded)

struct YeeGeom {
  long long q, m;  // interior and plus-grid extents
  double h2inv;
};

// Permittivity eps_c(x) = epsC exp(alpha/6 sum_l cos(P_l . x)), x_d = (i_d - n) h (+ h/2 on
// axis c).  When every row of P has at most one non-zero the sum splits over the axes and so
// does the exponential: eps_c = epsC E[s_0][0][i] E[s_1][1][j] E[s_2][2][k] with
// E[s][d][i] = exp(alpha/6 sum_{rows l on axis d} cos(P(l,d) x_d(i))), s_d = (d == c).
// Otherwise tab[((s*3 + d)*L + l)*q + i] = (cos, sin)(P(l,d) x_d(i)) and cos(a + b + c) is
// expanded into products of the three axis factors.
struct YeeMat {
  const double* sep;    // 6 x q separable factors (null: general form)
  const double2* tab;   // 6 x L x q (cos, sin)
  int L;
  double epsC, a6;
};

__device__ __forceinline__ double yee_eps(const YeeMat& mt, long long q, int comp, long long i, long long j,
                                          long long k) {
  if (mt.sep) {
    const double ex = mt.sep[((comp == 0 ? 3 : 0) + 0) * q + i];
    const double ey = mt.sep[((comp == 1 ? 3 : 0) + 1) * q + j];
    const double ez = mt.sep[((comp == 2 ? 3 : 0) + 2) * q + k];
    return mt.epsC * (ex * ey * ez);
  }
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

// Cropped curl-curl at one interior point from an accessor U(comp, dx, dy, dz) of the
// plus-grid fields at offsets in {-1, 0, 1}.  Operator order and signs as
// exp3/apply_cropped_A_exp3.m:11-28 with the difference operators of :32-56.
template <class Get>
__device__ __forceinline__ void yee_curlcurl(const Get& U, double h2inv, cpx<double>& A1, cpx<double>& A2,
                                             cpx<double>& A3) {
  const cpx<double> u1c = U(0, 0, 0, 0), u2c = U(1, 0, 0, 0), u3c = U(2, 0, 0, 0);
  const cpx<double> u1xm = U(0, -1, 0, 0), u2xm = U(1, -1, 0, 0), u3xm = U(2, -1, 0, 0);
  const cpx<double> u2xp = U(1, 1, 0, 0), u3xp = U(2, 1, 0, 0);
  const cpx<double> u1ym = U(0, 0, -1, 0), u1yp = U(0, 0, 1, 0), u2ym = U(1, 0, -1, 0), u3ym = U(2, 0, -1, 0),
                    u3yp = U(2, 0, 1, 0);
  const cpx<double> u1zm = U(0, 0, 0, -1), u1zp = U(0, 0, 0, 1), u2zm = U(1, 0, 0, -1), u2zp = U(1, 0, 0, 1),
                    u3zm = U(2, 0, 0, -1);
  {
    const cpx<double> dyy = (-u1ym + 2.0 * u1c) - u1yp;
    const cpx<double> dzz = (-u1zm + 2.0 * u1c) - u1zp;
    const cpx<double> dytdxp = ((U(1, 1, -1, 0) - u2ym) - u2xp) + u2c;
    const cpx<double> dztdxp = ((U(2, 1, 0, -1) - u3zm) - u3xp) + u3c;
    A1 = h2inv * (((dyy + dzz) - dytdxp) - dztdxp);
  }
  {
    const cpx<double> dxx = (-u2xm + 2.0 * u2c) - u2xp;
    const cpx<double> dzz = (-u2zm + 2.0 * u2c) - u2zp;
    const cpx<double> dxtdyp = ((-u1xm + u1c) + U(0, -1, 1, 0)) - u1yp;
    const cpx<double> dztdyp = ((-u3zm + U(2, 0, 1, -1)) + u3c) - u3yp;
    A2 = h2inv * (((dxx + dzz) - dxtdyp) - dztdyp);
  }
  {
    const cpx<double> dxx = (-u3xm + 2.0 * u3c) - u3xp;
    const cpx<double> dyy = (-u3ym + 2.0 * u3c) - u3yp;
    const cpx<double> dxtdzp = ((-u1xm + u1c) + U(0, -1, 0, 1)) - u1zp;
    const cpx<double> dytdzp = ((-u2ym + u2c) + U(1, 0, -1, 1)) - u2zp;
    A3 = h2inv * (((dxx + dyy) - dxtdzp) - dytdzp);
  }
}

__device__ __forceinline__ double abs2(cpx<double> z) { return z.re * z.re + z.im * z.im; }
__device__ __forceinline__ double re_dot(cpx<double> a, cpx<double> b) { return a.re * b.re + a.im * b.im; }
__device__ __forceinline__ int yee_tid() { return threadIdx.x + kYeeTX * threadIdx.y; }
__device__ __forceinline__ int yee_block() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ __forceinline__ int yee_nblocks() { return gridDim.x * gridDim.y * gridDim.z; }

// NV block sums at once; the totals end in v[] of thread 0.  red: NV * kYeeThreads.
template <int NV>
__device__ __forceinline__ void yee_block_sums(double (&v)[NV], double* red) {
  const int tid = yee_tid();
  for (int k = 0; k < NV; ++k) red[k * kYeeThreads + tid] = v[k];
  __syncthreads();
  if (tid < 64) {
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = tid; i < kYeeThreads; i += 64) s += red[k * kYeeThreads + i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      v[k] = s;
    }
  }
}

enum YeeMode : int { YEE_APPLY_A = 0, YEE_DENSITIES = 1, YEE_RESIDUALS = 2 };

// Plus-grid planes held in LDS by the point passes: the 64 x 4 interior tile of a workgroup
// with a one-point halo in x and y, three components; a plane is kYPE values, loaded as
// kYPR rounds of the 256 threads.
constexpr int kYPX = kYeeTX + 2, kYPY = kYeeTY + 2, kYPC = kYPX * kYPY, kYPE = 3 * kYPC;
constexpr int kYPR = (kYPE + kYeeThreads - 1) / kYeeThreads;

// One pass over the window interior, a 2.5D stencil: the workgroup owns a 64 x 4 tile of
// (i, j) and walks k through chunk blockIdx.z, holding the plus-grid planes below, at and
// above the current one (halo included) in an LDS ring, so every field value is read from
// memory about once.  The plane two steps ahead is in flight in registers while a plane is
// computed (two register sets, the loop unrolled by two), so the load latency overlaps the
// compute of two planes.
//   YEE_APPLY_A:   A1..A3 = cropped curl-curl of u
//   YEE_DENSITIES: a/den, b/den per point (a = Re(u' A u), b = Re(u' B u)); partials
//                  [sum a, sum b, |Au - l0 Bu|^2, Re<Au - l0 Bu, Bu>, |Bu|^2, |Au|^2]
//   YEE_RESIDUALS: partials [|Au - l0 Bu|^2]
template <int MODE>
__global__ __launch_bounds__(kYeeThreads, 2) void k_yee_points(YeeGeom g, const cpx<double>* __restrict__ u1,
                                                               const cpx<double>* __restrict__ u2,
                                                               const cpx<double>* __restrict__ u3, YeeMat mt,
                                                               cpx<double>* A1o, cpx<double>* A2o,
                                                               cpx<double>* A3o, const double* prof,
                                                               double* a_out, double* b_out, double lam0,
                                                               long long chunk, double* partials) {
  constexpr int NV = MODE == YEE_DENSITIES ? 6 : 1;
  __shared__ cpx<double> pl[3][kYPE];
  __shared__ double red[NV * kYeeThreads];
  const long long q = g.q, m = g.m, mm = g.m * g.m;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = yee_tid();
  const long long i0 = blockIdx.x * (long long)kYeeTX, j0 = blockIdx.y * (long long)kYeeTY;
  const long long i = i0 + tx, j = j0 + ty;
  const bool act = i < q && j < q;
  const long long k0 = blockIdx.z * chunk, k1 = min(q, k0 + chunk), kend = k1 + 1;  // last plus plane
  // This thread's share of a plane: element e = tid + 256 r (component, halo row, halo column).
  // Halo cells beyond the plus grid are read clamped to it: active points never use them
  // (i, j < q touch plus indices <= q + 1 = m - 1).
  const cpx<double>* src[kYPR];
#pragma unroll
  for (int r = 0; r < kYPR; ++r) {
    const int e = min(tid + kYeeThreads * r, kYPE - 1);
    const int comp = e / kYPC, rr = e - comp * kYPC, yy = rr / kYPX, xx = rr - yy * kYPX;
    const long long gx = min(i0 + xx, m - 1), gy = min(j0 + yy, m - 1);
    src[r] = (comp == 0 ? u1 : (comp == 1 ? u2 : u3)) + gx + m * gy;
  }
  // The in-flight planes are clang vectors: struct copies into a local array become memcpys
  // that keep the array in scratch memory instead of registers.
  typedef double d2 __attribute__((ext_vector_type(2)));
  auto gload = [&](d2(&v)[kYPR], long long z) {  // planes beyond kend: kend again, unused
    const long long off = min(z, kend) * mm;
#pragma unroll
    for (int r = 0; r < kYPR; ++r) v[r] = *reinterpret_cast<const d2*>(src[r] + off);
  };
  auto lstore = [&](int slot, const d2(&v)[kYPR], long long z) {
    if (z <= kend) {
#pragma unroll
      for (int r = 0; r < kYPR; ++r)
        if (tid + kYeeThreads * r < kYPE) *reinterpret_cast<d2*>(&pl[slot][tid + kYeeThreads * r]) = v[r];
    }
  };
  // Separable permittivity: the (i, j) factors of each component are fixed for the thread and
  // the k factor is uniform over the workgroup.
  double exy[3] = {0.0, 0.0, 0.0};
  if (mt.sep && act) {
#pragma unroll
    for (int c = 0; c < 3; ++c) exy[c] = mt.sep[((c == 0 ? 3 : 0) + 0) * q + i] * mt.sep[((c == 1 ? 3 : 0) + 1) * q + j];
  }
  const double pij = (prof && act) ? prof[i] * prof[j] : 1.0;
  double acc[NV];
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;

  // Interior plane k from the slots holding plus planes k, k + 1, k + 2.
  auto compute = [&](long long k, int s0, int s1, int s2) {
    if (!act) return;
    const int c = (ty + 1) * kYPX + (tx + 1);
    const cpx<double>* Pm = pl[s0];
    const cpx<double>* Pc = pl[s1];
    const cpx<double>* Pp = pl[s2];
    auto U = [&](int comp, int dx, int dy, int dz) -> cpx<double> {
      const cpx<double>* P = dz < 0 ? Pm : (dz == 0 ? Pc : Pp);
      return P[comp * kYPC + c + dx + kYPX * dy];
    };
    cpx<double> A1, A2, A3;
    yee_curlcurl(U, g.h2inv, A1, A2, A3);
    const long long t = i + q * (j + q * k);
    if constexpr (MODE == YEE_APPLY_A) {
      A1o[t] = A1;
      A2o[t] = A2;
      A3o[t] = A3;
    } else {
      const cpx<double> v1 = Pc[c], v2 = Pc[kYPC + c], v3 = Pc[2 * kYPC + c];
      double e1, e2, e3;
      if (mt.sep) {
        const double ez = mt.sep[2 * q + k], ez2 = mt.sep[5 * q + k];
        e1 = mt.epsC * (exy[0] * ez);
        e2 = mt.epsC * (exy[1] * ez);
        e3 = mt.epsC * (exy[2] * ez2);
      } else {
        e1 = yee_eps(mt, q, 0, i, j, k);
        e2 = yee_eps(mt, q, 1, i, j, k);
        e3 = yee_eps(mt, q, 2, i, j, k);
      }
      const cpx<double> B1 = e1 * v1, B2 = e2 * v2, B3 = e3 * v3;
      const cpx<double> R1 = A1 - lam0 * B1, R2 = A2 - lam0 * B2, R3 = A3 - lam0 * B3;
      if constexpr (MODE == YEE_DENSITIES) {
        const double a = (re_dot(v1, A1) + re_dot(v2, A2)) + re_dot(v3, A3);
        const double b = (re_dot(v1, B1) + re_dot(v2, B2)) + re_dot(v3, B3);
        const double den = pij * prof[k];
        a_out[t] = a / den;
        b_out[t] = b / den;
        acc[0] += a;
        acc[1] += b;
        acc[2] += abs2(R1) + abs2(R2) + abs2(R3);
        acc[3] += re_dot(R1, B1) + re_dot(R2, B2) + re_dot(R3, B3);
        acc[4] += abs2(B1) + abs2(B2) + abs2(B3);
        acc[5] += abs2(A1) + abs2(A2) + abs2(A3);
      } else {
        acc[0] += abs2(R1) + abs2(R2) + abs2(R3);
      }
    }
  };

  if (k0 < k1) {
    d2 va[kYPR], vb[kYPR];
    gload(va, k0);  // plus planes k0 .. k0 + 2 (interior k0 - 1 .. k0 + 1); k0 + 2 <= kend
    gload(vb, k0 + 1);
    lstore(0, va, k0);
    lstore(1, vb, k0 + 1);
    gload(va, k0 + 2);
    lstore(2, va, k0 + 2);
    gload(va, k0 + 3);  // in flight: plus planes k0 + 3 (va) and k0 + 4 (vb)
    gload(vb, k0 + 4);
    __syncthreads();
    int s0 = 0, s1 = 1, s2 = 2;
    for (long long k = k0; k < k1; k += 2) {
      // step k: slots (s0, s1, s2) hold plus planes k, k + 1, k + 2
      compute(k, s0, s1, s2);
      __syncthreads();
      lstore(s0, va, k + 3);
      gload(va, k + 5);
      __syncthreads();
      if (k + 1 < k1) {
        // step k + 1: slots (s1, s2, s0)
        compute(k + 1, s1, s2, s0);
        __syncthreads();
        lstore(s1, vb, k + 4);
        gload(vb, k + 6);
        __syncthreads();
      }
      const int sn = s0;  // two rotations of the ring
      s0 = s2;
      s2 = s1;
      s1 = sn;
    }
  }
  if constexpr (MODE != YEE_APPLY_A) {
    yee_block_sums<NV>(acc, red);
    if (tid == 0) {
      const int b = yee_block(), nb = yee_nblocks();
      for (int v = 0; v < NV; ++v) partials[v * nb + b] = acc[v];
    }
  }
}

// The three permittivity families on the interior (exp3/sample_scalar_yee_material_exp4.m:7-9).
__global__ __launch_bounds__(kYeeThreads) void k_yee_eps(long long q, YeeMat mt, double* e1, double* e2,
                                                         double* e3) {
  const long long i = blockIdx.x * (long long)kYeeTX + threadIdx.x, j = blockIdx.y * (long long)kYeeTY + threadIdx.y;
  if (i >= q || j >= q) return;
  for (long long k = blockIdx.z; k < q; k += gridDim.z) {
    const long long t = i + q * (j + q * k);
    e1[t] = yee_eps(mt, q, 0, i, j, k);
    e2[t] = yee_eps(mt, q, 1, i, j, k);
    e3[t] = yee_eps(mt, q, 2, i, j, k);
  }
}

struct Taps {
  double k[2 * kYeeMaxR + 1];
  int R;
};

// Zero-padded 'same' convolution of two arrays along x; the workgroup's 64 x 4 tile of (i, j)
// loops over k, staging the tile's rows with an R-point halo on both sides in LDS (one load
// per value; the zero padding is stored there), double-buffered with the next plane's values
// in flight in registers.  With `prof` every source value is first divided by
// den = prof(i') prof(j) prof(k) (the column normalisation of exp3/compute_lwrq_exp4.m:22-26)
// and partials [sum in_a, sum in_b] of the raw inputs are written.
template <int R>
__global__ __launch_bounds__(kYeeThreads) void k_lwrq_conv_x(long long q, Taps tp, const double* __restrict__ ina,
                                                             const double* __restrict__ inb, double* outa,
                                                             double* outb, const double* prof, double* partials) {
  constexpr int W = 2 * R + 1, RW = kYeeTX + 2 * R, NE = kYeeTY * RW;
  constexpr int NR = (NE + kYeeThreads - 1) / kYeeThreads;
  __shared__ double ra[2][NE], rb[2][NE];
  __shared__ double red[2 * kYeeThreads];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = yee_tid();
  const long long i0 = blockIdx.x * (long long)kYeeTX, j0 = blockIdx.y * (long long)kYeeTY;
  const long long i = i0 + tx, j = j0 + ty, qq = q * q, gz = gridDim.z;
  const bool act = i < q && j < q;
  // This thread's share of a staged plane: element e = tid + 256 r (row e / RW, column e % RW)
  // is source point (i0 - R + column, j0 + row); outside the window it is the zero padding.
  bool ok[NR];
  long long off[NR];
  double pi[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int e = tid + kYeeThreads * r, row = e / RW, col = e - row * RW;
    const long long gi = i0 - R + col, gj = j0 + row;
    ok[r] = e < NE && gi >= 0 && gi < q && gj < q;
    off[r] = ok[r] ? gi + q * gj : 0;
    pi[r] = (ok[r] && prof) ? prof[gi] * prof[gj] : 1.0;
  }
  auto gload = [&](double (&va)[NR], double (&vb)[NR], long long k) {
    if (k >= q) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double x = 0.0, y = 0.0;
      if (ok[r]) {
        x = ina[off[r] + qq * k];
        y = inb[off[r] + qq * k];
      }
      va[r] = x;
      vb[r] = y;
    }
  };
  auto lstore = [&](int slot, const double (&va)[NR], const double (&vb)[NR], long long k) {
    if (k >= q) return;
    const double pk = prof ? prof[k] : 1.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int e = tid + kYeeThreads * r;
      if (e < NE) {
        double x = va[r], y = vb[r];
        if (prof && ok[r]) {
          const double den = pi[r] * pk;
          x = x / den;
          y = y / den;
        }
        ra[slot][e] = x;
        rb[slot][e] = y;
      }
    }
  };
  double kk[W];
#pragma unroll
  for (int t = 0; t < W; ++t) kk[t] = tp.k[t];
  double acc[2] = {0.0, 0.0};
  // plane k is computed from LDS slot sl while planes k + gz (being stored into the other
  // slot) and k + 2 gz are in registers; the loop is unrolled by two for the two register sets
  auto step = [&](long long k, int sl, double (&xa)[NR], double (&xb)[NR]) {
    if (act) {
      const int c = ty * RW + tx + R;
      double sa = 0.0, sb = 0.0;
#pragma unroll
      for (int t = 0; t < W; ++t) {  // tap tau = t - R reads source i - tau
        sa += kk[t] * ra[sl][c - (t - R)];
        sb += kk[t] * rb[sl][c - (t - R)];
      }
      const long long pt = i + q * j + qq * k;
      outa[pt] = sa;
      outb[pt] = sb;
      if (partials) {
        acc[0] += ina[pt];
        acc[1] += inb[pt];
      }
    }
    lstore(sl ^ 1, xa, xb, k + gz);  // the other slot was last read before the previous barrier
    gload(xa, xb, k + 3 * gz);
    __syncthreads();
  };
  const long long kb = blockIdx.z;
  if (kb < q) {
    double a0[NR], b0[NR], a1[NR], b1[NR];
    gload(a0, b0, kb);
    lstore(0, a0, b0, kb);
    gload(a0, b0, kb + gz);
    gload(a1, b1, kb + 2 * gz);
    __syncthreads();
    for (long long k = kb; k < q; k += 2 * gz) {
      step(k, 0, a0, b0);
      if (k + gz < q) step(k + gz, 1, a1, b1);
    }
  }
  if (partials) {
    yee_block_sums<2>(acc, red);
    if (tid == 0) {
      const int b = yee_block(), nb = yee_nblocks();
      partials[b] = acc[0];
      partials[nb + b] = acc[1];
    }
  }
}

// Convolution along y (AXIS 1) or z (AXIS 2) of two arrays: thread (i, o) walks the axis
// through chunk blockIdx.z (length `chunk`) with the 2R+1 taps' inputs held in registers,
// one load per array per point.  o is k for AXIS 1 and j for AXIS 2.
// REDUCE (the z pass of the LWRQ): instead of writing, E and M are reduced into partials
// [sum E, sum M, sum (E - lam_a M)^2 / M, sum (E - lam_a M), non-finite count, max M, min M]
// (lam_a = sc[2]) and rho = E/M is written when asked.
template <int R, int AXIS, bool REDUCE>
__global__ __launch_bounds__(kYeeThreads) void k_lwrq_march(long long q, Taps tp, const double* __restrict__ ina,
                                                            const double* __restrict__ inb, double* outa,
                                                            double* outb, long long chunk, const double* sc,
                                                            double* rho, double* partials) {
  constexpr int W = 2 * R + 1;
  constexpr int NV = 5;
  __shared__ double red[REDUCE ? NV * kYeeThreads : 1];
  __shared__ double redm[REDUCE ? 2 * kYeeThreads : 1];
  const long long i = blockIdx.x * (long long)kYeeTX + threadIdx.x, o = blockIdx.y * (long long)kYeeTY + threadIdx.y;
  const long long s = AXIS == 1 ? q : q * q;
  const long long base = AXIS == 1 ? i + q * q * o : i + q * o;
  const long long p0 = blockIdx.z * chunk, p1 = p0 + chunk < q ? p0 + chunk : q;
  double kk[W];
#pragma unroll
  for (int t = 0; t < W; ++t) kk[t] = tp.k[t];
  double acc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double mx = -INFINITY, mn = INFINITY;
  const double lam = REDUCE ? sc[2] : 0.0;
  if (i < q && o < q && p0 < q) {
    // window w[t] = in(p - R + t); before the first step it holds in(p0 - R - 1 .. p0 + R - 1)
    double wa[W], wb[W];
#pragma unroll
    for (int t = 0; t < W; ++t) {
      const long long src = p0 - R - 1 + t;
      const bool ok = src >= 0 && src < q;
      wa[t] = ok ? ina[base + src * s] : 0.0;
      wb[t] = ok ? inb[base + src * s] : 0.0;
    }
    for (long long p = p0; p < p1; ++p) {
#pragma unroll
      for (int t = 0; t + 1 < W; ++t) {
        wa[t] = wa[t + 1];
        wb[t] = wb[t + 1];
      }
      const long long src = p + R;
      const bool ok = src < q;
      wa[W - 1] = ok ? ina[base + src * s] : 0.0;
      wb[W - 1] = ok ? inb[base + src * s] : 0.0;
      double E = 0.0, M = 0.0;
#pragma unroll
      for (int t = 0; t < W; ++t) {
        // in(p - R + t) carries tap index R - (R - t) ... the taps are symmetric: k[t] = k[W-1-t]
        E += kk[W - 1 - t] * wa[t];
        M += kk[W - 1 - t] * wb[t];
      }
      const long long pt = base + p * s;
      if constexpr (REDUCE) {
        const double d = E - lam * M;
        acc[0] += E;
        acc[1] += M;
        acc[2] += d * d / M;
        acc[3] += d;
        if (!isfinite(M) || !isfinite(E)) acc[4] += 1.0;
        mx = fmax(mx, M);
        mn = fmin(mn, M);
        if (rho) rho[pt] = E / M;
      } else {
        outa[pt] = E;
        outb[pt] = M;
      }
    }
  }
  if constexpr (REDUCE) {
    yee_block_sums<NV>(acc, red);
    const int tid = yee_tid();
    redm[tid] = mx;
    redm[kYeeThreads + tid] = mn;
    __syncthreads();
    if (tid < 64) {
      for (int t = tid + 64; t < kYeeThreads; t += 64) {
        mx = fmax(mx, redm[t]);
        mn = fmin(mn, redm[kYeeThreads + t]);
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_down(mx, off, 64));
        mn = fmin(mn, __shfl_down(mn, off, 64));
      }
    }
    if (tid == 0) {
      const int b = yee_block(), nb = yee_nblocks();
      for (int v = 0; v < NV; ++v) partials[v * nb + b] = acc[v];
      partials[5 * nb + b] = mx;
      partials[6 * nb + b] = mn;
    }
  }
}

struct Kinds {
  int k[8];
};

// out[v] = reduction of partials[v * nb ...]: kind 0 sum, 1 max, 2 min (one workgroup per value).
__global__ __launch_bounds__(256) void k_yee_finish(const double* partials, int nb, Kinds kinds, double* out) {
  __shared__ double red[256];
  const int v = blockIdx.x, kind = kinds.k[v];
  double s = kind == 0 ? 0.0 : (kind == 1 ? -INFINITY : INFINITY);
  for (int i = threadIdx.x; i < nb; i += 256) {
    const double x = partials[(long long)v * nb + i];
    s = kind == 0 ? s + x : (kind == 1 ? fmax(s, x) : fmin(s, x));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int i = 1; i < 256; ++i) r = kind == 0 ? r + red[i] : (kind == 1 ? fmax(r, red[i]) : fmin(r, red[i]));
    out[v] = r;
  }
}

// sc[2] = sc[0] / sc[1] (lambda of the raw densities, the centre of the deviation sums)
// The whole LWRQ in one pass over the two densities: the Gaussian convolution along x, y and
// z and the reductions of the z pass (as k_lwrq_march<R, 2, true>), without writing the x or
// y intermediate.  The workgroup owns a 64 x 8 tile of (i, j) (thread (tx, ty) the rows ty and
// ty + 4) and marches k through chunk blockIdx.z: for each input plane it stages the tile with
// an R-point halo in x and y in LDS (the next plane's values in flight in registers), convolves
// along x into a second LDS array and along y into registers, and pushes the result into a
// register window along z.  The halo rows and columns are read by the neighbouring tiles too
// (from L2); HBM sees each density value about once instead of five times.
constexpr int kLfTY = 8;  // output rows per workgroup (two per thread row)

template <int R>
__global__ __launch_bounds__(kYeeThreads, R <= 3 ? 3 : (R <= 6 ? 2 : 1)) void k_lwrq_fused(long long q, Taps tp, const double* __restrict__ ina,
                                                            const double* __restrict__ inb, long long chunk,
                                                            const double* sc, double* rho, double* partials) {
  constexpr int W = 2 * R + 1, IX = kYeeTX + 2 * R, IY = kLfTY + 2 * R, NI = IX * IY;
  constexpr int NRL = (NI + kYeeThreads - 1) / kYeeThreads;  // staged values per thread and array
  constexpr int NV = 5;
  __shared__ double sa[NI], sb[NI];                  // input plane with halo
  __shared__ double xa[IY * kYeeTX], xb[IY * kYeeTX];  // convolved along x
  __shared__ double red[NV * kYeeThreads];
  __shared__ double redm[2 * kYeeThreads];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = yee_tid();
  const long long i0 = blockIdx.x * (long long)kYeeTX, j0 = blockIdx.y * (long long)kLfTY;
  const long long qq = q * q;
  const long long p0 = blockIdx.z * chunk, p1 = min(q, p0 + chunk);
  double kk[W];
#pragma unroll
  for (int t = 0; t < W; ++t) kk[t] = tp.k[t];
  // This thread's share of a staged plane: element e = tid + 256 r, at (row e / IX, column e % IX)
  // = source (i0 - R + column, j0 - R + row); outside the window it is the zero padding.
  long long off[NRL];
  bool ok[NRL];
#pragma unroll
  for (int r = 0; r < NRL; ++r) {
    const int e = tid + kYeeThreads * r, row = e / IX, col = e - row * IX;
    const long long gi = i0 - R + col, gj = j0 - R + row;
    ok[r] = e < NI && gi >= 0 && gi < q && gj >= 0 && gj < q;
    off[r] = ok[r] ? gi + q * gj : 0;
  }
  double va[NRL], vb[NRL];
  auto gload = [&](long long pl) {  // plane pl (zeros outside 0..q-1)
    const bool in = pl >= 0 && pl < q;
#pragma unroll
    for (int r = 0; r < NRL; ++r) {
      double x = 0.0, y = 0.0;
      if (in && ok[r]) {
        x = ina[off[r] + qq * pl];
        y = inb[off[r] + qq * pl];
      }
      va[r] = x;
      vb[r] = y;
    }
  };
  // Y = the staged plane convolved along x then y at rows ty and ty + 4 of the tile.  Two
  // barriers: sa/sb are last read before the second, xa/xb before the next call's first.
  auto plane_y = [&](double (&ya)[2], double (&yb)[2]) {
#pragma unroll
    for (int r = 0; r < NRL; ++r) {
      const int e = tid + kYeeThreads * r;
      if (e < NI) {
        sa[e] = va[r];
        sb[e] = vb[r];
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int row = ty; row < IY; row += kYeeTY) {  // x: tap tau = t - R reads source i - tau
      const int c = row * IX + tx + R;
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int t = 0; t < W; ++t) {
        a += kk[t] * sa[c - (t - R)];
        b += kk[t] * sb[c - (t - R)];
      }
      xa[row * kYeeTX + tx] = a;
      xb[row * kYeeTX + tx] = b;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // y: window rows (j - R .. j + R) = tile rows jl .. jl + 2R
      const int jl = ty + kYeeTY * h;
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int t = 0; t < W; ++t) {
        a += kk[W - 1 - t] * xa[(jl + t) * kYeeTX + tx];
        b += kk[W - 1 - t] * xb[(jl + t) * kYeeTX + tx];
      }
      ya[h] = a;
      yb[h] = b;
    }
  };
  double acc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double mx = -INFINITY, mn = INFINITY;
  const double lam = sc[2];
  const long long i = i0 + tx;
  if (p0 < p1) {
    // window w[h][t] = Y(p - R + t) of row h; the first step needs planes p0 - R .. p0 + R
    double wa[2][W], wb[2][W];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < W; ++t) wa[h][t] = wb[h][t] = 0.0;
    gload(p0 - R);
#pragma unroll 1
    for (int t = 1; t < W; ++t) {  // planes p0 - R .. p0 + R - 1 into w[.][1..W-1]
      double ya[2], yb[2];
      plane_y(ya, yb);
      gload(p0 - R + t);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int u = 0; u + 1 < W; ++u) {
          wa[h][u] = wa[h][u + 1];
          wb[h][u] = wb[h][u + 1];
        }
        wa[h][W - 1] = ya[h];
        wb[h][W - 1] = yb[h];
      }
    }
#pragma unroll 1
    for (long long pz = p0; pz < p1; ++pz) {  // step pz adds plane pz + R (loaded already)
      double ya[2], yb[2];
      plane_y(ya, yb);
      gload(pz + R + 1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int u = 0; u + 1 < W; ++u) {
          wa[h][u] = wa[h][u + 1];
          wb[h][u] = wb[h][u + 1];
        }
        wa[h][W - 1] = ya[h];
        wb[h][W - 1] = yb[h];
        const long long j = j0 + ty + kYeeTY * h;
        if (i < q && j < q) {
          double E = 0.0, M = 0.0;
#pragma unroll
          for (int t = 0; t < W; ++t) {
            E += kk[W - 1 - t] * wa[h][t];
            M += kk[W - 1 - t] * wb[h][t];
          }
          const double d = E - lam * M;
          acc[0] += E;
          acc[1] += M;
          acc[2] += d * d / M;
          acc[3] += d;
          if (!isfinite(M) || !isfinite(E)) acc[4] += 1.0;
          mx = fmax(mx, M);
          mn = fmin(mn, M);
          if (rho) rho[i + q * j + qq * pz] = E / M;
        }
      }
    }
  }
  yee_block_sums<NV>(acc, red);
  redm[tid] = mx;
  redm[kYeeThreads + tid] = mn;
  __syncthreads();
  if (tid < 64) {
    for (int t = tid + 64; t < kYeeThreads; t += 64) {
      mx = fmax(mx, redm[t]);
      mn = fmin(mn, redm[kYeeThreads + t]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmax(mx, __shfl_down(mx, o, 64));
      mn = fmin(mn, __shfl_down(mn, o, 64));
    }
  }
  if (tid == 0) {
    const int b = yee_block(), nb = yee_nblocks();
    for (int v = 0; v < NV; ++v) partials[v * nb + b] = acc[v];
    partials[5 * nb + b] = mx;
    partials[6 * nb + b] = mn;
  }
}

__global__ void k_yee_ratio(double* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc[2] = sc[0] / sc[1];
}

// ----------------------------------------------------------------------------- host side
// Workgroup grid over (x, y) tiles with gz workgroups along the looped axis: at most
// kYeeMaxBlocks workgroups when the tile count allows (their partial sums are reduced after).
struct YeeLaunch {
  dim3 grid, block;
  int nb;
};

static YeeLaunch yee_launch(long long q, long long along) {
  const long long gx = (q + kYeeTX - 1) / kYeeTX, gy = (q + kYeeTY - 1) / kYeeTY;
  long long gz = kYeeMaxBlocks / (gx * gy);
  gz = std::max<long long>(1, std::min<long long>(gz, along));
  YeeLaunch l;
  l.grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)gz);
  l.block = dim3(kYeeTX, kYeeTY, 1);
  l.nb = (int)(gx * gy * gz);
  return l;
}

// The marching kernels split the walked axis into chunks; chunk >= min_chunk keeps the
// reloaded halo (2R planes for the convolutions, 2 for the point passes) a small fraction of
// each chunk.
static YeeLaunch yee_march_launch(long long q, long long min_chunk, long long& chunk) {
  YeeLaunch l = yee_launch(q, q);
  chunk = (q + l.grid.z - 1) / l.grid.z;
  chunk = std::max<long long>(chunk, std::max<long long>(min_chunk, 1));
  const long long gz = (q + chunk - 1) / chunk;
  l.grid.z = (unsigned)gz;
  l.nb = (int)(l.grid.x * l.grid.y * gz);
  return l;
}

static long long yee_max_partials(long long q) {
  const long long gx = (q + kYeeTX - 1) / kYeeTX, gy = (q + kYeeTY - 1) / kYeeTY;
  return std::max<long long>(kYeeMaxBlocks, gx * gy);
}

// Workspace layout (every part 256-byte aligned): 4 real (2n+1)^3 arrays | partials
// (8 x the largest workgroup count) | scalars (64) | convolution profile (2n+1) | separable
// permittivity factors 6 x (2n+1) | general tables 6 x kYeeMaxRows x (2n+1) (cos, sin).
struct YeeWs {
  double *w0, *w1, *w2, *w3, *part, *sc, *prof, *sep;
  double2* tab;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t yee_ws_bytes(long long q) {
  const size_t Q3 = (size_t)q * q * q;
  return 4 * align256(sizeof(double) * Q3) + align256(sizeof(double) * 8 * yee_max_partials(q)) +
         align256(sizeof(double) * 64) + align256(sizeof(double) * q) + align256(sizeof(double) * 6 * q) +
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
  p += align256(sizeof(double) * 8 * yee_max_partials(q));
  w.sc = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 64);
  w.prof = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * q);
  w.sep = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 6 * q);
  w.tab = reinterpret_cast<double2*>(p);
  return w;
}

// Gaussian taps k = exp(-(off/sigma)^2), off = (-R:R) h, sigma = sigma_factor h
// (exp3/compute_lwrq_exp4.m:12-15), and the 'same' convolution of ones along one axis,
// prof(i) = sum of the taps that land inside 0..q-1; den = prof(i) prof(j) prof(k).
static Taps lwrq_taps(double h, int R, double sigma_factor, long long q, std::vector<double>& prof) {
  if (R < 1 || R > kYeeMaxR) throw Error(FAME6D_EARG, "LWRQ neighbour radius must be in 1..8");
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

// Permittivity tables (see YeeMat).  `sep` and `tab` must stay alive until the stream has
// been synchronised.
static YeeMat upload_material(const YeeWs& w, int n, double h, long long q, const double* P, int L,
                              double epsC, double alpha, hipStream_t s, std::vector<double>& sep,
                              std::vector<double2>& tab) {
  if (L < 1 || L > kYeeMaxRows) throw Error(FAME6D_EARG, "the projection must have 1..16 rows");
  const double a6 = alpha / 6.0;
  bool separable = true;
  for (int l = 0; l < L; ++l) {
    int nz = 0;
    for (int d = 0; d < 3; ++d) nz += P[l * 3 + d] != 0.0;
    separable = separable && nz <= 1;
  }
  auto coord = [&](int st, long long i) { return (double)(i - n) * h + (st ? h / 2 : 0.0); };
  if (separable) {
    sep.assign((size_t)6 * q, 0.0);
    for (int st = 0; st < 2; ++st)
      for (int d = 0; d < 3; ++d)
        for (long long i = 0; i < q; ++i) {
          double S = 0.0;
          for (int l = 0; l < L; ++l) {
            int ax = 0;  // axis of row l (rows of zeros are constant and go to axis 0)
            for (int e = 0; e < 3; ++e)
              if (P[l * 3 + e] != 0.0) ax = e;
            if (ax == d) S += std::cos(P[l * 3 + d] * coord(st, i));
          }
          sep[((size_t)st * 3 + d) * q + i] = std::exp(a6 * S);
        }
    F6_HIP(hipMemcpyAsync(w.sep, sep.data(), sizeof(double) * sep.size(), hipMemcpyHostToDevice, s));
    return YeeMat{w.sep, nullptr, L, epsC, a6};
  }
  tab.assign((size_t)6 * L * q, double2{0.0, 0.0});
  for (int st = 0; st < 2; ++st)
    for (int d = 0; d < 3; ++d)
      for (int l = 0; l < L; ++l)
        for (long long i = 0; i < q; ++i) {
          const double ph = P[l * 3 + d] * coord(st, i);  // P row-major, L x 3
          tab[(((size_t)st * 3 + d) * L + l) * q + i] = double2{std::cos(ph), std::sin(ph)};
        }
  F6_HIP(hipMemcpyAsync(w.tab, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice, s));
  return YeeMat{nullptr, w.tab, L, epsC, a6};
}

struct LwrqSums {
  double sum_a, sum_b, S0, S1, S2, S3;           // pass-1 sums (fused path)
  double sum_E, sum_M, dev2, dev1, bad, max_M, min_M;
};

template <int R>
static void launch_march(long long q, const Taps& tp, const double* ina, const double* inb, double* outa,
                         double* outb, const double* sc, double* rho, double* part, bool reduce, hipStream_t s) {
  long long chunk = 0;
  const YeeLaunch l = yee_march_launch(q, 4 * R, chunk);
  if (reduce)
    k_lwrq_march<R, 2, true><<<l.grid, l.block, 0, s>>>(q, tp, ina, inb, nullptr, nullptr, chunk, sc, rho, part);
  else
    k_lwrq_march<R, 1, false><<<l.grid, l.block, 0, s>>>(q, tp, ina, inb, outa, outb, chunk, nullptr, nullptr,
                                                         nullptr);
  F6_CHECK_LAUNCH();
}

static void march(int R, long long q, const Taps& tp, const double* ina, const double* inb, double* outa,
                  double* outb, const double* sc, double* rho, double* part, bool reduce, hipStream_t s) {
  switch (R) {
    case 1: launch_march<1>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 2: launch_march<2>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 3: launch_march<3>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 4: launch_march<4>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 5: launch_march<5>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 6: launch_march<6>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 7: launch_march<7>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    case 8: launch_march<8>(q, tp, ina, inb, outa, outb, sc, rho, part, reduce, s); break;
    default: throw Error(FAME6D_EARG, "LWRQ neighbour radius must be in 1..8");
  }
}

static void conv_x(int R, const YeeLaunch& l, long long q, const Taps& tp, const double* ina, const double* inb,
                   double* outa, double* outb, const double* prof, double* part, hipStream_t s) {
#define F6_CONV_X(RR)                                                                                     \
  case RR:                                                                                                \
    k_lwrq_conv_x<RR><<<l.grid, l.block, 0, s>>>(q, tp, ina, inb, outa, outb, prof, part);                \
    break;
  switch (R) {
    F6_CONV_X(1)
    F6_CONV_X(2)
    F6_CONV_X(3)
    F6_CONV_X(4)
    F6_CONV_X(5)
    F6_CONV_X(6)
    F6_CONV_X(7)
    F6_CONV_X(8)
    default: throw Error(FAME6D_EARG, "LWRQ neighbour radius must be in 1..8");
  }
#undef F6_CONV_X
  F6_CHECK_LAUNCH();
}

// The three convolution passes of the LWRQ and its reductions; one synchronisation.
// normalise: the x pass divides a, b by den and sums them into sc[0], sc[1] (the standalone
// LWRQ); otherwise a, b are already a/den, b/den and pass 1 has put sc[0..5].  w2/w3 and
// w0/w1 are ping-ponged (a, b may be w0, w1: the x pass reads them before the y pass writes).
// The fused LWRQ pass: (x, y) tiles of 64 x 8 and about 1024 workgroups in all, each z chunk
// at least 4R planes long (its 2R warm-up planes are recomputed by the chunk).
static YeeLaunch lwrq_fused_launch(long long q, int R, long long& chunk) {
  const long long gx = (q + kYeeTX - 1) / kYeeTX, gy = (q + kLfTY - 1) / kLfTY;
  long long gz = std::max<long long>(1, std::min<long long>(q, (1024 + gx * gy - 1) / (gx * gy)));
  chunk = std::max<long long>((q + gz - 1) / gz, std::max(4 * R, 1));
  gz = (q + chunk - 1) / chunk;
  YeeLaunch l;
  l.grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)gz);
  l.block = dim3(kYeeTX, kYeeTY, 1);
  l.nb = (int)(gx * gy * gz);
  return l;
}

static void lwrq_fused(int R, const YeeLaunch& l, long long q, const Taps& tp, const double* ina, const double* inb,
                       long long chunk, const double* sc, double* rho, double* part, hipStream_t s) {
#define F6_LWRQ_FUSED(RR)                                                                              \
  case RR:                                                                                             \
    k_lwrq_fused<RR><<<l.grid, l.block, 0, s>>>(q, tp, ina, inb, chunk, sc, rho, part);                \
    break;
  switch (R) {
    F6_LWRQ_FUSED(1)
    F6_LWRQ_FUSED(2)
    F6_LWRQ_FUSED(3)
    F6_LWRQ_FUSED(4)
    F6_LWRQ_FUSED(5)
    F6_LWRQ_FUSED(6)
    F6_LWRQ_FUSED(7)
    F6_LWRQ_FUSED(8)
    default: throw Error(FAME6D_EARG, "LWRQ neighbour radius must be in 1..8");
  }
#undef F6_LWRQ_FUSED
  F6_CHECK_LAUNCH();
}

static LwrqSums lwrq_device(const YeeWs& w, long long q, const Taps& tp, const double* a, const double* b,
                            double* rho, hipStream_t s, bool normalise) {
  if (!normalise) {
    // a, b already normalised (the metrics pass): one fused pass; sc[2] = sum a / sum b is set
    k_yee_ratio<<<1, 64, 0, s>>>(w.sc);
    long long chunk = 0;
    const YeeLaunch lf = lwrq_fused_launch(q, tp.R, chunk);
    if (lf.nb > yee_max_partials(q)) throw Error(FAME6D_EARG, "LWRQ partials exceed the workspace");
    lwrq_fused(tp.R, lf, q, tp, a, b, chunk, w.sc, rho, w.part, s);
    k_yee_finish<<<7, 256, 0, s>>>(w.part, lf.nb, Kinds{{0, 0, 0, 0, 0, 1, 2}}, w.sc + 8);
    F6_CHECK_LAUNCH();
    double h[16];
    F6_HIP(hipMemcpyAsync(h, w.sc, sizeof(double) * 16, hipMemcpyDeviceToHost, s));
    F6_HIP(hipStreamSynchronize(s));
    return LwrqSums{h[0], h[1], h[3], h[4], h[5], h[6], h[8], h[9], h[10], h[11], h[12], h[13], h[14]};
  }
  const YeeLaunch lx = yee_launch(q, q);
  conv_x(tp.R, lx, q, tp, a, b, w.w2, w.w3, normalise ? w.prof : nullptr, normalise ? w.part : nullptr, s);
  if (normalise) k_yee_finish<<<2, 256, 0, s>>>(w.part, lx.nb, Kinds{{0, 0}}, w.sc);
  k_yee_ratio<<<1, 64, 0, s>>>(w.sc);
  F6_CHECK_LAUNCH();
  march(tp.R, q, tp, w.w2, w.w3, w.w0, w.w1, nullptr, nullptr, nullptr, false, s);
  long long chunk = 0;
  const YeeLaunch lz = yee_march_launch(q, 4 * tp.R, chunk);
  march(tp.R, q, tp, w.w0, w.w1, nullptr, nullptr, w.sc, rho, w.part, true, s);
  k_yee_finish<<<7, 256, 0, s>>>(w.part, lz.nb, Kinds{{0, 0, 0, 0, 0, 1, 2}}, w.sc + 8);
  F6_CHECK_LAUNCH();
  double h[16];
  F6_HIP(hipMemcpyAsync(h, w.sc, sizeof(double) * 16, hipMemcpyDeviceToHost, s));
  F6_HIP(hipStreamSynchronize(s));
  return LwrqSums{h[0], h[1], h[3], h[4], h[5], h[6], h[8], h[9], h[10], h[11], h[12], h[13], h[14]};
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
    long long chunk = 0;
    const YeeLaunch l = yee_march_launch(q, kYeeMinChunk, chunk);
    k_yee_points<YEE_APPLY_A><<<l.grid, l.block, 0, s>>>(
        g, (const cpx<double>*)u1, (const cpx<double>*)u2, (const cpx<double>*)u3, YeeMat{}, (cpx<double>*)A1,
        (cpx<double>*)A2, (cpx<double>*)A3, nullptr, nullptr, nullptr, 0.0, chunk, nullptr);
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
    std::vector<double> sep;
    std::vector<double2> tab;
    const YeeMat mt = upload_material(w, n, h, q, P_host, nrows, eps_c, alpha, s, sep, tab);
    const YeeLaunch l = yee_launch(q, q);
    k_yee_eps<<<l.grid, l.block, 0, s>>>(q, mt, (double*)e1, (double*)e2, (double*)e3);
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
    if (!std::isfinite(lambda6d) || lambda6d == 0.0) throw Error(FAME6D_EARG, "lambda6d must be finite and non-zero");
    hipStream_t s = (hipStream_t)stream;
    const long long q = 2LL * n + 1;
    std::vector<double> prof;
    const Taps tp = lwrq_taps(h, radius, sigma_factor, q, prof);
    YeeScratch scr;
    const YeeWs w = yee_ws_map(scr.get(work, yee_ws_bytes(q)), q);
    F6_HIP(hipMemcpyAsync(w.prof, prof.data(), sizeof(double) * q, hipMemcpyHostToDevice, s));
    std::vector<double> sep;
    std::vector<double2> tab;
    const YeeMat mt = upload_material(w, n, h, q, P_host, nrows, eps_c, alpha, s, sep, tab);
    const YeeGeom g{q, q + 2, (1.0 / h) * (1.0 / h)};
    const auto* f1 = (const cpx<double>*)u1;
    const auto* f2 = (const cpx<double>*)u2;
    const auto* f3 = (const cpx<double>*)u3;
    long long chunk = 0;
    const YeeLaunch l = yee_march_launch(q, kYeeMinChunk, chunk);
    // pass 1: a/den, b/den; sums of a, b and of the residual terms about lambda6d
    k_yee_points<YEE_DENSITIES><<<l.grid, l.block, 0, s>>>(g, f1, f2, f3, mt, nullptr, nullptr, nullptr, w.prof,
                                                           w.w0, w.w1, lambda6d, chunk, w.part);
    F6_CHECK_LAUNCH();
    k_yee_finish<<<1, 256, 0, s>>>(w.part, l.nb, Kinds{{0}}, w.sc);
    k_yee_finish<<<1, 256, 0, s>>>(w.part + l.nb, l.nb, Kinds{{0}}, w.sc + 1);
    k_yee_finish<<<4, 256, 0, s>>>(w.part + 2 * l.nb, l.nb, Kinds{{0, 0, 0, 0}}, w.sc + 3);
    F6_CHECK_LAUNCH();
    // LWRQ on the normalised densities; w0/w1 are free once the x pass has read them
    const LwrqSums r = lwrq_device(w, q, tp, w.w0, w.w1, rho, s, false);
    if (!(r.sum_b > 0)) throw Error(FAME6D_ENOCONV, "Cropped Yee mass is not positive.");
    const fame6d_lwrq_t L = lwrq_finish(r, mass_floor_rel);
    const double eps = 2.220446049250313e-16;
    fame6d_yee_metrics_t o{};
    o.lambda_crop = r.sum_a / r.sum_b;
    o.lambda_3d = L.lambda;
    o.delta_identity = std::fabs(o.lambda_3d - o.lambda_crop) / std::max(std::fabs(o.lambda_crop), eps);
    // residual at lambda_3d from the sums about lambda6d, or a second pass when they cancel
    const double d = o.lambda_3d - lambda6d;
    double r3 = r.S0 - 2.0 * d * r.S1 + d * d * r.S2;
    const double size = r.S0 + 2.0 * std::fabs(d * r.S1) + d * d * r.S2;
    if (!(r3 >= 1e-6 * size)) {
      k_yee_points<YEE_RESIDUALS><<<l.grid, l.block, 0, s>>>(g, f1, f2, f3, mt, nullptr, nullptr, nullptr, nullptr,
                                                             nullptr, nullptr, o.lambda_3d, chunk, w.part);
      F6_CHECK_LAUNCH();
      k_yee_finish<<<1, 256, 0, s>>>(w.part, l.nb, Kinds{{0}}, w.sc + 16);
      F6_CHECK_LAUNCH();
      F6_HIP(hipMemcpyAsync(&r3, w.sc + 16, sizeof(double), hipMemcpyDeviceToHost, s));
      F6_HIP(hipStreamSynchronize(s));
    }
    const double nA = std::sqrt(r.S3), nB = std::sqrt(r.S2);
    o.eta_3d = std::sqrt(std::max(r3, 0.0)) / (nA + std::fabs(o.lambda_3d) * nB);
    o.eta_6d = std::sqrt(r.S0) / (nA + std::fabs(lambda6d) * nB);
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
