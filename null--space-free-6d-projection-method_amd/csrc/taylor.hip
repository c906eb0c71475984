// Multi-center Taylor reconstruction of a Fourier field on Yee points.
//
// u(r) = sum_j c_j exp(i q_j . r)
//      ~ sum_{|alpha|<=p} (r - r_c)^alpha * m_alpha(r_c),
// m_alpha(r_c) = i^|alpha|/alpha! * sum_j q_j^alpha exp(i q_j . r_c) c_j
// (taylor_on_blocks_exp3.m:4-6; reconstruct_component_taylor_batch_exp4.m:14-38).
// Term order alpha = (a1, a2, a3) with a1 outermost (build_multi_index_list_3d_exp3.m:6-12).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

constexpr int kTaylorMaxP = 16;

__host__ __device__ constexpr int taylor_terms(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// ---------------------------------------------------------------------------- moments
// moments = coeff_alpha .* (Q_alpha^T B), B(j, c) = exp(i q_j . r_c) coeff_j: a real
// (L x NF) by complex (NF x C) product, done as one real GEMM with 2C columns
// (re/im interleaved per center) on the FP64 matrix cores.
//
// v_mfma_f64_16x16x4f64 (wave64): lane l supplies A[m = l%16][k = l/16] and
// B[k = l/16][n = l%16]; accumulator double r of lane l is D[(l/16) + 4*r][l%16]
// (row-interleaved; measured on gfx950 with a one-wave probe).
// Workgroup = 4 waves, CT 16-column tiles (8 CT centers) and one slice of the mode range;
// the operands are generated in LDS per chunk of kTChunk modes: powers q_x^s, q_y^s, q_z^s
// (A = products of three of them) and B = exp(i q.r_c) coeff.  Wave w owns row tiles
// w, w+4, ..., w+4(TPW-1) of the (L rounded up to 16) alpha rows, with TPW x CT accumulators
// in AGPRs: each A fragment feeds CT matrix-core ops and each B fragment TPW of them.  The
// power-row offsets of a lane's alphas are fixed for the whole launch and kept in registers;
// alpha rows beyond L read an all-zero power row instead of branching.  Up to 20 accumulators
// (and TPW < 10) two workgroups share a CU (registers capped at 256 per lane, MINB = 2), so one's operand
// generation (sincos, powers) overlaps the other's matrix-core work; more accumulators do not
// fit that cap without spilling and run one workgroup per CU.
constexpr int kTChunk = 64;
constexpr int kTZeroRow = kTaylorMaxP + 1;  // row of zeros in the power tables
constexpr int kTRow = kTChunk + 4;          // LDS row stride of the power tables
typedef double double4_t __attribute__((ext_vector_type(4)));

// Workgroups per CU an instance is compiled for: two when its accumulators leave room in
// 256 registers per lane (TPW = 10, CT = 2 would spill at two).
constexpr int taylor_minb(int tpw, int ct) { return (tpw * ct <= 20 && tpw < 10) ? 2 : 1; }

// Per-axis phase factors of centres that form a tensor grid, c = ix + nx (iy + ny iz) with
// centre (vx[ix], vy[iy], vz[iz]): exp(i q.r_c) = exp(i qx vx) exp(i qy vy) exp(i qz vz), so a
// mode needs nx + ny + nz phases instead of one per centre (216 -> 18 for the 6^3 block
// centres of a Yee window).  tab holds, per mode j, the nx x-factors times coeff_j, then the
// ny y-factors and the nz z-factors; tab == nullptr selects one phase per (mode, centre).
struct AxisPhases {
  const cpx<double>* tab;
  int nx, ny, na;  // na = nx + ny + nz
};

__global__ void k_taylor_axis_phases(long long NF, const double* __restrict__ qx, const double* __restrict__ qy,
                                     const double* __restrict__ qz, const cpx<double>* __restrict__ coeff, int nx,
                                     int ny, int na, const double* __restrict__ vals, cpx<double>* __restrict__ tab) {
  const long long tot = NF * na;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const long long j = e / na;
    const int k = (int)(e - j * na);
    const double q = k < nx ? qx[j] : (k < nx + ny ? qy[j] : qz[j]);
    double sn, cs;
    sincos(q * vals[k], &sn, &cs);
    cpx<double> v = mk(cs, sn);
    if (k < nx) v = v * coeff[j];
    tab[e] = v;
  }
}

// B operand of one chunk of kTChunk modes for the workgroup's 8 CT centres (column group cg):
// bt[jj * BROW + 2 cl + {0, 1}] = re/im of exp(i q_j . r_c) coeff_j, c = cg 8 CT + cl, zero
// for modes past the slice and centres past C.
template <int CT, int BROW>
__device__ __forceinline__ void taylor_build_b(double* bt, long long j0, long long j_end, int cg, int C,
                                               const double* qx, const double* qy, const double* qz,
                                               const cpx<double>* coeff, const double* centers, AxisPhases ax) {
  for (int e = threadIdx.x; e < kTChunk * 8 * CT; e += 256) {
    const int jj = e / (8 * CT), cl = e - jj * (8 * CT);
    const long long j = j0 + jj;
    const int c = cg * 8 * CT + cl;
    double re = 0.0, im = 0.0;
    if (j < j_end && c < C) {
      cpx<double> v;
      if (ax.tab) {
        const int ix = c % ax.nx, r = c / ax.nx;
        const int iy = r % ax.ny, iz = r / ax.ny;
        const cpx<double>* t = ax.tab + j * ax.na;
        v = t[ix] * t[ax.nx + iy] * t[ax.nx + ax.ny + iz];
      } else {
        double sn, cs;
        sincos(qx[j] * centers[3 * c] + qy[j] * centers[3 * c + 1] + qz[j] * centers[3 * c + 2], &sn, &cs);
        v = mk(cs, sn) * coeff[j];
      }
      re = v.re;
      im = v.im;
    }
    bt[jj * BROW + 2 * cl] = re;
    bt[jj * BROW + 2 * cl + 1] = im;
  }
}

template <int TPW, int CT, int MINB>
__global__ __launch_bounds__(256, MINB) void k_taylor_moments_mfma(long long NF, const double* qx, const double* qy,
                                                             const double* qz, const cpx<double>* coeff, int p,
                                                             int L, int C, const double* centers,
                                                             AxisPhases ax, const unsigned char* alist,
                                                             long long span, int Cpad2, double* part) {
  constexpr int NCOL = 16 * CT;  // real columns of this workgroup
  // Padded LDS strides.  A wave reads pw[.][s * kTRow + k4 + kl] for up to 11 powers s and 4
  // values of kl: with a stride of kTChunk doubles every s fell in the same banks (up to
  // 11-way conflicts); kTChunk + 4 spreads them over 8 bank pairs per kl.  B is read as
  // bt[kk * BROW + c * 16 + nl]: a row stride of NCOL + 16 doubles puts consecutive kk in the
  // other half of the banks.
  constexpr int BROW = NCOL + 16;
  const int Lpad = (L + 15) & ~15;
  const int T = Lpad / 16;
  __shared__ double pw[3][(kTZeroRow + 1) * kTRow];
  __shared__ double bt[kTChunk * BROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kl = lane >> 4, nl = lane & 15;
  const int cg = blockIdx.x;  // column group (8 CT centers)
  const long long j_begin = (long long)blockIdx.y * span;
  const long long j_end = min(NF, j_begin + span);

  int o0[TPW], o1[TPW], o2[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int alpha = (wave + 4 * t) * 16 + nl;
    const bool ok = alpha < L;
    o0[t] = (ok ? (int)alist[3 * alpha] : kTZeroRow) * kTRow;
    o1[t] = (ok ? (int)alist[3 * alpha + 1] : 0) * kTRow;
    o2[t] = (ok ? (int)alist[3 * alpha + 2] : 0) * kTRow;
  }
  for (int i = tid; i < kTChunk; i += 256) pw[0][kTZeroRow * kTRow + i] = 0.0;

  double4_t acc[TPW][CT];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = (double4_t){0.0, 0.0, 0.0, 0.0};

  for (long long j0 = j_begin; j0 < j_end; j0 += kTChunk) {
    __syncthreads();  // previous chunk fully consumed
    for (int jj = tid; jj < kTChunk; jj += 256) {
      const long long j = j0 + jj;
      const bool ok = j < j_end;
      const double a = ok ? qx[j] : 0.0, b = ok ? qy[j] : 0.0, d = ok ? qz[j] : 0.0;
      double x = ok ? 1.0 : 0.0, y = x, z = x;
      for (int s = 0; s <= p; ++s) {
        pw[0][s * kTRow + jj] = x;
        pw[1][s * kTRow + jj] = y;
        pw[2][s * kTRow + jj] = z;
        x *= a;
        y *= b;
        z *= d;
      }
    }
    taylor_build_b<CT, BROW>(bt, j0, j_end, cg, C, qx, qy, qz, coeff, centers, ax);
    __syncthreads();
    // two k-steps per trip, so the LDS reads of one overlap the matrix-core ops of the other
    // (one where the two sets of operands would not fit the registers of two workgroups)
#pragma unroll MINB == 2 && TPW * CT > 16 ? 1 : 2
    for (int k4 = 0; k4 < kTChunk; k4 += 4) {
      const int kk = k4 + kl;
      double bv[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) bv[c] = bt[kk * BROW + c * 16 + nl];
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        // row tiles past the last alpha (the fourth wave's last tile at p = 10) issue no
        // matrix-core ops; the test is uniform per wave
        if (wave + 4 * t >= T) continue;
        const double av = pw[0][o0[t] + kk] * pw[1][o1[t] + kk] * pw[2][o2[t] + kk];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[t][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[c], acc[t][c], 0, 0, 0);
      }
    }
  }
  // partial sums of this mode slice: part[(slice*Lpad + alpha)*Cpad2 + col]
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave + 4 * t;
    if (tile < T) {
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int alpha = tile * 16 + kl + 4 * r;
          const int col = cg * NCOL + c * 16 + nl;
          part[((long long)blockIdx.y * Lpad + alpha) * Cpad2 + col] = acc[t][c][r];
        }
    }
  }
}

// Balanced form for 4 column tiles (32 centres per workgroup) and up to 32 row tiles: the
// T = 4 TF + R row tiles of 16 alphas are split so every wave issues the same number of
// matrix-core ops per k-step.  Wave w owns the full row tiles w, w + 4, ..., w + 4 (TF - 1)
// over all 4 column tiles, and R of the 4 R (row tile, column tile) pairs of the last R row
// tiles: pairs g = w R .. w R + R - 1, row tile 4 TF + g / 4, column tile g % 4.  With the
// round-robin form (wave w owns tiles w, w + 4, ...) the waves of the last partial round of
// tiles sat idle at the chunk barrier (at p = 10, 18 tiles: two waves did 5 tiles, two did 4).
template <int TF, int R>
constexpr int taylor_bal_minb() {
  return TF * 4 + R <= 20 ? 2 : 1;
}

// Staged inputs of the balanced form: per-axis phase entries per mode (nx + ny + nz at most)
// and the chunk's three q components, copied to LDS by asynchronous global-to-LDS loads one
// chunk ahead.
constexpr int kStageNa = 18;

// Chunk loop of k_taylor_moments_bal for CTL live column tiles of the workgroup's 4 (the
// remainder pairs run whatever their column).  pw holds the power tables of the three axes
// PWN doubles apart.
//
// STAGE (centres on a tensor grid with nx + ny + nz <= kStageNa, whole chunks): the phase
// entries and q components of chunk i + 1 are loaded straight into LDS
// (global_load_lds_dwordx4, no registers) while the matrix cores work on chunk i; the table
// build of chunk i + 1 then reads LDS only.  Otherwise the build loads its inputs from global
// memory after the chunk barrier and waits for them there.
template <int TF, int R, int CTL, int BROW, int TFA, int RA, int PWN, bool STAGE>
__device__ __forceinline__ void taylor_bal_chunks(double4_t (&acc)[TFA][4], double4_t (&accr)[RA], double* pw,
                                                  double* bt, cpx<double>* tabs, double* qs,
                                                  const int (&o0)[TFA], const int (&o1)[TFA], const int (&o2)[TFA],
                                                  const int (&r0)[RA], const int (&r1)[RA], const int (&r2)[RA],
                                                  const int (&rc)[RA], long long j_begin, long long j_end, int p,
                                                  int cg, int C, const double* qx, const double* qy,
                                                  const double* qz, const cpx<double>* coeff, const double* centers,
                                                  AxisPhases ax) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, kl = lane >> 4;
  // STAGE: this thread's table column (every B element a thread builds has centre
  // cg 32 + (tid & 31)) and its three entries of each mode
  const int cl = tid & 31;
  int kx = 0, ky = 0, kz = 0;
  if constexpr (STAGE) {
    const int c = min(cg * 32 + cl, C - 1);
    const int ix = c % ax.nx, r = c / ax.nx;
    kx = ix;
    ky = ax.nx + r % ax.ny;
    kz = ax.nx + ax.ny + r / ax.ny;
  }
  // asynchronous copy of chunk j0's q components (waves 0, 1) and phase rows to LDS
  auto stage = [&](long long j0) {
    if (wave == 0)
      __builtin_amdgcn_global_load_lds((lane < 32 ? qx + j0 + 2 * lane : qy + j0 + 2 * (lane - 32)), qs, 16, 0, 0);
    else if (wave == 1 && lane < 32)
      __builtin_amdgcn_global_load_lds(qz + j0 + 2 * lane, qs + 2 * kTChunk, 16, 0, 0);
    const int tot = kTChunk * ax.na;
    const cpx<double>* src = ax.tab + j0 * ax.na;
    for (int e0 = 0; e0 < tot; e0 += 256)
      if (e0 + tid < tot) __builtin_amdgcn_global_load_lds(src + e0 + tid, tabs + e0 + wave * 64, 16, 0, 0);
  };
  if constexpr (STAGE) {
    if (j_begin < j_end) stage(j_begin);
  }
  for (long long j0 = j_begin; j0 < j_end; j0 += kTChunk) {
    if constexpr (STAGE) __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this thread's staged loads
    __syncthreads();  // previous chunk fully consumed; the staged inputs of this one visible
    if constexpr (STAGE) {
      for (int e = tid; e < 3 * kTChunk; e += 256) {
        const int d = e / kTChunk, jj = e - d * kTChunk;
        const double a = qs[e];
        double x = 1.0;
        for (int s = 0; s <= p; ++s) {
          pw[d * PWN + s * kTRow + jj] = x;
          x *= a;
        }
      }
      const bool cok = cg * 32 + cl < C;
#pragma unroll
      for (int k = 0; k < kTChunk * 32 / 256; ++k) {
        const int jj = (tid >> 5) + 8 * k;
        const cpx<double>* t = tabs + jj * ax.na;
        const cpx<double> v = t[kx] * t[ky] * t[kz];
        bt[jj * BROW + 2 * cl] = cok ? v.re : 0.0;
        bt[jj * BROW + 2 * cl + 1] = cok ? v.im : 0.0;
      }
    } else {
      // powers: thread (axis, mode) of 3 x kTChunk
      for (int e = tid; e < 3 * kTChunk; e += 256) {
        const int d = e / kTChunk, jj = e - d * kTChunk;
        const long long j = j0 + jj;
        const bool ok = j < j_end;
        const double* qd = d == 0 ? qx : (d == 1 ? qy : qz);
        const double a = ok ? qd[j] : 0.0;
        double x = ok ? 1.0 : 0.0;
        for (int s = 0; s <= p; ++s) {
          pw[d * PWN + s * kTRow + jj] = x;
          x *= a;
        }
      }
      taylor_build_b<4, BROW>(bt, j0, j_end, cg, C, qx, qy, qz, coeff, centers, ax);
    }
    __syncthreads();
    if constexpr (STAGE) {
      // the staged buffers are free again: the next chunk's inputs go in flight now
      if (j0 + kTChunk < j_end) stage(j0 + kTChunk);
    }
    const int nl = tid & 15;
    const double* pa = pw;
    const double* pb = pw + PWN;
    const double* pc = pw + 2 * PWN;
#pragma unroll 1
    for (int k4 = 0; k4 < kTChunk; k4 += 4) {
      const int kk = k4 + kl;
      double bv[CTL];
#pragma unroll
      for (int c = 0; c < CTL; ++c) bv[c] = bt[kk * BROW + c * 16 + nl];
      if constexpr (TF > 0) {
#pragma unroll
        for (int t = 0; t < TF; ++t) {
          const double av = pa[o0[t] + kk] * pb[o1[t] + kk] * pc[o2[t] + kk];
#pragma unroll
          for (int c = 0; c < CTL; ++c) acc[t][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[c], acc[t][c], 0, 0, 0);
        }
      }
      if constexpr (R > 0) {
        if constexpr (R == 2) {
          // both pairs of a wave lie in one row tile (g = 2w, 2w + 1)
          const double av = pa[r0[0] + kk] * pb[r1[0] + kk] * pc[r2[0] + kk];
#pragma unroll
          for (int r = 0; r < 2; ++r)
            accr[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bt[kk * BROW + rc[r]], accr[r], 0, 0, 0);
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double av = pa[r0[r] + kk] * pb[r1[r] + kk] * pc[r2[r] + kk];
            accr[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bt[kk * BROW + rc[r]], accr[r], 0, 0, 0);
          }
        }
      }
    }
  }
}

// PMAX: the largest order the instance serves (its power tables hold rows 0..PMAX and an
// all-zero row PMAX + 1).
template <int TF, int R, int MINB, int PMAX, bool STAGE>
__global__ __launch_bounds__(256, MINB) void k_taylor_moments_bal(long long NF, const double* qx, const double* qy,
                                                                  const double* qz, const cpx<double>* coeff, int p,
                                                                  int L, int C, const double* centers,
                                                                  AxisPhases ax, const unsigned char* alist,
                                                                  int gf, int nsl, long long span,
                                                                  long long span_last, int Cpad2, double* part) {
  constexpr int CT = 4;
  constexpr int NCOL = 16 * CT;
  constexpr int BROW = NCOL + 16;  // padded row stride of B (see k_taylor_moments_mfma)
  constexpr int TFA = TF > 0 ? TF : 1, RA = R > 0 ? R : 1;
  constexpr int ZR = PMAX + 1;              // the zero row of the power tables
  constexpr int PWN = (PMAX + 2) * kTRow;   // doubles per axis
  const int Lpad = (L + 15) & ~15;
  __shared__ double pw[3 * PWN];
  __shared__ double bt[kTChunk * BROW];
  __shared__ __attribute__((aligned(16))) cpx<double> tabs[STAGE ? kTChunk * kStageNa : 1];
  __shared__ __attribute__((aligned(16))) double qs[STAGE ? 3 * kTChunk : 1];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = tid >> 6;
  const int kl = lane >> 4, nl = lane & 15;
  // One-dimensional grid: workgroups 0 .. gf nsl - 1 take column group id % gf and mode slice
  // id / gf of length span; the rest (when the last column group holds fewer than 4 live
  // column tiles) take that group, gf, and slices of length span_last, so that they carry as
  // much work as the others.  The grid holds no idle workgroups: with the round-robin
  // assignment of workgroups to the XCDs, a grid of at most 8 x (resident per XCD)
  // workgroups runs in one round.
  const int id = blockIdx.x;
  const bool last = id >= gf * nsl;
  const int cg = last ? gf : id % gf;
  const long long sp = last ? span_last : span;
  const int slice = last ? id - gf * nsl : id / gf;
  const long long j_begin = (long long)slice * sp;
  const long long j_end = min(NF, j_begin + sp);

  // power-row offsets of this lane's alpha in each owned row tile (alphas past L read the
  // all-zero row)
  auto offsets = [&](int tile, int& a0, int& a1, int& a2) {
    const int alpha = tile * 16 + nl;
    const bool ok = alpha < L;
    a0 = (ok ? (int)alist[3 * alpha] : ZR) * kTRow;
    a1 = (ok ? (int)alist[3 * alpha + 1] : 0) * kTRow;
    a2 = (ok ? (int)alist[3 * alpha + 2] : 0) * kTRow;
  };
  int o0[TFA], o1[TFA], o2[TFA];
#pragma unroll
  for (int t = 0; t < TFA; ++t) offsets(wave + 4 * t, o0[t], o1[t], o2[t]);
  int r0[RA], r1[RA], r2[RA], rc[RA];
#pragma unroll
  for (int r = 0; r < RA; ++r) {
    const int g = wave * R + r;
    offsets(4 * TF + g / 4, r0[r], r1[r], r2[r]);
    rc[r] = (g % 4) * 16 + nl;
  }
  for (int i = tid; i < kTChunk; i += 256) pw[ZR * kTRow + i] = 0.0;

  double4_t acc[TFA][CT], accr[RA];
#pragma unroll
  for (int t = 0; t < TFA; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < RA; ++r) accr[r] = (double4_t){0.0, 0.0, 0.0, 0.0};

  // column tiles holding a centre: the last column group can hold fewer than 4, and its
  // full row tiles then issue no matrix-core ops for the empty ones
  const int ctl = min(CT, (C - cg * 8 * CT + 7) / 8);
  if (ctl >= 4)
    taylor_bal_chunks<TF, R, 4, BROW, TFA, RA, PWN, STAGE>(acc, accr, pw, bt, tabs, qs, o0, o1, o2, r0, r1, r2, rc,
                                                           j_begin, j_end, p, cg, C, qx, qy, qz, coeff, centers, ax);
  else
    taylor_bal_chunks<TF, R, 3, BROW, TFA, RA, PWN, STAGE>(acc, accr, pw, bt, tabs, qs, o0, o1, o2, r0, r1, r2, rc,
                                                           j_begin, j_end, p, cg, C, qx, qy, qz, coeff, centers, ax);
  // partial sums of this mode slice: part[(slice*Lpad + alpha)*Cpad2 + col]
  double* ps = part + (long long)slice * Lpad * Cpad2 + cg * NCOL;
  if constexpr (TF > 0) {
#pragma unroll
    for (int t = 0; t < TF; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) ps[(long long)((wave + 4 * t) * 16 + kl + 4 * q) * Cpad2 + c * 16 + nl] = acc[t][c][q];
  }
  if constexpr (R > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int g = wave * R + r;
      const int tile = 4 * TF + g / 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) ps[(long long)(tile * 16 + kl + 4 * q) * Cpad2 + rc[r]] = accr[r][q];
    }
  }
}

// Final moments: sum the mode slices, form re + i*im and multiply by i^|alpha|/alpha!.
__global__ void k_taylor_finish(int L, int C, int Cpad2, int nslices_all, int c_last, int nsl_last,
                                const double* part, const cpx<double>* calpha, cpx<double>* mom) {
  // consecutive threads take consecutive centres of one alpha row, so every slice's row is
  // read in order (16 bytes per thread); four running sums keep four loads in flight
  const int Lpad = (L + 15) & ~15;
  const long long tot = (long long)L * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(i / C), c = (int)(i % C);
    const int nslices = c >= c_last ? nsl_last : nslices_all;  // slices of c's column group
    const double* col = part + (long long)a * Cpad2 + 2 * c;
    const long long zs = (long long)Lpad * Cpad2;
    double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
    int z = 0;
    for (; z + 4 <= nslices; z += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 v = *reinterpret_cast<const double2*>(col + (z + u) * zs);
        re[u] += v.x;
        im[u] += v.y;
      }
    }
    for (; z < nslices; ++z) {
      const double2 v = *reinterpret_cast<const double2*>(col + z * zs);
      re[0] += v.x;
      im[0] += v.y;
    }
    const double sr = (re[0] + re[1]) + (re[2] + re[3]), si = (im[0] + im[1]) + (im[2] + im[3]);
    mom[(long long)c * L + a] = calpha[a] * mk(sr, si);  // column-major L x C
  }
}

// Nested Horner of sum_alpha dx^a1 dy^a2 dz^a3 m[alpha] (alpha ordered a1, a2, a3).
__device__ __forceinline__ cpx<double> taylor_horner(const cpx<double>* mc, int p, double dx, double dy,
                                                     double dz) {
  cpx<double> acc1 = mk(0.0, 0.0);
  double x1 = 1.0;
  int idx = 0;
  for (int a1 = 0; a1 <= p; ++a1) {
    cpx<double> acc2 = mk(0.0, 0.0);
    double y2 = 1.0;
    for (int a2 = 0; a2 <= p - a1; ++a2) {
      cpx<double> acc3 = mk(0.0, 0.0);
      double z3 = 1.0;
      for (int a3 = 0; a3 <= p - a1 - a2; ++a3) {
        acc3 = acc3 + z3 * mc[idx++];
        z3 *= dz;
      }
      acc2 = acc2 + y2 * acc3;
      y2 *= dy;
    }
    acc1 = acc1 + x1 * acc2;
    x1 *= dx;
  }
  return acc1;
}

// Evaluation: one thread per point, nested Horner over (a1, a2, a3).
__global__ void k_taylor_eval(long long M, const double* px, const double* py, const double* pz,
                              const int* cid, int p, int L, int C, const double* centers,
                              const cpx<double>* mom, cpx<double>* u) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = cid[i];
    if (c < 0 || c >= C) {  // invalid center id: flag the point instead of reading out of bounds
      u[i] = mk<double>(NAN, NAN);
      continue;
    }
    u[i] = taylor_horner(mom + (long long)c * L, p, px[i] - centers[3 * c], py[i] - centers[3 * c + 1],
                         pz[i] - centers[3 * c + 2]);
  }
}

// Block of a plus-grid index: nb nearly equal ranges, the first mod(m, nb) one longer
// (build_yee_block_partition_exp3.m:21-31).
__device__ __forceinline__ int block_of(int x, int base, int rem) {
  const int cut = rem * (base + 1);
  return x < cut ? x / (base + 1) : rem + (x - cut) / base;
}

// Taylor evaluation on the (2n+3)^3 plus-grid points of Yee component comp, generated on the
// fly: point (I,J,K) = h*((I,J,K) - (n+1) + 0.5 e_comp) (build_yee_edge_points_plus.m:11-17),
// column-major order, center = block (bI + nb*bJ + nb^2*bK).
__global__ void k_taylor_eval_yee(int n, double h, int comp, int nb, int p, int L, int C,
                                  const double* centers, const cpx<double>* mom, cpx<double>* u) {
  const int m = 2 * n + 3;
  const long long M = (long long)m * m * m;
  const int base = m / nb, rem = m % nb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const int I = (int)(i % m), J = (int)((i / m) % m), K = (int)(i / ((long long)m * m));
    const int c = block_of(I, base, rem) + nb * block_of(J, base, rem) + nb * nb * block_of(K, base, rem);
    if (c >= C) {
      u[i] = mk<double>(NAN, NAN);
      continue;
    }
    const double x = h * ((I - (n + 1)) + (comp == 0 ? 0.5 : 0.0));
    const double y = h * ((J - (n + 1)) + (comp == 1 ? 0.5 : 0.0));
    const double z = h * ((K - (n + 1)) + (comp == 2 ? 0.5 : 0.0));
    u[i] = taylor_horner(mom + (long long)c * L, p, x - centers[3 * c], y - centers[3 * c + 1],
                         z - centers[3 * c + 2]);
  }
}

// The same evaluation factored along x-rows.  Points of one plus-grid row (fixed J, K) share
// y and z, so for each block bI of the row and each a1 the partial sum
//   B[bI][a1] = sum_{a2, a3} dy^a2 dz^a3 m_c[a1, a2, a3],  c = bI + nb bJ + nb^2 bK,
// is common to all its points, and a point needs only the Horner sum over a1 of
// dx^a1 B[bI(I)][a1] (p + 1 terms instead of L).  One wave per row: its lanes form the
// nb (p + 1) partial sums (each over the contiguous run of terms with that a1, which the
// alpha order a1-outermost makes consecutive in the moment column), stage them in LDS, then
// evaluate the row's m points in order (coalesced stores).
constexpr int kEvalRowItems = 256;  // nb * (p + 1) at most (else the point-wise kernel runs)
__global__ __launch_bounds__(256) void k_taylor_eval_yee_rows(int n, double h, int comp, int nb, int p, int L,
                                                              int C, const double* __restrict__ centers,
                                                              const cpx<double>* __restrict__ mom,
                                                              cpx<double>* __restrict__ u) {
  __shared__ cpx<double> Bs[4][kEvalRowItems];
  const int m = 2 * n + 3;
  const int base = m / nb, rem = m % nb;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long row = (long long)blockIdx.x * 4 + wave;  // row = J + m K
  const bool live = row < (long long)m * m;
  const int J = live ? (int)(row % m) : 0, K = live ? (int)(row / m) : 0;
  const int cJK = nb * block_of(J, base, rem) + nb * nb * block_of(K, base, rem);
  const double y = h * ((J - (n + 1)) + (comp == 1 ? 0.5 : 0.0));
  const double z = h * ((K - (n + 1)) + (comp == 2 ? 0.5 : 0.0));
  const int P1 = p + 1, items = nb * P1;
  if (live) {
    for (int it = lane; it < items; it += 64) {
      const int bI = it / P1, a1 = it - bI * P1;
      const int c = bI + cJK;
      cpx<double> acc = mk<double>(NAN, NAN);
      if (c < C) {
        const double dy = y - centers[3 * c + 1], dz = z - centers[3 * c + 2];
        // first term of a1: sum over a < a1 of (p - a + 1)(p - a + 2) / 2
        int t = 0;
        for (int a = 0; a < a1; ++a) t += (p - a + 1) * (p - a + 2) / 2;
        const cpx<double>* mc = mom + (long long)c * L + t;
        const int q = p - a1;
        acc = mk(0.0, 0.0);
        double y2 = 1.0;
        for (int a2 = 0; a2 <= q; ++a2) {
          cpx<double> acc3 = mk(0.0, 0.0);
          double z3 = 1.0;
          for (int a3 = 0; a3 <= q - a2; ++a3) {
            acc3 = acc3 + z3 * mc[a3];
            z3 *= dz;
          }
          mc += q - a2 + 1;
          acc = acc + y2 * acc3;
          y2 *= dy;
        }
      }
      Bs[wave][it] = acc;
    }
  }
  __syncthreads();
  if (!live) return;
  cpx<double>* ur = u + row * m;
  for (int I = lane; I < m; I += 64) {
    const int bI = block_of(I, base, rem);
    const int c = bI + cJK;
    if (c >= C) {
      ur[I] = mk<double>(NAN, NAN);
      continue;
    }
    const double dx = h * ((I - (n + 1)) + (comp == 0 ? 0.5 : 0.0)) - centers[3 * c];
    const cpx<double>* B = Bs[wave] + bI * P1;
    cpx<double> acc = B[p];
    for (int a1 = p - 1; a1 >= 0; --a1) acc = dx * acc + B[a1];
    ur[I] = acc;
  }
}

// Direct sum u(r) = sum_j c_j exp(i q_j . r) (brute force, one thread per point).
__global__ void k_direct(long long NF, const double* qx, const double* qy, const double* qz,
                         const cpx<double>* coeff, long long M, const double* px, const double* py,
                         const double* pz, cpx<double>* u) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = px[i], y = py[i], z = pz[i];
    double sr = 0, si = 0;
    for (long long j = 0; j < NF; ++j) {
      double sn, cs;
      sincos(qx[j] * x + qy[j] * y + qz[j] * z, &sn, &cs);
      const cpx<double> cj = coeff[j];
      sr += cs * cj.re - sn * cj.im;
      si += cs * cj.im + sn * cj.re;
    }
    u[i] = mk(sr, si);
  }
}

// Device buffers of the moments call (tables, per-slice partial moments, per-axis phases),
// kept per device between calls and only grown, so repeated calls do not pay hipMalloc and
// hipFree; the mutex serialises the calls (each synchronises its stream before it returns,
// when the buffers are free again).  The buffers live until the process ends.
struct MomentsWorkspace {
  DevBuf tabs, part, axt;
};
constexpr int kMaxDevices = 64;
static std::mutex g_mom_mu;
static MomentsWorkspace* g_mom_ws[kMaxDevices];

// Largest per-axis phase table (NF x (nx + ny + nz) complex) the moments call allocates.
constexpr double kAxisTableMaxBytes = 2.0e9;

// Do the centres form a tensor grid in the order c = ix + nx (iy + ny iz), centre c =
// (vx[ix], vy[iy], vz[iz])?  On success vals = vx, vy, vz (nx + ny + nz values).  Every
// centre is compared exactly, so the per-axis factorisation of exp(i q . r_c) holds for each.
static bool tensor_grid(const double* cen, int C, int& nx, int& ny, int& nz, std::vector<double>& vals) {
  nx = 1;
  while (nx < C && cen[3 * nx + 1] == cen[1] && cen[3 * nx + 2] == cen[2]) ++nx;
  if (C % nx) return false;
  ny = 1;
  while ((long long)ny * nx < C && cen[3 * ny * nx + 2] == cen[2]) ++ny;
  if (C % (nx * ny)) return false;
  nz = C / (nx * ny);
  for (int c = 0; c < C; ++c) {
    const int ix = c % nx, iy = (c / nx) % ny, iz = c / (nx * ny);
    if (cen[3 * c] != cen[3 * ix] || cen[3 * c + 1] != cen[3 * iy * nx + 1] ||
        cen[3 * c + 2] != cen[3 * iz * nx * ny + 2])
      return false;
  }
  vals.clear();
  for (int i = 0; i < nx; ++i) vals.push_back(cen[3 * i]);
  for (int i = 0; i < ny; ++i) vals.push_back(cen[3 * i * nx + 1]);
  for (int i = 0; i < nz; ++i) vals.push_back(cen[3 * i * nx * ny + 2]);
  return true;
}

static void alpha_tables(int p, std::vector<int>& al, std::vector<cpx<double>>& ca) {
  al.clear();
  ca.clear();
  for (int a1 = 0; a1 <= p; ++a1)
    for (int a2 = 0; a2 <= p - a1; ++a2)
      for (int a3 = 0; a3 <= p - a1 - a2; ++a3) {
        al.push_back(a1);
        al.push_back(a2);
        al.push_back(a3);
        const int s = a1 + a2 + a3;
        const double f = std::exp(-(std::lgamma(a1 + 1.0) + std::lgamma(a2 + 1.0) + std::lgamma(a3 + 1.0)));
        // i^s
        const int r = s % 4;
        const double re = (r == 0) ? 1.0 : (r == 2 ? -1.0 : 0.0);
        const double im = (r == 1) ? 1.0 : (r == 3 ? -1.0 : 0.0);
        ca.push_back(mk(re * f, im * f));
      }
}

// ---------------------------------------------------------------------------- field errors
// Per-workgroup sums of |t_c - r_c|^2 and |r_c|^2 and maxima of |t_c - r_c| and |r_c| over
// the three components (compute_field_errors_exp3.m:3-10); the host combines the partials
// in workgroup order, so the result does not depend on scheduling.
constexpr int kErrBlocks = 512;
__global__ void __launch_bounds__(256) k_field_err(long long n, const cpx<double>* __restrict__ t1,
                                                   const cpx<double>* __restrict__ t2,
                                                   const cpx<double>* __restrict__ t3,
                                                   const cpx<double>* __restrict__ r1,
                                                   const cpx<double>* __restrict__ r2,
                                                   const cpx<double>* __restrict__ r3, double* __restrict__ part) {
  double sd = 0.0, sr = 0.0, md = 0.0, mr = 0.0;
  const cpx<double>* ts[3] = {t1, t2, t3};
  const cpx<double>* rs[3] = {r1, r2, r3};
  for (int c = 0; c < 3; ++c) {
    const cpx<double>* t = ts[c];
    const cpx<double>* r = rs[c];
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
      const cpx<double> a = t[i], b = r[i];
      const double dr = a.re - b.re, di = a.im - b.im;
      const double d2 = dr * dr + di * di, r2v = b.re * b.re + b.im * b.im;
      sd += d2;
      sr += r2v;
      md = fmax(md, d2);
      mr = fmax(mr, r2v);
    }
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = sd;
  red[1][threadIdx.x] = sr;
  red[2][threadIdx.x] = md;
  red[3][threadIdx.x] = mr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace fame6d

using namespace fame6d;

extern "C" {

int fame6d_taylor_moments(long long NF, const double* qmodes, const void* coeff, int p, int C,
                          const double* centers_host, void* moments, void* stream) {
  try {
    if (!qmodes || !coeff || !centers_host || !moments) throw Error(FAME6D_EARG, "null argument");
    if (p < 0 || p > kTaylorMaxP || C <= 0 || NF <= 0) throw Error(FAME6D_EARG, "bad Taylor sizes");
    hipStream_t s = (hipStream_t)stream;
    std::vector<int> al;
    std::vector<cpx<double>> ca;
    alpha_tables(p, al, ca);
    const int L = (int)ca.size();
    const int Lpad = (L + 15) & ~15;
    const int T = Lpad / 16;
    const bool balanced = T <= 32;         // 4 column tiles, balanced row tiles (p <= 12)
    const int tpw = (T + 3) / 4;           // row tiles per wave of the round-robin form
    const int ct = balanced ? 4 : 2;       // 16-column tiles per workgroup
    const int cgroups = (C + 8 * ct - 1) / (8 * ct);
    const int Cpad2 = cgroups * 16 * ct;
    // mode slices, each a multiple of the chunk: the grid is kept at whole rounds of the
    // resident workgroups rather than just over a multiple, which would add a round for a
    // few stragglers
    int dev = 0;
    F6_HIP(hipGetDevice(&dev));
    int cus = 256;
    {
      int n = 0;
      if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) cus = n;
    }
    const int minb = balanced ? (T / 4 * 4 + T % 4 <= 20 ? 2 : 1) : taylor_minb(tpw, ct);
    const int resident = cus * minb;
    const long long chunks = (NF + kTChunk - 1) / kTChunk;
    // one round for the balanced form (fewest partial sums: at N = 4, p = 10, 216 centres one
    // and two rounds measured the same and four 6% slower), four for the round-robin form
    const int rounds = balanced ? 1 : 4;
    long long target = std::max<long long>(1, ((long long)rounds * resident) / cgroups);
    int nslices = (int)std::max<long long>(1, std::min<long long>(chunks, target));
    long long span = ((chunks + nslices - 1) / nslices) * kTChunk;
    nslices = (int)((NF + span - 1) / span);
    // balanced form with a partly filled last column group (it runs 3 column tiles): its
    // slices are 4/3 as long, and the span is the shortest for which all workgroups fit the
    // rounds of resident workgroups
    long long span_last = span;
    int nsl_last = nslices;
    int gf = cgroups;  // column groups of the balanced grid that walk slices of length span
    const int ctl_last = std::min(4, (C - (cgroups - 1) * 32 + 7) / 8);
    if (balanced && cgroups > 1 && ctl_last < 4) {
      gf = cgroups - 1;
      const long long cap = (long long)rounds * resident;
      for (long long sc = 1; sc <= chunks; ++sc) {
        const long long nf = (chunks + sc - 1) / sc, scl = sc * 4 / 3, nl = (chunks + scl - 1) / scl;
        if ((cgroups - 1) * nf + nl <= cap || sc == chunks) {
          span = sc * kTChunk;
          span_last = scl * kTChunk;
          nslices = (int)((NF + span - 1) / span);
          nsl_last = (int)((NF + span_last - 1) / span_last);
          break;
        }
      }
    }
    // centres on a tensor grid: per-axis phase tables
    int nx = 0, ny = 0, nz = 0;
    std::vector<double> axv;
    const bool axis = tensor_grid(centers_host, C, nx, ny, nz, axv) && nx + ny + nz < C &&
                      (double)NF * (nx + ny + nz) * sizeof(cpx<double>) <= kAxisTableMaxBytes;
    std::vector<unsigned char> alb(al.begin(), al.end());
    if (dev < 0 || dev >= kMaxDevices) throw Error(FAME6D_EARG, "device index beyond the workspace table");
    std::lock_guard<std::mutex> lock(g_mom_mu);
    if (!g_mom_ws[dev]) g_mom_ws[dev] = new MomentsWorkspace();
    DevBuf& tabs = g_mom_ws[dev]->tabs;
    DevBuf& part = g_mom_ws[dev]->part;
    DevBuf& axt = g_mom_ws[dev]->axt;
    const size_t off_ca = ((alb.size() + 255) / 256) * 256;
    const size_t off_cen = off_ca + sizeof(cpx<double>) * ca.size();
    const size_t off_axv = off_cen + sizeof(double) * 3 * C;
    tabs.ensure(off_axv + sizeof(double) * axv.size());
    char* b = tabs.as<char>();
    F6_HIP(hipMemcpyAsync(b, alb.data(), alb.size(), hipMemcpyHostToDevice, s));
    F6_HIP(hipMemcpyAsync(b + off_ca, ca.data(), sizeof(cpx<double>) * ca.size(), hipMemcpyHostToDevice, s));
    F6_HIP(hipMemcpyAsync(b + off_cen, centers_host, sizeof(double) * 3 * C, hipMemcpyHostToDevice, s));
    part.ensure(sizeof(double) * (size_t)nslices * Lpad * Cpad2);
    dim3 grid(cgroups, nslices);
    const int grid_bal = gf * nslices + (cgroups - gf) * nsl_last;
    const auto* cf = (const cpx<double>*)coeff;
    const auto* cen = (const double*)(b + off_cen);
    const auto* alp = (const unsigned char*)b;
    AxisPhases ax{nullptr, 1, 1, 3};
    if (axis) {
      const int na = nx + ny + nz;
      F6_HIP(hipMemcpyAsync(b + off_axv, axv.data(), sizeof(double) * axv.size(), hipMemcpyHostToDevice, s));
      axt.ensure(sizeof(cpx<double>) * (size_t)NF * na);
      k_taylor_axis_phases<<<grid_for(NF * na), 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, cf, nx, ny,
                                                             na, (const double*)(b + off_axv), axt.as<cpx<double>>());
      F6_CHECK_LAUNCH();
      ax = AxisPhases{axt.as<cpx<double>>(), nx, ny, na};
    }
    // staged inputs: tensor-grid phase tables that fit the LDS staging area, whole chunks
    const bool stage = axis && ax.na <= kStageNa && NF % kTChunk == 0;
    if (balanced) {
      switch (p) {
#define F6_TAYLOR_BAL(P)                                                                                          \
  case P: {                                                                                                      \
    constexpr int TT = (taylor_terms(P) + 15) / 16;                                                              \
    constexpr int PM = P < 2 ? 2 : P; /* p = 0, 1, 2 share the one-row-tile instance */                        \
    if (stage)                                                                                                   \
      k_taylor_moments_bal<TT / 4, TT % 4, taylor_bal_minb<TT / 4, TT % 4>(), PM, true>                          \
          <<<grid_bal, 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, cf, p, L, C, cen, ax, alp, gf,    \
                                    nslices, span, span_last, Cpad2, part.as<double>());                        \
    else                                                                                                         \
      k_taylor_moments_bal<TT / 4, TT % 4, taylor_bal_minb<TT / 4, TT % 4>(), PM, false>                         \
          <<<grid_bal, 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, cf, p, L, C, cen, ax, alp, gf,    \
                                    nslices, span, span_last, Cpad2, part.as<double>());                        \
    break;                                                                                                       \
  }
        F6_TAYLOR_BAL(0) F6_TAYLOR_BAL(1) F6_TAYLOR_BAL(2) F6_TAYLOR_BAL(3) F6_TAYLOR_BAL(4) F6_TAYLOR_BAL(5)
        F6_TAYLOR_BAL(6) F6_TAYLOR_BAL(7) F6_TAYLOR_BAL(8) F6_TAYLOR_BAL(9) F6_TAYLOR_BAL(10) F6_TAYLOR_BAL(11)
        F6_TAYLOR_BAL(12)
#undef F6_TAYLOR_BAL
        default:
          throw Error(FAME6D_EARG, "Taylor order without a balanced instance");
      }
    } else {
      switch (tpw) {
#define F6_TAYLOR_CASE(TP, CTV)                                                                                   \
  case TP:                                                                                                       \
    k_taylor_moments_mfma<TP, CTV, taylor_minb(TP, CTV)><<<grid, 256, 0, s>>>(                                  \
        NF, qmodes, qmodes + NF, qmodes + 2 * NF, cf, p, L, C, cen, ax, alp, span, Cpad2, part.as<double>());   \
    break;
        F6_TAYLOR_CASE(9, 2) F6_TAYLOR_CASE(10, 2) F6_TAYLOR_CASE(11, 2) F6_TAYLOR_CASE(12, 2)
        F6_TAYLOR_CASE(13, 2) F6_TAYLOR_CASE(14, 2) F6_TAYLOR_CASE(15, 2) F6_TAYLOR_CASE(16, 2)
#undef F6_TAYLOR_CASE
        default:
          throw Error(FAME6D_EARG, "Taylor order too large");
      }
    }
    F6_CHECK_LAUNCH();
    k_taylor_finish<<<grid_for((long long)L * C), 256, 0, s>>>(L, C, Cpad2, nslices, gf * 8 * ct, nsl_last,
                                                               part.as<double>(),
                                                               (const cpx<double>*)(b + off_ca),
                                                               (cpx<double>*)moments);
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(s));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_taylor_evaluate(long long M, const double* points, const int* center_id, int p, int C,
                           const double* centers_host, const void* moments, void* u, void* stream) {
  try {
    if (!points || !center_id || !centers_host || !moments || !u) throw Error(FAME6D_EARG, "null argument");
    if (p < 0 || p > kTaylorMaxP || C <= 0 || M < 0) throw Error(FAME6D_EARG, "bad Taylor sizes");
    hipStream_t s = (hipStream_t)stream;
    DevBuf cen;
    cen.ensure(sizeof(double) * 3 * C);
    F6_HIP(hipMemcpyAsync(cen.ptr, centers_host, sizeof(double) * 3 * C, hipMemcpyHostToDevice, s));
    const int L = taylor_terms(p);
    if (M > 0) {
      k_taylor_eval<<<grid_for(M), 256, 0, s>>>(M, points, points + M, points + 2 * M, center_id, p, L, C,
                                                 cen.as<double>(), (const cpx<double>*)moments,
                                                 (cpx<double>*)u);
      F6_CHECK_LAUNCH();
    }
    F6_HIP(hipStreamSynchronize(s));
    cen.release();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_taylor_evaluate_yee(int n, double h, int comp, int nblocks, int p, int C,
                               const double* centers_host, const void* moments, void* u, void* stream) {
  try {
    if (!centers_host || !moments || !u) throw Error(FAME6D_EARG, "null argument");
    if (n < 0 || comp < 0 || comp > 2 || nblocks < 1 || nblocks > 2 * n + 3 || p < 0 || p > kTaylorMaxP ||
        C != nblocks * nblocks * nblocks)
      throw Error(FAME6D_EARG, "bad Yee/Taylor sizes");
    hipStream_t s = (hipStream_t)stream;
    DevBuf cen;
    cen.ensure(sizeof(double) * 3 * C);
    F6_HIP(hipMemcpyAsync(cen.ptr, centers_host, sizeof(double) * 3 * C, hipMemcpyHostToDevice, s));
    const long long m = 2LL * n + 3;
    if (nblocks * (p + 1) <= kEvalRowItems) {
      const long long rows = m * m;
      k_taylor_eval_yee_rows<<<(unsigned)((rows + 3) / 4), 256, 0, s>>>(
          n, h, comp, nblocks, p, taylor_terms(p), C, cen.as<double>(), (const cpx<double>*)moments,
          (cpx<double>*)u);
    } else {
      k_taylor_eval_yee<<<grid_for(m * m * m), 256, 0, s>>>(n, h, comp, nblocks, p, taylor_terms(p), C,
                                                             cen.as<double>(), (const cpx<double>*)moments,
                                                             (cpx<double>*)u);
    }
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(s));
    cen.release();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_field_errors(long long n, const void* t1, const void* t2, const void* t3, const void* r1,
                        const void* r2, const void* r3, double* e2, double* einf, void* stream) {
  try {
    if (!t1 || !t2 || !t3 || !r1 || !r2 || !r3 || !e2 || !einf) throw Error(FAME6D_EARG, "null argument");
    if (n < 0) throw Error(FAME6D_EARG, "negative point count");
    hipStream_t s = (hipStream_t)stream;
    const int blocks = (int)std::max<long long>(1, std::min<long long>(kErrBlocks, (n + 255) / 256));
    DevBuf part;
    part.ensure(sizeof(double) * 4 * blocks);
    k_field_err<<<blocks, 256, 0, s>>>(n, (const cpx<double>*)t1, (const cpx<double>*)t2, (const cpx<double>*)t3,
                                       (const cpx<double>*)r1, (const cpx<double>*)r2, (const cpx<double>*)r3,
                                       part.as<double>());
    F6_CHECK_LAUNCH();
    std::vector<double> h(4 * blocks);
    F6_HIP(hipMemcpyAsync(h.data(), part.ptr, sizeof(double) * 4 * blocks, hipMemcpyDeviceToHost, s));
    F6_HIP(hipStreamSynchronize(s));
    double sd = 0, sr = 0, md = 0, mr = 0;
    for (int b = 0; b < blocks; ++b) {
      sd += h[4 * b];
      sr += h[4 * b + 1];
      md = std::max(md, h[4 * b + 2]);
      mr = std::max(mr, h[4 * b + 3]);
    }
    const double eps = 2.220446049250313e-16;  // MATLAB eps, the floor of both denominators
    *e2 = std::sqrt(sd) / std::max(std::sqrt(sr), eps);
    *einf = std::sqrt(md) / std::max(std::sqrt(mr), eps);
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_direct_reconstruct(long long NF, const double* qmodes, const void* coeff, long long M,
                              const double* points, void* u, void* stream) {
  try {
    if (!qmodes || !coeff || !points || !u) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (M > 0) {
      k_direct<<<grid_for(M), 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, (const cpx<double>*)coeff,
                                            M, points, points + M, points + 2 * M, (cpx<double>*)u);
      F6_CHECK_LAUNCH();
    }
    F6_HIP(hipStreamSynchronize(s));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

}  // extern "C"
