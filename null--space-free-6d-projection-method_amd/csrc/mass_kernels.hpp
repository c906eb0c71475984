// Circulant 6D mass operator  M_s x = fftn(eps .* ifftn(x))  on the (2N)^6 torus.
//
// Reference: applyMij_fft (exp1/applyMij_fft.m:9-23); the rank-1 lattice form
// applyMij_rank1_exact (exp1/applyMij_rank1_exact.m:4-5) computes the same matrix
// (kernel indices are reduced mod nGrid, build_exact_rank1_scalar_kernel_fft.m:26-34).
//
// Layout: a grid vector is the column-major (2N)^6 array, linear index
//   i = k1 + n*k2 + n^2*k3 + S*(k4 + n*k5 + n^2*k6),  n = 2N, S = n^3,
// i.e. "inner" position a = i mod S (dims 1-3) and "outer" position b = i / S
// (dims 4-6).  Complex values are interleaved.
//
// The apply is three HBM sweeps:
//   pass A: per outer position b (a contiguous S-block): pre-op, inverse DFT over
//           dims 1,2,3 (register pencil along dim 1, LDS exchange for dims 2,3);
//   pass B: per tile of A consecutive inner positions: inverse DFT over dims 4,5,6,
//           multiply by eps/NF, forward DFT over dims 6,5,4 (register pencils,
//           LDS exchange), data stays on chip between the two transforms;
//   pass C: per S-block: forward DFT over dims 3,2,1 and post-op.
// The pre/post ops fuse the Method-C modewise work around the scalar applies
// (apply_Mhat_C_scalar_compact.m:8-21): one read of x feeds the three scalar
// applies M_s z, M_s(W1 z), M_s(W2 z), and pass C accumulates
// invC*(s0 + W1 s1 + W2 s2) in registers.
#pragma once

#include <cstddef>
#include <type_traits>

#include "common.hpp"
#include "dft.hpp"

namespace fame6d {

// Per-group device state of a conjugate-gradient solve (all scalars FP64).  The
// device runs MATLAB pcg's whole control flow (recurrence test, true-residual check,
// "moresteps" resume); the host only reads `finished` once per launched chunk.
struct CgState {
  double rho;         // r'*r (r'*z preconditioned) of the current residual
  double alpha;       // rho / (p'*q)
  double beta;        // rho_new / rho_old
  double pq;          // p'*q
  double normr;       // ||r|| (recurrence)
  double tolb;        // tol * ||b||
  double n2b;         // ||b||
  double rho_prev;    // rho before the last update
  double normr_true;  // ||b - A x|| of the last true-residual check
  int iter;           // completed iterations
  int done;           // 1: the iteration kernels skip this group
  int flag;           // MATLAB pcg flag convention (0 ok, 1 maxit, 3 stagnation, 4 breakdown, -1 NaN)
  int maxit;
  int check_every;    // every later iteration ends with a true-residual check
  int finished;       // 1: solve complete (confirmed by the true residual, or failed)
  int moresteps;      // failed true-residual checks so far
  int checks;         // true-residual checks (one A x each)
  int maxmsteps;      // failed checks allowed before flag 3
  int pad[5];
};
static_assert(sizeof(CgState) == 128, "CgState layout");

// Per-group early-exit mask: group g is skipped when p[g*stride] != 0.
struct Skip {
  const int* p;
  int stride;
};
__host__ __device__ inline Skip no_skip() { return Skip{nullptr, 0}; }
__host__ inline Skip skip_done(const CgState* st) {
  if (!st) return no_skip();
  return Skip{reinterpret_cast<const int*>(reinterpret_cast<const char*>(st) + offsetof(CgState, done)),
              (int)(sizeof(CgState) / sizeof(int))};
}
__device__ __forceinline__ bool skipped(Skip s, int g) { return s.p != nullptr && s.p[g * s.stride] != 0; }
// Group -> q-system slot (nullptr = identity).
__device__ __forceinline__ long long qslot(const int* qmap, int g) { return qmap ? qmap[g] : g; }

enum PreOp : int {
  PRE_COPY = 0,     // t = x                                  (1 output)
  PRE_MHAT = 1,     // z = invC*x; t = [z, W1 z, W2 z]        (3 outputs)
  PRE_MHAT_CG = 2,  // p = r + beta p (p = r first); then as PRE_MHAT on p
  // the modewise steps of K_r^{-1} around its Mhat solve, folded into the M_s apply that
  // follows each (FP64 only; group g reads x + g*ldx):
  PRE_V1H = 3,  // t = V1^{-*} [x1; x2] (x2 at x1 + NF)      (2 outputs)
  PRE_ZW = 4,   // z = invC*x; t = [W1 z, W2 z]             (2 outputs)
};
template <int OP>
constexpr int pre_outputs() {
  return OP == PRE_COPY ? 1 : (OP == PRE_V1H || OP == PRE_ZW) ? 2 : 3;
}
enum PostOp : int {
  POST_COPY = 0,     // y = s
  POST_MHAT = 1,     // y = invC*(s0 + W1 s1 + W2 s2)  [W real]
  POST_MHAT_CG = 2,  // q = as POST_MHAT, not stored: x += alpha p, r -= alpha q (CG update)
};

template <class T>
struct PassAArgs {
  const cpx<T>* x;  // input of group g at x + g*ldx (the residual r for PRE_MHAT_CG)
  long long ldx;
  cpx<T>* t;  // output o of group g at t + (g*nout + o)*ldt
  long long ldt;
  const T* W1;  // modewise diagonals of group g at W1 + g*ldq
  const T* W2;
  const T* invC;
  long long ldq;
  cpx<T>* p;  // PRE_MHAT_CG search direction of group g at p + g*ldx
  const CgState* st;  // per group (PRE_MHAT_CG)
  Skip skip;
  const int* qmap;
  // PRE_MHAT_CG, optional: the pcg step after the previous iteration (the work of
  // k_cg_beta) is folded into this pass.  Every workgroup reduces the nrr pass-C partials
  // of ||r||^2 of group g (rr + g*nrr) and updates a copy of st[g]; workgroup 0 stores the
  // updated state in st_next[g], which the rest of the iteration uses.  No workgroup writes
  // st, so all of them see the same old state.
  CgState* st_next;
  const double* rr;
  int nrr;
  // ... and for the preconditioned form (the work of k_pcg_beta) also the nrz partials of
  // r'z (rz + g*nrz), which become rho; ||r|| is then taken from rr alone.
  const double* rz;
  int nrz;
  // PRE_V1H: the 2x2 blocks [V11 V12; V21 V22] of group g at V11 + g*ldq, ...
  const double* V11;
  const double* V12;
  const double* V21;
  const double* V22;
};

template <class T>
struct PassBArgs {
  cpx<T>* t;  // vector r of group g at t + (g*NR + r)*ldt
  long long ldt;
  const T* eps;  // eps / NF
  Skip skip;
  // MODE 0, optional: partials[v*gridDim.x + tile] = sum eps/NF * |G u_v|^2 over the tile,
  // G u_v the unnormalised 6D inverse DFT of vector v.  Summed over the NR vectors of a
  // group this is u'(F D G)u = p' Mhat p for the Mhat pre-op vectors u_r = W_r invC p.
  double* partials;
};

template <class T>
struct PassCArgs {
  const cpx<T>* t;  // as PassAArgs::t
  long long ldt;
  cpx<T>* y;  // output of group g at y + g*ldy
  long long ldy;
  const T* W1;
  const T* W2;
  const T* invC;
  long long ldq;
  const cpx<T>* dotp;  // if set: partial sums of Re(dotp^H y) per workgroup
  long long lddot;
  const cpx<T>* resb;  // POST_MHAT, if set: y = resb - Mhat x (same layout as y), ||y||^2 partials
  // POST_MHAT, optional: the transformed outputs 1 and 2 before they are combined,
  // M_s(W1 z) and M_s(W2 z) with z = invC x, at zw + (2g)*ldy and zw + (2g+1)*ldy
  cpx<T>* zw;
  double* partials;  // partials[g*gridDim.x + blockIdx.x]
  Skip skip;
  const int* qmap;
  // POST_MHAT_CG: alpha = rho / p'q with p'q the sum of npq pass-B energy partials per
  // group; x, p, r of group g at base + g*ldx; partials receive ||r_new||^2.
  CgState* cg;
  const double* pq_partials;
  int npq;
  cpx<T>* x;
  const cpx<T>* p;
  cpx<T>* r;
  long long ldx;
};

// XCD-aware block order: the dispatcher deals consecutive linear block ids to the 8 XCDs
// round-robin, each XCD with its own L2.  This bijection of [0, total) gives the blocks
// an XCD runs consecutive logical ids, so neighbouring tiles that share cache lines
// meet in the same L2.
__device__ __forceinline__ int xcd_swizzle(int lin, int total) {
  constexpr int X = 8;
  const int per = (total + X - 1) / X;  // ids per XCD (the first `tall` XCDs get per)
  const int tall = (total % X == 0) ? X : total % X;
  const int xcd = lin % X, local = lin / X;
  return (xcd < tall) ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

template <int n>
struct Grid {
  static constexpr int P2 = n * n;
  static constexpr int S = n * n * n;
};

// Workgroup sum of one double per thread in a fixed order; result valid in
// thread 0.  `red` must hold THREADS doubles.  Values are staged through LDS so
// that partial waves never shuffle from inactive lanes; wave 0 (always full when
// THREADS >= 64) then folds them.  Callers synchronise before reusing `red`.
template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  double s = 0.0;
  if constexpr (THREADS >= 64) {
    if (threadIdx.x < 64) {
      for (int i = threadIdx.x; i < THREADS; i += 64) s += red[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    }
  } else {
    if (threadIdx.x == 0)
      for (int i = 0; i < THREADS; ++i) s += red[i];
  }
  return s;
}

// s + src[i0] + src[i0 + stride] + ... over i < n, added in that order (the order of the plain
// strided loop, so every caller that sums the same partials gets the same value), with the
// loads issued four at a time: the plain loop waited for each load before the next.
__device__ __forceinline__ double strided_sum(double s, const double* src, int i0, int n, int stride) {
  for (int i = i0; i < n; i += 4 * stride) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = i + u * stride < n ? src[i + u * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) s += v[u];
  }
  return s;
}

// MATLAB pcg's bookkeeping after an unpreconditioned iteration: iteration count, recurrence
// convergence test, maxit, breakdown, beta = rho_new / rho (make_explicit_krinv_compact.m:43
// calls pcg with x0 = 0).  rho_new = ||r_new||^2 is the fixed-order sum of the pass-C
// partials: thread stride kCgSumThreads, then a workgroup sum, the same value in k_cg_beta
// and in the pass-A fold below (extra threads add zeros).
constexpr int kCgSumThreads = 256;
// The fields a step changes, as scalars (a local copy of the whole struct would be kept in
// scratch memory).
struct CgStep {
  int iter, done, flag;
  double normr, beta, rho, rho_prev;
};
// The pcg step after an iteration from rr = ||r||^2 and rho_new = r'z (= rr unpreconditioned).
__device__ __forceinline__ CgStep cg_pbeta_step(const CgState* st, double rr, double rho_new) {
  CgStep c{st->iter + 1, st->done, st->flag, sqrt(rr), st->beta, rho_new, st->rho};
  if (!isfinite(rr) || !isfinite(rho_new)) {
    c.flag = -1;
    c.done = 1;
  } else if (c.normr <= st->tolb || st->check_every) {
    c.flag = 0;  // provisional: the true residual is checked after the chunk
    c.done = 1;
  } else if (c.iter >= st->maxit) {
    c.flag = 1;
    c.done = 1;
  } else {
    const double beta = rho_new / st->rho;
    if (beta == 0.0 || isinf(beta) || rho_new == 0.0) {
      c.flag = 4;
      c.done = 1;
    }
    c.beta = beta;
  }
  return c;
}
__device__ __forceinline__ CgStep cg_beta_step(const CgState* st, double rho_new) {
  return cg_pbeta_step(st, rho_new, rho_new);
}
__device__ __forceinline__ void cg_store_step(CgState* d, const CgStep& c) {
  d->iter = c.iter;
  d->done = c.done;
  d->flag = c.flag;
  d->normr = c.normr;
  d->beta = c.beta;
  d->rho = c.rho;
  d->rho_prev = c.rho_prev;
}

// State of group g for a PRE_MHAT_CG pass A: false when the group is done (the caller
// returns).  With a.st_next the previous iteration's pcg step is applied here (see
// PassAArgs); red is LDS scratch of THREADS doubles, free again on return.
template <int THREADS, class T>
__device__ __forceinline__ bool cg_pass_a_state(const PassAArgs<T>& a, int g, double* red, bool& use_beta,
                                                T& beta) {
  if (a.st_next == nullptr) {
    if (skipped(a.skip, g)) return false;
    use_beta = a.st[g].iter > 0;
    beta = static_cast<T>(a.st[g].beta);
    return true;
  }
  const int tid = threadIdx.x;
  // k_cg_beta's (k_pcg_beta's) fixed-order sum of the nb partials at src, on every workgroup
  auto fold_sum = [&](const double* src, int nb) {
    double v = 0.0;
    if constexpr (THREADS >= kCgSumThreads) {
      if (tid < kCgSumThreads)
        v = strided_sum(v, src, tid, nb, kCgSumThreads);
      v = block_sum<THREADS>(v, red);
    } else {
      // fewer threads than k_cg_beta has: each takes the lanes tid, tid + THREADS, ... of a
      // kCgSumThreads-wide sum (in LDS of their own: red may hold fewer doubles), and the
      // first wave adds them up as block_sum does
      static_assert(THREADS >= 64, "one full wave");
      __shared__ double lane[kCgSumThreads];
      for (int l = tid; l < kCgSumThreads; l += THREADS) {
        double u = 0.0;
        u = strided_sum(u, src, l, nb, kCgSumThreads);
        lane[l] = u;
      }
      __syncthreads();
      if (tid < 64) {
        for (int i = tid; i < kCgSumThreads; i += 64) v += lane[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      }
    }
    __syncthreads();  // red / lane free for the next sum
    return v;
  };
  const double v = fold_sum(a.rr + (long long)g * a.nrr, a.nrr);
  const double vz = a.rz ? fold_sum(a.rz + (long long)g * a.nrz, a.nrz) : v;
  if (tid == 0) {
    const CgState* o = a.st + g;
    const bool was_done = o->done != 0;
    const CgStep c = was_done ? CgStep{o->iter, o->done, o->flag, o->normr, o->beta, o->rho, o->rho_prev}
                              : cg_pbeta_step(o, v, vz);
    if (blockIdx.x == 0) {
      a.st_next[g] = *o;
      if (!was_done) cg_store_step(a.st_next + g, c);
    }
    red[0] = c.done ? 1.0 : 0.0;
    red[1] = c.iter > 0 ? 1.0 : 0.0;
    red[2] = c.beta;
  }
  __syncthreads();
  const bool done = red[0] != 0.0;
  use_beta = red[1] != 0.0;
  beta = static_cast<T>(red[2]);
  __syncthreads();  // red is LDS the caller reuses
  return !done;
}

// invC = 1/sqrt(1 + W1^2 + W2^2) (build_experiment1_q_system_C_compact.m:26-31) recomputed
// from W1 and W2 with the operations, order and rounding of the host's table (no
// contraction; IEEE square root and division), so it equals the table entry: the FP64 M-hat
// pass C, which has W1 and W2 in hand, reads 8 bytes per mode less.  The FP32 passes read
// their table, whose entries are the FP64 values rounded.  (Pass A keeps the table: there
// each of its three thread groups would load both W and repeat the square root, which
// measured no faster.)
__device__ __forceinline__ double inv_c_from(double w1, double w2) {
#pragma clang fp contract(off)
  const double korth = 1.0 + w1 * w1 + w2 * w2;
  return 1.0 / sqrt(korth);
}
// Weight of output o of the M-hat pre-op at diagonal index q: invC, invC W1, invC W2.  Both
// loads are issued by every thread (output 0 reads invC twice): a wave can hold threads of
// two outputs, and a load under `if (o > 0)` became a divergent branch that waited for every
// earlier load before its own, once per pencil element.
template <class T>
__device__ __forceinline__ T mhat_weight(const T* invC, const T* W1, const T* W2, int o, long long q) {
  const T* wp = o == 0 ? invC : (o == 1 ? W1 : W2);
  const T a = invC[q], b = wp[q];
  return o > 0 ? a * b : a;
}

// invC at q for the post-op: recomputed from the W1, W2 values in hand in FP64, read in FP32.
template <class T>
__device__ __forceinline__ T inv_c_or(const T* invC, T w1, T w2, long long q) {
  if constexpr (std::is_same<T, double>::value) return inv_c_from(w1, w2);
  else return invC[q];
}

// ---------------------------------------------------------------- pass A
template <class T, int n, int BPW, int OP>
__global__ __launch_bounds__(BPW* n* n) void k_pass_a(PassAArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int NOUT = (OP == PRE_COPY) ? 1 : 3;
  __shared__ cpx<T> lds[BPW * S];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int blk = tid / P2, pp = tid - blk * P2;
  const long long off = ((long long)blockIdx.x * BPW + blk) * S;
  cpx<T>* L = lds + blk * S;

  // dim-1 pencil of this thread: (k2, k3) = (pp % n, pp / n)
  const int base1 = (pp % n) * n + (pp / n) * P2;
  const cpx<T>* xg = a.x + g * a.ldx + off + base1;
  cpx<T> pn[n];
#pragma unroll
  for (int k = 0; k < n; ++k) pn[k] = xg[k];

  if constexpr (OP == PRE_MHAT_CG) {
    cpx<T>* pg = a.p + g * a.ldx + off + base1;
    if (a.st[g].iter > 0) {
      const T beta = static_cast<T>(a.st[g].beta);
#pragma unroll
      for (int k = 0; k < n; ++k) pn[k] = pn[k] + beta * pg[k];
    }
#pragma unroll
    for (int k = 0; k < n; ++k) pg[k] = pn[k];
  }

  const long long qoff = qslot(a.qmap, g) * a.ldq + off + base1;
#pragma unroll 1
  for (int o = 0; o < NOUT; ++o) {
    cpx<T> u[n];
    if constexpr (OP == PRE_COPY) {
#pragma unroll
      for (int k = 0; k < n; ++k) u[k] = pn[k];
    } else {
      const T* wo = (o == 1) ? a.W1 : a.W2;
#pragma unroll
      for (int k = 0; k < n; ++k) {
        T w = a.invC[qoff + k];
        if (o > 0) w *= wo[qoff + k];
        u[k] = w * pn[k];
      }
    }
    dft<T, n, +1>(u);
#pragma unroll
    for (int k = 0; k < n; ++k) L[base1 + k] = u[k];
    __syncthreads();
    {  // dim 2: (k1, k3) = (pp % n, pp / n)
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
      dft<T, n, +1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
    }
    __syncthreads();
    {  // dim 3: (k1, k2) = pp; store straight to HBM (coalesced across pp)
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[pp + P2 * k];
      dft<T, n, +1>(w);
      cpx<T>* tg = a.t + (g * NOUT + o) * a.ldt + off + pp;
#pragma unroll
      for (int k = 0; k < n; ++k) tg[P2 * k] = w[k];
    }
    if (o + 1 < NOUT) __syncthreads();
  }
}

// ---------------------------------------------------------------- pass A, coalesced I/O
// Same transform as k_pass_a with every HBM access coalesced: the register pencil of
// the first DFT is along dim 3 (stride n^2, so consecutive threads read consecutive
// modes), dims 2 and 1 go through LDS, and the result leaves LDS in linear order.
// THREADS = BPW*n^2 and each thread moves exactly n values per block set.
template <class T, int n, int BPW, int OP>
__global__ __launch_bounds__(BPW* n* n) void k_pass_a_lin(PassAArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int NOUT = pre_outputs<OP>();
  constexpr int THREADS = BPW * P2;
  __shared__ cpx<T> lds[BPW * S];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int blk = tid / P2, pp = tid - blk * P2;
  const long long wg0 = (long long)blockIdx.x * BPW * S;  // first mode of this workgroup
  const long long off = wg0 + (long long)blk * S;
  cpx<T>* L = lds + blk * S;

  // dim-3 pencil of this thread: (k1, k2) = pp, elements pp + n^2*k
  const cpx<T>* xg = a.x + g * a.ldx + off + pp;
  cpx<T> pn[n];
#pragma unroll
  for (int k = 0; k < n; ++k) pn[k] = xg[P2 * k];
  cpx<T> pn2[OP == PRE_V1H ? n : 1];  // second component (x2) for PRE_V1H
  if constexpr (OP == PRE_V1H) {
#pragma unroll
    for (int k = 0; k < n; ++k) pn2[k] = xg[(long long)S * S + P2 * k];
  }
  if constexpr (OP == PRE_MHAT_CG) {
    cpx<T>* pg = a.p + g * a.ldx + off + pp;
    if (a.st[g].iter > 0) {
      const T beta = static_cast<T>(a.st[g].beta);
#pragma unroll
      for (int k = 0; k < n; ++k) pn[k] = pn[k] + beta * pg[P2 * k];
    }
#pragma unroll
    for (int k = 0; k < n; ++k) pg[P2 * k] = pn[k];
  }
  const long long qoff = qslot(a.qmap, g) * a.ldq + off + pp;
#pragma unroll 1
  for (int o = 0; o < NOUT; ++o) {
    cpx<T> u[n];
    if constexpr (OP == PRE_COPY) {
#pragma unroll
      for (int k = 0; k < n; ++k) u[k] = pn[k];
    } else if constexpr (OP == PRE_V1H) {
      // y1 = (V22 x1 - V21 x2) / det, y2 = (V11 x2 - V12 x1) / det (solveV1H_blockdiag.m:7-8);
      // both at the first output, y2 kept in pn2 for the second (V is read once)
      if (o == 0) {
#pragma unroll
        for (int k = 0; k < n; ++k) {
          const long long j = qoff + P2 * k;
          const double v11 = a.V11[j], v12 = a.V12[j], v21 = a.V21[j], v22 = a.V22[j];
          const double det = v11 * v22 - v12 * v21;
          u[k] = (1.0 / det) * (v22 * pn[k] - v21 * pn2[k]);
          pn2[k] = (1.0 / det) * (v11 * pn2[k] - v12 * pn[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < n; ++k) u[k] = pn2[k];
      }
    } else if constexpr (OP == PRE_ZW) {
      // z = invC u, [W1 z; W2 z] (make_explicit_krinv_compact.m:60-62)
      const T* wo = (o == 0) ? a.W1 : a.W2;
#pragma unroll
      for (int k = 0; k < n; ++k) u[k] = wo[qoff + P2 * k] * (a.invC[qoff + P2 * k] * pn[k]);
    } else {
      const T* wo = (o == 1) ? a.W1 : a.W2;
#pragma unroll
      for (int k = 0; k < n; ++k) {
        T w = a.invC[qoff + P2 * k];
        if (o > 0) w *= wo[qoff + P2 * k];
        u[k] = w * pn[k];
      }
    }
    dft<T, n, +1>(u);
#pragma unroll
    for (int k = 0; k < n; ++k) L[pp + P2 * k] = u[k];
    __syncthreads();
    {  // dim 2: (k1, k3) = (pp % n, pp / n)
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
      dft<T, n, +1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
    }
    __syncthreads();
    {  // dim 1: (k2, k3) = (pp % n, pp / n)
      const int base1 = (pp % n) * n + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
      dft<T, n, +1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base1 + k] = w[k];
    }
    __syncthreads();
    cpx<T>* tg = a.t + (g * NOUT + o) * a.ldt + wg0;
#pragma unroll
    for (int i = 0; i < n; ++i) tg[tid + THREADS * i] = lds[tid + THREADS * i];
    if (o + 1 < NOUT) __syncthreads();
  }
}

// ---------------------------------------------------------------- pass B
// MODE 0: inverse DFT dims 4-6, * eps/NF, forward DFT dims 6-4 (the mass apply).
// MODE 1: inverse DFT dims 4-6 only (completes an unnormalised 6D ifft after pass A).
// grid = (tiles, vectors); vector v belongs to group v / NR (for the skip mask).  The
// block order is vector-fastest within a tile, so the NR (and batched-group) vectors of
// one tile run back to back on one XCD and share the eps tile in its L2.
template <class T, int n, int A, int NR, int MODE>
__global__ __launch_bounds__(A* n* n) void k_pass_b(PassBArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int THREADS = A * P2;
  __shared__ cpx<T> L[A * S];  // l = al + A*(k4 + n*k5 + n^2*k6)
  const int nt = gridDim.x, nv = gridDim.y;
  const int logical = xcd_swizzle(blockIdx.x + nt * blockIdx.y, nt * nv);
  const int tile = logical / nv, vidx = logical - tile * nv;
  if (skipped(a.skip, vidx / NR)) return;
  const int tid = threadIdx.x;
  const int al = tid % A, qq = tid / A;  // qq in [0, n^2)
  const long long a0 = (long long)tile * A + al;
  cpx<T>* v = a.t + vidx * a.ldt + a0;
  // FP32: the permittivity of this thread's dim-6 pencil is loaded with the input, so it is
  // in registers when the dim-6 stage needs it (a load there, after two LDS stages and their
  // barriers, waits its full latency).  FP64 loads it in the stage: held from the start it
  // made the FP64 kernel 3% slower at N = 5.
  constexpr bool PRE_EPS = MODE == 0 && std::is_same<T, float>::value;
  T ep[n];
  if constexpr (PRE_EPS) {
#pragma unroll
    for (int k = 0; k < n; ++k) ep[k] = a.eps[a0 + (long long)S * (qq + P2 * k)];
  }
  {  // dim 4 inverse: (k5, k6) = qq
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = v[(long long)S * (k + n * qq)];
    dft<T, n, +1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[al + A * (k + n * qq)] = w[k];
  }
  __syncthreads();
  const int b5 = al + A * ((qq % n) + P2 * (qq / n));  // (k4, k6) = qq
  {
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[b5 + A * n * k];
    dft<T, n, +1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[b5 + A * n * k] = w[k];
  }
  __syncthreads();
  double en = 0.0;  // sum eps/NF |G6 z|^2 = Re z^H F6 (eps/NF) G6 z of this thread's dim-6 pencil z
  {  // dim 6: inverse, eps, forward
    const int b6 = al + A * qq;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[b6 + A * P2 * k];
    dft<T, n, +1>(w);
    if constexpr (MODE == 1) {
#pragma unroll
      for (int k = 0; k < n; ++k) v[(long long)S * (qq + P2 * k)] = w[k];
      return;
    } else {
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = (PRE_EPS ? ep[k] : a.eps[a0 + (long long)S * (qq + P2 * k)]) * w[k];
      dft<T, n, -1>(w);
      if (a.partials != nullptr) {
        // z (still in LDS) against F6 eps G6 z: no extra registers over the plain apply
#pragma unroll
        for (int k = 0; k < n; ++k) {
          const cpx<T> z = L[b6 + A * P2 * k];
          en += static_cast<double>(z.re) * w[k].re + static_cast<double>(z.im) * w[k].im;
          L[b6 + A * P2 * k] = w[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < n; ++k) L[b6 + A * P2 * k] = w[k];
      }
    }
  }
  __syncthreads();
  {
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[b5 + A * n * k];
    dft<T, n, -1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[b5 + A * n * k] = w[k];
  }
  __syncthreads();
  {  // dim 4 forward, store
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[al + A * (k + n * qq)];
    dft<T, n, -1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) v[(long long)S * (k + n * qq)] = w[k];
  }
  if (a.partials != nullptr) {
    static_assert(sizeof(cpx<T>) * A * S >= sizeof(double) * THREADS, "LDS reuse for the reduction");
    __syncthreads();  // every thread has read L
    en = block_sum<THREADS>(en, reinterpret_cast<double*>(L));
    if (tid == 0) a.partials[(long long)vidx * nt + tile] = en;
  }
}

// Pass B with two adjacent inner positions per thread, FP32 only: a thread's pencil element
// is one 16-byte load or store (re, im of positions 2t and 2t + 1), where the one-position
// form issues 8-byte accesses, and a tile of A positions is a run of 8 A bytes at stride S
// with half the threads of the one-position form.  Same arithmetic per pencil.
template <int n, int A, int NR, int MODE>
__global__ __launch_bounds__(A / 2 * n * n) void k_pass_b2(PassBArgs<float> a) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int AT = A / 2;  // threads per pencil column
  constexpr int THREADS = AT * P2;
  static_assert(A % 2 == 0, "pairs of inner positions");
  __shared__ cpx<float> L[A * S];  // l = al + A*(k4 + n*k5 + n^2*k6)
  const int nt = gridDim.x, nv = gridDim.y;
  const int logical = xcd_swizzle(blockIdx.x + nt * blockIdx.y, nt * nv);
  const int tile = logical / nv, vidx = logical - tile * nv;
  if (skipped(a.skip, vidx / NR)) return;
  const int tid = threadIdx.x;
  const int ap = tid % AT, qq = tid / AT;  // qq in [0, n^2)
  const int al = 2 * ap;
  const long long a0 = (long long)tile * A + al;
  cpx<float>* v = a.t + vidx * a.ldt + a0;
  auto ld2 = [&](long long off, cpx<float>& x, cpx<float>& y) {
    const f4 u = *reinterpret_cast<const f4*>(v + off);
    x = mk(u.x, u.y);
    y = mk(u.z, u.w);
  };
  auto st2 = [&](long long off, const cpx<float>& x, const cpx<float>& y) {
    *reinterpret_cast<f4*>(v + off) = (f4){x.re, x.im, y.re, y.im};
  };
  auto lds2 = [&](int l, cpx<float>& x, cpx<float>& y) {
    const f4 u = *reinterpret_cast<const f4*>(L + l);
    x = mk(u.x, u.y);
    y = mk(u.z, u.w);
  };
  auto sts2 = [&](int l, const cpx<float>& x, const cpx<float>& y) {
    *reinterpret_cast<f4*>(L + l) = (f4){x.re, x.im, y.re, y.im};
  };
  // the permittivity of this thread's dim-6 pencils, loaded with the input (see k_pass_b)
  f2 ep[n];
  if constexpr (MODE == 0) {
#pragma unroll
    for (int k = 0; k < n; ++k) ep[k] = *reinterpret_cast<const f2*>(a.eps + a0 + (long long)S * (qq + P2 * k));
  }
  {  // dim 4 inverse: (k5, k6) = qq
    cpx<float> w0[n], w1[n];
#pragma unroll
    for (int k = 0; k < n; ++k) ld2((long long)S * (k + n * qq), w0[k], w1[k]);
    dft<float, n, +1>(w0);
    dft<float, n, +1>(w1);
#pragma unroll
    for (int k = 0; k < n; ++k) sts2(al + A * (k + n * qq), w0[k], w1[k]);
  }
  __syncthreads();
  const int b5 = al + A * ((qq % n) + P2 * (qq / n));  // (k4, k6) = qq
  {
    cpx<float> w0[n], w1[n];
#pragma unroll
    for (int k = 0; k < n; ++k) lds2(b5 + A * n * k, w0[k], w1[k]);
    dft<float, n, +1>(w0);
    dft<float, n, +1>(w1);
#pragma unroll
    for (int k = 0; k < n; ++k) sts2(b5 + A * n * k, w0[k], w1[k]);
  }
  __syncthreads();
  double en = 0.0;  // as in k_pass_b, over both positions
  {  // dim 6: inverse, eps, forward
    const int b6 = al + A * qq;
    cpx<float> w0[n], w1[n];
#pragma unroll
    for (int k = 0; k < n; ++k) lds2(b6 + A * P2 * k, w0[k], w1[k]);
    dft<float, n, +1>(w0);
    dft<float, n, +1>(w1);
    if constexpr (MODE == 1) {
#pragma unroll
      for (int k = 0; k < n; ++k) st2((long long)S * (qq + P2 * k), w0[k], w1[k]);
      return;
    } else {
#pragma unroll
      for (int k = 0; k < n; ++k) {
        w0[k] = ep[k].x * w0[k];
        w1[k] = ep[k].y * w1[k];
      }
      dft<float, n, -1>(w0);
      dft<float, n, -1>(w1);
      if (a.partials != nullptr) {
#pragma unroll
        for (int k = 0; k < n; ++k) {
          cpx<float> z0, z1;
          lds2(b6 + A * P2 * k, z0, z1);
          en += static_cast<double>(z0.re) * w0[k].re + static_cast<double>(z0.im) * w0[k].im;
          en += static_cast<double>(z1.re) * w1[k].re + static_cast<double>(z1.im) * w1[k].im;
          sts2(b6 + A * P2 * k, w0[k], w1[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < n; ++k) sts2(b6 + A * P2 * k, w0[k], w1[k]);
      }
    }
  }
  __syncthreads();
  {
    cpx<float> w0[n], w1[n];
#pragma unroll
    for (int k = 0; k < n; ++k) lds2(b5 + A * n * k, w0[k], w1[k]);
    dft<float, n, -1>(w0);
    dft<float, n, -1>(w1);
#pragma unroll
    for (int k = 0; k < n; ++k) sts2(b5 + A * n * k, w0[k], w1[k]);
  }
  __syncthreads();
  {  // dim 4 forward, store
    cpx<float> w0[n], w1[n];
#pragma unroll
    for (int k = 0; k < n; ++k) lds2(al + A * (k + n * qq), w0[k], w1[k]);
    dft<float, n, -1>(w0);
    dft<float, n, -1>(w1);
#pragma unroll
    for (int k = 0; k < n; ++k) st2((long long)S * (k + n * qq), w0[k], w1[k]);
  }
  if (a.partials != nullptr) {
    static_assert(sizeof(cpx<float>) * A * S >= sizeof(double) * THREADS, "LDS reuse for the reduction");
    __syncthreads();  // every thread has read L
    en = block_sum<THREADS>(en, reinterpret_cast<double*>(L));
    if (tid == 0) a.partials[(long long)vidx * nt + tile] = en;
  }
}

// ---------------------------------------------------------------- pass C
template <class T, int n, int BPW, int OP, int NOUT>
__global__ __launch_bounds__(BPW* n* n) void k_pass_c(PassCArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int THREADS = BPW * P2;
  __shared__ cpx<T> lds[BPW * S];
  __shared__ double red[THREADS];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int blk = tid / P2, pp = tid - blk * P2;
  const long long off = ((long long)blockIdx.x * BPW + blk) * S;
  cpx<T>* L = lds + blk * S;
  const int base1 = (pp % n) * n + (pp / n) * P2;
  const long long qoff = qslot(a.qmap, g) * a.ldq + off + base1;

  cpx<T> acc[n];
#pragma unroll 1
  for (int o = 0; o < NOUT; ++o) {
    {  // dim 3 forward, loaded straight from HBM
      const cpx<T>* tg = a.t + (g * NOUT + o) * a.ldt + off + pp;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = tg[P2 * k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[pp + P2 * k] = w[k];
    }
    __syncthreads();
    {
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
    }
    __syncthreads();
    {
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
      dft<T, n, -1>(w);
      if (o == 0) {
#pragma unroll
        for (int k = 0; k < n; ++k) acc[k] = w[k];
      } else {
        const T* wo = (o == 1) ? a.W1 : a.W2;
#pragma unroll
        for (int k = 0; k < n; ++k) acc[k] = acc[k] + wo[qoff + k] * w[k];
      }
    }
    if (o + 1 < NOUT) __syncthreads();
  }
  if constexpr (OP == POST_MHAT) {
#pragma unroll
    for (int k = 0; k < n; ++k) acc[k] = a.invC[qoff + k] * acc[k];
  }
  cpx<T>* yg = a.y + g * a.ldy + off + base1;
#pragma unroll
  for (int k = 0; k < n; ++k) yg[k] = acc[k];

  if (a.partials != nullptr) {
    const cpx<T>* dg = a.dotp + g * a.lddot + off + base1;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const cpx<T> d = dg[k];
      s += static_cast<double>(d.re * acc[k].re + d.im * acc[k].im);
    }
    s = block_sum<THREADS>(s, red);
    if (tid == 0) a.partials[(long long)g * gridDim.x + blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------- pass C, coalesced I/O
// Forward DFT over dims 3 (register pencil straight from HBM), 2, 1 (LDS); the
// post-op reads the transformed block from LDS in linear order, so the weights, the
// output and the p'q dot operand are all accessed coalesced.
template <class T, int n, int BPW, int OP, int NOUT>
__global__ __launch_bounds__(BPW* n* n) void k_pass_c_lin(PassCArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int THREADS = BPW * P2;
  __shared__ cpx<T> lds[BPW * S];
  __shared__ double red[THREADS];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int blk = tid / P2, pp = tid - blk * P2;
  const long long wg0 = (long long)blockIdx.x * BPW * S;
  const long long off = wg0 + (long long)blk * S;
  cpx<T>* L = lds + blk * S;
  const long long qoff = qslot(a.qmap, g) * a.ldq + wg0 + tid;

  cpx<T> acc[n];
#pragma unroll 1
  for (int o = 0; o < NOUT; ++o) {
    {
      const cpx<T>* tg = a.t + (g * NOUT + o) * a.ldt + off + pp;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = tg[P2 * k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[pp + P2 * k] = w[k];
    }
    __syncthreads();
    {
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
    }
    __syncthreads();
    {
      const int base1 = (pp % n) * n + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base1 + k] = w[k];
    }
    __syncthreads();
    if (o == 0) {
#pragma unroll
      for (int i = 0; i < n; ++i) acc[i] = lds[tid + THREADS * i];
    } else {
      const T* wo = (o == 1) ? a.W1 : a.W2;
#pragma unroll
      for (int i = 0; i < n; ++i) acc[i] = acc[i] + wo[qoff + THREADS * i] * lds[tid + THREADS * i];
    }
    if (o + 1 < NOUT) __syncthreads();
  }
  if constexpr (OP == POST_MHAT) {
#pragma unroll
    for (int i = 0; i < n; ++i) acc[i] = a.invC[qoff + THREADS * i] * acc[i];
  }
  cpx<T>* yg = a.y + g * a.ldy + wg0 + tid;
#pragma unroll
  for (int i = 0; i < n; ++i) yg[THREADS * i] = acc[i];
  if (a.partials != nullptr) {
    const cpx<T>* dg = a.dotp + g * a.lddot + wg0 + tid;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) {
      const cpx<T> d = dg[THREADS * i];
      s += static_cast<double>(d.re * acc[i].re + d.im * acc[i].im);
    }
    s = block_sum<THREADS>(s, red);
    if (tid == 0) a.partials[(long long)g * gridDim.x + blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------- Mhat passes, split
// Pass A and pass C of Mhat with the three scalar applies (z, W1 z, W2 z) carried by
// three thread groups of one workgroup (THREADS = 3*BPW*n^2): each group owns an LDS
// region and one output, so a 16 MB grid vector yields three times the waves in
// flight of a one-output workgroup.  All HBM accesses are coalesced (dim-3 register
// pencils, linear LDS-staged stores).
template <class T, int n, int BPW, int OP>
__global__ __launch_bounds__(3 * BPW * n * n) void k_pass_a3(PassAArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int TG = BPW * P2;  // threads per group
  __shared__ cpx<T> lds[3 * BPW * S];
  const int g = blockIdx.y;
  bool use_beta = false;
  T beta = T(0);
  if constexpr (OP == PRE_MHAT_CG) {
    if (!cg_pass_a_state<3 * BPW * n * n>(a, g, reinterpret_cast<double*>(lds), use_beta, beta)) return;
  } else {
    if (skipped(a.skip, g)) return;
  }
  const int tid = threadIdx.x;
  const int o = tid / TG, tl = tid - o * TG;
  const int blk = tl / P2, pp = tl - blk * P2;
  const long long wg0 = (long long)blockIdx.x * BPW * S;
  const long long off = wg0 + (long long)blk * S;
  cpx<T>* R = lds + o * (BPW * S);
  cpx<T>* L = R + blk * S;

  const cpx<T>* xg = a.x + g * a.ldx + off + pp;
  cpx<T> pn[n];
#pragma unroll
  for (int k = 0; k < n; ++k) pn[k] = xg[P2 * k];
  if constexpr (OP == PRE_MHAT_CG) {
    cpx<T>* pg = a.p + g * a.ldx + off + pp;
    if (use_beta) {
#pragma unroll
      for (int k = 0; k < n; ++k) pn[k] = pn[k] + beta * pg[P2 * k];
    }
    __syncthreads();  // every group has read the old p before group 0 overwrites it
    if (o == 0) {
#pragma unroll
      for (int k = 0; k < n; ++k) pg[P2 * k] = pn[k];
    }
  }
  const long long qoff = qslot(a.qmap, g) * a.ldq + off + pp;
  cpx<T> u[n];
#pragma unroll
  for (int k = 0; k < n; ++k) u[k] = mhat_weight(a.invC, a.W1, a.W2, o, qoff + P2 * k) * pn[k];
  dft<T, n, +1>(u);
#pragma unroll
  for (int k = 0; k < n; ++k) L[pp + P2 * k] = u[k];
  __syncthreads();
  {
    const int base2 = (pp % n) + (pp / n) * P2;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
    dft<T, n, +1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
  }
  __syncthreads();
  {
    const int base1 = (pp % n) * n + (pp / n) * P2;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
    dft<T, n, +1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[base1 + k] = w[k];
  }
  __syncthreads();
  cpx<T>* tg = a.t + (g * 3 + o) * a.ldt + wg0;
#pragma unroll
  for (int i = 0; i < n; ++i) tg[tl + TG * i] = R[tl + TG * i];
}

// Mhat pass A as a persistent, software-pipelined kernel (the counterpart of k_pass_c3p):
// the workgroup walks the tiles blockIdx.x, blockIdx.x + gridDim.x, ... of its group; the
// input pencils (x, and p for the CG form) of the next tile are loaded into registers once
// the current tile's first DFT has gone to LDS.  Same arithmetic per mode as k_pass_a3.
template <class T, int n, int BPW, int OP>
__global__ __launch_bounds__(3 * BPW * n * n) void k_pass_a3p(PassAArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int TG = BPW * P2;  // threads per group
  constexpr int E = BPW * S;    // modes per tile
  constexpr int NT = S / BPW;   // tiles per group
  constexpr bool CG = OP == PRE_MHAT_CG;
  typedef T v2 __attribute__((ext_vector_type(2)));
  __shared__ cpx<T> lds[3 * E];
  const int g = blockIdx.y;
  bool use_beta = false;
  T beta = T(0);
  if constexpr (CG) {
    if (!cg_pass_a_state<3 * TG>(a, g, reinterpret_cast<double*>(lds), use_beta, beta)) return;
  } else {
    if (skipped(a.skip, g)) return;
  }
  const int tid = threadIdx.x;
  const int o = tid / TG, tl = tid - o * TG;
  const int blk = tl / P2, pp = tl - blk * P2;
  cpx<T>* R = lds + o * E;
  cpx<T>* L = R + blk * S;
  const long long boff = (long long)blk * S + pp;
  const cpx<T>* xsrc = a.x + g * a.ldx + boff;
  cpx<T>* psrc = CG ? a.p + g * a.ldx + boff : nullptr;
  const long long qg = qslot(a.qmap, g) * a.ldq + boff;
  v2 px[n], pq[n];
  auto fetch = [&](int tile) {
    const long long to = (long long)tile * E;
#pragma unroll
    for (int k = 0; k < n; ++k) px[k] = *reinterpret_cast<const v2*>(xsrc + to + P2 * k);
    if (CG && use_beta) {
#pragma unroll
      for (int k = 0; k < n; ++k) pq[k] = *reinterpret_cast<const v2*>(psrc + to + P2 * k);
    }
  };
  int tile = blockIdx.x;
  fetch(tile);  // the launch has gridDim.x <= NT, so every workgroup owns a tile
  for (; tile < NT; tile += gridDim.x) {
    const long long to = (long long)tile * E;
    T w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = mhat_weight(a.invC, a.W1, a.W2, o, qg + to + P2 * k);
    cpx<T> u[n];
#pragma unroll
    for (int k = 0; k < n; ++k) u[k] = mk(px[k].x, px[k].y);
    if constexpr (CG) {
      if (use_beta) {
#pragma unroll
        for (int k = 0; k < n; ++k) u[k] = u[k] + beta * mk(pq[k].x, pq[k].y);
      }
      __syncthreads();  // every group holds its copy of the old p before group 0 overwrites it
      if (o == 0) {
#pragma unroll
        for (int k = 0; k < n; ++k) psrc[to + P2 * k] = u[k];
      }
    }
#pragma unroll
    for (int k = 0; k < n; ++k) u[k] = w[k] * u[k];
    dft<T, n, +1>(u);
#pragma unroll
    for (int k = 0; k < n; ++k) L[pp + P2 * k] = u[k];
    if (tile + (int)gridDim.x < NT) fetch(tile + gridDim.x);
    __syncthreads();
    {
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> v[n];
#pragma unroll
      for (int k = 0; k < n; ++k) v[k] = L[base2 + n * k];
      dft<T, n, +1>(v);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = v[k];
    }
    __syncthreads();
    {
      const int base1 = (pp % n) * n + (pp / n) * P2;
      cpx<T> v[n];
#pragma unroll
      for (int k = 0; k < n; ++k) v[k] = L[base1 + k];
      dft<T, n, +1>(v);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base1 + k] = v[k];
    }
    __syncthreads();
    cpx<T>* tg = a.t + (g * 3 + o) * a.ldt + to;
#pragma unroll
    for (int i = 0; i < n; ++i) tg[tl + TG * i] = R[tl + TG * i];
    __syncthreads();  // the next tile's first DFT rewrites LDS
  }
}

// Mhat pass A, persistent, with the pre-op shared: the tile's input (x, and p for the CG
// form) is read once, in linear order by all threads (the three thread groups of k_pass_a3
// each read every pencil), combined (p = r + beta p, stored), and staged in a fourth LDS
// region from which every group reads its dim-3 pencils.  The next tile's input is loaded
// into registers while the current tile goes through its DFTs.  Same arithmetic per mode as
// k_pass_a3.
template <class T, int n, int BPW, int OP>
__global__ __launch_bounds__(3 * BPW * n * n) void k_pass_a3s(PassAArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int TG = BPW * P2;  // threads per group
  constexpr int THREADS = 3 * TG;
  constexpr int E = BPW * S;    // modes per tile
  constexpr int NT = S / BPW;   // tiles per group
  constexpr int NE = (E + THREADS - 1) / THREADS;  // linear input modes per thread
  constexpr bool CG = OP == PRE_MHAT_CG;
  typedef T v2 __attribute__((ext_vector_type(2)));
  __shared__ cpx<T> lds[4 * E];
  const int g = blockIdx.y;
  bool use_beta = false;
  T beta = T(0);
  if constexpr (CG) {
    if (!cg_pass_a_state<THREADS>(a, g, reinterpret_cast<double*>(lds), use_beta, beta)) return;
  } else {
    if (skipped(a.skip, g)) return;
  }
  const int tid = threadIdx.x;
  const int o = tid / TG, tl = tid - o * TG;
  const int blk = tl / P2, pp = tl - blk * P2;
  cpx<T>* R = lds + o * E;
  cpx<T>* L = R + blk * S;
  cpx<T>* X = lds + 3 * E;  // the shared pre-op result of the tile
  const cpx<T>* xsrc = a.x + g * a.ldx;
  cpx<T>* pdst = CG ? a.p + g * a.ldx : nullptr;
  const long long qg = qslot(a.qmap, g) * a.ldq + (long long)blk * S + pp;
  v2 px[NE], pq[NE];
  auto fetch = [&](int tile) {
    const long long to = (long long)tile * E;
#pragma unroll
    for (int r = 0; r < NE; ++r) {
      const int e = tid + THREADS * r;
      if (e < E) {
        px[r] = *reinterpret_cast<const v2*>(xsrc + to + e);
        if (CG && use_beta) pq[r] = *reinterpret_cast<const v2*>(pdst + to + e);
      }
    }
  };
  int tile = blockIdx.x;
  fetch(tile);  // the launch has gridDim.x <= NT, so every workgroup owns a tile
  for (; tile < NT; tile += gridDim.x) {
    const long long to = (long long)tile * E;
    T w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = mhat_weight(a.invC, a.W1, a.W2, o, qg + to + P2 * k);
#pragma unroll
    for (int r = 0; r < NE; ++r) {
      const int e = tid + THREADS * r;
      if (e < E) {
        cpx<T> u = mk(px[r].x, px[r].y);
        if constexpr (CG) {
          if (use_beta) u = u + beta * mk(pq[r].x, pq[r].y);
          pdst[to + e] = u;
        }
        X[e] = u;
      }
    }
    if (tile + (int)gridDim.x < NT) fetch(tile + gridDim.x);
    __syncthreads();
    {
      const cpx<T>* Xb = X + blk * S;
      cpx<T> u[n];
#pragma unroll
      for (int k = 0; k < n; ++k) u[k] = w[k] * Xb[pp + P2 * k];
      dft<T, n, +1>(u);
#pragma unroll
      for (int k = 0; k < n; ++k) L[pp + P2 * k] = u[k];
    }
    __syncthreads();
    {
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> v[n];
#pragma unroll
      for (int k = 0; k < n; ++k) v[k] = L[base2 + n * k];
      dft<T, n, +1>(v);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = v[k];
    }
    __syncthreads();
    {
      const int base1 = (pp % n) * n + (pp / n) * P2;
      cpx<T> v[n];
#pragma unroll
      for (int k = 0; k < n; ++k) v[k] = L[base1 + k];
      dft<T, n, +1>(v);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base1 + k] = v[k];
    }
    __syncthreads();
    cpx<T>* tg = a.t + (g * 3 + o) * a.ldt + to;
#pragma unroll
    for (int i = 0; i < n; ++i) tg[tl + TG * i] = R[tl + TG * i];
    // the next tile's pre-op rewrites X only: X was last read before the barrier above
  }
}

// Epilogue of the Mhat pass C over the E modes of one tile, already transformed in LDS
// (component o at lds + o*E): q = invC*(s0 + W1 s1 + W2 s2) in linear mode order.  CG: x +=
// al*p, r -= al*q, s += |r_new|^2; otherwise y = q and s += Re(d^H q) when d is given.
// qg: the group's offset in the diagonals (qslot * ldq), read once per workgroup: a load here
// would make the tile wait for every memory operation issued before it.  (Issuing all of a
// thread's loads before its stores measured slower than this loop.)
template <class T, int E, int THREADS, bool CG>
__device__ __forceinline__ void mhat_c_epilogue(const PassCArgs<T>& a, const cpx<T>* lds, int g, long long qg,
                                                long long wg0, T al, double& s) {
  const int tid = threadIdx.x;
  const long long qbase = qg + wg0;
  if constexpr (CG) {
    const long long xo = g * a.ldx + wg0;
    for (int e = tid; e < E; e += THREADS) {
      const long long q = qbase + e;
      const T w1 = a.W1[q], w2 = a.W2[q];
      const cpx<T> v = inv_c_or(a.invC, w1, w2, q) * (lds[e] + w1 * lds[E + e] + w2 * lds[2 * E + e]);
      a.x[xo + e] = a.x[xo + e] + al * a.p[xo + e];
      const cpx<T> rn = a.r[xo + e] - al * v;
      a.r[xo + e] = rn;
      s += static_cast<double>(rn.re) * rn.re + static_cast<double>(rn.im) * rn.im;
    }
  } else {
    cpx<T>* yg = a.y + g * a.ldy + wg0;
    const cpx<T>* dg = (a.partials && a.dotp) ? a.dotp + g * a.lddot + wg0 : nullptr;
    const cpx<T>* bg = a.resb ? a.resb + g * a.ldy + wg0 : nullptr;
    cpx<T>* zg = a.zw ? a.zw + 2 * g * a.ldy + wg0 : nullptr;
    for (int e = tid; e < E; e += THREADS) {
      const long long q = qbase + e;
      const T w1 = a.W1[q], w2 = a.W2[q];
      const cpx<T> s1 = lds[E + e], s2 = lds[2 * E + e];
      cpx<T> v = inv_c_or(a.invC, w1, w2, q) * (lds[e] + w1 * s1 + w2 * s2);
      if (bg) {  // the residual b - Mhat x and its ||.||^2 partials
        v = bg[e] - v;
        s += static_cast<double>(v.re) * v.re + static_cast<double>(v.im) * v.im;
      }
      yg[e] = v;
      if (zg) {
        zg[e] = s1;
        zg[a.ldy + e] = s2;
      }
      if (dg) {
        const cpx<T> d = dg[e];
        s += static_cast<double>(d.re * v.re + d.im * v.im);
      }
    }
  }
}

// CG = true fuses the conjugate-gradient step of MATLAB pcg into the epilogue: every
// workgroup reduces the pass-B partials of p'q in the same fixed order, so all agree on
// alpha = rho/p'q (and on a breakdown), workgroup 0 records it, and each workgroup updates
// x and r on its own modes; q itself never reaches HBM.
template <class T, int n, int BPW, bool CG = false>
__global__ __launch_bounds__(3 * BPW * n * n) void k_pass_c3(PassCArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int TG = BPW * P2;
  constexpr int THREADS = 3 * TG;
  constexpr int E = BPW * S;  // modes per workgroup
  __shared__ cpx<T> lds[3 * E];
  __shared__ double red[THREADS];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int o = tid / TG, tl = tid - o * TG;
  const int blk = tl / P2, pp = tl - blk * P2;
  const long long wg0 = (long long)blockIdx.x * E;
  const long long off = wg0 + (long long)blk * S;
  cpx<T>* L = lds + o * E + blk * S;
  {
    const cpx<T>* tg = a.t + (g * 3 + o) * a.ldt + off + pp;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = tg[P2 * k];
    dft<T, n, -1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[pp + P2 * k] = w[k];
  }
  __syncthreads();
  {
    const int base2 = (pp % n) + (pp / n) * P2;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
    dft<T, n, -1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
  }
  __syncthreads();
  {
    const int base1 = (pp % n) * n + (pp / n) * P2;
    cpx<T> w[n];
#pragma unroll
    for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
    dft<T, n, -1>(w);
#pragma unroll
    for (int k = 0; k < n; ++k) L[base1 + k] = w[k];
  }
  __syncthreads();
  // combine y = invC*(s0 + W1 s1 + W2 s2) in linear mode order
  const long long qbase = qslot(a.qmap, g) * a.ldq + wg0;
  double s = 0.0;
  if constexpr (CG) {
    double pq = 0.0;
    {
      const double* pp = a.pq_partials + (long long)g * a.npq;
      pq = strided_sum(pq, pp, tid, a.npq, THREADS);
      pq = block_sum<THREADS>(pq, red);
      if (tid == 0) red[0] = pq;
      __syncthreads();
      pq = red[0];
      __syncthreads();  // red is reused below
    }
    const double alpha = a.cg[g].rho / pq;
    const bool ok = (pq > 0.0) && !isinf(pq) && !isinf(alpha) && !isnan(alpha);
    if (blockIdx.x == 0 && tid == 0) {
      a.cg[g].pq = pq;
      if (ok) {
        a.cg[g].alpha = alpha;
      } else {
        a.cg[g].flag = 4;
        a.cg[g].done = 1;
      }
    }
    if (!ok) return;
    const T al = static_cast<T>(alpha);
    const long long xo = g * a.ldx + wg0;
    for (int e = tid; e < E; e += THREADS) {
      const long long q = qbase + e;
      const T w1 = a.W1[q], w2 = a.W2[q];
      const cpx<T> v = inv_c_or(a.invC, w1, w2, q) * (lds[e] + w1 * lds[E + e] + w2 * lds[2 * E + e]);
      a.x[xo + e] = a.x[xo + e] + al * a.p[xo + e];
      const cpx<T> rn = a.r[xo + e] - al * v;
      a.r[xo + e] = rn;
      s += static_cast<double>(rn.re) * rn.re + static_cast<double>(rn.im) * rn.im;
    }
  } else {
    cpx<T>* yg = a.y + g * a.ldy + wg0;
    const cpx<T>* dg = (a.partials && a.dotp) ? a.dotp + g * a.lddot + wg0 : nullptr;
    const cpx<T>* bg = a.resb ? a.resb + g * a.ldy + wg0 : nullptr;
    cpx<T>* zg = a.zw ? a.zw + 2 * g * a.ldy + wg0 : nullptr;
    for (int e = tid; e < E; e += THREADS) {
      const long long q = qbase + e;
      const T w1 = a.W1[q], w2 = a.W2[q];
      const cpx<T> s1 = lds[E + e], s2 = lds[2 * E + e];
      cpx<T> v = inv_c_or(a.invC, w1, w2, q) * (lds[e] + w1 * s1 + w2 * s2);
      if (bg) {  // the residual b - Mhat x and its ||.||^2 partials
        v = bg[e] - v;
        s += static_cast<double>(v.re) * v.re + static_cast<double>(v.im) * v.im;
      }
      yg[e] = v;
      if (zg) {
        zg[e] = s1;
        zg[a.ldy + e] = s2;
      }
      if (dg) {
        const cpx<T> d = dg[e];
        s += static_cast<double>(d.re * v.re + d.im * v.im);
      }
    }
  }
  if (a.partials != nullptr) {
    s = block_sum<THREADS>(s, red);
    if (tid == 0) a.partials[(long long)g * gridDim.x + blockIdx.x] = s;
  }
}

// Mhat pass C as a persistent, software-pipelined kernel: the workgroup walks the tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... of its group, and the dim-3 pencils of the next
// tile are loaded into registers as soon as the current tile's first DFT has gone to LDS,
// so they are in flight through the current tile's LDS phases and epilogue.  With one
// workgroup per tile the load, LDS and store phases of a workgroup run one after the other
// and only other workgroups on the CU can fill the gaps.  Results are those of k_pass_c3
// (same arithmetic per mode); partials are per workgroup (gridDim.x per group).
template <class T, int n, int BPW, bool CG = false>
__global__ __launch_bounds__(3 * BPW * n * n) void k_pass_c3p(PassCArgs<T> a) {
  constexpr int P2 = Grid<n>::P2, S = Grid<n>::S;
  constexpr int TG = BPW * P2;
  constexpr int THREADS = 3 * TG;
  constexpr int E = BPW * S;   // modes per tile
  constexpr int NT = S / BPW;  // tiles per group
  // pencils in flight are clang vectors: a local array of structs copied from memory would be
  // kept in scratch across the loop
  typedef T v2 __attribute__((ext_vector_type(2)));
  __shared__ cpx<T> lds[3 * E];
  __shared__ double red[THREADS];
  const int g = blockIdx.y;
  if (skipped(a.skip, g)) return;
  const int tid = threadIdx.x;
  const int o = tid / TG, tl = tid - o * TG;
  const int blk = tl / P2, pp = tl - blk * P2;
  cpx<T>* L = lds + o * E + blk * S;
  const cpx<T>* tsrc = a.t + (g * 3 + o) * a.ldt + (long long)blk * S + pp;
  v2 pf[n];
  auto fetch = [&](int tile) {
    const cpx<T>* src = tsrc + (long long)tile * E;
#pragma unroll
    for (int k = 0; k < n; ++k) pf[k] = *reinterpret_cast<const v2*>(src + P2 * k);
  };
  int tile = blockIdx.x;
  fetch(tile);  // the launch has gridDim.x <= NT, so every workgroup owns a tile
  T al = T(0);
  if constexpr (CG) {
    double pq = 0.0;
    {
      const double* pqp = a.pq_partials + (long long)g * a.npq;
      pq = strided_sum(pq, pqp, tid, a.npq, THREADS);
      pq = block_sum<THREADS>(pq, red);
      if (tid == 0) red[0] = pq;
      __syncthreads();
      pq = red[0];
      __syncthreads();  // red is reused below
    }
    const double alpha = a.cg[g].rho / pq;
    const bool ok = (pq > 0.0) && !isinf(pq) && !isinf(alpha) && !isnan(alpha);
    if (blockIdx.x == 0 && tid == 0) {
      a.cg[g].pq = pq;
      if (ok) {
        a.cg[g].alpha = alpha;
      } else {
        a.cg[g].flag = 4;
        a.cg[g].done = 1;
      }
    }
    if (!ok) return;
    al = static_cast<T>(alpha);
  }
  const long long qg = qslot(a.qmap, g) * a.ldq;
  double s = 0.0;
  for (; tile < NT; tile += gridDim.x) {
    {
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = mk(pf[k].x, pf[k].y);
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[pp + P2 * k] = w[k];
    }
    if (tile + (int)gridDim.x < NT) fetch(tile + gridDim.x);
    __syncthreads();
    {
      const int base2 = (pp % n) + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base2 + n * k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base2 + n * k] = w[k];
    }
    __syncthreads();
    {
      const int base1 = (pp % n) * n + (pp / n) * P2;
      cpx<T> w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = L[base1 + k];
      dft<T, n, -1>(w);
#pragma unroll
      for (int k = 0; k < n; ++k) L[base1 + k] = w[k];
    }
    __syncthreads();
    mhat_c_epilogue<T, E, THREADS, CG>(a, lds, g, qg, (long long)tile * E, al, s);
    __syncthreads();  // the next tile's first DFT rewrites LDS
  }
  if (a.partials != nullptr) {
    s = block_sum<THREADS>(s, red);
    if (tid == 0) a.partials[(long long)g * gridDim.x + blockIdx.x] = s;
  }
}

}  // namespace fame6d
