// Modewise (diagonal) kernels of the null-space-free q-system, deterministic
// wave64 reductions, the device-side conjugate-gradient scalar recurrences and
// the tall-skinny kernels of the Lanczos basis.
#pragma once

#include "common.hpp"
#include "mass_kernels.hpp"

namespace fame6d {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 1024;
constexpr int kLzMaxP = 256;  // largest Lanczos basis (Krylov dimension p)

__host__ inline int red_blocks(long long n) {
  long long b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
  if (b < 1) b = 1;
  if (b > kMaxRedBlocks) b = kMaxRedBlocks;
  return static_cast<int>(b);
}

// ------------------------------------------------------------------ modewise
// All modewise kernels run over grid.y = groups (Bloch points solved together).
// Group g reads its 2NF/NF vectors at base + g*ld and its diagonals in q-system
// slot qslot(qmap, g) (diagonals of slot s at base + s*n).  V, W, invC are real
// because q_xi = P^T (xi + k_lift) is real for real P, xi and Bloch vector.
#define F6_MODEWISE_LOOP                                                              \
  const int g = blockIdx.y;                                                           \
  if (skipped(skip, g)) return;                                                       \
  const long long qs = qslot(qmap, g) * n;                                            \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;         \
       i += (long long)gridDim.x * blockDim.x)

// bhat = -invC .* (W1 .* mr1 + W2 .* mr2) (make_explicit_krinv_compact.m:34-35) and the start
// of the conjugate-gradient solve on it in one pass: x = 0, r = bhat and the partials of
// ||bhat||^2 in k_cg_start's element order and sums.
// grid (nb, groups) with kRedThreads threads, as k_cg_start.
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_bhat_cg_start(const cpx<T>* mr, long long ldm, long long n,
                                                               const T* W1, const T* W2, const T* invC,
                                                               const int* qmap, cpx<T>* bhat, cpx<T>* x,
                                                               cpx<T>* r, double* partials) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.y;
  const long long go = g * n, qs = qslot(qmap, g) * n;
  double sr = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    const long long j = qs + i;
    const cpx<T> u = (-invC[j]) * (W1[j] * mr[g * ldm + i] + W2[j] * mr[g * ldm + n + i]);
    bhat[go + i] = u;
    x[go + i] = mk(T(0), T(0));
    r[go + i] = u;
    sr += static_cast<double>(u.re) * u.re + static_cast<double>(u.im) * u.im;
  }
  sr = block_sum<kRedThreads>(sr, red);
  if (threadIdx.x == 0) partials[(long long)g * gridDim.x + blockIdx.x] = sr;
}

// y = V1^{-1} [mr1 + s1; mr2 + s2]   (make_explicit_krinv_compact.m:61-64, solveV1_blockdiag.m:7-8)
template <class T>
__global__ void k_v1_solve(const cpx<T>* mr, const cpx<T>* s, long long ld, long long n,
                           const double* V11, const double* V12, const double* V21,
                           const double* V22, cpx<double>* y, long long ldy, Skip skip,
                           const int* qmap) {
  F6_MODEWISE_LOOP {
    const long long j = qs + i;
    const double det = V11[j] * V22[j] - V12[j] * V21[j];
    const cpx<double> m1 = cvt<double>(mr[g * ld + i]) + cvt<double>(s[g * ld + i]);
    const cpx<double> m2 = cvt<double>(mr[g * ld + n + i]) + cvt<double>(s[g * ld + n + i]);
    y[g * ldy + i] = (1.0 / det) * (V22[j] * m1 - V12[j] * m2);
    y[g * ldy + n + i] = (1.0 / det) * (V11[j] * m2 - V21[j] * m1);
  }
}

// y(3NF) = V x(2NF)   (apply_V_blockdiag.m:12-16)
static __global__ void k_apply_v(const cpx<double>* x, long long ldx, long long n, const double* V11,
                          const double* V12, const double* V21, const double* V22,
                          const double* V31, const double* V32, cpx<double>* y, long long ldy,
                          Skip skip, const int* qmap) {
  F6_MODEWISE_LOOP {
    const long long j = qs + i;
    const cpx<double> x1 = x[g * ldx + i], x2 = x[g * ldx + n + i];
    y[g * ldy + i] = V11[j] * x1 + V12[j] * x2;
    y[g * ldy + n + i] = V21[j] * x1 + V22[j] * x2;
    y[g * ldy + 2 * n + i] = V31[j] * x1 + V32[j] * x2;
  }
}

// r = b - y (groups of length n; the true residual of a CG check)
template <class T>
__global__ void k_residual(const cpx<T>* b, const cpx<T>* y, cpx<T>* r, long long ld, long long n,
                           Skip skip) {
  const int g = blockIdx.y;
  if (skipped(skip, g)) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    r[g * ld + i] = b[g * ld + i] - y[g * ld + i];
  }
}

// ------------------------------------------------------------------ reductions
// Block-level partial sums of conj(a).*b over groups of length n: one partial
// per (group, block) for the real part and one for the imaginary part.
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_dot_partials(const cpx<T>* a, long long lda,
                                                              const cpx<T>* b, long long ldb,
                                                              long long n, double* pre,
                                                              double* pim, Skip skip = no_skip()) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.y;
  if (skipped(skip, g)) return;
  const cpx<T>* ag = a + g * lda;
  const cpx<T>* bg = b + g * ldb;
  double sr = 0.0, si = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    const cpx<T> u = ag[i], v = bg[i];
    sr += static_cast<double>(u.re) * v.re + static_cast<double>(u.im) * v.im;
    si += static_cast<double>(u.re) * v.im - static_cast<double>(u.im) * v.re;
  }
  sr = block_sum<kRedThreads>(sr, red);
  __syncthreads();
  si = block_sum<kRedThreads>(si, red);
  if (threadIdx.x == 0) {
    pre[(long long)g * gridDim.x + blockIdx.x] = sr;
    if (pim != nullptr) pim[(long long)g * gridDim.x + blockIdx.x] = si;
  }
}

// Start of a conjugate-gradient solve with x0 = 0: x = 0, r = b and the partials of ||b||^2
// (the same element order and sums as k_dot_partials(b, b)), one pass over b.
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_cg_start(const cpx<T>* b, cpx<T>* x, cpx<T>* r, long long n,
                                                          double* partials) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.y;
  const long long go = g * n;
  double sr = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    const cpx<T> u = b[go + i];
    x[go + i] = mk(T(0), T(0));
    r[go + i] = u;
    sr += static_cast<double>(u.re) * u.re + static_cast<double>(u.im) * u.im;
  }
  sr = block_sum<kRedThreads>(sr, red);
  if (threadIdx.x == 0) partials[(long long)g * gridDim.x + blockIdx.x] = sr;
}

// Fixed-order sum of nb partials per group; out[g] = sum.
static __global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double* partials, int nb,
                                                              double* out) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  double s = 0.0;
  s = strided_sum(s, partials + (long long)g * nb, threadIdx.x, nb, kRedThreads);
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) out[g] = s;
}

// Two sets of jj reductions in one launch (grid 2*jj): out_re[g] from partials_re, out_im[g]
// from partials_im; skipped as a whole with `sk`.
static __global__ __launch_bounds__(kRedThreads) void k_reduce_final_ri(const double* pre, const double* pim,
                                                                 int nb, int jj, double* out_re, double* out_im,
                                                                 Skip sk) {
  __shared__ double red[kRedThreads];
  if (skipped(sk, 0)) return;
  const bool im = blockIdx.x >= (unsigned)jj;
  const int g = im ? blockIdx.x - jj : blockIdx.x;
  const double* partials = im ? pim : pre;
  double s = 0.0;
  s = strided_sum(s, partials + (long long)g * nb, threadIdx.x, nb, kRedThreads);
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) (im ? out_im : out_re)[g] = s;
}

// ------------------------------------------------------------------ CG scalars
// Follows MATLAB pcg without preconditioner and x0 = 0 (call sites
// make_explicit_krinv_compact.m:43): rho = r'r, p = r + (rho/rho_old) p,
// alpha = rho/(p'q), stop on the recurrence residual ||r|| <= tol*||b||
// (the host then verifies with the true residual b - A x).
__device__ __forceinline__ double sum_partials(const double* partials, int nb, double* red) {
  const double s = strided_sum(0.0, partials, threadIdx.x, nb, kRedThreads);
  return block_sum<kRedThreads>(s, red);
}

static __global__ __launch_bounds__(kRedThreads) void k_cg_init(CgState* st, const double* partials,
                                                         int nb, double tol, int maxit, int maxmsteps) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  const double s = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    CgState c{};
    c.n2b = sqrt(s);
    c.tolb = tol * c.n2b;
    c.rho = s;
    c.normr = c.n2b;
    c.iter = 0;
    c.maxit = maxit;
    c.maxmsteps = maxmsteps;
    c.flag = 0;
    c.done = (!(c.n2b > 0.0) || c.normr <= c.tolb || maxit <= 0) ? 1 : 0;
    if (!isfinite(c.n2b)) c.flag = -1;
    if (maxit <= 0 && c.done && c.n2b > 0.0 && c.normr > c.tolb) c.flag = 1;
    st[g] = c;
  }
}

// Preconditioned start: n2b from ||r||^2 partials, rho = r'z from the r'z partials.
static __global__ __launch_bounds__(kRedThreads) void k_pcg_init(CgState* st, const double* prr, int nbr,
                                                                 const double* prz, int nbz, double tol,
                                                                 int maxit, int maxmsteps) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  const double s = sum_partials(prr + (long long)g * nbr, nbr, red);
  __syncthreads();
  const double rz = sum_partials(prz + (long long)g * nbz, nbz, red);
  if (threadIdx.x == 0) {
    CgState c{};
    c.n2b = sqrt(s);
    c.tolb = tol * c.n2b;
    c.rho = rz;
    c.normr = c.n2b;
    c.maxit = maxit;
    c.maxmsteps = maxmsteps;
    c.done = (!(c.n2b > 0.0) || c.normr <= c.tolb || maxit <= 0) ? 1 : 0;
    if (!isfinite(c.n2b) || !isfinite(rz)) c.flag = -1;
    if (maxit <= 0 && c.done && c.n2b > 0.0 && c.normr > c.tolb) c.flag = 1;
    st[g] = c;
  }
}

// Preconditioned step end: normr from ||r||^2 partials, rho_new = r'z from r'z partials.
static __global__ __launch_bounds__(kRedThreads) void k_pcg_beta(CgState* st, const double* prr, int nbr,
                                                                 const double* prz, int nbz) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (st[g].done) return;
  const double rr = sum_partials(prr + (long long)g * nbr, nbr, red);
  __syncthreads();
  const double rho_new = sum_partials(prz + (long long)g * nbz, nbz, red);
  if (threadIdx.x == 0) {
    CgState c = st[g];
    c.iter += 1;
    c.normr = sqrt(rr);
    if (!isfinite(rr) || !isfinite(rho_new)) {
      c.flag = -1;
      c.done = 1;
    } else if (c.normr <= c.tolb || c.check_every) {
      c.flag = 0;
      c.done = 1;
    } else if (c.iter >= c.maxit) {
      c.flag = 1;
      c.done = 1;
    } else {
      const double beta = rho_new / c.rho;
      if (beta == 0.0 || isinf(beta) || rho_new == 0.0) {
        c.flag = 4;
        c.done = 1;
      }
      c.beta = beta;
    }
    c.rho_prev = c.rho;
    c.rho = rho_new;
    st[g] = c;
  }
}

static __global__ __launch_bounds__(kRedThreads) void k_cg_alpha(CgState* st, const double* partials,
                                                          int nb) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (st[g].done) return;
  const double pq = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    st[g].pq = pq;
    const double alpha = st[g].rho / pq;
    if (!(pq > 0.0) || isinf(pq) || isinf(alpha) || isnan(alpha)) {
      st[g].flag = 4;
      st[g].done = 1;
    } else {
      st[g].alpha = alpha;
    }
  }
}

// x += alpha p; r -= alpha q; partials of ||r||^2
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_cg_update(cpx<T>* x, const cpx<T>* p, cpx<T>* r,
                                                           const cpx<T>* q, long long ld,
                                                           long long n, const CgState* st,
                                                           double* partials) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.y;
  if (st[g].done) return;
  const T alpha = static_cast<T>(st[g].alpha);
  cpx<T>* xg = x + g * ld;
  const cpx<T>* pg = p + g * ld;
  cpx<T>* rg = r + g * ld;
  const cpx<T>* qg = q + g * ld;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    xg[i] = xg[i] + alpha * pg[i];
    const cpx<T> rn = rg[i] - alpha * qg[i];
    rg[i] = rn;
    s += static_cast<double>(rn.re) * rn.re + static_cast<double>(rn.im) * rn.im;
  }
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) partials[(long long)g * gridDim.x + blockIdx.x] = s;
}

static __global__ __launch_bounds__(kRedThreads) void k_cg_beta(CgState* st, const double* partials,
                                                         int nb, int* need = nullptr);

// True-residual check of MATLAB pcg, on the device.  k_cg_need_check marks the groups
// that stopped on the recurrence with flag 0 (need[g] = 0 means "check g"; every other
// group gets 1, and groups that stopped with a failure flag are finished here).  The host
// then applies A to x for the marked groups and reduces ||b - A x||^2 into partials, and
// k_cg_true_check decides: converged (finished, flag 0), failed (flag 3 after maxmsteps
// failed checks, flag 1 at maxit) or resume ("moresteps": continue from the true
// residual, checking every later iteration).  resume[g] = 0 marks the groups that resume
// (for the preconditioned form, which needs z = P^{-1} r first); 1 for all others.
static __global__ void k_cg_need_check(CgState* st, int G, int* need) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    CgState& c = st[g];
    int skip = 1;
    if (!c.finished && c.done) {
      if (c.flag == 0) skip = 0;
      else c.finished = 1;
    }
    need[g] = skip;
  }
}

// The pcg step after the last iteration of a chunk; with `need`, k_cg_need_check of the
// same group in the same launch.
static __global__ __launch_bounds__(kRedThreads) void k_cg_beta(CgState* st, const double* partials,
                                                         int nb, int* need) {
  static_assert(kRedThreads == kCgSumThreads, "the pass-A fold sums the partials in the same order");
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (!st[g].done) {  // uniform in the workgroup
    const double rho_new = sum_partials(partials + (long long)g * nb, nb, red);
    if (threadIdx.x == 0) cg_store_step(st + g, cg_beta_step(st + g, rho_new));
  }
  if (need && threadIdx.x == 0) {
    CgState& c = st[g];
    int skip = 1;
    if (!c.finished && c.done) {
      if (c.flag == 0) skip = 0;
      else c.finished = 1;
    }
    need[g] = skip;
  }
}

static __global__ __launch_bounds__(kRedThreads) void k_cg_true_check(CgState* st, const double* partials,
                                                                      int nb, const int* need, int* resume,
                                                                      int precond) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (need[g]) {
    if (threadIdx.x == 0) resume[g] = 1;
    return;
  }
  const double rr = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    CgState c = st[g];
    const double nr = sqrt(rr);
    c.checks += 1;
    c.normr_true = nr;
    int res = 1;
    if (!isfinite(rr)) {
      c.flag = -1;
      c.finished = 1;
    } else if (nr <= c.tolb) {
      c.flag = 0;
      c.finished = 1;
    } else {
      c.moresteps += 1;
      if (c.moresteps >= c.maxmsteps || c.iter >= c.maxit) {
        c.flag = (c.moresteps >= c.maxmsteps) ? 3 : 1;
        c.finished = 1;
      } else {
        // continue from r = b - A x: beta uses the true residual
        c.flag = 0;
        c.done = 0;
        c.check_every = 1;
        c.normr = nr;
        if (precond) {
          res = 0;  // rho = r'z after z = P^{-1} r (k_cg_resume)
        } else {
          c.beta = rr / c.rho_prev;
          c.rho = rr;
        }
      }
    }
    st[g] = c;
    resume[g] = res;
  }
}

// Preconditioned resume: rho = r'z from the partials, beta = rho / rho_prev.
static __global__ __launch_bounds__(kRedThreads) void k_cg_resume(CgState* st, const double* partials,
                                                                  int nb, const int* resume) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (resume[g]) return;
  const double rho_new = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    st[g].beta = rho_new / st[g].rho_prev;
    st[g].rho = rho_new;
  }
}

// A solve whose recurrence stopped with flag 0 is accepted without the true-residual check
// (inner solves of the mixed form, whose residual the FP64 refinement measures itself).
static __global__ void k_cg_accept(CgState* st, int G) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
    if (st[g].done) st[g].finished = 1;
}

// ------------------------------------------------------------------ FP64 refinement (mixed form)
// Per-group state of M u = b solved by FP64 iterative refinement around FP32 inner solves.
struct RefState {
  double nb;      // ||b||
  double nr;      // ||b - M u|| of the current iterate
  int inactive;   // 1: converged (or ||b|| = 0); the group's refinement kernels are skipped
  int nonfinite;  // 1: NaN/Inf met in ||b|| or a residual
  int pad[2];
};
static_assert(sizeof(RefState) == 32, "RefState layout");
__host__ inline Skip skip_inactive(const RefState* rs) {
  return Skip{reinterpret_cast<const int*>(reinterpret_cast<const char*>(rs) + offsetof(RefState, inactive)),
              (int)(sizeof(RefState) / sizeof(int))};
}

static __global__ __launch_bounds__(kRedThreads) void k_ref_init(RefState* rs, const double* partials, int nb) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  const double s = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    RefState r{};
    r.nb = sqrt(s);
    r.nr = r.nb;
    r.nonfinite = isfinite(s) ? 0 : 1;
    r.inactive = (r.nb > 0.0 && !r.nonfinite) ? 0 : 1;
    rs[g] = r;
  }
}

// After a refinement cycle: ||r||^2 partials -> nr; converged when nr <= tol*nb.
static __global__ __launch_bounds__(kRedThreads) void k_ref_check(RefState* rs, const double* partials, int nb,
                                                                  double tol) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.x;
  if (rs[g].inactive) return;
  const double s = sum_partials(partials + (long long)g * nb, nb, red);
  if (threadIdx.x == 0) {
    const double nr = sqrt(s);
    rs[g].nr = nr;
    if (!isfinite(s)) {
      rs[g].nonfinite = 1;
      rs[g].inactive = 1;
    } else if (nr <= tol * rs[g].nb) {
      rs[g].inactive = 1;
    }
  }
}

// y = x / nr for the groups still refining, 0 for the others (grid.y = group).
template <class To, class Ti>
__global__ void k_convert_ref(long long n, const cpx<Ti>* x, long long ldx, const RefState* rs, cpx<To>* y,
                              long long ldy) {
  const int g = blockIdx.y;
  const bool on = !rs[g].inactive;
  const double scale = on ? 1.0 / rs[g].nr : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cpx<Ti> v = on ? x[g * ldx + i] : mk<Ti>(0, 0);
    y[g * ldy + i] = mk(static_cast<To>(scale * v.re), static_cast<To>(scale * v.im));
  }
}

// k_convert_ref into the start of the inner solve on y in one pass: r = x / nr (0 for the
// groups no longer refining), z = 0 and the partials of ||r||^2 in k_cg_start's element order
// and sums.  grid (nb, groups) with kRedThreads threads, as k_cg_start.
template <class To, class Ti>
__global__ __launch_bounds__(kRedThreads) void k_ref_cg_start(long long n, const cpx<Ti>* x, long long ldx,
                                                              const RefState* rs, cpx<To>* z, cpx<To>* r,
                                                              double* partials) {
  __shared__ double red[kRedThreads];
  const int g = blockIdx.y;
  const bool on = !rs[g].inactive;
  const double scale = on ? 1.0 / rs[g].nr : 0.0;
  const long long go = g * n;
  double sr = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    const cpx<Ti> v = on ? x[g * ldx + i] : mk<Ti>(0, 0);
    const cpx<To> u = mk(static_cast<To>(scale * v.re), static_cast<To>(scale * v.im));
    z[go + i] = mk(To(0), To(0));
    r[go + i] = u;
    sr += static_cast<double>(u.re) * u.re + static_cast<double>(u.im) * u.im;
  }
  sr = block_sum<kRedThreads>(sr, red);
  if (threadIdx.x == 0) partials[(long long)g * gridDim.x + blockIdx.x] = sr;
}

// acc += nr * x for the groups still refining (grid.y = group).
template <class Ti>
__global__ void k_accumulate_ref(long long n, const cpx<Ti>* x, long long ldx, const RefState* rs,
                                 cpx<double>* acc, long long lda) {
  const int g = blockIdx.y;
  if (rs[g].inactive) return;
  const double scale = rs[g].nr;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    acc[g * lda + i] = acc[g * lda + i] + scale * cvt<double>(x[g * ldx + i]);
  }
}

// ------------------------------------------------------------------ vectors
template <class T>
__global__ void k_axpby(long long n, cpx<double> a, const cpx<T>* x, cpx<double> b,
                        const cpx<T>* y, cpx<T>* z) {
  // z = a*x + b*y
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    z[i] = cvt<T>(a * cvt<double>(x[i]) + b * cvt<double>(y[i]));
  }
}

template <class To, class Ti>
__global__ void k_convert(long long n, const cpx<Ti>* x, double scale, cpx<To>* y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cpx<Ti> v = x[i];
    y[i] = mk(static_cast<To>(scale * v.re), static_cast<To>(scale * v.im));
  }
}

// acc += scale * x  (x possibly lower precision)
template <class Ti>
__global__ void k_accumulate(long long n, const cpx<Ti>* x, double scale, cpx<double>* acc) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    acc[i] = acc[i] + scale * cvt<double>(x[i]);
  }
}

template <class T>
__global__ void k_fill_zero(long long n, cpx<T>* x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] = mk<T>(0, 0);
  }
}

template <class T>
__global__ void k_fill_zero_real(long long n, T* x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] = T(0);
  }
}

// ------------------------------------------------------------------ Lanczos basis
// Partials of h_k = Q(:,k)' r for all k < jj in one launch.  Block (rb, c) of the
// logical grid covers rows of row block rb for columns [c*KB, c*KB + KB); the logical
// order (column chunk fastest) is mapped XCD-aware, so the chunks of one row block run
// on one XCD and re-read r from its L2.  pre/pim[k*nrb + rb].  A thread takes two rows per
// step and issues all 2(KB+1) loads before the first product: a short last chunk re-reads
// its last column instead of branching around the loads, which would serialise them.
// THREE: r is the three-term update w - alpha q_{jj-1} - beta q_{jj-2} (tt.r = w), formed
// on the fly from columns the sweep reads anyway, and the chunk-0 blocks also write the
// ||r||^2 partials to tt.norm[rb]; r itself is never stored.
struct ThreeTerm {
  const double* alpha;  // device scalar
  double beta;          // 0 on the first step
  double* norm;         // nrb partials of ||r||^2
};
// A basis element read once per sweep: a non-temporal load (the sweeps stream a basis larger
// than the last-level cache, so keeping its lines there only displaces the solve's vectors).
__device__ __forceinline__ cpx<double> load_stream(const cpx<double>* p) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
  return mk(v.x, v.y);
}

// The work of block (rb, ch) (row block, column chunk) of k_gemv_c_all.
template <int KB, bool THREE>
__device__ __forceinline__ void gemv_c_block(const cpx<double>* Q, long long ld, int jj, const cpx<double>* r,
                                             long long n, int nrb, int rb, int ch, double* pre, double* pim,
                                             const ThreeTerm& tt, double* red) {
  constexpr int RPT = 2;
  const int k0 = ch * KB, kcount = min(KB, jj - k0);
  double ar[KB], ai[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) ar[k] = ai[k] = 0.0;
  // every column slot accumulates (a short last chunk re-reads its last column into the
  // slots past kcount, whose sums are never written): a test per slot became a branch into
  // which the compiler sank that slot's load, behind a wait for every earlier load
  auto accumulate = [&](const cpx<double>& v, const cpx<double>* q) {
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      ar[k] += q[k].re * v.re + q[k].im * v.im;
      ai[k] += q[k].re * v.im - q[k].im * v.re;
    }
  };
  const long long stride = (long long)nrb * kRedThreads;
  // three-term form: alpha, beta and the two newest columns (jj >= 2 for beta)
  double alpha = 0.0, nrm = 0.0;
  const cpx<double>* qa = nullptr;
  const cpx<double>* qb = nullptr;
  if constexpr (THREE) {
    alpha = *tt.alpha;
    qa = Q + (long long)(jj - 1) * ld;
    qb = jj >= 2 ? Q + (long long)(jj - 2) * ld : qa;
  }
  // beta = 0 before the second column exists: v - 0 q = v, so the loads stay unconditional
  const double beta = (THREE && jj >= 2) ? tt.beta : 0.0;
  auto row = [&](long long ii) {
    cpx<double> v = r[ii];
    if constexpr (THREE) v = (v - alpha * qa[ii]) - beta * qb[ii];
    return v;
  };
  long long i = rb * (long long)kRedThreads + threadIdx.x;
  for (; i + (RPT - 1) * stride < n; i += RPT * stride) {
    // every load of the step first (the three-term inputs as loaded, not yet combined: an
    // update formed between the loads made the column loads wait for it)
    cpx<double> rv[RPT], av[RPT], bv[RPT], q[RPT][KB];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      rv[u] = r[i + u * stride];
      if constexpr (THREE) {
        av[u] = qa[i + u * stride];
        bv[u] = qb[i + u * stride];
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) q[u][k] = load_stream(Q + (long long)(k0 + min(k, kcount - 1)) * ld + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      cpx<double> v = rv[u];
      if constexpr (THREE) v = (v - alpha * av[u]) - beta * bv[u];
      accumulate(v, q[u]);
      if constexpr (THREE) nrm += v.re * v.re + v.im * v.im;
    }
  }
  if (i < n) {  // at most one row is left: the loop above stops within RPT strides of n
    cpx<double> q[KB];
    const cpx<double> v = row(i);
#pragma unroll
    for (int k = 0; k < KB; ++k) q[k] = load_stream(Q + (long long)(k0 + min(k, kcount - 1)) * ld + i);
    accumulate(v, q);
    if constexpr (THREE) nrm += v.re * v.re + v.im * v.im;
  }
  if constexpr (THREE) {
    if (ch == 0) {  // uniform per block
      const double s = block_sum<kRedThreads>(nrm, red);
      __syncthreads();
      if (threadIdx.x == 0) tt.norm[rb] = s;
    }
  }
  for (int k = 0; k < kcount; ++k) {
    const double s = block_sum<kRedThreads>(ar[k], red);
    __syncthreads();
    const double t = block_sum<kRedThreads>(ai[k], red);
    __syncthreads();
    if (threadIdx.x == 0) {
      pre[(long long)(k0 + k) * nrb + rb] = s;
      pim[(long long)(k0 + k) * nrb + rb] = t;
    }
  }
}

template <int KB, bool THREE = false>
__global__ __launch_bounds__(kRedThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_gemv_c_all(const cpx<double>* Q, long long ld, int jj,
                                                            const cpx<double>* r, long long n, int nrb,
                                                            double* pre, double* pim, Skip sk,
                                                            ThreeTerm tt = ThreeTerm{nullptr, 0.0, nullptr}) {
  __shared__ double red[kRedThreads];
  if (skipped(sk, 0)) return;
  const int nch = (jj + KB - 1) / KB;
  const int logical = xcd_swizzle(blockIdx.x, gridDim.x);
  const int rb = logical / nch, ch = logical - rb * nch;
  gemv_c_block<KB, THREE>(Q, ld, jj, r, n, nrb, rb, ch, pre, pim, tt, red);
}

// r -= Q(:, 0:j) h   (h complex, on device)
static __global__ void k_gemv_n_sub(const cpx<double>* Q, long long ld, int j, const double* hre,
                             const double* him, cpx<double>* r, long long n) {
  __shared__ cpx<double> hs[kLzMaxP];
  for (int k = threadIdx.x; k < j; k += blockDim.x) hs[k] = mk(hre[k], him[k]);
  __syncthreads();
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    cpx<double> acc = r[i];
    for (int k = 0; k < j; ++k) acc = acc - Q[(long long)k * ld + i] * hs[k];
    r[i] = acc;
  }
}

// In-place restart Q(:, 0:kk) = Q(:, 0:m) U, m <= 64, kk <= CB, with Q as the real (2n x m)
// matrix of its interleaved parts: a thread owns one real row, reads all m of its entries
// (eight loads in flight at a time) before it writes any of its kk outputs, so no other
// thread's input is overwritten; U (real, m x kk, column-major) is broadcast from LDS.  The
// sums run over l = 0..m-1 in order.
template <int CB>
__global__ __launch_bounds__(256) void k_restart_rows(double* __restrict__ Q, long long ldq, long long nr, int m,
                                                      int kk, const double* __restrict__ U) {
  __shared__ double Us[64 * CB];  // Us[l * CB + c] = U(l, c), zero past kk
  for (int t = threadIdx.x; t < m * CB; t += 256) {
    const int l = t / CB, c = t - (t / CB) * CB;
    Us[t] = c < kk ? U[(long long)c * m + l] : 0.0;
  }
  __syncthreads();
  const long long row = blockIdx.x * 256LL + threadIdx.x;
  if (row >= nr) return;
  double acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = 0.0;
  int l = 0;
  for (; l + 8 <= m; l += 8) {
    double q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = Q[(long long)(l + u) * ldq + row];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[c] += q[u] * Us[(l + u) * CB + c];
  }
  for (; l < m; ++l) {
    const double q = Q[(long long)l * ldq + row];
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[c] += q * Us[l * CB + c];
  }
#pragma unroll
  for (int c = 0; c < CB; ++c)  // compile-time indices keep acc[] in registers
    if (c < kk) Q[(long long)c * ldq + row] = acc[c];
}

// out(:, c) = Q(:, 0:m) U(:, c) for c < kk, out-of-place: complex Q (ld complex rows) times
// real U (m x kk, column-major), i.e. the real (2n x m) matrix of Q's interleaved parts times
// U.  A thread owns one real row and CB output columns in registers; U is staged through LDS
// in slices of 32 rows, so Q is read ceil(kk / CB) times and U comes from LDS.  Used for the
// thick restart when the basis is too wide for the in-place kernel above, and for the Ritz
// vectors Y = Q U.
template <int CB>
__global__ __launch_bounds__(256) void k_basis_gemm(const double* __restrict__ Q, long long ldq, long long nr, int m,
                                                    int kk, const double* __restrict__ U, double* __restrict__ out,
                                                    long long ldo) {
  __shared__ double Us[32 * CB];
  const long long row = blockIdx.x * 256LL + threadIdx.x;
  const int c0 = blockIdx.y * CB;
  const int kc = min(CB, kk - c0);
  double acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = 0.0;
  for (int l0 = 0; l0 < m; l0 += 32) {
    const int lc = min(32, m - l0);
    __syncthreads();
    for (int t = threadIdx.x; t < 32 * CB; t += 256) {
      const int l = t / CB, c = t - (t / CB) * CB;
      Us[t] = (l < lc && c < kc) ? U[(long long)(c0 + c) * m + l0 + l] : 0.0;
    }
    __syncthreads();
    if (row < nr) {
      int l = 0;
      for (; l + 8 <= lc; l += 8) {  // eight loads in flight, then their products in order
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = Q[(long long)(l0 + l + u) * ldq + row];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int c = 0; c < CB; ++c) acc[c] += q[u] * Us[(l + u) * CB + c];
      }
      for (; l < lc; ++l) {
        const double qv = Q[(long long)(l0 + l) * ldq + row];
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[c] += qv * Us[l * CB + c];
      }
    }
  }
  if (row < nr) {
#pragma unroll
    for (int c = 0; c < CB; ++c)  // compile-time indices keep acc[] in registers
      if (c < kc) out[(long long)(c0 + c) * ldo + row] = acc[c];
  }
}

// out = in - Q(:, 0:j) h (h complex, on device); optional ||out||^2 partials per workgroup.
// With `alpha` set, h_{j-1} += alpha and h_{j-2} += beta (j >= 2): out = in - alpha q_{j-1}
// - beta q_{j-2} - Q h, the three-term update and the sweep in one pass.
// The work of block blk of nblk of k_gemv_n_out.
__device__ __forceinline__ void gemv_n_block(const cpx<double>* Q, long long ld, int j, const double* hre,
                                             const double* him, const cpx<double>* in, cpx<double>* out,
                                             long long n, double* norm_partials, const double* alpha,
                                             double beta, int blk, int nblk, cpx<double>* hs, double* red) {
  for (int k = threadIdx.x; k < j; k += blockDim.x) {
    cpx<double> h = mk(hre[k], him[k]);
    if (alpha != nullptr) {
      if (k == j - 1) h.re += *alpha;
      if (k == j - 2) h.re += beta;
    }
    hs[k] = h;
  }
  __syncthreads();
  double s = 0.0;
  for (long long i = blk * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)nblk * blockDim.x) {
    cpx<double> acc = in[i];
    for (int k = 0; k < j; ++k) acc = acc - load_stream(Q + (long long)k * ld + i) * hs[k];
    out[i] = acc;
    s += acc.re * acc.re + acc.im * acc.im;
  }
  if (norm_partials != nullptr) {
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) norm_partials[blk] = s;
  }
}

static __global__ __launch_bounds__(256) void k_gemv_n_out(const cpx<double>* Q, long long ld, int j,
                                                           const double* hre, const double* him,
                                                           const cpx<double>* in, cpx<double>* out,
                                                           long long n, double* norm_partials, Skip sk,
                                                           const double* alpha = nullptr, double beta = 0.0) {
  __shared__ cpx<double> hs[kLzMaxP];
  __shared__ double red[256];
  if (skipped(sk, 0)) return;
  gemv_n_block(Q, ld, j, hre, him, in, out, n, norm_partials, alpha, beta, blockIdx.x, gridDim.x, hs, red);
}

// Norm bookkeeping of the re-orthogonalisation (one workgroup).  sc[1] = ||r||^2 of the
// current vector, sc[2] = ||r||^2 before the first sweep, the int at sc + 4 = 1 when the
// second sweep is not needed.  stage 0: after the three-term update; stage 1: after the
// first sweep, which decides the second one by the DGKS test ||r_after||^2 < eta2 ||r_before||^2
// (eta2 = 1/2 by default)
// (with `pre`, the stage-0 sum of the same step first); stage 2: after the second sweep (taken
// only when it ran).
__device__ __forceinline__ void lz_dgks_block(double* sc, const double* partials, int nb, int stage,
                                              const double* pre, int npre, double eta2, double* red) {
  int* skip = reinterpret_cast<int*>(sc + 4);
  if (stage == 2 && *skip) return;
  if (pre) {  // stage 0 in the same launch: ||r||^2 before the sweep
    double s0 = strided_sum(0.0, pre, threadIdx.x, npre, kRedThreads);
    s0 = block_sum<kRedThreads>(s0, red);
    if (threadIdx.x == 0) sc[2] = s0;
    __syncthreads();
  }
  double s = strided_sum(0.0, partials, threadIdx.x, nb, kRedThreads);
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) {
    if (stage == 0) {
      sc[2] = s;
    } else {
      sc[1] = s;
      if (stage == 1) skip[0] = (eta2 >= 1.0 || s < eta2 * sc[2]) ? 0 : 1;  // eta >= 1: always
    }
  }
}

static __global__ __launch_bounds__(kRedThreads) void k_lz_dgks(double* sc, const double* partials, int nb,
                                                                int stage, const double* pre = nullptr,
                                                                int npre = 0, double eta2 = 0.5) {
  __shared__ double red[kRedThreads];
  lz_dgks_block(sc, partials, nb, stage, pre, npre, eta2, red);
}

}  // namespace fame6d
