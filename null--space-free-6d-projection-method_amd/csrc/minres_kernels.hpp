// Device side of the preconditioned MINRES of Method A: the shifted operator
// A = K - sigma*M on the full 3NF Fourier space, its spectral preconditioner
// P = K + tau*I applied exactly mode by mode, and the Paige-Saunders recurrences with
// every scalar kept on the device.
//
// Reference: make_A_eigs_top_preconditioned.m:59-60 (A, P), :284-286 (MATLAB minres call),
// build_original_shift_preconditioner_A_only.m:30-56 (closed-form P^{-1}).
#pragma once

#include "common.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

// Per-solve MINRES state (FP64).  Names follow the Paige-Saunders recurrence.
struct MrState {
  double beta1;   // sqrt(r0' P^{-1} r0)
  double beta;    // current beta
  double oldb;    // previous beta
  double alfa;
  double dbar, epsln, oldeps, delta, gbar, gamma, cs, sn, phi, phibar;
  double tolb;    // stop when phibar <= tolb
  int iter;
  int done;       // 1: the iteration kernels skip the solve
  int flag;       // 0 converged (recurrence), 1 maxit, 4 breakdown, -1 NaN
  int maxit;
};
static_assert(sizeof(MrState) == 136, "MrState layout");

__host__ inline Skip skip_mr(const MrState* st) {
  return Skip{reinterpret_cast<const int*>(reinterpret_cast<const char*>(st) + offsetof(MrState, done)), 0};
}

// Curl-curl symbol K_q = |q|^2 I - q q^T applied to mode i of the three components.
__device__ __forceinline__ void kq_apply(double a, double b, double c, cpx<double> x1, cpx<double> x2,
                                         cpx<double> x3, cpx<double>& y1, cpx<double>& y2, cpx<double>& y3) {
  y1 = (b * b + c * c) * x1 - (a * b) * x2 - (a * c) * x3;
  y2 = (a * a + c * c) * x2 - (a * b) * x1 - (b * c) * x3;
  y3 = (a * a + b * b) * x3 - (a * c) * x1 - (b * c) * x2;
}

// (K_q + tau I)^{-1} x = x/(Q+tau) + (1/tau - 1/(Q+tau)) q (q^T x)/Q, Q = |q|^2 (q real).
__device__ __forceinline__ void pinv_apply(double a, double b, double c, double tau, cpx<double> x1,
                                           cpx<double> x2, cpx<double> x3, cpx<double>& y1, cpx<double>& y2,
                                           cpx<double>& y3) {
  const double Q = a * a + b * b + c * c;
  const double it = 1.0 / (Q + tau);
  const double corr = (1.0 / tau - it) / Q;
  const cpx<double> qx = corr * (a * x1 + b * x2 + c * x3);
  y1 = it * x1 + a * qx;
  y2 = it * x2 + b * qx;
  y3 = it * x3 + c * qx;
}

// Shared blocks of the per-mode loops below (n = NF modes, 3 components each).
#define F6_MR_LOOP                                                                     \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;         \
       i += (long long)gridDim.x * blockDim.x)

// out = K x - sigma Mx - coef*sub (sub may be null); the true-residual form uses b.
// With `partials`: Re(v' out) per workgroup (alpha of the recurrence, v = x).
static __global__ __launch_bounds__(kRedThreads) void k_mr_shifted(
    const double* q1, const double* q2, const double* q3, double sigma, const cpx<double>* x,
    const cpx<double>* Mx, const cpx<double>* sub, const MrState* st, int use_ratio, cpx<double>* out,
    const cpx<double>* b, long long n, double* partials) {
  __shared__ double red[kRedThreads];
  if (st && st->done) return;
  const double coef = (st && use_ratio && st->iter >= 1) ? st->beta / st->oldb : 0.0;
  double s = 0.0;
  F6_MR_LOOP {
    cpx<double> k1, k2, k3;
    kq_apply(q1[i], q2[i], q3[i], x[i], x[n + i], x[2 * n + i], k1, k2, k3);
    cpx<double> y1 = k1 - sigma * Mx[i], y2 = k2 - sigma * Mx[n + i], y3 = k3 - sigma * Mx[2 * n + i];
    if (coef != 0.0) {
      y1 = y1 - coef * sub[i];
      y2 = y2 - coef * sub[n + i];
      y3 = y3 - coef * sub[2 * n + i];
    }
    if (b) {  // residual b - A x
      y1 = b[i] - y1;
      y2 = b[n + i] - y2;
      y3 = b[2 * n + i] - y3;
    }
    out[i] = y1;
    out[n + i] = y2;
    out[2 * n + i] = y3;
    if (partials) {
      const cpx<double> v1 = b ? y1 : x[i], v2 = b ? y2 : x[n + i], v3 = b ? y3 : x[2 * n + i];
      s += v1.re * y1.re + v1.im * y1.im + v2.re * y2.re + v2.im * y2.im + v3.re * y3.re + v3.im * y3.im;
    }
  }
  if (partials) {
    s = block_sum<kRedThreads>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

// z = P^{-1} y (and, with `partials`, Re(y' z) per workgroup).  With `st`: first
// y -= (alfa/beta) r2 (the Lanczos step of the recurrence), y updated in place.
static __global__ __launch_bounds__(kRedThreads) void k_mr_precond(
    const double* q1, const double* q2, const double* q3, double tau, cpx<double>* y, const cpx<double>* r2,
    const MrState* st, cpx<double>* z, long long n, double* partials) {
  __shared__ double red[kRedThreads];
  if (st && st->done) return;
  const double c = st ? st->alfa / st->beta : 0.0;
  double s = 0.0;
  F6_MR_LOOP {
    cpx<double> y1 = y[i], y2 = y[n + i], y3 = y[2 * n + i];
    if (st) {
      y1 = y1 - c * r2[i];
      y2 = y2 - c * r2[n + i];
      y3 = y3 - c * r2[2 * n + i];
      y[i] = y1;
      y[n + i] = y2;
      y[2 * n + i] = y3;
    }
    cpx<double> z1, z2, z3;
    pinv_apply(q1[i], q2[i], q3[i], tau, y1, y2, y3, z1, z2, z3);
    z[i] = z1;
    z[n + i] = z2;
    z[2 * n + i] = z3;
    s += y1.re * z1.re + y1.im * z1.im + y2.re * z2.re + y2.im * z2.im + y3.re * z3.re + y3.im * z3.im;
  }
  if (partials) {
    s = block_sum<kRedThreads>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

__device__ __forceinline__ double mr_sum(const double* partials, int nb, double* red) {
  const double s = strided_sum(0.0, partials, threadIdx.x, nb, kRedThreads);
  return block_sum<kRedThreads>(s, red);
}

// Start of a MINRES call: beta1^2 = r0' P^{-1} r0 from the partials; the recurrence stops
// when |phibar| <= tol_abs, or <= tol_rel * beta1 when tol_abs < 0.
static __global__ __launch_bounds__(kRedThreads) void k_mr_init(MrState* st, const double* partials, int nb,
                                                                double tol_rel, double tol_abs, int maxit) {
  __shared__ double red[kRedThreads];
  const double b2 = mr_sum(partials, nb, red);
  if (threadIdx.x == 0) {
    MrState c{};
    c.beta1 = sqrt(fmax(b2, 0.0));
    c.beta = c.beta1;
    c.oldb = 0.0;
    c.phibar = c.beta1;
    c.cs = -1.0;
    c.sn = 0.0;
    c.tolb = tol_abs >= 0.0 ? tol_abs : tol_rel * c.beta1;
    c.maxit = maxit;
    c.flag = isfinite(b2) ? 0 : -1;
    c.done = (!(c.beta1 > 0.0) || !isfinite(b2) || maxit <= 0) ? 1 : 0;
    *st = c;
  }
}

static __global__ __launch_bounds__(kRedThreads) void k_mr_alpha(MrState* st, const double* partials, int nb) {
  __shared__ double red[kRedThreads];
  if (st->done) return;
  const double a = mr_sum(partials, nb, red);
  if (threadIdx.x == 0) st->alfa = a;
}

// beta update, Givens rotation and the stopping test (Paige-Saunders).
static __global__ __launch_bounds__(kRedThreads) void k_mr_givens(MrState* st, const double* partials, int nb) {
  __shared__ double red[kRedThreads];
  if (st->done) return;
  const double b2 = mr_sum(partials, nb, red);
  if (threadIdx.x == 0) {
    MrState c = *st;
    c.oldb = c.beta;
    c.beta = sqrt(fmax(b2, 0.0));
    c.oldeps = c.epsln;
    c.delta = c.cs * c.dbar + c.sn * c.alfa;
    c.gbar = c.sn * c.dbar - c.cs * c.alfa;
    c.epsln = c.sn * c.beta;
    c.dbar = -c.cs * c.beta;
    c.gamma = fmax(sqrt(c.gbar * c.gbar + c.beta * c.beta), 2.220446049250313e-16);
    c.cs = c.gbar / c.gamma;
    c.sn = c.beta / c.gamma;
    c.phi = c.cs * c.phibar;
    c.phibar = c.sn * c.phibar;
    c.iter += 1;
    if (!isfinite(b2) || !isfinite(c.alfa)) {
      c.flag = -1;
      c.done = 1;
    } else if (fabs(c.phibar) <= c.tolb || !(c.beta > 0.0)) {
      c.flag = 0;
      c.done = 1;  // the last x update below still runs (done_pending)
    } else if (c.iter >= c.maxit) {
      c.flag = 1;
      c.done = 1;
    }
    *st = c;
  }
}

// w = (v - oldeps w1 - delta w2)/gamma; x += phi w; v_next = z/beta.  Runs for the
// iteration that set `done` as well (its x update belongs to that iteration), so it tests
// `iter` against the iteration the host issued rather than `done`.
static __global__ void k_mr_update(const MrState* st, int issued_iter, const cpx<double>* v,
                                   const cpx<double>* w1, const cpx<double>* w2, cpx<double>* w, cpx<double>* x,
                                   const cpx<double>* z, cpx<double>* vnext, long long n3) {
  if (st->iter != issued_iter + 1) return;  // this iteration was skipped
  const double ig = 1.0 / st->gamma, oe = st->oldeps, de = st->delta, ph = st->phi;
  const double ib = st->beta > 0.0 ? 1.0 / st->beta : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3;
       i += (long long)gridDim.x * blockDim.x) {
    const cpx<double> wn = ig * (v[i] - oe * w1[i] - de * w2[i]);
    w[i] = wn;
    x[i] = x[i] + ph * wn;
    vnext[i] = ib * z[i];
  }
}

// Continues an ended recurrence: the stopping threshold becomes factor*|phibar| (0: no
// threshold) and the iteration kernels run again from the state the last iteration left.
static __global__ void k_mr_rearm(MrState* st, double factor) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->tolb = fabs(st->phibar) * factor;
    st->flag = 0;
    st->done = 0;
  }
}

// v = z / beta1 at the start of a segment.
static __global__ void k_mr_first_v(const MrState* st, const cpx<double>* z, cpx<double>* v, long long n3) {
  if (st->done) return;
  const double ib = 1.0 / st->beta1;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3;
       i += (long long)gridDim.x * blockDim.x)
    v[i] = ib * z[i];
}

#undef F6_MR_LOOP

}  // namespace fame6d
