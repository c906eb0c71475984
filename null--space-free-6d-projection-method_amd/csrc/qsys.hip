// Null-space-free q-system setup, the fused Mhat operator, the device-side
// PCG on Mhat and the Method-C explicit reduced inverse K_r^{-1}
// (FP64, and FP32 inner solves with FP64 iterative refinement).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

// ------------------------------------------------------------------ q-system
// Restates assemble_K_and_nullspace_free_basis.m:18-105 and
// build_experiment1_q_system_C_compact.m:9-31 for real P (6x3), T = 2*pi.  The per-mode
// quantities are formed on the device (one thread per mode, every Bloch point of the set at
// once); the host forms the lift vectors and checks the per-point minima.

struct QsysBuildArgs {
  double P[18];          // P (6x3, row-major)
  const double* klift;   // 6 per point
  int N, n;
  long long NF, GN;      // modes per point, G * NF
  double* V;             // V11, V12, V21, V22, V31, V32 (GN each)
  double* W;             // W1, W2, invC (GN each)
  float* Wf;             // W in FP32
  double* Qm;            // q1, q2, q3 (GN each)
  double* minqq;         // per workgroup and point: min |q|^2
  double* mindet;        //                         min |V11 V22 - V12 V21|
  int* nsing;            //                         modes on the singular branch
};

constexpr int kQsysThreads = 256, kQsysBlocks = 256;

// One thread per mode of point blockIdx.y: q = P^T (xi + k_lift), the null-space-free basis
// V = |q| [u1 u2] (the regular branch, or the singular one of
// assemble_K_and_nullspace_free_basis.m:75-91 where 3|q|^2 - (q1+q2+q3)^2 <= 1e-13), W and
// invC.  The operations and their order are those of the per-mode formulas in the reference;
// contraction into fused multiply-adds is off, so every value is the correctly rounded
// result of the same IEEE operations on any conforming machine.
__global__ __launch_bounds__(kQsysThreads) void k_qsys_build(QsysBuildArgs a) {
#pragma clang fp contract(off)
  __shared__ double rq[kQsysThreads], rd[kQsysThreads];
  __shared__ int rs[kQsysThreads];
  const int g = blockIdx.y;
  const double* kl = a.klift + 6 * g;
  const long long off = (long long)g * a.NF;
  double minqq = INFINITY, mindet = INFINITY;
  int nsing = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < a.NF;
       i += (long long)gridDim.x * blockDim.x) {
    long long t = i;
    double qv[3] = {0, 0, 0};
    for (int l = 0; l < 6; ++l) {
      const int k = (int)(t % a.n);
      t /= a.n;
      const double xi = (k < a.N) ? k : k - a.n;  // build_fourier_grid_6d.m:12
      const double sx = xi + kl[l];
#pragma unroll
      for (int c = 0; c < 3; ++c) qv[c] += a.P[l * 3 + c] * sx;
    }
    const long long j = off + i;
    a.Qm[j] = qv[0];
    a.Qm[a.GN + j] = qv[1];
    a.Qm[2 * a.GN + j] = qv[2];
    const double Qq = qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2];
    minqq = fmin(minqq, Qq);
    const double sq = sqrt(Qq);
    const double Qs = qv[0] + qv[1] + qv[2];
    const double den = 3 * Qq - Qs * Qs;
    double u1[3], u2[3];
    if (den > 1e-13) {
      const double s1 = 1.0 / sqrt(Qq * den), s2 = 1.0 / sqrt(den);
#pragma unroll
      for (int c = 0; c < 3; ++c) u1[c] = (Qq - qv[c] * Qs) * s1;
      u2[0] = (qv[2] - qv[1]) * s2;
      u2[1] = (qv[0] - qv[2]) * s2;
      u2[2] = (qv[1] - qv[0]) * s2;
    } else {
      // singular branch (assemble_K_and_nullspace_free_basis.m:75-91)
      ++nsing;
      const double h[3] = {qv[0] / sq, qv[1] / sq, qv[2] / sq};
      int imin = 0;
      for (int c = 1; c < 3; ++c)
        if (fabs(h[c]) < fabs(h[imin])) imin = c;
      double av[3] = {0, 0, 0};
      av[imin] = 1;
      const double ha = h[imin];
      double z1[3] = {av[0] - h[0] * ha, av[1] - h[1] * ha, av[2] - h[2] * ha};
      double nz = sqrt(z1[0] * z1[0] + z1[1] * z1[1] + z1[2] * z1[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) z1[c] /= nz;
      const double z2[3] = {h[1] * z1[2] - h[2] * z1[1], h[2] * z1[0] - h[0] * z1[2], h[0] * z1[1] - h[1] * z1[0]};
      nz = sqrt(z2[0] * z2[0] + z2[1] * z2[1] + z2[2] * z2[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        u1[c] = z1[c];
        u2[c] = z2[c] / nz;
      }
    }
    const double v11 = u1[0] * sq, v21 = u1[1] * sq, v31 = u1[2] * sq;
    const double v12 = u2[0] * sq, v22 = u2[1] * sq, v32 = u2[2] * sq;
    a.V[j] = v11;
    a.V[a.GN + j] = v12;
    a.V[2 * a.GN + j] = v21;
    a.V[3 * a.GN + j] = v22;
    a.V[4 * a.GN + j] = v31;
    a.V[5 * a.GN + j] = v32;
    const double d1 = v11 * v22 - v12 * v21;
    mindet = fmin(mindet, fabs(d1));
    const double w1 = (v22 * v31 - v21 * v32) / d1;
    const double w2 = (v11 * v32 - v12 * v31) / d1;
    const double korth = 1.0 + w1 * w1 + w2 * w2;
    const double ic = 1.0 / sqrt(korth);
    a.W[j] = w1;
    a.W[a.GN + j] = w2;
    a.W[2 * a.GN + j] = ic;
    a.Wf[j] = (float)w1;
    a.Wf[a.GN + j] = (float)w2;
    a.Wf[2 * a.GN + j] = (float)ic;
  }
  const int tid = threadIdx.x;
  rq[tid] = minqq;
  rd[tid] = mindet;
  rs[tid] = nsing;
  __syncthreads();
  for (int h = kQsysThreads / 2; h > 0; h >>= 1) {
    if (tid < h) {
      rq[tid] = fmin(rq[tid], rq[tid + h]);
      rd[tid] = fmin(rd[tid], rd[tid + h]);
      rs[tid] += rs[tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const long long b = (long long)g * gridDim.x + blockIdx.x;
    a.minqq[b] = rq[0];
    a.mindet[b] = rd[0];
    a.nsing[b] = rs[0];
  }
}

// Per point: the minima and the count over the workgroups of k_qsys_build (one workgroup
// per point), into out[3 g .. 3 g + 2] (the count as a double).
__global__ __launch_bounds__(kQsysThreads) void k_qsys_minima(const double* minqq, const double* mindet,
                                                              const int* nsing, int nb, double* out) {
  __shared__ double rq[kQsysThreads], rd[kQsysThreads];
  __shared__ int rs[kQsysThreads];
  const int g = blockIdx.x, tid = threadIdx.x;
  double q = INFINITY, d = INFINITY;
  int c = 0;
  for (int b = tid; b < nb; b += kQsysThreads) {
    q = fmin(q, minqq[(long long)g * nb + b]);
    d = fmin(d, mindet[(long long)g * nb + b]);
    c += nsing[(long long)g * nb + b];
  }
  rq[tid] = q;
  rd[tid] = d;
  rs[tid] = c;
  __syncthreads();
  for (int h = kQsysThreads / 2; h > 0; h >>= 1) {
    if (tid < h) {
      rq[tid] = fmin(rq[tid], rq[tid + h]);
      rd[tid] = fmin(rd[tid], rd[tid + h]);
      rs[tid] += rs[tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[3 * g] = rq[0];
    out[3 * g + 1] = rd[0];
    out[3 * g + 2] = (double)rs[0];
  }
}
fame6d_qsys* qsys_build(fame6d_mass* m, int G, const double* q, const double* P) {
  if (G <= 0) throw Error(FAME6D_EARG, "need at least one Bloch point");
  const int N = m->N, n = m->n;
  const long long NF = m->NF;
  auto* S = new fame6d_qsys();
  S->m = m;
  S->G = G;
  S->NF = NF;
  std::memcpy(S->P, P, sizeof(double) * 18);
  S->q.assign(q, q + 3 * G);
  S->klift.assign(6 * G, 0.0);
  S->min_abs_det.assign(G, 0.0);
  S->min_qq.assign(G, 0.0);
  S->n_singular.assign(G, 0);

  // P^T P (3x3) and its inverse (k_lift = P (P^T P)^{-1} q)
  double A[9] = {0};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int l = 0; l < 6; ++l) A[r * 3 + c] += P[l * 3 + r] * P[l * 3 + c];
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                     A[2] * (A[3] * A[7] - A[4] * A[6]);
  if (!(std::fabs(det) > 0)) throw Error(FAME6D_EARG, "P^T P is singular");
  double Ai[9];
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) / det;
  Ai[1] = (A[2] * A[7] - A[1] * A[8]) / det;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) / det;
  Ai[3] = (A[5] * A[6] - A[3] * A[8]) / det;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) / det;
  Ai[5] = (A[2] * A[3] - A[0] * A[5]) / det;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) / det;
  Ai[7] = (A[1] * A[6] - A[0] * A[7]) / det;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) / det;

  const size_t GN = (size_t)G * NF;
  for (int g = 0; g < G; ++g) {
    const double* qq = q + 3 * g;
    if (!(std::isfinite(qq[0]) && std::isfinite(qq[1]) && std::isfinite(qq[2]))) {
      delete S;
      throw Error(FAME6D_ENONFINITE, "Bloch vector contains NaN/Inf");
    }
    double y[3];
    for (int r = 0; r < 3; ++r) y[r] = Ai[r * 3 + 0] * qq[0] + Ai[r * 3 + 1] * qq[1] + Ai[r * 3 + 2] * qq[2];
    double* kl = &S->klift[6 * g];
    for (int l = 0; l < 6; ++l) kl[l] = P[l * 3 + 0] * y[0] + P[l * 3 + 1] * y[1] + P[l * 3 + 2] * y[2];
  }
  S->Vd.ensure(sizeof(double) * 6 * GN);
  S->Wd.ensure(sizeof(double) * 3 * GN);
  S->Wf.ensure(sizeof(float) * 3 * GN);
  S->Qd.ensure(sizeof(double) * 3 * GN);
  const int nb = kQsysBlocks;
  DevBuf tmp;  // k_lift, then the per-workgroup minima and counts, then the per-point results
  const size_t o_kl = 0, o_q = o_kl + 6 * G, o_d = o_q + (size_t)nb * G, o_s = o_d + (size_t)nb * G;
  const size_t o_out = o_s + ((size_t)nb * G + 1) / 2;  // the int counts take half as many doubles
  tmp.ensure(sizeof(double) * (o_out + 3 * G));
  double* td = tmp.as<double>();
  h2d_async(m, td + o_kl, S->klift.data(), sizeof(double) * 6 * G);
  QsysBuildArgs a{};
  std::memcpy(a.P, P, sizeof(double) * 18);
  a.klift = td + o_kl;
  a.N = N;
  a.n = n;
  a.NF = NF;
  a.GN = (long long)GN;
  a.V = S->Vd.as<double>();
  a.W = S->Wd.as<double>();
  a.Wf = S->Wf.as<float>();
  a.Qm = S->Qd.as<double>();
  a.minqq = td + o_q;
  a.mindet = td + o_d;
  a.nsing = reinterpret_cast<int*>(td + o_s);
  k_qsys_build<<<dim3(nb, G), kQsysThreads, 0, m->stream>>>(a);
  k_qsys_minima<<<G, kQsysThreads, 0, m->stream>>>(a.minqq, a.mindet, a.nsing, nb, td + o_out);
  F6_CHECK_LAUNCH();
  std::vector<double> mins(3 * G);
  d2h_sync(m, mins.data(), td + o_out, sizeof(double) * 3 * G);
  for (int g = 0; g < G; ++g) {
    const double minqq = mins[3 * g];
    if (minqq <= 1e-13) {
      delete S;
      throw Error(FAME6D_EGAMMA,
                  "A zero or nearly zero projected Fourier mode was detected; the null-space-free "
                  "reduction requires Gamma-point regularization (min |q|^2 = " +
                      std::to_string(minqq) + ")");
    }
    S->min_qq[g] = minqq;
    S->min_abs_det[g] = mins[3 * g + 1];
    S->n_singular[g] = (int)mins[3 * g + 2];
  }
  F6_HIP(hipStreamSynchronize(m->stream));
  return S;
}

// ------------------------------------------------------------------ helpers
template <class T>
static const T* Wt(const fame6d_qsys* S, int i) {
  if constexpr (std::is_same<T, double>::value) return S->W64(i);
  else return S->W32(i);
}

template <class T>
void mhat_apply(fame6d_qsys* S, int nact, const int* qmap, const cpx<T>* x, cpx<T>* y, Skip skip,
                hipStream_t s, bool inverse_eps, double* energy_partials, const cpx<T>* resb,
                double* res_partials, cpx<T>* zw) {
  fame6d_mass* m = S->m;
  const long long NF = m->NF;
  DevBuf& w = work_buf<T>(m);
  w.ensure(sizeof(cpx<T>) * (size_t)NF * 3 * nact);
  cpx<T>* t = w.as<cpx<T>>();
  PassAArgs<T> a{};
  a.x = x;
  a.ldx = NF;
  a.t = t;
  a.ldt = NF;
  a.W1 = Wt<T>(S, 0);
  a.W2 = Wt<T>(S, 1);
  a.invC = Wt<T>(S, 2);
  a.ldq = NF;
  a.skip = skip;
  a.qmap = qmap;
  launch_pass_a<T>(m, PRE_MHAT, nact, a, s);
  launch_pass_b<T>(m, 3, 0, nact, t, NF, skip, s, inverse_eps, energy_partials);
  PassCArgs<T> c{};
  c.t = t;
  c.ldt = NF;
  c.y = y;
  c.ldy = NF;
  c.W1 = a.W1;
  c.W2 = a.W2;
  c.invC = a.invC;
  c.ldq = NF;
  c.skip = skip;
  c.qmap = qmap;
  c.resb = resb;
  c.partials = resb ? res_partials : nullptr;
  c.zw = zw;
  launch_pass_c<T>(m, POST_MHAT, nact, c, s);
}

template <class T>
void dot_groups(const fame6d_mass* m, int G, const cpx<T>* a, long long lda, const cpx<T>* b,
                long long ldb, long long n, double* out_re_host, double* out_im_host) {
  fame6d_mass* mm = const_cast<fame6d_mass*>(m);
  const int nb = red_blocks(n);
  mm->partials.ensure(sizeof(double) * (size_t)nb * G * 2);
  mm->scal.ensure(sizeof(double) * (size_t)G * 2 + 64);
  double* pre = mm->partials.as<double>();
  double* pim = pre + (size_t)nb * G;
  double* out = mm->scal.as<double>();
  k_dot_partials<T><<<dim3(nb, G), kRedThreads, 0, m->stream>>>(a, lda, b, ldb, n, pre, pim);
  F6_CHECK_LAUNCH();
  k_reduce_final<<<G, kRedThreads, 0, m->stream>>>(pre, nb, out);
  k_reduce_final<<<G, kRedThreads, 0, m->stream>>>(pim, nb, out + G);
  F6_CHECK_LAUNCH();
  std::vector<double> h(2 * G);
  d2h_sync(mm, h.data(), out, sizeof(double) * 2 * G);
  for (int g = 0; g < G; ++g) {
    if (out_re_host) out_re_host[g] = h[g];
    if (out_im_host) out_im_host[g] = h[G + g];
  }
}

// ------------------------------------------------------------------ PCG on Mhat
// MATLAB pcg (x0 = 0) for nact independent right-hand sides, each kept in lock-step
// but with its own scalars; finished solves are skipped on the device.  Semantics
// restated: convergence when the recurrence residual meets tol*||b||, confirmed by
// the true residual b - Mhat x; a failed check continues from the true residual and
// checks every later iteration ("moresteps", at most 5); flag 1 at maxit, 4 on
// breakdown (p'q <= 0 or beta 0/Inf).  Not restated: the stagnation test
// ||p||*|alpha| < eps*||x|| and the minimum-residual iterate returned on failure.
// The whole control flow runs on the device (k_cg_* kernels); the host launches a
// chunk of iterations followed by the check kernels and reads the states once per
// chunk.  p'q (and r'z) come from pass B as sum eps/NF |G u|^2 (PassBArgs::partials).
// With `precond` the iteration is preconditioned by P^{-1} = Mhat evaluated with
// 1/eps (exact inverse for constant eps); the reference's solves are unpreconditioned,
// and `precond` is used only for the FP32 inner solves of the mixed form.
// `true_check` false accepts the recurrence residual without the true-residual check (for
// the inner solves of the mixed form).  `hint` is the first chunk length, updated to the
// iterations this call needed plus one.
// `zw` (optional, with true_check): every true-residual apply also stores M_s(W1 z), M_s(W2 z)
// of the checked x (z = invC x) there, so a solve that ends on a passing check (flag 0) leaves
// them for its final x.
template <class T>
struct PcgRun {
  fame6d_qsys* S;
  int nact;
  const int* qmap_dev;
  bool precond, true_check;
  int maxit;
  cpx<T>* zw;
  fame6d_mass* m;
  hipStream_t s;
  long long NF;
  int nbE, nbC, nbD;
  CgState* st0;
  int cur = 0;  // slot of the current state (the unpreconditioned loop alternates two)
  double *part, *part2;
  int *need, *resume;
  cpx<T> *r, *z;
  PassAArgs<T> pa{};
  PassCArgs<T> pc{};
  double bytesA, bytesB, bytesC;
  int timer_first = 0, issued = 0, rounds = 0;
  int owner;  // timer tag owner of this run's launches

  PcgRun(fame6d_qsys* S_, int nact_, const int* qmap_, cpx<T>* x, cpx<T>* r_, cpx<T>* p, cpx<T>* q, cpx<T>* z_,
         bool precond_, int maxit_, bool true_check_, cpx<T>* zw_, int set = 0)
      : S(S_), nact(nact_), qmap_dev(qmap_), precond(precond_), true_check(true_check_), maxit(maxit_), zw(zw_),
        m(S_->m), s(S_->m->stream), NF(S_->m->NF), r(r_), z(z_) {
    owner = m->timer.next_owner++ & 0x3fff;
    constexpr bool f32 = std::is_same<T, float>::value;
    nbE = 3 * pass_b_blocks(m, f32, nact);  // pass-B energy partials per group
    nbC = pass_c_blocks(m, f32, nact);       // pass-C ||r||^2 partials per group
    nbD = red_blocks(NF);
    S->st_set(set).ensure(sizeof(CgState) * nact * 2);
    S->part_set(set).ensure(sizeof(double) * (size_t)nbE * nact);
    S->part2_set(set).ensure(sizeof(double) * ((size_t)std::max(nbC, nbD) * nact + 8));
    S->ints_set(set).ensure(sizeof(int) * (size_t)nact * 2 + 64);
    st0 = S->st_set(set).as<CgState>();
    part = S->part_set(set).as<double>();
    part2 = S->part2_set(set).as<double>();
    need = S->ints_set(set).as<int>();
    resume = need + nact;
    DevBuf& w = work_buf<T>(m);
    w.ensure(sizeof(cpx<T>) * (size_t)NF * 3 * nact);
    cpx<T>* t = w.as<cpx<T>>();
    CgState* st = slot(0);
    pa.x = precond ? z : r;  // the vector the search direction is built from
    pa.ldx = NF;
    pa.t = t;
    pa.ldt = NF;
    pa.W1 = Wt<T>(S, 0);
    pa.W2 = Wt<T>(S, 1);
    pa.invC = Wt<T>(S, 2);
    pa.ldq = NF;
    pa.p = p;
    pa.st = st;
    pa.skip = skip_done(st);
    pa.qmap = qmap_dev;
    pc.t = t;
    pc.ldt = NF;
    pc.y = q;
    pc.ldy = NF;
    pc.W1 = pa.W1;
    pc.W2 = pa.W2;
    pc.invC = pa.invC;
    pc.ldq = NF;
    pc.skip = skip_done(st);
    pc.qmap = qmap_dev;
    pc.partials = part2;  // ||r_new||^2
    pc.cg = st;
    pc.pq_partials = part;
    pc.npq = nbE;
    pc.x = x;
    pc.p = p;
    pc.r = r;
    pc.ldx = NF;
    // algorithmic HBM bytes per launch (cs: complex element, rs: real element)
    const double cs = sizeof(cpx<T>), rs = sizeof(T), gb = (double)NF * nact;
    bytesA = gb * (6 * cs + 3 * rs);  // zr, p, invC, W1, W2 in; p, 3 vectors out
    bytesB = gb * (6 * cs + rs);      // 3 vectors + eps in; 3 vectors out
    // 3 vectors, W1, W2 (and in FP32 invC; FP64 recomputes it from W1, W2), x, p, r in; x, r out
    bytesC = gb * (8 * cs + (f32 ? 3 : 2) * rs);
  }
  CgState* slot(int k) const { return st0 + (size_t)k * nact; }
  int tag() const { return owner * KernelTimer::kTagSpan + std::min(issued, KernelTimer::kTagSpan - 1); }
  CgState* state() const { return slot(cur); }

  // x = 0, r = b, ||b||, the initial states (MATLAB pcg with x0 = 0); with `mr` (FP64 K_r^{-1}),
  // b = bhat is formed from mr in the same pass and stored at b
  // with `rs` (the inner solves of the mixed form), b = res / ||res|| is formed from the FP64
  // refinement residual `res` in the same pass (b itself is not stored)
  void start(const cpx<T>* b, double tol, const cpx<T>* mr = nullptr, const RefState* rs = nullptr,
             const cpx<double>* res = nullptr) {
    const int maxmsteps = (int)std::min<long long>({NF / 50, 5, std::max<long long>(NF - maxit, 0)});
    CgState* st = slot(0);
    if (rs)
      k_ref_cg_start<T, double><<<dim3(nbD, nact), kRedThreads, 0, s>>>(NF, res, NF, rs, pc.x, r, part2);
    else if (mr)
      k_bhat_cg_start<T><<<dim3(nbD, nact), kRedThreads, 0, s>>>(mr, 2 * NF, NF, pa.W1, pa.W2, pa.invC, qmap_dev,
                                                                 const_cast<cpx<T>*>(b), pc.x, r, part2);
    else
      k_cg_start<T><<<dim3(nbD, nact), kRedThreads, 0, s>>>(b, pc.x, r, NF, part2);
    if (precond) {
      mhat_apply<T>(S, nact, qmap_dev, r, z, no_skip(), s, true, part);
      k_pcg_init<<<nact, kRedThreads, 0, s>>>(st, part2, nbD, part, nbE, tol, maxit, maxmsteps);
    } else {
      k_cg_init<<<nact, kRedThreads, 0, s>>>(st, part2, nbD, tol, maxit, maxmsteps);
    }
    F6_CHECK_LAUNCH();
    timer_first = m->timer.pending;
  }

  // `chunk` iterations, the pcg step of the last, then the true-residual check of the groups
  // that stopped on the recurrence (or their acceptance without it)
  void issue(int chunk) {
    if (++rounds > maxit + 8) throw Error(FAME6D_ENOCONV, "Mhat CG control loop did not terminate");
    CgState* st = state();
    for (int c = 0; c < chunk; ++c, ++issued) {
      // from the second iteration of a chunk on, pass A applies the pcg step of the previous
      // iteration (reading slot cur, writing slot cur^1); preconditioned, with the r'z
      // partials of the previous preconditioner apply
      const bool fold = c > 0;
      pa.st = slot(cur);
      pa.skip = skip_done(slot(cur));
      pa.st_next = fold ? slot(cur ^ 1) : nullptr;
      pa.rr = part2;
      pa.nrr = nbC;
      pa.rz = precond ? part : nullptr;
      pa.nrz = nbE;
      timer_begin(m, TS_MHAT_PASS_A);
      launch_pass_a<T>(m, PRE_MHAT_CG, nact, pa, s);
      timer_end(m, bytesA, tag());
      if (fold) cur ^= 1;
      st = state();
      pc.cg = st;
      pc.skip = skip_done(st);
      timer_begin(m, TS_MHAT_PASS_B);
      launch_pass_b<T>(m, 3, 0, nact, pa.t, NF, skip_done(st), s, false, part);
      timer_end(m, bytesB, tag());
      timer_begin(m, TS_MHAT_PASS_C);
      launch_pass_c<T>(m, POST_MHAT_CG, nact, pc, s);  // alpha, x += alpha p, r -= alpha q
      timer_end(m, bytesC, tag());
      if (precond) mhat_apply<T>(S, nact, qmap_dev, r, z, skip_done(st), s, true, part);
      F6_CHECK_LAUNCH();
    }
    // the pcg step of the chunk's last iteration (a group that stopped earlier is skipped)
    if (precond) k_pcg_beta<<<nact, kRedThreads, 0, s>>>(st, part2, nbC, part, nbE);
    else k_cg_beta<<<nact, kRedThreads, 0, s>>>(st, part2, nbC, true_check ? need : nullptr);
    if (true_check) {
      // true residual r = b - Mhat x of the groups that stopped on the recurrence (marked in
      // the beta launch when unpreconditioned)
      if (precond) k_cg_need_check<<<1, 64, 0, s>>>(st, nact, need);
      const Skip sk{need, 1};
      // r = b - Mhat x and its ||r||^2 partials in the last pass of the apply
      mhat_apply<T>(S, nact, qmap_dev, pc.x, r, sk, s, false, nullptr, bcheck, part2, zw);
      k_cg_true_check<<<nact, kRedThreads, 0, s>>>(st, part2, nbC, need, resume, precond ? 1 : 0);
      if (precond) {
        // z = P^{-1} r_true and rho = r'z for the groups that resume
        mhat_apply<T>(S, nact, qmap_dev, r, z, Skip{resume, 1}, s, true, part);
        k_cg_resume<<<nact, kRedThreads, 0, s>>>(st, part, nbE, resume);
      }
    } else {
      k_cg_accept<<<1, 64, 0, s>>>(st, nact);
    }
    F6_CHECK_LAUNCH();
  }
  const cpx<T>* bcheck = nullptr;  // right-hand side of the true-residual check

  // length of the next chunk after a read that found unfinished groups
  static int next_chunk(const std::vector<CgState>& hs) {
    for (const CgState& c : hs)
      if (!c.finished && c.check_every) return 1;
    return 2;
  }
  static bool all_finished(const std::vector<CgState>& hs) {
    for (const CgState& c : hs)
      if (!c.finished) return false;
    return true;
  }
  // reads the current states (one synchronisation); true when every group has finished
  bool read(std::vector<CgState>& hs) {
    hs.resize(nact);
    d2h_sync(m, hs.data(), state(), sizeof(CgState) * nact);
    return all_finished(hs);
  }
  void collect(const std::vector<CgState>& hs, std::vector<PcgOut>& out, int& hint_ref) {
    out.assign(nact, PcgOut{});
    int max_iter = 0;
    for (int g = 0; g < nact; ++g) {
      const CgState& c = hs[g];
      out[g].flag = c.flag;
      out[g].iter = c.iter;
      out[g].checks = c.checks;
      const double nr = c.checks > 0 ? c.normr_true : c.normr;
      out[g].relres = c.n2b > 0 ? nr / c.n2b : (std::isnan(c.n2b) ? c.n2b : 0.0);
      max_iter = std::max(max_iter, c.iter);
    }
    // iterations issued after the last group finished ran as skipped launches
    m->timer.discard_tags(timer_first, owner, max_iter);
    hint_ref = std::max(2, max_iter + 1);
  }
  // issues chunks until every group has finished, starting from the host copy hs
  void finish(std::vector<CgState>& hs, int chunk) {
    while (!all_finished(hs)) {
      issue(chunk);
      read(hs);
      chunk = next_chunk(hs);
    }
  }
};

template <class T>
static void pcg_mhat(fame6d_qsys* S, int nact, const int* qmap_dev, const cpx<T>* b, cpx<T>* x,
                     cpx<T>* r, cpx<T>* p, cpx<T>* q, cpx<T>* z, bool precond, double tol, int maxit,
                     bool true_check, int& hint_ref, std::vector<PcgOut>& out, cpx<T>* zw = nullptr,
                     const cpx<T>* mr = nullptr) {
  PcgRun<T> R(S, nact, qmap_dev, x, r, p, q, z, precond, maxit, true_check, zw);
  R.bcheck = b;
  R.start(b, tol, mr);
  std::vector<CgState> hs;
  R.issue(hint_ref > 0 ? hint_ref : 8);
  R.read(hs);
  R.finish(hs, PcgRun<T>::next_chunk(hs));
  R.collect(hs, out, hint_ref);
}

// ------------------------------------------------------------------ K_r^{-1}
struct KrinvWs64 {
  cpx<double>*r, *mr, *s, *bhat, *u, *cr, *cp, *cq, *res, *y;
};

// page-locked host copy of one CgState (allocated on first use, freed with the q-system)
static CgState* cg_stage(fame6d_qsys* S) {
  if (!S->cg_pin) F6_HIP(hipHostMalloc(&S->cg_pin, sizeof(CgState), hipHostMallocDefault));
  return static_cast<CgState*>(S->cg_pin);
}

// page-locked host copies of the inner solves' states of one mixed K_r^{-1}: one CgState per
// group and refinement cycle (allocated on first use, freed with the q-system)
static CgState* inner_stage(fame6d_qsys* S, int nact) {
  const size_t need = sizeof(CgState) * (size_t)nact * std::max(1, S->mixed_max_cycles);
  if (S->inner_pin_bytes < need) {
    if (S->inner_pin) F6_HIP(hipHostFree(S->inner_pin));
    S->inner_pin = nullptr;
    S->inner_pin_bytes = 0;
    F6_HIP(hipHostMalloc(&S->inner_pin, need, hipHostMallocDefault));
    S->inner_pin_bytes = need;
  }
  return static_cast<CgState*>(S->inner_pin);
}

static KrinvWs64 krinv_ws(fame6d_qsys* S, int nact, int set = 0) {
  const size_t gn = (size_t)S->m->NF * nact;
  // FP64 workspace: r, mr, s (2NF each) + bhat, u, cr, cp, cq, res, y (NF each)
  S->kr64_set(set).ensure(sizeof(cpx<double>) * gn * 13);
  cpx<double>* base = S->kr64_set(set).as<cpx<double>>();
  return KrinvWs64{base,          base + 2 * gn, base + 4 * gn,  base + 6 * gn,  base + 7 * gn,
                   base + 8 * gn, base + 9 * gn, base + 10 * gn, base + 11 * gn, base + 12 * gn};
}

// Pass-A arguments of the modewise steps around the Mhat solve (V1 solves, W products).
static PassAArgs<double> krinv_pre_args(fame6d_qsys* S, const int* qmap) {
  PassAArgs<double> pa{};
  pa.V11 = S->V(0);
  pa.V12 = S->V(1);
  pa.V21 = S->V(2);
  pa.V22 = S->V(3);
  pa.W1 = S->W64(0);
  pa.W2 = S->W64(1);
  pa.invC = S->W64(2);
  pa.ldq = S->m->NF;
  pa.skip = no_skip();
  pa.qmap = qmap;
  return pa;
}

// Counts the iterations of an FP64 solve and checks its outcome (make_explicit_krinv_compact.m:
// "Mhat CG failed" unless flag 0 or the relative residual within 10 tol).
static void krinv_account(fame6d_qsys* S, const std::vector<PcgOut>& res, double tol, bool fp64_solve) {
  for (const PcgOut& o : res) {
    if (fp64_solve) {
      S->stats.total_iterations += o.iter;
      S->stats.total_fp64_iterations += o.iter;
      S->stats.mass_applies_fp64 += 3LL * (o.iter + o.checks);
    }
    if (!std::isfinite(o.relres) || o.flag < 0 || (o.flag != 0 && o.relres > 10 * tol))
      throw Error(FAME6D_ENOCONV, "Mhat CG failed: flag=" + std::to_string(o.flag) +
                                      ", relres=" + std::to_string(o.relres) +
                                      ", iter=" + std::to_string(o.iter) + ".");
    S->stats.max_iterations = std::max(S->stats.max_iterations, o.iter);
    S->stats.max_relres = std::max(S->stats.max_relres, o.relres);
    S->stats.last_iterations = o.iter;
    S->stats.last_flag = o.flag;
    S->stats.last_relres = o.relres;
  }
}

// z = invC u; s = M_s [W1 z; W2 z] (z and W z formed in the apply's first pass, unless the
// solve's last residual apply left s); y = V1^{-1}(mr + s)
static void krinv_tail(fame6d_qsys* S, int nact, const int* qmap, const KrinvWs64& w, bool zw_ready,
                       cpx<double>* yout) {
  fame6d_mass* m = S->m;
  const long long NF = m->NF;
  if (!zw_ready) {
    PassAArgs<double> pa = krinv_pre_args(S, qmap);
    pa.x = w.u;
    pa.ldx = NF;
    mass_apply_pre(m, PRE_ZW, nact, pa, w.s, NF, m->stream);
    S->stats.mass_applies_fp64 += 2 * nact;
  }
  k_v1_solve<double><<<dim3(grid_for(NF), nact), 256, 0, m->stream>>>(
      w.mr, w.s, 2 * NF, NF, S->V(0), S->V(1), S->V(2), S->V(3), yout, 2 * NF, no_skip(), qmap);
  F6_CHECK_LAUNCH();
}

// M_s [W1 z; W2 z] of the final u came with the passing true-residual check (flag 0 after at
// least one check); otherwise the tail forms it
static bool zw_from_check(const std::vector<PcgOut>& res) {
  for (const PcgOut& o : res)
    if (o.flag != 0 || o.checks < 1) return false;
  return true;
}

void krinv_groups(fame6d_qsys* S, int precision, int nact, const int* qmap_host,
                  const cpx<double>* qin, cpx<double>* yout, double tol, int maxit,
                  std::vector<PcgOut>* outs, KrinvDeferred* def) {
  if (nact <= 0) return;
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long NF = m->NF;
  const size_t gn = (size_t)NF * nact;
  const auto t0 = std::chrono::steady_clock::now();
  // device q-slot map
  S->sc.ensure(sizeof(int) * 256);
  int* qmap = S->sc.as<int>();
  std::vector<int> qm(qmap_host, qmap_host + nact);
  for (int v : qm)
    if (v < 0 || v >= S->G) throw Error(FAME6D_EARG, "bad q-system slot");
  if (nact > 64) throw Error(FAME6D_EARG, "at most 64 Bloch points per batched K_r^{-1}");
  if (def && (precision != FAME6D_FP64 || nact != 1))
    throw Error(FAME6D_EARG, "a deferred K_r^{-1} is FP64 with one group");
  if (S->qmap_uploaded != qm) {
    h2d_async(m, qmap, qm.data(), sizeof(int) * nact);
    S->qmap_uploaded = qm;
  }

  const KrinvWs64 w = krinv_ws(S, nact, def ? def->set : 0);
  const int gridm = grid_for(NF);

  // mr = M_s r with r = V1^{-*} q (both families), r formed in the apply's first pass
  PassAArgs<double> pa = krinv_pre_args(S, qmap);
  pa.x = qin;
  pa.ldx = 2 * NF;
  mass_apply_pre(m, PRE_V1H, nact, pa, w.mr, NF, s);
  S->stats.mass_applies_fp64 += 2 * nact;
  // bhat = -invC (W1 mr1 + W2 mr2): formed with the FP64 solve's start, or here for the mixed form

  std::vector<PcgOut> res;
  bool zw_ready = false;  // w.s holds M_s [W1 z; W2 z] of the final u
  if (def) {
    // first chunk and its check, then the tail as if the check passed; the caller reads the
    // state together with its own scalars and calls krinv_complete
    PcgRun<double> R(S, 1, qmap, w.u, w.cr, w.cp, w.cq, nullptr, false, maxit, true, w.s, def->set);
    R.bcheck = w.bhat;
    R.start(w.bhat, tol, w.mr);
    R.issue(S->chunk_hint > 0 ? S->chunk_hint : 8);
    F6_HIP(hipMemcpyAsync(def->host_stage ? def->host_stage : cg_stage(S), R.state(), sizeof(CgState),
                          hipMemcpyDeviceToHost, s));
    krinv_tail(S, 1, qmap, w, true, yout);
    def->pending = true;
    def->cur = R.cur;
    def->issued = R.issued;
    def->rounds = R.rounds;
    def->timer_first = R.timer_first;
    def->timer_owner = R.owner;
    def->tol = tol;
    def->maxit = maxit;
    def->yout = yout;
    def->state = R.state();
    def->t0 = t0;
    return;
  }
  if (precision == FAME6D_FP64) {
    pcg_mhat<double>(S, nact, qmap, w.bhat, w.u, w.cr, w.cp, w.cq, nullptr, false, tol, maxit, true,
                     S->chunk_hint, res, w.s, w.mr);
    zw_ready = zw_from_check(res);
    krinv_account(S, res, tol, true);
  } else if (precision == FAME6D_MIXED) {
    // FP64 iterative refinement of Mhat u = bhat with FP32 inner PCG solves.  The refinement
    // state (||b||, ||r||, converged) lives on the device; the host reads it once the
    // cycles that cannot have converged yet are done (after every cycle from then on).
    S->kr32.ensure(sizeof(cpx<float>) * gn * 6);
    cpx<float>* b32 = S->kr32.as<cpx<float>>();
    cpx<float>* x32 = b32 + gn;
    cpx<float>* r32 = b32 + 2 * gn;
    cpx<float>* p32 = b32 + 3 * gn;
    cpx<float>* q32 = b32 + 4 * gn;
    cpx<float>* z32 = b32 + 5 * gn;
    S->ref.ensure(sizeof(RefState) * nact);
    RefState* rs = S->ref.as<RefState>();
    const int nbD = red_blocks(NF);
    // sized for every user up front (the FP64 residual's pass-C partials, the inner FP32
    // solves' own), so no later ensure moves the buffer under this pointer
    const int nbP = std::max({nbD, pass_c_blocks(m, false, nact), pass_c_blocks(m, true, nact)});
    S->part2.ensure(sizeof(double) * ((size_t)nbP * nact + 8));
    double* part2 = S->part2.as<double>();
    // bhat, u = 0, res = bhat and the ||bhat||^2 partials in one pass over mr
    k_bhat_cg_start<double><<<dim3(nbD, nact), kRedThreads, 0, s>>>(w.mr, 2 * NF, NF, S->W64(0), S->W64(1),
                                                                    S->W64(2), qmap, w.bhat, w.u, w.res, part2);
    k_ref_init<<<nact, kRedThreads, 0, s>>>(rs, part2, nbD);
    F6_CHECK_LAUNCH();
    res.assign(nact, PcgOut{});
    const double tol_in = S->mixed_inner_tol;
    // cycles before the first possible convergence: each reduces the residual by about tol_in
    const int blind = (tol_in > tol && tol_in < 1.0)
                          ? std::max(1, (int)std::floor(std::log(tol) / std::log(tol_in)))
                          : 1;
    const Skip skr = skip_inactive(rs);
    std::vector<RefState> hr(nact);
    struct InnerRun {
      int budget, slot, timer_first, owner;
    };
    std::vector<InnerRun> runs;
    int cycles = 0, seen = 0;
    while (true) {
      if (cycles >= S->mixed_max_cycles)
        throw Error(FAME6D_ENOCONV, "Mhat iterative refinement did not converge");
      // The inner solve: one chunk of `budget` iterations and no host read (an inner solve
      // that stops short of tol_in still improves u, and the FP64 residual decides).  Its
      // right-hand side r / ||r|| is formed in its start pass (0 for converged groups, whose
      // inner solve then stops at once); its states are copied out for the host to read with
      // the refinement state.
      const int ci = std::min(cycles, (int)S->chunk_hint32.size() - 1);
      const int budget = std::max(2, std::min(maxit, S->chunk_hint32[ci] > 0 ? S->chunk_hint32[ci] : 16));
      {
        PcgRun<float> R(S, nact, qmap, x32, r32, p32, q32, z32, S->mixed_precond, budget, false, nullptr);
        R.start(nullptr, tol_in, nullptr, rs, w.res);
        R.issue(budget);
        F6_HIP(hipMemcpyAsync(inner_stage(S, nact) + (size_t)cycles * nact, R.state(), sizeof(CgState) * nact,
                              hipMemcpyDeviceToHost, s));
        runs.push_back({budget, ci, R.timer_first, R.owner});
      }
      ++cycles;
      k_accumulate_ref<float><<<dim3(gridm, nact), 256, 0, s>>>(NF, x32, NF, rs, w.u, NF);
      // FP64 residual of every group still refining
      // res = bhat - Mhat u and its ||res||^2 partials in the last pass of the apply
      mhat_apply<double>(S, nact, qmap, w.u, w.res, skr, s, false, nullptr, w.bhat, part2, w.s);
      k_ref_check<<<nact, kRedThreads, 0, s>>>(rs, part2, pass_c_blocks(m, false, nact), tol);
      F6_CHECK_LAUNCH();
      if (cycles < blind) continue;
      d2h_sync(m, hr.data(), rs, sizeof(RefState) * nact);
      // the inner solves of the cycles since the last read
      for (; seen < cycles; ++seen) {
        const CgState* hs = inner_stage(S, nact) + (size_t)seen * nact;
        const InnerRun& ir = runs[seen];
        int need = 2, max_iter = 0;
        for (int g = 0; g < nact; ++g) {
          const CgState& c = hs[g];
          if (c.flag != 0 && c.flag != 1 && c.flag != 3)
            throw Error(FAME6D_ENOCONV, "FP32 Mhat CG broke down (flag " + std::to_string(c.flag) + ")");
          S->stats.total_iterations += c.iter;
          if (c.iter > 0) {
            S->stats.mass_applies_fp32 +=
                3LL * ((S->mixed_precond ? 2 : 1) * c.iter + c.checks + (S->mixed_precond ? 1 : 0));
            S->stats.mass_applies_fp64 += 3;
          }
          res[g].iter += c.iter;
          max_iter = std::max(max_iter, c.iter);
          // the next budget of this cycle: the iterations it took plus one, or two more when
          // it ran out of budget before tol_in
          need = std::max(need, c.flag == 1 ? ir.budget + 2 : c.iter + 1);
        }
        // launches after the last group stopped ran as skipped launches
        m->timer.discard_tags(ir.timer_first, ir.owner, max_iter);
        S->chunk_hint32[ir.slot] = need;
      }
      bool any = false;
      for (int g = 0; g < nact; ++g) {
        if (hr[g].nonfinite) throw Error(FAME6D_ENONFINITE, "Mhat refinement residual is NaN/Inf.");
        if (!hr[g].inactive) any = true;
      }
      if (!any) break;
    }
    // every group's last FP64 residual apply was at its final u (converged groups are
    // skipped afterwards), so M_s [W1 z; W2 z] of the final u is in w.s
    zw_ready = true;
    for (int g = 0; g < nact; ++g) {
      res[g].flag = 0;
      res[g].relres = hr[g].nb > 0 ? hr[g].nr / hr[g].nb : 0.0;
      if (!(hr[g].nb > 0)) zw_ready = false;
    }
    S->stats.refinement_cycles += cycles;
    krinv_account(S, res, tol, false);
  } else {
    throw Error(FAME6D_EARG, "K_r^{-1} precision must be FAME6D_FP64 or FAME6D_MIXED");
  }
  krinv_tail(S, nact, qmap, w, zw_ready, yout);
  S->stats.calls += nact;
  if (outs) *outs = res;
  S->stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool krinv_complete(fame6d_qsys* S, KrinvDeferred* def) {
  if (!def || !def->pending) throw Error(FAME6D_EARG, "no deferred K_r^{-1} to complete");
  def->pending = false;
  std::vector<CgState> hs(1);
  std::memcpy(hs.data(), def->host_stage ? def->host_stage : cg_stage(S), sizeof(CgState));
  const int* qmap = S->sc.as<int>();
  const KrinvWs64 w = krinv_ws(S, 1, def->set);
  PcgRun<double> R(S, 1, qmap, w.u, w.cr, w.cp, w.cq, nullptr, false, def->maxit, true, w.s, def->set);
  R.owner = def->timer_owner;
  R.bcheck = w.bhat;
  R.cur = def->cur;
  R.issued = def->issued;
  R.rounds = def->rounds;
  R.timer_first = def->timer_first;
  const bool valid = PcgRun<double>::all_finished(hs) && hs[0].flag == 0 && hs[0].checks >= 1;
  std::vector<PcgOut> res;
  if (!valid) R.finish(hs, PcgRun<double>::next_chunk(hs));
  R.collect(hs, res, S->chunk_hint);
  krinv_account(S, res, def->tol, true);
  if (!valid) krinv_tail(S, 1, qmap, w, zw_from_check(res), def->yout);
  S->stats.calls += 1;
  S->stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - def->t0).count();
  return valid;
}

template void mhat_apply<double>(fame6d_qsys*, int, const int*, const cpx<double>*, cpx<double>*, Skip,
                                 hipStream_t, bool, double*, const cpx<double>*, double*, cpx<double>*);
template void mhat_apply<float>(fame6d_qsys*, int, const int*, const cpx<float>*, cpx<float>*, Skip,
                                hipStream_t, bool, double*, const cpx<float>*, double*, cpx<float>*);
template void dot_groups<double>(const fame6d_mass*, int, const cpx<double>*, long long,
                                 const cpx<double>*, long long, long long, double*, double*);
template void dot_groups<float>(const fame6d_mass*, int, const cpx<float>*, long long,
                                const cpx<float>*, long long, long long, double*, double*);

}  // namespace fame6d
