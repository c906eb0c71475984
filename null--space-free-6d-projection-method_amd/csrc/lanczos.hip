// Thick-restart Lanczos for the largest eigenvalues theta of K_r^{-1}
// (lambda = 1/theta), eigenvector recovery and Fourier diagnostics.
//
// Restates Lanczos_hybrid_identity_gpuop (exp1/Lanczos_hybrid_identity_gpuop.m:11-130):
// the recurrence, the extra projections, two full re-orthogonalisation passes
// per step, the projected matrix H = [diag(d) c'; c tridiag], the Ritz residual
// |beta * U(end,:)|, the convergence threshold tol*max(eps^(2/3), |theta|) and
// the restart size k0 + min(nconv, floor((p-k0)/2)).  The basis stays on the
// device.  The orthogonalisation of each new vector w = K_r^{-1} q_j is two full
// classical Gram-Schmidt projections r = (I - QQ^H)(I - QQ^H) w (four streaming
// sweeps over the basis); the reference's three-term recurrence, its two extra
// single-vector projections and its two column-by-column passes compute the same
// projection in exact arithmetic (alpha = Re(q_j^H w), beta = ||r||), so results
// agree up to rounding.  Several Bloch points step together and share one batched
// K_r^{-1} per step.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <cmath>
#include <functional>
#include <numeric>

#include "dense_eig.hpp"
#include "internal.hpp"
#include "krylov_kernels.hpp"
#include "minres_kernels.hpp"
#include "shift_schedule.hpp"

namespace fame6d {

// y += c * x with c = c_re[0] + i*c_im[0] read from device memory, times scale.
__global__ void k_axpy_dev(long long n, cpx<double>* y, const cpx<double>* x, const double* c_re,
                           const double* c_im, double scale) {
  const cpx<double> c = mk(scale * c_re[0], c_im ? scale * c_im[0] : 0.0);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + c * x[i];
}

// y = x * (1/sqrt(s[0]))
__global__ void k_scale_rsqrt_dev(long long n, cpx<double>* y, const cpx<double>* x, const double* s) {
  const double f = 1.0 / sqrt(s[0]);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = f * x[i];
}

// v = r / ||r|| with ||r||^2 = s[0], enqueued before the host has read ||r||: it writes nothing
// when the step breaks down (||r|| <= 0 or < 1e-14, where the host ends the cycle instead);
// with `cg`, when the K_r^{-1} solve behind r has not passed its true-residual check (the
// host then finishes that solve and forms the step again); with `no_second`, when the DGKS
// test asked for a second sweep (the host then runs it and scales again).
__global__ void k_scale_next(long long n, cpx<double>* y, const cpx<double>* x, const double* s,
                             const CgState* cg, const int* no_second) {
  if (cg && !(cg->finished && cg->flag == 0 && cg->checks >= 1)) return;
  if (no_second && *no_second == 0) return;  // a second Gram-Schmidt sweep comes first
  const double b2 = s[0];
  if (!(b2 > 0.0) || sqrt(b2) < 1e-14) return;
  const double f = 1.0 / sqrt(b2);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = f * x[i];
}

// Kop: y = K x with K_xi = |q|^2 I - q q^T (assemble_K_and_nullspace_free_basis.m:110-119)
__global__ void k_apply_k(const double* q1, const double* q2, const double* q3, const cpx<double>* x,
                          cpx<double>* y, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double a = q1[i], b = q2[i], c = q3[i];
    const cpx<double> x1 = x[i], x2 = x[n + i], x3 = x[2 * n + i];
    y[i] = (b * b + c * c) * x1 - (a * b) * x2 - (a * c) * x3;
    y[n + i] = (a * a + c * c) * x2 - (a * b) * x1 - (b * c) * x3;
    y[2 * n + i] = (a * a + b * b) * x3 - (a * c) * x1 - (b * c) * x2;
  }
}

// D0 .* (U0^* x) = sum_c q_c x_c  (compute_fourier_diagnostics_exp2.m:16-17; real q)
__global__ void k_div(const double* q1, const double* q2, const double* q3, const cpx<double>* x,
                      cpx<double>* y, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = q1[i] * x[i] + q2[i] * x[n + i] + q3[i] * x[2 * n + i];
}

// ------------------------------------------------------------------ device basis ops
struct Basis {
  cpx<double>* Q = nullptr;     // n x p
  cpx<double>* v = nullptr;     // next Lanczos vector
  cpx<double>* r = nullptr;     // residual work
  cpx<double>* Qalt = nullptr;  // second n x p basis for out-of-place restarts (p > 64)
  double* sc = nullptr;         // device scalars: [0] alpha, [1] beta2, [2] ||r||^2 before the
                                // sweep, [4] DGKS flag, then h, h2 (re, im) of kLzMaxP each
  double* part = nullptr;       // partials: h (re, im) rows of nb each for kLzMaxP columns,
                                // then nb of ||r||^2
};

constexpr int kScH = 8, kScHim = 8 + kLzMaxP;
constexpr int kScStride = 8 + 4 * kLzMaxP;  // doubles of device scalars per Lanczos run

// dst_re/dst_im = sum(conj(a) .* b)
static void dot_dev(long long n, const cpx<double>* a, const cpx<double>* b, double* part,
                    double* dst_re, double* dst_im, hipStream_t s) {
  const int nb = red_blocks(n);
  double* pim = part + nb;
  k_dot_partials<double><<<dim3(nb, 1), kRedThreads, 0, s>>>(a, n, b, n, n, part, pim);
  k_reduce_final<<<1, kRedThreads, 0, s>>>(part, nb, dst_re);
  if (dst_im) k_reduce_final<<<1, kRedThreads, 0, s>>>(pim, nb, dst_im);
  F6_CHECK_LAUNCH();
}

// One Lanczos orthogonalisation step: h1 = Q^H w (its entry jj-1 is alpha = q_j^H w),
// r = w - Q h1, h2 = Q^H r, r -= Q h2 with ||r||^2 fused into the last update.
// h1 in sc[kScH..], h2 in sc[kScH2..], ||r||^2 in sc[1].
constexpr int kScH2 = 8 + 2 * kLzMaxP, kScH2im = 8 + 3 * kLzMaxP;
constexpr int kGemvKB = 4;
// With `sk` set the projection is skipped (and not timed: its launch does no work).
// With `tt` set, v is w and the projection is of the three-term update w - alpha q_{jj-1} -
// beta q_{jj-2} (||.||^2 partials to tt->norm).
static void gemv_c(fame6d_mass* m, long long n, const Basis& B, int jj, const cpx<double>* v, double* dre,
                   double* dim, hipStream_t s, Skip sk = no_skip(), const ThreeTerm* tt = nullptr) {
  const int nb = red_blocks(n);
  const int nch = (jj + kGemvKB - 1) / kGemvKB;
  double* pre = B.part;
  double* pim = B.part + (size_t)nb * kLzMaxP;
  if (sk.p == nullptr) timer_begin(m, TS_GEMV_C);
  if (tt)
    launch_k(m, k_gemv_c_all<kGemvKB, true>, dim3(nb * nch), dim3(kRedThreads), s, B.Q, n, jj, v, n, nb, pre,
             pim, sk, *tt);
  else
    launch_k(m, k_gemv_c_all<kGemvKB, false>, dim3(nb * nch), dim3(kRedThreads), s, B.Q, n, jj, v, n, nb, pre,
             pim, sk, ThreeTerm{nullptr, 0.0, nullptr});
  if (sk.p == nullptr) timer_end(m, (jj + 1) * 16.0 * n);
  k_reduce_final_ri<<<2 * jj, kRedThreads, 0, s>>>(pre, pim, nb, jj, dre, dim, sk);
  F6_CHECK_LAUNCH();
}

static void cgs2_step(fame6d_mass* m, long long n, const Basis& B, int jj, const cpx<double>* w,
                      hipStream_t s) {
  const int gn = (int)std::min<long long>(2048, (n + 255) / 256);
  const double vb = 16.0 * n;  // bytes of one basis vector
  gemv_c(m, n, B, jj, w, B.sc + kScH, B.sc + kScHim, s);
  timer_begin(m, TS_GEMV_N);
  launch_k(m, k_gemv_n_out, dim3(gn), dim3(256), s, B.Q, n, jj, B.sc + kScH, B.sc + kScHim, w, B.r, n,
           nullptr, no_skip(), nullptr, 0.0);
  timer_end(m, (jj + 2) * vb);
  gemv_c(m, n, B, jj, B.r, B.sc + kScH2, B.sc + kScH2im, s);
  timer_begin(m, TS_GEMV_N);
  launch_k(m, k_gemv_n_out, dim3(gn), dim3(256), s, B.Q, n, jj, B.sc + kScH2, B.sc + kScH2im, B.r, B.r, n,
           B.part, no_skip(), nullptr, 0.0);
  timer_end(m, (jj + 2) * vb);
  k_reduce_final<<<1, kRedThreads, 0, s>>>(B.part, gn, B.sc + 1);
  F6_CHECK_LAUNCH();
}

// The second sweep of lanczos_orth_step, r -= Q (Q^H r) with ||r||^2 into sc[1]; with `sk`
// taken only when the DGKS test of the first sweep asked for it (the int at sc + 4 is 0).
static void lanczos_second_sweep(fame6d_mass* m, long long n, const Basis& B, int jj, hipStream_t s, Skip sk) {
  const int gn = (int)std::min<long long>(2048, (n + 255) / 256);
  gemv_c(m, n, B, jj, B.r, B.sc + kScH2, B.sc + kScH2im, s, sk);
  k_gemv_n_out<<<gn, 256, 0, s>>>(B.Q, n, jj, B.sc + kScH2, B.sc + kScH2im, B.r, B.r, n, B.part, sk);
  k_lz_dgks<<<1, kRedThreads, 0, s>>>(B.sc, B.part, gn, 2);
  F6_CHECK_LAUNCH();
}

// A Lanczos step away from a restart (Lanczos_hybrid_identity_gpuop.m:55-63): alpha =
// Re q_j^H w into sc[0]; the three-term update r = w - alpha q_j - beta q_{j-1}; one
// classical Gram-Schmidt sweep r -= Q (Q^H r) over the whole basis; and a second sweep only
// when the first left less than eta of ||r|| (the DGKS test, eta^2 = eta2, 1/2 by default),
// decided on the device (`second`) or left to the caller.  The three-term update removes the
// two large components of w, so the full sweep only cleans rounding-level ones and seldom
// needs repeating; the reference always runs two modified Gram-Schmidt sweeps (eta >= 1
// here).  ||r||^2 ends in sc[1].
static void lanczos_orth_step(fame6d_mass* m, long long n, const Basis& B, int jj, const cpx<double>* w,
                              double beta, hipStream_t s, bool second = true, double eta2 = 0.5) {
  const int gn = (int)std::min<long long>(2048, (n + 255) / 256);
  const double vb = 16.0 * n;
  const cpx<double>* qj = B.Q + (size_t)n * (jj - 1);
  dot_dev(n, qj, w, B.part, B.sc, nullptr, s);
  // the three-term update is formed inside the first sweep: Q^H r from w and the two newest
  // columns (with ||r||^2 for the DGKS test), then r = w - Q (h + alpha e_j + beta e_{j-1})
  const int nb = red_blocks(n);
  double* npart = B.part + (size_t)2 * nb * kLzMaxP;
  const ThreeTerm tt{B.sc, beta, npart};
  gemv_c(m, n, B, jj, w, B.sc + kScH, B.sc + kScHim, s, no_skip(), &tt);
  timer_begin(m, TS_GEMV_N);
  launch_k(m, k_gemv_n_out, dim3(gn), dim3(256), s, B.Q, n, jj, B.sc + kScH, B.sc + kScHim, w, B.r, n,
           B.part, no_skip(), (const double*)B.sc, beta);
  timer_end(m, (jj + 2) * vb);
  // ||r||^2 before (the three-term update) and after the sweep, and the DGKS test
  k_lz_dgks<<<1, kRedThreads, 0, s>>>(B.sc, B.part, gn, 1, npart, nb, eta2);
  F6_CHECK_LAUNCH();
  if (second) lanczos_second_sweep(m, n, B, jj, s, Skip{reinterpret_cast<const int*>(B.sc + 4), 0});
}

// ------------------------------------------------------------------ batched steps
// The orthogonalisation of one Lanczos step for several runs (the Bloch points of a batched
// eigensolve) with one launch per kernel, blockIdx.y = entry of a per-step table: the same
// kernels and arithmetic as lanczos_orth_step, each entry with its own basis, column count
// and scalars.
struct LzPt {
  cpx<double>* Q;         // basis (columns of n)
  cpx<double>* r;         // the orthogonalised vector
  double* sc;             // device scalars of the run
  double* part;           // partials of the run
  const cpx<double>* w;   // the K_r^{-1} output of the step (k_scale_pts: unused)
  cpx<double>* dst;       // k_scale_pts: where v = r / ||r|| goes
  double beta;            // beta of the three-term update (0 before the second column)
  int jj;                 // columns set; column jj - 1 is the current q_j
  int pad;
};
constexpr int kLzPtMax = 64;

// out(:, e) = Q_e(:, jj_e - 1), the K_r^{-1} inputs of the step packed for the batched solve
__global__ void k_gather_pts(const LzPt* pts, cpx<double>* out, long long n) {
  const LzPt P = pts[blockIdx.y];
  const cpx<double>* src = P.Q + (long long)(P.jj - 1) * n;
  cpx<double>* dst = out + blockIdx.y * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// partials of alpha = Re q_j^H w (k_dot_partials' element order and sums)
__global__ __launch_bounds__(kRedThreads) void k_dot_pts(const LzPt* pts, long long n) {
  __shared__ double red[kRedThreads];
  const LzPt P = pts[blockIdx.y];
  const cpx<double>* a = P.Q + (long long)(P.jj - 1) * n;
  double sr = 0.0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kRedThreads) {
    const cpx<double> u = a[i], v = P.w[i];
    sr += u.re * v.re + u.im * v.im;
  }
  sr = block_sum<kRedThreads>(sr, red);
  if (threadIdx.x == 0) P.part[blockIdx.x] = sr;
}

// sc[0] = alpha (grid = entries)
__global__ __launch_bounds__(kRedThreads) void k_reduce_alpha_pts(const LzPt* pts, int nb) {
  __shared__ double red[kRedThreads];
  const LzPt P = pts[blockIdx.x];
  double s = strided_sum(0.0, P.part, threadIdx.x, nb, kRedThreads);
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) P.sc[0] = s;
}

// h = Q^H r of the three-term update, grid (nrb * nchmax, entries); chunks past an entry's
// columns return
template <int KB>
__global__ __launch_bounds__(kRedThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_gemv_c_pts(
    const LzPt* pts, long long n, int nrb, int nchmax) {
  __shared__ double red[kRedThreads];
  const LzPt P = pts[blockIdx.y];
  const int logical = xcd_swizzle(blockIdx.x, gridDim.x);
  const int rb = logical / nchmax, ch = logical - rb * nchmax;
  if (ch * KB >= P.jj) return;
  const ThreeTerm tt{P.sc, P.beta, P.part + (size_t)2 * nrb * kLzMaxP};
  gemv_c_block<KB, true>(P.Q, n, P.jj, P.w, n, nrb, rb, ch, P.part, P.part + (size_t)nrb * kLzMaxP, tt, red);
}

// sc[kScH + k], sc[kScHim + k] for k < jj, grid (2 * jjmax, entries)
__global__ __launch_bounds__(kRedThreads) void k_reduce_h_pts(const LzPt* pts, int nb) {
  __shared__ double red[kRedThreads];
  const LzPt P = pts[blockIdx.y];
  if ((int)blockIdx.x >= 2 * P.jj) return;
  const bool im = (int)blockIdx.x >= P.jj;
  const int k = im ? blockIdx.x - P.jj : blockIdx.x;
  const double* partials = (im ? P.part + (size_t)nb * kLzMaxP : P.part) + (size_t)k * nb;
  double s = strided_sum(0.0, partials, threadIdx.x, nb, kRedThreads);
  s = block_sum<kRedThreads>(s, red);
  if (threadIdx.x == 0) P.sc[(im ? kScHim : kScH) + k] = s;
}

// r = w - Q (h + alpha e_j + beta e_{j-1}) with ||r||^2 partials, grid (gn, entries)
__global__ __launch_bounds__(256) void k_gemv_n_pts(const LzPt* pts, long long n) {
  __shared__ cpx<double> hs[kLzMaxP];
  __shared__ double red[256];
  const LzPt P = pts[blockIdx.y];
  gemv_n_block(P.Q, n, P.jj, P.sc + kScH, P.sc + kScHim, P.w, P.r, n, P.part, P.sc, P.beta, blockIdx.x,
               gridDim.x, hs, red);
}

// the DGKS test of each entry (k_lz_dgks stage 1), then its scalars sc[0 .. kScH + jj) into
// out + e * per for one read by the host (grid = entries)
__global__ __launch_bounds__(kRedThreads) void k_dgks_pts(const LzPt* pts, int gn, int nb, double eta2,
                                                          double* out, int per) {
  __shared__ double red[kRedThreads];
  const LzPt P = pts[blockIdx.x];
  lz_dgks_block(P.sc, P.part, gn, 1, P.part + (size_t)2 * nb * kLzMaxP, nb, eta2, red);
  __syncthreads();  // thread 0's scalars are visible to the block
  for (int t = threadIdx.x; t < kScH + P.jj; t += kRedThreads) out[(size_t)blockIdx.x * per + t] = P.sc[t];
}

// dst = r / sqrt(sc[1]), grid (blocks, entries)
__global__ void k_scale_pts(const LzPt* pts, long long n) {
  const LzPt P = pts[blockIdx.y];
  const double f = 1.0 / sqrt(P.sc[1]);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    P.dst[i] = f * P.r[i];
}

// Host-side state of one Lanczos run.
struct LzState {
  int slot = 0;
  Basis B;
  std::vector<double> Alpha, Beta, d, c, Us, res;
  std::vector<int> ind;
  double beta = 0;
  bool justRestarted = false;
  int sizeV = 1, jj = 1, k = 0, kk = 0, mm = 1, qcols = 0;
  bool done = false, converged = false, hitmax = false, breakdown = false;
  int nconv = 0;
  double maxres = 0;
  std::vector<double> theta;
  cpx<double>* Z = nullptr;  // M-weighted basis Z = M Q (Method A), restarted with Q
  int alpha_at = 0;          // index of alpha in the device scalars of the last step
  int second_sweeps = 0;     // steps whose DGKS test asked for a second Gram-Schmidt sweep
  cpx<double>* vloc = nullptr;  // where the newest unit vector is: B.v or the next basis column
};

// Q(:, 0:k) = Q(:, 0:m) U in place (n complex rows, column stride n; U real m x k, column-major,
// on the device; m, k <= 64)
static void restart_in_place(cpx<double>* Q, long long n, int m, int k, const double* U, hipStream_t s) {
  if (m > 64 || k > 64) throw Error(FAME6D_EARG, "in-place basis product of more than 64 columns");
  const long long nr = 2 * n;  // Q as a real (2n x m) matrix
  const int blocks = (int)((nr + 255) / 256);
  double* Qd = reinterpret_cast<double*>(Q);
  if (k <= 16) k_restart_rows<16><<<blocks, 256, 0, s>>>(Qd, nr, nr, m, k, U);
  else if (k <= 32) k_restart_rows<32><<<blocks, 256, 0, s>>>(Qd, nr, nr, m, k, U);
  else k_restart_rows<64><<<blocks, 256, 0, s>>>(Qd, nr, nr, m, k, U);
  F6_CHECK_LAUNCH();
}

// out(:, c) = Q(:, 0:m) U(:, c) for c < kk, out of place (n complex rows, column stride n for Q
// and out); the accumulator count follows kk
static void basis_product(const cpx<double>* Q, long long n, int m, int kk, const double* U, cpx<double>* out,
                          hipStream_t s) {
  const long long nr = 2 * n;
  const unsigned gx = (unsigned)((nr + 255) / 256);
  const double* Qd = reinterpret_cast<const double*>(Q);
  double* Od = reinterpret_cast<double*>(out);
  if (kk <= 16) k_basis_gemm<16><<<dim3(gx, 1), 256, 0, s>>>(Qd, nr, nr, m, kk, U, Od, nr);
  else if (kk <= 32) k_basis_gemm<32><<<dim3(gx, 1), 256, 0, s>>>(Qd, nr, nr, m, kk, U, Od, nr);
  else k_basis_gemm<64><<<dim3(gx, (unsigned)((kk + 63) / 64)), 256, 0, s>>>(Qd, nr, nr, m, kk, U, Od, nr);
  F6_CHECK_LAUNCH();
}

// Builds H, solves it, tests convergence and restarts (Lanczos_hybrid_identity_gpuop.m:79-120).
// Returns true when the run is finished.
static bool restart_step(LzState& L, int k0, int p, double tol, int maxRestart, long long n2,
                         DevBuf& ubuf, hipStream_t s) {
  const int nd = (int)L.d.size(), na = (int)L.Alpha.size();
  const int m = nd + na;
  if (m == 0) throw Error(FAME6D_ENOCONV, "Lanczos produced no projected data.");
  std::vector<double> H((size_t)m * m, 0.0);
  for (int i = 0; i < nd; ++i) H[(size_t)i * m + i] = L.d[i];
  for (int i = 0; i < na; ++i) H[(size_t)(nd + i) * m + nd + i] = L.Alpha[i];
  for (int i = 0; i + 1 < na; ++i) {
    H[(size_t)(nd + i + 1) * m + nd + i] = L.Beta[i];
    H[(size_t)(nd + i) * m + nd + i + 1] = L.Beta[i];
  }
  if (nd > 0) {
    for (int i = 0; i < L.k; ++i) {
      H[(size_t)L.k * m + i] = L.c[i];  // H(1:k, k+1) = c'
      H[(size_t)i * m + L.k] = L.c[i];  // H(k+1, 1:k) = c
    }
  }
  L.Alpha.clear();
  L.Beta.clear();
  std::vector<double> w, U;
  try {
    sym_eig(m, H, w, U);
  } catch (const std::runtime_error& e) {
    throw Error(FAME6D_ENOCONV, e.what());
  }
  std::vector<int> ind(m);
  std::iota(ind.begin(), ind.end(), 0);
  std::stable_sort(ind.begin(), ind.end(), [&](int a, int b) { return w[a] > w[b]; });
  std::vector<double> ds(m), res(m);
  for (int i = 0; i < m; ++i) {
    ds[i] = w[ind[i]];
    res[i] = std::fabs(L.beta * U[(size_t)ind[i] * m + (m - 1)]);
  }
  L.Us = U;
  L.ind = ind;
  L.res = res;
  L.qcols = m;
  const int nw = std::min(k0, m);
  const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
  int nconv = 0;
  double mx = 0;
  for (int i = 0; i < nw; ++i) {
    const double thr = tol * std::max(eps23, std::fabs(ds[i]));
    if (res[i] < thr) ++nconv;
    mx = std::max(mx, res[i]);
  }
  L.nconv = nconv;
  L.maxres = mx;
  L.d = ds;
  if (nconv >= k0) {
    L.converged = true;
    return true;
  }
  if (L.mm >= maxRestart) {
    L.hitmax = true;
    return true;
  }
  int k = k0 + std::min(nconv, (p - k0) / 2);
  if (k == 1 && p > 3) k = p / 2;
  if (k > m) k = m;
  // Q(:,1:k) = Q(:,1:m) * Ukeep ; d = d(1:k) ; c = beta * Ukeep(end,:)
  std::vector<double> Uk((size_t)m * k);
  L.c.assign(k, 0.0);
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i < m; ++i) Uk[(size_t)j * m + i] = U[(size_t)ind[j] * m + i];
    L.c[j] = L.beta * U[(size_t)ind[j] * m + (m - 1)];
  }
  ubuf.ensure(sizeof(double) * Uk.size());
  F6_HIP(hipMemcpyAsync(ubuf.ptr, Uk.data(), sizeof(double) * Uk.size(), hipMemcpyHostToDevice, s));
  if (m <= 64) {
    restart_in_place(L.B.Q, n2, m, k, ubuf.as<double>(), s);
    if (L.Z) restart_in_place(L.Z, n2, m, k, ubuf.as<double>(), s);
  } else {
    // wide basis: out-of-place product into the second basis, then swap the two
    if (!L.B.Qalt || L.Z) throw Error(FAME6D_EARG, "Lanczos restart of a basis wider than 64 needs the second basis");
    basis_product(L.B.Q, n2, m, k, ubuf.as<double>(), L.B.Qalt, s);
    std::swap(L.B.Q, L.B.Qalt);
  }
  F6_CHECK_LAUNCH();
  F6_HIP(hipStreamSynchronize(s));  // ubuf is reused by the next restart
  L.d.resize(k);
  L.k = k;
  L.justRestarted = true;
  L.sizeV = k + 1;
  L.jj = k + 1;
  L.mm += 1;
  return false;
}

// Eigensolve of all G Bloch points of S.  v0: G x 2NF host start vectors.
// Nested-method settings when the eigensolve runs Method B (tolM, maxitM, epsC).
struct NestedCfg {
  bool on = false;
  double tolM = 1e-12, epsC = 1.0;
  int maxitM = 1600;
};
static thread_local NestedCfg g_nested;

void krinv_nested(fame6d_qsys* S, const cpx<double>* bin, cpx<double>* xout, double tolKr, double tolM,
                  int maxitKr, int maxitM, double epsC, fame6d_krinv_stats_t* st);

void set_nested(bool on, double tolM, int maxitM, double epsC) {
  g_nested.on = on;
  g_nested.tolM = tolM;
  g_nested.maxitM = maxitM;
  g_nested.epsC = epsC;
}

// The Lanczos loop of one FP64 run with the host one step ahead of the device.  Each step's
// K_r^{-1} solve is enqueued with its first PCG chunk, true-residual check and tail (the
// deferred form of krinv_groups), followed by the orthogonalisation and the scaling of the
// next Lanczos vector into its column, which does nothing unless the solve passed its check
// and the step did not break down.  While the cycle continues, the next step's solve is
// enqueued at once on that column (in the other workspace set) before the host waits for
// this step's scalars, so the device never waits for the host between steps.  A solve that
// needed more iterations is continued from its own workspace set and its step is formed
// again; the next step enqueued meanwhile read a column that was not yet written and is
// dropped, as it is after a breakdown.  The arithmetic is that of the one-read loop.
static void lanczos_run_ahead(fame6d_qsys* S, LzState& Lg, int k0, int p, double tol, int maxRestart,
                              double tol_cg, int maxit_cg, cpx<double>* qout, DevBuf& ubuf, int& steps) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long n2 = 2 * m->NF;
  const int gr = grid_for(n2);
  struct Stage {
    CgState cg;
    double sc[kScH + kLzMaxP];
  };
  struct HostStages {
    Stage* p = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~HostStages() {
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
      if (p) (void)hipHostFree(p);
    }
  } hst;
  F6_HIP(hipHostMalloc(reinterpret_cast<void**>(&hst.p), sizeof(Stage) * 2, hipHostMallocDefault));
  for (hipEvent_t& e : hst.ev) F6_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  KrinvDeferred kd[2];
  auto enqueue_solve = [&](int k, const cpx<double>* vin) {
    kd[k] = KrinvDeferred{};
    kd[k].set = k;
    kd[k].host_stage = &hst.p[k].cg;
    int slot = Lg.slot;
    krinv_groups(S, FAME6D_FP64, 1, &slot, vin, qout, tol_cg, maxit_cg, nullptr, &kd[k]);
  };
  auto next_dst = [&]() { return Lg.jj < p ? Lg.B.Q + (size_t)n2 * Lg.jj : Lg.B.v; };
  // the second DGKS sweep is not enqueued ahead: the host runs it when the test asked for it
  const int* no_second = reinterpret_cast<const int*>(Lg.B.sc + 4);
  auto enqueue_orth = [&](const CgState* gate) {
    const int jj = Lg.jj;
    if (Lg.justRestarted) {
      // w couples to every kept Ritz vector: CGS2 over the whole basis (alpha = Re h1[j])
      cgs2_step(m, n2, Lg.B, jj, qout, s);
      Lg.alpha_at = kScH + jj - 1;
    } else {
      lanczos_orth_step(m, n2, Lg.B, jj, qout, jj >= 2 ? Lg.beta : 0.0, s, false, S->dgks_eta2);
      Lg.alpha_at = 0;
    }
    k_scale_next<<<gr, 256, 0, s>>>(n2, next_dst(), Lg.B.r, Lg.B.sc + 1, gate, Lg.alpha_at == 0 ? no_second : nullptr);
    F6_CHECK_LAUNCH();
  };
  auto stage_scalars = [&](int k) {
    F6_HIP(hipMemcpyAsync(hst.p[k].sc, Lg.B.sc, sizeof(double) * (kScH + Lg.jj), hipMemcpyDeviceToHost, s));
    F6_HIP(hipEventRecord(hst.ev[k], s));
  };

  int k = 0;
  bool queued = false;  // this step's solve is already in the stream (slot k)
  while (!Lg.done) {
    const int jj = Lg.jj;  // columns 0..jj-1 of Q are set, column jj-1 is the current q_j
    if (!queued) {
      cpx<double>* Qj = Lg.B.Q + (size_t)n2 * (jj - 1);
      if (Lg.vloc != Qj)  // after the start and after restarts
        F6_HIP(hipMemcpyAsync(Qj, Lg.vloc, sizeof(cpx<double>) * n2, hipMemcpyDeviceToDevice, s));
      enqueue_solve(k, Qj);
    }
    enqueue_orth(kd[k].state);
    stage_scalars(k);
    // the next step's solve, on the column this step's scaling writes
    bool ahead = jj < p;
    if (ahead) enqueue_solve(k ^ 1, Lg.B.Q + (size_t)n2 * jj);
    F6_HIP(hipEventSynchronize(hst.ev[k]));
    ++steps;
    if (!krinv_complete(S, &kd[k])) {
      // the solve was continued and its tail formed again: so is the step
      enqueue_orth(nullptr);
      stage_scalars(k);
      F6_HIP(hipEventSynchronize(hst.ev[k]));
      ahead = false;
    }
    if (Lg.alpha_at == 0) {
      int skipped2 = 0;
      std::memcpy(&skipped2, hst.p[k].sc + 4, sizeof(int));
      if (!skipped2) {
        // the DGKS test asked for the second sweep: run it, then scale
        Lg.second_sweeps += 1;
        lanczos_second_sweep(m, n2, Lg.B, Lg.jj, s, no_skip());
        k_scale_next<<<gr, 256, 0, s>>>(n2, next_dst(), Lg.B.r, Lg.B.sc + 1, nullptr, nullptr);
        F6_CHECK_LAUNCH();
        stage_scalars(k);
        F6_HIP(hipEventSynchronize(hst.ev[k]));
        ahead = false;
      }
    }
    const double* v = hst.p[k].sc;
    const double alpha = v[Lg.alpha_at], beta2 = v[1];
    Lg.justRestarted = false;
    if (!std::isfinite(alpha) || !std::isfinite(beta2))
      throw Error(FAME6D_ENONFINITE, "Lanczos inverse action returned invalid data.");
    bool end_cycle = false;
    if (beta2 <= 0 || std::sqrt(beta2) < 1e-14) {  // the scaling wrote nothing
      Lg.beta = std::sqrt(std::max(beta2, 0.0));
      Lg.breakdown = true;
      end_cycle = true;
      ahead = false;
    } else {
      Lg.beta = std::sqrt(beta2);
      Lg.vloc = next_dst();
      Lg.Alpha.push_back(alpha);
      Lg.Beta.push_back(Lg.beta);
      Lg.kk += 1;
      if (Lg.jj >= p) end_cycle = true;
      else Lg.jj += 1;
    }
    if (!ahead) kd[k ^ 1].pending = false;  // dropped: its kernels only touch workspace set k^1
    if (end_cycle) {
      // the restart rewrites the basis columns in place: the newest v must be in B.v
      if (Lg.vloc != Lg.B.v) {
        F6_HIP(hipMemcpyAsync(Lg.B.v, Lg.vloc, sizeof(cpx<double>) * n2, hipMemcpyDeviceToDevice, s));
        Lg.vloc = Lg.B.v;
      }
      if (restart_step(Lg, k0, p, tol, maxRestart, n2, ubuf, s)) Lg.done = true;
    }
    queued = ahead;
    k ^= 1;
  }
  F6_HIP(hipStreamSynchronize(s));
}

void eigs_method_c(fame6d_qsys* S, int precision, int k0, int p, double tol, int maxRestart,
                   double tol_cg, int maxit_cg, const cpx<double>* v0, double* lambda, cpx<double>* Y,
                   fame6d_lanczos_stats_t* stats, const RitzSelect* sel) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const int G = S->G;
  const long long NF = m->NF, n2 = 2 * NF;
  if (k0 < 1 || p <= k0 || p > kLzMaxP) throw Error(FAME6D_EARG, "need 1 <= k < p <= 256");
  if (sel && G != 1) throw Error(FAME6D_EARG, "Ritz selection needs a single Bloch point");
  if (G > 64) throw Error(FAME6D_EARG, "at most 64 Bloch points");
  if (precision != FAME6D_FP64 && precision != FAME6D_MIXED)
    throw Error(FAME6D_EARG, "eigensolver precision must be FAME6D_FP64 or FAME6D_MIXED");
  const auto t0 = std::chrono::steady_clock::now();
  const fame6d_krinv_stats_t st0 = S->stats;

  // device memory: per group Q (n2 x p) + v + r; scalars; partials; packed in/out
  const int nb = red_blocks(n2);
  DevBuf basis, scal, part, io, ubuf;
  DevBuf ptab, sstage;  // batched steps: per-step tables, the scalars staged for one read
  const int cols = p + 2 + (p > 64 ? p : 0);  // Q, v, r (+ the second basis of wide restarts)
  const size_t part_per = (size_t)std::max(nb * 2 * kLzMaxP, 2048) + nb;
  basis.ensure(sizeof(cpx<double>) * (size_t)n2 * cols * G);
  scal.ensure(sizeof(double) * kScStride * G);
  part.ensure(sizeof(double) * part_per * G);
  io.ensure(sizeof(cpx<double>) * (size_t)n2 * 2 * G);
  std::vector<LzState> L(G);
  for (int g = 0; g < G; ++g) {
    L[g].slot = g;
    cpx<double>* b0 = basis.as<cpx<double>>() + (size_t)n2 * cols * g;
    L[g].B.Q = b0;
    L[g].B.v = b0 + (size_t)n2 * p;
    L[g].B.r = b0 + (size_t)n2 * (p + 1);
    L[g].B.Qalt = p > 64 ? b0 + (size_t)n2 * (p + 2) : nullptr;
    L[g].B.sc = scal.as<double>() + (size_t)kScStride * g;
    L[g].B.part = part.as<double>() + part_per * g;
    L[g].k = k0;
  }
  cpx<double>* qin = io.as<cpx<double>>();
  cpx<double>* qout = qin + (size_t)n2 * G;
  const int gr = grid_for(n2);

  // v = v0 / ||v0||  (Lanczos_hybrid_identity_gpuop.m:14-16)
  for (int g = 0; g < G; ++g) {
    F6_HIP(hipMemcpyAsync(L[g].B.v, v0 + (size_t)n2 * g, sizeof(cpx<double>) * n2, hipMemcpyDefault, s));
    dot_dev(n2, L[g].B.v, L[g].B.v, L[g].B.part, L[g].B.sc + 1, nullptr, s);
  }
  {
    std::vector<double> nv(G);
    for (int g = 0; g < G; ++g)
      F6_HIP(hipMemcpyAsync(&nv[g], L[g].B.sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    F6_HIP(hipStreamSynchronize(s));
    for (int g = 0; g < G; ++g) {
      if (!(std::isfinite(nv[g]) && nv[g] > 0)) throw Error(FAME6D_EARG, "Invalid Lanczos start vector.");
      k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n2, L[g].B.v, L[g].B.v, L[g].B.sc + 1);
      L[g].vloc = L[g].B.v;
    }
    F6_CHECK_LAUNCH();
  }

  int steps = 0;
  if (G == 1 && !g_nested.on && precision == FAME6D_FP64)
    lanczos_run_ahead(S, L[0], k0, p, tol, maxRestart, tol_cg, maxit_cg, qout, ubuf, steps);
  while (true) {
    std::vector<int> act;
    for (int g = 0; g < G; ++g)
      if (!L[g].done) act.push_back(g);
    if (act.empty()) break;
    const int na = (int)act.size();
    std::vector<int> slots(na);
    // K_r^{-1} reads its input only, so a single active run passes its basis column itself;
    // several are gathered into qin by one launch over the step's table
    const cpx<double>* kin = qin;
    const bool batch = na >= 2;
    std::vector<LzPt> tab(na);
    for (int i = 0; i < na; ++i) {
      LzState& Lg = L[act[i]];
      cpx<double>* Qj = Lg.B.Q + (size_t)n2 * (Lg.jj - 1);
      if (Lg.vloc != Qj)  // after the start and after restarts; a step's own v is written in place
        F6_HIP(hipMemcpyAsync(Qj, Lg.vloc, sizeof(cpx<double>) * n2, hipMemcpyDeviceToDevice, s));
      if (na == 1) kin = Qj;
      tab[i] = LzPt{Lg.B.Q, Lg.B.r, Lg.B.sc, Lg.B.part, qout + (size_t)n2 * i, nullptr,
                    Lg.jj >= 2 ? Lg.beta : 0.0, Lg.jj, 0};
      slots[i] = Lg.slot;
    }
    LzPt* tab_d = nullptr;  // three tables of kLzPtMax entries: the step, its batched sweep, the scaling
    if (batch) {
      ptab.ensure(sizeof(LzPt) * 3 * kLzPtMax);
      tab_d = ptab.as<LzPt>();
      h2d_async(m, tab_d, tab.data(), sizeof(LzPt) * na);
      k_gather_pts<<<dim3(gr, na), 256, 0, s>>>(tab_d, qin, n2);
      F6_CHECK_LAUNCH();
    }
    // a single FP64 run defers the solve's confirmation to the step's one read of scalars
    const bool defer = na == 1 && !g_nested.on && precision == FAME6D_FP64;
    KrinvDeferred kd;
    if (g_nested.on) {
      for (int i = 0; i < na; ++i)
        krinv_nested(S, kin + (size_t)n2 * i, qout + (size_t)n2 * i, tol_cg, g_nested.tolM, maxit_cg,
                     g_nested.maxitM, g_nested.epsC, nullptr);
    } else {
      krinv_groups(S, precision, na, slots.data(), kin, qout, tol_cg, maxit_cg, nullptr, defer ? &kd : nullptr);
    }
    ++steps;
    // where v = r / beta goes: the column it takes next, or B.v when the basis is full and a
    // restart follows
    auto next_dst = [&](LzState& Lg) { return Lg.jj < p ? Lg.B.Q + (size_t)n2 * Lg.jj : Lg.B.v; };
    const int per = kScH + p;  // doubles of one run's staged scalars
    std::vector<int> sub_of(na, -1);  // entry of each active run in the batched sweep, or -1
    int nsub = 0;
    auto orth_batched = [&]() {
      // runs just restarted take the full CGS2 sweep, one by one; the others one launch each
      std::vector<LzPt> sub;
      for (int i = 0; i < na; ++i) {
        LzState& Lg = L[act[i]];
        if (Lg.justRestarted) {
          cgs2_step(m, n2, Lg.B, Lg.jj, qout + (size_t)n2 * i, s);
          Lg.alpha_at = kScH + Lg.jj - 1;
          sub_of[i] = -1;
        } else {
          sub_of[i] = (int)sub.size();
          sub.push_back(tab[i]);
          Lg.alpha_at = 0;
        }
      }
      nsub = (int)sub.size();
      if (nsub == 0) return;
      LzPt* t2 = tab_d + kLzPtMax;
      h2d_async(m, t2, sub.data(), sizeof(LzPt) * nsub);
      int jjmax = 0;
      for (const LzPt& e : sub) jjmax = std::max(jjmax, e.jj);
      const int nchmax = (jjmax + kGemvKB - 1) / kGemvKB;
      const int gn = (int)std::min<long long>(2048, (n2 + 255) / 256);
      sstage.ensure(sizeof(double) * (size_t)kLzPtMax * per);
      k_dot_pts<<<dim3(nb, nsub), kRedThreads, 0, s>>>(t2, n2);
      k_reduce_alpha_pts<<<nsub, kRedThreads, 0, s>>>(t2, nb);
      k_gemv_c_pts<kGemvKB><<<dim3(nb * nchmax, nsub), kRedThreads, 0, s>>>(t2, n2, nb, nchmax);
      k_reduce_h_pts<<<dim3(2 * jjmax, nsub), kRedThreads, 0, s>>>(t2, nb);
      k_gemv_n_pts<<<dim3(gn, nsub), 256, 0, s>>>(t2, n2);
      k_dgks_pts<<<nsub, kRedThreads, 0, s>>>(t2, gn, nb, S->dgks_eta2, sstage.as<double>(), per);
      F6_CHECK_LAUNCH();
    };
    auto orth = [&](const CgState* gate) {
      if (batch) {
        orth_batched();
        return;
      }
      for (int i = 0; i < na; ++i) {
        LzState& Lg = L[act[i]];
        const Basis& B = Lg.B;
        const int jj = Lg.jj;  // columns 0..jj-1 of Q are set, column jj-1 is the current q_j
        const cpx<double>* w = qout + (size_t)n2 * i;
        if (Lg.justRestarted) {
          // w couples to every kept Ritz vector: CGS2 over the whole basis (alpha = Re h1[j])
          cgs2_step(m, n2, B, jj, w, s);
          Lg.alpha_at = kScH + jj - 1;
        } else {
          // the deferred single run keeps the device-side skip of the second sweep (its
          // k_scale_next waits on the same flag); otherwise the host issues the second sweep
          // after the step's read of scalars, only when the DGKS test asked for it
          lanczos_orth_step(m, n2, B, jj, w, jj >= 2 ? Lg.beta : 0.0, s, defer, S->dgks_eta2);
          Lg.alpha_at = 0;
        }
        if (defer) {
          k_scale_next<<<gr, 256, 0, s>>>(n2, next_dst(Lg), B.r, B.sc + 1, gate, nullptr);
          F6_CHECK_LAUNCH();
        }
      }
    };
    // alpha = Re h1[jj-1] and beta^2 of every active run: one staged read, one sync
    std::vector<double> ab(2 * na);
    std::vector<int> second(na, 0);  // the DGKS second sweep ran (steps away from a restart)
    // with `staged`, the batched entries' scalars come from the stage of k_dgks_pts in one copy
    auto read_scalars = [&](bool staged) {
      const size_t bper = sizeof(double) * per;
      char* stage = static_cast<char*>(m->pin.take(bper * na, s));
      char* sst = staged && nsub > 0 ? static_cast<char*>(m->pin.take(bper * nsub, s)) : nullptr;
      if (sst) F6_HIP(hipMemcpyAsync(sst, sstage.ptr, bper * nsub, hipMemcpyDeviceToHost, s));
      for (int i = 0; i < na; ++i)
        if (!sst || sub_of[i] < 0)
          F6_HIP(hipMemcpyAsync(stage + bper * i, L[act[i]].B.sc, sizeof(double) * (kScH + L[act[i]].jj),
                                hipMemcpyDeviceToHost, s));
      F6_HIP(hipStreamSynchronize(s));
      m->pin.used = 0;
      for (int i = 0; i < na; ++i) {
        const double* v = reinterpret_cast<const double*>(sst && sub_of[i] >= 0 ? sst + bper * sub_of[i]
                                                                                 : stage + bper * i);
        ab[2 * i] = v[L[act[i]].alpha_at];
        ab[2 * i + 1] = v[1];
        second[i] = 0;
        if (L[act[i]].alpha_at == 0) {
          int skipped2 = 0;
          std::memcpy(&skipped2, v + 4, sizeof(int));
          second[i] = skipped2 ? 0 : 1;
        }
      }
    };
    orth(defer ? kd.state : nullptr);
    read_scalars(batch);
    if (defer && !krinv_complete(S, &kd)) {
      // the solve needed more iterations: its tail has been formed again, and so is the step
      orth(nullptr);
      read_scalars(false);
    }
    if (!defer) {
      // second Gram-Schmidt sweeps the DGKS tests asked for, then beta^2 read again
      bool any = false;
      for (int i = 0; i < na; ++i)
        if (second[i]) {
          lanczos_second_sweep(m, n2, L[act[i]].B, L[act[i]].jj, s, no_skip());
          any = true;
        }
      if (any) {
        const std::vector<int> asked = second;
        read_scalars(false);
        second = asked;
      }
    }
    for (int i = 0; i < na; ++i) {
      L[act[i]].second_sweeps += second[i];
      L[act[i]].justRestarted = false;
    }
    // v = r / beta of every run (one launch when batched, before any restart reads it), then
    // the restarts of the runs whose cycle ended
    std::vector<char> ends(na, 0);
    std::vector<LzPt> scl;
    for (int i = 0; i < na; ++i) {
      LzState& Lg = L[act[i]];
      const double alpha = ab[2 * i], beta2 = ab[2 * i + 1];
      if (!std::isfinite(alpha) || !std::isfinite(beta2))
        throw Error(FAME6D_ENONFINITE, "Lanczos inverse action returned invalid data.");
      bool end_cycle = false;
      if (beta2 <= 0 || std::sqrt(beta2) < 1e-14) {
        Lg.beta = std::sqrt(std::max(beta2, 0.0));
        Lg.breakdown = true;
        end_cycle = true;
      } else {
        Lg.beta = std::sqrt(beta2);
        cpx<double>* dst = next_dst(Lg);
        if (batch) {
          scl.push_back(LzPt{nullptr, Lg.B.r, Lg.B.sc, nullptr, nullptr, dst, 0.0, 0, 0});
        } else if (!defer) {
          k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n2, dst, Lg.B.r, Lg.B.sc + 1);
          F6_CHECK_LAUNCH();
        }
        Lg.vloc = dst;
        Lg.Alpha.push_back(alpha);
        Lg.Beta.push_back(Lg.beta);
        Lg.kk += 1;
        if (Lg.jj >= p) end_cycle = true;
        else Lg.jj += 1;
      }
      ends[i] = end_cycle;
    }
    if (!scl.empty()) {
      LzPt* t3 = tab_d + 2 * kLzPtMax;
      h2d_async(m, t3, scl.data(), sizeof(LzPt) * scl.size());
      k_scale_pts<<<dim3(gr, (unsigned)scl.size()), 256, 0, s>>>(t3, n2);
      F6_CHECK_LAUNCH();
    }
    for (int i = 0; i < na; ++i) {
      LzState& Lg = L[act[i]];
      if (ends[i]) {
        // the restart rewrites the basis columns in place: the newest v must be in B.v
        if (Lg.vloc != Lg.B.v) {
          F6_HIP(hipMemcpyAsync(Lg.B.v, Lg.vloc, sizeof(cpx<double>) * n2, hipMemcpyDeviceToDevice, s));
          Lg.vloc = Lg.B.v;
        }
        if (restart_step(Lg, k0, p, tol, maxRestart, n2, ubuf, s)) Lg.done = true;
      }
    }
  }

  // outputs: theta = d(1:k0) (descending), lambda = sort(1./theta) ascending;
  // Ritz vectors Y(:,i) = Q(:,1:qcols) * U(:, ind(i)) in the same order as lambda.
  for (int g = 0; g < G; ++g) {
    LzState& Lg = L[g];
    if ((int)Lg.d.size() < k0) throw Error(FAME6D_ENOCONV, "Lanczos ended without k inverse Ritz values.");
    std::vector<std::pair<double, int>> lam(k0);
    for (int i = 0; i < k0; ++i) {
      const double th = Lg.d[i];
      if (!std::isfinite(th) || th <= 0) throw Error(FAME6D_ENOCONV, "Method C invalid inverse Ritz values.");
      lam[i] = {1.0 / th, i};
    }
    std::stable_sort(lam.begin(), lam.end());
    for (int i = 0; i < k0; ++i) lambda[(size_t)g * k0 + i] = lam[i].first;
    // Ritz vectors to form: all k0 (in the order of lambda), or the tracking selection
    // (exp3/Lanczos_tracking.m:134-165), whose indices count the inverse Ritz values from the
    // largest, i.e. the eigenvalues from the smallest
    std::vector<int> want;  // positions in d order
    if (sel) {
      if (sel->rel_res)
        for (int i = 0; i < k0; ++i) sel->rel_res[i] = Lg.res[i] / std::max(std::fabs(Lg.d[i]), 2.220446049250313e-16);
      std::vector<int> pick;
      auto add = [&](int v) {
        if (std::find(pick.begin(), pick.end(), v) == pick.end()) pick.push_back(v);
      };
      if (sel->nidx > 0) {
        for (int t = 0; t < sel->nidx; ++t) {
          if (sel->idx[t] < 1 || sel->idx[t] > k0)
            throw Error(FAME6D_EARG, "Requested fixed tracking index is outside the scan.");
          add(sel->idx[t]);
        }
      } else {
        for (int t = 0; t < sel->ntarget; ++t) {
          const double tg = sel->target[t];
          int ic = 0;
          double best = INFINITY;
          for (int i = 0; i < k0; ++i) {
            const double dv = std::fabs(lam[i].first - tg) / std::max(std::fabs(tg), 2.220446049250313e-16);
            if (dv < best) {
              best = dv;
              ic = i;
            }
          }
          for (int i = std::max(0, ic - sel->half_width); i <= std::min(k0 - 1, ic + sel->half_width); ++i)
            add(i + 1);
        }
        if ((int)pick.size() < 2)
          throw Error(FAME6D_EARG, "Tracking candidate extraction produced only " + std::to_string(pick.size()) +
                                       " candidate(s).");
      }
      if ((int)pick.size() > sel->idx_capacity) throw Error(FAME6D_EARG, "selection exceeds its output capacity");
      for (size_t t = 0; t < pick.size(); ++t) {
        sel->idx_out[t] = pick[t];
        want.push_back(lam[pick[t] - 1].second);
      }
      *sel->nidx_out = (int)pick.size();
    } else {
      for (int i = 0; i < k0; ++i) want.push_back(lam[i].second);
    }
    if (Y) {
      const int mq = Lg.qcols, ny = (int)want.size();
      std::vector<double> Uy((size_t)mq * ny);
      for (int i = 0; i < ny; ++i) {
        const int col = Lg.ind[want[i]];
        for (int r = 0; r < mq; ++r) Uy[(size_t)i * mq + r] = Lg.Us[(size_t)col * mq + r];
      }
      ubuf.ensure(sizeof(double) * Uy.size());
      F6_HIP(hipMemcpyAsync(ubuf.ptr, Uy.data(), sizeof(double) * Uy.size(), hipMemcpyHostToDevice, s));
      basis_product(Lg.B.Q, n2, mq, ny, ubuf.as<double>(), Y + (size_t)n2 * ny * g, s);
      F6_HIP(hipStreamSynchronize(s));
    }
    if (stats) {
      fame6d_lanczos_stats_t& o = stats[g];
      o.converged = Lg.converged ? 1 : 0;
      o.num_converged = Lg.nconv;
      o.outer_steps = Lg.kk;
      o.second_sweeps = Lg.second_sweeps;
      o.restarts = Lg.mm;
      o.hit_max_restart = Lg.hitmax ? 1 : 0;
      o.max_ritz_residual = Lg.maxres;
    }
  }
  F6_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    fame6d_krinv_stats_t d = S->stats;
    d.calls -= st0.calls;
    d.total_iterations -= st0.total_iterations;
    d.total_fp64_iterations -= st0.total_fp64_iterations;
    d.refinement_cycles -= st0.refinement_cycles;
    d.mass_applies_fp64 -= st0.mass_applies_fp64;
    d.mass_applies_fp32 -= st0.mass_applies_fp32;
    d.seconds -= st0.seconds;
    for (int g = 0; g < G; ++g) {
      stats[g].seconds = secs;
      stats[g].krinv = d;
    }
  }
  basis.release();
  scal.release();
  part.release();
  io.release();
  ubuf.release();
  for (int g = 0; g < G; ++g)
    if (!L[g].converged)
      throw Error(FAME6D_ENOCONV, "Method C outer Lanczos did not converge: " +
                                      std::to_string(L[g].nconv) + "/" + std::to_string(k0) + ".");
}

// ------------------------------------------------------------------ host-stepped PCG
// MATLAB pcg (x0 = 0) with an optional preconditioner z = Minv(r), host-stepped: one
// synchronisation per iteration.  Used where the matvec itself is an inner solve
// (Method B) and for the eigenvector recovery.
using Matvec = std::function<void(const cpx<double>*, cpx<double>*)>;

static PcgOut pcg_host(fame6d_mass* m, long long n, const Matvec& A, const Matvec* Minv,
                       const cpx<double>* b, cpx<double>* x, cpx<double>* r, cpx<double>* p,
                       cpx<double>* q, cpx<double>* z, double tol, int maxit) {
  hipStream_t s = m->stream;
  const int gr = grid_for(n);
  auto dot = [&](const cpx<double>* a, const cpx<double>* c) {
    double re;
    dot_groups<double>(m, 1, a, n, c, n, n, &re, nullptr);
    return re;
  };
  PcgOut o;
  k_fill_zero<double><<<gr, 256, 0, s>>>(n, x);
  F6_HIP(hipMemcpyAsync(r, b, sizeof(cpx<double>) * n, hipMemcpyDeviceToDevice, s));
  const double n2b = std::sqrt(dot(b, b));
  if (!std::isfinite(n2b)) throw Error(FAME6D_ENONFINITE, "CG right-hand side contains NaN/Inf");
  if (n2b == 0) return o;
  const double tolb = tol * n2b;
  double rho = 1, normr = n2b;
  int moresteps = 0;
  const int maxmsteps = (int)std::min<long long>({n / 50, 5, std::max<long long>(n - maxit, 0)});
  o.flag = 1;
  o.iter = maxit;
  for (int ii = 1; ii <= maxit; ++ii) {
    const cpx<double>* zz = r;
    if (Minv) {
      (*Minv)(r, z);
      zz = z;
    }
    const double rho1 = rho;
    rho = dot(r, zz);
    if (rho == 0 || std::isinf(rho)) {
      o.flag = 4;
      o.iter = ii - 1;
      break;
    }
    if (ii == 1) {
      F6_HIP(hipMemcpyAsync(p, zz, sizeof(cpx<double>) * n, hipMemcpyDeviceToDevice, s));
    } else {
      const double beta = rho / rho1;
      if (beta == 0 || std::isinf(beta)) {
        o.flag = 4;
        o.iter = ii - 1;
        break;
      }
      k_axpby<double><<<gr, 256, 0, s>>>(n, mk(1.0, 0.0), zz, mk(beta, 0.0), p, p);
    }
    A(p, q);
    const double pq = dot(p, q);
    if (!(pq > 0) || std::isinf(pq)) {
      o.flag = 4;
      o.iter = ii - 1;
      break;
    }
    const double alpha = rho / pq;
    k_axpby<double><<<gr, 256, 0, s>>>(n, mk(1.0, 0.0), x, mk(alpha, 0.0), p, x);
    k_axpby<double><<<gr, 256, 0, s>>>(n, mk(1.0, 0.0), r, mk(-alpha, 0.0), q, r);
    normr = std::sqrt(dot(r, r));
    o.iter = ii;
    if (normr <= tolb || moresteps) {
      A(x, q);
      k_axpby<double><<<gr, 256, 0, s>>>(n, mk(1.0, 0.0), b, mk(-1.0, 0.0), q, r);
      const double act = std::sqrt(dot(r, r));
      ++o.checks;
      normr = act;
      if (act <= tolb) {
        o.flag = 0;
        break;
      }
      moresteps += 1;
      if (moresteps >= maxmsteps) {
        o.flag = 3;
        break;
      }
    }
  }
  o.relres = normr / n2b;
  return o;
}

static void mass3(fame6d_mass* m, const cpx<double>* x, cpx<double>* y, Skip skip, hipStream_t s);

// p = r on the first iteration, p = r + beta p after it (MATLAB pcg without preconditioner);
// skipped once the solve is done.
static __global__ void k_cg_direction(const CgState* st, const cpx<double>* r, cpx<double>* p, long long n) {
  if (st->done) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = first ? r[i] : r[i] + beta * p[i];
}

// MATLAB pcg (x0 = 0, no preconditioner) on M = I3 (x) M_s with its control on the device:
// the iteration kernels read and update one CgState and are skipped once it is done; the
// host issues `hint` iterations at a time and reads the state once per chunk, and the
// true-residual confirmation and "moresteps" continuation run on the device too.  `hint`
// carries the last solve's iteration count to the next one.
static PcgOut pcg_mass3(fame6d_mass* m, const cpx<double>* b, cpx<double>* x, cpx<double>* r,
                        cpx<double>* p, cpx<double>* q, double tol, int maxit, int& hint) {
  hipStream_t s = m->stream;
  const long long n = 3 * m->NF;
  const int nb = red_blocks(n);
  const int gr = grid_for(n);
  DevBuf ctl;
  ctl.ensure(256 + sizeof(double) * (size_t)nb + 256);
  CgState* st = ctl.as<CgState>();
  double* part = reinterpret_cast<double*>(ctl.as<char>() + 256);
  int* need = reinterpret_cast<int*>(part + nb);
  int* resume = need + 1;
  const Skip sk{reinterpret_cast<const int*>(reinterpret_cast<const char*>(st) + offsetof(CgState, done)), 0};
  const int maxmsteps = (int)std::min<long long>({n / 50, 5, std::max<long long>(n - maxit, 0)});
  k_fill_zero<double><<<gr, 256, 0, s>>>(n, x);
  F6_HIP(hipMemcpyAsync(r, b, sizeof(cpx<double>) * n, hipMemcpyDeviceToDevice, s));
  k_dot_partials<double><<<dim3(nb, 1), kRedThreads, 0, s>>>(b, n, b, n, n, part, nullptr);
  k_cg_init<<<1, kRedThreads, 0, s>>>(st, part, nb, tol, maxit, maxmsteps);
  F6_CHECK_LAUNCH();
  int chunk = std::max(2, hint);
  CgState hs{};
  for (int rounds = 0;; ++rounds) {
    if (rounds > maxit + 8) throw Error(FAME6D_ENOCONV, "M CG control loop did not terminate");
    for (int c = 0; c < chunk; ++c) {
      k_cg_direction<<<gr, 256, 0, s>>>(st, r, p, n);
      mass3(m, p, q, sk, s);
      k_dot_partials<double><<<dim3(nb, 1), kRedThreads, 0, s>>>(p, n, q, n, n, part, nullptr, sk);
      k_cg_alpha<<<1, kRedThreads, 0, s>>>(st, part, nb);
      k_cg_update<double><<<dim3(nb, 1), kRedThreads, 0, s>>>(x, p, r, q, n, n, st, part);
      k_cg_beta<<<1, kRedThreads, 0, s>>>(st, part, nb);
      F6_CHECK_LAUNCH();
    }
    // true residual r = b - M x once the recurrence has stopped (MATLAB pcg's confirmation)
    k_cg_need_check<<<1, 64, 0, s>>>(st, 1, need);
    const Skip skn{need, 0};
    mass3(m, x, q, skn, s);
    k_residual<double><<<dim3(gr, 1), 256, 0, s>>>(b, q, r, n, n, skn);
    k_dot_partials<double><<<dim3(nb, 1), kRedThreads, 0, s>>>(r, n, r, n, n, part, nullptr, skn);
    k_cg_true_check<<<1, kRedThreads, 0, s>>>(st, part, nb, need, resume, 0);
    F6_CHECK_LAUNCH();
    d2h_sync(m, &hs, st, sizeof(CgState));
    if (hs.finished) break;
    chunk = hs.check_every ? 1 : 2;
  }
  hint = std::max(2, hs.iter + 1);
  PcgOut o;
  o.flag = hs.flag;
  o.iter = hs.iter;
  o.checks = hs.checks;
  const double nr = hs.checks > 0 ? hs.normr_true : hs.normr;
  o.relres = hs.n2b > 0 ? nr / hs.n2b : (std::isnan(hs.n2b) ? hs.n2b : 0.0);
  return o;
}

// ------------------------------------------------------------------ Method B
// make_nested_krinv_instrumented (exp1/make_nested_krinv_instrumented.m:33-132): outer PCG
// on K_r x = b with preconditioner epsC * r ./ sigma_r (sigma_r = [|q|^2; |q|^2], :47),
// every K_r z = V^* (M^{-1} (V z)) with an inner PCG on M = I3 (x) M_s (tolM, maxitM).
__global__ void k_sigma_precond(const double* q1, const double* q2, const double* q3, long long n,
                                double scale, const cpx<double>* r, cpx<double>* z) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double qq = q1[i] * q1[i] + q2[i] * q2[i] + q3[i] * q3[i];
    z[i] = (scale / qq) * r[i];
    z[n + i] = (scale / qq) * r[n + i];
  }
}

// y(2NF) = V^* x(3NF)   (apply_VH_blockdiag.m:13-15; V real)
__global__ void k_apply_vh(const cpx<double>* x, long long n, const double* V11, const double* V12,
                           const double* V21, const double* V22, const double* V31, const double* V32,
                           cpx<double>* y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cpx<double> x1 = x[i], x2 = x[n + i], x3 = x[2 * n + i];
    y[i] = V11[i] * x1 + V21[i] * x2 + V31[i] * x3;
    y[n + i] = V12[i] * x1 + V22[i] * x2 + V32[i] * x3;
  }
}

void krinv_nested(fame6d_qsys* S, const cpx<double>* bin, cpx<double>* xout, double tolKr, double tolM,
                  int maxitKr, int maxitM, double epsC, fame6d_krinv_stats_t* st) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long NF = m->NF, n2 = 2 * NF, n3 = 3 * NF;
  if (S->G != 1) throw Error(FAME6D_EARG, "Method B expects a single-point q-system");
  DevBuf ws;
  ws.ensure(sizeof(cpx<double>) * (size_t)(5 * n2 + 6 * n3));
  cpx<double>* r = ws.as<cpx<double>>();
  cpx<double>* p = r + n2;
  cpx<double>* q = p + n2;
  cpx<double>* z = q + n2;
  cpx<double>* xo = z + n2;
  cpx<double>* vz = xo + n2;
  cpx<double>* mi = vz + n3;
  cpx<double>* ir = mi + n3;
  cpx<double>* ip = ir + n3;
  cpx<double>* iq = ip + n3;
  cpx<double>* tmp = iq + n3;
  const int gr = grid_for(NF);
  long long inner_its = 0, inner_calls = 0;
  int inner_max = 0, inner_hint = 0;
  double inner_relres = 0;
  Matvec Kr = [&](const cpx<double>* zin, cpx<double>* yo) {
    k_apply_v<<<dim3(gr, 1), 256, 0, s>>>(zin, n2, NF, S->V(0), S->V(1), S->V(2), S->V(3), S->V(4), S->V(5),
                                          vz, n3, no_skip(), nullptr);
    F6_CHECK_LAUNCH();
    PcgOut o = pcg_mass3(m, vz, mi, ir, ip, iq, tolM, maxitM, inner_hint);
    if (!std::isfinite(o.relres) || (o.flag != 0 && o.relres > 10 * tolM))
      throw Error(FAME6D_ENOCONV, "Nested M solve failed: flag=" + std::to_string(o.flag));
    inner_its += o.iter;
    inner_calls += 1;
    inner_max = std::max(inner_max, o.iter);
    inner_relres = std::max(inner_relres, o.relres);
    k_apply_vh<<<gr, 256, 0, s>>>(mi, NF, S->V(0), S->V(1), S->V(2), S->V(3), S->V(4), S->V(5), yo);
    F6_CHECK_LAUNCH();
  };
  Matvec Pc = [&](const cpx<double>* rin, cpx<double>* zo) {
    k_sigma_precond<<<gr, 256, 0, s>>>(S->Qm(0), S->Qm(1), S->Qm(2), NF, epsC, rin, zo);
    F6_CHECK_LAUNCH();
  };
  const auto t0 = std::chrono::steady_clock::now();
  PcgOut o = pcg_host(m, n2, Kr, &Pc, bin, xout, r, p, q, z, tolKr, maxitKr);
  (void)xo;
  (void)tmp;
  if (!std::isfinite(o.relres) || (o.flag != 0 && o.relres > 10 * tolKr))
    throw Error(FAME6D_ENOCONV, "Nested K_r solve failed: flag=" + std::to_string(o.flag) +
                                    ", relres=" + std::to_string(o.relres));
  F6_HIP(hipStreamSynchronize(s));
  ws.release();
  S->stats.calls += 1;
  S->stats.total_iterations += o.iter;
  S->stats.total_fp64_iterations += inner_its;
  S->stats.max_iterations = std::max(S->stats.max_iterations, o.iter);
  S->stats.last_iterations = o.iter;
  S->stats.last_flag = o.flag;
  S->stats.last_relres = o.relres;
  S->stats.max_relres = std::max(S->stats.max_relres, o.relres);
  S->stats.mass_applies_fp64 += 3 * (inner_its + inner_calls);
  S->stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = S->stats;
  (void)inner_max;
  (void)inner_relres;
}

void recover_eigenvectors(fame6d_qsys* S, const cpx<double>* Y, int k, double tol, int maxit,
                          cpx<double>* X, cpx<double>* MX, fame6d_recover_stats_t* stats) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long NF = m->NF, n3 = 3 * NF;
  if (k < 0 || k > 64) throw Error(FAME6D_EARG, "k must be in [0, 64]");
  DevBuf ws;
  ws.ensure(sizeof(cpx<double>) * (size_t)n3 * 4);
  cpx<double>* rhs = ws.as<cpx<double>>();
  cpx<double>* r = rhs + n3;
  cpx<double>* p = rhs + 2 * n3;
  cpx<double>* q = rhs + 3 * n3;
  const int gr = grid_for(NF);
  int hint = 0;
  for (int j = 0; j < k; ++j) {
    k_apply_v<<<dim3(gr, 1), 256, 0, s>>>(Y + (size_t)2 * NF * j, 2 * NF, NF, S->V(0), S->V(1), S->V(2),
                                          S->V(3), S->V(4), S->V(5), rhs, n3, no_skip(), nullptr);
    F6_CHECK_LAUNCH();
    cpx<double>* xj = X + (size_t)n3 * j;
    cpx<double>* mxj = MX + (size_t)n3 * j;
    PcgOut o = pcg_mass3(m, rhs, xj, r, p, q, tol, maxit, hint);
    if (!std::isfinite(o.relres) || (o.flag != 0 && o.relres > 10 * tol))
      throw Error(FAME6D_ENOCONV, "Recovery CG failed for eigenvector " + std::to_string(j + 1));
    mass_apply<double>(m, xj, NF, mxj, NF, 3, s);
    double nj2;
    dot_groups<double>(m, 1, xj, n3, mxj, n3, n3, &nj2, nullptr);
    if (!std::isfinite(nj2) || nj2 <= 0)
      throw Error(FAME6D_EMASSNORM, "Nonpositive/nonfinite M-norm for eigenvector " + std::to_string(j + 1));
    const double sc = 1.0 / std::sqrt(nj2);
    k_axpby<double><<<grid_for(n3), 256, 0, s>>>(n3, mk(sc, 0.0), xj, mk(0.0, 0.0), xj, xj);
    k_axpby<double><<<grid_for(n3), 256, 0, s>>>(n3, mk(sc, 0.0), mxj, mk(0.0, 0.0), mxj, mxj);
    F6_CHECK_LAUNCH();
    if (stats) {
      stats->iterations[j] = o.iter;
      stats->relres[j] = o.relres;
      stats->mass_norm2[j] = nj2;
    }
  }
  F6_HIP(hipStreamSynchronize(s));
  ws.release();
}

void fourier_diagnostics(fame6d_qsys* S, const cpx<double>* X, const cpx<double>* MX,
                         const double* lambda, int k, double* res, double* div, double* ray) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long NF = m->NF, n3 = 3 * NF;
  DevBuf ws;
  ws.ensure(sizeof(cpx<double>) * (size_t)n3 * 2);
  cpx<double>* Kx = ws.as<cpx<double>>();
  cpx<double>* t = Kx + n3;
  for (int j = 0; j < k; ++j) {
    const cpx<double>* x = X + (size_t)n3 * j;
    const cpx<double>* mx = MX + (size_t)n3 * j;
    k_apply_k<<<grid_for(NF), 256, 0, s>>>(S->Qm(0), S->Qm(1), S->Qm(2), x, Kx, NF);
    F6_CHECK_LAUNCH();
    double xKx, xMx, nK, nM;
    dot_groups<double>(m, 1, x, n3, Kx, n3, n3, &xKx, nullptr);
    dot_groups<double>(m, 1, x, n3, mx, n3, n3, &xMx, nullptr);
    dot_groups<double>(m, 1, Kx, n3, Kx, n3, n3, &nK, nullptr);
    dot_groups<double>(m, 1, mx, n3, mx, n3, n3, &nM, nullptr);
    nK = std::sqrt(nK);
    nM = std::sqrt(nM);
    if (ray) ray[j] = xKx / xMx;
    k_axpby<double><<<grid_for(n3), 256, 0, s>>>(n3, mk(1.0, 0.0), Kx, mk(-lambda[j], 0.0), mx, t);
    double nr;
    dot_groups<double>(m, 1, t, n3, t, n3, n3, &nr, nullptr);
    const double denom = nK + std::fabs(lambda[j]) * nM;
    if (res) res[j] = std::sqrt(nr) / std::max(denom, 2.220446049250313e-16);
    k_div<<<grid_for(NF), 256, 0, s>>>(S->Qm(0), S->Qm(1), S->Qm(2), mx, t, NF);
    F6_CHECK_LAUNCH();
    double nd;
    dot_groups<double>(m, 1, t, NF, t, NF, NF, &nd, nullptr);
    if (div) div[j] = std::sqrt(nd) / std::max(nM, 2.220446049250313e-16);
  }
  ws.release();
}

// ================================================================== Method A
// Shift-invert on the full 3NF Fourier space: T x = (K - sigma M)^{-1} M x, whose largest
// eigenvalues theta give lambda = sigma + 1/theta (make_A_eigs_top_preconditioned.m:1-60,
// run_method_A_fig1a.m:29-50).  T is self-adjoint in the M inner product, so the outer
// solver is the thick-restart Lanczos above run in that inner product, with Z = M Q kept
// beside the basis (the M q of each step is the right-hand side of its shifted solve); the
// reference hands T to MATLAB eigs (Krylov-Schur) instead.  Each shifted system is solved by
// preconditioned MINRES with P = K + tau I, tau = sigma * mean(eps).

// y = M x for the three components of a 3NF vector (skip: stride 0 over the components).
static void mass3(fame6d_mass* m, const cpx<double>* x, cpx<double>* y, Skip skip, hipStream_t s) {
  const long long NF = m->NF;
  DevBuf& w = work_buf<double>(m);
  w.ensure(sizeof(cpx<double>) * (size_t)NF * 3);
  cpx<double>* t = w.as<cpx<double>>();
  PassAArgs<double> a{};
  a.x = x;
  a.ldx = NF;
  a.t = t;
  a.ldt = NF;
  a.skip = skip;
  launch_pass_a<double>(m, PRE_COPY, 3, a, s);
  launch_pass_b<double>(m, 1, 0, 3, t, NF, skip, s);
  PassCArgs<double> c{};
  c.t = t;
  c.ldt = NF;
  c.y = y;
  c.ldy = NF;
  c.skip = skip;
  launch_pass_c<double>(m, POST_COPY, 3, c, s);
}

// Workspace of one MINRES solve (3NF complex vectors) and its device scalars.  x is the
// solution of the shifted system; dx and rc are the iterate and right-hand side of a
// residual-correction call.
struct MrWs {
  cpx<double>*x, *v, *vn, *Mv, *r1, *r2, *y, *z, *w, *w1, *w2, *res, *dx, *rc;
  MrState* st;
  double* part;
  double* red;  // reduced scalars
};

struct MrAcc {
  long long solves = 0, iterations = 0, segments = 0, restarts = 0, corrections = 0, correction_iterations = 0;
  int max_iterations = 0, plateau_accepts = 0;
  double max_relres = 0.0, seconds = 0.0;
};

// sum over the partials, read to the host (one synchronisation)
static double reduce_read(fame6d_mass* m, const double* part, int nb, double* red) {
  k_reduce_final<<<1, kRedThreads, 0, m->stream>>>(part, nb, red);
  F6_CHECK_LAUNCH();
  double h = 0.0;
  d2h_sync(m, &h, red, sizeof(double));
  return h;
}

// One MINRES iteration moves the recurrence buffers round: (r1, r2, y) <- (r2, y, r1),
// (w1, w2, w) <- (w2, w, w1), (v, vn) <- (vn, v).
static void mr_rotate(MrWs& W) {
  std::swap(W.r1, W.r2);
  std::swap(W.r2, W.y);
  cpx<double>* ow1 = W.w1;
  W.w1 = W.w2;
  W.w2 = W.w;
  W.w = ow1;
  std::swap(W.v, W.vn);
}

// ||b - A x|| / bn with A = K - sigma M; leaves b - A x in out.
static double mr_explicit_relres(fame6d_qsys* S, double sigma, const cpx<double>* b, double bn,
                                 const cpx<double>* x, cpx<double>* out, MrWs& W) {
  fame6d_mass* m = S->m;
  const long long NF = m->NF;
  const int nb = red_blocks(NF);
  mass3(m, x, W.Mv, no_skip(), m->stream);
  k_mr_shifted<<<nb, kRedThreads, 0, m->stream>>>(S->Qm(0), S->Qm(1), S->Qm(2), sigma, x, W.Mv, nullptr, nullptr,
                                                   0, out, b, NF, W.part);
  F6_CHECK_LAUNCH();
  const double rel = std::sqrt(reduce_read(m, W.part, nb, W.red)) / bn;
  if (!std::isfinite(rel)) throw Error(FAME6D_ENONFINITE, "MINRES true residual is NaN/Inf.");
  return rel;
}

// One call of MATLAB minres(A, b, tol, maxit, P, [], x0) as the reference makes it
// (make_A_eigs_top_preconditioned.m:284-286 from x0, :619-620 from zero), with A = K - sigma M
// and P = K + tau I.  x holds x0 (from_x) or is cleared, and receives the iterate.  A call
// whose start already meets tol returns flag 0 after no iterations.  The Paige-Saunders
// recurrence runs until its estimate of the P^{-1}-norm residual reaches tol * ||b||_{P^{-1}};
// the explicit residual ||b - A x|| / ||b|| is then checked.  On a miss the recurrence goes on
// one step at a time with a check after each, five checks in all (MATLAB's continuation past
// a failed true-residual check), and then returns flag 3.  Flag 1: the maxit budget ran out;
// flag 4: the Lanczos process broke down (beta = 0) without meeting the check.
static MrCallResult mr_call(fame6d_qsys* S, double sigma, double tau, const cpx<double>* b, double bn, double tol,
                            int maxit, bool from_x, cpx<double>* x, MrWs& W, int& hint) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long NF = m->NF, n3 = 3 * NF;
  const double *q1 = S->Qm(0), *q2 = S->Qm(1), *q3 = S->Qm(2);
  const int nb = red_blocks(NF);
  const int gr = grid_for(n3);
  constexpr int kChecks = 5;
  MrCallResult out;
  double tol_abs = -1.0;  // target of the estimate: tol * beta1 from zero, else tol * ||b||_P^{-1}
  if (!from_x) {
    k_fill_zero<double><<<gr, 256, 0, s>>>(n3, x);
    F6_HIP(hipMemcpyAsync(W.r1, b, sizeof(cpx<double>) * n3, hipMemcpyDeviceToDevice, s));
  } else {
    k_mr_precond<<<nb, kRedThreads, 0, s>>>(q1, q2, q3, tau, const_cast<cpx<double>*>(b), nullptr, nullptr, W.z,
                                            NF, W.part);
    F6_CHECK_LAUNCH();
    tol_abs = tol * std::sqrt(std::max(reduce_read(m, W.part, nb, W.red), 0.0));
    out.relres = mr_explicit_relres(S, sigma, b, bn, x, W.r1, W);
    if (out.relres <= tol) return out;
  }
  k_mr_precond<<<nb, kRedThreads, 0, s>>>(q1, q2, q3, tau, W.r1, nullptr, nullptr, W.z, NF, W.part);
  k_mr_init<<<1, kRedThreads, 0, s>>>(W.st, W.part, nb, tol, tol_abs, maxit);
  F6_HIP(hipMemcpyAsync(W.r2, W.r1, sizeof(cpx<double>) * n3, hipMemcpyDeviceToDevice, s));
  k_fill_zero<double><<<gr, 256, 0, s>>>(n3, W.w);
  k_fill_zero<double><<<gr, 256, 0, s>>>(n3, W.w2);
  k_mr_first_v<<<gr, 256, 0, s>>>(W.st, W.z, W.v, n3);
  F6_CHECK_LAUNCH();
  const Skip sk = skip_mr(W.st);
  int issued = 0;
  // issues `count` iterations; those after the one that ends the recurrence are skipped on
  // the device, and the buffer roles are put back to where that iteration left them
  auto issue = [&](int count) {
    const MrWs W0 = W;
    const int base = issued;
    for (int c = 0; c < count; ++c, ++issued) {
      mass3(m, W.v, W.Mv, sk, s);
      k_mr_shifted<<<nb, kRedThreads, 0, s>>>(q1, q2, q3, sigma, W.v, W.Mv, W.r1, W.st, 1, W.y, nullptr, NF,
                                               W.part);
      k_mr_alpha<<<1, kRedThreads, 0, s>>>(W.st, W.part, nb);
      k_mr_precond<<<nb, kRedThreads, 0, s>>>(q1, q2, q3, tau, W.y, W.r2, W.st, W.z, NF, W.part);
      k_mr_givens<<<1, kRedThreads, 0, s>>>(W.st, W.part, nb);
      mr_rotate(W);  // W.w is now the free slot; W.w1, W.w2 hold the two previous directions
      k_mr_update<<<gr, 256, 0, s>>>(W.st, issued, W.vn, W.w1, W.w2, W.w, x, W.z, W.v, n3);
      F6_CHECK_LAUNCH();
    }
    MrState hs{};
    d2h_sync(m, &hs, W.st, sizeof(MrState));
    if (hs.flag == -1) throw Error(FAME6D_ENONFINITE, "MINRES produced nonfinite scalars.");
    if (hs.iter < issued) {
      W = W0;
      for (int j = base; j < hs.iter; ++j) mr_rotate(W);
      issued = hs.iter;
    }
    return hs;
  };
  MrState hs{};
  for (int chunk = std::max(4, hint); !hs.done; chunk = std::max(4, chunk / 2)) hs = issue(chunk);
  if (!from_x) hint = std::max(4, hs.iter + 1);
  for (int check = 1;; ++check) {
    out.iter = hs.iter;
    out.relres = mr_explicit_relres(S, sigma, b, bn, x, W.res, W);
    if (out.relres <= tol) {
      out.flag = 0;
      return out;
    }
    if (hs.flag == 1) {
      out.flag = 1;
      return out;
    }
    if (!(hs.beta > 0.0)) {
      out.flag = 4;
      return out;
    }
    if (hs.iter >= maxit) {
      out.flag = 1;
      return out;
    }
    if (check == kChecks) {
      out.flag = 3;
      return out;
    }
    k_mr_rearm<<<1, 1, 0, s>>>(W.st, 0.0);  // no estimate target: one step, then the next check
    F6_CHECK_LAUNCH();
    hs = issue(1);
  }
}

// Solves (K - sigma M) x = b to the true relative residual tol under the reference's
// recovery schedule (make_A_eigs_top_preconditioned.m:203-560, run_shift_schedule):
// MINRES calls from zero and then from the current iterate, residual-correction cycles
// when a call stalls, and acceptance at the 1e-9 plateau or guard ceilings.  The explicit
// residual has a floor near eps*max|q|^2*||x||/||b|| because the gradient part of x grows
// like 1/sigma: a shifted system whose floor lies above 1e-9 ends with FAME6D_ENOCONV,
// as the reference's controlled stop.
static void minres_solve(fame6d_qsys* S, double sigma, double tau, const cpx<double>* b, double tol, int maxit,
                         MrWs& W, int& hint, MrAcc& acc) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long n3 = 3 * m->NF;
  const auto t0 = std::chrono::steady_clock::now();
  k_dot_partials<double><<<dim3(red_blocks(n3), 1), kRedThreads, 0, s>>>(b, n3, b, n3, n3, W.part, nullptr);
  const double bn = std::sqrt(reduce_read(m, W.part, red_blocks(n3), W.red));
  if (!std::isfinite(bn)) throw Error(FAME6D_ENONFINITE, "Invalid shifted-system right-hand side.");
  acc.solves += 1;
  if (!(bn > 0)) {
    k_fill_zero<double><<<grid_for(n3), 256, 0, s>>>(n3, W.x);
    F6_CHECK_LAUNCH();
    return;
  }
  double rc_norm = 0.0;
  auto call = [&](bool from_current) {
    return mr_call(S, sigma, tau, b, bn, tol, maxit, from_current, W.x, W, hint);
  };
  auto residual = [&]() {
    const double rel = mr_explicit_relres(S, sigma, b, bn, W.x, W.rc, W);
    rc_norm = rel * bn;
    return rel;
  };
  auto correct = [&](double corr_tol) {
    int chint = 4;
    MrCallResult r = mr_call(S, sigma, tau, W.rc, rc_norm, corr_tol, maxit, false, W.dx, W, chint);
    k_accumulate<double><<<grid_for(n3), 256, 0, s>>>(n3, W.dx, 1.0, W.x);
    F6_CHECK_LAUNCH();
    r.relres = mr_explicit_relres(S, sigma, b, bn, W.x, W.res, W);
    return r;
  };
  const ShiftOutcome o = run_shift_schedule(tol, ShiftRecovery{}, call, residual, correct, [](int) {});
  acc.iterations += o.iterations;
  acc.segments += o.calls;
  acc.restarts += o.stagnation_restarts + o.maxit_restarts;
  acc.corrections += o.corrections;
  acc.correction_iterations += o.correction_iterations;
  acc.max_iterations = std::max(acc.max_iterations, o.iterations);
  acc.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (o.end == kShiftFailed) {
    char msg[280];
    std::snprintf(msg, sizeof msg, "Method A MINRES did not converge: %s (%d calls, %d iterations)", o.message,
                  o.calls, o.iterations);
    throw Error(FAME6D_ENOCONV, msg);
  }
  if (o.end == kShiftPlateau) acc.plateau_accepts += 1;
  acc.max_relres = std::max(acc.max_relres, o.relres);
}

static void mr_alloc(fame6d_mass* m, DevBuf& vec, DevBuf& aux, MrWs& W) {
  const size_t n3 = (size_t)3 * m->NF;
  constexpr int kSlots = 14;
  vec.ensure(sizeof(cpx<double>) * n3 * kSlots);
  cpx<double>* b0 = vec.as<cpx<double>>();
  cpx<double>** slots[kSlots] = {&W.x, &W.v,  &W.vn, &W.Mv, &W.r1,  &W.r2,  &W.y,
                                 &W.z, &W.w,  &W.w1, &W.w2, &W.res, &W.dx, &W.rc};
  for (int i = 0; i < kSlots; ++i) *slots[i] = b0 + n3 * i;
  const int nb = red_blocks((long long)n3);
  aux.ensure(sizeof(MrState) + sizeof(double) * ((size_t)nb + 64));
  W.st = aux.as<MrState>();
  W.part = reinterpret_cast<double*>(reinterpret_cast<char*>(aux.ptr) + sizeof(MrState));
  W.red = W.part + nb;
}

static void mr_stats_out(const MrAcc& a, fame6d_minres_stats_t* o) {
  if (!o) return;
  o->solves = a.solves;
  o->iterations = a.iterations;
  o->segments = a.segments;
  o->restarts = a.restarts;
  o->max_iterations = a.max_iterations;
  o->plateau_accepts = a.plateau_accepts;
  o->corrections = a.corrections;
  o->correction_iterations = a.correction_iterations;
  o->max_relres = a.max_relres;
  o->seconds = a.seconds;
}

// y = T x = (K - sigma M)^{-1} M x  (the reference's Top, make_A_eigs_top_preconditioned.m:73-113)
void method_a_apply(fame6d_qsys* S, double sigma, double tol, int maxit, const cpx<double>* x, cpx<double>* y,
                    fame6d_minres_stats_t* stats) {
  fame6d_mass* m = S->m;
  const double tau = sigma * m->eps_mean;
  if (!(sigma > 0) || !(tau > 0)) throw Error(FAME6D_EARG, "sigma and tau must be positive");
  DevBuf vec, aux, rhs;
  MrWs W{};
  mr_alloc(m, vec, aux, W);
  rhs.ensure(sizeof(cpx<double>) * 3 * (size_t)m->NF);
  mass3(m, x, rhs.as<cpx<double>>(), no_skip(), m->stream);
  MrAcc acc;
  int hint = 0;
  minres_solve(S, sigma, tau, rhs.as<cpx<double>>(), tol, maxit, W, hint, acc);
  F6_HIP(hipMemcpyAsync(y, W.x, sizeof(cpx<double>) * 3 * m->NF, hipMemcpyDeviceToDevice, m->stream));
  F6_HIP(hipStreamSynchronize(m->stream));
  mr_stats_out(acc, stats);
}

void eigs_method_a(fame6d_qsys* S, int k0, int p, double tol, int maxRestart, double sigma, double tol_mr,
                   int maxit_mr, const cpx<double>* v0, double* lambda, cpx<double>* Y,
                   fame6d_lanczos_stats_t* stats, fame6d_minres_stats_t* mstats) {
  fame6d_mass* m = S->m;
  hipStream_t s = m->stream;
  const long long n3 = 3 * m->NF;
  if (k0 < 1 || p <= k0 || p > 64) throw Error(FAME6D_EARG, "need 1 <= k < p <= 64");
  const double tau = sigma * m->eps_mean;
  if (!(sigma > 0) || !(tau > 0)) throw Error(FAME6D_EARG, "sigma and tau must be positive");
  const auto t0 = std::chrono::steady_clock::now();
  const int gr = grid_for(n3);
  const int nb3 = red_blocks(n3);
  DevBuf basis, zb, scal, part, ubuf, vec, aux, wout;
  basis.ensure(sizeof(cpx<double>) * (size_t)n3 * (p + 2));  // Q(:,0..p-1), v, r
  zb.ensure(sizeof(cpx<double>) * (size_t)n3 * (p + 2));     // Z(:,0..p-1), zv, zr
  scal.ensure(sizeof(double) * kScStride);
  part.ensure(sizeof(double) * (size_t)std::max(nb3 * 2 * kLzMaxP, 2048));
  wout.ensure(sizeof(cpx<double>) * (size_t)n3);
  MrWs W{};
  mr_alloc(m, vec, aux, W);
  LzState L;
  L.B.Q = basis.as<cpx<double>>();
  L.B.v = L.B.Q + (size_t)n3 * p;
  L.B.r = L.B.Q + (size_t)n3 * (p + 1);
  L.B.sc = scal.as<double>();
  L.B.part = part.as<double>();
  L.Z = zb.as<cpx<double>>();
  L.k = k0;
  cpx<double>* zv = L.Z + (size_t)n3 * p;
  cpx<double>* zr = L.Z + (size_t)n3 * (p + 1);
  Basis BZ = L.B;
  BZ.Q = L.Z;
  const int gn = (int)std::min<long long>(2048, (n3 + 255) / 256);

  // v = v0 / ||v0||_M, zv = M v  (run_method_A_fig1a.m:12-14 normalises in the 2-norm; the
  // M-norm normalisation here is the same direction)
  F6_HIP(hipMemcpyAsync(L.B.v, v0, sizeof(cpx<double>) * n3, hipMemcpyDefault, s));
  mass3(m, L.B.v, zv, no_skip(), s);
  k_dot_partials<double><<<dim3(nb3, 1), kRedThreads, 0, s>>>(L.B.v, n3, zv, n3, n3, L.B.part, nullptr);
  const double b0 = reduce_read(m, L.B.part, nb3, L.B.sc + 1);
  if (!(std::isfinite(b0) && b0 > 0)) throw Error(FAME6D_EARG, "Invalid Lanczos start vector.");
  k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n3, L.B.v, L.B.v, L.B.sc + 1);
  k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n3, zv, zv, L.B.sc + 1);
  F6_CHECK_LAUNCH();

  MrAcc acc;
  int hint = 0;
  while (!L.done) {
    const int jj = L.jj;
    F6_HIP(hipMemcpyAsync(L.B.Q + (size_t)n3 * (jj - 1), L.B.v, sizeof(cpx<double>) * n3,
                          hipMemcpyDeviceToDevice, s));
    F6_HIP(hipMemcpyAsync(L.Z + (size_t)n3 * (jj - 1), zv, sizeof(cpx<double>) * n3, hipMemcpyDeviceToDevice,
                          s));
    // w = (K - sigma M)^{-1} (M q_j)
    minres_solve(S, sigma, tau, zv, tol_mr, maxit_mr, W, hint, acc);
    F6_HIP(hipMemcpyAsync(wout.ptr, W.x, sizeof(cpx<double>) * n3, hipMemcpyDeviceToDevice, s));
    const cpx<double>* w = wout.as<cpx<double>>();
    // classical Gram-Schmidt twice in the M inner product: h = Z^H w, r = w - Q h
    gemv_c(m, n3, BZ, jj, w, L.B.sc + kScH, L.B.sc + kScHim, s);
    k_gemv_n_out<<<gn, 256, 0, s>>>(L.B.Q, n3, jj, L.B.sc + kScH, L.B.sc + kScHim, w, L.B.r, n3, nullptr,
                                    no_skip());
    gemv_c(m, n3, BZ, jj, L.B.r, L.B.sc + kScH2, L.B.sc + kScH2im, s);
    k_gemv_n_out<<<gn, 256, 0, s>>>(L.B.Q, n3, jj, L.B.sc + kScH2, L.B.sc + kScH2im, L.B.r, L.B.r, n3,
                                    nullptr, no_skip());
    // beta^2 = r^H M r
    mass3(m, L.B.r, zr, no_skip(), s);
    k_dot_partials<double><<<dim3(nb3, 1), kRedThreads, 0, s>>>(L.B.r, n3, zr, n3, n3, L.B.part, nullptr);
    k_reduce_final<<<1, kRedThreads, 0, s>>>(L.B.part, nb3, L.B.sc + 1);
    F6_CHECK_LAUNCH();
    std::vector<double> sc(kScH + jj);
    d2h_sync(m, sc.data(), L.B.sc, sizeof(double) * (kScH + jj));
    const double alpha = sc[kScH + jj - 1], beta2 = sc[1];
    if (!std::isfinite(alpha) || !std::isfinite(beta2))
      throw Error(FAME6D_ENONFINITE, "Method A operator returned invalid data.");
    L.justRestarted = false;
    bool end_cycle = false;
    if (beta2 <= 0 || std::sqrt(beta2) < 1e-14) {
      L.beta = std::sqrt(std::max(beta2, 0.0));
      L.breakdown = true;
      end_cycle = true;
    } else {
      L.beta = std::sqrt(beta2);
      k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n3, L.B.v, L.B.r, L.B.sc + 1);
      k_scale_rsqrt_dev<<<gr, 256, 0, s>>>(n3, zv, zr, L.B.sc + 1);
      F6_CHECK_LAUNCH();
      L.Alpha.push_back(alpha);
      L.Beta.push_back(L.beta);
      L.kk += 1;
      if (L.jj >= p) end_cycle = true;
      else L.jj += 1;
    }
    if (end_cycle && restart_step(L, k0, p, tol, maxRestart, n3, ubuf, s)) L.done = true;
  }
  if ((int)L.d.size() < k0) throw Error(FAME6D_ENOCONV, "Method A ended without k Ritz values.");
  // lambda = sigma + 1/theta, ascending (run_method_A_fig1a.m:47-49)
  std::vector<std::pair<double, int>> lam(k0);
  for (int i = 0; i < k0; ++i) {
    const double th = L.d[i];
    if (!std::isfinite(th) || th == 0) throw Error(FAME6D_ENOCONV, "Method A invalid Ritz values.");
    lam[i] = {sigma + 1.0 / th, i};
  }
  std::stable_sort(lam.begin(), lam.end());
  for (int i = 0; i < k0; ++i) lambda[i] = lam[i].first;
  if (Y) {
    const int mq = L.qcols;
    std::vector<double> Uy((size_t)mq * k0);
    for (int i = 0; i < k0; ++i) {
      const int col = L.ind[lam[i].second];
      for (int r = 0; r < mq; ++r) Uy[(size_t)i * mq + r] = L.Us[(size_t)col * mq + r];
    }
    ubuf.ensure(sizeof(double) * Uy.size());
    F6_HIP(hipMemcpyAsync(ubuf.ptr, Uy.data(), sizeof(double) * Uy.size(), hipMemcpyHostToDevice, s));
    restart_in_place(L.B.Q, n3, mq, k0, ubuf.as<double>(), s);
    F6_HIP(hipMemcpyAsync(Y, L.B.Q, sizeof(cpx<double>) * n3 * k0, hipMemcpyDeviceToDevice, s));
  }
  F6_HIP(hipStreamSynchronize(s));
  if (stats) {
    *stats = fame6d_lanczos_stats_t{};
    stats->converged = L.converged ? 1 : 0;
    stats->num_converged = L.nconv;
    stats->outer_steps = L.kk;
    stats->restarts = L.mm;
    stats->hit_max_restart = L.hitmax ? 1 : 0;
    stats->max_ritz_residual = L.maxres;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  mr_stats_out(acc, mstats);
}

}  // namespace fame6d
