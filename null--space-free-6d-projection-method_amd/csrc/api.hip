// C ABI of libfame6d (declared in include/fame6d.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

void eigs_method_c(fame6d_qsys* S, int precision, int k0, int p, double tol, int maxRestart,
                   double tol_cg, int maxit_cg, const cpx<double>* v0, double* lambda, cpx<double>* Y,
                   fame6d_lanczos_stats_t* stats, const RitzSelect* sel = nullptr);
void krinv_nested(fame6d_qsys* S, const cpx<double>* bin, cpx<double>* xout, double tolKr, double tolM,
                  int maxitKr, int maxitM, double epsC, fame6d_krinv_stats_t* st);
void set_nested(bool on, double tolM, int maxitM, double epsC);
void method_a_apply(fame6d_qsys* S, double sigma, double tol, int maxit, const cpx<double>* x, cpx<double>* y,
                    fame6d_minres_stats_t* stats);
void eigs_method_a(fame6d_qsys* S, int k0, int p, double tol, int maxRestart, double sigma, double tol_mr,
                   int maxit_mr, const cpx<double>* v0, double* lambda, cpx<double>* Y,
                   fame6d_lanczos_stats_t* stats, fame6d_minres_stats_t* mstats);
void recover_eigenvectors(fame6d_qsys* S, const cpx<double>* Y, int k, double tol, int maxit,
                          cpx<double>* X, cpx<double>* MX, fame6d_recover_stats_t* stats);
void fourier_diagnostics(fame6d_qsys* S, const cpx<double>* X, const cpx<double>* MX,
                         const double* lambda, int k, double* res, double* div, double* ray);

// geps[j] += sum_g w[g] * Re(conj(a_g[j]) * b_g[j])   (real grid vector, groups of length n)
__global__ void k_eps_grad(long long n, int G, const cpx<double>* a, const cpx<double>* b, long long ld,
                           const double* w, double* geps) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < G; ++g) {
      const cpx<double> u = a[g * ld + i], v = b[g * ld + i];
      s += w[g] * (u.re * v.re + u.im * v.im);
    }
    geps[i] += s;
  }
}

// Unnormalised inverse 6D DFT of G grid vectors: t = F^{-1}_unnorm x (passes A and B mode 1).
static void ifft6_unnorm(fame6d_mass* m, const cpx<double>* x, long long ldx, cpx<double>* t, int G) {
  PassAArgs<double> a{};
  a.x = x;
  a.ldx = ldx;
  a.t = t;
  a.ldt = m->NF;
  launch_pass_a<double>(m, PRE_COPY, G, a, m->stream);
  launch_pass_b<double>(m, 1, 1, G, t, m->NF, no_skip(), m->stream);
}

template <class T>
static void cols12(fame6d_mass* m, const cpx<T>* x, cpx<T>* Y, int nrhs) {
  const long long NF = m->NF;
  for (int j = 0; j < nrhs; ++j) {
    mass_apply<T>(m, x + 2 * NF * j, NF, Y + 3 * NF * j, NF, 2, m->stream);
    k_fill_zero<T><<<grid_for(NF), 256, 0, m->stream>>>(NF, Y + 3 * NF * j + 2 * NF);
  }
  F6_CHECK_LAUNCH();
}

template <class T>
static void col3(fame6d_mass* m, const cpx<T>* x, cpx<T>* Y, int nrhs) {
  const long long NF = m->NF;
  if (nrhs <= 0) return;
  for (int j = 0; j < nrhs; ++j)
    k_fill_zero<T><<<grid_for(2 * NF), 256, 0, m->stream>>>(2 * NF, Y + 3 * NF * j);
  F6_CHECK_LAUNCH();
  mass_apply<T>(m, x, NF, Y + 2 * NF, 3 * NF, nrhs, m->stream);
}

__global__ void k_apply_k_api(const double* q1, const double* q2, const double* q3, const cpx<double>* x,
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

// Times a callable `reps` times on the mass stream (after one warm-up call).
template <class F>
static float time_ms(fame6d_mass* m, hipEvent_t e0, hipEvent_t e1, int reps, F&& f) {
  f();
  F6_HIP(hipStreamSynchronize(m->stream));
  F6_HIP(hipEventRecord(e0, m->stream));
  for (int r = 0; r < reps; ++r) f();
  F6_HIP(hipEventRecord(e1, m->stream));
  F6_HIP(hipEventSynchronize(e1));
  float ms = 0;
  F6_HIP(hipEventElapsedTime(&ms, e0, e1));
  return ms;
}

constexpr int kTuneGroups = 8;  // groups of the several-group tile choice's timing

template <class T>
static void autotune_t(fame6d_mass* m, int reps) {
  const long long NF = m->NF;
  TileChoice& tc = std::is_same<T, float>::value ? m->tile32 : m->tile64;
  DevBuf vec, wts;
  vec.ensure(sizeof(cpx<T>) * NF * 4);
  wts.ensure(sizeof(T) * NF * 3);
  cpx<T>* x = vec.as<cpx<T>>();
  k_fill_zero<T><<<grid_for(NF * 4), 256, 0, m->stream>>>(NF * 4, x);
  k_fill_zero_real<T><<<grid_for(NF * 3), 256, 0, m->stream>>>(NF * 3, wts.as<T>());
  F6_CHECK_LAUNCH();
  hipEvent_t e0, e1;
  F6_HIP(hipEventCreate(&e0));
  F6_HIP(hipEventCreate(&e1));
  // 1) Mhat (three scalar applies): pass-B tile and the thread-split pass A/C tile
  DevBuf& w = work_buf<T>(m);
  w.ensure(sizeof(cpx<T>) * NF * 3);
  cpx<T>* t = w.as<cpx<T>>();
  PassAArgs<T> pa{};
  pa.x = x;
  pa.ldx = NF;
  pa.t = t;
  pa.ldt = NF;
  pa.W1 = wts.as<T>();
  pa.W2 = wts.as<T>() + NF;
  pa.invC = wts.as<T>() + 2 * NF;
  pa.ldq = NF;
  PassCArgs<T> pc{};
  pc.t = t;
  pc.ldt = NF;
  pc.y = x + NF;
  pc.ldy = NF;
  pc.W1 = pa.W1;
  pc.W2 = pa.W2;
  pc.invC = pa.invC;
  pc.ldq = NF;
  // Every candidate is timed in kRounds interleaved rounds and keeps its fastest round,
  // so that a transient slowdown of one round cannot decide the choice.
  constexpr int kRounds = 5;
  auto pick = [&](std::vector<TileChoice> cand, auto&& run) {
    std::vector<float> best(cand.size(), 1e30f);
    for (int round = 0; round < kRounds; ++round)
      for (size_t c = 0; c < cand.size(); ++c) {
        tc = cand[c];
        best[c] = std::min(best[c], time_ms(m, e0, e1, reps, run));
      }
    size_t arg = 0;
    for (size_t c = 1; c < cand.size(); ++c)
      if (best[c] < best[arg]) arg = c;
    tc = cand[arg];
  };
  auto mhat = [&]() {
    launch_pass_a<T>(m, PRE_MHAT, 1, pa, m->stream);
    launch_pass_b<T>(m, 3, 0, 1, t, NF, no_skip(), m->stream);
    launch_pass_c<T>(m, POST_MHAT, 1, pc, m->stream);
  };
  std::vector<TileChoice> cand;
  for (int vb = 0; vb < 2; ++vb)
    for (int v3 = 0; v3 < 2; ++v3) {
      TileChoice c = tc;
      c.b = vb;
      c.a3 = v3;
      c.c3 = v3;
      c.c3p = 0;
      c.a3p = 0;
      c.a3s = 0;
      cand.push_back(c);
    }
  pick(cand, mhat);
  // 1b, 1c) the Mhat passes A and C on their own terms (tile, persistent workgroups per CU),
  // each timed alone in the conjugate-gradient form the solves launch: pass A forms
  // p = r + beta p, pass C takes alpha from pass-B partials and updates x and r.  The state
  // and partials are stand-ins (iter 1, beta 1/2, rho 1, unit partials) so that no branch
  // is skipped; the vectors are zero.
  CgState hst{};
  hst.iter = 1;
  hst.beta = 0.5;
  hst.rho = 1.0;
  DevBuf cgs, pqp, cpart;
  cgs.ensure(sizeof(CgState));
  F6_HIP(hipMemcpyAsync(cgs.ptr, &hst, sizeof(CgState), hipMemcpyHostToDevice, m->stream));
  const int npq = 3 * pass_b_blocks(m, std::is_same<T, float>::value, 1);
  const std::vector<double> ones(npq, 1.0);
  pqp.ensure(sizeof(double) * npq);
  F6_HIP(hipMemcpyAsync(pqp.ptr, ones.data(), sizeof(double) * npq, hipMemcpyHostToDevice, m->stream));
  cpart.ensure(sizeof(double) * (size_t)m->n * m->n * m->n);  // one partial per S-block at most
  PassAArgs<T> pa_cg = pa;
  pa_cg.p = x + NF;
  pa_cg.st = cgs.as<CgState>();
  PassCArgs<T> pc_cg = pc;
  pc_cg.y = nullptr;
  pc_cg.cg = cgs.as<CgState>();
  pc_cg.pq_partials = pqp.as<double>();
  pc_cg.npq = npq;
  pc_cg.x = x + 2 * NF;
  pc_cg.p = x + NF;
  pc_cg.r = x + 3 * NF;
  pc_cg.ldx = NF;
  pc_cg.partials = cpart.as<double>();
  F6_HIP(hipStreamSynchronize(m->stream));
  cand.clear();
  for (int v3 = 0; v3 < 2; ++v3)
    for (int per = 0; per < 4; ++per)
      for (int sh = 0; sh < (per > 0 ? 2 : 1); ++sh) {
        TileChoice c = tc;
        c.a3 = v3;
        c.a3p = per;
        c.a3s = sh;
        cand.push_back(c);
      }
  pick(cand, [&]() { launch_pass_a<T>(m, PRE_MHAT_CG, 1, pa_cg, m->stream); });
  cand.clear();
  for (int v3 = 0; v3 < 2; ++v3)
    for (int per = 0; per < 4; ++per) {
      TileChoice c = tc;
      c.c3 = v3;
      c.c3p = per;
      cand.push_back(c);
    }
  pick(cand, [&]() { launch_pass_c<T>(m, POST_MHAT_CG, 1, pc_cg, m->stream); });
  cgs.release();
  pqp.release();
  cpart.release();
  // 2) plain applies: pass A/C tile and I/O form and the one-vector pass-B tile, timed on one
  // and on two right-hand sides (the API's single applies and the pairs around each K_r^{-1})
  cand.clear();
  for (int vio = 0; vio < 2; ++vio)
    for (int va = 0; va < 2; ++va)
      for (int vb1 = 0; vb1 < 2; ++vb1) {
        TileChoice c = tc;
        c.io = vio;
        c.a = va;
        c.b1 = vb1;
        cand.push_back(c);
      }
  pick(cand, [&]() {
    mass_apply<T>(m, x, NF, x + 2 * NF, NF, 2, m->stream);
    mass_apply<T>(m, x, NF, x + 2 * NF, NF, 1, m->stream);
  });
  // 3) launches over several groups (batched Bloch points): one tile per workgroup, tuned on
  // kTuneGroups groups, every pass-A/C blocks-per-workgroup pair with either pass-B tile for the
  // Mhat passes, then the plain passes on 2 kTuneGroups right-hand sides.  The groups share
  // q-system slot 0 (zero vectors and weights, as above).
  {
    constexpr int GT = kTuneGroups;
    TileChoice& tg = std::is_same<T, float>::value ? m->tile32g : m->tile64g;
    DevBuf gvec, gmap;
    gvec.ensure(sizeof(cpx<T>) * NF * 4 * GT);
    gmap.ensure(sizeof(int) * GT);
    cpx<T>* gx = gvec.as<cpx<T>>();
    k_fill_zero<T><<<grid_for(NF * 4 * GT), 256, 0, m->stream>>>(NF * 4 * GT, gx);
    F6_HIP(hipMemsetAsync(gmap.ptr, 0, sizeof(int) * GT, m->stream));
    F6_CHECK_LAUNCH();
    w.ensure(sizeof(cpx<T>) * NF * 3 * GT);
    PassAArgs<T> ga = pa;
    ga.x = gx;
    ga.t = w.as<cpx<T>>();
    ga.qmap = gmap.as<int>();
    PassCArgs<T> gc = pc;
    gc.t = ga.t;
    gc.y = gx + NF * GT;
    gc.qmap = ga.qmap;
    auto pick_g = [&](std::vector<TileChoice> cnd, auto&& run) {
      std::vector<float> best(cnd.size(), 1e30f);
      for (int round = 0; round < kRounds; ++round)
        for (size_t c = 0; c < cnd.size(); ++c) {
          tg = cnd[c];
          best[c] = std::min(best[c], time_ms(m, e0, e1, reps, run));
        }
      size_t arg = 0;
      for (size_t c = 1; c < cnd.size(); ++c)
        if (best[c] < best[arg]) arg = c;
      tg = cnd[arg];
    };
    std::vector<TileChoice> cg;
    for (int vb = 0; vb < 2; ++vb)
      for (int va3 = 0; va3 < 2; ++va3)
        for (int vc3 = 0; vc3 < 2; ++vc3) {
          TileChoice c = tc;
          c.b = vb;
          c.a3 = va3;
          c.c3 = vc3;
          c.a3p = c.c3p = c.a3s = 0;
          cg.push_back(c);
        }
    pick_g(cg, [&]() {
      launch_pass_a<T>(m, PRE_MHAT, GT, ga, m->stream);
      launch_pass_b<T>(m, 3, 0, GT, ga.t, NF, no_skip(), m->stream);
      launch_pass_c<T>(m, POST_MHAT, GT, gc, m->stream);
    });
    cg.clear();
    for (int vio = 0; vio < 2; ++vio)
      for (int va = 0; va < 2; ++va)
        for (int vb1 = 0; vb1 < 2; ++vb1) {
          TileChoice c = tg;
          c.io = vio;
          c.a = va;
          c.b1 = vb1;
          cg.push_back(c);
        }
    pick_g(cg, [&]() { mass_apply<T>(m, gx, NF, gx + 2 * NF * GT, NF, 2 * GT, m->stream); });
    F6_HIP(hipStreamSynchronize(m->stream));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  vec.release();
  wts.release();
}

}  // namespace fame6d

using namespace fame6d;

#define F6_API_BEGIN try {
#define F6_API_END                              \
  return FAME6D_OK;                             \
  }                                             \
  catch (const Error& e) {                      \
    set_last_error(e.what());                   \
    return e.code;                              \
  }                                             \
  catch (const std::exception& e) {             \
    set_last_error(e.what());                   \
    return FAME6D_EARG;                         \
  }

#define F6_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) throw Error(code, msg); \
  } while (0)

extern "C" {

const char* fame6d_last_error(void) { return g_last_error.c_str(); }
int fame6d_version(void) { return 10000; }

// ---------------------------------------------------------------- allocation-free apply
// Pass intermediates of fame6d_mass_apply_scalar_noalloc: a fixed array of the library, one
// copy per GPU, held until the process ends.  Host code never uses the name itself (on the
// host it is only a stand-in for the symbol); it uses the device address found once per
// device by hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {nullptr};
static hipError_t g_scratch_status[kMaxDevices] = {hipSuccess};
static bool g_scratch_done[kMaxDevices] = {false};

// Device address of `scratch` on `device` (the current device), looked up on the first call
// per device.  Returns nullptr and sets *status on failure.
static void* scratch_address(int device, hipError_t* status) {
  if (device < 0 || device >= kMaxDevices) {
    *status = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_status[device] = e;
    g_scratch_ptr[device] = (e == hipSuccess) ? p : nullptr;
    g_scratch_done[device] = true;
  }
  *status = g_scratch_status[device];
  return g_scratch_ptr[device];
}

// Per-workgroup min, max and sum of the permittivity samples, and a non-finite count
// (block b writes stats[4b .. 4b+3]); the host combines the kEpsBlocks partials.
constexpr int kEpsBlocks = 256;
__global__ void __launch_bounds__(256) k_eps_stats(long long n, const double* __restrict__ eps,
                                                   double* __restrict__ stats) {
  double mn = INFINITY, mx = -INFINITY, sum = 0.0, bad = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
    const double e = eps[i];
    if (!isfinite(e)) {
      bad += 1.0;
    } else {
      mn = fmin(mn, e);
      mx = fmax(mx, e);
      sum += e;
    }
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = mn;
  red[1][threadIdx.x] = mx;
  red[2][threadIdx.x] = sum;
  red[3][threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + w]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
      red[2][threadIdx.x] += red[2][threadIdx.x + w];
      red[3][threadIdx.x] += red[3][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) stats[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The four operator tables of the samples: eps/NF and 1/(eps NF), FP64 and FP32.
__global__ void k_eps_tables(long long n, const double* __restrict__ eps, double* __restrict__ e64,
                             float* __restrict__ e32, double* __restrict__ i64, float* __restrict__ i32) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double e = eps[i];
    const double a = e / (double)n;
    const double b = 1.0 / (e * (double)n);
    e64[i] = a;
    e32[i] = (float)a;
    i64[i] = b;
    i32[i] = (float)b;
  }
}

static void check_eps_stats(fame6d_mass* m, double mn, double mx, double sum, bool finite) {
  if (!finite) throw Error(FAME6D_ENONFINITE, "permittivity contains NaN/Inf");
  if (!(mn > 0))
    throw Error(FAME6D_EARG, "permittivity must be positive: M_s = F diag(eps) F^{-1} is then Hermitian "
                             "positive definite, which the PCG solves require");
  m->eps_min = mn;
  m->eps_max = mx;
  m->eps_mean = sum / (double)m->NF;
}

// Device form of upload_eps: the samples are validated and the tables formed on the
// device; only kEpsBlocks x 4 partial statistics cross to the host.
static void upload_eps_device(fame6d_mass* m, const double* eps_dev) {
  const long long NF = m->NF;
  m->eps64.ensure(sizeof(double) * NF);
  m->eps32.ensure(sizeof(float) * NF);
  m->ieps64.ensure(sizeof(double) * NF);
  m->ieps32.ensure(sizeof(float) * NF);
  m->scal.ensure(std::max<size_t>(m->scal.bytes, sizeof(double) * 4 * kEpsBlocks));
  k_eps_stats<<<kEpsBlocks, 256, 0, m->stream>>>(NF, eps_dev, m->scal.as<double>());
  F6_CHECK_LAUNCH();
  std::vector<double> st(4 * kEpsBlocks);
  d2h_sync(m, st.data(), m->scal.ptr, sizeof(double) * 4 * kEpsBlocks);
  double mn = INFINITY, mx = -INFINITY, sum = 0, bad = 0;
  for (int b = 0; b < kEpsBlocks; ++b) {
    mn = std::min(mn, st[4 * b]);
    mx = std::max(mx, st[4 * b + 1]);
    sum += st[4 * b + 2];
    bad += st[4 * b + 3];
  }
  check_eps_stats(m, mn, mx, sum, bad == 0);
  k_eps_tables<<<grid_for(NF), 256, 0, m->stream>>>(NF, eps_dev, m->eps64.as<double>(), m->eps32.as<float>(),
                                                    m->ieps64.as<double>(), m->ieps32.as<float>());
  F6_CHECK_LAUNCH();
}

static void upload_eps(fame6d_mass* m, const double* eps_host) {
  const long long NF = m->NF;
  double mn = INFINITY, mx = -INFINITY, sum = 0;
  for (long long i = 0; i < NF; ++i) {
    const double e = eps_host[i];
    if (!std::isfinite(e)) throw Error(FAME6D_ENONFINITE, "permittivity contains NaN/Inf");
    mn = std::min(mn, e);
    mx = std::max(mx, e);
    sum += e;
  }
  check_eps_stats(m, mn, mx, sum, true);
  std::vector<double> e64(NF), i64(NF);
  std::vector<float> e32(NF), i32(NF);
  for (long long i = 0; i < NF; ++i) {
    e64[i] = eps_host[i] / (double)NF;
    e32[i] = (float)e64[i];
    i64[i] = 1.0 / (eps_host[i] * (double)NF);
    i32[i] = (float)i64[i];
  }
  m->eps64.ensure(sizeof(double) * NF);
  m->eps32.ensure(sizeof(float) * NF);
  m->ieps64.ensure(sizeof(double) * NF);
  m->ieps32.ensure(sizeof(float) * NF);
  F6_HIP(hipMemcpyAsync(m->eps64.ptr, e64.data(), sizeof(double) * NF, hipMemcpyHostToDevice, m->stream));
  F6_HIP(hipMemcpyAsync(m->eps32.ptr, e32.data(), sizeof(float) * NF, hipMemcpyHostToDevice, m->stream));
  F6_HIP(hipMemcpyAsync(m->ieps64.ptr, i64.data(), sizeof(double) * NF, hipMemcpyHostToDevice, m->stream));
  F6_HIP(hipMemcpyAsync(m->ieps32.ptr, i32.data(), sizeof(float) * NF, hipMemcpyHostToDevice, m->stream));
  F6_HIP(hipStreamSynchronize(m->stream));
}

int fame6d_mass_create(int N, const double* eps_host, void* stream, fame6d_mass** out) {
  F6_API_BEGIN
  F6_REQUIRE(out != nullptr && eps_host != nullptr, FAME6D_EARG, "null argument");
  *out = nullptr;
  F6_REQUIRE(N >= 1 && N <= 6, FAME6D_EARG, "N must be in 1..6 (nGrid = 2N in {2,...,12})");
  auto* m = new fame6d_mass();
  try {
    m->N = N;
    m->n = 2 * N;
    long long NF = 1;
    for (int i = 0; i < 6; ++i) NF *= m->n;
    m->NF = NF;
    m->stream = static_cast<hipStream_t>(stream);
    F6_HIP(hipGetDevice(&m->device));
    F6_HIP(hipDeviceGetAttribute(&m->cus, hipDeviceAttributeMultiprocessorCount, m->device));
    if (m->cus <= 0) throw Error(FAME6D_EHIP, "device reports no compute units");
    m->pin.init(1 << 16);
    {
      // the allocation-free apply's array on this device (the current one)
      hipError_t st = hipSuccess;
      m->noalloc_scratch = scratch_address(m->device, &st);
      m->noalloc_status = (int)st;
    }
    upload_eps(m, eps_host);
  } catch (...) {
    m->eps64.release();
    m->eps32.release();
    m->pin.release();
    delete m;
    throw;
  }
  *out = m;
  F6_API_END
}

int fame6d_mass_set_eps(fame6d_mass* m, const double* eps_host) {
  F6_API_BEGIN
  F6_REQUIRE(m && eps_host, FAME6D_EARG, "null argument");
  upload_eps(m, eps_host);
  F6_API_END
}

int fame6d_mass_set_eps_device(fame6d_mass* m, const double* eps_dev) {
  F6_API_BEGIN
  F6_REQUIRE(m && eps_dev, FAME6D_EARG, "null argument");
  upload_eps_device(m, eps_dev);
  F6_API_END
}

int fame6d_mass_destroy(fame6d_mass* m) {
  F6_API_BEGIN
  if (!m) return FAME6D_OK;
  (void)hipStreamSynchronize(m->stream);
  m->eps64.release();
  m->eps32.release();
  m->ieps64.release();
  m->ieps32.release();
  m->work64.release();
  m->work32.release();
  m->io64.release();
  m->partials.release();
  m->scal.release();
  m->timer.destroy();
  m->pin.release();
  delete m;
  F6_API_END
}

int fame6d_mass_info(const fame6d_mass* m, fame6d_mass_info_t* info) {
  F6_API_BEGIN
  F6_REQUIRE(m && info, FAME6D_EARG, "null argument");
  info->N = m->N;
  info->nGrid = m->n;
  info->NF = m->NF;
  info->eps_min = m->eps_min;
  info->eps_max = m->eps_max;
  info->eps_mean = m->eps_mean;
  int th;
  tile_variant_desc(m->n, false, false, m->tile64.a, &info->tile_a64, &th);
  tile_variant_desc(m->n, false, true, m->tile64.b, &info->tile_b64, &th);
  tile_variant_desc(m->n, true, false, m->tile32.a, &info->tile_a32, &th);
  tile_variant_desc(m->n, true, true, m->tile32.b, &info->tile_b32, &th);
  auto code = [](const TileChoice& t) {
    return (t.a & 1) | ((t.io == 1 ? 0 : 1) << 1) | ((t.a3 & 1) << 2) | ((t.c3p & 3) << 3) | ((t.c3 & 1) << 5) |
           ((t.a3p & 3) << 6) | ((t.a3s & 1) << 8);
  };
  info->variant_ac64 = code(m->tile64);
  info->variant_b64 = m->tile64.b | (m->tile64.b1 << 1);
  info->variant_ac32 = code(m->tile32);
  info->variant_b32 = m->tile32.b | (m->tile32.b1 << 1);
  F6_API_END
}

int fame6d_mass_set_tiles(fame6d_mass* m, int precision, int va, int vb) {
  F6_API_BEGIN
  F6_REQUIRE(m, FAME6D_EARG, "null handle");
  F6_REQUIRE(va >= 0 && va < 512 && vb >= 0 && vb < 4, FAME6D_EARG, "tile variant out of range");
  TileChoice& t = (precision == FAME6D_FP32) ? m->tile32 : m->tile64;
  t.a = va & 1;
  t.io = ((va >> 1) & 1) ? 0 : 1;
  t.a3 = (va >> 2) & 1;
  t.c3p = (va >> 3) & 3;
  t.c3 = (va >> 5) & 1;
  t.a3p = (va >> 6) & 3;
  t.a3s = (va >> 8) & 1;
  t.b = vb & 1;
  t.b1 = (vb >> 1) & 1;
  F6_API_END
}

int fame6d_mass_profile(fame6d_mass* m, int enable) {
  F6_API_BEGIN
  F6_REQUIRE(m, FAME6D_EARG, "null handle");
  F6_REQUIRE(enable >= 0, FAME6D_EARG, "profile stride must be >= 0");
  m->timer.enable(enable);
  F6_API_END
}

int fame6d_mass_kernel_times(fame6d_mass* m, fame6d_kernel_time_t* out) {
  F6_API_BEGIN
  F6_REQUIRE(m && out, FAME6D_EARG, "null argument");
  F6_HIP(hipStreamSynchronize(m->stream));
  m->timer.harvest();
  for (int i = 0; i < FAME6D_NUM_TIMER_SLOTS; ++i) {
    out[i].launches = m->timer.count[i];
    out[i].milliseconds = m->timer.ms[i];
    out[i].bytes = m->timer.bytes[i];
  }
  F6_API_END
}

int fame6d_mass_autotune(fame6d_mass* m, int precision, int reps) {
  F6_API_BEGIN
  F6_REQUIRE(m, FAME6D_EARG, "null handle");
  F6_REQUIRE(precision == FAME6D_FP64 || precision == FAME6D_FP32, FAME6D_EARG, "precision");
  if (reps < 1) reps = 1;
  if (precision == FAME6D_FP32) autotune_t<float>(m, reps);
  else autotune_t<double>(m, reps);
  F6_API_END
}

int fame6d_mass_apply_scalar(fame6d_mass* m, int precision, const void* x, void* y, int nrhs) {
  F6_API_BEGIN
  F6_REQUIRE(m && x && y, FAME6D_EARG, "null argument");
  F6_REQUIRE(nrhs >= 0, FAME6D_ESIZE, "nrhs must be >= 0");
  if (precision == FAME6D_FP64)
    mass_apply<double>(m, (const cpx<double>*)x, m->NF, (cpx<double>*)y, m->NF, nrhs, m->stream);
  else if (precision == FAME6D_FP32)
    mass_apply<float>(m, (const cpx<float>*)x, m->NF, (cpx<float>*)y, m->NF, nrhs, m->stream);
  else
    throw Error(FAME6D_EARG, "precision must be FAME6D_FP64 or FAME6D_FP32");
  F6_API_END
}

int fame6d_mass_apply_scalar_noalloc(fame6d_mass* m, int precision, const void* x, void* y, int nrhs) {
  F6_API_BEGIN
  F6_REQUIRE(m && x && y, FAME6D_EARG, "null argument");
  F6_REQUIRE(nrhs >= 0, FAME6D_ESIZE, "nrhs must be >= 0");
  F6_REQUIRE(precision == FAME6D_FP64 || precision == FAME6D_FP32, FAME6D_EARG,
             "precision must be FAME6D_FP64 or FAME6D_FP32");
  if (!m->noalloc_scratch)
    throw Error(FAME6D_EHIP, std::string("hipGetSymbolAddress of the allocation-free apply's array failed: ") +
                                 hipGetErrorString((hipError_t)m->noalloc_status));
  const size_t per = (precision == FAME6D_FP64 ? sizeof(cpx<double>) : sizeof(cpx<float>)) * (size_t)m->NF;
  if ((size_t)nrhs * per > kNoAllocScratchBytes)
    throw Error(FAME6D_ESIZE, "fame6d_mass_apply_scalar_noalloc: nrhs * NF * " + std::to_string(per / m->NF) +
                                  " bytes = " + std::to_string((size_t)nrhs * per) +
                                  " exceeds its fixed 64 MiB intermediate array (at most " +
                                  std::to_string(kNoAllocScratchBytes / per) + " columns at this N and precision)");
  if (precision == FAME6D_FP64)
    mass_apply_with<double>(m, (const cpx<double>*)x, m->NF, (cpx<double>*)y, m->NF, nrhs,
                            (cpx<double>*)m->noalloc_scratch, m->stream);
  else
    mass_apply_with<float>(m, (const cpx<float>*)x, m->NF, (cpx<float>*)y, m->NF, nrhs,
                           (cpx<float>*)m->noalloc_scratch, m->stream);
  F6_CHECK_LAUNCH();
  F6_API_END
}

int fame6d_mass_apply_3comp(fame6d_mass* m, int precision, const void* U, void* MU, int nrhs) {
  // [U1;U2;U3] columns are 3*nrhs contiguous NF blocks with the same layout in MU
  return fame6d_mass_apply_scalar(m, precision, U, MU, 3 * nrhs);
}

int fame6d_mass_apply_cols12(fame6d_mass* m, int precision, const void* x12, void* Y, int nrhs) {
  F6_API_BEGIN
  F6_REQUIRE(m && x12 && Y, FAME6D_EARG, "null argument");
  F6_REQUIRE(nrhs >= 0, FAME6D_ESIZE, "nrhs must be >= 0");
  F6_REQUIRE(x12 != Y, FAME6D_EARG, "x12 and Y must not alias");
  if (precision == FAME6D_FP64) cols12<double>(m, (const cpx<double>*)x12, (cpx<double>*)Y, nrhs);
  else if (precision == FAME6D_FP32) cols12<float>(m, (const cpx<float>*)x12, (cpx<float>*)Y, nrhs);
  else throw Error(FAME6D_EARG, "precision must be FAME6D_FP64 or FAME6D_FP32");
  F6_API_END
}

int fame6d_mass_apply_col3(fame6d_mass* m, int precision, const void* x, void* Y, int nrhs) {
  F6_API_BEGIN
  F6_REQUIRE(m && x && Y, FAME6D_EARG, "null argument");
  F6_REQUIRE(nrhs >= 0, FAME6D_ESIZE, "nrhs must be >= 0");
  F6_REQUIRE(x != Y, FAME6D_EARG, "x and Y must not alias");
  if (precision == FAME6D_FP64) col3<double>(m, (const cpx<double>*)x, (cpx<double>*)Y, nrhs);
  else if (precision == FAME6D_FP32) col3<float>(m, (const cpx<float>*)x, (cpx<float>*)Y, nrhs);
  else throw Error(FAME6D_EARG, "precision must be FAME6D_FP64 or FAME6D_FP32");
  F6_API_END
}

int fame6d_qsys_create(fame6d_mass* m, const double* q_host, const double* P_host, fame6d_qsys** out) {
  F6_API_BEGIN
  F6_REQUIRE(m && q_host && P_host && out, FAME6D_EARG, "null argument");
  *out = qsys_build(m, 1, q_host, P_host);
  F6_API_END
}

int fame6d_qsys_destroy(fame6d_qsys* s) {
  F6_API_BEGIN
  if (!s) return FAME6D_OK;
  (void)hipStreamSynchronize(s->m->stream);
  for (DevBuf* b : {&s->Vd, &s->Wd, &s->Wf, &s->Qd, &s->kr64, &s->kr32, &s->st, &s->part, &s->part2,
                    &s->ints, &s->sc, &s->ref, &s->kr64_1, &s->st_1, &s->part_1, &s->part2_1, &s->ints_1})
    b->release();
  if (s->cg_pin) (void)hipHostFree(s->cg_pin);
  if (s->inner_pin) (void)hipHostFree(s->inner_pin);
  delete s;
  F6_API_END
}

int fame6d_qsys_info(const fame6d_qsys* s, fame6d_qsys_info_t* info) {
  F6_API_BEGIN
  F6_REQUIRE(s && info, FAME6D_EARG, "null argument");
  info->NF = s->NF;
  info->min_abs_det_v1 = s->min_abs_det[0];
  info->min_qq = s->min_qq[0];
  info->n_singular = s->n_singular[0];
  for (int l = 0; l < 6; ++l) info->k_lift[l] = s->klift[l];
  F6_API_END
}

int fame6d_qsys_diagonals(const fame6d_qsys* s, double* V11, double* V12, double* V21, double* V22,
                          double* V31, double* V32, double* W1, double* W2, double* invC) {
  F6_API_BEGIN
  F6_REQUIRE(s, FAME6D_EARG, "null handle");
  const size_t bytes = sizeof(double) * s->NF;
  double* outs[6] = {V11, V12, V21, V22, V31, V32};
  for (int i = 0; i < 6; ++i)
    if (outs[i]) F6_HIP(hipMemcpy(outs[i], s->V(i), bytes, hipMemcpyDeviceToHost));
  double* wo[3] = {W1, W2, invC};
  for (int i = 0; i < 3; ++i)
    if (wo[i]) F6_HIP(hipMemcpy(wo[i], s->W64(i), bytes, hipMemcpyDeviceToHost));
  F6_API_END
}

int fame6d_mhat_apply(fame6d_qsys* s, int precision, const void* x, void* y) {
  F6_API_BEGIN
  F6_REQUIRE(s && x && y, FAME6D_EARG, "null argument");
  if (precision == FAME6D_FP64)
    mhat_apply<double>(s, 1, nullptr, (const cpx<double>*)x, (cpx<double>*)y, no_skip(), s->m->stream);
  else if (precision == FAME6D_FP32)
    mhat_apply<float>(s, 1, nullptr, (const cpx<float>*)x, (cpx<float>*)y, no_skip(), s->m->stream);
  else
    throw Error(FAME6D_EARG, "precision must be FAME6D_FP64 or FAME6D_FP32");
  F6_API_END
}

int fame6d_apply_v(fame6d_qsys* s, const void* x, void* y) {
  F6_API_BEGIN
  F6_REQUIRE(s && x && y, FAME6D_EARG, "null argument");
  const long long NF = s->NF;
  k_apply_v<<<dim3(grid_for(NF), 1), 256, 0, s->m->stream>>>((const cpx<double>*)x, 2 * NF, NF, s->V(0), s->V(1),
                                                              s->V(2), s->V(3), s->V(4), s->V(5),
                                                              (cpx<double>*)y, 3 * NF, no_skip(), nullptr);
  F6_CHECK_LAUNCH();
  F6_API_END
}

int fame6d_apply_k(fame6d_qsys* s, const void* x, void* y) {
  F6_API_BEGIN
  F6_REQUIRE(s && x && y && x != y, FAME6D_EARG, "null or aliased argument");
  k_apply_k_api<<<grid_for(s->NF), 256, 0, s->m->stream>>>(s->Qm(0), s->Qm(1), s->Qm(2),
                                                            (const cpx<double>*)x, (cpx<double>*)y, s->NF);
  F6_CHECK_LAUNCH();
  F6_API_END
}

int fame6d_krinv_apply(fame6d_qsys* s, int precision, const void* q, void* y, double tol, int maxit,
                       fame6d_krinv_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && q && y, FAME6D_EARG, "null argument");
  F6_REQUIRE(tol > 0 && maxit >= 0, FAME6D_EARG, "tol must be > 0 and maxit >= 0");
  const fame6d_krinv_stats_t st0 = s->stats;
  const int slot = 0;
  krinv_groups(s, precision, 1, &slot, (const cpx<double>*)q, (cpx<double>*)y, tol, maxit, nullptr);
  if (stats) {
    *stats = s->stats;
    stats->calls -= st0.calls;
    stats->total_iterations -= st0.total_iterations;
    stats->total_fp64_iterations -= st0.total_fp64_iterations;
    stats->refinement_cycles -= st0.refinement_cycles;
    stats->mass_applies_fp64 -= st0.mass_applies_fp64;
    stats->mass_applies_fp32 -= st0.mass_applies_fp32;
    stats->seconds -= st0.seconds;
  }
  F6_API_END
}

int fame6d_qsys_set_mixed(fame6d_qsys* s, double inner_tol, int max_cycles, int precondition) {
  F6_API_BEGIN
  F6_REQUIRE(s && inner_tol > 0 && inner_tol < 1 && max_cycles > 0, FAME6D_EARG, "bad mixed parameters");
  s->mixed_inner_tol = inner_tol;
  s->mixed_max_cycles = max_cycles;
  s->mixed_precond = precondition != 0;
  F6_API_END
}

int fame6d_qsys_set_reorth(fame6d_qsys* s, double eta) {
  F6_API_BEGIN
  F6_REQUIRE(s && eta > 0 && eta <= 1e3, FAME6D_EARG, "bad re-orthogonalisation threshold");
  s->dgks_eta2 = eta * eta;
  F6_API_END
}

int fame6d_eigs_method_c(fame6d_qsys* s, int precision, int k, int p, double tol, int max_restart,
                         double tol_cg, int maxit_cg, const void* v0, double* lambda_host, void* Y,
                         fame6d_lanczos_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && v0 && lambda_host, FAME6D_EARG, "null argument");
  F6_REQUIRE(s->G == 1, FAME6D_EARG, "single-point handle expected");
  eigs_method_c(s, precision, k, p, tol, max_restart, tol_cg, maxit_cg, (const cpx<double>*)v0,
                lambda_host, (cpx<double>*)Y, stats);
  F6_API_END
}

int fame6d_eigs_tracking(fame6d_qsys* s, int precision, int k, int p, double tol, int max_restart, double tol_cg,
                         int maxit_cg, const void* v0, int nfixed, const int* fixed_host, int ntarget,
                         const double* target_host, int half_width, double* lambda_host, double* rel_res_host,
                         int* sel_host, int sel_capacity, int* nsel, void* Y, fame6d_lanczos_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && v0 && lambda_host && sel_host && nsel, FAME6D_EARG, "null argument");
  F6_REQUIRE(s->G == 1, FAME6D_EARG, "single-point handle expected");
  F6_REQUIRE((nfixed > 0 && fixed_host) || (ntarget > 0 && target_host), FAME6D_EARG,
             "Either fixedIndices or targetLambda must be supplied.");
  F6_REQUIRE(half_width >= 0 && sel_capacity > 0, FAME6D_EARG, "bad selection sizes");
  RitzSelect sel;
  if (nfixed > 0 && fixed_host) {
    sel.idx = fixed_host;
    sel.nidx = nfixed;
  } else {
    sel.target = target_host;
    sel.ntarget = ntarget;
    sel.half_width = half_width;
  }
  sel.idx_out = sel_host;
  sel.idx_capacity = sel_capacity;
  sel.nidx_out = nsel;
  sel.rel_res = rel_res_host;
  eigs_method_c(s, precision, k, p, tol, max_restart, tol_cg, maxit_cg, (const cpx<double>*)v0, lambda_host,
                (cpx<double>*)Y, stats, &sel);
  F6_API_END
}

int fame6d_krinv_nested_apply(fame6d_qsys* s, const void* b, void* x, double tol_kr, double tol_m,
                               int maxit_kr, int maxit_m, double eps_c, fame6d_krinv_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && b && x, FAME6D_EARG, "null argument");
  F6_REQUIRE(tol_kr > 0 && tol_m > 0 && eps_c > 0, FAME6D_EARG, "tolerances and eps_c must be positive");
  const fame6d_krinv_stats_t st0 = s->stats;
  krinv_nested(s, (const cpx<double>*)b, (cpx<double>*)x, tol_kr, tol_m, maxit_kr, maxit_m, eps_c, nullptr);
  if (stats) {
    *stats = s->stats;
    stats->calls -= st0.calls;
    stats->total_iterations -= st0.total_iterations;
    stats->total_fp64_iterations -= st0.total_fp64_iterations;
    stats->mass_applies_fp64 -= st0.mass_applies_fp64;
    stats->seconds -= st0.seconds;
  }
  F6_API_END
}

int fame6d_eigs_method_b(fame6d_qsys* s, int k, int p, double tol, int max_restart, double tol_kr,
                         double tol_m, int maxit_kr, int maxit_m, double eps_c, const void* v0,
                         double* lambda_host, void* Y, fame6d_lanczos_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && v0 && lambda_host, FAME6D_EARG, "null argument");
  F6_REQUIRE(s->G == 1, FAME6D_EARG, "single-point handle expected");
  F6_REQUIRE(tol_m > 0 && eps_c > 0, FAME6D_EARG, "tol_m and eps_c must be positive");
  set_nested(true, tol_m, maxit_m, eps_c);
  try {
    eigs_method_c(s, FAME6D_FP64, k, p, tol, max_restart, tol_kr, maxit_kr, (const cpx<double>*)v0,
                  lambda_host, (cpx<double>*)Y, stats);
  } catch (...) {
    set_nested(false, 0, 0, 0);
    throw;
  }
  set_nested(false, 0, 0, 0);
  F6_API_END
}

int fame6d_method_a_apply(fame6d_qsys* s, double sigma, double tol, int maxit, const void* x, void* y,
                          fame6d_minres_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && x && y, FAME6D_EARG, "null argument");
  F6_REQUIRE(s->G == 1, FAME6D_EARG, "single-point handle expected");
  F6_REQUIRE(sigma > 0 && tol > 0 && maxit > 0, FAME6D_EARG, "sigma, tol and maxit must be positive");
  method_a_apply(s, sigma, tol, maxit, (const cpx<double>*)x, (cpx<double>*)y, stats);
  F6_API_END
}

int fame6d_eigs_method_a(fame6d_qsys* s, int k, int p, double tol, int max_restart, double sigma,
                         double tol_minres, int maxit_minres, const void* v0, double* lambda_host, void* Y,
                         fame6d_lanczos_stats_t* stats, fame6d_minres_stats_t* mstats) {
  F6_API_BEGIN
  F6_REQUIRE(s && v0 && lambda_host, FAME6D_EARG, "null argument");
  F6_REQUIRE(s->G == 1, FAME6D_EARG, "single-point handle expected");
  F6_REQUIRE(sigma > 0 && tol_minres > 0 && maxit_minres > 0, FAME6D_EARG, "bad Method-A parameters");
  eigs_method_a(s, k, p, tol, max_restart, sigma, tol_minres, maxit_minres, (const cpx<double>*)v0,
                lambda_host, (cpx<double>*)Y, stats, mstats);
  F6_API_END
}

int fame6d_eigs_method_c_batch(fame6d_mass* m, int nq, const double* q_host, const double* P_host,
                               int precision, int k, int p, double tol, int max_restart, double tol_cg,
                               int maxit_cg, const void* v0, double* lambda_host, void* Y,
                               fame6d_lanczos_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(m && q_host && P_host && v0 && lambda_host, FAME6D_EARG, "null argument");
  F6_REQUIRE(nq >= 1 && nq <= 64, FAME6D_EARG, "nq must be in 1..64");
  fame6d_qsys* S = qsys_build(m, nq, q_host, P_host);
  try {
    eigs_method_c(S, precision, k, p, tol, max_restart, tol_cg, maxit_cg, (const cpx<double>*)v0,
                  lambda_host, (cpx<double>*)Y, stats);
  } catch (...) {
    fame6d_qsys_destroy(S);
    throw;
  }
  fame6d_qsys_destroy(S);
  F6_API_END
}

int fame6d_recover_eigenvectors(fame6d_qsys* s, const void* Y, int k, double tol, int maxit, void* X,
                                void* MX, fame6d_recover_stats_t* stats) {
  F6_API_BEGIN
  F6_REQUIRE(s && Y && X && MX, FAME6D_EARG, "null argument");
  recover_eigenvectors(s, (const cpx<double>*)Y, k, tol, maxit, (cpx<double>*)X, (cpx<double>*)MX, stats);
  F6_API_END
}

int fame6d_fourier_diagnostics(fame6d_qsys* s, const void* X, const void* MX, const double* lambda_host,
                               int k, double* res_gevp_host, double* div_host, double* rayleigh_host) {
  F6_API_BEGIN
  F6_REQUIRE(s && X && MX && lambda_host, FAME6D_EARG, "null argument");
  fourier_diagnostics(s, (const cpx<double>*)X, (const cpx<double>*)MX, lambda_host, k, res_gevp_host,
                      div_host, rayleigh_host);
  F6_API_END
}

int fame6d_mass_adjoint_eps(fame6d_mass* m, const void* x, const void* gy, int nrhs, double* geps) {
  F6_API_BEGIN
  F6_REQUIRE(m && x && gy && geps, FAME6D_EARG, "null argument");
  F6_REQUIRE(nrhs >= 0, FAME6D_ESIZE, "nrhs must be >= 0");
  const long long NF = m->NF;
  DevBuf t;
  t.ensure(sizeof(cpx<double>) * NF * 2);
  DevBuf w;
  w.ensure(sizeof(double) * 8);
  // NF * Re(conj(ifftn gy) ifftn x) = (1/NF) Re(conj(F^-1_u gy) F^-1_u x)
  const double wv = 1.0 / (double)NF;
  F6_HIP(hipMemcpyAsync(w.ptr, &wv, sizeof(double), hipMemcpyHostToDevice, m->stream));
  for (int j = 0; j < nrhs; ++j) {
    ifft6_unnorm(m, (const cpx<double>*)gy + NF * j, NF, t.as<cpx<double>>(), 1);
    ifft6_unnorm(m, (const cpx<double>*)x + NF * j, NF, t.as<cpx<double>>() + NF, 1);
    k_eps_grad<<<grid_for(NF), 256, 0, m->stream>>>(NF, 1, t.as<cpx<double>>(), t.as<cpx<double>>() + NF, NF,
                                                    w.as<double>(), geps);
    F6_CHECK_LAUNCH();
  }
  F6_HIP(hipStreamSynchronize(m->stream));
  t.release();
  w.release();
  F6_API_END
}

int fame6d_eig_grad_eps(fame6d_mass* m, const void* X, int k, const double* lambda_host,
                        const double* glambda_host, double* geps) {
  F6_API_BEGIN
  F6_REQUIRE(m && X && lambda_host && glambda_host && geps, FAME6D_EARG, "null argument");
  const long long NF = m->NF;
  DevBuf t, w;
  t.ensure(sizeof(cpx<double>) * NF * 3);
  w.ensure(sizeof(double) * 3);
  // d lambda / d eps_j = -lambda * NF * sum_c |ifftn(x_c)_j|^2 = -(lambda/NF) sum_c |F^-1_u x_c|^2
  for (int j = 0; j < k; ++j) {
    const double c = -glambda_host[j] * lambda_host[j] / (double)NF;
    const double wv[3] = {c, c, c};
    F6_HIP(hipMemcpyAsync(w.ptr, wv, sizeof(wv), hipMemcpyHostToDevice, m->stream));
    ifft6_unnorm(m, (const cpx<double>*)X + 3 * NF * j, NF, t.as<cpx<double>>(), 3);
    k_eps_grad<<<grid_for(NF), 256, 0, m->stream>>>(NF, 3, t.as<cpx<double>>(), t.as<cpx<double>>(), NF,
                                                    w.as<double>(), geps);
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(m->stream));
  }
  t.release();
  w.release();
  F6_API_END
}

}  // extern "C"
