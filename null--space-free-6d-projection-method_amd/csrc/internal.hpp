// Library-internal declarations shared by the translation units of libfame6d.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"
#include "mass_kernels.hpp"
#include "../../include/fame6d.h"

namespace fame6d {

// Error carrying a FAME6D_* status code; converted to a return code at the C-ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define F6_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::fame6d::Error(FAME6D_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define F6_CHECK_LAUNCH() F6_HIP(hipGetLastError())

// Growable device buffer owned by a context or a scope; freed by its destructor, so a
// buffer of a function that throws (ENOCONV, EBUDGET, ...) does not outlive the call.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void ensure(size_t need) {
    if (need <= bytes) return;
    if (ptr) F6_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) throw Error(FAME6D_ENOMEM, "hipMalloc of " + std::to_string(need) + " bytes failed");
    bytes = need;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class U>
  U* as() const {
    return static_cast<U*>(ptr);
  }
};

// Page-locked host staging for the small control transfers (CG state, dot products,
// Lanczos scalars, group maps): a device-to-host read then costs one DMA and one stream
// synchronisation, and host-to-device copies are asynchronous.  Bump allocation; the
// space is recycled after each synchronisation of the owning stream.
struct HostPin {
  char* ptr = nullptr;
  size_t bytes = 0, used = 0;
  void init(size_t n) {
    F6_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr), n, hipHostMallocDefault));
    bytes = n;
    used = 0;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = used = 0;
  }
  // n bytes of staging; synchronises `s` first when the buffer is exhausted
  void* take(size_t n, hipStream_t s) {
    n = (n + 63) & ~size_t(63);
    if (n > bytes) throw Error(FAME6D_EARG, "host staging request of " + std::to_string(n) + " bytes");
    if (used + n > bytes) {
      F6_HIP(hipStreamSynchronize(s));
      used = 0;
    }
    void* p = ptr + used;
    used += n;
    return p;
  }
};

// Optional per-kernel-family timing (enabled through fame6d_mass_profile).  A timed
// launch goes through hipExtLaunchKernel with a start/stop event pair, so the events
// hold that kernel's own start and end times; durations are harvested when the pool
// fills or when the sums are read.  `bytes` accumulates the algorithmic HBM bytes of
// each timed launch as stated by its launcher.
struct KernelTimer {
  static constexpr int kSlots = FAME6D_NUM_TIMER_SLOTS;
  static constexpr int kPool = 16384;
  bool on = false;
  int stride = 1;                      // time every stride-th launch of each family
  bool armed = false;                  // the next launch_k of the current family is timed
  bool fired = false;                  // ... and has been issued
  long long seen[kSlots] = {0};
  std::vector<hipEvent_t> ev;          // 2*kPool events
  std::vector<int> slot;               // slot of pending pair i
  std::vector<double> bytes_pending;   // bytes of pending pair i (< 0: discarded)
  std::vector<int> tag;                // caller tag of pending pair i (-1: none)
  int pending = 0;
  long long count[kSlots] = {0};
  double ms[kSlots] = {0}, bytes[kSlots] = {0};
  void enable(int stride_or_off);
  void begin(int sl);
  void end(double alg_bytes, int tag = -1);
  hipEvent_t ev0() const { return ev[2 * pending]; }
  hipEvent_t ev1() const { return ev[2 * pending + 1]; }
  // discards pending pairs from index `first` on whose tag belongs to `owner` and whose
  // iteration is >= `threshold` (launches that the device skipped, e.g. CG iterations after
  // convergence); tags are owner * kTagSpan + iteration
  static constexpr int kTagSpan = 1 << 16;
  void discard_tags(int first, int owner, int threshold);
  int next_owner = 0;
  void harvest();  // waits for the pending events and accumulates
  void reset();
  void destroy();
};

// Tile variant indices of the three passes (chosen by the autotuner or defaults).
struct TileChoice {
  int a = 0;  // pass A / pass C variant
  int b = 0;  // pass B variant (M-hat applies, three vectors per group)
  int b1 = 0;  // pass B variant of the plain applies (one vector per group)
  int io = 1;  // passes A/C: 1 = coalesced dim-3-first I/O, 0 = dim-1 register pencils
  int a3 = 0;  // Mhat pass A (three thread groups) variant
  int c3 = 0;   // Mhat pass C (three thread groups) variant
  int c3p = 0;  // Mhat pass C: 0 = one workgroup per tile, k > 0 = persistent, k workgroups per CU
  int a3p = 0;  // Mhat pass A: the same
  int a3s = 0;  // persistent Mhat pass A: 1 = input read once and shared through LDS
};

}  // namespace fame6d

// Opaque handle of build_mass_backend_large_scalar (exp1/build_mass_backend_large_scalar.m:1-49).
struct fame6d_mass {
  int N = 0, n = 0;
  long long NF = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 0;  // compute units of the device
  double eps_min = 0, eps_max = 0, eps_mean = 0;
  fame6d::DevBuf eps64, eps32;  // eps / NF
  fame6d::DevBuf ieps64, ieps32;  // (1/eps) / NF: the inverse mass M_s^{-1} = F diag(1/eps) F^{-1}
  fame6d::DevBuf work64, work32;   // pass intermediates
  fame6d::DevBuf io64;             // staging for component wrappers
  fame6d::DevBuf partials;         // reduction partials
  fame6d::DevBuf scal;             // small device scalars
  void* noalloc_scratch = nullptr;  // this device's copy of the allocation-free apply's array
  int noalloc_status = 0;           // hipGetSymbolAddress status of that lookup
  fame6d::TileChoice tile64, tile32;
  // launches over several groups (batched Bloch points, several right-hand sides): one tile
  // per workgroup, tuned on several groups
  fame6d::TileChoice tile64g, tile32g;
  fame6d::HostPin pin;             // page-locked staging of small transfers
  mutable fame6d::KernelTimer timer;
};

namespace fame6d {

// Pass launchers (mass.hip).  Vectors of group g live at base + g*ld.
template <class T>
void launch_pass_a(const fame6d_mass* m, int op, int ngroups, const PassAArgs<T>& a, hipStream_t s);
// mode 0: fused inverse / eps / forward over dims 4-6; mode 1: inverse only.
template <class T>
void launch_pass_b(const fame6d_mass* m, int nr, int mode, int ngroups, cpx<T>* t, long long ldt,
                   Skip skip, hipStream_t s, bool inverse_eps = false, double* partials = nullptr);
template <class T>
void launch_pass_c(const fame6d_mass* m, int op, int ngroups, const PassCArgs<T>& a, hipStream_t s);
// Number of workgroups pass C uses per group (size of its partials per group).
int pass_c_blocks(const fame6d_mass* m, bool fp32, int ngroups);  // for a launch over ngroups groups
// Number of pass-B tiles per vector (its partials per vector; NR vectors per group).
int pass_b_blocks(const fame6d_mass* m, bool fp32, int ngroups);  // for a launch over ngroups groups
int num_tile_variants(int n, bool fp32, bool pass_b);
void tile_variant_desc(int n, bool fp32, bool pass_b, int v, int* param, int* threads);

// Plain scalar apply y = M_s x for nrhs columns (applyMij_fft semantics).
template <class T>
void mass_apply(fame6d_mass* m, const cpx<T>* x, long long ldx, cpx<T>* y, long long ldy, int nrhs,
                hipStream_t s);

// The same with the nrhs * NF pass intermediates in a caller-supplied buffer t (nothing is
// allocated).
template <class T>
void mass_apply_with(fame6d_mass* m, const cpx<T>* x, long long ldx, cpx<T>* y, long long ldy, int nrhs,
                     cpx<T>* t, hipStream_t s);

template <class T>
DevBuf& work_buf(fame6d_mass* m);

// y = M_s t with t the two outputs of the modewise step `op` (PRE_V1H or PRE_ZW) of each of
// ngroups groups: the step runs in the M_s apply's first pass.
void mass_apply_pre(fame6d_mass* m, int op, int ngroups, const PassAArgs<double>& a, cpx<double>* y,
                    long long ldy, hipStream_t s);

// Small transfers through the page-locked staging of m.  d2h_sync synchronises the
// stream and recycles the staging; h2d_async returns at once.
void d2h_sync(fame6d_mass* m, void* host, const void* dev, size_t bytes);
void h2d_async(fame6d_mass* m, void* dev, const void* host, size_t bytes);

// Kernel-family timer scopes (no-ops unless profiling is enabled on m).
enum TimerSlot : int {
  TS_MHAT_PASS_A = 0,  // Mhat pass A inside PCG (p = r + beta p fused)
  TS_MHAT_PASS_B = 1,  // Mhat pass B (three vectors, p'q partials)
  TS_MHAT_PASS_C = 2,  // Mhat pass C with alpha, x += alpha p, r -= alpha q, ||r||^2 partials
  TS_GEMV_C = 3,       // Lanczos h = Q^H w sweeps
  TS_GEMV_N = 4,       // Lanczos r = w - Q h sweeps (+||r||^2)
};
inline void timer_begin(const fame6d_mass* m, int slot) {
  if (m->timer.on) m->timer.begin(slot);
}
inline void timer_end(const fame6d_mass* m, double bytes, int tag = -1) {
  if (m->timer.on) m->timer.end(bytes, tag);
}

// kernel<<<grid, block, 0, s>>>(args...), with the arguments converted to the kernel's
// parameter types; a launch inside an armed timer scope carries the timer's events.
template <typename... KArgs, typename... Args>
inline void launch_ks(const fame6d_mass* m, void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shmem,
                      hipStream_t s, Args&&... args) {
  if (m->timer.armed && !m->timer.fired) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, m->timer.ev0(), m->timer.ev1(), 0,
                          static_cast<KArgs>(args)...);
    m->timer.fired = true;
  } else {
    kernel<<<grid, block, shmem, s>>>(static_cast<KArgs>(args)...);
  }
}
template <typename... KArgs, typename... Args>
inline void launch_k(const fame6d_mass* m, void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t s,
                     Args&&... args) {
  launch_ks(m, kernel, grid, block, 0u, s, static_cast<Args&&>(args)...);
}

// Per-solve results of a (batched) PCG on Mhat.
struct PcgOut {
  int flag = 0;
  int iter = 0;
  double relres = 0;
  int checks = 0;  // true-residual checks (one Mhat apply each)
};

// A one-group FP64 K_r^{-1} whose first PCG chunk, true-residual check and tail are enqueued
// as if that check passes; its state is copied to the host with the caller's own reads, and
// krinv_complete confirms it or finishes the solve.
struct KrinvDeferred {
  bool pending = false;
  int set = 0;                  // workspace set (fame6d_qsys::*_set)
  CgState* host_stage = nullptr;  // page-locked destination of the state copy (null: the q-system's own)
  int cur = 0, issued = 0, rounds = 0, timer_first = 0, timer_owner = 0, maxit = 0;
  double tol = 0.0;
  cpx<double>* yout = nullptr;
  const CgState* state = nullptr;  // device state after the check (valid: finished, flag 0, checked)
  std::chrono::steady_clock::time_point t0;
};

}  // namespace fame6d

// Opaque handle of build_experiment1_q_system_C_compact for G Bloch points
// (exp1/build_experiment1_q_system_C_compact.m:7-38).
struct fame6d_qsys {
  fame6d_mass* m = nullptr;
  int G = 0;
  long long NF = 0;
  double P[18];
  std::vector<double> q, klift, min_abs_det, min_qq;
  std::vector<int> n_singular;
  // diagonal blocks, each G*NF: V11 V12 V21 V22 V31 V32 | W1 W2 invC | q1 q2 q3
  fame6d::DevBuf Vd, Wd, Wf, Qd;
  // Method-C workspaces (grown to the number of active groups)
  fame6d::DevBuf kr64, kr32, st, part, part2, ints, sc, ref;
  double mixed_inner_tol = 3e-5;
  bool mixed_precond = true;  // FP32 inner PCG preconditioned by Mhat evaluated with 1/eps
  double dgks_eta2 = 0.5;     // squared DGKS threshold of the Lanczos re-orthogonalisation
  int mixed_max_cycles = 30;
  int chunk_hint = 0;                 // first chunk of the FP64 PCG (iterations of the last call + 1)
  std::vector<int> chunk_hint32 = std::vector<int>(4, 0);  // per refinement cycle of the mixed form
  std::vector<int> qmap_uploaded;  // q-slot map last copied to sc (skips repeated uploads)
  void* cg_pin = nullptr;          // page-locked host copy of a deferred solve's CgState
  void* inner_pin = nullptr;       // page-locked copies of the mixed form's inner-solve states
  size_t inner_pin_bytes = 0;
  // second workspace set of the one-group FP64 K_r^{-1}: a solve enqueued while the one before
  // it is still unconfirmed works in the other set, so that one can still be continued
  fame6d::DevBuf kr64_1, st_1, part_1, part2_1, ints_1;
  fame6d::DevBuf& kr64_set(int k) { return k ? kr64_1 : kr64; }
  fame6d::DevBuf& st_set(int k) { return k ? st_1 : st; }
  fame6d::DevBuf& part_set(int k) { return k ? part_1 : part; }
  fame6d::DevBuf& part2_set(int k) { return k ? part2_1 : part2; }
  fame6d::DevBuf& ints_set(int k) { return k ? ints_1 : ints; }
  fame6d_krinv_stats_t stats{};
  const double* V(int i) const { return Vd.as<double>() + (size_t)i * G * NF; }
  const double* W64(int i) const { return Wd.as<double>() + (size_t)i * G * NF; }
  const float* W32(int i) const { return Wf.as<float>() + (size_t)i * G * NF; }
  const double* Qm(int i) const { return Qd.as<double>() + (size_t)i * G * NF; }
};

namespace fame6d {

// Builds the q-system of G Bloch points on the host (FP64) and uploads it.
fame6d_qsys* qsys_build(fame6d_mass* m, int G, const double* q, const double* P);
// y = Mhat x for nact groups in q-system slots qmap (device ints, nullptr = identity).
// energy_partials (optional): pass-B partials whose per-group sum is x' Mhat x
// (3 * pass_b_blocks per group).  With resb set, y = resb - Mhat x instead, and res_partials
// receive ||y||^2 (pass_c_blocks per group).
// zw (optional, POST form): the pass-C outputs M_s(W1 z), M_s(W2 z), z = invC x, of group g
// at zw + 2g*NF and zw + (2g+1)*NF.
template <class T>
void mhat_apply(fame6d_qsys* S, int nact, const int* qmap, const cpx<T>* x, cpx<T>* y, Skip skip,
                hipStream_t s, bool inverse_eps = false, double* energy_partials = nullptr,
                const cpx<T>* resb = nullptr, double* res_partials = nullptr, cpx<T>* zw = nullptr);
// Batched K_r^{-1} of make_explicit_krinv_compact (one call per group slot).  With `def`
// (FP64, one group) the call returns without synchronising: the caller synchronises the
// stream and then calls krinv_complete, which returns false when the enqueued tail was not
// valid and has been formed again (the consumers of yout enqueued meanwhile must be redone).
void krinv_groups(fame6d_qsys* S, int precision, int nact, const int* qmap_host,
                  const cpx<double>* qin, cpx<double>* yout, double tol, int maxit,
                  std::vector<PcgOut>* outs, KrinvDeferred* def = nullptr);
bool krinv_complete(fame6d_qsys* S, KrinvDeferred* def);
template <class T>
void dot_groups(const fame6d_mass* m, int G, const cpx<T>* a, long long lda, const cpx<T>* b,
                long long ldb, long long n, double* out_re_host, double* out_im_host);

// Ritz-vector selection of the tracking scan (exp3/Lanczos_tracking.m:134-165): fixed
// 1-based indices, or the nearest eigenvalue to each target widened by half_width.
struct RitzSelect {
  const int* idx = nullptr;
  int nidx = 0;
  const double* target = nullptr;
  int ntarget = 0;
  int half_width = 0;
  int* idx_out = nullptr;     // selected 1-based indices (capacity given by the caller)
  int idx_capacity = 0;
  int* nidx_out = nullptr;
  double* rel_res = nullptr;  // k relative inverse-Ritz residuals |beta U(end, i)| / |theta_i|
};

inline int grid_for(long long n, int threads = 256) {
  long long b = (n + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

}  // namespace fame6d
