/*
 * fame6d.h - C ABI of the MI355X-native hot path of the null-space-free 6D
 * Fourier projection eigensolver (scalar-permittivity mass operator, Method-C
 * explicit reduced inverse, thick-restart Lanczos, Taylor reconstruction).
 *
 * Every entry point replaces one function of the reference MATLAB code; the
 * reference file:line is cited above each declaration (paths relative to the
 * reference repository root).
 *
 * Conventions
 *   - Grid vectors are the column-major (2N)^6 arrays of the reference
 *     (linear index k1 + 2N*k2 + ... , first index fastest), NF = (2N)^6.
 *   - Complex vectors are interleaved (re, im): complex128 for FAME6D_FP64,
 *     complex64 for FAME6D_FP32.  Multi-component vectors are stacked
 *     component-first ([x1; x2; x3]); several right-hand sides are columns
 *     with leading dimension equal to the row count.
 *   - Vector arguments are DEVICE pointers unless the name ends in _host.
 *   - Work is enqueued on the stream given at fame6d_mass_create; calls that
 *     return host scalars synchronise that stream.  A handle is not re-entrant:
 *     use one handle per stream.
 *   - The caller owns input/output buffers; handles own their workspaces.
 *   - Return value 0 = success, otherwise a FAME6D_E* code; the message of the
 *     last failure on the calling thread is fame6d_last_error().
 */
#ifndef FAME6D_H
#define FAME6D_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; map the reference error identifiers. */
#define FAME6D_OK 0
#define FAME6D_ESIZE 1      /* rank1_exact_fft_scatter:BadSize (exp1/rank1_exact_fft_scatter.m:5-7) */
#define FAME6D_ENONFINITE 2 /* NaN/Inf guards (exp1/make_explicit_krinv_compact.m:27,37,67) */
#define FAME6D_ENOCONV 3    /* PCG failure / EXP1:NotConverged (make_explicit_krinv_compact.m:55-58, run_method_C_fig1bc.m:15-17) */
#define FAME6D_EBUDGET 4    /* EXP1:IterationLimit / TimeLimit (exp1/make_run_budget.m:28-46) */
#define FAME6D_ENOMEM 5     /* EXP1:MemoryLimit (exp1/classify_exp1_exception.m:22-36) */
#define FAME6D_EHIP 6       /* HIP runtime failure */
#define FAME6D_EARG 7       /* invalid argument */
#define FAME6D_EGAMMA 8     /* zero projected mode, Gamma point (exp1/assemble_K_and_nullspace_free_basis.m:36-42) */
#define FAME6D_EMASSNORM 9  /* recover_original_eigenvectors_exp2:BadMassNorm (exp2/recover_original_eigenvectors_exp2.m:36-41) */

/* Arithmetic of an operation. */
#define FAME6D_FP64 0  /* complex128 vectors, FP64 FFTs (cfg.gpu.precision = 'double') */
#define FAME6D_FP32 1  /* complex64 vectors, FP32 FFTs with FP32 accumulation */
#define FAME6D_MIXED 2 /* FP32 inner PCG + FP64 iterative refinement; FP64 Lanczos */

typedef struct fame6d_mass fame6d_mass;
typedef struct fame6d_qsys fame6d_qsys;

const char* fame6d_last_error(void);
int fame6d_version(void);

/* ----------------------------------------------------------------- mass backend */

typedef struct {
  int N;
  int nGrid;
  long long NF;
  double eps_min, eps_max, eps_mean;
  int tile_a64, tile_b64, tile_a32, tile_b32; /* autotuned tile parameters */
  int variant_ac64, variant_b64;              /* fame6d_mass_set_tiles codes of the choice */
  int variant_ac32, variant_b32;
} fame6d_mass_info_t;

/* build_mass_backend_large_scalar (exp1/build_mass_backend_large_scalar.m:1-49) with
 * the 'batchfft-scalar' operator M_s = F diag(eps) F^{-1} (exp1/applyMij_fft.m:11),
 * which equals the 'rank1-exact-scalar' operator (exp1/applyMij_rank1_exact.m:4-5).
 * eps_host: NF real permittivity samples, column-major. stream: hipStream_t or NULL. */
int fame6d_mass_create(int N, const double* eps_host, void* stream, fame6d_mass** out);
int fame6d_mass_destroy(fame6d_mass* m);
/* Replace the permittivity samples of an existing backend (same N); used by gradient
 * checks and optimisation loops over eps.  No reference counterpart (the reference
 * rebuilds the backend, exp1/build_experiment1_common.m:12-18). */
int fame6d_mass_set_eps(fame6d_mass* m, const double* eps_host);
/* The same with the NF samples in device memory (float64, column-major): validated and
 * scaled on the device, so a gradient loop over a device-resident eps moves no samples
 * through the host.  The permittivity enters the operator as in exp1/applyMij_fft.m:11. */
int fame6d_mass_set_eps_device(fame6d_mass* m, const double* eps_dev);
int fame6d_mass_info(const fame6d_mass* m, fame6d_mass_info_t* info);

/* Project-written tile autotuner: times every compiled tile variant of the three
 * passes at this N and precision and keeps the fastest, once for launches over one group
 * (one Bloch point; one or two right-hand sides of the plain apply) and once, among the
 * one-tile-per-workgroup forms, for launches over several (batched Bloch points, more
 * right-hand sides); fame6d_mass_set_tiles sets the one-group choice (no reference
 * counterpart; the reference has one implicit cuFFT plan). */
int fame6d_mass_autotune(fame6d_mass* m, int precision, int reps);
/* Set tile variants explicitly: variant_ac in 0..511 (bit 0: blocks per workgroup of the
 * plain passes A/C, bit 1 set: their dim-1 register-pencil I/O instead of the coalesced
 * dim-3-first I/O, bit 2: blocks per workgroup of the three-group Mhat pass A, bits 3-4:
 * k > 0 runs the Mhat pass C persistent with k workgroups per compute unit, 0 one
 * workgroup per tile, bit 5: blocks per workgroup of the Mhat pass C, bits 6-7: the same
 * as bits 3-4 for the Mhat pass A, bit 8: the persistent Mhat pass A reads its input once
 * and shares it through LDS), variant_b in 0..3 (bit 0: inner positions per workgroup of
 * pass B in the M-hat applies, bit 1: the same in the plain applies). */
int fame6d_mass_set_tiles(fame6d_mass* m, int precision, int variant_ac, int variant_b);

/* Kernel-family timing (no reference counterpart; the reference times whole operator
 * calls with tic/toc, Lanczos_hybrid_identity_gpuop.m:37-39).  While enabled, every
 * `enable`-th launch of a timed family on the handle's stream is bracketed by HIP
 * events (enable = 0 switches timing off); the durations and the algorithmic HBM bytes
 * of the timed launches accumulate per slot until read.  Slots: 0 Mhat pass A in PCG,
 * 1 Mhat pass B, 2 Mhat pass C with the fused CG update, 3 Lanczos Q^H w sweeps,
 * 4 Lanczos w - Q h sweeps.  Every call of fame6d_mass_profile resets the sums. */
#define FAME6D_NUM_TIMER_SLOTS 5
typedef struct {
  long long launches;
  double milliseconds;  /* summed kernel durations */
  double bytes;         /* summed algorithmic HBM bytes */
} fame6d_kernel_time_t;
int fame6d_mass_profile(fame6d_mass* m, int enable);
int fame6d_mass_kernel_times(fame6d_mass* m, fame6d_kernel_time_t* out /* FAME6D_NUM_TIMER_SLOTS */);

/* mass.MscalarOp: applyMij_rank1_exact_lowmem (exp1/applyMij_rank1_exact_lowmem.m:8-22),
 * applyMij_fft (exp1/applyMij_fft.m:9-23).  x, y: NF x nrhs.  y may alias x. */
int fame6d_mass_apply_scalar(fame6d_mass* m, int precision, const void* x, void* y, int nrhs);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * operator of fame6d_mass_apply_scalar, with its pass intermediates in a fixed 64 MiB device
 * array of the library (one copy per GPU, held until the process ends) instead of the
 * handle's growable workspace, so the call allocates nothing.  Limit: nrhs * NF * 16 bytes
 * (FP64) or nrhs * NF * 8 bytes (FP32) <= 64 MiB, else FAME6D_ESIZE (N = 5: up to 4 FP64
 * columns; N = 6: 1).  The array is fully overwritten by each call before it is read.  All
 * calls on one GPU share it: calls that overlap on one GPU (several streams or host threads)
 * race. */
int fame6d_mass_apply_scalar_noalloc(fame6d_mass* m, int precision, const void* x, void* y, int nrhs);
/* mass.Mop: applyM_component_first_rank1_scalar (exp1/applyM_component_first_rank1_scalar.m:4-19).
 * U, MU: 3NF x nrhs. */
int fame6d_mass_apply_3comp(fame6d_mass* m, int precision, const void* U, void* MU, int nrhs);
/* mass.Mcols12op: applyM_cols12_rank1_scalar (exp1/applyM_cols12_rank1_scalar.m:3-11).
 * x12: 2NF x nrhs -> Y: 3NF x nrhs with block 3 = 0. */
int fame6d_mass_apply_cols12(fame6d_mass* m, int precision, const void* x12, void* Y, int nrhs);
/* mass.Mcol3op: applyM_col3_rank1_scalar (exp1/applyM_col3_rank1_scalar.m:3-9).
 * x: NF x nrhs -> Y: 3NF x nrhs with blocks 1-2 = 0. */
int fame6d_mass_apply_col3(fame6d_mass* m, int precision, const void* x, void* Y, int nrhs);

/* ----------------------------------------------------------------- q-system */

typedef struct {
  long long NF;
  double min_abs_det_v1; /* build_experiment1_q_system_C_compact.m:15-19 (warn if < 1e-12) */
  double min_qq;         /* min |q_xi|^2 (assemble_K_and_nullspace_free_basis.m:33-42) */
  int n_singular;        /* modes on the singular branch (assemble_K_and_nullspace_free_basis.m:75-91) */
  double k_lift[6];      /* P (P^T P)^{-1} q (build_experiment1_q_system_C_compact.m:9) */
} fame6d_qsys_info_t;

/* build_experiment1_q_system_C_compact (exp1/build_experiment1_q_system_C_compact.m:7-38)
 * + assemble_K_and_nullspace_free_basis (exp1/assemble_K_and_nullspace_free_basis.m:18-105).
 * q_host: Bloch vector (3); P_host: 6x3 projection, row-major (P[l*3+c]); T = 2*pi. */
int fame6d_qsys_create(fame6d_mass* m, const double* q_host, const double* P_host, fame6d_qsys** out);
int fame6d_qsys_destroy(fame6d_qsys* s);
int fame6d_qsys_info(const fame6d_qsys* s, fame6d_qsys_info_t* info);
/* Modewise diagonals as host arrays (each NF, may be NULL): V11,V12,V21,V22,V31,V32,W1,W2,invC. */
int fame6d_qsys_diagonals(const fame6d_qsys* s, double* V11, double* V12, double* V21, double* V22,
                          double* V31, double* V32, double* W1, double* W2, double* invC);

/* Parameters of the FAME6D_MIXED Mhat solve: relative tolerance of each FP32 inner PCG,
 * the maximum number of FP64 refinement cycles, and whether the FP32 inner PCG is
 * preconditioned by Mhat evaluated with 1/eps (exact for constant eps; the FP64
 * refinement keeps the result independent of this choice).  No reference counterpart:
 * the reference solves in FP64 only (cfg.gpu.precision = 'double'). */
int fame6d_qsys_set_mixed(fame6d_qsys* s, double inner_tol, int max_cycles, int precondition);

/* Re-orthogonalisation threshold of the Method-C Lanczos steps: after the first classical
 * Gram-Schmidt sweep a second one runs when ||r_after|| < eta ||r_before|| (the DGKS test;
 * default eta = 1/sqrt(2)); eta >= 1 runs it at every step, as the reference's two fixed
 * modified Gram-Schmidt sweeps (exp1/Lanczos_hybrid_identity_gpuop.m:55-63). */
int fame6d_qsys_set_reorth(fame6d_qsys* s, double eta);

/* apply_Mhat_C_scalar_compact (exp1/apply_Mhat_C_scalar_compact.m:8-21): x, y NF. */
int fame6d_mhat_apply(fame6d_qsys* s, int precision, const void* x, void* y);

/* apply_V_blockdiag (exp1/apply_V_blockdiag.m:9-16): x 2NF -> y 3NF (FP64). */
int fame6d_apply_v(fame6d_qsys* s, const void* x, void* y);
/* Kop (exp1/assemble_K_and_nullspace_free_basis.m:110-119): x 3NF -> y 3NF (FP64). */
int fame6d_apply_k(fame6d_qsys* s, const void* x, void* y);

typedef struct {
  long long calls;
  long long total_iterations;  /* Mhat CG iterations (FP32 ones in MIXED) */
  long long total_fp64_iterations;
  int max_iterations;
  int last_iterations;
  int last_flag;
  double last_relres;
  double max_relres;
  long long refinement_cycles; /* MIXED only */
  long long mass_applies_fp64; /* scalar M_s applies issued */
  long long mass_applies_fp32;
  double seconds;
} fame6d_krinv_stats_t;

/* make_explicit_krinv_compact applyOne (exp1/make_explicit_krinv_compact.m:24-80):
 * y = K_r^{-1} q with PCG on Mhat (tol, maxit; pcg call at :43).  q, y: 2NF complex128.
 * precision FAME6D_MIXED runs the Mhat solve in FP32 with FP64 iterative refinement. */
int fame6d_krinv_apply(fame6d_qsys* s, int precision, const void* q, void* y, double tol, int maxit,
                       fame6d_krinv_stats_t* stats);

/* ----------------------------------------------------------------- eigensolver */

typedef struct {
  int converged;
  int num_converged;
  int outer_steps;      /* total Lanczos steps (Lanczos_hybrid_identity_gpuop.m:75) */
  int restarts;
  int hit_max_restart;
  double max_ritz_residual;
  double seconds;
  fame6d_krinv_stats_t krinv;
  int second_sweeps;    /* steps whose DGKS test ran a second Gram-Schmidt sweep (Methods B, C) */
} fame6d_lanczos_stats_t;

/* run_method_C_fig1bc (exp1/run_method_C_fig1bc.m:4-34) driving
 * Lanczos_hybrid_identity_gpuop (exp1/Lanczos_hybrid_identity_gpuop.m:11-130) with the
 * basis kept on the device.  v0: 2NF complex128 start vector (normalised inside), in host or
 * device memory (a device start vector is copied device to device).
 * lambda_host: k ascending eigenvalues 1/theta.  Y: optional device 2NF x k complex128
 * reduced Ritz vectors (ordered as lambda). */
int fame6d_eigs_method_c(fame6d_qsys* s, int precision, int k, int p, double tol, int max_restart,
                         double tol_cg, int maxit_cg, const void* v0, double* lambda_host,
                         void* Y, fame6d_lanczos_stats_t* stats);


/* Lanczos_tracking (exp3/Lanczos_tracking.m:1-188) as solve_tracking_scan runs it
 * (exp3/solve_tracking_scan.m:14-40): the Lanczos of fame6d_eigs_method_c for the k lowest
 * eigenvalues (k = 140, p = k + 48 for the N = 5 scan, exp3/make_config_experiment4.m:48-50;
 * p <= 256), after which only the selected Ritz vectors are formed.  Selection: the nfixed
 * 1-based indices fixed_host (i-th smallest eigenvalue), or, with nfixed = 0, for each of
 * the ntarget values target_host the nearest eigenvalue widened by half_width on each side
 * (at least two candidates, as the reference requires).  Outputs: lambda_host (k, ascending),
 * rel_res_host (k relative inverse-Ritz residuals, optional), sel_host (the selected indices,
 * capacity sel_capacity), *nsel, Y (device, 2NF x *nsel complex128, optional). */
int fame6d_eigs_tracking(fame6d_qsys* s, int precision, int k, int p, double tol, int max_restart, double tol_cg,
                         int maxit_cg, const void* v0, int nfixed, const int* fixed_host, int ntarget,
                         const double* target_host, int half_width, double* lambda_host, double* rel_res_host,
                         int* sel_host, int sel_capacity, int* nsel, void* Y, fame6d_lanczos_stats_t* stats);

/* The same for nq Bloch points sharing one mass backend, solved together: every
 * Lanczos step issues one batched K_r^{-1} over all still-active points, so each
 * kernel launch covers all of them (q_host: 3*nq; v0: nq x 2NF, host or device; lambda_host:
 * nq x k; Y: optional device nq x (2NF x k)).  No reference counterpart: the
 * reference loops over Bloch points sequentially (exp1/run_fig1a_BC_method.m:37-80). */
int fame6d_eigs_method_c_batch(fame6d_mass* m, int nq, const double* q_host, const double* P_host,
                               int precision, int k, int p, double tol, int max_restart,
                               double tol_cg, int maxit_cg, const void* v0,
                               double* lambda_host, void* Y, fame6d_lanczos_stats_t* stats);

/* make_nested_krinv_instrumented applyOne (exp1/make_nested_krinv_instrumented.m:33-132), Method B:
 * x = K_r^{-1} b by an outer PCG on K_r (tol_kr, maxit_kr) preconditioned with
 * eps_c * r ./ sigma_r (:47), each K_r z = V^*(M^{-1}(V z)) with an inner PCG on M
 * (tol_m, maxit_m).  b, x: 2NF complex128 (FP64). */
int fame6d_krinv_nested_apply(fame6d_qsys* s, const void* b, void* x, double tol_kr, double tol_m,
                              int maxit_kr, int maxit_m, double eps_c, fame6d_krinv_stats_t* stats);
/* run_method_B_fig1bc: the Lanczos of fame6d_eigs_method_c driving the Method-B inverse
 * (exp1/make_nested_krinv_instrumented.m; cfg.nested in exp1/make_config_exp1bc.m:83-87). */
int fame6d_eigs_method_b(fame6d_qsys* s, int k, int p, double tol, int max_restart, double tol_kr,
                         double tol_m, int maxit_kr, int maxit_m, double eps_c, const void* v0,
                         double* lambda_host, void* Y, fame6d_lanczos_stats_t* stats);

/* Method A statistics (make_A_eigs_top_preconditioned.m:53-54 and the per-solve info of
 * :562-580: totalMINRESIterations, stagnationRestarts, residualCorrections, ...). */
typedef struct {
  long long solves;      /* shifted systems solved (topVectorApplications) */
  long long iterations;  /* MINRES iterations over all calls, correction calls included */
  long long segments;    /* MINRES calls (segmentCount; totalMINRESCalls) */
  long long restarts;    /* stagnation restarts and MAXIT continuations from the current iterate */
  int max_iterations;    /* most iterations of one shifted solve */
  int plateau_accepts;   /* solves accepted at the 1e-9 plateau / guard ceiling */
  double max_relres;     /* largest true relative residual accepted */
  double seconds;
  long long corrections;            /* residual-correction cycles (residualCorrections) */
  long long correction_iterations;  /* their MINRES iterations (correctionMINRESIterations) */
} fame6d_minres_stats_t;
/* Method A Top (make_A_eigs_top_preconditioned.m:73-113): y = (K - sigma M)^{-1} M x by
 * preconditioned MINRES with P = K + tau I, tau = sigma * mean(eps)
 * (build_original_shift_preconditioner_A_only.m:1-56), true relative residual <= tol.
 * x, y: 3NF complex128 (component-first). */
int fame6d_method_a_apply(fame6d_qsys* s, double sigma, double tol, int maxit, const void* x, void* y,
                          fame6d_minres_stats_t* stats);
/* run_method_A_fig1a (exp1/run_method_A_fig1a.m:6-50): the k largest theta of
 * T = (K - sigma M)^{-1} M, lambda = sigma + 1/theta ascending.  The outer solver is
 * thick-restart Lanczos in the M inner product (T is M-self-adjoint) where the reference
 * calls eigs; k, p, tol, max_restart as cfg.eigs (make_config_exp1bc.m:36-47).
 * v0: 3NF start vector, host or device; Y (device, 3NF x k) or NULL. */
int fame6d_eigs_method_a(fame6d_qsys* s, int k, int p, double tol, int max_restart, double sigma,
                         double tol_minres, int maxit_minres, const void* v0, double* lambda_host, void* Y,
                         fame6d_lanczos_stats_t* stats, fame6d_minres_stats_t* mstats);

typedef struct {
  int iterations[64];
  double relres[64];
  double mass_norm2[64];
} fame6d_recover_stats_t;

/* recover_original_eigenvectors_exp2 (exp2/recover_original_eigenvectors_exp2.m:18-51):
 * X(:,j) = M^{-1} V Y(:,j) by PCG on M (tol, maxit), M-normalised; MX = M X.
 * Y: 2NF x k, X, MX: 3NF x k (complex128, device). k <= 64. */
int fame6d_recover_eigenvectors(fame6d_qsys* s, const void* Y, int k, double tol, int maxit, void* X,
                                void* MX, fame6d_recover_stats_t* stats);

/* compute_fourier_diagnostics_exp2 (exp2/compute_fourier_diagnostics_exp2.m:10-26). */
int fame6d_fourier_diagnostics(fame6d_qsys* s, const void* X, const void* MX,
                               const double* lambda_host, int k, double* res_gevp_host,
                               double* div_host, double* rayleigh_host);

/* ----------------------------------------------------------------- Taylor */

/* Moments of reconstruct_component_taylor_batch_exp4 (exp3/reconstruct_component_taylor_batch_exp4.m:14-25):
 * moments(alpha, c) = i^|alpha|/alpha! * sum_j q_j^alpha exp(i q_j . r_c) coeff_j
 * for |alpha| <= p (ordering of build_multi_index_list_3d_exp3.m:6-12).
 * qmodes: NF x 3 real (column-major, device), coeff: NF complex128 (device),
 * centers_host: C x 3 (row-major), moments: L x C complex128 (device, column-major). */
int fame6d_taylor_moments(long long NF, const double* qmodes, const void* coeff, int p, int C,
                          const double* centers_host, void* moments, void* stream);
/* Evaluation loop of the same function (:33-38): u(pt) = sum_alpha (r - r_c)^alpha
 * moments(alpha, c(pt)).  points: M x 3 real (column-major, device), center_id: M int32
 * (device), u: M complex128 (device). */
int fame6d_taylor_evaluate(long long M, const double* points, const int* center_id, int p, int C,
                           const double* centers_host, const void* moments, void* u, void* stream);
/* The same evaluation on the (2n+3)^3 plus-grid Yee points of component comp (0,1,2),
 * generated on the device (exp2/build_yee_edge_points_plus.m:11-17) and split into
 * nblocks^3 centers (exp2/build_yee_block_partition_exp3.m:1-31); u: (2n+3)^3 complex128
 * in column-major point order; C = nblocks^3. */
int fame6d_taylor_evaluate_yee(int n, double h, int comp, int nblocks, int p, int C,
                               const double* centers_host, const void* moments, void* u, void* stream);
/* reconstruct_field_direct_blocked_exp3 (exp2/reconstruct_field_direct_blocked_exp3.m:10-15):
 * u = exp(i R q^T) coeff. */
int fame6d_direct_reconstruct(long long NF, const double* qmodes, const void* coeff, long long M,
                              const double* points, void* u, void* stream);

/* compute_field_errors_exp3 (exp2/compute_field_errors_exp3.m:3-10): relative 2-norm and
 * max-norm errors of a three-component field against a reference,
 * e2 = ||t - r||_2 / max(||r||_2, eps), einf = max|t - r| / max(max|r|, eps), the norms taken
 * over all three components together.  t1..t3, r1..r3: n complex128 each (device);
 * e2, einf: host scalars (the call synchronises the stream). */
int fame6d_field_errors(long long n, const void* t1, const void* t2, const void* t3, const void* r1,
                        const void* r2, const void* r3, double* e2, double* einf, void* stream);

/* ----------------------------------------------------------------- Yee closure */

/* Window of half-width n*h: plus-grid fields u1, u2, u3 are (2n+3)^3 complex128 arrays
 * vectorised column-major (x fastest; the output of fame6d_taylor_evaluate_yee), the
 * interior and every density or rho array is (2n+1)^3, column-major.  Entry points that
 * take `work` use a device workspace of fame6d_yee_workspace_bytes(n) bytes, or allocate
 * one for the call when it is NULL. */
size_t fame6d_yee_workspace_bytes(int n);

/* apply_cropped_A_exp3 (exp3/apply_cropped_A_exp3.m:8-29): cropped Yee curl-curl of the
 * plus-grid field on the interior.  A1..A3: (2n+1)^3 complex128 (device). */
int fame6d_yee_apply_cropped_A(int n, double h, const void* u1, const void* u2, const void* u3, void* A1,
                               void* A2, void* A3, void* stream);

/* sample_scalar_yee_material_exp4 (exp3/sample_scalar_yee_material_exp4.m:4-21):
 * eps_c = eps_c * exp(alpha/6 sum_l cos(P_l . x)) on the three edge families of the interior.
 * P_host: nrows x 3 row-major (the 6 x 3 cfg.problem.P); e1..e3: (2n+1)^3 float64 (device). */
int fame6d_yee_sample_eps(int n, double h, const double* P_host, int nrows, double eps_c, double alpha, void* e1,
                          void* e2, void* e3, void* work, void* stream);

typedef struct {
  double lambda;                 /* sum E / sum M */
  double V, nu;                  /* weighted variance of rho = E/M about lambda, sqrt(V) */
  double partition_error;        /* max of the two below */
  double energy_partition_error; /* |sum E - sum a| / |sum a| */
  double mass_partition_error;   /* |sum M - sum b| / |sum b| */
  double sum_a, sum_b;
} fame6d_lwrq_t;

/* compute_lwrq_exp4 (exp3/compute_lwrq_exp4.m:12-56): E = conv(a ./ den), M = conv(b ./ den)
 * with the separable Gaussian of radius `radius` cells and width sigma_factor*h,
 * den = conv(ones); FAME6D_ENOCONV when M is non-finite or <= mass_floor_rel * max M.
 * a, b: (2n+1)^3 float64 (device); rho: optional (2n+1)^3 float64 output E./M. */
int fame6d_lwrq(int n, double h, const double* a, const double* b, int radius, double sigma_factor,
                double mass_floor_rel, double* rho, void* work, fame6d_lwrq_t* out, void* stream);

typedef struct {
  double lambda_crop, lambda_3d;      /* sum a / sum b; LWRQ lambda */
  double eta_3d, eta_6d;              /* ||Au - l Bu|| / (||Au|| + |l| ||Bu||) at lambda_3d, lambda6d */
  double delta_lambda;                /* |lambda_3d - lambda6d| / |lambda6d| */
  double V, nu;                       /* LWRQ variance and its square root */
  double delta_part, delta_identity;  /* partition error; |lambda_3d - lambda_crop| / |lambda_crop| */
  double cropped_mass;                /* sum b */
  double energy_partition_error, mass_partition_error;
} fame6d_yee_metrics_t;

/* compute_yee_lwrq_metrics_exp4 (exp3/compute_yee_lwrq_metrics_exp4.m:16-70): the cropped
 * curl-curl, the Yee material (P_host, eps_c, alpha as above), the densities
 * a = Re(u' A u), b = Re(u' B u), the LWRQ (radius, sigma_factor, mass_floor_rel as
 * cfg.lwrq, exp3/make_config_experiment4.m:72-74) and the residuals at lambda_3d and the
 * 6D eigenvalue lambda6d.  rho: optional (2n+1)^3 float64 output of the local quotient. */
int fame6d_yee_lwrq_metrics(int n, double h, const void* u1, const void* u2, const void* u3, const double* P_host,
                            int nrows, double eps_c, double alpha, double lambda6d, int radius,
                            double sigma_factor, double mass_floor_rel, double* rho, void* work,
                            fame6d_yee_metrics_t* out, void* stream);

/* ----------------------------------------------------------------- adjoint */

/* Adjoint of y = M_s x with respect to the permittivity samples: for each rhs,
 * geps += NF * Re(conj(ifftn(gy)) .* ifftn(x))  (derivative of applyMij_fft.m:11 in eps;
 * no reference counterpart).  x, gy: NF x nrhs complex128; geps: NF real (device). */
int fame6d_mass_adjoint_eps(fame6d_mass* m, const void* x, const void* gy, int nrhs, double* geps);
/* Hellmann-Feynman eigenvalue gradient: geps += sum_j glambda_j * (-lambda_j * NF *
 * sum_c |ifftn(X_c(:,j))|^2) for M-normalised X (3NF x k, device). */
int fame6d_eig_grad_eps(fame6d_mass* m, const void* X, int k, const double* lambda_host,
                        const double* glambda_host, double* geps);

#ifdef __cplusplus
}
#endif

#endif /* FAME6D_H */
