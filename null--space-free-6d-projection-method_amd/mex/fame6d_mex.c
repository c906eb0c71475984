/*
 * fame6d_mex.c - MATLAB MEX gateway onto the C ABI of libfame6d (include/fame6d.h).
 *
 * MATLAB's gpuArray/mxGPUArray API targets NVIDIA devices only, so this gateway
 * takes host interleaved-complex arrays (R2018a API, compile with -R2018a) and keeps
 * device state behind uint64 handles.  Build where MATLAB is installed:
 *
 *   mex -R2018a fame6d_mex.c -I<repo>/include -L<pkg dir> -lfame6d -L/opt/rocm/lib -lamdhip64
 *
 * Usage (mirrors the reference handles of build_mass_backend_large_scalar,
 * exp1/build_mass_backend_large_scalar.m:27-36, and run_method_C_fig1bc,
 * exp1/run_method_C_fig1bc.m:4-21):
 *   h  = fame6d_mex('mass_create', N, eps)            % eps: NF real (column-major grid)
 *   y  = fame6d_mex('mscalar', h, x)                    % MscalarOp, x: NF x r complex
 *   y  = fame6d_mex('mop', h, U)                        % Mop, U: 3NF x r complex
 *   s  = fame6d_mex('qsys_create', h, q, P)             % q: 3x1, P: 6x3
 *   y  = fame6d_mex('krinv', s, x, tol, maxit)          % KrInvOp (Method C), x: 2NF x 1
 *   [lambda, Y] = fame6d_mex('eigs', s, k, p, tol, maxRestart, tolCG, maxitCG, v0, precision)
 *   y  = fame6d_mex('topA', s, x, sigma, tol, maxit)    % Method-A Top (make_A_eigs_top_preconditioned.m)
 *   [lambda, Y] = fame6d_mex('eigsA', s, k, p, tol, maxRestart, sigma, tolMINRES, maxitMINRES, v0)
 *   [lambda, Y, idx, relRes] = fame6d_mex('tracking', s, k, p, tol, maxRestart, tolCG, maxitCG, v0, fixedIdx)
 *                                                     % Lanczos_tracking (exp3/Lanczos_tracking.m)
 *   m  = fame6d_mex('yee_metrics', n, h, u1, u2, u3, lambda6D, P, epsC, alpha)
 *                                                     % compute_yee_lwrq_metrics_exp4 (exp3/...:16-70),
 *                                                     % u*: (2n+3)^3 complex plus-grid fields
 *   fame6d_mex('qsys_destroy', s); fame6d_mex('mass_destroy', h)
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include "fame6d.h"
#include "mex.h"

static void fail(int code) {
  if (code != FAME6D_OK) mexErrMsgIdAndTxt("fame6d:error", "fame6d status %d: %s", code, fame6d_last_error());
}

static void* handle_in(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("fame6d:handle", "expected a uint64 handle");
  return (void*)(uintptr_t)(*mxGetUint64s(a));
}

static mxArray* handle_out(void* p) {
  mxArray* a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *mxGetUint64s(a) = (uint64_t)(uintptr_t)p;
  return a;
}

/* device copy of a host complex matrix (interleaved complex128) */
static void* to_device(const mxArray* a, size_t* n_out) {
  size_t n = mxGetNumberOfElements(a);
  void* d = NULL;
  if (hipMalloc(&d, n * 16) != hipSuccess) mexErrMsgIdAndTxt("fame6d:nomem", "hipMalloc failed");
  if (mxIsComplex(a)) {
    hipMemcpy(d, mxGetComplexDoubles(a), n * 16, hipMemcpyHostToDevice);
  } else {
    /* promote real input to complex */
    double* tmp = (double*)mxCalloc(2 * n, sizeof(double));
    const double* re = mxGetDoubles(a);
    for (size_t i = 0; i < n; ++i) tmp[2 * i] = re[i];
    hipMemcpy(d, tmp, n * 16, hipMemcpyHostToDevice);
    mxFree(tmp);
  }
  *n_out = n;
  return d;
}

static mxArray* from_device(const void* d, size_t rows, size_t cols) {
  mxArray* a = mxCreateDoubleMatrix(rows, cols, mxCOMPLEX);
  hipMemcpy(mxGetComplexDoubles(a), d, rows * cols * 16, hipMemcpyDeviceToHost);
  return a;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[64];
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof(cmd)) != 0)
    mexErrMsgIdAndTxt("fame6d:usage", "first argument must be a command string");

  if (!strcmp(cmd, "mass_create")) {
    int N = (int)mxGetScalar(prhs[1]);
    fame6d_mass* m = NULL;
    fail(fame6d_mass_create(N, mxGetDoubles(prhs[2]), NULL, &m));
    plhs[0] = handle_out(m);
  } else if (!strcmp(cmd, "mass_destroy")) {
    fail(fame6d_mass_destroy((fame6d_mass*)handle_in(prhs[1])));
  } else if (!strcmp(cmd, "mscalar") || !strcmp(cmd, "mop")) {
    fame6d_mass* m = (fame6d_mass*)handle_in(prhs[1]);
    size_t rows = mxGetM(prhs[2]), cols = mxGetN(prhs[2]), n;
    void* x = to_device(prhs[2], &n);
    void* y = NULL;
    hipMalloc(&y, n * 16);
    if (!strcmp(cmd, "mscalar")) fail(fame6d_mass_apply_scalar(m, FAME6D_FP64, x, y, (int)cols));
    else fail(fame6d_mass_apply_3comp(m, FAME6D_FP64, x, y, (int)cols));
    hipDeviceSynchronize();
    plhs[0] = from_device(y, rows, cols);
    hipFree(x);
    hipFree(y);
  } else if (!strcmp(cmd, "qsys_create")) {
    fame6d_mass* m = (fame6d_mass*)handle_in(prhs[1]);
    const double* Pcm = mxGetDoubles(prhs[3]); /* MATLAB 6x3 column-major */
    double P[18];
    for (int l = 0; l < 6; ++l)
      for (int c = 0; c < 3; ++c) P[l * 3 + c] = Pcm[l + 6 * c];
    fame6d_qsys* s = NULL;
    fail(fame6d_qsys_create(m, mxGetDoubles(prhs[2]), P, &s));
    plhs[0] = handle_out(s);
  } else if (!strcmp(cmd, "qsys_destroy")) {
    fail(fame6d_qsys_destroy((fame6d_qsys*)handle_in(prhs[1])));
  } else if (!strcmp(cmd, "krinv")) {
    fame6d_qsys* s = (fame6d_qsys*)handle_in(prhs[1]);
    size_t n;
    void* x = to_device(prhs[2], &n);
    void* y = NULL;
    hipMalloc(&y, n * 16);
    fame6d_krinv_stats_t st;
    fail(fame6d_krinv_apply(s, FAME6D_FP64, x, y, mxGetScalar(prhs[3]), (int)mxGetScalar(prhs[4]), &st));
    plhs[0] = from_device(y, n, 1);
    hipFree(x);
    hipFree(y);
  } else if (!strcmp(cmd, "eigs")) {
    fame6d_qsys* s = (fame6d_qsys*)handle_in(prhs[1]);
    int k = (int)mxGetScalar(prhs[2]), p = (int)mxGetScalar(prhs[3]);
    double tol = mxGetScalar(prhs[4]);
    int maxRestart = (int)mxGetScalar(prhs[5]);
    double tolCG = mxGetScalar(prhs[6]);
    int maxitCG = (int)mxGetScalar(prhs[7]);
    const mxArray* v0 = prhs[8];
    int prec = (nrhs > 9) ? (int)mxGetScalar(prhs[9]) : FAME6D_FP64;
    size_t n2 = mxGetNumberOfElements(v0);
    if (!mxIsComplex(v0)) mexErrMsgIdAndTxt("fame6d:usage", "v0 must be complex");
    plhs[0] = mxCreateDoubleMatrix(k, 1, mxREAL);
    void* Y = NULL;
    if (nlhs > 1) hipMalloc(&Y, n2 * k * 16);
    fame6d_lanczos_stats_t st;
    fail(fame6d_eigs_method_c(s, prec, k, p, tol, maxRestart, tolCG, maxitCG, mxGetComplexDoubles(v0),
                              mxGetDoubles(plhs[0]), Y, &st));
    if (nlhs > 1) {
      plhs[1] = from_device(Y, n2, k);
      hipFree(Y);
    }
  } else if (!strcmp(cmd, "topA")) {
    fame6d_qsys* s = (fame6d_qsys*)handle_in(prhs[1]);
    size_t n;
    void* x = to_device(prhs[2], &n);
    void* y = NULL;
    hipMalloc(&y, n * 16);
    fame6d_minres_stats_t st;
    fail(fame6d_method_a_apply(s, mxGetScalar(prhs[3]), mxGetScalar(prhs[4]), (int)mxGetScalar(prhs[5]), x, y,
                               &st));
    plhs[0] = from_device(y, n, 1);
    hipFree(x);
    hipFree(y);
  } else if (!strcmp(cmd, "eigsA")) {
    fame6d_qsys* s = (fame6d_qsys*)handle_in(prhs[1]);
    int k = (int)mxGetScalar(prhs[2]), p = (int)mxGetScalar(prhs[3]);
    double tol = mxGetScalar(prhs[4]);
    int maxRestart = (int)mxGetScalar(prhs[5]);
    double sigma = mxGetScalar(prhs[6]), tolM = mxGetScalar(prhs[7]);
    int maxitM = (int)mxGetScalar(prhs[8]);
    const mxArray* v0 = prhs[9];
    size_t n3 = mxGetNumberOfElements(v0);
    if (!mxIsComplex(v0)) mexErrMsgIdAndTxt("fame6d:usage", "v0 must be complex");
    plhs[0] = mxCreateDoubleMatrix(k, 1, mxREAL);
    void* Y = NULL;
    if (nlhs > 1) hipMalloc(&Y, n3 * k * 16);
    fame6d_lanczos_stats_t st;
    fame6d_minres_stats_t ms;
    fail(fame6d_eigs_method_a(s, k, p, tol, maxRestart, sigma, tolM, maxitM, mxGetComplexDoubles(v0),
                              mxGetDoubles(plhs[0]), Y, &st, &ms));
    if (nlhs > 1) {
      plhs[1] = from_device(Y, n3, k);
      hipFree(Y);
    }
  } else if (!strcmp(cmd, "tracking")) {
    fame6d_qsys* s = (fame6d_qsys*)handle_in(prhs[1]);
    const int k = (int)mxGetScalar(prhs[2]), p = (int)mxGetScalar(prhs[3]);
    const double tol = mxGetScalar(prhs[4]);
    const int maxRestart = (int)mxGetScalar(prhs[5]);
    const double tolCG = mxGetScalar(prhs[6]);
    const int maxitCG = (int)mxGetScalar(prhs[7]);
    const mxArray* v0 = prhs[8];
    if (!mxIsComplex(v0)) mexErrMsgIdAndTxt("fame6d:usage", "v0 must be complex");
    const size_t n2 = mxGetNumberOfElements(v0);
    const int nfix = (int)mxGetNumberOfElements(prhs[9]);
    if (nfix < 1 || nfix > 256) mexErrMsgIdAndTxt("fame6d:usage", "fixedIdx must hold 1..256 indices");
    int fixed[256], sel[256];
    for (int i = 0; i < nfix; ++i) fixed[i] = (int)mxGetDoubles(prhs[9])[i];
    plhs[0] = mxCreateDoubleMatrix(k, 1, mxREAL);
    mxArray* rel = mxCreateDoubleMatrix(k, 1, mxREAL);
    void* Y = NULL;
    hipMalloc(&Y, n2 * nfix * 16);
    int nsel = 0;
    fame6d_lanczos_stats_t st;
    fail(fame6d_eigs_tracking(s, FAME6D_FP64, k, p, tol, maxRestart, tolCG, maxitCG, mxGetComplexDoubles(v0), nfix,
                              fixed, 0, NULL, 0, mxGetDoubles(plhs[0]), mxGetDoubles(rel), sel, nfix, &nsel, Y, &st));
    if (nlhs > 1) plhs[1] = from_device(Y, n2, nsel);
    hipFree(Y);
    if (nlhs > 2) {
      plhs[2] = mxCreateDoubleMatrix(nsel, 1, mxREAL);
      for (int i = 0; i < nsel; ++i) mxGetDoubles(plhs[2])[i] = sel[i];
    }
    if (nlhs > 3) plhs[3] = rel;
  } else if (!strcmp(cmd, "yee_metrics")) {
    /* cfg.lwrq of exp3/make_config_experiment4.m:72-74 */
    const int n = (int)mxGetScalar(prhs[1]);
    const double h = mxGetScalar(prhs[2]);
    size_t m3[3];
    void* u[3];
    for (int c = 0; c < 3; ++c) u[c] = to_device(prhs[3 + c], &m3[c]);
    const size_t mm = (size_t)(2 * n + 3);
    if (m3[0] != mm * mm * mm || m3[1] != m3[0] || m3[2] != m3[0])
      mexErrMsgIdAndTxt("fame6d:usage", "fieldPlus size does not match requested n");
    const double* P = mxGetDoubles(prhs[7]); /* 6 x 3 column-major in MATLAB */
    double Prow[48];
    const int rows = (int)mxGetM(prhs[7]);
    if (rows < 1 || rows > 16 || mxGetN(prhs[7]) != 3) mexErrMsgIdAndTxt("fame6d:usage", "P must be rows x 3");
    for (int r = 0; r < rows; ++r)
      for (int d = 0; d < 3; ++d) Prow[r * 3 + d] = P[r + rows * d];
    fame6d_yee_metrics_t o;
    fail(fame6d_yee_lwrq_metrics(n, h, u[0], u[1], u[2], Prow, rows, mxGetScalar(prhs[8]), mxGetScalar(prhs[9]),
                                 mxGetScalar(prhs[6]), 3, 2.0, 1e-14, NULL, NULL, &o, NULL));
    for (int c = 0; c < 3; ++c) hipFree(u[c]);
    const char* names[] = {"lambdaCrop", "lambda3D", "eta3D", "etaUsing6DLambda", "deltaLambda", "VLWRQ", "nuLWRQ",
                           "deltaPart", "deltaLWRQIdentity", "croppedMass"};
    const double vals[] = {o.lambda_crop, o.lambda_3d, o.eta_3d, o.eta_6d, o.delta_lambda,
                           o.V, o.nu, o.delta_part, o.delta_identity, o.cropped_mass};
    plhs[0] = mxCreateStructMatrix(1, 1, 10, names);
    for (int i = 0; i < 10; ++i) mxSetField(plhs[0], 0, names[i], mxCreateDoubleScalar(vals[i]));
  } else {
    mexErrMsgIdAndTxt("fame6d:usage", "unknown command %s", cmd);
  }
}
