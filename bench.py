"""Benchmark of the 6D Fourier projection eigensolver hot path on one MI355X.

Metric: "N=5 time-to-10-eigenpairs + M-apply HBM GB/s; mixed vs FP64 at 1e-8 rel".

Workload (the N=5 configuration the metric is quoted on, plain FP64 on the GPU): 6D
Fourier grid 10^6 modes, 2e6 complex unknowns, eps(y) = 1 + (0.4/6)*sum_k cos(2*pi*y_k +
phi_k) with seeded phases, q = (0.05, 0.04, 0.03), Method-C shift-invert Lanczos
(k = 10, p = 40, tol 1e-10) with PCG(Mhat) inner solves to 1e-10.  One step = one
complete FP64 solve for the 10 lowest positive eigenpairs.  Beside it, the
mixed-precision configuration (amplitude 0.8, N = 5 and 6) is solved in FP64 and in the
mixed form (FP32 mass FFTs and inner PCG, FP64 refinement, FP64 Lanczos) and the
per-mode eigenvalue differences and the speedup are reported.  Inputs are synthetic and
device-resident before the timed region.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
Q = (0.05, 0.04, 0.03)
N_BENCH = 5
AMP = 0.4
EPS_SEED = 0
LANCZOS = dict(k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=1200)
# a timed launch carries a start/stop event pair, which costs a few microseconds of device
# time; one launch in TIMER_STRIDE of each kernel family is timed
TIMER_STRIDE = 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=N_BENCH)
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware-counter collection: a 64 MiB device copy (calibration), "
                         "then one FP64 solve of the benchmark configuration")
    ap.add_argument("--probe-taylor", action="store_true",
                    help="short run for hardware-counter collection of the Taylor moments contraction: "
                         "N=4 (NF = 8^6), p = 10 (286 monomials), 216 centers, 4 launches")
    ap.add_argument("--probe-batch", action="store_true",
                    help="short run for the kernel trace of the batched Bloch path: N=4, 12 points, "
                         "FP64, two batched solves and the same 12 points one after the other")
    ap.add_argument("--tiles", default=None,
                    help="FP64 tile variants 'ac,b' (fame6d_mass_set_tiles codes) instead of the autotuner; "
                         "counter runs use the choice the timed run printed (its config.tile_variants)")
    ap.add_argument("--extras", action="store_true",
                    help="also time the batched 12-point Bloch path at N=4 and the Taylor reconstruction")
    return ap.parse_args()


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def time_applies(torch, mass, dtype, reps, comps=1):
    """Average duration of one scalar M_s apply (3 pass kernels; comps=3: the three-component
    Mop, whose passes take the three vectors together) from HIP events recorded on the stream
    the library launches on."""
    x = torch.randn(comps * mass.NF, dtype=dtype, device=mass.device)
    op = mass.MscalarOp if comps == 1 else mass.Mop
    for _ in range(3):
        y = op(x)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(mass.stream)
    for _ in range(reps):
        y = op(x)
    e1.record(mass.stream)
    e1.synchronize()
    del y
    return e0.elapsed_time(e1) / reps * 1e-3


def cpu_baseline(pkg, eps, N, gpu_stats, n3_gpu=None):
    """Oracle (numpy, FP64; its FFTs threaded over the host's cores) on a bounded sample:
    3 scalar M_s applies and 6 PCG(Mhat) iterations at the benchmark size, extrapolated to the
    GPU FP64 run's counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fame6d_oracle as O

    workers = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    O.set_fft_workers(workers)
    NF = (2 * N) ** 6
    rng = np.random.default_rng(0)
    x = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    O.mass_apply(eps, x, N)  # warm-up (FFT plans, thread pool)
    t = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.mass_apply(eps, x, N)
        t.append(time.perf_counter() - t0)
    t_apply = float(np.median(t))
    sys_ = O.qsystem(N, Q)
    b = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    iters = 6
    t0 = time.perf_counter()
    O.pcg(lambda v: O.mhat_apply(sys_, eps, v, N), b, 1e-30, iters)
    t_iter = (time.perf_counter() - t0) / iters
    kr = gpu_stats["krinv"]
    calls = kr["calls"]
    its = kr["total_fp64_iterations"]
    # the oracle follows the reference: 4 scalar applies around each K_r^-1 call, one Mhat apply
    # per PCG iteration and one true-residual check per call
    checks = calls
    est = calls * 4 * t_apply + (its + checks) * t_iter
    # a complete solve, not extrapolated, at the plain-method size N = 3 (about half a minute)
    full = None
    if n3_gpu is not None:
        N3 = 3
        eps3, _ = _cosine_eps(N3)
        v3 = np.random.default_rng(20260905)
        v3 = v3.standard_normal(2 * (2 * N3) ** 6) + 1j * v3.standard_normal(2 * (2 * N3) ** 6)
        t0 = time.perf_counter()
        lam3, _, info3, _ = O.eigs_method_c(N3, eps3, Q, v3, k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10,
                                            maxit_cg=1200)
        t3 = time.perf_counter() - t0
        full = dict(N=N3, cpu_seconds=t3, gpu_seconds=n3_gpu["seconds"], gpu_speedup=t3 / n3_gpu["seconds"],
                    max_rel_diff_lambda=float(np.max(np.abs(lam3 - n3_gpu["lam"]) / lam3)),
                    krinv_calls=info3["calls"], pcg_iterations=info3["iters"],
                    what="oracle Method-C solve for the 10 lowest eigenvalues, same inputs as the GPU solve")
    O.set_fft_workers(None)
    return dict(value=est, unit="s", cores=workers, kind="port",
                sample="numpy oracle at N=%d, its 6D FFTs on %d threads (scipy.fft; the elementwise steps are "
                       "numpy, one thread): median of 3 M_s applies (%.3f s) and 6 PCG(Mhat) iterations "
                       "(%.3f s/iter), extrapolated to the GPU FP64 solve's %d K_r^-1 calls and %d PCG "
                       "iterations (extrapolated)" % (N, workers, t_apply, t_iter, calls, its),
                cpu_model=cpu_model(), apply_seconds=t_apply, apply_gbs=40.0 * NF / t_apply / 1e9,
                full_solve_n3=full)


def _cosine_eps(N):
    """The benchmark permittivity restated for the oracle: 1 + (AMP/6) sum_k cos(2 pi y_k + phi_k),
    phases from numpy default_rng(EPS_SEED)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fame6d_oracle as O
    ph = np.random.default_rng(EPS_SEED).uniform(0.0, 2.0 * np.pi, 6)
    return O.eps_cosine(N, AMP, ph), ph


def gpu_solve_n3(torch, pkg):
    """Plain FP64 Method-C solve at N = 3 on the GPU (the configuration of the full CPU solve)."""
    eps, _ = pkg.cosine_material(3, AMP, EPS_SEED)
    mass = pkg.MassBackend(3, eps)
    mass.autotune(reps=3)
    s = pkg.QSystem(mass, Q)
    v0 = pkg.start_vector(s.nr, 20260905)
    pkg.run_method_C(s, v0=v0, precision="double", **LANCZOS)  # warm-up
    t, out = solve_timed(torch, pkg, s, v0, "double", reps=3)
    s.close()
    mass.close()
    return dict(seconds=t, lam=out["lambda_"])


def torch_mod():
    import torch
    return torch


def bloch_path(pkg, N=4, npts=12):
    """The batched-path workload: N, points t*(0.5,0.4,0.3), t = k/npts, k = 1..npts (q = 0
    excluded because the null-space-free reduction is undefined there), seeded start vectors."""
    eps, _ = pkg.cosine_material(N, AMP, EPS_SEED)
    mass = pkg.MassBackend(N, eps)
    mass.autotune(reps=3)
    qs = np.array([[0.5 * t, 0.4 * t, 0.3 * t] for t in np.arange(1, npts + 1) / float(npts)])
    # start vectors resident on the device before any timed solve
    v0 = torch_mod().as_tensor(np.stack([pkg.start_vector(2 * mass.NF, 100 + i) for i in range(npts)]),
                               device="cuda")
    return mass, qs, v0


def probe_batch(torch, pkg):
    mass, qs, v0 = bloch_path(pkg)
    rows = []
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = pkg.run_method_C_batch(mass, qs, v0=v0, precision="double", **LANCZOS)
        torch.cuda.synchronize()
        rows.append(("batched", time.perf_counter() - t0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(qs)):
        s_i = pkg.QSystem(mass, qs[i])
        pkg.run_method_C(s_i, v0=v0[i], precision="double", **LANCZOS)
        s_i.close()
    torch.cuda.synchronize()
    rows.append(("sequential", time.perf_counter() - t0))
    print(json.dumps({"probe": "batch", "N": 4, "points": len(qs), "seconds": rows,
                      "outer_steps": [st["outer_steps"] for st in out["stats"]]}))
    mass.close()


def extras(torch, pkg):
    """Batched Bloch path (N=4, 12 points on t*(0.5,0.4,0.3), t = k/12, k = 1..12; q = 0
    excluded because the null-space-free reduction is undefined there) and the Taylor
    reconstruction of mode 10 of the last point: p = 10 (286 terms), 6^3 = 216 centers,
    h = 0.0125, (2n+3)^3 Yee points per component for n = 40..200 step 40."""
    import ctypes as C
    from importlib import import_module
    lib = import_module("fame6d_amd._lib")
    N = 4
    mass, qs, v0 = bloch_path(pkg, N)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outb = pkg.run_method_C_batch(mass, qs, v0=v0, precision="double", **LANCZOS)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # same 12 points one after the other (the reference's sequential loop)
    t0 = time.perf_counter()
    lam_seq = []
    for i in range(12):
        s_i = pkg.QSystem(mass, qs[i])
        lam_seq.append(pkg.run_method_C(s_i, v0=v0[i], precision="double", **LANCZOS)["lambda_"])
        s_i.close()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    maxdiff = float(np.max(np.abs(outb["lambda_"] - np.array(lam_seq)) / np.array(lam_seq)))
    # mode 10 of the last point
    sq = pkg.QSystem(mass, qs[-1])
    out = pkg.run_method_C(sq, v0=v0[-1], precision="double", want_vectors=True, **LANCZOS)
    X, MX, _ = pkg.recover_original_eigenvectors(out["Y"][9:10], sq)
    qinfo = sq.info
    # projected wave vectors of the grid modes (q_xi = P^T (xi + k_lift))
    f = np.concatenate([np.arange(N), np.arange(-N, 0)]).astype(float)
    grids = np.meshgrid(*([f] * 6), indexing="ij")
    Xi = [g.reshape(-1, order="F") for g in grids]
    P = pkg.P_DEFAULT
    qm = np.zeros((mass.NF, 3))
    for l in range(6):
        for c in range(3):
            qm[:, c] += P[l, c] * (Xi[l] + qinfo["kLift"][l])
    qm_d = torch.as_tensor(qm.T.copy(), device="cuda")
    p_t, nb, h = 10, 6, 0.0125
    L = (p_t + 1) * (p_t + 2) * (p_t + 3) // 6
    rows = []
    for n in range(40, 201, 40):
        m = 2 * n + 3
        base, rem = divmod(m, nb)
        bounds = []
        st_ = 0
        for b in range(nb):
            ln = base + (1 if b < rem else 0)
            bounds.append((st_, st_ + ln - 1))
            st_ += ln
        u = torch.empty(m ** 3, dtype=torch.complex128, device="cuda")
        mom = torch.empty((nb ** 3, L), dtype=torch.complex128, device="cuda")
        t_mom = t_eval = 0.0
        for comp in range(3):
            mid = [h * ((lo + hi) / 2.0 - (n + 1)) for lo, hi in bounds]
            cen = np.array([[mid[i] + (0.5 * h if comp == 0 else 0.0), mid[j] + (0.5 * h if comp == 1 else 0.0),
                             mid[k] + (0.5 * h if comp == 2 else 0.0)]
                            for k in range(nb) for j in range(nb) for i in range(nb)])
            coeff = X[0, comp * mass.NF:(comp + 1) * mass.NF].contiguous()
            torch.cuda.synchronize()
            a0 = time.perf_counter()
            lib.call("fame6d_taylor_moments", mass.NF, lib.ptr(qm_d), lib.ptr(coeff), p_t, nb ** 3,
                     cen.ctypes.data_as(C.c_void_p), lib.ptr(mom), C.c_void_p(torch.cuda.current_stream().cuda_stream))
            torch.cuda.synchronize()
            a1 = time.perf_counter()
            lib.call("fame6d_taylor_evaluate_yee", n, h, comp, nb, p_t, nb ** 3, cen.ctypes.data_as(C.c_void_p),
                     lib.ptr(mom), lib.ptr(u), C.c_void_p(torch.cuda.current_stream().cuda_stream))
            torch.cuda.synchronize()
            a2 = time.perf_counter()
            t_mom += a1 - a0
            t_eval += a2 - a1
        flops = 3 * 2.0 * L * mass.NF * 2 * nb ** 3
        rows.append(dict(n=n, points_per_component=m ** 3, moments_s=t_mom, eval_s=t_eval,
                         moments_TFLOPs=flops / t_mom / 1e12, mfma_fp64_util=flops / t_mom / 78.6e12))
        del u
    sq.close()
    mass.close()
    return dict(batched_12pt_N4_seconds=t_batch, sequential_12pt_N4_seconds=t_seq,
                batched_vs_sequential_max_rel_diff=maxdiff, taylor=rows)


def plane_wave_fields(torch, n, h, kvec, p):
    """Staggered plane wave u_c = p_c exp(i k . x_c) on the plus grid, flat column-major."""
    I = torch.arange(-(n + 1), n + 2, dtype=torch.float64, device="cuda") * h
    u = []
    for c in range(3):
        ax = [I + (h / 2 if d == c else 0.0) for d in range(3)]
        ph = kvec[0] * ax[0][None, None, :] + kvec[1] * ax[1][None, :, None] + kvec[2] * ax[2][:, None, None]
        u.append((complex(p[c]) * torch.exp(1j * ph)).reshape(-1))
    return u


def yee_closure(torch, pkg):
    """Experiment 4 (exp3/run_experiment4.m:40-95, exp3/make_config_experiment4.m:17-85): N = 5,
    q = (0.13, 0.21, 0.17), eps = 4 exp(1.2/6 sum cos); the selected-mode scan (k = 140,
    p = 188) for modes 11 and 114 with the reference's eigenvalue check; Taylor p = 10 with the
    automatic phase-radius partition; the cropped Yee / LWRQ metrics of the mesh study
    (h = 0.1, 0.05, 0.025, 0.0125 at half-width 0.5).  Then the closure kernels alone on the
    largest window of the window study (n = 160, h = 0.025) with an exact discrete plane wave:
    algorithmic bytes 80 per interior point (fields 48 in and normalised densities 16 out in
    pass 1, the densities 16 in for the fused x/y/z convolution and its reductions)."""
    N = 5
    eps = pkg.smooth_scalar_material(N)
    mass = pkg.MassBackend(N, eps)
    mass.autotune(reps=3)
    sq = pkg.QSystem(mass, (0.13, 0.21, 0.17))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sel = pkg.solve_tracked_modes(sq)
    torch.cuda.synchronize()
    t_eig = time.perf_counter() - t0
    modes = sel["modes"]
    lam = sel["lambda6D"]
    X = sel["X"]
    T = pkg.taylor
    q_d = torch.as_tensor(T.projected_q_modes(N, sq.info["kLift"], pkg.P_DEFAULT).T.copy(), device="cuda")
    rows = []
    for h, n in ((0.1, 5), (0.05, 10), (0.025, 20), (0.0125, 40)):
        for i, mode in enumerate(modes):
            torch.cuda.synchronize()
            a0 = time.perf_counter()
            u, nb = T.reconstruct_yee_field(X[i], q_d, n, h)
            torch.cuda.synchronize()
            a1 = time.perf_counter()
            met = pkg.compute_yee_lwrq_metrics(*u, float(lam[i]), n, h)
            torch.cuda.synchronize()
            a2 = time.perf_counter()
            rows.append(dict(mode=mode, h=h, n=n, blocks=nb, reconstruct_s=a1 - a0, metrics_s=a2 - a1,
                             lambda_6d=float(lam[i]), lambda_3d=met["lambda_3d"], eta_3d=met["eta_3d"],
                             eta_6d=met["eta_6d"], nu=met["nu"], delta_lambda=met["delta_lambda"],
                             delta_identity=met["delta_identity"], delta_part=met["delta_part"]))
            del u
    sq.close()
    mass.close()
    # closure kernels on the n = 160 window
    n, h = 160, 0.025
    kvec = np.array([1.3, -0.7, 2.1])
    sig = 2.0 / h * np.sin(kvec * h / 2)
    pw = np.cross(sig, [0.3, 1.0, -0.4])
    u = plane_wave_fields(torch, n, h, kvec, pw / np.linalg.norm(pw))
    work = pkg.yee.workspace(n)
    lam_pw = float(sig @ sig) / 4.0
    pkg.compute_yee_lwrq_metrics(*u, lam_pw, n, h, eps_c=4.0, alpha=0.0, work=work)
    reps = 5
    torch.cuda.synchronize()
    a0 = time.perf_counter()
    for _ in range(reps):
        met = pkg.compute_yee_lwrq_metrics(*u, lam_pw, n, h, eps_c=4.0, alpha=1.2, work=work)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - a0) / reps
    q3 = (2 * n + 1) ** 3
    ref = np.array([1.200090550277e-02, 4.462939906027e-02])
    st = sel["scan"]["stats"]
    return dict(tracking_scan=dict(k=140, p=188, seconds=t_eig, outer_steps=st["outer_steps"],
                                   restarts=st["restarts"], lambda_11_114=[float(x) for x in lam],
                                   rel_diff_vs_reference=[float(x) for x in np.abs(lam - ref) / ref]),
                mesh=rows, window_n160=dict(n=n, h=h, interior_points=q3, seconds=dt,
                                           algorithmic_GBs=80.0 * q3 / dt / 1e9))


def method_comparison(torch, pkg, N):
    """Methods A (MINRES shift-invert on the full 3NF space), B (nested PCG) and C on the
    reference's own Fig. 1(b,c) configuration (exp1/make_config_exp1bc.m:17-87: N = 5,
    eps = 4 exp(1.2/6 sum cos), q = (0.05, 0.04, 0.03), k = 10, p = 40), each against
    Method C (the reference warns above 1e-7, exp1/check_fig1bc_results.m:23-25)."""
    mass = pkg.MassBackend(N, pkg.smooth_scalar_material(N))
    mass.autotune(reps=3)
    sysq = pkg.QSystem(mass, Q)
    out = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    c = pkg.run_method_C(sysq, v0=pkg.start_vector(sysq.nr, 20260905), **LANCZOS)
    torch.cuda.synchronize()
    lam_c = c["lambda_"]
    out["C"] = dict(seconds=time.perf_counter() - t0, outer_steps=c["stats"]["outer_steps"],
                    pcg_iterations=c["stats"]["krinv"]["total_iterations"], lambda_=lam_c.tolist())
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a = pkg.run_method_A(sysq, k=10, p=40, tol=1e-10, max_restart=300, sigma=1e-4, tol_minres=1e-10,
                             maxit_minres=4000, v0=pkg.start_vector(sysq.ndof, 20260906))
        torch.cuda.synchronize()
        out["A"] = dict(seconds=time.perf_counter() - t0, outer_steps=a["stats"]["outer_steps"],
                        minres=a["minres"],
                        max_rel_diff_vs_C=float(np.max(np.abs(a["lambda_"] - lam_c) / lam_c)))
    except pkg.Fame6dError as e:  # the reference's controlled stop
        out["A"] = dict(error=str(e))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = pkg.run_method_B(sysq, v0=pkg.start_vector(sysq.nr, 20260905), **{k: v for k, v in LANCZOS.items()
                                                                          if k in ("k", "p", "tol", "max_restart")})
    torch.cuda.synchronize()
    out["B"] = dict(seconds=time.perf_counter() - t0, outer_steps=b["stats"]["outer_steps"],
                    max_rel_diff_vs_C=float(np.max(np.abs(b["lambda_"] - lam_c) / lam_c)))
    sysq.close()
    mass.close()
    return out


# kernel family (library timer slot) -> kernel-name pattern in rocprofv3 output (%d: nGrid)
FAMILY_KERNEL = {
    "mhat_pass_a": r"k_pass_a3[ps]?<double, %d, \d+, 2>",
    "mhat_pass_b": r"k_pass_b<double, %d, \d+, 3, 0>",
    "mhat_pass_c": r"k_pass_c3p?<double, %d, \d+, true>",
    "lanczos_gemv_c": r"k_gemv_c_all<\d+, (true|false)>",
    "lanczos_gemv_n": r"k_gemv_n_out",
}
FAMILY_DESC = {
    "mhat_pass_a": "Mhat pass A in PCG (p = r + beta p; z, W1 z, W2 z; inverse DFT dims 1-3): "
                   "operand bytes NF*(6*16 + 3*8) per launch",
    "mhat_pass_b": "Mhat pass B (inverse DFT dims 4-6, *eps/NF, forward DFT dims 6-4, p'q partials; 3 vectors): "
                   "operand bytes NF*(6*16 + 8) per launch",
    "mhat_pass_c": "Mhat pass C in PCG (forward DFT dims 3-1, q = invC(s0 + W1 s1 + W2 s2) in registers, "
                   "alpha, x += alpha p, r -= alpha q, ||r||^2 partials; invC recomputed from W1, W2): "
                   "operand bytes NF*(8*16 + 2*8) per launch",
    "lanczos_gemv_c": "Lanczos CGS2 sweep h = Q(:,1:j)^H w: (j+1)*2NF*16 bytes per sweep",
    "lanczos_gemv_n": "Lanczos CGS2 sweep r = w - Q(:,1:j) h (+||r||^2): (j+2)*2NF*16 bytes per sweep",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "method_c_N5_counters.md")
MHAT_FAMILIES = ("mhat_pass_a", "mhat_pass_b", "mhat_pass_c")
ALG_MHAT_ITER_PER_NF = 3 * 40 + 5 * 16  # bytes per mode of one PCG(Mhat) iteration


def pmc_traffic(pattern):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by its 64 MiB copy calibration; the
    variant with the most dispatches when several tile variants match), or None."""
    import re
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith("traffic-json:"):
            tab = json.loads(line.split(":", 1)[1])
            rows = [(name, row) for name, row in tab["kernels"].items() if re.fullmatch(pattern, name)]
            if not rows:
                return None
            name, row = max(rows, key=lambda kv: kv[1]["dispatches"])
            return dict(bytes=row["fetch_kib"] * 1024 * tab["fetch_scale"] + row["write_kib"] * 1024 * tab["write_scale"],
                        kernel=name, source=os.path.relpath(TRAFFIC_FILE, ROOT))
    return None


def solve_timed(torch, pkg, sysq, v0, precision, reps=1):
    best = None
    out = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = pkg.run_method_C(sysq, v0=v0, precision=precision, **LANCZOS)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return best, out


def vector_errors(torch, lam, Y64, Ymx, cluster_rel=1e-3):
    """Per-mode eigenvector error of the mixed run against FP64: sin of the angle between
    the mixed Ritz vector and the span of the FP64 Ritz vectors of its eigenvalue cluster
    (modes within cluster_rel relative; the Bloch modes come in near-degenerate pairs whose
    individual vectors are not determined).  Ritz vectors are orthonormal."""
    G = (Y64.conj() @ Ymx.transpose(0, 1)).abs().pow(2).cpu().numpy()  # G[j, i] = |<y64_j, ymx_i>|^2
    out = []
    for i in range(len(lam)):
        cl = [j for j in range(len(lam)) if abs(lam[j] - lam[i]) <= cluster_rel * abs(lam[i])]
        out.append(float(np.sqrt(max(0.0, 1.0 - G[cl, i].sum()))))
    return out


def differentiation(torch, pkg, N=5, k=5, ndir=64, h=1e-6):
    """Gradients of the k lowest eigenvalues at q = (0.05, 0.04, 0.03) with respect to all NF
    permittivity samples (Hellmann-Feynman through the adjoint mass kernel, behind
    torch.autograd), checked against central differences with step h on ndir seeded random
    directions of the weighted sum of the eigenvalues."""
    from importlib import import_module
    ag = import_module("fame6d_amd.autograd")
    eps0, _ = pkg.cosine_material(N, AMP, EPS_SEED)
    mass = pkg.MassBackend(N, eps0)
    mass.autotune(reps=3)
    opts = dict(p=40, tol=1e-12, max_restart=60, tol_cg=1e-13, maxit_cg=5000)
    wts = torch.tensor([1.0, 0.5, -0.25, 0.125, 2.0][:k], dtype=torch.float64, device="cuda")
    eps = torch.as_tensor(eps0, device="cuda").requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam = ag.lowest_eigenvalues(eps, mass, Q, k=k, **opts)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    (lam * wts).sum().backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    g = eps.grad.cpu().numpy()
    rng = np.random.default_rng(64)
    devs, fds = [], []
    t0 = time.perf_counter()
    for _ in range(ndir):
        d = rng.standard_normal(mass.NF)
        lp = ag.lowest_eigenvalues(torch.as_tensor(eps0 + h * d, device="cuda"), mass, Q, k=k, **opts)
        lm = ag.lowest_eigenvalues(torch.as_tensor(eps0 - h * d, device="cuda"), mass, Q, k=k, **opts)
        fd = float(((lp - lm) / (2 * h) * wts).sum())
        fds.append(fd)
        devs.append(abs(fd - float(g @ d)))
    t_fd = time.perf_counter() - t0
    mass.close()
    fds = np.array(fds)
    devs = np.array(devs)
    rms = float(np.sqrt(np.mean(fds ** 2)))
    return dict(N=N, k=k, directions=ndir, step=h, lambda_=lam.detach().cpu().numpy().tolist(),
                forward_seconds=t_fwd, backward_seconds=t_bwd, fd_seconds=t_fd,
                median_rel_dev_fd=float(np.median(devs / np.abs(fds))),
                max_dev_over_rms_fd=float(devs.max() / rms), rms_fd=rms)


def mixed_comparison(torch, pkg, N, amp, reps):
    """Mixed vs FP64 on the amplitude-0.8 configuration: per-mode relative eigenvalue
    differences and eigenvector angles."""
    eps, _ = pkg.cosine_material(N, amp, EPS_SEED)
    mass = pkg.MassBackend(N, eps)
    mass.autotune(reps=5)
    sysq = pkg.QSystem(mass, Q)
    v0 = torch.as_tensor(pkg.start_vector(sysq.nr, 20260905), device="cuda")  # resident before timing
    solve_timed(torch, pkg, sysq, v0, "mixed")  # warm-up (workspaces)
    t64, o64 = solve_timed(torch, pkg, sysq, v0, "double", reps)
    tmx, omx = solve_timed(torch, pkg, sysq, v0, "mixed", reps)
    rel = (np.abs(omx["lambda_"] - o64["lambda_"]) / np.abs(o64["lambda_"])).tolist()
    # Ritz vectors of both forms (untimed repeats of the same solves)
    v64 = pkg.run_method_C(sysq, v0=v0, precision="double", want_vectors=True, **LANCZOS)
    vmx = pkg.run_method_C(sysq, v0=v0, precision="mixed", want_vectors=True, **LANCZOS)
    vec = vector_errors(torch, v64["lambda_"], v64["Y"], vmx["Y"])
    del v64, vmx
    sysq.close()
    mass.close()
    return dict(N=N, amp=amp, fp64_seconds=t64, mixed_seconds=tmx, speedup=t64 / tmx,
                rel_err_per_mode=rel, max_rel_err=max(rel), vec_sin_per_mode=vec, max_vec_sin=max(vec),
                lambda_fp64=o64["lambda_"].tolist(),
                fp64_krinv=o64["stats"]["krinv"], mixed_krinv=omx["stats"]["krinv"],
                outer_steps=o64["stats"]["outer_steps"])


def tune(mass, tiles):
    if tiles:
        va, vb = (int(v) for v in tiles.split(","))
        mass.set_tiles("double", va, vb)
    else:
        mass.autotune(reps=5)
    return mass.info["tiles"]


def probe(torch, pkg, N, tiles):
    """Fixed kernel sequence for counter runs: a 64 MiB device copy (calibration), then one
    FP64 solve of the benchmark configuration (the kernels and launch shapes of a step)."""
    eps, _ = pkg.cosine_material(N, AMP, EPS_SEED)
    mass = pkg.MassBackend(N, eps)
    tune(mass, tiles)
    sysq = pkg.QSystem(mass, Q)
    v0 = pkg.start_vector(sysq.nr, 20260905)
    src = torch.ones(64 * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    out = pkg.run_method_C(sysq, v0=v0, precision="double", **LANCZOS)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "done", "N": N, "copy_bytes": 64 * 2 ** 20, "tiles": mass.info["tiles"],
                      "outer_steps": out["stats"]["outer_steps"]}), flush=True)


def probe_taylor(torch, pkg):
    """Fixed Taylor workload for profiler runs: the moment GEMM of the reconstruction
    configuration (N=4, p=10, 6^3 centers) on synthetic projected modes and coefficients, four
    times with the block centres of a Yee window (a tensor grid: per-axis phase tables) and
    four times with scattered centres (one phase per mode and centre), then four Yee
    evaluations of one component on the n = 200 window."""
    import numpy as np
    NF, p, nb = 8 ** 6, 10, 6
    rng = np.random.default_rng(0)
    q = torch.as_tensor(rng.uniform(-10, 10, (NF, 3)), device="cuda")
    coeff = torch.as_tensor(rng.standard_normal(NF) + 1j * rng.standard_normal(NF), device="cuda")
    cen_s = rng.uniform(-1, 1, (nb ** 3, 3)) * 0.05
    cen_g = np.ascontiguousarray(pkg.yee_block_centers(40, 0.0125, 0, nb))
    L = (p + 1) * (p + 2) * (p + 3) // 6
    flops = 2.0 * L * NF * 2 * nb ** 3  # complex coefficients against real monomials

    def timed(cen):
        for _ in range(2):  # warm-up: workspace, clocks
            pkg.taylor_moments(q, coeff, p, cen)
        ts = []
        for _ in range(4):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pkg.taylor_moments(q, coeff, p, cen)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    tg = timed(cen_g)
    ts_ = timed(cen_s)
    # the Yee evaluation of one component on the n = 200 window (403^3 plus-grid points,
    # h = 0.0125) with the moments of that window's 6^3 block centres
    n, h = 200, 0.0125
    m = 2 * n + 3
    cen_y = np.ascontiguousarray(pkg.yee_block_centers(n, h, 0, nb))
    mom = pkg.taylor_moments(q, coeff, p, cen_y)
    u = torch.empty(m ** 3, dtype=torch.complex128, device="cuda")
    from importlib import import_module
    import ctypes as C
    lib = import_module("fame6d_amd._lib")
    te = []
    for _ in range(4):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lib.call("fame6d_taylor_evaluate_yee", n, h, 0, nb, p, nb ** 3, cen_y.ctypes.data_as(C.c_void_p),
                 lib.ptr(mom.T.contiguous()), lib.ptr(u), C.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        te.append(time.perf_counter() - t0)
    print(json.dumps({"probe": "taylor", "NF": NF, "p": p, "monomials": L, "centers": nb ** 3,
                      "min_ms": tg * 1e3, "TFLOPs": flops / tg / 1e12,
                      "scattered_centres_min_ms": ts_ * 1e3, "scattered_centres_TFLOPs": flops / ts_ / 1e12,
                      "yee_eval_n": n, "yee_points": m ** 3, "yee_eval_min_ms": min(te) * 1e3}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(Bloch points would shard across GPUs, each GPU solving its own points with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__

    pkg = __graft_entry__.package()
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    N = a.N
    if a.probe:
        probe(torch, pkg, N, a.tiles)
        return
    if a.probe_taylor:
        probe_taylor(torch, pkg)
        return
    if a.probe_batch:
        probe_batch(torch, pkg)
        return
    NF = (2 * N) ** 6
    eps, phases = pkg.cosine_material(N, AMP, EPS_SEED)
    mass = pkg.MassBackend(N, eps)
    tiles = tune(mass, a.tiles)
    variants = mass.info["variants"]
    sysq = pkg.QSystem(mass, Q)
    # the start vector is an input: resident on the device before the timed region
    v0 = torch.as_tensor(pkg.start_vector(sysq.nr, 20260905), device="cuda")

    # ---- M-apply roofline (FP64 and FP32), HIP events on the library stream
    t64 = time_applies(torch, mass, torch.complex128, a.apply_reps)
    t32 = time_applies(torch, mass, torch.complex64, a.apply_reps)
    t64_3 = time_applies(torch, mass, torch.complex128, a.apply_reps, comps=3)
    bytes64 = 40.0 * NF
    bytes32 = 20.0 * NF
    gbs64 = bytes64 / t64 / 1e9
    gbs32 = bytes32 / t32 / 1e9

    # ---- FP64 steps (the configuration the metric is quoted on)
    for _ in range(a.warmup):
        pkg.run_method_C(sysq, v0=v0, precision="double", **LANCZOS)
    # kernel families timed live with HIP events on the library stream (every TIMER_STRIDE-th launch)
    mass.profile(TIMER_STRIDE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = []
    for _ in range(a.steps):
        outs.append(pkg.run_method_C(sysq, v0=v0, precision="double", **LANCZOS))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_step = elapsed / max(a.steps, 1)
    ktimes = mass.kernel_times()
    mass.profile(0)
    # Roofline unit: one PCG(Mhat) iteration, the three Mhat pass kernels A, B, C (about two
    # thirds of the step).  Algorithmic bytes are the survey's minimal accounting, regardless
    # of implementation: three scalar M_s applies at 40*NF bytes each plus the PCG vectors,
    # 5 * 16*NF bytes, i.e. 200*NF per iteration.  Time per iteration: the sum of the three
    # passes' average launch durations measured live (HIP events on the library stream).
    iter_us = sum(ktimes[f]["avg_us"] for f in MHAT_FAMILIES)
    alg_iter = ALG_MHAT_ITER_PER_NF * NF
    achieved = alg_iter / (iter_us * 1e-6) / 1e9
    traffic = None
    if N == 5:  # the committed counter summary is for N = 5
        parts = [pmc_traffic(FAMILY_KERNEL[f] % (2 * N)) for f in MHAT_FAMILIES]
        if all(parts):
            traffic = dict(bytes=sum(t["bytes"] for t in parts), kernels=[t["kernel"] for t in parts],
                           source=parts[0]["source"])
    pass_operand = {f: dict(avg_us=ktimes[f]["avg_us"], operand_bytes=ktimes[f]["bytes"] / ktimes[f]["launches"],
                            operand_GBs=ktimes[f]["GBs"]) for f in MHAT_FAMILIES}
    lam64 = outs[-1]["lambda_"]
    st64 = outs[-1]["stats"]
    # mixed form on the same configuration (reported, untimed by the step contract)
    tmix, omix = solve_timed(torch, pkg, sysq, v0, "mixed")
    rel_same = (np.abs(omix["lambda_"] - lam64) / np.abs(lam64)).tolist()
    sysq.close()
    mass.close()

    comparisons = []
    if not a.no_fp64:
        comparisons.append(mixed_comparison(torch, pkg, 5, 0.8, 1))
        comparisons.append(mixed_comparison(torch, pkg, 6, 0.8, 1))

    cpu = None
    if not a.no_cpu:
        n3 = gpu_solve_n3(torch, pkg)
        cpu = cpu_baseline(pkg, eps, N, st64, n3)
    ext = extras(torch, pkg) if a.extras else None
    if ext is not None:
        for key, fn in (("methods_vs_C", lambda: method_comparison(torch, pkg, N)),
                        ("yee_closure", lambda: yee_closure(torch, pkg)),
                        ("differentiation", lambda: differentiation(torch, pkg))):
            try:
                ext[key] = fn()
            except Exception as e:  # an extra that fails is reported, not fatal to the line
                ext[key] = dict(error="%s: %s" % (type(e).__name__, e))

    line = {
        "metric": "N=5 time-to-10-eigenpairs + M-apply HBM GB/s; mixed vs FP64 at 1e-8 rel",
        "value": per_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": per_step * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (f64)",
        "data": "synthetic (seeded cosine permittivity, seeded complex Gaussian start vector)",
        "config": {
            "workload": "N=%d 6D Fourier grid (%d modes, %d complex unknowns), eps=1+(%.1f/6)sum cos(2pi y_k+phi_k), "
                        "q=(0.05,0.04,0.03), Method C Lanczos k=10 p=40 tol=1e-10, PCG(Mhat) tol=1e-10, FP64"
                        % (N, NF, 2 * NF, AMP),
            "N": N, "NF": NF, "eps_phases": [float(p) for p in phases], "lanczos": LANCZOS,
            "tile_variants": {"double": "%d,%d" % tuple(variants["double"]),
                              "single": "%d,%d" % tuple(variants["single"])},
            "parallelism": "single GPU",
        },
        "lambda_fp64": lam64.tolist(),
        "fp64_stats": st64,
        "mixed_same_config": {"seconds": tmix, "speedup": per_step / tmix, "rel_err_per_mode": rel_same,
                              "max_rel_err": max(rel_same), "krinv": omix["stats"]["krinv"]},
        "mixed_vs_fp64": comparisons,
        "m_apply": {"fp64_us": t64 * 1e6, "fp64_alg_GBs": gbs64, "fp32_us": t32 * 1e6, "fp32_alg_GBs": gbs32,
                    "alg_bytes_fp64": bytes64, "alg_bytes_fp32": bytes32, "tiles": tiles},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_over_algorithmic": traffic["bytes"] / alg_iter if traffic else None,
                     "kernel": "one PCG(Mhat) iteration = Mhat passes A + B + C (%s)" % "; ".join(
                         FAMILY_DESC[f] for f in MHAT_FAMILIES),
                     "unit_of_work": "one PCG(Mhat) iteration: 3 scalar M_s applies x 40*NF B + PCG vectors "
                                     "5 x 16*NF B = 200*NF B",
                     "avg_launch_us": iter_us, "algorithmic_bytes_per_launch": alg_iter,
                     "traffic_source": traffic,
                     "m_apply_fp64": {"us": t64 * 1e6, "algorithmic_bytes": bytes64, "achieved": gbs64,
                                      "frac": gbs64 / HBM_PEAK_GBS},
                     "m_apply_fp32": {"us": t32 * 1e6, "algorithmic_bytes": bytes32, "achieved": gbs32,
                                      "frac": gbs32 / HBM_PEAK_GBS},
                     "m_apply_3comp_fp64": {"us": t64_3 * 1e6, "algorithmic_bytes": 3 * bytes64,
                                            "achieved": 3 * bytes64 / t64_3 / 1e9,
                                            "frac": 3 * bytes64 / t64_3 / 1e9 / HBM_PEAK_GBS,
                                            "what": "Mop: M_s on the three components of a 3NF vector, "
                                                    "one set of pass launches for the three"},
                     "pass_operand_bandwidth": pass_operand,
                     "timing": "start/stop HIP events of hipExtLaunchKernel on the library stream, every %dth "
                               "launch of each kernel family during the timed steps; the M_s apply from events "
                               "around %d back-to-back applies" % (TIMER_STRIDE, a.apply_reps)},
        "kernel_families": ktimes,
        "cpu_baseline": cpu,
        "extras": ext,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
