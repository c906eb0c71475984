"""Benchmark of the 6D Fourier projection eigensolver hot path on one MI355X.

Metric: "N=5 time-to-10-eigenpairs + M-apply HBM GB/s; mixed vs FP64 at 1e-8 rel".

Workload (the N=5 configuration the metric is quoted on): 6D Fourier grid 10^6 modes,
2e6 complex unknowns, eps(y) = 1 + (0.4/6)*sum_k cos(2*pi*y_k + phi_k) with seeded
phases, q = (0.05, 0.04, 0.03), Method-C shift-invert Lanczos (k = 10, p = 40,
tol 1e-10) with PCG(Mhat) inner solves to 1e-10.  One step = one complete solve for the
10 lowest positive eigenpairs in the mixed-precision form (FP32 mass FFTs and inner
PCG, FP64 refinement, FP64 Lanczos); the FP64 form is solved once beside it and the
per-mode eigenvalue differences are reported.  Inputs are synthetic and device-resident
before the timed region.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=N_BENCH)
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
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


def time_applies(torch, mass, dtype, reps):
    """Average duration of one scalar M_s apply (3 pass kernels) from HIP events recorded
    on the stream the library launches on."""
    x = torch.randn(mass.NF, dtype=dtype, device=mass.device)
    for _ in range(3):
        y = mass.MscalarOp(x)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(mass.stream)
    for _ in range(reps):
        y = mass.MscalarOp(x)
    e1.record(mass.stream)
    e1.synchronize()
    del y
    return e0.elapsed_time(e1) / reps * 1e-3


def cpu_baseline(pkg, eps, N, gpu_stats):
    """Oracle (numpy, FP64) on a bounded sample: 3 scalar M_s applies and 6 PCG(Mhat)
    iterations at the benchmark size, extrapolated to the GPU FP64 run's counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fame6d_oracle as O

    NF = (2 * N) ** 6
    rng = np.random.default_rng(0)
    x = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
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
    checks = max(kr["mass_applies_fp64"] - 4 * calls - 3 * its, 0) / 3.0
    est = calls * 4 * t_apply + (its + checks) * t_iter
    return dict(value=est, unit="s", cores=1, kind="port",
                sample="numpy oracle at N=%d: median of 3 M_s applies (%.3f s) and 6 PCG(Mhat) iterations "
                       "(%.3f s/iter), extrapolated to the GPU FP64 solve's %d K_r^-1 calls and %d PCG "
                       "iterations (extrapolated)" % (N, t_apply, t_iter, calls, its),
                cpu_model=cpu_model(), apply_seconds=t_apply, apply_gbs=40.0 * NF / t_apply / 1e9)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(Bloch points would shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__

    pkg = __graft_entry__.package()
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    N = a.N
    NF = (2 * N) ** 6
    eps, phases = pkg.cosine_material(N, AMP, EPS_SEED)
    mass = pkg.MassBackend(N, eps)
    tiles = mass.autotune(reps=5)
    sysq = pkg.QSystem(mass, Q)
    v0 = pkg.start_vector(sysq.nr, 20260905)

    # ---- M-apply roofline (FP64 and FP32), HIP events on the library stream
    t64 = time_applies(torch, mass, torch.complex128, a.apply_reps)
    t32 = time_applies(torch, mass, torch.complex64, a.apply_reps)
    bytes64 = 40.0 * NF
    bytes32 = 20.0 * NF
    gbs64 = bytes64 / t64 / 1e9
    gbs32 = bytes32 / t32 / 1e9

    # ---- FP64 reference solve (untimed by the step contract, reported)
    lam64 = None
    st64 = None
    t_fp64 = None
    if not a.no_fp64:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out64 = pkg.run_method_C(sysq, v0=v0, precision="double", **LANCZOS)
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        lam64 = out64["lambda_"]
        st64 = out64["stats"]

    # ---- mixed-precision steps
    for _ in range(a.warmup):
        pkg.run_method_C(sysq, v0=v0, precision="mixed", **LANCZOS)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = []
    for _ in range(a.steps):
        outs.append(pkg.run_method_C(sysq, v0=v0, precision="mixed", **LANCZOS))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_step = elapsed / max(a.steps, 1)
    lamm = outs[-1]["lambda_"]
    stm = outs[-1]["stats"]
    rel = None
    if lam64 is not None:
        rel = (np.abs(lamm - lam64) / np.abs(lam64)).tolist()

    cpu = None
    if not a.no_cpu and st64 is not None:
        cpu = cpu_baseline(pkg, eps, N, st64)

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
        "dtype": "c64 inner / f64 outer (mixed)",
        "data": "synthetic (seeded cosine permittivity, seeded complex Gaussian start vector)",
        "config": {
            "workload": "N=%d 6D Fourier grid (%d modes, %d complex unknowns), eps=1+(%.1f/6)sum cos(2pi y_k+phi_k), "
                        "q=(0.05,0.04,0.03), Method C Lanczos k=10 p=40 tol=1e-10, PCG(Mhat) tol=1e-10"
                        % (N, NF, 2 * NF, AMP),
            "N": N, "NF": NF, "eps_phases": [float(p) for p in phases], "lanczos": LANCZOS,
            "parallelism": "single GPU",
        },
        "fp64_seconds": t_fp64,
        "mixed_over_fp64_speedup": (t_fp64 / per_step) if t_fp64 else None,
        "lambda_fp64": None if lam64 is None else lam64.tolist(),
        "lambda_mixed": lamm.tolist(),
        "rel_err_mixed_vs_fp64_per_mode": rel,
        "max_rel_err_mixed_vs_fp64": None if rel is None else max(rel),
        "mixed_stats": stm,
        "fp64_stats": st64,
        "m_apply": {"fp64_us": t64 * 1e6, "fp64_alg_GBs": gbs64, "fp32_us": t32 * 1e6, "fp32_alg_GBs": gbs32,
                    "alg_bytes_fp64": bytes64, "alg_bytes_fp32": bytes32, "tiles": tiles},
        "roofline": {"bound": "hbm", "achieved": gbs64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs64 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "scalar M_s apply FP64 (passes A+B+C), 40*NF algorithmic bytes per apply"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
