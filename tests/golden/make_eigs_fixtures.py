"""Writes the eigenvalue fixtures of the benchmark workloads with the numpy FP64 oracle.

Each fixture is the oracle's Method-C solve (exp1/run_method_C_fig1bc.m:4-21: thick-restart
Lanczos on K_r^{-1}, exp1/Lanczos_hybrid_identity_gpuop.m:11-130, with PCG on Mhat,
exp1/make_explicit_krinv_compact.m:24-64) for the k lowest positive eigenvalues of one
workload.  The GPU tests compare the library's FP64 and mixed solves with these values per
mode, the way the reference checks its held eigenvalues
(exp3/solve_tracked_modes_exp4.m:37-42).

Inputs restated here, so the fixture does not depend on the product package:
  permittivity eps = 1 + (amp/6) sum_k cos(2 pi y_k + phi_k), phases phi_k drawn as
      numpy default_rng(seed).uniform(0, 2 pi, 6);
  Lanczos start vector v0 = a + i b, a and b standard normal from default_rng(v0_seed);
  P and T = 2 pi of exp1/make_config_exp1bc.m:17-26.

Usage: python tests/golden/make_eigs_fixtures.py [--workers W] [case ...]
Runs on the CPU only; the N = 5 cases take several minutes each, the N = 6 case longer.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import fame6d_oracle as O  # noqa: E402

Q_PATH_END = (0.5, 0.4, 0.3)

# name -> workload (N, permittivity amplitude, Bloch vector, number of eigenvalues)
CASES = {
    # plain method at N = 3 (46,656 modes)
    "eigs_N3_a04": dict(N=3, amp=0.4, q=(0.05, 0.04, 0.03), k=10),
    # plain FP64 at N = 5 (10^6 modes), the headline workload
    "eigs_N5_a04": dict(N=5, amp=0.4, q=(0.05, 0.04, 0.03), k=10),
    # mixed-precision workloads: amplitude 0.8 at N = 5 and N = 6
    "eigs_N5_a08": dict(N=5, amp=0.8, q=(0.05, 0.04, 0.03), k=10),
    "eigs_N6_a08": dict(N=6, amp=0.8, q=(0.05, 0.04, 0.03), k=10),
    # batched Bloch path at N = 4: first and last of the 12 points t*(0.5,0.4,0.3), t = j/12
    "eigs_N4_a04_t01": dict(N=4, amp=0.4, q=tuple(c / 12.0 for c in Q_PATH_END), k=10),
    "eigs_N4_a04_t12": dict(N=4, amp=0.4, q=Q_PATH_END, k=10),
}
EPS_SEED = 0
V0_SEED = 20260905
LANCZOS = dict(p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=1200)


def phases(seed):
    return np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, 6)


def start_vector(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def run_case(name, spec):
    N, amp, q, k = spec["N"], spec["amp"], spec["q"], spec["k"]
    ph = phases(EPS_SEED)
    eps = O.eps_cosine(N, amp, ph)
    nr = 2 * (2 * N) ** 6
    v0 = start_vector(nr, V0_SEED)
    t0 = time.time()
    lam, _, info, _ = O.eigs_method_c(N, eps, q, v0, k=k, **LANCZOS)
    dt = time.time() - t0
    return {
        "case": name,
        "N": N, "NF": (2 * N) ** 6, "amp": amp, "eps_seed": EPS_SEED, "eps_phases": ph.tolist(),
        "q": list(q), "k": k, "v0_seed": V0_SEED, "lanczos": LANCZOS,
        "lambda": [float(v) for v in lam],
        "oracle": {"seconds": dt, "krinv_calls": info["calls"], "pcg_iterations": info["iters"],
                   "lanczos_steps": info["steps"], "restarts": info["restarts"],
                   "max_ritz_residual": info.get("max_ritz_residual")},
        "generator": "tests/golden/make_eigs_fixtures.py (numpy FP64 oracle, oracle/fame6d_oracle.py)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cases", nargs="*", default=list(CASES))
    ap.add_argument("--workers", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    O.set_fft_workers(a.workers)
    for name in a.cases:
        out = run_case(name, CASES[name])
        path = os.path.join(HERE, name + ".json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")
        print("%s: %.1f s, lambda_1 = %.15e, lambda_k = %.15e" % (name, out["oracle"]["seconds"],
                                                                  out["lambda"][0], out["lambda"][-1]),
              flush=True)


if __name__ == "__main__":
    main()
