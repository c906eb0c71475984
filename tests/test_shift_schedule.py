"""The Method-A shifted-solve recovery schedule (exp1/make_A_eigs_top_preconditioned.m:251-560,
residual correction :583-660, settings exp1/make_config_exp1bc.m:50-72) on scripted MINRES
outcomes: the C++ schedule that drives the GPU solver (csrc/shift_schedule.hpp, through
tests/shift_schedule_driver.cpp) and the oracle's restatement, each against action traces
derived by hand from the reference's branches, and against each other on random scripts."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "null--space-free-6d-projection-method_amd", "csrc")
TOL = 1e-10
ENDS = {0: "converged", 1: "plateau", 2: "failed"}


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("sched") / "shift_schedule")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", CSRC,
                    os.path.join(ROOT, "tests", "shift_schedule_driver.cpp"), "-o", exe], check=True)
    return exe


def run_cpp(exe, calls, residuals, corrections, tol=TOL):
    lines = ["%.17g" % tol, str(len(calls))] + ["%d %d %.17g" % c for c in calls]
    lines += [str(len(residuals))] + ["%.17g" % r for r in residuals]
    lines += [str(len(corrections))] + ["%d %d %.17g" % c for c in corrections]
    p = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    actions, ctols = [], []
    for ln in p.stdout.split("\n"):
        f = ln.split()
        if not f:
            continue
        if f[0] == "T":
            ctols.append(float(f[1]))
        elif f[0] == "END":
            out = dict(end=ENDS[int(f[1])], relres=float(f[2]))
            keys = ("iterations", "calls", "stagnation_restarts", "maxit_restarts", "corrections",
                    "correction_iterations", "recovery_events", "plateau_events")
            out.update({k: int(v) for k, v in zip(keys, f[3:])})
        else:
            actions.append({"C": "call", "S": "stagnation", "M": "maxit", "R": "correct"}[f[0]])
    return actions, ctols, out


def run_oracle(oracle, calls, residuals, corrections, tol=TOL):
    it_c, it_r, it_k = iter(calls), iter(residuals), iter(corrections)
    actions, ctols = [], []

    def correct(t):
        ctols.append(t)
        return next(it_k)

    out = oracle.shift_recovery_schedule(tol, lambda cur: next(it_c), lambda: next(it_r), correct,
                                         log=actions.append)
    return actions, ctols, out


def both(driver, oracle, calls, residuals=(), corrections=()):
    a = run_cpp(driver, calls, residuals, corrections)
    b = run_oracle(oracle, calls, residuals, corrections)
    assert a[0] == b[0] and np.allclose(a[1], b[1], rtol=1e-15, atol=0)
    for k, v in a[2].items():
        assert (v == b[2][k]) if not isinstance(v, float) else math.isclose(v, b[2][k], rel_tol=1e-15), k
    return a


C, S, M, R = "call", "stagnation", "maxit", "correct"


def test_first_call_converges(driver, oracle):
    acts, ct, o = both(driver, oracle, [(0, 120, 5e-11)])
    assert acts == [C] and ct == [] and o["end"] == "converged" and o["iterations"] == 120 and o["calls"] == 1


def test_stagnation_restart_then_converged(driver, oracle):
    # :360-384: flag 3 with iterations on the first termination restarts from the iterate
    acts, ct, o = both(driver, oracle, [(3, 200, 3e-10), (0, 15, 8e-11)])
    assert acts == [C, S, C] and o["end"] == "converged"
    assert (o["stagnation_restarts"], o["recovery_events"], o["iterations"], o["calls"]) == (1, 1, 215, 2)


def test_zero_iteration_stagnation_goes_to_correction(driver, oracle):
    # :389-400 and :583-620: corrTol = 0.25 * tol / relres inside [1e-6, 1e-1]
    acts, ct, o = both(driver, oracle, [(3, 0, 4e-10)], [4e-10], [(0, 30, 6e-11)])
    assert acts == [C, R] and ct == [pytest.approx(0.0625, rel=1e-15)]
    assert o["end"] == "converged" and (o["corrections"], o["calls"], o["iterations"]) == (1, 2, 30)


@pytest.mark.parametrize("before,expect", [(2e-10, 0.1), (1e-3, 1e-6), (1e-8, 2.5e-3)])
def test_correction_tolerance_clamp(driver, oracle, before, expect):
    acts, ct, o = both(driver, oracle, [(3, 0, before)], [before], [(0, 7, 5e-11)])
    assert ct == [pytest.approx(expect, rel=1e-14)] and o["end"] == "converged"


def test_correction_skipped_when_measured_residual_meets_tol(driver, oracle):
    # :601-611: a measured residual within tol returns the iterate with no correction solve
    acts, ct, o = both(driver, oracle, [(3, 0, 2e-10)], [9e-11], [])
    assert acts == [C, R] and ct == [] and o["end"] == "converged" and o["corrections"] == 1
    assert o["relres"] == 9e-11 and o["iterations"] == 0


def test_plateau_accepted_after_stagnation_restarts(driver, oracle):
    # :363-375: two restarts that each reduce the residual by less than 1% below 1e-9
    acts, ct, o = both(driver, oracle, [(3, 100, 5e-10), (3, 5, 4.999e-10), (3, 5, 4.998e-10)])
    assert acts == [C, S, C, S, C] and o["end"] == "plateau" and o["relres"] == 4.998e-10
    assert o["plateau_events"] == 2 and o["iterations"] == 110


@pytest.mark.parametrize("floor,end", [(1.06e-9, "failed"), (9e-10, "plateau")])
def test_residual_floor(driver, oracle, floor, end):
    # a residual that no call or correction lowers: correction cycles until the plateau
    # (:436-446) or the zero-progress guard (:480-496) decides at the 1e-9 ceiling
    calls = [(3, 5, floor)] * 3
    acts, ct, o = both(driver, oracle, calls, [floor] * 2, [(0, 20, floor)] * 2)
    assert acts == [C, S, C, R, C, R] and o["end"] == end and o["relres"] == floor
    assert (o["calls"], o["iterations"], o["corrections"], o["recovery_events"]) == (5, 55, 2, 3)


def test_correction_progress_returns_to_main_minres(driver, oracle):
    # :460-470: a correction that lowers the residual resumes MINRES from the corrected iterate
    acts, ct, o = both(driver, oracle, [(3, 0, 4e-10), (0, 40, 9e-11)], [4e-10], [(3, 10, 2e-10)])
    assert acts == [C, R, C] and o["end"] == "converged" and o["plateau_events"] == 0


def test_maxit_continuation(driver, oracle):
    # :521-545
    acts, ct, o = both(driver, oracle, [(1, 4000, 2e-8), (0, 300, 9e-11)])
    assert acts == [C, M, C] and o["maxit_restarts"] == 1 and o["end"] == "converged"


@pytest.mark.parametrize("flag,it,rel,end", [(1, 0, 5e-10, "plateau"), (1, 0, 2e-9, "failed"),
                                             (4, 12, 8e-10, "plateau"), (4, 12, 3e-9, "failed")])
def test_zero_iteration_maxit_and_breakdown(driver, oracle, flag, it, rel, end):
    # :503-515 (flag 1 with no iterations) and :547-560 (other flags) accept only within 1e-9
    acts, ct, o = both(driver, oracle, [(flag, it, rel)])
    assert acts == [C] and o["end"] == end


@pytest.mark.parametrize("start,end", [(5e-9, "failed"), (2.5e-9, "plateau")])
def test_recovery_event_guard(driver, oracle, start, end):
    # :251-268: restarts that each gain 2% (progress, no plateau) until 50 recovery events
    calls = [(3, 5, start * 0.98 ** k) for k in range(60)]
    acts, ct, o = both(driver, oracle, calls)
    assert o["calls"] == 51 and o["stagnation_restarts"] == 51 and o["end"] == end
    assert o["relres"] == start * 0.98 ** 50


def test_random_scripts_agree(driver, oracle):
    rng = np.random.default_rng(2026)
    for trial in range(150):
        n = 80
        base = 10.0 ** rng.uniform(-10.5, -8.0)
        calls = [(int(rng.choice([0, 1, 3, 3, 3, 4])) if rng.random() < 0.9 else 3, int(rng.integers(0, 3) * 7),
                  base * 10 ** rng.uniform(-0.3, 0.3)) for _ in range(n)]
        res = [base * 10 ** rng.uniform(-0.3, 0.3) for _ in range(n)]
        corr = [(0, int(rng.integers(0, 40)), base * 10 ** rng.uniform(-0.6, 0.1)) for _ in range(n)]
        both(driver, oracle, calls, res, corr)
