// Recovery schedule of one Method-A shifted solve (K - sigma M) y = rhs: which MINRES call
// runs next, when a stalled solve switches to residual correction, and when a solve that
// stops short of its tolerance is accepted or given up.  Host-only and independent of the
// device: the caller supplies the MINRES call, the explicit residual and the correction
// solve as callables, so the same schedule drives the GPU solver and a CPU trace driver.
//
// Reference: exp1/make_A_eigs_top_preconditioned.m:203-560 (the solve loop with stagnation
// restarts, MAXIT continuations, residual-correction cycles and the plateau / guard
// acceptance), :583-660 (residual_correction: corrTol = safety * tol / relres clamped,
// MINRES on A delta = rhs - A y from zero, y += delta); settings cfg.shift of
// exp1/make_config_exp1bc.m:50-72.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>

namespace fame6d {

// cfg.shift of exp1/make_config_exp1bc.m:50-72.  The two acceptance ceilings are raised to
// tol when tol exceeds them (the reference rejects such a configuration outright).
struct ShiftRecovery {
  int max_recovery_events = 50;      // maxRecoveryEventsPerTop
  int max_corrections = 8;           // maxResidualCorrectionCycles
  int max_zero_progress = 2;         // maxZeroProgressEvents
  int plateau_min_events = 2;        // plateauMinFailedRecoveryCycles
  double min_progress = 1e-8;        // minResidualProgressFraction
  double corr_safety = 0.25;         // correctionSafetyFactor
  double corr_max_tol = 1e-1;        // correctionMaxRelativeTolerance
  double corr_min_tol = 1e-6;        // correctionMinRelativeTolerance
  double plateau_ceiling = 1e-9;     // plateauAcceptanceCeiling
  double plateau_reduction = 1e-2;   // plateauMinRelativeReduction
  double guard_ceiling = 1e-9;       // guardNearToleranceCeiling
};

// What one MATLAB-minres call returns: its flag (0 converged, 1 maxit, 3 the true residual
// missed after the recurrence converged, 4 breakdown), its iterations and the explicit
// relative residual of the returned iterate against that call's right-hand side.
struct MrCallResult {
  int flag = 0;
  int iter = 0;
  double relres = 0.0;
};

// Why a schedule ended.
enum ShiftEnd : int {
  kShiftConverged = 0,      // true relres <= tol
  kShiftPlateau = 1,        // accepted at the plateau or guard ceiling
  kShiftFailed = 2,         // a guard fired above the ceiling, or a genuine breakdown
};

struct ShiftOutcome {
  int end = kShiftConverged;
  double relres = INFINITY;          // explicit relres of the final iterate
  int iterations = 0;                // MINRES iterations, main and correction calls
  int calls = 0;                     // MINRES calls (segmentCount)
  int stagnation_restarts = 0;
  int maxit_restarts = 0;
  int corrections = 0;               // residual-correction cycles
  int correction_iterations = 0;
  int recovery_events = 0;
  int plateau_events = 0;
  char message[200] = {0};
};

// One action of the schedule, recorded for trace comparison.
enum ShiftAction : int {
  kActCall = 0,          // MINRES on A y = rhs (from zero at the start, else from y)
  kActStagnation = 1,    // restart after flag 3 with progress
  kActMaxit = 2,         // continuation after flag 1
  kActCorrect = 3,       // a residual-correction cycle
};

// call(bool from_current) -> MrCallResult: MINRES on A y = rhs from zero (first call) or
//   from the current y, leaving the result in y.
// residual() -> double: ||rhs - A y|| / ||rhs||, leaving rhs - A y for the correction.
// correct(double corr_tol) -> MrCallResult: MINRES on A delta = r from zero to corr_tol,
//   y += delta; its relres is that of the updated y against rhs.
// log(int action) is called once per action taken (trace comparison; may do nothing).
template <class Call, class Residual, class Correct, class Log>
ShiftOutcome run_shift_schedule(double tol, ShiftRecovery c, Call&& call, Residual&& residual, Correct&& correct,
                                Log&& log) {
  if (c.guard_ceiling < tol) c.guard_ceiling = tol;
  if (c.plateau_ceiling < c.guard_ceiling) c.plateau_ceiling = c.guard_ceiling;
  ShiftOutcome o;
  double prev = INFINITY;  // explicit relres at the previous recovery decision
  int zero_progress = 0;
  bool from_current = false;

  auto finish = [&](int end, double rel, const char* fmt, ...) {
    o.end = end;
    o.relres = rel;
    if (fmt) {
      va_list ap;
      va_start(ap, fmt);
      std::vsnprintf(o.message, sizeof o.message, fmt, ap);
      va_end(ap);
    }
    return o;
  };
  // counts one more event on the same residual level when rel fell by less than 1 %
  auto note_plateau = [&](double rel, double ref) {
    o.plateau_events = (rel >= ref * (1.0 - c.plateau_reduction)) ? o.plateau_events + 1 : 0;
  };
  auto plateau_met = [&](double rel) {
    return o.plateau_events >= c.plateau_min_events && std::isfinite(rel) && rel <= c.plateau_ceiling;
  };
  // acceptance at a guard: a plateau below its ceiling, or a residual within the guard ceiling
  auto guard_accepts = [&](double rel) { return plateau_met(rel) || (std::isfinite(rel) && rel <= c.guard_ceiling); };

  double rel = INFINITY;
  for (;;) {
    if (o.recovery_events > c.max_recovery_events) {
      if (guard_accepts(rel)) return finish(kShiftPlateau, rel, nullptr);
      return finish(kShiftFailed, rel, "recovery guard after %d events: true relres %.3e", o.recovery_events, rel);
    }
    log(kActCall);
    const MrCallResult m = call(from_current);
    o.calls += 1;
    o.iterations += m.iter;
    rel = m.relres;
    if (!std::isfinite(rel)) return finish(kShiftFailed, rel, "explicit residual is not finite");
    if (m.flag == 0 || rel <= tol) return finish(kShiftConverged, rel, nullptr);
    const bool first = !std::isfinite(prev);
    const bool progress = !first && rel < prev * (1.0 - c.min_progress);

    if (m.flag == 3) {
      if (m.iter > 0 && (first || progress)) {
        if (!first) {
          note_plateau(rel, prev);
          if (plateau_met(rel)) return finish(kShiftPlateau, rel, nullptr);
        }
        log(kActStagnation);
        o.stagnation_restarts += 1;
        o.recovery_events += 1;
        zero_progress = 0;
        prev = rel;
        from_current = true;
        continue;
      }
      // no iterations, or no material progress: one residual-correction cycle
      if (o.corrections >= c.max_corrections) {
        if (guard_accepts(rel)) return finish(kShiftPlateau, rel, nullptr);
        return finish(kShiftFailed, rel, "correction guard after %d cycles: true relres %.3e", o.corrections, rel);
      }
      log(kActCorrect);
      const double before = residual();
      MrCallResult cr;
      cr.relres = before;
      if (!std::isfinite(before)) return finish(kShiftFailed, before, "correction residual is not finite");
      if (before > tol) {
        double ct = c.corr_safety * tol / std::fmax(before, DBL_MIN);
        ct = std::fmin(c.corr_max_tol, std::fmax(c.corr_min_tol, ct));
        cr = correct(ct);
      }
      o.corrections += 1;
      o.correction_iterations += cr.iter;
      o.iterations += cr.iter;
      o.calls += 1;
      o.recovery_events += 1;
      const double after = cr.relres;
      if (!std::isfinite(after)) return finish(kShiftFailed, after, "corrected residual is not finite");
      if (after <= tol) return finish(kShiftConverged, after, nullptr);
      note_plateau(after, rel);
      if (plateau_met(after)) return finish(kShiftPlateau, after, nullptr);
      if (after < rel * (1.0 - c.min_progress)) {
        zero_progress = 0;
      } else {
        zero_progress += 1;
        if (zero_progress >= c.max_zero_progress) {
          if (guard_accepts(after)) return finish(kShiftPlateau, after, nullptr);
          return finish(kShiftFailed, after, "residual correction made no progress %d times: true relres %.3e",
                        zero_progress, after);
        }
      }
      rel = after;
      prev = after;
      from_current = true;
      continue;
    }

    if (m.flag == 1) {
      if (m.iter <= 0) {
        if (std::isfinite(rel) && rel <= c.guard_ceiling) return finish(kShiftPlateau, rel, nullptr);
        return finish(kShiftFailed, rel, "MINRES reached maxit with no iterations: true relres %.3e", rel);
      }
      if (!first) {
        note_plateau(rel, prev);
        if (plateau_met(rel)) return finish(kShiftPlateau, rel, nullptr);
      }
      log(kActMaxit);
      o.maxit_restarts += 1;
      o.recovery_events += 1;
      prev = rel;
      from_current = true;
      continue;
    }

    if (rel <= c.guard_ceiling) return finish(kShiftPlateau, rel, nullptr);
    return finish(kShiftFailed, rel, "MINRES breakdown (flag %d): true relres %.3e", m.flag, rel);
  }
}

}  // namespace fame6d
