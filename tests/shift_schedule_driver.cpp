// CPU driver of run_shift_schedule (csrc/shift_schedule.hpp) on scripted MINRES outcomes.
// stdin: tol, then three lists, each a count followed by its entries:
//   calls:       flag iter relres   (what each MINRES call on the full system returns)
//   residuals:   relres             (what each explicit residual before a correction measures)
//   corrections: flag iter relres   (what each correction call returns, relres of y + delta)
// stdout: one line per action (C call, S stagnation restart, M maxit continuation,
// R <corr_tol> correction; the tolerance is printed only when a correction solve runs),
// then "END <end> <relres> <iterations> <calls> <stagnation> <maxit> <corrections>
// <correction_iterations> <recovery_events> <plateau_events>".  Exit 3 when a list runs out.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "shift_schedule.hpp"

using fame6d::MrCallResult;

static std::vector<MrCallResult> read_results() {
  int n = 0;
  if (std::scanf("%d", &n) != 1) std::exit(2);
  std::vector<MrCallResult> v(n);
  for (auto& r : v)
    if (std::scanf("%d %d %lf", &r.flag, &r.iter, &r.relres) != 3) std::exit(2);
  return v;
}

int main() {
  double tol = 0.0;
  if (std::scanf("%lf", &tol) != 1) return 2;
  const std::vector<MrCallResult> calls = read_results();
  int nres = 0;
  if (std::scanf("%d", &nres) != 1) return 2;
  std::vector<double> res(nres);
  for (double& r : res)
    if (std::scanf("%lf", &r) != 1) return 2;
  const std::vector<MrCallResult> corrs = read_results();
  size_t ic = 0, ir = 0, ik = 0;
  auto exhausted = []() {
    std::printf("EXHAUSTED\n");
    std::exit(3);
  };
  auto call = [&](bool) {
    if (ic >= calls.size()) exhausted();
    return calls[ic++];
  };
  auto residual = [&]() {
    if (ir >= res.size()) exhausted();
    return res[ir++];
  };
  auto correct = [&](double corr_tol) {
    if (ik >= corrs.size()) exhausted();
    std::printf("T %.17g\n", corr_tol);
    return corrs[ik++];
  };
  auto log = [](int a) {
    static const char* names = "CSMR";
    std::printf("%c\n", names[a]);
  };
  const fame6d::ShiftOutcome o = fame6d::run_shift_schedule(tol, fame6d::ShiftRecovery{}, call, residual, correct, log);
  std::printf("END %d %.17g %d %d %d %d %d %d %d %d\n", o.end, o.relres, o.iterations, o.calls, o.stagnation_restarts,
              o.maxit_restarts, o.corrections, o.correction_iterations, o.recovery_events, o.plateau_events);
  return 0;
}
