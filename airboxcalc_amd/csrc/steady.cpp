// Time to steady behind the C ABI (no reference counterpart: the reference
// always runs a fixed number of steps and averages the last `avg` of them,
// engine.js:341-358, 613-630).
//
// The solve runs in windows of `window` steps whose fields are averaged on the
// GPU.  After each window the probe checks for non-finite flow (the
// reference's instability error, engine.js:335) and evaluate's metrics of the
// window's average are taken (engine.js:632-691).  The solve is steady when
// gamma-core moved by less than tol_gamma and zeta by at most tol_zeta
// (relative) between two consecutive windows; its result is then the mean of
// those two windows' metrics (zeta is linear in the averaged density, so its
// mean is zeta of the two-window average).
#include <math.h>

#include "abx_common.hpp"

using namespace abx;

namespace {

void mean_metrics(const abx_metrics& a, const abx_metrics& b, abx_metrics* m) {
  *m = b;  // section indices and counts of the last window
  m->gamma = 0.5 * (b.gamma + a.gamma);
  m->gammaCore = 0.5 * (b.gammaCore + a.gammaCore);
  m->lowCore = 0.5 * (b.lowCore + a.lowCore);
  m->highCore = 0.5 * (b.highCore + a.highCore);
  m->reverse = 0.5 * (b.reverse + a.reverse);
  m->maxToMean = 0.5 * (b.maxToMean + a.maxToMean);
  m->zeta = 0.5 * (b.zeta + a.zeta);
  m->inlet_cm2 = 0.5 * (b.inlet_cm2 + a.inlet_cm2);
  m->filter_cm2 = 0.5 * (b.filter_cm2 + a.filter_cm2);
  m->rhoRef = 0.5 * (b.rhoRef + a.rhoRef);
  m->pin = 0.5 * (b.pin + a.pin);
}

}  // namespace

extern "C" int abx_sim_solve_to_steady(abx_sim* sim, const abx_eval_geom* g, const abx_steady_params* p,
                                       abx_steady_cb cb, void* user, abx_steady_result* out) {
  if (!sim || !g || !p || !out) {
    set_error("abx_sim_solve_to_steady: null argument");
    return -1;
  }
  if (p->window < 1 || p->max_steps < p->window || p->step0 < 0 || !(p->tol_gamma >= 0) || !(p->tol_zeta >= 0)) {
    set_error("abx_sim_solve_to_steady: bad parameters (window >= 1, max_steps >= window, step0 >= 0, tolerances "
              ">= 0)");
    return -1;
  }
  out->converged = 0;
  out->steps = 0;
  out->windows = 0;
  abx_metrics prev{};
  bool have_prev = false;
  int steps = 0;
  while (steps + p->window <= p->max_steps) {
    if (abx_sim_clear_acc(sim) || abx_sim_steps(sim, p->window, 1)) return -1;
    steps += p->window;
    abx_steady_window w{};
    w.step = p->step0 + steps;
    double pr[3];
    if (abx_sim_probe(sim, w.step, pr)) return -1;  // the reference's instability error
    w.gamma = pr[0];
    w.massRatio = pr[1];
    if (abx_evaluate(sim, g, 0, (double)p->window, &w.metrics, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr))
      return -1;
    out->windows++;
    out->steps = steps;
    const bool steady = have_prev && fabs(w.metrics.gammaCore - prev.gammaCore) < p->tol_gamma &&
                        fabs(w.metrics.zeta - prev.zeta) <= p->tol_zeta * fabs(prev.zeta);
    if (steady) {
      mean_metrics(prev, w.metrics, &out->metrics);
      out->converged = 1;
    } else {
      out->metrics = w.metrics;
    }
    if (cb && cb(user, &w)) break;  // the caller stopped the solve
    if (steady) break;
    prev = w.metrics;
    have_prev = true;
  }
  return 0;
}
