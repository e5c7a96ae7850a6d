// powell_minimize.hpp — Powell's direction-set method (Powell 1964) with Brent line
// minimisation (Brent 1973) on the host, for the distortion calibration's POWELL and
// POWELL_ITER procedures (src/distortion_correction.cpp:241-400). The reference calls the
// toofus Powell(func, 1e-12).minimize(X, perturb), which is not part of this tree; this is
// the published method: a set of directions, one line minimisation along each per sweep,
// and the direction of largest decrease replaced by the sweep's net displacement when the
// extrapolation test allows it; it stops when a sweep lowers f by no more than
// ftol (|f_old| + |f|) / 2. The objective here is the equiprojectivity criterion evaluated on
// the GPU (ptg_equiproj_eval), so every f is one kernel launch.
#pragma once

#include <cmath>
#include <functional>
#include <utility>
#include <vector>

namespace ptgpu {

using Objective = std::function<double(const std::vector<double>&)>;

namespace powell_detail {

// f along x + t d.
struct Line {
  const Objective& f;
  const std::vector<double>& x;
  const std::vector<double>& d;
  std::vector<double> p;
  int* evals;
  double operator()(double t) {
    p.resize(x.size());
    for (size_t i = 0; i < x.size(); i++) p[i] = x[i] + t * d[i];
    ++*evals;
    return f(p);
  }
};

// A bracket of a minimum along the line: three steps a, b, c in order with f(b) below f(a)
// and f(c), found by stepping from (0, 1) downhill with steps growing by the golden ratio.
inline void bracket(Line& g, double& a, double& b, double& c, double& fb) {
  const double grow = 1.618033988749895;
  a = 0.0;
  b = 1.0;
  double fa = g(a);
  fb = g(b);
  if (fb > fa) {  // walk the other way
    std::swap(a, b);
    std::swap(fa, fb);
  }
  c = b + grow * (b - a);
  double fc = g(c);
  for (int it = 0; it < 400 && fc < fb; it++) {
    a = b;
    fa = fb;
    b = c;
    fb = fc;
    c = b + grow * (b - a);
    fc = g(c);
  }
}

// Brent's method on a bracket: the best, second-best and previous second-best points
// (best, second, third) define a parabola; its vertex is taken when it lies inside the
// interval and the step is below half the one before last, a golden-section step into the
// larger side otherwise. Stops when the interval around the best point is within
// tol |best| + 1e-20 on each side (the floor in units of the direction, which the caller may
// scale far from the step needed: the calibration's perturbations of 1e-6 for coefficients of
// 1e-8 and 1e-14).
inline double brent(Line& g, double a, double b, double c, double fb, double tol, double& tmin) {
  const double golden = 0.3819660112501051;
  double lo = std::min(a, c), hi = std::max(a, c);
  double best = b, second = b, third = b, f_best = fb, f_second = fb, f_third = fb;
  double step = 0.0, step_before = 0.0;
  for (int it = 0; it < 300; it++) {
    const double mid = 0.5 * (lo + hi);
    const double eps = tol * std::fabs(best) + 1e-20;
    if (std::fabs(best - mid) <= 2.0 * eps - 0.5 * (hi - lo)) break;
    bool parabolic = false;
    if (std::fabs(step_before) > eps) {
      const double r = (best - second) * (f_best - f_third), q0 = (best - third) * (f_best - f_second);
      double num = (best - third) * q0 - (best - second) * r, den = 2.0 * (q0 - r);
      if (den > 0.0) num = -num;
      den = std::fabs(den);
      const double limit = step_before;
      step_before = step;
      if (std::fabs(num) < std::fabs(0.5 * den * limit) && num > den * (lo - best) && num < den * (hi - best)) {
        step = num / den;
        const double u = best + step;
        if (u - lo < 2.0 * eps || hi - u < 2.0 * eps) step = std::copysign(eps, mid - best);
        parabolic = true;
      }
    }
    if (!parabolic) {
      step_before = best >= mid ? lo - best : hi - best;
      step = golden * step_before;
    }
    const double u = std::fabs(step) >= eps ? best + step : best + std::copysign(eps, step);
    const double fu = g(u);
    if (fu <= f_best) {
      (u >= best ? lo : hi) = best;
      third = second;
      f_third = f_second;
      second = best;
      f_second = f_best;
      best = u;
      f_best = fu;
    } else {
      (u < best ? lo : hi) = u;
      if (fu <= f_second || second == best) {
        third = second;
        f_third = f_second;
        second = u;
        f_second = fu;
      } else if (fu <= f_third || third == best || third == second) {
        third = u;
        f_third = fu;
      }
    }
  }
  tmin = best;
  return f_best;
}

// Minimises f along d from x; moves x to the minimum, scales d by the step taken and
// returns f there.
inline double line_minimize(const Objective& f, std::vector<double>& x, std::vector<double>& d, double tol,
                            int* evals) {
  Line g{f, x, d, {}, evals};
  double a, b, c, fb;
  bracket(g, a, b, c, fb);
  double t = 0.0;
  const double fmin = brent(g, a, b, c, fb, tol, t);
  for (size_t i = 0; i < x.size(); i++) {
    d[i] *= t;
    x[i] += d[i];
  }
  return fmin;
}

}  // namespace powell_detail

// Minimises f from x along the initial directions dirs (each a vector of x's size);
// returns f at the minimum and updates x. ftol: the stopping test above; max_iter sweeps.
inline double powell_minimize(const Objective& f, std::vector<double>& x, std::vector<std::vector<double>> dirs,
                              double ftol, int max_iter = 200, int* sweeps = nullptr, int* evals = nullptr) {
  int ne = 0, it = 0;
  const size_t n = x.size();
  double f_now = f(x);
  ne++;
  std::vector<double> start = x, extrap(n), net(n);
  const double line_tol = 1e-8;
  for (; it < max_iter && !dirs.empty(); it++) {
    const double f_start = f_now;
    size_t i_big = 0;  // the direction of the sweep's largest decrease, and that decrease
    double big_drop = 0.0;
    for (size_t i = 0; i < dirs.size(); i++) {
      const double f_before = f_now;
      f_now = powell_detail::line_minimize(f, x, dirs[i], line_tol, &ne);
      if (f_before - f_now > big_drop) {
        big_drop = f_before - f_now;
        i_big = i;
      }
    }
    if (2.0 * (f_start - f_now) <= ftol * (std::fabs(f_start) + std::fabs(f_now)) + 1e-300) break;
    // the sweep's net displacement, and f one more such step ahead
    for (size_t j = 0; j < n; j++) {
      net[j] = x[j] - start[j];
      extrap[j] = x[j] + net[j];
      start[j] = x[j];
    }
    const double f_extrap = f(extrap);
    ne++;
    if (f_extrap < f_start) {
      // Powell's test: take the net displacement as a new direction (in place of the one of
      // largest decrease) unless the decrease along it is mostly that one direction's
      const double gap = f_start - f_now - big_drop, rise = f_start - f_extrap;
      if (2.0 * (f_start - 2.0 * f_now + f_extrap) * gap * gap - big_drop * rise * rise < 0.0) {
        f_now = powell_detail::line_minimize(f, x, net, line_tol, &ne);
        dirs[i_big] = dirs.back();
        dirs.back() = net;
      }
    }
  }
  if (sweeps) *sweeps = it;
  if (evals) *evals = ne;
  return f_now;
}

// Powell's method restarted from its own result with the initial directions, short runs of
// n + 1 sweeps, until a run lowers f by no more than ftol (|f_old| + |f|) / 2. A single run
// of the method whose direction of largest decrease is replaced by the net displacement can
// keep directions far from conjugate in a narrow valley and creep along it; restarting
// resets them. Returns f at the result; runs receives the runs made.
inline double powell_minimize_restarted(const Objective& f, std::vector<double>& x,
                                        const std::vector<std::vector<double>>& dirs, double ftol,
                                        int max_runs = 200, int* runs = nullptr) {
  double f_now = f(x);
  int r = 0;
  for (; r < max_runs && !dirs.empty(); r++) {
    const double f_before = f_now;
    f_now = powell_minimize(f, x, dirs, ftol, (int)dirs.size() + 1);
    if (2.0 * (f_before - f_now) <= ftol * (std::fabs(f_before) + std::fabs(f_now)) + 1e-300) {
      r++;
      break;
    }
  }
  if (runs) *runs = r;
  return f_now;
}

}  // namespace ptgpu
