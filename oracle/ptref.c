/*
 * ptref.c — CPU restatement ("oracle") of p-tracker's per-grain tracking path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libptgpu, used by tests/, the smoke check
 * and the CPU-baseline leg of bench.py. Nothing in the product path links it.
 *
 * Built with -O2 -ffp-contract=off (no FMA, like the reference on x86-64 without
 * -march), so every double expression below rounds exactly as the reference's does.
 * Each function cites the reference lines (paths relative to the reference root) whose
 * loop order and arithmetic it restates.
 */
#include "ptref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PX(f, pitch, x, y) ((f)[(int64_t)(y) * (pitch) + (x)])

/* src/p-tracker.cpp:684-686 (also :771/:824 and :908/:961): the steps are produced by
 * repeated addition in double, so their count and values differ from -rotmax + k*inc. */
int ptref_rotation_steps(double inc_rot, int32_t num_rot, double* out, int32_t cap) {
  if (!(inc_rot > 0.0)) return -1;
  double rotmax = fabs(inc_rot * num_rot);
  int n = 0;
  for (double drot = -rotmax; drot <= rotmax; drot += inc_rot) {
    if (out && n < cap) out[n] = drot;
    n++;
  }
  return n;
}

/* make_circ_pattern, src/p-tracker.cpp:1386-1408. */
int ptref_make_circ_pattern(int32_t radius_pix, int32_t* out_xy) {
  int n = 0;
  for (int dy = -radius_pix; dy <= radius_pix; ++dy)
    for (int dx = -radius_pix; dx <= radius_pix; ++dx) {
      double dst = sqrt((double)dx * (double)dx + (double)dy * (double)dy);
      if (dst <= (double)radius_pix) {
        if (out_xy) {
          out_xy[2 * n] = dx;
          out_xy[2 * n + 1] = dy;
        }
        n++;
      }
    }
  return n;
}

/* make_rect_pattern, src/p-tracker.cpp:1369-1384. */
int ptref_make_rect_pattern(int32_t half_width, int32_t half_height, int32_t* out_xy) {
  int n = 0;
  for (int dy = -half_height; dy <= half_height; ++dy)
    for (int dx = -half_width; dx <= half_width; ++dx) {
      if (out_xy) {
        out_xy[2 * n] = dx;
        out_xy[2 * n + 1] = dy;
      }
      n++;
    }
  return n;
}

/* make_ring_pattern, src/p-tracker.cpp:1410-1433. */
int ptref_make_ring_pattern(int32_t radius_in, int32_t radius_out, int32_t* out_xy) {
  int n = 0;
  for (int dy = -radius_out; dy <= radius_out; ++dy)
    for (int dx = -radius_out; dx <= radius_out; ++dx) {
      double dst = sqrt((double)dx * (double)dx + (double)dy * (double)dy);
      if (dst <= (double)radius_out && dst >= (double)radius_in) {
        if (out_xy) {
          out_xy[2 * n] = dx;
          out_xy[2 * n + 1] = dy;
        }
        n++;
      }
    }
  return n;
}

/* rotate_pixel_pattern, src/p-tracker.cpp:615-621: y points down, so the rotation
 * matrix is [[c, s], [-s, c]]; std::round rounds half away from zero. */
static inline void rot_px(int dx, int dy, double c, double s, int32_t* xo, int32_t* yo) {
  double xpix = c * dx + s * dy;
  double ypix = -s * dx + c * dy;
  *xo = (int32_t)round(xpix);
  *yo = (int32_t)round(ypix);
}

void ptref_rotate_pattern(int32_t n, const int32_t* xy, double c, double s, int32_t* out_xy) {
  for (int i = 0; i < n; i++) rot_px(xy[2 * i], xy[2 * i + 1], c, s, &out_xy[2 * i], &out_xy[2 * i + 1]);
}

/* do_precomputations, src/p-tracker.cpp:624-656 (one grain). */
void ptref_precompute(const uint16_t* I0, int64_t pitch, ptg_grain* g, int32_t n,
                      const int32_t* pattern_xy, int32_t* p0_xy) {
  int rx = g->refcoord_xpix, ry = g->refcoord_ypix;
  double c0 = cos(g->refrot), s0 = sin(g->refrot);
  double mean0 = 0.0, C0C0 = 0.0;
  for (int i = 0; i < n; i++) {
    rot_px(pattern_xy[2 * i], pattern_xy[2 * i + 1], c0, s0, &p0_xy[2 * i], &p0_xy[2 * i + 1]);
    mean0 += (double)PX(I0, pitch, p0_xy[2 * i] + rx, p0_xy[2 * i + 1] + ry);
  }
  mean0 /= (double)n;
  for (int i = 0; i < n; i++) {
    double diffC0 = (double)PX(I0, pitch, p0_xy[2 * i] + rx, p0_xy[2 * i + 1] + ry) - mean0;
    C0C0 += diffC0 * diffC0;
  }
  g->mean0 = mean0;
  g->C0C0 = C0C0;
}

/* The two-pass ZNCC of one candidate, src/p-tracker.cpp:698-716 (main) and
 * :834-850 / :971-987 (rescues, same arithmetic after the implicit conversions). */
static double candidate_ncc(const uint16_t* I0, const uint16_t* I1, int64_t pitch,
                            const ptg_grain* g, int32_t n, const int32_t* p0_xy,
                            const int32_t* p1_xy, int bx, int by) {
  int rx = g->refcoord_xpix, ry = g->refcoord_ypix;
  double mean1 = 0.0, C0C1 = 0.0, C1C1 = 0.0;
  for (int i = 0; i < n; i++) mean1 += (double)PX(I1, pitch, p1_xy[2 * i] + bx, p1_xy[2 * i + 1] + by);
  mean1 /= (double)n;
  for (int i = 0; i < n; i++) {
    double diffC1 = (double)PX(I1, pitch, p1_xy[2 * i] + bx, p1_xy[2 * i + 1] + by) - mean1;
    C0C1 += ((double)PX(I0, pitch, p0_xy[2 * i] + rx, p0_xy[2 * i + 1] + ry) - g->mean0) * diffC1;
    C1C1 += diffC1 * diffC1;
  }
  return C0C1 / sqrt(g->C0C0 * C1C1);
}

/* follow_pattern_pixel (:659-743), follow_pattern_rescue_pixel (:745-881) and
 * follow_pattern_super_rescue_pixel (:883-1016) of src/p-tracker.cpp. With
 * num_neighbour always 0 (:1136) the rescues' allowed-shift lists hold every shift
 * of the zone in the order ix outer, iy inner (:813-821), the same order as the main
 * search's u outer, v inner (:695-696). */
int ptref_follow_pattern(const uint16_t* I0, const uint16_t* I1, int64_t pitch, ptg_grain* g,
                         int32_t n, const int32_t* pattern_xy, const int32_t* p0_xy,
                         const ptg_search_zone* z, int stage, double NCC_min,
                         double NCC_min_super, int32_t* p1_xy) {
  int grain_dx = (int)round(g->dx);
  int grain_dy = (int)round(g->dy);
  double rest_dx = g->dx - grain_dx;
  double rest_dy = g->dy - grain_dy;
  if (stage == PTG_STAGE_MAIN) g->upix = g->vpix = g->rot_inc = 0.0; /* :670 */

  double best_NCC = 0.0, best_drot = 0.0;
  int best_upix = 0, best_vpix = 0;
  double rotmax = fabs(z->inc_rot * z->num_rot);
  if (!(z->inc_rot > 0.0)) return 0;
  int bx0 = g->refcoord_xpix + grain_dx, by0 = g->refcoord_ypix + grain_dy;

  for (double drot = -rotmax; drot <= rotmax; drot += z->inc_rot) {
    double total_rot = g->refrot + g->drot + drot;
    double c = cos(total_rot), s = sin(total_rot);
    ptref_rotate_pattern(n, pattern_xy, c, s, p1_xy);
    for (int upix = -z->left; upix <= z->right; upix++)
      for (int vpix = -z->up; vpix <= z->down; vpix++) {
        double NCC_test = candidate_ncc(I0, I1, pitch, g, n, p0_xy, p1_xy, bx0 + upix, by0 + vpix);
        if (best_NCC < NCC_test) { /* strict: the first maximum wins, NaN never does */
          best_NCC = NCC_test;
          best_upix = upix;
          best_vpix = vpix;
          best_drot = drot;
        }
      }
  }

  int queued = 0, update = 0;
  if (stage == PTG_STAGE_MAIN) { /* :729-742 */
    g->NCC = best_NCC;
    g->NCC_rescue = best_NCC;
    if (best_NCC < NCC_min) queued = 1; else update = 1;
  } else if (stage == PTG_STAGE_RESCUE) { /* :868-880 */
    g->NCC_rescue = best_NCC;
    if (best_NCC < NCC_min_super) queued = 1; else update = 1;
  } else { /* :1005-1015 */
    g->NCC_rescue = best_NCC;
    if (best_NCC > NCC_min_super) update = 1;
  }
  if (update) {
    g->upix = best_upix - rest_dx;
    g->vpix = best_vpix - rest_dy;
    g->rot_inc = best_drot;
    g->dx = grain_dx + best_upix;
    g->dy = grain_dy + best_vpix;
    g->drot += best_drot;
  }
  return queued;
}

int ptref_candidate_ncc(const uint16_t* I0, const uint16_t* I1, int64_t pitch,
                        const ptg_grain* g, int32_t n, const int32_t* pattern_xy,
                        const int32_t* p0_xy, const ptg_search_zone* z, double* ncc_out,
                        int64_t cap) {
  int grain_dx = (int)round(g->dx), grain_dy = (int)round(g->dy);
  double rotmax = fabs(z->inc_rot * z->num_rot);
  if (!(z->inc_rot > 0.0)) return -1;
  int32_t* p1 = (int32_t*)malloc(sizeof(int32_t) * 2 * (n > 0 ? n : 1));
  int64_t k = 0;
  int bx0 = g->refcoord_xpix + grain_dx, by0 = g->refcoord_ypix + grain_dy;
  for (double drot = -rotmax; drot <= rotmax; drot += z->inc_rot) {
    double total_rot = g->refrot + g->drot + drot;
    ptref_rotate_pattern(n, pattern_xy, cos(total_rot), sin(total_rot), p1);
    for (int upix = -z->left; upix <= z->right; upix++)
      for (int vpix = -z->up; vpix <= z->down; vpix++) {
        double v = candidate_ncc(I0, I1, pitch, g, n, p0_xy, p1, bx0 + upix, by0 + vpix);
        if (k < cap) ncc_out[k] = v;
        k++;
      }
  }
  free(p1);
  return (int)k;
}

/* image_interpCubic::CubicInterpolate, src/interpol.cpp:42-53 (Catmull-Rom in
 * Hermite form). */
static inline double cubic1(const double p[4], double x) {
  double f0 = p[1];
  double f1 = p[2];
  double fp0 = 0.5 * (p[2] - p[0]);
  double fp1 = 0.5 * (p[3] - p[1]);
  double d = f0;
  double c = fp0;
  double b = -3.0 * f0 + 3.0 * f1 - 2.0 * fp0 - fp1;
  double a = 2.0 * f0 - 2.0 * f1 + fp0 + fp1;
  double x2 = x * x;
  return (a * x2 * x + b * x2 + c * x + d);
}

/* image_interpCubic::getValue, src/interpol.cpp:55-86: interpolate along y for each
 * of the 4 columns ix-1..ix+2, then along x. */
double ptref_interp_cubic(const uint16_t* im, int64_t pitch, double x, double y) {
  int ix = (int)floor(x);
  int iy = (int)floor(y);
  double y01 = y - (double)iy;
  double x01 = x - (double)ix;
  double p[4][4], arr[4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) p[i][j] = PX(im, pitch, ix - 1 + i, iy - 1 + j);
  for (int i = 0; i < 4; i++) arr[i] = cubic1(p[i], y01);
  return cubic1(arr, x01);
}

void ptref_interp_cubic_batch(const uint16_t* im, int64_t pitch, int32_t n, const double* xy,
                              double* out) {
  for (int k = 0; k < n; k++) out[k] = ptref_interp_cubic(im, pitch, xy[2 * k], xy[2 * k + 1]);
}

/* NCC_to_minimize_xyR, src/p-tracker.cpp:1018-1088. pattern0_rotated already carries
 * refrot, so the rotation here is drot + X[2] only (:1035). */
double ptref_ncc_objective(const uint16_t* I0, const uint16_t* I1, int64_t pitch,
                           const ptg_grain* g, int32_t n, const int32_t* p0_xy, const double* X,
                           double subpix_displacement_max, double overflow_stiffness) {
  double inv_num_pix = 1.0 / (double)n;
  int xpix0 = g->refcoord_xpix, ypix0 = g->refcoord_ypix;
  double xpix1 = (double)g->refcoord_xpix + (double)g->dx + X[0];
  double ypix1 = (double)g->refcoord_ypix + (double)g->dy + X[1];
  double rot = g->drot + X[2];
  double c1 = cos(rot), s1 = sin(rot);
  double* vals = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  double mean1 = 0.0;
  for (int i = 0; i < n; i++) {
    double x = p0_xy[2 * i], y = p0_xy[2 * i + 1];
    double xs = xpix1 + c1 * x + s1 * y;
    double ys = ypix1 - s1 * x + c1 * y;
    double v = ptref_interp_cubic(I1, pitch, xs, ys);
    vals[i] = v;
    mean1 += v;
  }
  mean1 *= inv_num_pix;
  double C0C1 = 0.0, C1C1 = 0.0;
  for (int i = 0; i < n; i++) {
    double diffC1 = vals[i] - mean1;
    double diffC0 = (double)PX(I0, pitch, xpix0 + p0_xy[2 * i], ypix0 + p0_xy[2 * i + 1]) - g->mean0;
    C0C1 += diffC0 * diffC1;
    C1C1 += diffC1 * diffC1;
  }
  free(vals);
  double NCC = C0C1 / sqrt(g->C0C0 * C1C1);
  double W = 1.0;
  double disp = sqrt(X[0] * X[0] + X[1] * X[1]);
  if (disp > subpix_displacement_max) W = overflow_stiffness * (disp - subpix_displacement_max) + 1.0;
  return (1.0 - NCC) * W;
}

/* ---------------------------------------------------------------------------------
 * Powell's method (M.J.D. Powell, Computer Journal 7, 1964) with the direction-
 * replacement test of that paper, and Brent's derivative-free line minimisation
 * (R.P. Brent, "Algorithms for Minimization without Derivatives", 1973, ch. 5),
 * preceded by a golden-ratio expansion that brackets the minimum. The GPU optimizer
 * runs the same control flow, so this is its reference trajectory. */
#define PW_GOLD 1.618033988749895
#define PW_CGOLD 0.3819660112501051
#define PW_LINE_TOL 3.0e-8
#define PW_ZEPS 1.0e-12
#define PW_TINY 1.0e-25
#define PW_BRACKET_MAX 60
#define PW_BRENT_MAX 100

typedef struct {
  ptref_fn3 f;
  void* user;
  const double* p;
  const double* xi;
  int64_t nevals;
} line_ctx;

static double line_eval(line_ctx* L, double a) {
  double x[3];
  for (int j = 0; j < 3; j++) x[j] = L->p[j] + a * L->xi[j];
  L->nevals++;
  return L->f(x, L->user);
}

/* Minimise f(p + a*xi) over a, starting from a = 0 with f(p) = f0. Returns the
 * minimising a; *fmin receives its value. */
static double line_minimize(line_ctx* L, double f0, double* fmin) {
  /* bracket: a < b < c (or reversed) with f(b) <= f(a), f(b) < f(c) */
  double a = 0.0, fa = f0, b = 1.0, fb = line_eval(L, b);
  if (fb > fa) {
    double t = a; a = b; b = t;
    t = fa; fa = fb; fb = t;
  }
  double c = b + PW_GOLD * (b - a), fc = line_eval(L, c);
  for (int it = 0; it < PW_BRACKET_MAX && fc < fb; it++) {
    a = b; fa = fb;
    b = c; fb = fc;
    c = b + PW_GOLD * (b - a);
    fc = line_eval(L, c);
  }
  (void)fa;
  /* Brent's localmin on [lo, hi] from the best point b */
  double lo = a < c ? a : c, hi = a < c ? c : a;
  double x = b, w = b, v = b, fx = fb, fw = fb, fv = fb;
  double d = 0.0, e = 0.0;
  for (int it = 0; it < PW_BRENT_MAX; it++) {
    double m = 0.5 * (lo + hi);
    double tol1 = PW_LINE_TOL * fabs(x) + PW_ZEPS;
    double t2 = 2.0 * tol1;
    if (fabs(x - m) <= t2 - 0.5 * (hi - lo)) break;
    int golden = 1;
    if (fabs(e) > tol1) {
      double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p; else q = -q;
      double etemp = e;
      e = d;
      if (fabs(p) < fabs(0.5 * q * etemp) && p > q * (lo - x) && p < q * (hi - x)) {
        d = p / q;
        double u = x + d;
        if (u - lo < t2 || hi - u < t2) d = (x < m) ? tol1 : -tol1;
        golden = 0;
      }
    }
    if (golden) {
      e = (x < m) ? hi - x : lo - x;
      d = PW_CGOLD * e;
    }
    double u = x + (fabs(d) >= tol1 ? d : (d > 0.0 ? tol1 : -tol1));
    double fu = line_eval(L, u);
    if (fu <= fx) {
      if (u < x) hi = x; else lo = x;
      v = w; fv = fw;
      w = x; fw = fx;
      x = u; fx = fu;
    } else {
      if (u < x) lo = u; else hi = u;
      if (fu <= fw || w == x) {
        v = w; fv = fw;
        w = u; fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u; fv = fu;
      }
    }
  }
  *fmin = fx;
  return x;
}

double ptref_powell(ptref_fn3 f, void* user, double* p, const double* dx, double ftol,
                    int32_t max_iter, int64_t* nevals) {
  double u[3][3] = {{dx[0], 0.0, 0.0}, {0.0, dx[1], 0.0}, {0.0, 0.0, dx[2]}};
  if (max_iter <= 0) max_iter = 200;
  line_ctx L = {f, user, NULL, NULL, 0};
  L.nevals = 1;
  double fret = f(p, user);
  if (fret != fret) { /* undefined ZNCC (flat pattern): nothing to minimise */
    if (nevals) *nevals += L.nevals;
    return fret;
  }
  double pt[3] = {p[0], p[1], p[2]};
  for (int iter = 0; iter < max_iter; iter++) {
    double fp = fret, del = 0.0;
    int ibig = 0;
    for (int i = 0; i < 3; i++) {
      double fstart = fret, fmin;
      L.p = p;
      L.xi = u[i];
      double a = line_minimize(&L, fret, &fmin);
      for (int j = 0; j < 3; j++) {
        u[i][j] *= a;
        p[j] += u[i][j];
      }
      fret = fmin;
      if (fstart - fret > del) {
        del = fstart - fret;
        ibig = i;
      }
    }
    if (2.0 * (fp - fret) <= ftol * (fabs(fp) + fabs(fret)) + PW_TINY) break;
    double ptt[3], xit[3];
    for (int j = 0; j < 3; j++) {
      ptt[j] = 2.0 * p[j] - pt[j];
      xit[j] = p[j] - pt[j];
      pt[j] = p[j];
    }
    L.nevals++;
    double fptt = f(ptt, user);
    if (fptt < fp) {
      double t1 = fp - fret - del, t3 = fp - fptt;
      double t = 2.0 * (fp - 2.0 * fret + fptt) * t1 * t1 - del * t3 * t3;
      if (t < 0.0) {
        double fmin;
        L.p = p;
        L.xi = xit;
        double a = line_minimize(&L, fret, &fmin);
        for (int j = 0; j < 3; j++) {
          xit[j] *= a;
          p[j] += xit[j];
        }
        fret = fmin;
        for (int j = 0; j < 3; j++) {
          u[ibig][j] = u[2][j];
          u[2][j] = xit[j];
        }
      }
    }
  }
  if (nevals) *nevals += L.nevals;
  return fret;
}

typedef struct {
  const uint16_t *I0, *I1;
  int64_t pitch;
  const ptg_grain* g;
  int32_t n;
  const int32_t* p0;
  double dmax, stiff;
} obj_ctx;

double ptref_minimize_fp64(ptref_fn3 f, void* user, double* x, const double* dx, double ftol,
                           int32_t max_iter, int32_t max_restarts, int64_t* nevals) {
  double fret = ptref_powell(f, user, x, dx, ftol, max_iter, nevals);
  if (fret != fret) return fret;
  for (int r = 0; r < max_restarts; r++) {
    double xr[3] = {x[0], x[1], x[2]};
    const double fr = ptref_powell(f, user, xr, dx, ftol, max_iter, nevals);
    if (!(fr < fret)) break;
    x[0] = xr[0];
    x[1] = xr[1];
    x[2] = xr[2];
    fret = fr;
  }
  return fret;
}

static double obj_fn(const double* X, void* user) {
  obj_ctx* o = (obj_ctx*)user;
  return ptref_ncc_objective(o->I0, o->I1, o->pitch, o->g, o->n, o->p0, X, o->dmax, o->stiff);
}

/* follow_pattern_subpixel_xyR, src/p-tracker.cpp:1090-1128. */
void ptref_subpixel(const uint16_t* I0, const uint16_t* I1, int64_t pitch, ptg_grain* g,
                    int32_t n, const int32_t* p0_xy, const ptg_subpix_params* prm,
                    int64_t* nevals) {
  double X[3] = {0.0, 0.0, 0.0};
  obj_ctx o = {I0, I1, pitch, g, n, p0_xy, prm->subpix_displacement_max, prm->overflow_stiffness};
  double fret = prm->method == PTG_SUBPIX_LM
                    ? ptref_minimize_fp64(obj_fn, &o, X, prm->initial_direction, prm->subpix_tol, prm->max_iter, 20,
                                          nevals)
                    : ptref_powell(obj_fn, &o, X, prm->initial_direction, prm->subpix_tol, prm->max_iter, nevals);
  g->NCC_subpix = 1.0 - fret;
  g->upix += X[0];
  g->vpix += X[1];
  g->rot_inc += X[2];
  g->dx += X[0];
  g->dy += X[1];
  g->drot += X[2];
}

/* particle_tracking's per-pair body, src/p-tracker.cpp:460-581. */
int ptref_track_pair(const uint16_t* I0, const uint16_t* I1, int64_t pitch, int32_t G,
                     ptg_grain* grains, const int32_t* pattern_xy, const int64_t* pattern_begin,
                     const ptg_track_params* prm, int32_t nthreads, int32_t* counts,
                     int64_t* nevals_out) {
  int64_t total = pattern_begin[G];
  int32_t* p0 = (int32_t*)malloc(sizeof(int32_t) * 2 * (total > 0 ? total : 1));
  int32_t* rescue = (int32_t*)malloc(sizeof(int32_t) * (G > 0 ? G : 1));
  int32_t* super = (int32_t*)malloc(sizeof(int32_t) * (G > 0 ? G : 1));
  int32_t nres = 0, nsup = 0;
  int64_t nevals = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  /* do_precomputations (:460-465) */
#pragma omp parallel for schedule(dynamic)
  for (int g = 0; g < G; g++) {
    int64_t b = pattern_begin[g];
    ptref_precompute(I0, pitch, &grains[g], (int32_t)(pattern_begin[g + 1] - b), pattern_xy + 2 * b, p0 + 2 * b);
  }
  if (prm->rescue_level >= 0) { /* stage 1, :472-484 */
#pragma omp parallel
    {
      int32_t* scratch = NULL;
      int64_t cap = 0;
#pragma omp for schedule(dynamic)
      for (int g = 0; g < G; g++) {
        ptg_grain* gr = &grains[g];
        gr->NCC = gr->NCC_rescue = gr->NCC_subpix = 0.0; /* reset(), src/grain_type_2D.cpp:57 */
        gr->dx_prev = gr->dx;                            /* backup(), :51-55 */
        gr->dy_prev = gr->dy;
        gr->drot_prev = gr->drot;
        if (gr->masked) continue;
        int64_t b = pattern_begin[g], n = pattern_begin[g + 1] - b;
        if (n > cap) {
          free(scratch);
          cap = n;
          scratch = (int32_t*)malloc(sizeof(int32_t) * 2 * cap);
        }
        if (ptref_follow_pattern(I0, I1, pitch, gr, (int32_t)n, pattern_xy + 2 * b, p0 + 2 * b,
                                 &prm->search_zone, PTG_STAGE_MAIN, prm->NCC_min, prm->NCC_min_super, scratch)) {
#pragma omp critical
          rescue[nres++] = g;
        }
      }
      free(scratch);
    }
  }
  if (prm->rescue_level >= 1) { /* stage 2, :488-504 */
#pragma omp parallel
    {
      int32_t* scratch = NULL;
      int64_t cap = 0;
#pragma omp for schedule(dynamic)
      for (int k = 0; k < nres; k++) {
        int g = rescue[k];
        int64_t b = pattern_begin[g], n = pattern_begin[g + 1] - b;
        if (n > cap) {
          free(scratch);
          cap = n;
          scratch = (int32_t*)malloc(sizeof(int32_t) * 2 * cap);
        }
        if (ptref_follow_pattern(I0, I1, pitch, &grains[g], (int32_t)n, pattern_xy + 2 * b, p0 + 2 * b,
                                 &prm->search_zone_rescue, PTG_STAGE_RESCUE, prm->NCC_min,
                                 prm->NCC_min_super, scratch)) {
#pragma omp critical
          super[nsup++] = g;
        }
      }
      free(scratch);
    }
  }
  if (prm->rescue_level >= 2) { /* stage 3, :521-537 */
#pragma omp parallel
    {
      int32_t* scratch = NULL;
      int64_t cap = 0;
#pragma omp for schedule(dynamic)
      for (int k = 0; k < nsup; k++) {
        int g = super[k];
        int64_t b = pattern_begin[g], n = pattern_begin[g + 1] - b;
        if (n > cap) {
          free(scratch);
          cap = n;
          scratch = (int32_t*)malloc(sizeof(int32_t) * 2 * cap);
        }
        ptref_follow_pattern(I0, I1, pitch, &grains[g], (int32_t)n, pattern_xy + 2 * b, p0 + 2 * b,
                             &prm->search_zone_super_rescue, PTG_STAGE_SUPER, prm->NCC_min,
                             prm->NCC_min_super, scratch);
      }
      free(scratch);
    }
  }
  if (prm->subpixel) { /* stage 4, :556-572 */
#pragma omp parallel for schedule(dynamic) reduction(+ : nevals)
    for (int g = 0; g < G; g++) {
      if (grains[g].masked) continue;
      int64_t b = pattern_begin[g], n = pattern_begin[g + 1] - b, ne = 0;
      ptref_subpixel(I0, I1, pitch, &grains[g], (int32_t)n, p0 + 2 * b, &prm->subpix, &ne);
      nevals += ne;
    }
  }
  for (int g = 0; g < G; g++) { /* :577-581 */
    grains[g].upix = grains[g].dx - grains[g].dx_prev;
    grains[g].vpix = grains[g].dy - grains[g].dy_prev;
    grains[g].rot_inc = grains[g].drot - grains[g].drot_prev;
  }
  if (counts) {
    counts[0] = nres;
    counts[1] = nsup;
  }
  if (nevals_out) *nevals_out = nevals;
  free(p0);
  free(rescue);
  free(super);
  return 0;
}

/* src/p-tracker.cpp:598-609. */
void ptref_change_reference(int32_t G, ptg_grain* grains) {
  for (int g = 0; g < G; g++) {
    ptg_grain* gr = &grains[g];
    double actual_pos = (double)(gr->refcoord_xpix) + gr->dx;
    gr->refcoord_xpix = (int32_t)round(actual_pos);
    gr->dx = actual_pos - gr->refcoord_xpix;
    actual_pos = (double)(gr->refcoord_ypix) + gr->dy;
    gr->refcoord_ypix = (int32_t)round(actual_pos);
    gr->dy = actual_pos - gr->refcoord_ypix;
    gr->refrot += gr->drot;
    gr->drot = 0.0;
  }
}

/* undistor, src/distortion_correction.cpp:20-61: Brown-Conrady radial (K1..K3) and
 * tangential (P1..P3) terms, then the two-axis parallax scaling. */
void ptref_undistor(const double* X, double xd, double yd, double* xu, double* yu) {
  double Xc = X[0], Yc = X[1], K1 = X[2], K2 = X[3], K3 = X[4], P1 = X[5], P2 = X[6], P3 = X[7];
  double dx = xd - Xc, dy = yd - Yc;
  double r2 = dx * dx + dy * dy;
  double r4 = r2 * r2;
  double r6 = r4 * r2;
  double x = xd + dx * (K1 * r2 + K2 * r4 + K3 * r6) + (P1 * (r2 + 2.0 * dx * dx) + 2.0 * P2 * dx * dy) * (1.0 + P3 * r2);
  double y = yd + dy * (K1 * r2 + K2 * r4 + K3 * r6) + (2.0 * P1 * dx * dy + P2 * (r2 + 2.0 * dy * dy)) * (1.0 + P3 * r2);
  double Xcp = X[8], Ycp = X[9], theta1 = X[10], parallax1 = X[11], theta2 = X[12], parallax2 = X[13];
  double nx1 = cos(theta1), ny1 = -sin(theta1);
  double nx2 = -sin(theta2), ny2 = cos(theta2);
  double dX = x - Xcp, dY = y - Ycp;
  double S1 = (1.0 + parallax1) * (dX * nx1 + dY * ny1);
  double S2 = (1.0 + parallax2) * (dX * nx2 + dY * ny2);
  *xu = Xcp + S1 * nx1 + S2 * nx2;
  *yu = Ycp + S1 * ny1 + S2 * ny2;
}

int64_t ptref_precompute_pairs(int32_t G, const double* ref_xy, double dmin, double dmax, int32_t* pairs, int64_t cap) {
  int64_t n = 0;
  for (int32_t i = 0; i < G; i++) {
    const double xi = ref_xy[2 * i], yi = ref_xy[2 * i + 1];
    for (int32_t j = i + 1; j < G; j++) {
      const double xij = ref_xy[2 * j] - xi, yij = ref_xy[2 * j + 1] - yi;
      const double dij = sqrt(xij * xij + yij * yij);
      if (dij >= dmin && dij <= dmax) {
        if (n < cap) {
          pairs[2 * n] = i;
          pairs[2 * n + 1] = j;
        }
        n++;
      }
    }
  }
  return n;
}

void ptref_equiproj_residuals(const double* X, int32_t G, const double* ref_xy, int32_t nimg, const double* disp,
                              const double* ncc, int32_t npairs, const int32_t* pairs, double NCC_min, double* res) {
  for (int32_t k = 0; k < nimg; k++) {
    const double* d = disp + 2 * (int64_t)k * G;
    const double* nc = ncc + (int64_t)k * G;
    for (int32_t l = 0; l < npairs; l++) {
      const int32_t i = pairs[2 * l], j = pairs[2 * l + 1];
      double* r = &res[(int64_t)k * npairs + l];
      *r = 0.0;
      if (nc[i] < NCC_min || nc[j] < NCC_min) continue;
      double A1x, A1y, A2x, A2y, B1x, B1y, B2x, B2y;
      ptref_undistor(X, ref_xy[2 * i], ref_xy[2 * i + 1], &A1x, &A1y);
      ptref_undistor(X, ref_xy[2 * i] + d[2 * i], ref_xy[2 * i + 1] + d[2 * i + 1], &A2x, &A2y);
      ptref_undistor(X, ref_xy[2 * j], ref_xy[2 * j + 1], &B1x, &B1y);
      ptref_undistor(X, ref_xy[2 * j] + d[2 * j], ref_xy[2 * j + 1] + d[2 * j + 1], &B2x, &B2y);
      double ABx = B1x - A1x, ABy = B1y - A1y;
      const double AB = sqrt(ABx * ABx + ABy * ABy);
      ABx /= AB;
      ABy /= AB;
      const double BBx = B2x - B1x, BBy = B2y - B1y, AAx = A2x - A1x, AAy = A2y - A1y;
      *r = (ABx * (BBx - AAx)) + (ABy * (BBy - AAy));
    }
  }
}

/* src/distortion_correction.cpp:316-354 (POWELL_ITER): A1 = orig_i, B1 = orig_j. */
void ptref_equiproj_residuals_orig(const double* X, int32_t G, const double* orig_xy, const double* ref_xy,
                                   int32_t nimg, const double* disp, const double* ncc, int32_t npairs,
                                   const int32_t* pairs, double NCC_min, double* res) {
  for (int32_t k = 0; k < nimg; k++) {
    const double* d = disp + 2 * (int64_t)k * G;
    const double* nc = ncc + (int64_t)k * G;
    for (int32_t l = 0; l < npairs; l++) {
      const int32_t i = pairs[2 * l], j = pairs[2 * l + 1];
      double* r = &res[(int64_t)k * npairs + l];
      *r = 0.0;
      if (nc[i] < NCC_min || nc[j] < NCC_min) continue;
      double A2x, A2y, B2x, B2y;
      const double A1x = orig_xy[2 * i], A1y = orig_xy[2 * i + 1];
      ptref_undistor(X, ref_xy[2 * i] + d[2 * i], ref_xy[2 * i + 1] + d[2 * i + 1], &A2x, &A2y);
      const double B1x = orig_xy[2 * j], B1y = orig_xy[2 * j + 1];
      ptref_undistor(X, ref_xy[2 * j] + d[2 * j], ref_xy[2 * j + 1] + d[2 * j + 1], &B2x, &B2y);
      double ABx = B1x - A1x, ABy = B1y - A1y;
      const double AB = sqrt(ABx * ABx + ABy * ABy);
      ABx /= AB;
      ABy /= AB;
      const double BBx = B2x - B1x, BBy = B2y - B1y, AAx = A2x - A1x, AAy = A2y - A1y;
      *r = (ABx * (BBx - AAx)) + (ABy * (BBy - AAy));
    }
  }
}

/* Jacobian of undistor with respect to (xd, yd), by differentiating the expressions
 * above: J = Jparallax * Jbrown. */
static void undistor_jac(const double* X, double xd, double yd, double J[4]) {
  double Xc = X[0], Yc = X[1], K1 = X[2], K2 = X[3], K3 = X[4], P1 = X[5], P2 = X[6], P3 = X[7];
  double dx = xd - Xc, dy = yd - Yc;
  double r2 = dx * dx + dy * dy;
  double r4 = r2 * r2;
  double Kr = K1 * r2 + K2 * r4 + K3 * (r4 * r2);
  double Kp = K1 + 2.0 * K2 * r2 + 3.0 * K3 * r4; /* d(Kr)/d(r2) */
  double A = P1 * (r2 + 2.0 * dx * dx) + 2.0 * P2 * dx * dy;
  double B = 2.0 * P1 * dx * dy + P2 * (r2 + 2.0 * dy * dy);
  double S = 1.0 + P3 * r2;
  double Ax = 6.0 * P1 * dx + 2.0 * P2 * dy, Ay = 2.0 * P1 * dy + 2.0 * P2 * dx;
  double Bx = 2.0 * P1 * dy + 2.0 * P2 * dx, By = 2.0 * P1 * dx + 6.0 * P2 * dy;
  double Sx = 2.0 * P3 * dx, Sy = 2.0 * P3 * dy;
  double bxx = 1.0 + Kr + dx * Kp * 2.0 * dx + (Ax * S + A * Sx);
  double bxy = dx * Kp * 2.0 * dy + (Ay * S + A * Sy);
  double byx = dy * Kp * 2.0 * dx + (Bx * S + B * Sx);
  double byy = 1.0 + Kr + dy * Kp * 2.0 * dy + (By * S + B * Sy);
  double theta1 = X[10], parallax1 = X[11], theta2 = X[12], parallax2 = X[13];
  double nx1 = cos(theta1), ny1 = -sin(theta1);
  double nx2 = -sin(theta2), ny2 = cos(theta2);
  double a1 = 1.0 + parallax1, a2 = 1.0 + parallax2;
  double pxx = a1 * nx1 * nx1 + a2 * nx2 * nx2, pxy = a1 * ny1 * nx1 + a2 * ny2 * nx2;
  double pyx = a1 * nx1 * ny1 + a2 * nx2 * ny2, pyy = a1 * ny1 * ny1 + a2 * ny2 * ny2;
  J[0] = pxx * bxx + pxy * byx;
  J[1] = pxx * bxy + pxy * byy;
  J[2] = pyx * bxx + pyy * byx;
  J[3] = pyx * bxy + pyy * byy;
}

#define REMAP_NEWTON_ITERS 6

void ptref_distort_point(const double* X, double xu, double yu, double* xd, double* yd) {
  double x = xu, y = yu;
  for (int it = 0; it < REMAP_NEWTON_ITERS; it++) {
    double fx, fy, J[4];
    ptref_undistor(X, x, y, &fx, &fy);
    fx -= xu;
    fy -= yu;
    undistor_jac(X, x, y, J);
    double det = J[0] * J[3] - J[1] * J[2];
    x -= (J[3] * fx - J[1] * fy) / det;
    y -= (J[0] * fy - J[2] * fx) / det;
  }
  *xd = x;
  *yd = y;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void ptref_remap(const uint16_t* src, int32_t dimx, int32_t dimy, int64_t pitch, int32_t bits,
                 const double* Xin, uint16_t* dst, int64_t dst_pitch) {
  double X[14];
  memcpy(X, Xin, sizeof(X));
  /* centres left negative default to the image centre, src/distortion_correction.cpp:230-233 */
  if (X[0] < 0.0) X[0] = 0.5 * (double)dimx;
  if (X[1] < 0.0) X[1] = 0.5 * (double)dimy;
  if (X[8] < 0.0) X[8] = 0.5 * (double)dimx;
  if (X[9] < 0.0) X[9] = 0.5 * (double)dimy;
  double maxv = (double)((1 << bits) - 1);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dimy; y++)
    for (int x = 0; x < dimx; x++) {
      double xd, yd;
      ptref_distort_point(X, (double)x, (double)y, &xd, &yd);
      double v = 0.0;
      if (xd > -1.0e6 && xd < 1.0e6 && yd > -1.0e6 && yd < 1.0e6) {
        int ix = (int)floor(xd), iy = (int)floor(yd);
        double y01 = yd - (double)iy, x01 = xd - (double)ix;
        double p[4][4], arr[4];
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++)
            p[i][j] = PX(src, pitch, clampi(ix - 1 + i, 0, dimx - 1), clampi(iy - 1 + j, 0, dimy - 1));
        for (int i = 0; i < 4; i++) arr[i] = cubic1(p[i], y01);
        v = cubic1(arr, x01);
      }
      v = round(v);
      if (!(v > 0.0)) v = 0.0;
      if (v > maxv) v = maxv;
      dst[(int64_t)y * dst_pitch + x] = (uint16_t)v;
    }
}

/* ---------------------------------------------------------------------------------
 * The differentiation form's joint objective (test infrastructure; the reference has no
 * counterpart: it fits the lens only through the equiprojectivity criterion). Pattern pixel
 * q = ref + p0_i of grain g maps to undistorted coordinates by undistor()
 * (src/distortion_correction.cpp:20-61, radial K1 = lens[0], K2 = lens[1] about
 * (lens[2], lens[3])), moves rigidly there by T = (P[0], P[1]) and the rotation P[2] about
 * undistor(ref), with the rotation convention of NCC_to_minimize_xyR ((x, y) -> (c x + s y,
 * -s x + c y), src/p-tracker.cpp:1040-1045), and is sampled in I1 at the distorted position
 * of the result (ptref_distort_point: Newton on the general undistor). The value is
 * 1 - NCC with the reference's two-pass sums and sampler (:1018-1088). With K1 = K2 = 0 it is
 * NCC_to_minimize_xyR at X = (P[0] - dx, P[1] - dy, P[2] - drot). */
double ptref_joint_objective(const uint16_t* I0, const uint16_t* I1, int64_t pitch, const ptg_grain* g,
                             int32_t n, const int32_t* p0_xy, const double* P, const double* lens) {
  double X[14] = {lens[2], lens[3], lens[0], lens[1], 0.0, 0.0, 0.0, 0.0, lens[2], lens[3], 0.0, 0.0, 0.0, 0.0};
  int xr = g->refcoord_xpix, yr = g->refcoord_ypix;
  double wrx, wry;
  ptref_undistor(X, (double)xr, (double)yr, &wrx, &wry);
  double c1 = cos(P[2]), s1 = sin(P[2]);
  double* vals = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  double inv_num_pix = 1.0 / (double)n, mean1 = 0.0;
  for (int i = 0; i < n; i++) {
    double wx, wy;
    ptref_undistor(X, (double)(xr + p0_xy[2 * i]), (double)(yr + p0_xy[2 * i + 1]), &wx, &wy);
    double ax = wx - wrx, ay = wy - wry;
    double tx = wrx + P[0] + c1 * ax + s1 * ay;
    double ty = wry + P[1] - s1 * ax + c1 * ay;
    double xs, ys;
    ptref_distort_point(X, tx, ty, &xs, &ys);
    double v = ptref_interp_cubic(I1, pitch, xs, ys);
    vals[i] = v;
    mean1 += v;
  }
  mean1 *= inv_num_pix;
  double C0C1 = 0.0, C1C1 = 0.0;
  for (int i = 0; i < n; i++) {
    double diffC1 = vals[i] - mean1;
    double diffC0 = (double)PX(I0, pitch, xr + p0_xy[2 * i], yr + p0_xy[2 * i + 1]) - g->mean0;
    C0C1 += diffC0 * diffC1;
    C1C1 += diffC1 * diffC1;
  }
  free(vals);
  return 1.0 - C0C1 / sqrt(g->C0C0 * C1C1);
}

/* ptref_joint_objective for grains list[0..n-1] (OpenMP over grains): f[k] for grain list[k] at
 * P[3k..3k+2]; pattern0_rotated offsets p0_xy with pattern_begin as the tracker's CSR layout. */
void ptref_joint_objective_batch(const uint16_t* I0, const uint16_t* I1, int64_t pitch, const ptg_grain* grains,
                                 const int32_t* p0_xy, const int64_t* pattern_begin, int32_t n,
                                 const int32_t* list, const double* P, const double* lens, double* f,
                                 int32_t nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; k++) {
    const int g = list[k];
    const int64_t b = pattern_begin[g];
    f[k] = ptref_joint_objective(I0, I1, pitch, &grains[g], (int32_t)(pattern_begin[g + 1] - b), p0_xy + 2 * b,
                                 P + 3 * k, lens);
  }
}
