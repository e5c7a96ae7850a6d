// ptg_subpixel.hip — the subpixel stage of libptgpu (gfx950): the ZNCC objective
// NCC_to_minimize_xyR (src/p-tracker.cpp:1018-1088) evaluated by one wave per grain,
// Powell's method with Brent line searches (the derivative-free optimizer of
// follow_pattern_subpixel_xyR, :1090-1128), and a Levenberg-Marquardt form driven by
// the analytic derivative of the ZNCC residual through the bicubic sampler.
//
// One wave owns one grain. Every lane runs the optimizer's scalar control flow on
// identical values; the objective's pixel sums are split across the 64 lanes and
// combined with butterfly reductions whose result is bit-identical in every lane, so
// all branches stay wave-uniform.
#include <algorithm>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

struct Obj {
  const uint16_t* __restrict__ I1;
  int dimx, dimy;
  int64_t pitch;
  // the grain's neighbourhood of I1 copied into LDS (taps outside it read global memory)
  const uint32_t* tile;  // one 32-bit word per pixel: aligned LDS reads
  int tx0, ty0, tw, th;
  const int32_t* __restrict__ p0;  // pattern0_rotated pairs of this grain
  const double* __restrict__ d0;   // I0(p0) - mean0 of this grain
  int n;
  double xr, yr;   // refcoord + dx, refcoord + dy (:1031-1032 before adding X)
  double drot;
  double C0C0, sd0, shift;
  double dmax, stiff;
  bool clamped;
  long long evals;
};

__device__ __forceinline__ void load_obj(Obj& o, const ptg_grain& g, const uint16_t* I1, int dimx, int dimy,
                                         int64_t pitch, const int32_t* p0, const double* d0, int n, double sd0,
                                         double dmax, double stiff) {
  o.I1 = I1;
  o.dimx = dimx;
  o.dimy = dimy;
  o.pitch = pitch;
  o.tile = nullptr;
  o.tx0 = o.ty0 = 0;
  o.tw = o.th = 0;
  o.p0 = p0;
  o.d0 = d0;
  o.n = n;
  o.xr = (double)g.refcoord_xpix + (double)g.dx;
  o.yr = (double)g.refcoord_ypix + (double)g.dy;
  o.drot = g.drot;
  o.C0C0 = g.C0C0;
  o.sd0 = sd0;
  o.shift = g.mean0;  // sums run on I1 - mean0 to keep the one-pass variance well conditioned
  o.dmax = dmax;
  o.stiff = stiff;
  o.clamped = false;
  o.evals = 0;
}

// image_interpCubic::getValue (src/interpol.cpp:55-86) reading the 4x4 stencil from the
// LDS tile when it lies inside it (the tile holds the frame's values, edge-clamped as the
// global path clamps), else from global memory.
__device__ __forceinline__ double sample(const Obj& o, double x, double y, bool& clamped) {
  const int ix = (int)floor(x), iy = (int)floor(y);
  const int lx = ix - 1 - o.tx0, ly = iy - 1 - o.ty0;
  if (o.tile && lx >= 0 && ly >= 0 && lx < o.tw - 3 && ly < o.th - 3) {
    if (!(ix >= 1 && iy >= 1 && ix < o.dimx - 2 && iy < o.dimy - 2)) clamped = true;
    const double y01 = y - (double)iy, x01 = x - (double)ix;
    const uint32_t* r0 = o.tile + ly * o.tw + lx;
    double arr[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      arr[i] = cubic1((double)r0[i], (double)r0[o.tw + i], (double)r0[2 * o.tw + i], (double)r0[3 * o.tw + i], y01);
    return cubic1(arr[0], arr[1], arr[2], arr[3], x01);
  }
  return interp_cubic(o.I1, o.dimx, o.dimy, o.pitch, x, y, clamped);
}

__device__ __forceinline__ double weight(const Obj& o, const double* X) {
  double W = 1.0;
  const double disp = sqrt(X[0] * X[0] + X[1] * X[1]);
  if (disp > o.dmax) W = o.stiff * (disp - o.dmax) + 1.0;
  return W;
}

// (1 - NCC) * W at X. The reference's two passes (mean, then centred products) are
// replaced by one pass over shifted sums: C1C1 = Sgg - Sg^2/N, C0C1 = Sdg - Sg*Sd0/N.
__device__ __noinline__ double eval(Obj& o, const double* X) {
  o.evals++;
  const int lane = threadIdx.x & 63;
  const double xpix1 = o.xr + X[0], ypix1 = o.yr + X[1];
  const double rot = o.drot + X[2];
  const double c1 = cos(rot), s1 = sin(rot);
  double Sg = 0.0, Sgg = 0.0, Sdg = 0.0;
  bool cl = false;
  for (int i = lane; i < o.n; i += 64) {
    const double x = (double)o.p0[2 * i], y = (double)o.p0[2 * i + 1];
    const double xs = xpix1 + c1 * x + s1 * y;
    const double ys = ypix1 - s1 * x + c1 * y;
    const double v = sample(o, xs, ys, cl) - o.shift;
    Sg += v;
    Sgg = fma(v, v, Sgg);
    Sdg = fma(o.d0[i], v, Sdg);
  }
  if (cl) o.clamped = true;
  Sg = wave_sum(Sg);
  Sgg = wave_sum(Sgg);
  Sdg = wave_sum(Sdg);
  const double N = (double)o.n;
  const double C1C1 = Sgg - Sg * Sg / N;
  const double C0C1 = Sdg - Sg * o.sd0 / N;
  const double NCC = C0C1 / sqrt(o.C0C0 * C1C1);
  return (1.0 - NCC) * weight(o, X);
}

// Value and gradient at X in one pass. g_i = I1(p_i(X)); its derivatives come from
// differentiating the Catmull-Rom sampler: d/dx of cubic1(arr, x01) and cubic1 along y
// of the y-derivatives. Returns f; writes grad[3]; optionally the Gauss-Newton matrix
// of the ZNCC residual r_i = (g_i - m1)/sqrt(C1C1) - d0_i/sqrt(C0C0), whose squared
// norm is 2 (1 - NCC): H = (sum h h^T - beta beta^T) / C1C1 with h_i = dg_i - mean(dg)
// and beta = sum ghat_i h_i.
__device__ __forceinline__ double cubic1_d(double p0, double p1, double p2, double p3, double x) {
  const double fp0 = 0.5 * (p2 - p0);
  const double fp1 = 0.5 * (p3 - p1);
  const double b = -3.0 * p1 + 3.0 * p2 - 2.0 * fp0 - fp1;
  const double a = 2.0 * p1 - 2.0 * p2 + fp0 + fp1;
  return (3.0 * a * x + 2.0 * b) * x + fp0;
}

__device__ __forceinline__ void interp_cubic_grad(const Obj& o, double x, double y, double& v, double& gx,
                                                  double& gy, bool& clamped) {
  const uint16_t* __restrict__ im = o.I1;
  const int dimx = o.dimx, dimy = o.dimy;
  const int64_t pitch = o.pitch;
  const int ix = (int)floor(x), iy = (int)floor(y);
  const double y01 = y - (double)iy, x01 = x - (double)ix;
  double arr[4], darr[4];
  const bool inside = ix >= 1 && iy >= 1 && ix < dimx - 2 && iy < dimy - 2;
  if (!inside) clamped = true;
  const int lx = ix - 1 - o.tx0, ly = iy - 1 - o.ty0;
  const bool in_tile = o.tile && lx >= 0 && ly >= 0 && lx < o.tw - 3 && ly < o.th - 3;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double p[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (in_tile) {
        p[j] = (double)o.tile[(ly + j) * o.tw + lx + i];
      } else {
        const int X = inside ? ix - 1 + i : clampi(ix - 1 + i, 0, dimx - 1);
        const int Y = inside ? iy - 1 + j : clampi(iy - 1 + j, 0, dimy - 1);
        p[j] = (double)im[(int64_t)Y * pitch + X];
      }
    }
    arr[i] = cubic1(p[0], p[1], p[2], p[3], y01);
    darr[i] = cubic1_d(p[0], p[1], p[2], p[3], y01);
  }
  v = cubic1(arr[0], arr[1], arr[2], arr[3], x01);
  gx = cubic1_d(arr[0], arr[1], arr[2], arr[3], x01);
  gy = cubic1(darr[0], darr[1], darr[2], darr[3], x01);
}

struct Derivs {
  double f, NCC;
  double grad[3];
  double H[6];  // xx, xy, xt, yy, yt, tt of the Gauss-Newton matrix (W = 1)
  double JTr[3];
};

__device__ __noinline__ void eval_derivs(Obj& o, const double* X, Derivs& d) {
  o.evals++;
  const int lane = threadIdx.x & 63;
  const double xpix1 = o.xr + X[0], ypix1 = o.yr + X[1];
  const double rot = o.drot + X[2];
  const double c1 = cos(rot), s1 = sin(rot);
  // sums: g, gg, dg, dG[3], gdG[3], d0 dG[3], dG dG^T[6]
  double S[18];
#pragma unroll
  for (int k = 0; k < 18; k++) S[k] = 0.0;
  bool cl = false;
  for (int i = lane; i < o.n; i += 64) {
    const double x = (double)o.p0[2 * i], y = (double)o.p0[2 * i + 1];
    const double xs = xpix1 + c1 * x + s1 * y;
    const double ys = ypix1 - s1 * x + c1 * y;
    double v, gx, gy;
    interp_cubic_grad(o, xs, ys, v, gx, gy, cl);
    v -= o.shift;
    const double dd = o.d0[i];
    const double gt = gx * (-s1 * x + c1 * y) + gy * (-c1 * x - s1 * y);
    const double G[3] = {gx, gy, gt};
    S[0] += v;
    S[1] = fma(v, v, S[1]);
    S[2] = fma(dd, v, S[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      S[3 + k] += G[k];
      S[6 + k] = fma(v, G[k], S[6 + k]);
      S[9 + k] = fma(dd, G[k], S[9 + k]);
    }
    S[12] = fma(G[0], G[0], S[12]);
    S[13] = fma(G[0], G[1], S[13]);
    S[14] = fma(G[0], G[2], S[14]);
    S[15] = fma(G[1], G[1], S[15]);
    S[16] = fma(G[1], G[2], S[16]);
    S[17] = fma(G[2], G[2], S[17]);
  }
  if (cl) o.clamped = true;
#pragma unroll
  for (int k = 0; k < 18; k++) S[k] = wave_sum(S[k]);
  const double N = (double)o.n;
  const double m1 = S[0] / N;
  const double C1C1 = S[1] - S[0] * m1;
  const double C0C1 = S[2] - m1 * o.sd0;
  const double sig1 = sqrt(C1C1), sig0 = sqrt(o.C0C0);
  const double NCC = C0C1 / (sig0 * sig1);
  double mG[3], beta[3], hf[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    mG[k] = S[3 + k] / N;
    beta[k] = (S[6 + k] - m1 * S[3 + k]) / sig1;   // sum ghat_i h_i
    hf[k] = (S[9 + k] - mG[k] * o.sd0) / sig0;     // sum h_i fhat_i
  }
  // dNCC/dX = (sum h fhat - beta NCC) / sig1
  double dN[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dN[k] = (hf[k] - beta[k] * NCC) / sig1;
  const double W = weight(o, X);
  double dW[3] = {0.0, 0.0, 0.0};
  const double disp = sqrt(X[0] * X[0] + X[1] * X[1]);
  if (disp > o.dmax && disp > 0.0) {
    dW[0] = o.stiff * X[0] / disp;
    dW[1] = o.stiff * X[1] / disp;
  }
  d.NCC = NCC;
  d.f = (1.0 - NCC) * W;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    d.grad[k] = -dN[k] * W + (1.0 - NCC) * dW[k];
    d.JTr[k] = -dN[k];
  }
  const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int a = ij[q][0], b = ij[q][1];
    const double hh = S[12 + q] - N * mG[a] * mG[b];
    d.H[q] = (hh - beta[a] * beta[b]) / C1C1;
  }
}

// ---------------------------------------------------------------------------------
// Powell's method with Brent line searches; the same control flow and constants as
// the oracle's restatement, with eval() as the objective.
#define PW_GOLD 1.618033988749895
#define PW_CGOLD 0.3819660112501051
#define PW_LINE_TOL 3.0e-8
#define PW_ZEPS 1.0e-12
#define PW_TINY 1.0e-25
#define PW_BRACKET_MAX 60
#define PW_BRENT_MAX 100

__device__ double line_eval(Obj& o, const double* p, const double* xi, double a) {
  double x[3];
#pragma unroll
  for (int j = 0; j < 3; j++) x[j] = p[j] + a * xi[j];
  return eval(o, x);
}

__device__ __noinline__ double line_minimize(Obj& o, const double* p, const double* xi, double f0, double* fmin,
                                double line_tol = PW_LINE_TOL) {
  double a = 0.0, fa = f0, b = 1.0, fb = line_eval(o, p, xi, b);
  if (fb > fa) {
    double t = a; a = b; b = t;
    t = fa; fa = fb; fb = t;
  }
  double c = b + PW_GOLD * (b - a), fc = line_eval(o, p, xi, c);
  for (int it = 0; it < PW_BRACKET_MAX && fc < fb; it++) {
    a = b; fa = fb;
    b = c; fb = fc;
    c = b + PW_GOLD * (b - a);
    fc = line_eval(o, p, xi, c);
  }
  double lo = a < c ? a : c, hi = a < c ? c : a;
  double x = b, w = b, v = b, fx = fb, fw = fb, fv = fb;
  double d = 0.0, e = 0.0;
  for (int it = 0; it < PW_BRENT_MAX; it++) {
    const double m = 0.5 * (lo + hi);
    const double tol1 = line_tol * fabs(x) + PW_ZEPS;
    const double t2 = 2.0 * tol1;
    if (fabs(x - m) <= t2 - 0.5 * (hi - lo)) break;
    bool golden = true;
    if (fabs(e) > tol1) {
      double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double pp = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) pp = -pp; else q = -q;
      const double etemp = e;
      e = d;
      if (fabs(pp) < fabs(0.5 * q * etemp) && pp > q * (lo - x) && pp < q * (hi - x)) {
        d = pp / q;
        const double u = x + d;
        if (u - lo < t2 || hi - u < t2) d = (x < m) ? tol1 : -tol1;
        golden = false;
      }
    }
    if (golden) {
      e = (x < m) ? hi - x : lo - x;
      d = PW_CGOLD * e;
    }
    const double u = x + (fabs(d) >= tol1 ? d : (d > 0.0 ? tol1 : -tol1));
    const double fu = line_eval(o, p, xi, u);
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

__device__ double powell(Obj& o, double* p, const double* dx, double ftol, int max_iter) {
  double u[3][3] = {{dx[0], 0.0, 0.0}, {0.0, dx[1], 0.0}, {0.0, 0.0, dx[2]}};
  if (max_iter <= 0) max_iter = 200;
  double fret = eval(o, p);
  if (fret != fret) return fret;  // undefined ZNCC (flat pattern): nothing to minimise
  double pt[3] = {p[0], p[1], p[2]};
  for (int iter = 0; iter < max_iter; iter++) {
    const double fp = fret;
    double del = 0.0;
    int ibig = 0;
    for (int i = 0; i < 3; i++) {
      const double fstart = fret;
      double fmin;
      const double a = line_minimize(o, p, u[i], fret, &fmin);
#pragma unroll
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
#pragma unroll
    for (int j = 0; j < 3; j++) {
      ptt[j] = 2.0 * p[j] - pt[j];
      xit[j] = p[j] - pt[j];
      pt[j] = p[j];
    }
    const double fptt = eval(o, ptt);
    if (fptt < fp) {
      const double t1 = fp - fret - del, t3 = fp - fptt;
      const double t = 2.0 * (fp - 2.0 * fret + fptt) * t1 * t1 - del * t3 * t3;
      if (t < 0.0) {
        double fmin;
        const double a = line_minimize(o, p, xit, fret, &fmin);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          xit[j] *= a;
          p[j] += xit[j];
        }
        fret = fmin;
#pragma unroll
        for (int j = 0; j < 3; j++) {
          // u[ibig] = u[2] with static indices (no scratch for a dynamically indexed array)
          if (ibig == 0) u[0][j] = u[2][j];
          if (ibig == 1) u[1][j] = u[2][j];
          u[2][j] = xit[j];
        }
      }
    }
  }
  return fret;
}

// The LM form: the first sweep of Powell's method (line minimisations along the three
// initial directions, which is where Powell's run picks its basin when the objective has
// several minima in theta), then Levenberg-Marquardt on (u, v, theta) with the
// Gauss-Newton matrix of the ZNCC residual until the step is below 1e-8 px and 1e-10 rad.
__device__ __forceinline__ bool solve3(const double* H, const double* g, double lambda, double* dx) {
  // (H + lambda diag(H)) dx = -g, symmetric 3x3 by Cramer's rule
  double a = H[0] * (1.0 + lambda), b = H[1], c = H[2];
  double d = H[3] * (1.0 + lambda), e = H[4], f = H[5] * (1.0 + lambda);
  const double A = d * f - e * e, B = c * e - b * f, C = b * e - c * d;
  const double det = a * A + b * B + c * C;
  if (!(fabs(det) > 0.0) || !isfinite(det)) return false;
  const double D = a * f - c * c, E = b * c - a * e, F = a * d - b * b;
  dx[0] = -(A * g[0] + B * g[1] + C * g[2]) / det;
  dx[1] = -(B * g[0] + D * g[1] + E * g[2]) / det;
  dx[2] = -(C * g[0] + E * g[1] + F * g[2]) / det;
  return isfinite(dx[0]) && isfinite(dx[1]) && isfinite(dx[2]);
}

#define LM_SWEEP_TOL 1.0e-4
#define LM_STEP_PX 1.0e-8
#define LM_STEP_RAD 1.0e-10

__device__ double lm(Obj& o, double* X, const double* dx, int max_iter) {
  if (max_iter <= 0) max_iter = 40;
  double fret = eval(o, X);
  if (fret != fret) return fret;
  for (int i = 0; i < 3; i++) {
    double u[3] = {0.0, 0.0, 0.0};
    u[i] = dx[i];
    double fmin;
    const double a = line_minimize(o, X, u, fret, &fmin, LM_SWEEP_TOL);
    X[i] += a * u[i];
    fret = fmin;
  }
  Derivs d;
  eval_derivs(o, X, d);
  double lambda = 1e-3;
  for (int it = 0; it < max_iter; it++) {
    double step[3];
    if (!solve3(d.H, d.grad, lambda, step)) break;
    const bool small = fabs(step[0]) < LM_STEP_PX && fabs(step[1]) < LM_STEP_PX && fabs(step[2]) < LM_STEP_RAD;
    double Xn[3] = {X[0] + step[0], X[1] + step[1], X[2] + step[2]};
    Derivs dn;
    eval_derivs(o, Xn, dn);
    if (dn.f <= d.f) {
      X[0] = Xn[0];
      X[1] = Xn[1];
      X[2] = Xn[2];
      d = dn;
      lambda = fmax(lambda * 0.1, 1e-12);
      if (small) break;
    } else {
      lambda *= 10.0;
      if (small || lambda > 1e10) break;
    }
  }
  return d.f;
}

// follow_pattern_subpixel_xyR, src/p-tracker.cpp:1090-1128, one wave per grain of the
// unmasked list.
__global__ __launch_bounds__(64) void k_subpixel(const uint16_t* __restrict__ I1, int dimx, int dimy, int64_t pitch,
                                                 ptg_grain* __restrict__ grains,
                                                 const int64_t* __restrict__ pbegin,
                                                 const int32_t* __restrict__ p0, const double* __restrict__ d0,
                                                 const double* __restrict__ sd0, const int32_t* __restrict__ list,
                                                 int nlist, ptg_subpix_params prm,
                                                 unsigned long long* __restrict__ evals,
                                                 const int32_t* __restrict__ prad, int tile_margin,
                                                 int32_t* __restrict__ evals_g) {
  const int li = blockIdx.x;
  if (li >= nlist) return;
  const int g = list[li];
  ptg_grain& gr = grains[g];
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  Obj o;
  load_obj(o, gr, I1, dimx, dimy, pitch, p0 + 2 * b, d0 + b, n, sd0[g], prm.subpix_displacement_max,
           prm.overflow_stiffness);
  // the tile: the grain's pattern extent plus tile_margin pixels, clamped to the frame edge
  extern __shared__ uint32_t tile[];
  {
    const int h = prad[g] + tile_margin;
    o.tx0 = (int)floor(o.xr) - h;
    o.ty0 = (int)floor(o.yr) - h;
    o.tw = o.th = 2 * h + 2;
    for (int t = threadIdx.x; t < o.tw * o.th; t += 64) {
      const int yy = t / o.tw, xx = t - yy * o.tw;
      const int X = clampi(o.tx0 + xx, 0, dimx - 1), Y = clampi(o.ty0 + yy, 0, dimy - 1);
      tile[t] = I1[(int64_t)Y * pitch + X];
    }
    o.tile = tile;
    __syncthreads();
  }
  double X[3] = {0.0, 0.0, 0.0};
  double fret;
  if (prm.method == PTG_SUBPIX_LM)
    fret = lm(o, X, prm.initial_direction, prm.max_iter);
  else
    fret = powell(o, X, prm.initial_direction, prm.subpix_tol, prm.max_iter);
  if ((threadIdx.x & 63) == 0) {
    gr.NCC_subpix = 1.0 - fret;
    gr.upix += X[0];
    gr.vpix += X[1];
    gr.rot_inc += X[2];
    gr.dx += X[0];
    gr.dy += X[1];
    gr.drot += X[2];
    if (o.clamped) gr.status |= PTG_STATUS_CLAMPED;
    atomicAdd(evals, (unsigned long long)o.evals);
    evals_g[g] = (int32_t)o.evals;
  }
}

hipError_t launch_subpixel(Context& c, const ptg_subpix_params& p) {
  if (c.n_unmasked == 0) return hipSuccess;
  const Image& im = c.img[c.icur];
  // LDS tile: pattern radius bound + margin for the subpixel excursion and the 4x4 stencil
  const int margin = 12;
  int rmax = 0;
  for (int g : c.h_unmasked_ref()) rmax = std::max(rmax, c.h_prad[g]);
  const int tw = 2 * (rmax + margin) + 2;
  const size_t lds = (size_t)tw * tw * sizeof(uint32_t);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_subpixel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  k_subpixel<<<c.n_unmasked, 64, lds, c.stream>>>(im.d, im.dimx, im.dimy, im.pitch, c.d_grains, c.d_pbegin, c.d_p0,
                                                   c.d_d0, c.d_sd0, c.d_list[0], c.n_unmasked, p,
                                                   reinterpret_cast<unsigned long long*>(c.d_evals), c.d_prad, margin,
                                                   c.d_evals_g);
  return hipGetLastError();
}

// NCC_to_minimize_xyR for a batch of (grain, X); with grad != nullptr also the
// analytic gradient. One wave per evaluation.
__global__ __launch_bounds__(64) void k_objective(const uint16_t* __restrict__ I1, int dimx, int dimy, int64_t pitch,
                                                  const ptg_grain* __restrict__ grains,
                                                  const int64_t* __restrict__ pbegin,
                                                  const int32_t* __restrict__ p0, const double* __restrict__ d0,
                                                  const double* __restrict__ sd0, int n_eval,
                                                  const int32_t* __restrict__ ig, const double* __restrict__ Xs,
                                                  double dmax, double stiff, double* __restrict__ f,
                                                  double* __restrict__ grad) {
  const int e = blockIdx.x;
  if (e >= n_eval) return;
  const int g = ig[e];
  const ptg_grain& gr = grains[g];
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  Obj o;
  load_obj(o, gr, I1, dimx, dimy, pitch, p0 + 2 * b, d0 + b, n, sd0[g], dmax, stiff);
  const double X[3] = {Xs[3 * e], Xs[3 * e + 1], Xs[3 * e + 2]};
  if (grad) {
    Derivs d;
    eval_derivs(o, X, d);
    if ((threadIdx.x & 63) == 0) {
      f[e] = d.f;
      grad[3 * e] = d.grad[0];
      grad[3 * e + 1] = d.grad[1];
      grad[3 * e + 2] = d.grad[2];
    }
  } else {
    const double v = eval(o, X);
    if ((threadIdx.x & 63) == 0) f[e] = v;
  }
}

hipError_t launch_objective_batch(Context& c, int32_t n, const int32_t* d_ig, const double* d_X, double dmax,
                                  double stiff, double* d_f, double* d_grad) {
  if (n == 0) return hipSuccess;
  const Image& im = c.img[c.icur];
  k_objective<<<n, 64, 0, c.stream>>>(im.d, im.dimx, im.dimy, im.pitch, c.d_grains, c.d_pbegin, c.d_p0, c.d_d0,
                                      c.d_sd0, n, d_ig, d_X, dmax, stiff, d_f, d_grad);
  return hipGetLastError();
}

}  // namespace ptg
