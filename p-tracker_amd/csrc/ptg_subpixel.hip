// ptg_subpixel.hip — the subpixel stage of libptgpu (gfx950): the ZNCC objective
// NCC_to_minimize_xyR (src/p-tracker.cpp:1018-1088), Powell's method with Brent line
// searches (the derivative-free optimizer of follow_pattern_subpixel_xyR, :1090-1128),
// and a Levenberg-Marquardt form driven by the analytic derivative of the ZNCC residual
// through the bicubic sampler.
//
// One workgroup of kSpWaves waves owns one grain. At its start the workgroup stages a
// tile of the current frame around the grain (16-bit pixels) in LDS; the pattern is read
// from global memory as one 32-bit word per pixel, written by the precomputation (x, y
// offsets as int8 and the reference image value I0, from which d0 = I0 - mean0 is
// re-formed bit-exactly). The optimizers are
// written in reverse-communication form: wave 0 holds their state in LDS, consumes the
// objective at the last requested point and publishes the next point (or "done"); every
// wave then sums its share of the pattern pixels for that point. So the objective is
// evaluated at one place in the code, the workgroup's registers hold only the pixel loop,
// and the optimizer's scalar control flow runs once per workgroup.
#include <algorithm>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

constexpr int kSpWaves = 4;
constexpr int kSpThreads = 64 * kSpWaves;
constexpr int kSpSums = 24;             // per-pixel sums of the value-and-derivatives pass
constexpr int kSpTileBytesCap = 65536;  // the frame tile is skipped above this size

enum { RQ_VALUE = 0, RQ_DERIVS = 1, RQ_DONE = 2 };

// The workgroup's dynamic LDS. The tile is addressed as an offset into it (not through a
// generic pointer) so that its loads are LDS loads.
extern __shared__ __align__(16) unsigned char sp_lds[];

// The point an evaluation is requested at, as the pixel loop needs it.
struct Request {
  double xpix1, ypix1;  // refcoord + dx + X[0], refcoord + dy + X[1]
  double c1, s1;        // cos, sin of drot + X[2]
  // weights of the Hessian's second-order sums, w_i = wa d0_i - wb (v_i - wm): the ZNCC
  // residual of the latest evaluation (wb = NCC / sig1, wm = mean of v there)
  double wa, wb, wm;
  int32_t kind;
  int32_t pad;
};

// Per-grain inputs of the objective, uniform across the workgroup.
struct Obj {
  const uint16_t* __restrict__ I1;
  int dimx, dimy;
  int64_t pitch;
  bool tile;           // LDS tile of I1 present: 16-bit pixels, row stride tw, at kSpOffTile
  int tx0, ty0, tw, th;
  const uint32_t* __restrict__ words;  // packed pattern words of this grain (used when nw == n)
  double m0;                            // the mean0 of d0 = I0 - m0
  int nw;
  const int32_t* __restrict__ p0;  // pattern0_rotated pairs of this grain (used when nw == 0)
  const double* __restrict__ d0;   // I0(p0) - mean0 of this grain
  int n;
  double xr, yr;  // refcoord + dx, refcoord + dy (:1031-1032 before adding X)
  double drot;
  double C0C0, sd0, shift;
  double dmax, stiff;
};


// One LDS pixel: a plain 16-bit load (kept from being merged into wider, unaligned
// loads of neighbouring taps).
__device__ __forceinline__ unsigned int tile_u16(int idx);

// Tile coordinates of the 4x4 stencil at (ix, iy) when it lies inside the tile.
__device__ __forceinline__ bool in_tile(const Obj& o, int ix, int iy, int& off) {
  const unsigned lx = (unsigned)ix - 1u - (unsigned)o.tx0, ly = (unsigned)iy - 1u - (unsigned)o.ty0;
  off = (int)(ly * (unsigned)o.tw + lx);
  return o.tile && lx < (unsigned)(o.tw - 3) && ly < (unsigned)(o.th - 3);
}

// The Catmull-Rom segment of src/interpol.cpp:42-53 written as weights of its four taps:
// cubic1(p0, p1, p2, p3, t) = sum_i w_i(t) p_i, with
//   w0 = t (-1/2 + t (1 - t/2)),  w1 = 1 + t^2 (-5/2 + 3t/2),
//   w2 = t (1/2 + t (2 - 3t/2)),  w3 = t^2 (-1/2 + t/2),
// and their derivatives in t. The same interpolant as the reference's, evaluated in another
// order (the subpixel stage is checked to a tolerance, not bit for bit).
__device__ __forceinline__ void cr_weights(double t, double* w) {
  const double t2 = t * t;
  w[0] = t * fma(t, fma(-0.5, t, 1.0), -0.5);
  w[1] = fma(t2, fma(1.5, t, -2.5), 1.0);
  w[2] = t * fma(t, fma(-1.5, t, 2.0), 0.5);
  w[3] = t2 * fma(0.5, t, -0.5);
}

__device__ __forceinline__ void cr_dweights(double t, double* d) {
  d[0] = fma(t, fma(-1.5, t, 2.0), -0.5);
  d[1] = t * fma(4.5, t, -5.0);
  d[2] = fma(t, fma(-4.5, t, 4.0), 0.5);
  d[3] = t * fma(1.5, t, -1.0);
}

// The 4x4 stencil of image_interpCubic::getValue (src/interpol.cpp:55-86) at (ix, iy):
// p[j][i] = I1(ix - 1 + i, iy - 1 + j), from the LDS tile when the stencil lies inside it
// (the tile holds the frame's values, edge-clamped as the global path clamps), else from
// global memory with clamped taps.
__device__ __forceinline__ void stencil(const Obj& o, int ix, int iy, unsigned int (*p)[4], bool& clamped) {
  const bool inside = ix >= 1 && iy >= 1 && ix < o.dimx - 2 && iy < o.dimy - 2;
  if (!inside) clamped = true;
  int off;
  if (in_tile(o, ix, iy, off)) {
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int i = 0; i < 4; i++) p[j][i] = tile_u16(off + j * o.tw + i);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int Y = clampi(iy - 1 + j, 0, o.dimy - 1);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int X = clampi(ix - 1 + i, 0, o.dimx - 1);
        p[j][i] = o.I1[(int64_t)Y * o.pitch + X];
      }
    }
  }
}

// The sampler's value at (x, y).
__device__ __forceinline__ double sample(const Obj& o, double x, double y, bool& clamped) {
  const int ix = (int)floor(x), iy = (int)floor(y);
  const double y01 = y - (double)iy, x01 = x - (double)ix;
  unsigned int p[4][4];
  stencil(o, ix, iy, p, clamped);
  double wx[4], wy[4];
  cr_weights(x01, wx);
  cr_weights(y01, wy);
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double r = fma(wx[3], (double)p[j][3],
                         fma(wx[2], (double)p[j][2], fma(wx[1], (double)p[j][1], wx[0] * (double)p[j][0])));
    v = fma(wy[j], r, v);
  }
  return v;
}

// The sampler's value and its x and y derivatives.
__device__ __forceinline__ void sample_grad(const Obj& o, double x, double y, double& v, double& gx, double& gy,
                                            bool& clamped) {
  const int ix = (int)floor(x), iy = (int)floor(y);
  const double y01 = y - (double)iy, x01 = x - (double)ix;
  unsigned int p[4][4];
  stencil(o, ix, iy, p, clamped);
  double wx[4], wy[4], dx[4], dy[4];
  cr_weights(x01, wx);
  cr_weights(y01, wy);
  cr_dweights(x01, dx);
  cr_dweights(y01, dy);
  v = 0.0;
  gx = 0.0;
  gy = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double q0 = (double)p[j][0], q1 = (double)p[j][1], q2 = (double)p[j][2], q3 = (double)p[j][3];
    const double r = fma(wx[3], q3, fma(wx[2], q2, fma(wx[1], q1, wx[0] * q0)));
    const double sd = fma(dx[3], q3, fma(dx[2], q2, fma(dx[1], q1, dx[0] * q0)));
    v = fma(wy[j], r, v);
    gx = fma(wy[j], sd, gx);
    gy = fma(dy[j], r, gy);
  }
}

// Weight, derivative and second derivative of Catmull-Rom tap j at t.
__device__ __forceinline__ void cr_tap(int j, double t, double& w, double& d, double& e) {
  const double t2 = t * t;
  if (j == 0) {
    w = t * fma(t, fma(-0.5, t, 1.0), -0.5);
    d = fma(t, fma(-1.5, t, 2.0), -0.5);
    e = fma(-3.0, t, 2.0);
  } else if (j == 1) {
    w = fma(t2, fma(1.5, t, -2.5), 1.0);
    d = t * fma(4.5, t, -5.0);
    e = fma(9.0, t, -5.0);
  } else if (j == 2) {
    w = t * fma(t, fma(-1.5, t, 2.0), 0.5);
    d = fma(t, fma(-4.5, t, 4.0), 0.5);
    e = fma(-9.0, t, 4.0);
  } else {
    w = t2 * fma(0.5, t, -0.5);
    d = t * fma(1.5, t, -1.0);
    e = fma(3.0, t, -1.0);
  }
}

// The sampler's value, gradient and Hessian (xx, xy, yy) at (x, y); the y weights are
// formed row by row to keep fewer values live.
__device__ __forceinline__ void sample_hess(const Obj& o, double x, double y, double& v, double& gx, double& gy,
                                            double& gxx, double& gxy, double& gyy, bool& clamped) {
  const int ix = (int)floor(x), iy = (int)floor(y);
  const double y01 = y - (double)iy, x01 = x - (double)ix;
  unsigned int p[4][4];
  stencil(o, ix, iy, p, clamped);
  double wx[4], dx[4], ex[4];
#pragma unroll
  for (int i = 0; i < 4; i++) cr_tap(i, x01, wx[i], dx[i], ex[i]);
  v = gx = gy = gxx = gxy = gyy = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double q0 = (double)p[j][0], q1 = (double)p[j][1], q2 = (double)p[j][2], q3 = (double)p[j][3];
    const double r = fma(wx[3], q3, fma(wx[2], q2, fma(wx[1], q1, wx[0] * q0)));
    const double sd = fma(dx[3], q3, fma(dx[2], q2, fma(dx[1], q1, dx[0] * q0)));
    const double se = fma(ex[3], q3, fma(ex[2], q2, fma(ex[1], q1, ex[0] * q0)));
    double wy, dy, ey;
    cr_tap(j, y01, wy, dy, ey);
    v = fma(wy, r, v);
    gx = fma(wy, sd, gx);
    gy = fma(dy, r, gy);
    gxx = fma(wy, se, gxx);
    gxy = fma(dy, sd, gxy);
    gyy = fma(ey, r, gyy);
  }
}

// Calls f(x, y, d0) for this thread's share of the grain's pattern pixels.
template <class F>
__device__ __forceinline__ void for_pixels(const Obj& o, F& f) {
  for (int i = threadIdx.x; i < o.nw; i += kSpThreads) {
    const uint32_t w = o.words[i];
    f((int)(int8_t)(w & 0xffu), (int)(int8_t)((w >> 8) & 0xffu), (double)(w >> 16) - o.m0);
  }
  for (int i = o.nw + (int)threadIdx.x; i < o.n; i += kSpThreads) f(o.p0[2 * i], o.p0[2 * i + 1], o.d0[i]);
}

// Per-pixel sums at a request. Value pass: sum g, g^2, d0 g with g = I1(p(X)) - mean0.
// Derivative pass adds sum dG, g dG, d0 dG and dG dG^T for dG = dg/d(u, v, theta).
__device__ __forceinline__ void pixel_sums_value(const Obj& o, const Request& rq, double* S, bool& cl) {
  S[0] = S[1] = S[2] = 0.0;
  const double xpix1 = rq.xpix1, ypix1 = rq.ypix1, c1 = rq.c1, s1 = rq.s1;
  auto px = [&](int xi, int yi, double dd) {
    const double x = (double)xi, y = (double)yi;
    const double xs = xpix1 + c1 * x + s1 * y;
    const double ys = ypix1 - s1 * x + c1 * y;
    const double v = sample(o, xs, ys, cl) - o.shift;
    S[0] += v;
    S[1] = fma(v, v, S[1]);
    S[2] = fma(dd, v, S[2]);
  };
  for_pixels(o, px);
}

// Derivative pass: with g(X) = I1(p(X)) - mean0 and its derivatives G = dg/d(u, v, theta)
// and G2 = d2g (xx, xy, xt, yy, yt, tt), the sums g, g^2, d0 g, G, g G, d0 G, G G^T and
// w G2 (w the residual weight of the request).
__device__ __forceinline__ void pixel_sums_derivs(const Obj& o, const Request& rq, double* S, bool& cl) {
#pragma unroll
  for (int k = 0; k < kSpSums; k++) S[k] = 0.0;
  const double xpix1 = rq.xpix1, ypix1 = rq.ypix1, c1 = rq.c1, s1 = rq.s1;
  const double wa = rq.wa, wb = rq.wb, wm = rq.wm;
  auto px = [&](int xi, int yi, double dd) {
    const double x = (double)xi, y = (double)yi;
    const double xs = xpix1 + c1 * x + s1 * y;
    const double ys = ypix1 - s1 * x + c1 * y;
    double v, gx, gy, gxx, gxy, gyy;
    sample_hess(o, xs, ys, v, gx, gy, gxx, gxy, gyy, cl);
    v -= o.shift;
    // d(xs, ys)/dtheta and their second derivatives (xtt = yt, ytt = -xt)
    const double xt = -s1 * x + c1 * y, yt = -c1 * x - s1 * y;
    const double gt = gx * xt + gy * yt;
    const double G[3] = {gx, gy, gt};
    const double gxt = fma(gxx, xt, gxy * yt), gyt = fma(gxy, xt, gyy * yt);
    const double gtt = fma(gxt, xt, fma(gyt, yt, fma(gx, yt, -gy * xt)));
    const double w = fma(wa, dd, -wb * (v - wm));
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
    S[18] = fma(w, gxx, S[18]);
    S[19] = fma(w, gxy, S[19]);
    S[20] = fma(w, gxt, S[20]);
    S[21] = fma(w, gyy, S[21]);
    S[22] = fma(w, gyt, S[22]);
    S[23] = fma(w, gtt, S[23]);
  };
  for_pixels(o, px);
}

// Workgroup sums in two halves around a barrier: each wave's butterfly sum goes to
// red[wave * NS + k]; after the barrier one thread adds the waves' partials in wave order.
template <int NS>
__device__ __forceinline__ void wave_partials(double* S, double* red) {
#pragma unroll
  for (int k = 0; k < NS; k++) S[k] = wave_sum(S[k]);
  if ((threadIdx.x & 63) == 0) {
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NS; k++) red[wv * NS + k] = S[k];
  }
}

template <int NS>
__device__ __forceinline__ void reduce_partials(const double* red, double* S) {
#pragma unroll
  for (int k = 0; k < NS; k++) {
    double t = red[k];
#pragma unroll
    for (int w = 1; w < kSpWaves; w++) t += red[w * NS + k];
    S[k] = t;
  }
}

// The grain's constants of the objective.
struct ObjConst {
  double N, sd0, C0C0, dmax, stiff;
  double xr, yr, drot;
  double wa;  // 1 / sqrt(C0C0), the residual weight of d0
};

__device__ __forceinline__ ObjConst obj_const(const Obj& o) {
  ObjConst k;
  k.N = (double)o.n;
  k.sd0 = o.sd0;
  k.C0C0 = o.C0C0;
  k.dmax = o.dmax;
  k.stiff = o.stiff;
  k.xr = o.xr;
  k.yr = o.yr;
  k.drot = o.drot;
  k.wa = 1.0 / sqrt(o.C0C0);
  return k;
}

__device__ __forceinline__ double weight(const ObjConst& o, const double* X) {
  double W = 1.0;
  const double disp = sqrt(X[0] * X[0] + X[1] * X[1]);
  if (disp > o.dmax) W = o.stiff * (disp - o.dmax) + 1.0;
  return W;
}

// rc (optional): the last angle and its cosine and sine, reused while the angle is
// unchanged (the line searches along u and v keep it).
__device__ __forceinline__ void make_request(const ObjConst& o, const double* X, int kind, double wb, double wm,
                                             Request* rq, double* rc = nullptr) {
  const double rot = o.drot + X[2];
  rq->xpix1 = o.xr + X[0];
  rq->ypix1 = o.yr + X[1];
  if (rc && rot == rc[0]) {
    rq->c1 = rc[1];
    rq->s1 = rc[2];
  } else {
    double sv, cv;
    sincos(rot, &sv, &cv);  // (one argument reduction for both)
    rq->c1 = cv;
    rq->s1 = sv;
    if (rc) {
      rc[0] = rot;
      rc[1] = rq->c1;
      rc[2] = rq->s1;
    }
  }
  rq->wa = o.wa;
  rq->wb = wb;
  rq->wm = wm;
  rq->kind = kind;
}

// A sampled window whose one-pass variance C1C1 = Sgg - Sg^2 / N is zero up to the
// rounding of Sgg: the samples are all equal, where the reference's two passes give
// C1C1 = C0C1 = 0 exactly and NCC = 0 / 0, not a number (with integer samples, as at an
// integer position, any other window has C1C1 >= 1/2, far above this bound). The optimizers
// stop at such a value as the reference's do.
__device__ __forceinline__ bool flat_window(double C1C1, double Sgg) { return !(C1C1 > 64.0 * 2.220446049250313e-16 * Sgg); }

// (1 - NCC) * W at X from the value sums. The reference's two passes (mean, then
// centred products) are replaced by one pass over shifted sums:
// C1C1 = Sgg - Sg^2/N, C0C1 = Sdg - Sg*Sd0/N.
__device__ __forceinline__ double finish_value(const ObjConst& o, const double* S, const double* X, double* wb,
                                               double* wm) {
  const double N = o.N;
  const double C1C1 = S[1] - S[0] * S[0] / N;
  const double C0C1 = S[2] - S[0] * o.sd0 / N;
  const double NCC = flat_window(C1C1, S[1]) ? __builtin_nan("") : C0C1 / sqrt(o.C0C0 * C1C1);
  *wb = NCC / sqrt(C1C1);
  *wm = S[0] / N;
  return (1.0 - NCC) * weight(o, X);
}

// Value, gradient and Hessian of f = (1 - NCC) W at X. With u = g - mean(g), sig1 = |u|,
// fhat = d0 / sig0, beta_a = u.G_a / sig1, hf_a = fhat.G_a and D_ab = sum (G_a - mean G_a)
// (G_b - mean G_b):
//   dNCC_a = (hf_a - NCC beta_a) / sig1
//   d2NCC_ab = T_ab / sig1 - (hf_a beta_b + hf_b beta_a + NCC (D_ab - 3 beta_a beta_b)) / sig1^2
// where T_ab = sum (fhat - NCC u / sig1) G2_ab, taken with the residual weights of the
// request (those of the latest evaluation, so T is exact only at a fixed point; the gradient
// is exact everywhere, so the optimizer's fixed points are those of f).
struct Derivs {
  double f, NCC;
  double grad[3];
  double H[6];  // xx, xy, xt, yy, yt, tt
};

__device__ __forceinline__ void finish_derivs(const ObjConst& o, const double* S, const double* X, Derivs& d,
                                              double* wb, double* wm) {
  const double N = o.N;
  const double m1 = S[0] / N;
  const double C1C1 = S[1] - S[0] * m1;
  const double C0C1 = S[2] - m1 * o.sd0;
  const double sig1 = sqrt(C1C1), sig0 = sqrt(o.C0C0);
  const double NCC = flat_window(C1C1, S[1]) ? __builtin_nan("") : C0C1 / (sig0 * sig1);
  double mG[3], beta[3], hf[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    mG[k] = S[3 + k] / N;
    beta[k] = (S[6 + k] - m1 * S[3 + k]) / sig1;  // u . G_k / sig1
    hf[k] = (S[9 + k] - mG[k] * o.sd0) / sig0;    // fhat . G_k
  }
  double dN[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dN[k] = (hf[k] - beta[k] * NCC) / sig1;
  const double W = weight(o, X);
  double dW[3] = {0.0, 0.0, 0.0};
  double d2W[3] = {0.0, 0.0, 0.0};  // xx, xy, yy of the penalty weight (zero in theta)
  const double disp = sqrt(X[0] * X[0] + X[1] * X[1]);
  if (disp > o.dmax && disp > 0.0) {
    dW[0] = o.stiff * X[0] / disp;
    dW[1] = o.stiff * X[1] / disp;
    const double i3 = o.stiff / (disp * disp * disp);
    d2W[0] = i3 * X[1] * X[1];
    d2W[1] = -i3 * X[0] * X[1];
    d2W[2] = i3 * X[0] * X[0];
  }
  d.NCC = NCC;
  d.f = (1.0 - NCC) * W;
#pragma unroll
  for (int k = 0; k < 3; k++) d.grad[k] = -dN[k] * W + (1.0 - NCC) * dW[k];
  const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  const int w2[6] = {0, 1, -1, 2, -1, -1};  // index into d2W, -1: zero
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int a = ij[q][0], b = ij[q][1];
    const double D = S[12 + q] - N * mG[a] * mG[b];
    const double d2N = S[18 + q] / sig1 - (hf[a] * beta[b] + hf[b] * beta[a] + NCC * (D - 3.0 * beta[a] * beta[b])) / C1C1;
    double h = -d2N * W - (dN[a] * dW[b] + dN[b] * dW[a]);
    if (w2[q] >= 0) h += (1.0 - NCC) * d2W[w2[q]];
    d.H[q] = h;
  }
  *wb = NCC / sig1;
  *wm = m1;
}

// ---------------------------------------------------------------------------------
// The optimizers, in reverse communication. Their state lives in the workgroup's LDS
// (SpState) and only thread 0 runs them: each step consumes the objective at the point
// last requested (st.Xq) and writes the next point into st.Xq, or finishes. The steps
// are out-of-line functions without pointer arguments, so that their code and registers
// stay apart from the pixel loop's.
//
// Brent line minimisation with golden-section bracketing (Brent 1973): line_start
// requests f(p + xi); line_feed consumes the value at the last requested point and either
// requests the next one or finishes with (x, fx). The sequence of points is that of a
// plain loop with the same constants.
#define PW_GOLD 1.618033988749895
#define PW_CGOLD 0.3819660112501051
#define PW_LINE_TOL 3.0e-8
#define PW_ZEPS 1.0e-12
#define PW_TINY 1.0e-25
#define PW_BRACKET_MAX 60
#define PW_BRENT_MAX 100

struct Line {
  double p[3], xi[3];
  double a, b, c, fa, fb, fc;
  double lo, hi, x, w, v, fx, fw, fv, d, e, u;
  double tol;
  int32_t phase, it;
};

// Powell's method (Powell 1964) over (u, v, theta): line minimisations along a set of
// directions, replacing the direction of largest decrease by the average direction when
// the extrapolation test allows it. The same control flow and constants as the oracle's
// restatement.
enum { PS_INIT = 0, PS_LINE, PS_PTT, PS_XLINE };

struct Powell {
  double u[3][3];
  double p[3], pt[3], xit[3];
  double fret, fp, del, fstart, ftol;
  int32_t ibig, i, iter, max_iter, state, pad;
};

// The LM form: the first sweep of Powell's method (line minimisations along the three
// initial directions, which is where Powell's run picks its basin when the objective has
// several minima in theta), then damped Newton (Levenberg-Marquardt on the objective's
// Hessian) on (u, v, theta). It stops at the current point, without evaluating the next,
// once the undamped Newton step from it (the damped step times 1 + lambda, a bound for the
// diagonal damping used here) is below 1e-6 px and 1e-8 rad: the point is then that close
// to the minimum (Newton converges quadratically there), far inside the stage's 1e-3 px
// tolerance, and its objective is already known. A step below 1e-8 px and 1e-10 rad (by
// damping) also ends the run after its evaluation. The sweep's line searches only have to
// land in the same basin as Powell's (Brent tolerance 3e-8), but that needs more than a
// coarse bracket: on 2000 grains of a 6000x4000 frame pair 1e-3 and 1e-4 reach the
// oracle's minimum everywhere while 1e-2 sends one grain to another minimum.
#define LM_SWEEP_TOL 1.0e-3
#define LM_DONE_PX 1.0e-6
#define LM_DONE_RAD 1.0e-8
#define LM_STEP_PX 1.0e-8
#define LM_STEP_RAD 1.0e-10

enum { LS_INIT = 0, LS_SWEEP, LS_D0, LS_DN };

struct LevMar {
  double X[3], Xn[3], dx[3];
  double fret, lambda;
  int32_t i, it, max_iter, state, small, pad;
  Derivs d;
};

// The workgroup's shared state, at the start of its LDS.
struct SpState {
  Request rq;     // the published request
  ObjConst k;     // the grain's constants
  double Xq[3];   // the point of the request
  int32_t lm, nevals;
  Line L;         // the line search in progress
  Derivs dq;      // derivatives at Xq, for derivative requests
  double hwb, hwm;  // residual weights (NCC / sig1, mean of v) of the latest evaluation
  double rc[3];     // the last requested angle, its cosine and sine
  double sums[kSpSums];  // the workgroup's pixel sums at Xq
  union {
    Powell P;
    LevMar M;
  } opt;
};

__device__ __forceinline__ SpState& sp_state() { return *reinterpret_cast<SpState*>(sp_lds); }

__device__ __forceinline__ void line_point(SpState& st, double t) {
#pragma unroll
  for (int j = 0; j < 3; j++) st.Xq[j] = st.L.p[j] + t * st.L.xi[j];
}

__device__ __forceinline__ void line_start(SpState& st, double p0, double p1, double p2, double x0, double x1,
                                           double x2, double f0, double tol) {
  Line& L = st.L;
  L.p[0] = p0; L.p[1] = p1; L.p[2] = p2;
  L.xi[0] = x0; L.xi[1] = x1; L.xi[2] = x2;
  L.a = 0.0;
  L.fa = f0;
  L.b = 1.0;
  L.tol = tol;
  L.phase = 0;
  line_point(st, L.b);
}

// Consumes f at the last requested point; returns true when the line search has ended
// (st.L.x, st.L.fx), else requests the next point.
__device__ __noinline__ bool line_feed(double f) {
  SpState& st = sp_state();
  Line& L = st.L;
  if (L.phase == 0) {  // f(b = 1)
    L.fb = f;
    if (L.fb > L.fa) {
      double t = L.a; L.a = L.b; L.b = t;
      t = L.fa; L.fa = L.fb; L.fb = t;
    }
    L.c = L.b + PW_GOLD * (L.b - L.a);
    L.phase = 1;
    L.it = 0;
    line_point(st, L.c);
    return false;
  }
  if (L.phase == 1) {  // f(c) of the bracketing walk
    L.fc = f;
    if (L.it < PW_BRACKET_MAX && L.fc < L.fb) {
      L.a = L.b; L.fa = L.fb;
      L.b = L.c; L.fb = L.fc;
      L.c = L.b + PW_GOLD * (L.b - L.a);
      L.it++;
      line_point(st, L.c);
      return false;
    }
    L.lo = L.a < L.c ? L.a : L.c;
    L.hi = L.a < L.c ? L.c : L.a;
    L.x = L.w = L.v = L.b;
    L.fx = L.fw = L.fv = L.fb;
    L.d = L.e = 0.0;
    L.it = 0;
    L.phase = 2;
  } else {  // f(u) of a Brent step
    const double u = L.u, fu = f;
    if (fu <= L.fx) {
      if (u < L.x) L.hi = L.x; else L.lo = L.x;
      L.v = L.w; L.fv = L.fw;
      L.w = L.x; L.fw = L.fx;
      L.x = u; L.fx = fu;
    } else {
      if (u < L.x) L.lo = u; else L.hi = u;
      if (fu <= L.fw || L.w == L.x) {
        L.v = L.w; L.fv = L.fw;
        L.w = u; L.fw = fu;
      } else if (fu <= L.fv || L.v == L.x || L.v == L.w) {
        L.v = u; L.fv = fu;
      }
    }
    L.it++;
  }
  // the head of Brent's loop: convergence test, then a parabolic or golden step
  if (L.it >= PW_BRENT_MAX) return true;
  const double x = L.x, lo = L.lo, hi = L.hi;
  const double m = 0.5 * (lo + hi);
  const double tol1 = L.tol * fabs(x) + PW_ZEPS;
  const double t2 = 2.0 * tol1;
  if (fabs(x - m) <= t2 - 0.5 * (hi - lo)) return true;
  bool golden = true;
  double d = L.d, e = L.e;
  if (fabs(e) > tol1) {
    double r = (x - L.w) * (L.fx - L.fv);
    double q = (x - L.v) * (L.fx - L.fw);
    double pp = (x - L.v) * q - (x - L.w) * r;
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
  L.d = d;
  L.e = e;
  L.u = x + (fabs(d) >= tol1 ? d : (d > 0.0 ? tol1 : -tol1));
  line_point(st, L.u);
  return false;
}

__device__ __forceinline__ void pw_begin_line(SpState& st) {
  Powell& P = st.opt.P;
  P.fstart = P.fret;
  line_start(st, P.p[0], P.p[1], P.p[2], P.u[P.i][0], P.u[P.i][1], P.u[P.i][2], P.fret, PW_LINE_TOL);
  P.state = PS_LINE;
}

__device__ __forceinline__ void pw_begin_iteration(SpState& st) {
  Powell& P = st.opt.P;
  P.fp = P.fret;
  P.del = 0.0;
  P.ibig = 0;
  P.i = 0;
  pw_begin_line(st);
}

// Returns false when finished (P.p, P.fret hold the result), else a value request is
// pending at st.Xq.
__device__ __noinline__ bool pw_feed(double f) {
  SpState& st = sp_state();
  Powell& P = st.opt.P;
  switch (P.state) {
    case PS_INIT:
      P.fret = f;
      if (f != f) return false;  // undefined ZNCC (flat pattern): nothing to minimise
#pragma unroll
      for (int j = 0; j < 3; j++) P.pt[j] = P.p[j];
      P.iter = 0;
      pw_begin_iteration(st);
      return true;
    case PS_LINE: {
      if (!line_feed(f)) return true;
      const double a = st.L.x;
      const int i = P.i;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        P.u[i][j] *= a;
        P.p[j] += P.u[i][j];
      }
      P.fret = st.L.fx;
      if (P.fstart - P.fret > P.del) {
        P.del = P.fstart - P.fret;
        P.ibig = i;
      }
      P.i = i + 1;
      if (P.i < 3) {
        pw_begin_line(st);
        return true;
      }
      if (2.0 * (P.fp - P.fret) <= P.ftol * (fabs(P.fp) + fabs(P.fret)) + PW_TINY) return false;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        st.Xq[j] = 2.0 * P.p[j] - P.pt[j];
        P.xit[j] = P.p[j] - P.pt[j];
        P.pt[j] = P.p[j];
      }
      P.state = PS_PTT;
      return true;
    }
    case PS_PTT: {
      const double fptt = f;
      if (fptt < P.fp) {
        const double t1 = P.fp - P.fret - P.del, t3 = P.fp - fptt;
        const double t = 2.0 * (P.fp - 2.0 * P.fret + fptt) * t1 * t1 - P.del * t3 * t3;
        if (t < 0.0) {
          line_start(st, P.p[0], P.p[1], P.p[2], P.xit[0], P.xit[1], P.xit[2], P.fret, PW_LINE_TOL);
          P.state = PS_XLINE;
          return true;
        }
      }
      break;
    }
    default: {  // PS_XLINE
      if (!line_feed(f)) return true;
      const double a = st.L.x;
      const int ib = P.ibig;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        P.xit[j] *= a;
        P.p[j] += P.xit[j];
        P.u[ib][j] = P.u[2][j];
        P.u[2][j] = P.xit[j];
      }
      P.fret = st.L.fx;
      break;
    }
  }
  P.iter++;
  if (P.iter >= P.max_iter) return false;
  pw_begin_iteration(st);
  return true;
}

__device__ __forceinline__ bool solve3(const double* H, const double* g, double lambda, double* dx) {
  // (H + lambda D) dx = -g with D = |diag(H)| (floored), symmetric 3x3 by Cramer's rule
  const double dmax = fmax(fabs(H[0]), fmax(fabs(H[3]), fabs(H[5])));
  const double floor_d = 1e-12 * dmax + 1e-300;
  double a = H[0] + lambda * fmax(fabs(H[0]), floor_d), b = H[1], c = H[2];
  double d = H[3] + lambda * fmax(fabs(H[3]), floor_d), e = H[4], f = H[5] + lambda * fmax(fabs(H[5]), floor_d);
  const double A = d * f - e * e, B = c * e - b * f, C = b * e - c * d;
  const double det = a * A + b * B + c * C;
  if (!(fabs(det) > 0.0) || !isfinite(det)) return false;
  const double D = a * f - c * c, E = b * c - a * e, F = a * d - b * b;
  dx[0] = -(A * g[0] + B * g[1] + C * g[2]) / det;
  dx[1] = -(B * g[0] + D * g[1] + E * g[2]) / det;
  dx[2] = -(C * g[0] + E * g[1] + F * g[2]) / det;
  return isfinite(dx[0]) && isfinite(dx[1]) && isfinite(dx[2]);
}

// The next Levenberg-Marquardt trial point (a derivative request), or RQ_DONE.
__device__ __forceinline__ int lm_step(SpState& st) {
  LevMar& M = st.opt.M;
  if (M.it >= M.max_iter) return RQ_DONE;
  double H[6], g[3], step[3];
#pragma unroll
  for (int q = 0; q < 6; q++) H[q] = M.d.H[q];
#pragma unroll
  for (int j = 0; j < 3; j++) g[j] = M.d.grad[j];
  if (!solve3(H, g, M.lambda, step)) return RQ_DONE;
  const double sc = 1.0 + M.lambda;
  if (fabs(step[0]) * sc < LM_DONE_PX && fabs(step[1]) * sc < LM_DONE_PX && fabs(step[2]) * sc < LM_DONE_RAD)
    return RQ_DONE;
  M.small = fabs(step[0]) < LM_STEP_PX && fabs(step[1]) < LM_STEP_PX && fabs(step[2]) < LM_STEP_RAD;
#pragma unroll
  for (int j = 0; j < 3; j++) st.Xq[j] = M.Xn[j] = M.X[j] + step[j];
  M.state = LS_DN;
  return RQ_DERIVS;
}

__device__ __forceinline__ int lm_begin_sweep_line(SpState& st) {
  LevMar& M = st.opt.M;
  const int i = M.i;
  line_start(st, M.X[0], M.X[1], M.X[2], i == 0 ? M.dx[0] : 0.0, i == 1 ? M.dx[1] : 0.0, i == 2 ? M.dx[2] : 0.0,
             M.fret, LM_SWEEP_TOL);
  M.state = LS_SWEEP;
  return RQ_VALUE;
}

// Consumes the value f (value requests) or st.dq (derivative requests) at the last
// request; returns the kind of the next request, or RQ_DONE (M.X holds the result, its
// objective in M.d.f).
__device__ __noinline__ int lm_feed(double f) {
  SpState& st = sp_state();
  LevMar& M = st.opt.M;
  switch (M.state) {
    case LS_INIT:
      M.fret = f;
      M.d.f = f;
      if (f != f) return RQ_DONE;
      M.i = 0;
      return lm_begin_sweep_line(st);
    case LS_SWEEP: {
      if (!line_feed(f)) return RQ_VALUE;
      const int i = M.i;
      M.X[i] += st.L.x * M.dx[i];
      M.fret = st.L.fx;
      M.i = i + 1;
      if (M.i < 3) return lm_begin_sweep_line(st);
#pragma unroll
      for (int j = 0; j < 3; j++) st.Xq[j] = M.X[j];
      M.state = LS_D0;
      return RQ_DERIVS;
    }
    case LS_D0:
      M.d = st.dq;
      M.lambda = 1e-3;
      M.it = 0;
      return lm_step(st);
    default:  // LS_DN
      if (st.dq.f <= M.d.f) {
#pragma unroll
        for (int j = 0; j < 3; j++) M.X[j] = M.Xn[j];
        M.d = st.dq;
        M.lambda = fmax(M.lambda * 0.1, 1e-12);
        if (M.small) return RQ_DONE;
      } else {
        M.lambda *= 10.0;
        if (M.small || M.lambda > 1e10) return RQ_DONE;
      }
      M.it++;
      return lm_step(st);
  }
}

// ---------------------------------------------------------------------------------
// Workgroup staging of one grain: the frame tile in LDS.
struct SpLayout {
  int32_t tile_elems;  // 16-bit elements reserved for the tile (0: no tile)
  int32_t margin;      // tile pixels beyond the pattern's extent
};

// LDS: [SpState][kSpWaves x kSpSums doubles][tile]
constexpr size_t sp_align(size_t b) { return (b + 15) & ~(size_t)15; }
constexpr size_t kSpOffRed = sp_align(sizeof(SpState));
constexpr size_t kSpOffTile = kSpOffRed + sp_align(sizeof(double) * kSpWaves * kSpSums);

// LDS-qualified: a volatile access through a generic pointer would become a flat load.
typedef __attribute__((address_space(3))) const volatile uint16_t lds_u16;

__device__ __forceinline__ unsigned int tile_u16(int idx) { return *(lds_u16*)(sp_lds + kSpOffTile + 2 * idx); }

__device__ __forceinline__ void stage_grain(Obj& o, const ptg_grain& gr, int prad, const SpLayout& lay) {
  uint16_t* tile = reinterpret_cast<uint16_t*>(sp_lds + kSpOffTile);
  // the tile: the grain's pattern extent plus margin pixels, clamped to the frame edge
  const int h = prad + lay.margin;
  const int tw = 2 * h + 2;
  o.tile = false;
  o.tx0 = (int)floor(o.xr) - h;
  o.ty0 = (int)floor(o.yr) - h;
  o.tw = o.th = tw;
  if (lay.tile_elems > 0 && tw * tw <= lay.tile_elems && o.xr == o.xr && o.yr == o.yr) {
    // kB frame loads in flight per thread, then their LDS stores
    constexpr int kB = 8;
    const int nt = tw * tw;
    for (int t0 = threadIdx.x; t0 < nt; t0 += kB * kSpThreads) {
      unsigned int v[kB];
#pragma unroll
      for (int j = 0; j < kB; j++) {
        const int t = t0 + j * kSpThreads;
        v[j] = 0u;
        if (t < nt) {
          const int yy = t / tw, xx = t - yy * tw;
          const int X = clampi(o.tx0 + xx, 0, o.dimx - 1), Y = clampi(o.ty0 + yy, 0, o.dimy - 1);
          v[j] = frame_px(o.I1, (int64_t)Y * o.pitch + X);
        }
      }
#pragma unroll
      for (int j = 0; j < kB; j++)
        if (t0 + j * kSpThreads < nt) tile[t0 + j * kSpThreads] = (uint16_t)v[j];
    }
    o.tile = true;
  }
  // the packed words hold the rotated offsets as int8: exact up to a pattern radius of 127
  o.nw = prad <= 127 ? o.n : 0;
  __syncthreads();  // the tile is written
}

__device__ __forceinline__ void load_obj(Obj& o, const ptg_grain& g, const uint16_t* I1, int dimx, int dimy,
                                         int64_t pitch, const uint32_t* words, double m0, const int32_t* p0,
                                         const double* d0, int n, double sd0, double dmax, double stiff) {
  o.I1 = I1;
  o.dimx = dimx;
  o.dimy = dimy;
  o.pitch = pitch;
  o.words = words;
  o.m0 = m0;
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
}

// One optimizer step, run by thread 0 once the workgroup's sums for the request in
// st.rq are in st.sums: finishes the objective (or its derivatives) at st.Xq, feeds
// it to the optimizer and publishes the next request or RQ_DONE.
__device__ __noinline__ void sp_step() {
  SpState& st = sp_state();
  const double X[3] = {st.Xq[0], st.Xq[1], st.Xq[2]};
  st.nevals++;
  int next;
  if (st.rq.kind == RQ_DERIVS) {
    finish_derivs(st.k, st.sums, X, st.dq, &st.hwb, &st.hwm);
    next = lm_feed(st.dq.f);
  } else {
    const double f = finish_value(st.k, st.sums, X, &st.hwb, &st.hwm);
    next = st.lm ? lm_feed(f) : (pw_feed(f) ? RQ_VALUE : RQ_DONE);
  }
  if (next == RQ_DONE)
    st.rq.kind = RQ_DONE;
  else
    make_request(st.k, st.Xq, next, st.hwb, st.hwm, &st.rq, st.rc);
}

// follow_pattern_subpixel_xyR, src/p-tracker.cpp:1090-1128: one workgroup per grain of the
// unmasked list, taken in the order of `list` (largest patterns first).
__global__ __launch_bounds__(kSpThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_subpixel(const FrameDesc* __restrict__ frames,
                                                         ptg_grain* __restrict__ grains,
                                                         const int64_t* __restrict__ pbegin,
                                                         const uint32_t* __restrict__ words,
                                                         const double* __restrict__ m0,
                                                         const int32_t* __restrict__ p0, const double* __restrict__ d0,
                                                         const double* __restrict__ sd0,
                                                         const int32_t* __restrict__ list, int nlist,
                                                         ptg_subpix_params prm, unsigned long long* __restrict__ evals,
                                                         const int32_t* __restrict__ prad, SpLayout lay,
                                                         int32_t* __restrict__ evals_g, int sel, double thr) {
  const int li = blockIdx.x;
  if (li >= nlist) return;
  double* red = reinterpret_cast<double*>(sp_lds + kSpOffRed);
  const int g = list[li];
  ptg_grain& gr = grains[g];
  // sel 1: only the grains whose main-stage NCC reached thr (those no rescue stage will move);
  // 2: only the others (the same value in every thread of the workgroup)
  if ((sel == 1 && !(gr.NCC >= thr)) || (sel == 2 && gr.NCC >= thr)) return;
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  Obj o;
  const FrameDesc fr = *frames;
  load_obj(o, gr, fr.I1, fr.dimx, fr.dimy, fr.pitch1, words + b, m0[g], p0 + 2 * b, d0 + b, n, sd0[g],
           prm.subpix_displacement_max, prm.overflow_stiffness);
  stage_grain(o, gr, prad[g], lay);
  SpState& st = sp_state();
  if (threadIdx.x == 0) {
    st.k = obj_const(o);
    st.lm = prm.method == PTG_SUBPIX_LM;
    st.nevals = 0;
    st.Xq[0] = st.Xq[1] = st.Xq[2] = 0.0;
    if (st.lm) {
      LevMar& M = st.opt.M;
      M.X[0] = M.X[1] = M.X[2] = 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++) M.dx[j] = prm.initial_direction[j];
      M.max_iter = prm.max_iter > 0 ? prm.max_iter : 40;
      M.state = LS_INIT;
    } else {
      Powell& P = st.opt.P;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) P.u[i][j] = i == j ? prm.initial_direction[j] : 0.0;
      P.p[0] = P.p[1] = P.p[2] = 0.0;
      P.ftol = prm.subpix_tol;
      P.max_iter = prm.max_iter > 0 ? prm.max_iter : 200;
      P.state = PS_INIT;
    }
    st.hwb = st.hwm = 0.0;
    st.rc[0] = __builtin_nan("");
    make_request(st.k, st.Xq, RQ_VALUE, 0.0, 0.0, &st.rq, st.rc);
  }
  bool cl = false;
  for (;;) {
    __syncthreads();  // the request is published
    const Request q = st.rq;
    if (q.kind == RQ_DONE) break;
    if (q.kind == RQ_DERIVS) {
      double S[kSpSums];
      pixel_sums_derivs(o, q, S, cl);
      wave_partials<kSpSums>(S, red);
    } else {
      double S[3];
      pixel_sums_value(o, q, S, cl);
      wave_partials<3>(S, red);
    }
    __syncthreads();  // the partials are written
    if (threadIdx.x < 64) {
      // wave 0: lane k adds the waves' partials of sum k in wave order, then thread 0
      // steps the optimizer (LDS operations of one wave complete in order)
      const int ns = q.kind == RQ_DERIVS ? kSpSums : 3;
      if ((int)threadIdx.x < ns) {
        double t = red[threadIdx.x];
#pragma unroll
        for (int w = 1; w < kSpWaves; w++) t += red[w * ns + threadIdx.x];
        st.sums[threadIdx.x] = t;
      }
      if (threadIdx.x == 0) sp_step();
    }
  }
  const bool any_clamped = __syncthreads_or(cl ? 1 : 0);
  if (threadIdx.x == 0) {
    double X[3], fret;
    if (st.lm) {
      X[0] = st.opt.M.X[0]; X[1] = st.opt.M.X[1]; X[2] = st.opt.M.X[2];
      fret = st.opt.M.d.f;
    } else {
      X[0] = st.opt.P.p[0]; X[1] = st.opt.P.p[1]; X[2] = st.opt.P.p[2];
      fret = st.opt.P.fret;
    }
    gr.NCC_subpix = 1.0 - fret;
    gr.upix += X[0];
    gr.vpix += X[1];
    gr.rot_inc += X[2];
    gr.dx += X[0];
    gr.dy += X[1];
    gr.drot += X[2];
    if (any_clamped) gr.status |= PTG_STATUS_CLAMPED;
    atomicAdd(evals, (unsigned long long)st.nevals);
    evals_g[g] = st.nevals;
  }
}

static SpLayout sp_layout(int max_prad, double dmax) {
  SpLayout lay;
  // the tile covers the pattern's extent, the optimizer's excursion up to the displacement
  // bound, and the 4x4 stencil; samples outside it read global memory
  const double dm = dmax == dmax ? std::min(std::max(dmax, 0.0), 16.0) : 16.0;
  lay.margin = 3 + (int)ceil(dm);
  const int tw = 2 * (max_prad + lay.margin) + 2;
  const size_t tile_bytes = (size_t)tw * tw * sizeof(uint16_t);
  lay.tile_elems = tile_bytes <= (size_t)kSpTileBytesCap ? tw * tw : 0;
  return lay;
}

static size_t sp_lds_bytes(const SpLayout& lay) {
  return kSpOffTile + sp_align(sizeof(uint16_t) * lay.tile_elems);
}


hipError_t launch_subpixel(Context& c, const ptg_subpix_params& p, int sel, double thr, hipStream_t st) {
  if (c.n_unmasked == 0) return hipSuccess;
  const SpLayout lay = sp_layout(c.max_prad_unmasked, p.subpix_displacement_max);
  const size_t lds = sp_lds_bytes(lay);
  hipError_t e = ensure_kernel_lds(c.device, (const void*)k_subpixel, lds);
  if (e != hipSuccess) return e;
  k_subpixel<<<c.n_unmasked, kSpThreads, lds, st ? st : c.stream>>>(
      c.d_frames, c.d_grains, c.d_pbegin, c.d_spwords, c.d_m0, c.d_p0, c.d_d0, c.d_sd0,
      c.d_list_lpt, c.n_unmasked,
      p, reinterpret_cast<unsigned long long*>(c.d_evals), c.d_prad, lay, c.d_evals_g, sel, thr);
  return hipGetLastError();
}

// NCC_to_minimize_xyR for a batch of (grain, X); with grad != nullptr also the
// analytic gradient. One workgroup per evaluation, staged as the subpixel stage stages
// its grains, so both run the same objective code.
__global__ __launch_bounds__(kSpThreads) void k_objective(const uint16_t* __restrict__ I1, int dimx, int dimy,
                                                          int64_t pitch, const ptg_grain* __restrict__ grains,
                                                          const int64_t* __restrict__ pbegin,
                                                          const uint32_t* __restrict__ words,
                                                          const double* __restrict__ m0,
                                                          const int32_t* __restrict__ p0,
                                                          const double* __restrict__ d0,
                                                          const double* __restrict__ sd0, int n_eval,
                                                          const int32_t* __restrict__ ig,
                                                          const double* __restrict__ Xs, double dmax, double stiff,
                                                          const int32_t* __restrict__ prad, SpLayout lay,
                                                          double* __restrict__ f, double* __restrict__ grad) {
  const int e = blockIdx.x;
  if (e >= n_eval) return;
  double* red = reinterpret_cast<double*>(sp_lds + kSpOffRed);
  const int g = ig[e];
  const ptg_grain& gr = grains[g];
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  Obj o;
  load_obj(o, gr, I1, dimx, dimy, pitch, words + b, m0[g], p0 + 2 * b, d0 + b, n, sd0[g], dmax, stiff);
  stage_grain(o, gr, prad[g], lay);
  const double X[3] = {Xs[3 * e], Xs[3 * e + 1], Xs[3 * e + 2]};
  const ObjConst k = obj_const(o);
  Request q;
  make_request(k, X, grad ? RQ_DERIVS : RQ_VALUE, 0.0, 0.0, &q);
  double wb_, wm_;
  bool cl = false;
  double S[kSpSums];
  if (grad) {
    pixel_sums_derivs(o, q, S, cl);
    wave_partials<kSpSums>(S, red);
    __syncthreads();
    if (threadIdx.x == 0) {
      reduce_partials<kSpSums>(red, S);
      Derivs d;
      finish_derivs(k, S, X, d, &wb_, &wm_);
      f[e] = d.f;
      grad[3 * e] = d.grad[0];
      grad[3 * e + 1] = d.grad[1];
      grad[3 * e + 2] = d.grad[2];
    }
  } else {
    pixel_sums_value(o, q, S, cl);
    wave_partials<3>(S, red);
    __syncthreads();
    if (threadIdx.x == 0) {
      reduce_partials<3>(red, S);
      f[e] = finish_value(k, S, X, &wb_, &wm_);
    }
  }
}

hipError_t launch_objective_batch(Context& c, int32_t n, const int32_t* d_ig, const double* d_X, double dmax,
                                  double stiff, double* d_f, double* d_grad) {
  if (n == 0) return hipSuccess;
  const Image& im = c.img[c.icur];
  int max_prad = 0;
  for (int32_t r : c.h_prad) max_prad = std::max(max_prad, r);
  const SpLayout lay = sp_layout(max_prad, dmax);
  const size_t lds = sp_lds_bytes(lay);
  hipError_t e = ensure_kernel_lds(c.device, (const void*)k_objective, lds);
  if (e != hipSuccess) return e;
  k_objective<<<n, kSpThreads, lds, c.stream>>>(im.d, im.dimx, im.dimy, im.pitch, c.d_grains, c.d_pbegin,
                                                c.d_spwords, c.d_m0, c.d_p0, c.d_d0, c.d_sd0, n, d_ig, d_X, dmax,
                                                stiff, c.d_prad, lay, d_f, d_grad);
  return hipGetLastError();
}

}  // namespace ptg
