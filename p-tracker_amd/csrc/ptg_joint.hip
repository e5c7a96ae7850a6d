// ptg_joint.hip — the differentiation form's joint refinement of libptgpu (gfx950): every
// grain's motion and a global radial lens, by Gauss-Newton / Levenberg-Marquardt on the ZNCC
// residual with its adjoint sums computed per grain on the GPU.
//
// Model (no counterpart in the reference, which fits the lens only through the
// equiprojectivity criterion of src/distortion_correction.cpp:135-479): undistor's radial
// part U(p) = p + (p - c)(K1 r^2 + K2 r^4), r = |p - c| (src/distortion_correction.cpp:20-39),
// maps frame pixels to scene coordinates. Pattern pixel q = ref + p0_i of a grain (its
// reference pixel plus the pattern0_rotated offset) moves rigidly in the scene by T and the
// rotation theta about U(ref), with NCC_to_minimize_xyR's convention (x, y) -> (c x + s y,
// -s x + c y) (src/p-tracker.cpp:1040-1045):
//   w' = U(ref) + T + R(theta) (U(q) - U(ref)),
// and is sampled in the current frame at x' = U^-1(w') with the reference's Catmull-Rom
// sampler (src/interpol.cpp:42-86). With K1 = K2 = 0 this is NCC_to_minimize_xyR at
// X = (T - (dx, dy), theta - drot). The per-grain objective is f = 1 - NCC.
//
// One workgroup per grain accumulates, in one pass over the pattern, the sums that give f,
// its gradient over p = (Tx, Ty, theta, k1, k2) and the Gauss-Newton matrix of the
// normalised residual r = f^ - g^ (f = |r|^2 / 2, f^ and g^ the unit, centred pattern and
// sampled vectors): with D the pixels' derivatives dg/dp centred over the pattern and
// sigma^2 = C1C1, J^T J = (D^T D - (D^T g^)(D^T g^)^T) / sigma^2 and
// df/dp = -(D^T f^ - NCC D^T g^) / sigma. The derivative of a sample goes through the
// sampler's gradient and dx'/dp = J_U(x')^-1 (dw'/dp - dU/dp(x')): the adjoint of the
// inverse lens without iterating it.
#include <math.h>

#include <algorithm>
#include <vector>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

// outputs per grain: f, NCC, df/dp (5), J^T J upper triangle (15, rows a <= b), C1C1, clamped
constexpr int kJointOut = 24;
constexpr int kJointThreads = 256;
constexpr int kJointSums = 34;  // n-free sums: Sg, Sgg, Sdg, Sa[5], Sga[5], Sda[5], Sab[15], spare

struct JointArgs {
  const uint16_t* I1;
  int32_t dimx, dimy;
  int64_t pitch;
  const ptg_grain* grains;
  const int64_t* pbegin;
  const int32_t* p0;   // pattern0_rotated (dx, dy) pairs
  const double* d0;    // I0(ref + p0) - mean0, same indexing
  const double* sd0;   // per grain: sum of d0
  const double* m0;    // per grain: mean0 (the samples are summed about it)
  const int32_t* list; // grains of the call
  int32_t nlist;
  const double* P;     // per list entry: Tx, Ty, theta
  double k1, k2, xc, yc;
  double* out;         // per list entry: kJointOut doubles
};

struct RadialLens {
  double k1, k2, xc, yc;
  __device__ __forceinline__ void U(double px, double py, double& ux, double& uy) const {
    const double dx = px - xc, dy = py - yc;
    const double r2 = dx * dx + dy * dy;
    const double K = k1 * r2 + k2 * (r2 * r2);
    ux = px + dx * K;
    uy = py + dy * K;
  }
  // J_U(p)^-1 as (a, b; b, d)
  __device__ __forceinline__ void Jinv(double px, double py, double& a, double& b, double& d) const {
    const double dx = px - xc, dy = py - yc;
    const double r2 = dx * dx + dy * dy;
    const double K = k1 * r2 + k2 * (r2 * r2), Kp = 2.0 * (k1 + 2.0 * k2 * r2);
    const double jxx = 1.0 + K + Kp * dx * dx, jxy = Kp * dx * dy, jyy = 1.0 + K + Kp * dy * dy;
    const double inv = 1.0 / (jxx * jyy - jxy * jxy);
    a = jyy * inv;
    b = -jxy * inv;
    d = jxx * inv;
  }
  // dU/dk1 = (p - c) r^2, dU/dk2 = (p - c) r^4
  __device__ __forceinline__ void dUdk(double px, double py, double& x1, double& y1, double& x2, double& y2) const {
    const double dx = px - xc, dy = py - yc;
    const double r2 = dx * dx + dy * dy;
    x1 = dx * r2;
    y1 = dy * r2;
    x2 = x1 * r2;
    y2 = y1 * r2;
  }
  // U^-1(w): one fixed-point step, then Newton on the analytic Jacobian
  __device__ __forceinline__ void Uinv(double wx, double wy, double& x, double& y) const {
    double ux, uy;
    U(wx, wy, ux, uy);
    x = 2.0 * wx - ux;
    y = 2.0 * wy - uy;
#pragma unroll
    for (int it = 0; it < 3; it++) {
      U(x, y, ux, uy);
      const double fx = ux - wx, fy = uy - wy;
      double a, b, d;
      Jinv(x, y, a, b, d);
      x -= a * fx + b * fy;
      y -= b * fx + d * fy;
    }
  }
};

// The reference's Catmull-Rom sampler (its Hermite form, as interp_cubic) with the gradient
// of the same interpolant; taps outside the frame are clamped to its edge.
__device__ __forceinline__ double cubic1_d(double p0, double p1, double p2, double p3, double x) {
  const double f0 = p1, f1 = p2, fp0 = 0.5 * (p2 - p0), fp1 = 0.5 * (p3 - p1);
  const double b = -3.0 * f0 + 3.0 * f1 - 2.0 * fp0 - fp1;
  const double a = 2.0 * f0 - 2.0 * f1 + fp0 + fp1;
  return (3.0 * a * x + 2.0 * b) * x + fp0;
}

__device__ __forceinline__ double interp_cubic_grad(const uint16_t* __restrict__ im, int dimx, int dimy, int64_t pitch,
                                                    double x, double y, double& gx, double& gy, bool& clamped) {
  // a trial step far off the frame (or not a number) samples the frame's edge, in range of int
  x = fmin(fmax(x, -1.0e6), 1.0e6);
  y = fmin(fmax(y, -1.0e6), 1.0e6);
  const int ix = (int)floor(x), iy = (int)floor(y);
  const double ty = y - (double)iy, tx = x - (double)ix;
  const bool inside = ix >= 1 && iy >= 1 && ix < dimx - 2 && iy < dimy - 2;
  if (!inside) clamped = true;
  double arr[4], darr[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int xi = clampi(ix - 1 + i, 0, dimx - 1);
    double p[4];
#pragma unroll
    for (int j = 0; j < 4; j++) p[j] = (double)im[(int64_t)clampi(iy - 1 + j, 0, dimy - 1) * pitch + xi];
    arr[i] = cubic1(p[0], p[1], p[2], p[3], ty);
    darr[i] = cubic1_d(p[0], p[1], p[2], p[3], ty);
  }
  gx = cubic1_d(arr[0], arr[1], arr[2], arr[3], tx);
  gy = cubic1(darr[0], darr[1], darr[2], darr[3], tx);
  return cubic1(arr[0], arr[1], arr[2], arr[3], tx);
}

template <bool kDerivs>
__global__ __launch_bounds__(kJointThreads) void k_joint(JointArgs a) {
  __shared__ double red[kJointThreads / 64][kJointSums];
  __shared__ int clamp_any;
  const int li = blockIdx.x;
  if (li >= a.nlist) return;
  const int g = a.list[li];
  const ptg_grain& gr = a.grains[g];
  const int64_t b = a.pbegin[g];
  const int n = (int)(a.pbegin[g + 1] - b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const RadialLens L{a.k1, a.k2, a.xc, a.yc};
  const double rx = (double)gr.refcoord_xpix, ry = (double)gr.refcoord_ypix;
  double wrx, wry;
  L.U(rx, ry, wrx, wry);
  double urx1 = 0.0, ury1 = 0.0, urx2 = 0.0, ury2 = 0.0;
  if (kDerivs) L.dUdk(rx, ry, urx1, ury1, urx2, ury2);
  const double Tx = a.P[3 * li], Ty = a.P[3 * li + 1], th = a.P[3 * li + 2];
  const double c = cos(th), s = sin(th);
  const double shift = a.m0[g];
  if (threadIdx.x == 0) clamp_any = 0;
  __syncthreads();
  double acc[kJointSums];
#pragma unroll
  for (int q = 0; q < kJointSums; q++) acc[q] = 0.0;
  bool clamped = false;
  for (int i = threadIdx.x; i < n; i += kJointThreads) {
    const double qx = rx + (double)a.p0[2 * (b + i)], qy = ry + (double)a.p0[2 * (b + i) + 1];
    double wqx, wqy;
    L.U(qx, qy, wqx, wqy);
    const double ax = wqx - wrx, ay = wqy - wry;
    const double wx = wrx + Tx + (c * ax + s * ay), wy = wry + Ty + (-s * ax + c * ay);
    double x, y;
    L.Uinv(wx, wy, x, y);
    double gx, gy;
    const double v = interp_cubic_grad(a.I1, a.dimx, a.dimy, a.pitch, x, y, gx, gy, clamped) - shift;
    const double d0 = a.d0[b + i];
    acc[0] += v;
    acc[1] += v * v;
    acc[2] += d0 * v;
    if (kDerivs) {
      double ja, jb, jd;
      L.Jinv(x, y, ja, jb, jd);
      // w' derivatives: T (unit vectors), theta (R' a), k1, k2 (through U(q) and U(ref)), minus
      // dU/dk at x' for the inverse lens
      double qx1, qy1, qx2, qy2, xx1, xy1, xx2, xy2;
      L.dUdk(qx, qy, qx1, qy1, qx2, qy2);
      L.dUdk(x, y, xx1, xy1, xx2, xy2);
      const double e1x = urx1 + (c * (qx1 - urx1) + s * (qy1 - ury1)) - xx1;
      const double e1y = ury1 + (-s * (qx1 - urx1) + c * (qy1 - ury1)) - xy1;
      const double e2x = urx2 + (c * (qx2 - urx2) + s * (qy2 - ury2)) - xx2;
      const double e2y = ury2 + (-s * (qx2 - urx2) + c * (qy2 - ury2)) - xy2;
      const double tdx = -s * ax + c * ay, tdy = -c * ax - s * ay;
      // sample gradient through J_U(x')^-1 (symmetric)
      const double hx = ja * gx + jb * gy, hy = jb * gx + jd * gy;
      double dg[5];
      dg[0] = hx;
      dg[1] = hy;
      dg[2] = hx * tdx + hy * tdy;
      dg[3] = hx * e1x + hy * e1y;
      dg[4] = hx * e2x + hy * e2y;
#pragma unroll
      for (int q = 0; q < 5; q++) {
        acc[3 + q] += dg[q];
        acc[8 + q] += v * dg[q];
        acc[13 + q] += d0 * dg[q];
      }
      int t = 18;
#pragma unroll
      for (int p = 0; p < 5; p++)
#pragma unroll
        for (int q = p; q < 5; q++) acc[t++] += dg[p] * dg[q];
    }
  }
  if (clamped) clamp_any = 1;
#pragma unroll
  for (int q = 0; q < kJointSums; q++) {
    const double t = wave_sum(acc[q]);
    if (lane == 0) red[wave][q] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S[kJointSums];
  for (int q = 0; q < kJointSums; q++) {
    double t = 0.0;
    for (int w = 0; w < kJointThreads / 64; w++) t += red[w][q];
    S[q] = t;
  }
  const double nd = (double)n;
  const double m = S[0] / nd;
  const double C11 = S[1] - nd * m * m;
  const double Sd = a.sd0[g];
  const double den = sqrt(gr.C0C0 * C11);
  const double ncc = (S[2] - m * Sd) / den;
  double* o = a.out + (int64_t)li * kJointOut;
  o[0] = 1.0 - ncc;
  o[1] = ncc;
  if (kDerivs) {
    double ga[5];
    for (int q = 0; q < 5; q++) {
      ga[q] = S[8 + q] - m * S[3 + q];  // sum (g - mean g) dg_q
      o[2 + q] = -((S[13 + q] - Sd * S[3 + q] / nd) / den - ncc * ga[q] / C11);
    }
    int t = 18, u = 7;
    for (int p = 0; p < 5; p++)
      for (int q = p; q < 5; q++, t++, u++) o[u] = (S[t] - S[3 + p] * S[3 + q] / nd - ga[p] * ga[q] / C11) / C11;
  } else {
    for (int q = 2; q < 22; q++) o[q] = 0.0;
  }
  o[22] = C11;
  o[23] = clamp_any ? 1.0 : 0.0;
}

hipError_t launch_joint(Context& c, const int32_t* d_list, int32_t nlist, const double* d_P, const double* lens,
                        bool derivs, double* d_out) {
  if (nlist == 0) return hipSuccess;
  const Image& im = c.img[c.icur];
  JointArgs a;
  a.I1 = im.d;
  a.dimx = im.dimx;
  a.dimy = im.dimy;
  a.pitch = im.pitch;
  a.grains = c.d_grains;
  a.pbegin = c.d_pbegin;
  a.p0 = c.d_p0;
  a.d0 = c.d_d0;
  a.sd0 = c.d_sd0;
  a.m0 = c.d_m0;
  a.list = d_list;
  a.nlist = nlist;
  a.P = d_P;
  a.k1 = lens[0];
  a.k2 = lens[1];
  a.xc = lens[2];
  a.yc = lens[3];
  a.out = d_out;
  if (derivs) k_joint<true><<<nlist, kJointThreads, 0, c.stream>>>(a);
  else k_joint<false><<<nlist, kJointThreads, 0, c.stream>>>(a);
  return hipGetLastError();
}

int joint_out_stride() { return kJointOut; }

}  // namespace ptg
