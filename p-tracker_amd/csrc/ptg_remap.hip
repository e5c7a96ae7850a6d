// ptg_remap.hip — point sampling, undistor() and the distortion-correction remap of
// libptgpu (gfx950).
//
// undistor (src/distortion_correction.cpp:20-61) maps a distorted point to its
// undistorted position. The remap produces the undistorted frame: for each output
// pixel it inverts undistor() by Newton iteration (analytic 2x2 Jacobian) and samples
// the distorted frame with the Catmull-Rom sampler of src/interpol.cpp:42-86. The
// expressions are written as the reference's, compiled unfused, with the parallax
// directions' cos/sin taken from the host's libm, so each pixel's value is
// bit-identical to the CPU restatement.
#include <algorithm>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

struct DistoParams {
  double Xc, Yc, K1, K2, K3, P1, P2, P3;
  double Xcp, Ycp, parallax1, parallax2;
  double nx1, ny1, nx2, ny2;  // cos(theta1), -sin(theta1), -sin(theta2), cos(theta2)
};

__device__ __forceinline__ void undistor_d(const DistoParams& P, double xd, double yd, double& xu, double& yu) {
  const double Xc = P.Xc, Yc = P.Yc, K1 = P.K1, K2 = P.K2, K3 = P.K3, P1 = P.P1, P2 = P.P2, P3 = P.P3;
  double dx = xd - Xc, dy = yd - Yc;
  double r2 = dx * dx + dy * dy;
  double r4 = r2 * r2;
  double r6 = r4 * r2;
  double x = xd + dx * (K1 * r2 + K2 * r4 + K3 * r6) + (P1 * (r2 + 2.0 * dx * dx) + 2.0 * P2 * dx * dy) * (1.0 + P3 * r2);
  double y = yd + dy * (K1 * r2 + K2 * r4 + K3 * r6) + (2.0 * P1 * dx * dy + P2 * (r2 + 2.0 * dy * dy)) * (1.0 + P3 * r2);
  const double Xcp = P.Xcp, Ycp = P.Ycp, parallax1 = P.parallax1, parallax2 = P.parallax2;
  const double nx1 = P.nx1, ny1 = P.ny1, nx2 = P.nx2, ny2 = P.ny2;
  double dX = x - Xcp, dY = y - Ycp;
  double S1 = (1.0 + parallax1) * (dX * nx1 + dY * ny1);
  double S2 = (1.0 + parallax2) * (dX * nx2 + dY * ny2);
  xu = Xcp + S1 * nx1 + S2 * nx2;
  yu = Ycp + S1 * ny1 + S2 * ny2;
}

__device__ __forceinline__ void undistor_jac_d(const DistoParams& P, double xd, double yd, double* J) {
  const double Xc = P.Xc, Yc = P.Yc, K1 = P.K1, K2 = P.K2, K3 = P.K3, P1 = P.P1, P2 = P.P2, P3 = P.P3;
  double dx = xd - Xc, dy = yd - Yc;
  double r2 = dx * dx + dy * dy;
  double r4 = r2 * r2;
  double Kr = K1 * r2 + K2 * r4 + K3 * (r4 * r2);
  double Kp = K1 + 2.0 * K2 * r2 + 3.0 * K3 * r4;
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
  const double nx1 = P.nx1, ny1 = P.ny1, nx2 = P.nx2, ny2 = P.ny2;
  double a1 = 1.0 + P.parallax1, a2 = 1.0 + P.parallax2;
  double pxx = a1 * nx1 * nx1 + a2 * nx2 * nx2, pxy = a1 * ny1 * nx1 + a2 * ny2 * nx2;
  double pyx = a1 * nx1 * ny1 + a2 * nx2 * ny2, pyy = a1 * ny1 * ny1 + a2 * ny2 * ny2;
  J[0] = pxx * bxx + pxy * byx;
  J[1] = pxx * bxy + pxy * byy;
  J[2] = pyx * bxx + pyy * byx;
  J[3] = pyx * bxy + pyy * byy;
}

#define REMAP_NEWTON_ITERS 6

__device__ __forceinline__ void distort_point_d(const DistoParams& P, double xu, double yu, double& xd, double& yd) {
  double x = xu, y = yu;
#pragma unroll 1
  for (int it = 0; it < REMAP_NEWTON_ITERS; it++) {
    double fx, fy, J[4];
    undistor_d(P, x, y, fx, fy);
    fx -= xu;
    fy -= yu;
    undistor_jac_d(P, x, y, J);
    double det = J[0] * J[3] - J[1] * J[2];
    x -= (J[3] * fx - J[1] * fy) / det;
    y -= (J[0] * fy - J[2] * fx) / det;
  }
  xd = x;
  yd = y;
}

static DistoParams make_params(const double* X, int dimx, int dimy) {
  DistoParams P;
  // centres left negative default to the image centre, src/distortion_correction.cpp:230-233
  P.Xc = X[0] < 0.0 ? 0.5 * (double)dimx : X[0];
  P.Yc = X[1] < 0.0 ? 0.5 * (double)dimy : X[1];
  P.K1 = X[2];
  P.K2 = X[3];
  P.K3 = X[4];
  P.P1 = X[5];
  P.P2 = X[6];
  P.P3 = X[7];
  P.Xcp = X[8] < 0.0 ? 0.5 * (double)dimx : X[8];
  P.Ycp = X[9] < 0.0 ? 0.5 * (double)dimy : X[9];
  P.parallax1 = X[11];
  P.parallax2 = X[13];
  P.nx1 = cos(X[10]);
  P.ny1 = -sin(X[10]);
  P.nx2 = -sin(X[12]);
  P.ny2 = cos(X[12]);
  return P;
}

__global__ __launch_bounds__(256) void k_remap(const uint16_t* __restrict__ src, int dimx, int dimy, int64_t pitch,
                                               uint16_t* __restrict__ dst, int64_t dpitch, DistoParams P,
                                               double maxv) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= dimx || y >= dimy) return;
  double xd, yd;
  distort_point_d(P, (double)x, (double)y, xd, yd);
  double v = 0.0;
  if (xd > -1.0e6 && xd < 1.0e6 && yd > -1.0e6 && yd < 1.0e6) {
    bool cl = false;
    v = interp_cubic(src, dimx, dimy, pitch, xd, yd, cl);
  }
  v = round(v);
  if (!(v > 0.0)) v = 0.0;
  if (v > maxv) v = maxv;
  dst[(int64_t)y * dpitch + x] = (uint16_t)v;
}

hipError_t launch_remap(Context& c, int src, int dst, const double* params) {
  const Image& s = c.img[src];
  const Image& d = c.img[dst];
  DistoParams P = make_params(params, s.dimx, s.dimy);
  dim3 grid((s.dimx + 63) / 64, (s.dimy + 3) / 4);
  k_remap<<<grid, 256, 0, c.stream>>>(s.d, s.dimx, s.dimy, s.pitch, d.d, d.pitch, P,
                                      (double)((1 << s.bits) - 1));
  return hipGetLastError();
}

__global__ void k_undistort_points(DistoParams P, int n, const double* __restrict__ in, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double xu, yu;
  undistor_d(P, in[2 * k], in[2 * k + 1], xu, yu);
  out[2 * k] = xu;
  out[2 * k + 1] = yu;
}

hipError_t launch_undistort_points(Context& c, const double* params, int32_t n, const double* d_in, double* d_out) {
  if (n == 0) return hipSuccess;
  // explicit parameters: no centre defaulting for the point map (undistor takes X as given)
  DistoParams P;
  P.Xc = params[0];
  P.Yc = params[1];
  P.K1 = params[2];
  P.K2 = params[3];
  P.K3 = params[4];
  P.P1 = params[5];
  P.P2 = params[6];
  P.P3 = params[7];
  P.Xcp = params[8];
  P.Ycp = params[9];
  P.parallax1 = params[11];
  P.parallax2 = params[13];
  P.nx1 = cos(params[10]);
  P.ny1 = -sin(params[10]);
  P.nx2 = -sin(params[12]);
  P.ny2 = cos(params[12]);
  k_undistort_points<<<(n + 255) / 256, 256, 0, c.stream>>>(P, n, d_in, d_out);
  return hipGetLastError();
}

__global__ void k_interp_batch(const uint16_t* __restrict__ im, int dimx, int dimy, int64_t pitch, int n,
                               const double* __restrict__ xy, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  bool cl = false;
  out[k] = interp_cubic(im, dimx, dimy, pitch, xy[2 * k], xy[2 * k + 1], cl);
}

hipError_t launch_interp_batch(Context& c, int index, int32_t n, const double* d_xy, double* d_out) {
  if (n == 0) return hipSuccess;
  const Image& im = c.img[index];
  k_interp_batch<<<(n + 255) / 256, 256, 0, c.stream>>>(im.d, im.dimx, im.dimy, im.pitch, n, d_xy, d_out);
  return hipGetLastError();
}

// ---- the equiprojectivity criterion of the distortion calibration ---------------------
// For a pair of grains (i, j) and an image k after the reference, with A1, B1 the
// undistorted reference pixels of i and j and A2, B2 the undistorted tracked positions,
//   e = (B1 - A1) / |B1 - A1|,  r = e . ((B2 - B1) - (A2 - A1))
// (src/distortion_correction.cpp:277-303, :409-475: a rigid motion moves every pair
// equally along its own direction). The reference minimises sum r^2 with Powell or with
// MINPACK's lmdif (forward-difference Jacobian); here each residual carries its analytic
// gradient with respect to the 14 parameters, a vector-Jacobian product through
// undistor() and the pair normalisation, and the workgroups reduce sum r^2, J^T r and the
// upper triangle of J^T J.
constexpr int kEqParams = 14;
constexpr int kEqSums = 1 + kEqParams + kEqParams * (kEqParams + 1) / 2;  // 120
constexpr int kEqThreads = 256;

// g += cx d(xu)/dp + cy d(yu)/dp for the 14 parameters p of undistor at (xd, yd).
__device__ __forceinline__ void undistor_vjp(const DistoParams& P, double xd, double yd, double cx, double cy,
                                             double* g) {
  const double K1 = P.K1, K2 = P.K2, K3 = P.K3, P1 = P.P1, P2 = P.P2, P3 = P.P3;
  const double dx = xd - P.Xc, dy = yd - P.Yc;
  const double r2 = dx * dx + dy * dy, r4 = r2 * r2, r6 = r4 * r2;
  const double R = K1 * r2 + K2 * r4 + K3 * r6;
  const double Rp = K1 + 2.0 * K2 * r2 + 3.0 * K3 * r4;  // dR / d(r2)
  const double Tx = P1 * (r2 + 2.0 * dx * dx) + 2.0 * P2 * dx * dy;
  const double Ty = 2.0 * P1 * dx * dy + P2 * (r2 + 2.0 * dy * dy);
  const double Q = 1.0 + P3 * r2;
  const double xb = xd + dx * R + Tx * Q, yb = yd + dy * R + Ty * Q;
  const double nx1 = P.nx1, ny1 = P.ny1, nx2 = P.nx2, ny2 = P.ny2;
  const double a1 = 1.0 + P.parallax1, a2 = 1.0 + P.parallax2;
  const double dX = xb - P.Xcp, dY = yb - P.Ycp;
  const double p1 = dX * nx1 + dY * ny1, p2 = dX * nx2 + dY * ny2;
  const double S1 = a1 * p1, S2 = a2 * p2;
  // xu = Xcp + S1 nx1 + S2 nx2, yu = Ycp + S1 ny1 + S2 ny2
  const double gS1 = cx * nx1 + cy * ny1, gS2 = cx * nx2 + cy * ny2;
  g[8] += cx - a1 * gS1 * nx1 - a2 * gS2 * nx2;
  g[9] += cy - a1 * gS1 * ny1 - a2 * gS2 * ny2;
  g[11] += gS1 * p1;
  g[13] += gS2 * p2;
  // theta1: d(nx1) = ny1, d(ny1) = -nx1; theta2: d(nx2) = -ny2, d(ny2) = nx2
  g[10] += gS1 * (a1 * (dX * ny1 - dY * nx1)) + S1 * (cx * ny1 - cy * nx1);
  g[12] += gS2 * (a2 * (-dX * ny2 + dY * nx2)) + S2 * (-cx * ny2 + cy * nx2);
  // back to the Brown-Conrady point (xb, yb)
  const double gxb = a1 * gS1 * nx1 + a2 * gS2 * nx2, gyb = a1 * gS1 * ny1 + a2 * gS2 * ny2;
  const double gr = gxb * dx + gyb * dy;
  g[2] += gr * r2;
  g[3] += gr * r4;
  g[4] += gr * r6;
  g[5] += (gxb * (r2 + 2.0 * dx * dx) + gyb * (2.0 * dx * dy)) * Q;
  g[6] += (gxb * (2.0 * dx * dy) + gyb * (r2 + 2.0 * dy * dy)) * Q;
  g[7] += (gxb * Tx + gyb * Ty) * r2;
  // the centre enters through dx = xd - Xc, dy = yd - Yc
  const double bxx = R + dx * Rp * 2.0 * dx + Q * (6.0 * P1 * dx + 2.0 * P2 * dy) + Tx * P3 * 2.0 * dx;
  const double bxy = dx * Rp * 2.0 * dy + Q * (2.0 * P1 * dy + 2.0 * P2 * dx) + Tx * P3 * 2.0 * dy;
  const double byx = dy * Rp * 2.0 * dx + Q * (2.0 * P1 * dy + 2.0 * P2 * dx) + Ty * P3 * 2.0 * dx;
  const double byy = R + dy * Rp * 2.0 * dy + Q * (2.0 * P1 * dx + 6.0 * P2 * dy) + Ty * P3 * 2.0 * dy;
  g[0] -= gxb * bxx + gyb * byx;
  g[1] -= gxb * bxy + gyb * byy;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// One residual per thread (index (k, il), image-major as the reference's fvec); residual,
// Jacobian row and workgroup partial sums. With a.orig set (the POWELL_ITER form,
// src/distortion_correction.cpp:316-354) the first positions A1, B1 are those given, which do
// not depend on the parameters.
__global__ __launch_bounds__(kEqThreads) void k_equiproj(EquiprojArgs a, DistoParams P) {
  __shared__ double wsum[kEqThreads / 64][kEqSums];
  const int64_t t = (int64_t)blockIdx.x * kEqThreads + threadIdx.x;
  const int64_t m = (int64_t)a.nimg * a.npairs;
  double r = 0.0;
  double g[kEqParams];
#pragma unroll
  for (int p = 0; p < kEqParams; p++) g[p] = 0.0;
  if (t < m) {
    const int k = (int)(t / a.npairs), il = (int)(t - (int64_t)k * a.npairs);
    const int i = a.pairs[2 * il], j = a.pairs[2 * il + 1];
    const double* ncc = a.ncc + (int64_t)k * a.G;
    if (!(ncc[i] < a.NCC_min || ncc[j] < a.NCC_min)) {
      const double* d = a.disp + 2 * (int64_t)k * a.G;
      const double xi = a.ref[2 * i], yi = a.ref[2 * i + 1], xj = a.ref[2 * j], yj = a.ref[2 * j + 1];
      const double xi2 = xi + d[2 * i], yi2 = yi + d[2 * i + 1], xj2 = xj + d[2 * j], yj2 = yj + d[2 * j + 1];
      double A1x, A1y, A2x, A2y, B1x, B1y, B2x, B2y;
      if (a.orig) {
        A1x = a.orig[2 * i];
        A1y = a.orig[2 * i + 1];
        B1x = a.orig[2 * j];
        B1y = a.orig[2 * j + 1];
      } else {
        undistor_d(P, xi, yi, A1x, A1y);
        undistor_d(P, xj, yj, B1x, B1y);
      }
      undistor_d(P, xi2, yi2, A2x, A2y);
      undistor_d(P, xj2, yj2, B2x, B2y);
      double ABx = B1x - A1x, ABy = B1y - A1y;
      const double AB = sqrt(ABx * ABx + ABy * ABy);
      ABx /= AB;
      ABy /= AB;
      const double BBx = B2x - B1x, BBy = B2y - B1y, AAx = A2x - A1x, AAy = A2y - A1y;
      const double Dx = BBx - AAx, Dy = BBy - AAy;
      r = (ABx * Dx) + (ABy * Dy);
      // dr/d(B1 - A1) = (D - e r) / |AB|, dr/dD = e
      const double wx = (Dx - ABx * r) / AB, wy = (Dy - ABy * r) / AB;
      if (!a.orig) {
        undistor_vjp(P, xi, yi, -wx + ABx, -wy + ABy, g);
        undistor_vjp(P, xj, yj, wx - ABx, wy - ABy, g);
      }
      undistor_vjp(P, xi2, yi2, -ABx, -ABy, g);
      undistor_vjp(P, xj2, yj2, ABx, ABy, g);
    }
    if (a.res) a.res[t] = r;
    if (a.jac)
#pragma unroll
      for (int p = 0; p < kEqParams; p++) a.jac[t * kEqParams + p] = g[p];
  }
  // sum r^2, J^T r, J^T J (upper triangle, row major) over the workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s = wave_sum_d(r * r);
  if (lane == 0) wsum[wave][0] = s;
#pragma unroll
  for (int p = 0; p < kEqParams; p++) {
    s = wave_sum_d(g[p] * r);
    if (lane == 0) wsum[wave][1 + p] = s;
  }
  int q = 1 + kEqParams;
#pragma unroll
  for (int p = 0; p < kEqParams; p++)
#pragma unroll
    for (int p2 = p; p2 < kEqParams; p2++, q++) {
      s = wave_sum_d(g[p] * g[p2]);
      if (lane == 0) wsum[wave][q] = s;
    }
  __syncthreads();
  if (threadIdx.x < kEqSums) {
    double v = 0.0;
    for (int w = 0; w < kEqThreads / 64; w++) v += wsum[w][threadIdx.x];
    a.partial[(int64_t)blockIdx.x * kEqSums + threadIdx.x] = v;
  }
}

// Sum of the workgroup partials in workgroup order (the same total on every run).
__global__ void k_equiproj_reduce(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kEqSums) return;
  double v = 0.0;
  for (int b = 0; b < nblocks; b++) v += partial[(int64_t)b * kEqSums + q];
  out[q] = v;
}

hipError_t launch_equiproj(Context& c, const EquiprojArgs& a, const double* params14, double* d_out) {
  const int64_t m = (int64_t)a.nimg * a.npairs;
  const int nblocks = (int)std::max<int64_t>(1, (m + kEqThreads - 1) / kEqThreads);
  // undistor takes the parameters as given; cos/sin of the parallax angles from the host
  DistoParams P;
  P.Xc = params14[0];
  P.Yc = params14[1];
  P.K1 = params14[2];
  P.K2 = params14[3];
  P.K3 = params14[4];
  P.P1 = params14[5];
  P.P2 = params14[6];
  P.P3 = params14[7];
  P.Xcp = params14[8];
  P.Ycp = params14[9];
  P.parallax1 = params14[11];
  P.parallax2 = params14[13];
  P.nx1 = cos(params14[10]);
  P.ny1 = -sin(params14[10]);
  P.nx2 = -sin(params14[12]);
  P.ny2 = cos(params14[12]);
  k_equiproj<<<nblocks, kEqThreads, 0, c.stream>>>(a, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_equiproj_reduce<<<1, 128, 0, c.stream>>>(a.partial, nblocks, d_out);
  return hipGetLastError();
}

int equiproj_partial_blocks(int64_t m) { return (int)std::max<int64_t>(1, (m + kEqThreads - 1) / kEqThreads); }

}  // namespace ptg
