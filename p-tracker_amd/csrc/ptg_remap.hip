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

}  // namespace ptg
