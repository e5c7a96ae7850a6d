// ptg_device.hpp — device helpers shared by the gfx950 kernels of libptgpu.
//
// Every translation unit is compiled with -ffp-contract=off, so each double
// expression below rounds as written, matching the reference's x86-64 build (which
// has no FMA). Fused multiply-adds appear only where written explicitly with fma().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ptg_internal.hpp"

namespace ptg {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// A frame pixel read as a global-memory load. Frame pointers reach the kernels through a device
// record, so the compiler sees them as generic addresses: their loads become flat loads, which
// it keeps in order with LDS accesses (a generic address could be LDS) and which also count
// against the LDS wait counter, so a staging loop waits for each load before its LDS store.
typedef __attribute__((address_space(1))) const uint16_t glb_u16;
__device__ __forceinline__ unsigned int frame_px(const uint16_t* base, int64_t idx) {
  return ((const glb_u16*)base)[idx];
}

// rotate_pixel_pattern, src/p-tracker.cpp:615-621: x' = round(c*dx + s*dy),
// y' = round(-s*dx + c*dy), rounding half away from zero as std::round does.
__device__ __forceinline__ void rotate_px(int dx, int dy, double c, double s, int& xo, int& yo) {
  double xpix = c * (double)dx + s * (double)dy;
  double ypix = -s * (double)dx + c * (double)dy;
  xo = (int)round(xpix);
  yo = (int)round(ypix);
}

// Whether the rotation of (dx, dy) by (c, s) could round differently with a cosine and sine that
// differ from c and s by up to 2^-50 (8 units in the last place of values in [0.5, 1): the device
// library and the host C library are each accurate to a few units): x = c dx + s dy then moves by
// at most (|dx| + |dy|) 2^-50 plus the roundings of the products and the sum (each below
// (|dx| + |dy|) 2^-52), well inside m = (|dx| + |dy| + 2) 2^-44, and round(x) can change only if x
// lies within m of a half-integer. Likewise y = -s dx + c dy.
__device__ __forceinline__ bool rot_near_half(int dx, int dy, double c, double s) {
  // (fused: x moves by less than one more rounding, far inside m)
  const double x = fma(c, (double)dx, s * (double)dy);
  const double y = fma(-s, (double)dx, c * (double)dy);
  const double m = (double)(abs(dx) + abs(dy) + 2) * 0x1p-44;
  const double fx = x - floor(x), fy = y - floor(y);
  return fabs(fx - 0.5) <= m || fabs(fy - 0.5) <= m;
}

// Appends one entry to the rotation guard list (entries beyond its capacity are counted only:
// the host then redoes the work with its own tables).
__device__ __forceinline__ void rot_guard_push(RotGuardEntry* guard, int32_t* count, int g, int k, double theta,
                                               double c, double s) {
  const int pos = atomicAdd(count, 1);
  if (pos < kRotGuardCap) {
    RotGuardEntry e;
    e.g = g;
    e.k = k;
    e.theta = theta;
    e.c = c;
    e.s = s;
    guard[pos] = e;
  }
}

// CubicInterpolate, src/interpol.cpp:42-53, in the reference's operation order.
__device__ __forceinline__ double cubic1(double p0, double p1, double p2, double p3, double x) {
  double f0 = p1;
  double f1 = p2;
  double fp0 = 0.5 * (p2 - p0);
  double fp1 = 0.5 * (p3 - p1);
  double d = f0;
  double c = fp0;
  double b = -3.0 * f0 + 3.0 * f1 - 2.0 * fp0 - fp1;
  double a = 2.0 * f0 - 2.0 * f1 + fp0 + fp1;
  double x2 = x * x;
  return (a * x2 * x + b * x2 + c * x + d);
}

// image_interpCubic::getValue, src/interpol.cpp:55-86, on a row-major frame. Taps
// outside the frame are clamped to its edge (the reference reads out of bounds); the
// return flag says whether any tap was clamped.
__device__ __forceinline__ double interp_cubic(const uint16_t* __restrict__ im, int dimx, int dimy,
                                               int64_t pitch, double x, double y, bool& clamped) {
  int ix = (int)floor(x);
  int iy = (int)floor(y);
  double y01 = y - (double)iy;
  double x01 = x - (double)ix;
  // written without ix + 2 so that saturated conversions of huge coordinates cannot wrap
  bool inside = ix >= 1 && iy >= 1 && ix < dimx - 2 && iy < dimy - 2;
  double arr[4];
  if (inside) {
    const uint16_t* r0 = im + (int64_t)(iy - 1) * pitch + (ix - 1);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      arr[i] = cubic1((double)r0[i], (double)r0[pitch + i], (double)r0[2 * pitch + i],
                      (double)r0[3 * pitch + i], y01);
    }
  } else {
    clamped = true;
    int xs[4], ys[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      xs[i] = clampi(ix - 1 + i, 0, dimx - 1);
      ys[i] = clampi(iy - 1 + i, 0, dimy - 1);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      arr[i] = cubic1((double)im[(int64_t)ys[0] * pitch + xs[i]], (double)im[(int64_t)ys[1] * pitch + xs[i]],
                      (double)im[(int64_t)ys[2] * pitch + xs[i]], (double)im[(int64_t)ys[3] * pitch + xs[i]], y01);
    }
  }
  return cubic1(arr[0], arr[1], arr[2], arr[3], x01);
}

// Wave-wide (64-lane) butterfly sums. Each step adds a pair in both lanes of the pair
// (a + b == b + a in IEEE), so every lane ends with the same bits: control flow that
// depends on the sum stays uniform across the wave.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

}  // namespace ptg
