// ptg_search.hip — precomputation, the plain integer ZNCC search and the per-pair
// bookkeeping kernels of libptgpu (gfx950).
//
// Plain search: one workgroup per (grain, rotation step). The workgroup stages the
// current-frame search window and the step's rotated pattern offsets in LDS, then
// each lane scores whole candidates with the reference's two-pass FP64 sums in the
// reference's pixel order, unfused (src/p-tracker.cpp:698-716), so every candidate's
// NCC is bit-identical to follow_pattern_pixel(). A commit kernel reduces the steps in
// order and applies the reference's update rules.
#include <limits.h>

#include <algorithm>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

// grain.reset(); grain.backup(), src/grain_type_2D.cpp:51-57, as called at
// src/p-tracker.cpp:478-479.
__global__ void k_begin_pair(ptg_grain* __restrict__ g, int G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  ptg_grain& r = g[i];
  r.NCC = r.NCC_rescue = r.NCC_subpix = 0.0;
  r.dx_prev = r.dx;
  r.dy_prev = r.dy;
  r.drot_prev = r.drot;
  r.status = 0;
}

hipError_t launch_begin_pair(Context& c) {
  if (c.G == 0) return hipSuccess;
  k_begin_pair<<<(c.G + 255) / 256, 256, 0, c.stream>>>(c.d_grains, c.G);
  return hipGetLastError();
}

// upix = dx - dx_prev, ..., src/p-tracker.cpp:577-581.
__global__ void k_end_pair(ptg_grain* __restrict__ g, int G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  ptg_grain& r = g[i];
  r.upix = r.dx - r.dx_prev;
  r.vpix = r.dy - r.dy_prev;
  r.rot_inc = r.drot - r.drot_prev;
}

hipError_t launch_end_pair(Context& c) {
  if (c.G == 0) return hipSuccess;
  k_end_pair<<<(c.G + 255) / 256, 256, 0, c.stream>>>(c.d_grains, c.G);
  return hipGetLastError();
}

// The reference swap, src/p-tracker.cpp:598-609.
__global__ void k_change_reference(ptg_grain* __restrict__ g, int G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  ptg_grain& r = g[i];
  double actual_pos = (double)(r.refcoord_xpix) + r.dx;
  r.refcoord_xpix = (int32_t)round(actual_pos);
  r.dx = actual_pos - r.refcoord_xpix;
  actual_pos = (double)(r.refcoord_ypix) + r.dy;
  r.refcoord_ypix = (int32_t)round(actual_pos);
  r.dy = actual_pos - r.refcoord_ypix;
  r.refrot += r.drot;
  r.drot = 0.0;
}

hipError_t launch_change_reference(Context& c) {
  if (c.G == 0) return hipSuccess;
  k_change_reference<<<(c.G + 255) / 256, 256, 0, c.stream>>>(c.d_grains, c.G);
  return hipGetLastError();
}

// do_precomputations, src/p-tracker.cpp:624-656: one workgroup of 4 waves per grain. The
// sum of I0 is a sum of integers, exact in any order, so mean0 matches the reference's
// bits. C0C0 is accumulated by one lane in the reference's pixel order: waves 1-3 stage
// chunk c of the differences in LDS while lane 0 of wave 0 sums chunk c - 1 (double
// buffered, 16 values per read batch). Also stores d0[i] = I0(p0_i) - mean0 (the
// reference's `(double)image[...] - mean0`, :708), the per-grain sum of d0 for the one-pass
// subpixel objective, and two packed words per pattern pixel (offsets as int8, I0 as u16):
// unrotated for the exact search, rotated (p0) for the subpixel stage.
constexpr int kPreThreads = 256;
constexpr int kPreChunk = 1024;

__global__ __launch_bounds__(kPreThreads) void k_precompute(const FrameDesc* __restrict__ frames,
                                                            ptg_grain* __restrict__ grains,
                                                            const int32_t* __restrict__ pattern,
                                                            const int64_t* __restrict__ pbegin,
                                                            const double* __restrict__ cs0, int32_t* __restrict__ p0,
                                                            double* __restrict__ d0, double* __restrict__ sd0,
                                                            uint32_t* __restrict__ pwords,
                                                            uint32_t* __restrict__ spwords, double* __restrict__ m0,
                                                            int dev_trig, int guard_all,
                                                            RotGuardEntry* __restrict__ guard,
                                                            int32_t* __restrict__ guard_count, int do_chain) {
  __shared__ __align__(16) double chunk[2][kPreChunk];
  __shared__ long long part[kPreThreads / 64];
  __shared__ int near_flag;
  const int g = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  ptg_grain& gr = grains[g];
  const int rx = gr.refcoord_xpix, ry = gr.refcoord_ypix;
  const FrameDesc fr = *frames;
  const uint16_t* __restrict__ I0 = fr.I0;
  const int dimx = fr.dimx0, dimy = fr.dimy0;
  const int64_t pitch = fr.pitch0;
  // cos/sin(refrot) (:636-637): the device library's, guarded as the rotation tables are,
  // or the host's
  const double c0 = dev_trig ? cos(gr.refrot) : cs0[2 * g], s0 = dev_trig ? sin(gr.refrot) : cs0[2 * g + 1];
  if (tid == 0) near_flag = dev_trig && guard_all;
  __syncthreads();
  bool near = false;
  long long s = 0;
  bool clamped = false;
  // four pixels per thread and pass, their loads issued together
  constexpr int kU = 4;
  for (int i0 = tid; i0 < n; i0 += kU * kPreThreads) {
    int px[kU], py[kU], x[kU], y[kU];
    uint32_t v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int i = min(i0 + u * kPreThreads, n - 1);
      px[u] = pattern[2 * (b + i)];
      py[u] = pattern[2 * (b + i) + 1];
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      if (dev_trig && rot_near_half(px[u], py[u], c0, s0)) near = true;
      rotate_px(px[u], py[u], c0, s0, x[u], y[u]);
      int X = rx + x[u], Y = ry + y[u];
      if (i0 + u * kPreThreads < n && (X < 0 || Y < 0 || X >= dimx || Y >= dimy)) clamped = true;
      X = clampi(X, 0, dimx - 1);
      Y = clampi(Y, 0, dimy - 1);
      v[u] = I0[(int64_t)Y * pitch + X];
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int i = i0 + u * kPreThreads;
      if (i < n) {
        p0[2 * (b + i)] = x[u];
        p0[2 * (b + i) + 1] = y[u];
        s += v[u];
        // the exact search's compact copy: offsets as int8 (its patterns are well inside
        // +-127) and the reference value
        pwords[b + i] = ((uint32_t)px[u] & 0xffu) | (((uint32_t)py[u] & 0xffu) << 8) | (v[u] << 16);
        // the subpixel stage's: the rotated offsets (int8 when the pattern radius is at
        // most 127, the only case it reads them) and the same value, from which
        // d0 = v - m0 is re-formed bit for bit
        spwords[b + i] = ((uint32_t)x[u] & 0xffu) | (((uint32_t)y[u] & 0xffu) << 8) | (v[u] << 16);
      }
    }
  }
  if (clamped) atomicOr(&gr.status, (int)PTG_STATUS_CLAMPED);
  if (near) near_flag = 1;  // (every writer stores the same value)
  s = wave_sum_i64(s);
  if (lane == 0) part[wave] = s;
  __syncthreads();  // also makes every thread's packed words visible to the whole workgroup
  long long tot = 0;
#pragma unroll
  for (int w = 0; w < kPreThreads / 64; w++) tot += part[w];
  const double mean0 = (double)tot / (double)n;
  if (!do_chain) {
    // the ordered sums come from k_precompute_chain, launched where they can overlap other work
    if (tid == 0) {
      gr.mean0 = mean0;
      m0[g] = mean0;
      if (near_flag) rot_guard_push(guard, guard_count, g, -1, gr.refrot, c0, s0);
    }
    return;
  }
  double C0C0 = 0.0, sd = 0.0;
  const int nch = (n + kPreChunk - 1) / kPreChunk;
  for (int ch = 0; ch <= nch; ch++) {
    if (wave > 0 && ch < nch) {
      // stage chunk ch: d0 from the packed words' I0
      const int base = ch * kPreChunk, m = min(kPreChunk, n - base);
      double* dst = chunk[ch & 1];
      for (int t = tid - 64; t < m; t += kPreThreads - 64) {
        const double diff = (double)(spwords[b + base + t] >> 16) - mean0;
        dst[t] = diff;
        d0[b + base + t] = diff;
      }
    }
    if (tid == 0 && ch > 0) {
      // sum chunk ch - 1 in pixel order
      const int base = (ch - 1) * kPreChunk, m = min(kPreChunk, n - base);
      const double2* src = reinterpret_cast<const double2*>(chunk[(ch - 1) & 1]);
      int t = 0;
      for (; t + 16 <= m; t += 16) {
        double2 q[8];
#pragma unroll
        for (int j = 0; j < 8; j++) q[j] = src[t / 2 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          C0C0 += q[j].x * q[j].x;
          sd += q[j].x;
          C0C0 += q[j].y * q[j].y;
          sd += q[j].y;
        }
      }
      const double* srcd = chunk[(ch - 1) & 1];
      for (; t < m; t++) {
        const double diff = srcd[t];
        C0C0 += diff * diff;
        sd += diff;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    gr.mean0 = mean0;
    gr.C0C0 = C0C0;
    sd0[g] = sd;
    m0[g] = mean0;
    if (near_flag) rot_guard_push(guard, guard_count, g, -1, gr.refrot, c0, s0);
  }
}

hipError_t launch_precompute(Context& c, bool dev_trig, bool chain) {
  if (c.G == 0) return hipSuccess;
  k_precompute<<<c.G, kPreThreads, 0, c.stream>>>(c.d_frames, c.d_grains, c.d_pattern,
                                                  c.d_pbegin, c.d_cs0, c.d_p0, c.d_d0, c.d_sd0, c.d_pwords,
                                                  c.d_spwords, c.d_m0, dev_trig ? 1 : 0,
                                                  c.rot_guard_test ? 1 : 0, c.d_guard, c.d_guard_count,
                                                  chain ? 1 : 0);
  return hipGetLastError();
}

// The ordered sums of do_precomputations (:645-652) after k_precompute without them: one wave
// per grain (four per workgroup, so that few wave slots are held while other kernels run):
// its lanes stage a chunk of d0 = I0 - mean0 in LDS (and store d0), then lane 0 adds the
// chunk's squares and values in the reference's pixel order.
constexpr int kChainChunk = 1024;

__global__ __launch_bounds__(256) void k_precompute_chain(ptg_grain* __restrict__ grains,
                                                          const int64_t* __restrict__ pbegin,
                                                          const uint32_t* __restrict__ spwords,
                                                          const double* __restrict__ m0, double* __restrict__ d0,
                                                          double* __restrict__ sd0, int G) {
  __shared__ __align__(16) double buf[4][kChainChunk];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + wave;
  if (g >= G) return;  // (no workgroup barrier below)
  const int64_t b = pbegin[g];
  const int n = (int)(pbegin[g + 1] - b);
  const double mean0 = m0[g];
  double* cb = buf[wave];
  double C0C0 = 0.0, sd = 0.0;
  for (int base = 0; base < n; base += kChainChunk) {
    const int m = min(kChainChunk, n - base);
    constexpr int kB = 8;  // words in flight per lane before their stores
    for (int t0 = lane; t0 < m; t0 += 64 * kB) {
      uint32_t w[kB];
#pragma unroll
      for (int j = 0; j < kB; j++) w[j] = t0 + 64 * j < m ? spwords[b + base + t0 + 64 * j] : 0u;
#pragma unroll
      for (int j = 0; j < kB; j++) {
        const int t = t0 + 64 * j;
        if (t < m) {
          const double diff = (double)(w[j] >> 16) - mean0;
          cb[t] = diff;
          d0[b + base + t] = diff;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) {
      const double2* src = reinterpret_cast<const double2*>(cb);
      int t = 0;
      for (; t + 16 <= m; t += 16) {
        double2 q[8];
#pragma unroll
        for (int j = 0; j < 8; j++) q[j] = src[t / 2 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          C0C0 += q[j].x * q[j].x;
          sd += q[j].x;
          C0C0 += q[j].y * q[j].y;
          sd += q[j].y;
        }
      }
      for (; t < m; t++) {
        const double diff = cb[t];
        C0C0 += diff * diff;
        sd += diff;
      }
    }
    // the next chunk overwrites the buffer after lane 0 has read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane == 0) {
    grains[g].C0C0 = C0C0;
    sd0[g] = sd;
  }
}

hipError_t launch_precompute_chain(Context& c, hipStream_t st) {
  if (c.G == 0) return hipSuccess;
  k_precompute_chain<<<(c.G + 3) / 4, 256, 0, st>>>(c.d_grains, c.d_pbegin, c.d_spwords, c.d_m0, c.d_d0,
                                                     c.d_sd0, c.G);
  return hipGetLastError();
}

// The pair's frame record (FrameDesc), written in stream order.
__global__ void k_set_frames(FrameDesc* d, FrameDesc f) {
  if (threadIdx.x == 0) *d = f;
}

hipError_t launch_set_frames(Context& c, const FrameDesc& f) {
  k_set_frames<<<1, 64, 0, c.stream>>>(c.d_frames, f);
  return hipGetLastError();
}

// Rotation tables of a stage on the device: cos and sin of refrot + drot + step_k, summed left to
// right as src/p-tracker.cpp:687-689 does, from the device library. One workgroup per (list
// entry, block of kRotSteps steps): the first lanes form the block's values, then every lane
// checks its pattern pixels against them (rot_near_half) and a step with a pixel near a rounding
// boundary goes on the guard list.
constexpr int kRotThreads = 256;
constexpr int kRotSteps = 8;  // (a block of 32 steps per workgroup measured slower: 65 against 51 us per pair)

__global__ __launch_bounds__(kRotThreads) void k_rot_tables(RotArgs r) {
  __shared__ double tc[kRotSteps], ts[kRotSteps], tth[kRotSteps];
  __shared__ int flag[kRotSteps];
  const int li = blockIdx.x;
  if (li >= *r.count) return;
  const int g = r.list[li];
  const int k0 = blockIdx.y * kRotSteps;
  const int nk = min(kRotSteps, r.nsteps - k0);
  const int t = threadIdx.x;
  if (t < nk) {
    const ptg_grain& gr = r.grains[g];
    const double th = gr.refrot + gr.drot + r.steps[k0 + t];
    const double c = cos(th), s = sin(th);
    tc[t] = c;
    ts[t] = s;
    tth[t] = th;
    flag[t] = r.all;
    r.cs[2 * ((int64_t)g * r.nsteps + k0 + t)] = c;
    r.cs[2 * ((int64_t)g * r.nsteps + k0 + t) + 1] = s;
  }
  __syncthreads();
  const int64_t b = r.pbegin[g];
  const int n = (int)(r.pbegin[g + 1] - b);
  unsigned int near = 0u;  // bit j: step k0 + j has a pixel near a boundary
  for (int i = t; i < n; i += kRotThreads) {
    const int dx = r.pattern[2 * (b + i)], dy = r.pattern[2 * (b + i) + 1];
    for (int j = 0; j < nk; j++)
      if (rot_near_half(dx, dy, tc[j], ts[j])) near |= 1u << j;
  }
  if (near)
    for (int j = 0; j < nk; j++)
      if ((near >> j) & 1) flag[j] = 1;  // (every writer stores the same value)
  __syncthreads();
  if (t < nk && flag[t]) rot_guard_push(r.guard, r.guard_count, g, k0 + t, tth[t], tc[t], ts[t]);
}

hipError_t launch_rot_tables(Context& c, const RotArgs& r, int list_cap) {
  if (list_cap == 0 || r.nsteps == 0) return hipSuccess;
  k_rot_tables<<<dim3(list_cap, (r.nsteps + kRotSteps - 1) / kRotSteps), kRotThreads, 0, c.stream>>>(r);
  return hipGetLastError();
}

// Candidate order of the reference: rotation step, then u (outer), then v (inner),
// src/p-tracker.cpp:686,695-696. Within a step, idx = (u + left) * V + (v + up).
__device__ __forceinline__ bool better(double ncc, int idx, double best, int bidx) {
  // strict improvement as `best_NCC < NCC_test` (:718); among equal values the lower
  // index (earlier in the reference loop) wins; NaN and values <= 0 never win
  return ncc > best || (ncc == best && ncc > 0.0 && idx < bidx);
}

constexpr int kPlainThreads = 256;
constexpr int kPlainCPT = 4;  // candidates per lane per round

__device__ __forceinline__ void plain_argmax_write(double best, int bidx, int k, int S, StepBest* out);

// One (list slot li, step k): the workgroup's best candidate to out.
__device__ __forceinline__ void plain_slot_step(const SearchArgs& a, int li, int k, StepBest* out) {
  const FrameDesc fr = *a.fr;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nsteps = a.nsteps;
  const int g = a.list[li];
  const ptg_grain& gr = a.grains[g];
  const int64_t b = a.pbegin[g];
  const int n = (int)(a.pbegin[g + 1] - b);
  const int rho = a.prad[g];
  const int gdx = (int)round(gr.dx), gdy = (int)round(gr.dy);
  const int U = a.left + a.right + 1, V = a.up + a.down + 1;
  // shifts in tiles of TU x TV (one tile unless the window exceeds the LDS); the window
  // of a tile is (tu + 2 rho) x (tv + 2 rho) cells stored with row stride ww
  const int TU = min(a.tile_u, U), TV = min(a.tile_v, V);
  const int ww = TU + 2 * rho;
  const int X0 = gr.refcoord_xpix + gdx - a.left - rho;
  const int Y0 = gr.refcoord_ypix + gdy - a.up - rho;
  const double C0C0 = gr.C0C0;

  uint16_t* win = reinterpret_cast<uint16_t*>(smem);
  int32_t* off = reinterpret_cast<int32_t*>(smem + ((size_t)a.win_cap * 2 + 15) / 16 * 16);

  const double c = a.cs[2 * ((int64_t)g * nsteps + k)];
  const double s = a.cs[2 * ((int64_t)g * nsteps + k) + 1];
  for (int i = threadIdx.x; i < n; i += kPlainThreads) {
    int x, y;
    rotate_px(a.pattern[2 * (b + i)], a.pattern[2 * (b + i) + 1], c, s, x, y);
    off[i] = (y + rho) * ww + (x + rho);
  }

  const double* __restrict__ d0 = a.d0 + b;
  const double dn = (double)n;
  double best = 0.0;
  int bidx = INT_MAX;
  bool clamped = false;
  for (int tu0 = 0; tu0 < U; tu0 += TU) {
    for (int tv0 = 0; tv0 < V; tv0 += TV) {
      const int tu = min(TU, U - tu0), tv = min(TV, V - tv0);
      const int tw = tu + 2 * rho, th = tv + 2 * rho;
      if (tu0 + tv0 > 0) __syncthreads();  // every lane is done with the previous tile
      for (int t = threadIdx.x; t < tw * th; t += kPlainThreads) {
        const int wy = t / tw, wx = t - wy * tw;
        int X = X0 + tu0 + wx, Y = Y0 + tv0 + wy;
        if (X < 0 || Y < 0 || X >= fr.dimx || Y >= fr.dimy) clamped = true;
        X = clampi(X, 0, fr.dimx - 1);
        Y = clampi(Y, 0, fr.dimy - 1);
        win[wy * ww + wx] = fr.I1[(int64_t)Y * fr.pitch1 + X];
      }
      __syncthreads();
      const int St = tu * tv;
      for (int r0 = 0; r0 < St; r0 += kPlainThreads * kPlainCPT) {
        int base[kPlainCPT], cidx[kPlainCPT];
        bool valid[kPlainCPT];
#pragma unroll
        for (int q = 0; q < kPlainCPT; q++) {
          // work item w walks u fastest so neighbouring lanes read neighbouring LDS words
          const int w = r0 + q * kPlainThreads + threadIdx.x;
          valid[q] = w < St;
          const int ww_ = valid[q] ? w : 0;
          const int v = ww_ / tu, u = ww_ - v * tu;
          base[q] = v * ww + u;
          cidx[q] = (tu0 + u) * V + (tv0 + v);
        }
        unsigned int isum[kPlainCPT];
#pragma unroll
        for (int q = 0; q < kPlainCPT; q++) isum[q] = 0u;
        // four pattern pixels per iteration: gathers in flight (measured 51.7 -> 39.8 ms on the
        // 2,000-grain 21-step search against the compiler's choice)
#pragma unroll 4
        for (int i = 0; i < n; i++) {
          const int o = off[i];
#pragma unroll
          for (int q = 0; q < kPlainCPT; q++) isum[q] += win[o + base[q]];
        }
        // mean1 += (double)I1 over integers is exact in any order, then one division (:705)
        double m1[kPlainCPT], c01[kPlainCPT], c11[kPlainCPT];
#pragma unroll
        for (int q = 0; q < kPlainCPT; q++) {
          m1[q] = (double)isum[q] / dn;
          c01[q] = 0.0;
          c11[q] = 0.0;
        }
#pragma unroll 4
        for (int i = 0; i < n; i++) {
          const int o = off[i];
          const double dd = d0[i];
#pragma unroll
          for (int q = 0; q < kPlainCPT; q++) {
            const double diffC1 = (double)win[o + base[q]] - m1[q];
            c01[q] = c01[q] + dd * diffC1;
            c11[q] = c11[q] + diffC1 * diffC1;
          }
        }
#pragma unroll
        for (int q = 0; q < kPlainCPT; q++) {
          if (!valid[q]) continue;
          const double ncc = c01[q] / sqrt(C0C0 * c11[q]);
          if (better(ncc, cidx[q], best, bidx)) {
            best = ncc;
            bidx = cidx[q];
          }
        }
      }
    }
  }
  if (clamped) atomicOr(&a.grains[g].status, (int)PTG_STATUS_CLAMPED);
  plain_argmax_write(best, bidx, k, U * V, out);
}

// The workgroup's argmax of its lanes' (best, bidx), in the reference's order, written to
// out with the global candidate index k S + bidx.
__device__ __forceinline__ void plain_argmax_write(double best, int bidx, int k, int S, StepBest* out) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double ob = __shfl_xor(best, m, 64);
    const int oi = __shfl_xor(bidx, m, 64);
    if (better(ob, oi, best, bidx)) {
      best = ob;
      bidx = oi;
    }
  }
  __shared__ double wb[kPlainThreads / 64];
  __shared__ int wi[kPlainThreads / 64];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wb[wave] = best;
    wi[wave] = bidx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPlainThreads / 64; w++)
      if (better(wb[w], wi[w], best, bidx)) {
        best = wb[w];
        bidx = wi[w];
      }
    StepBest sb;
    sb.ncc = best;
    sb.idx = bidx == INT_MAX ? INT_MAX : k * S + bidx;  // global candidate index
    sb.pad = 0;
    *out = sb;
  }
}

// The same candidates, in the same order and arithmetic, for a pattern whose rotated offsets
// and window exceed the LDS (radius above about 85 px): each lane rotates the pattern pixels
// itself and reads the current frame through the caches, clamped at its edges as the staged
// window is; the grain is marked clamped when the window's box leaves the frame, as there.
__device__ __forceinline__ void plain_slot_step_global(const SearchArgs& a, int li, int k, StepBest* out) {
  const FrameDesc fr = *a.fr;
  const int nsteps = a.nsteps;
  const int g = a.list[li];
  const ptg_grain& gr = a.grains[g];
  const int64_t b = a.pbegin[g];
  const int n = (int)(a.pbegin[g + 1] - b);
  const int rho = a.prad[g];
  const int gdx = (int)round(gr.dx), gdy = (int)round(gr.dy);
  const int U = a.left + a.right + 1, V = a.up + a.down + 1;
  const int X0 = gr.refcoord_xpix + gdx - a.left, Y0 = gr.refcoord_ypix + gdy - a.up;  // shift (0, 0)
  const double C0C0 = gr.C0C0;
  const double c = a.cs[2 * ((int64_t)g * nsteps + k)];
  const double s = a.cs[2 * ((int64_t)g * nsteps + k) + 1];
  const int32_t* __restrict__ pat = a.pattern + 2 * b;
  const double* __restrict__ d0 = a.d0 + b;
  const double dn = (double)n;
  auto px = [&](int X, int Y) -> unsigned int {
    return fr.I1[(int64_t)clampi(Y, 0, fr.dimy - 1) * fr.pitch1 + clampi(X, 0, fr.dimx - 1)];
  };
  double best = 0.0;
  int bidx = INT_MAX;
  const int S = U * V;
  for (int r0 = 0; r0 < S; r0 += kPlainThreads * kPlainCPT) {
    int cx[kPlainCPT], cy[kPlainCPT], cidx[kPlainCPT];
    bool valid[kPlainCPT];
#pragma unroll
    for (int q = 0; q < kPlainCPT; q++) {
      const int w = r0 + q * kPlainThreads + threadIdx.x;
      valid[q] = w < S;
      const int ww_ = valid[q] ? w : 0;
      const int v = ww_ / U, u = ww_ - v * U;
      cx[q] = X0 + u;
      cy[q] = Y0 + v;
      cidx[q] = u * V + v;
    }
    unsigned int isum[kPlainCPT];
#pragma unroll
    for (int q = 0; q < kPlainCPT; q++) isum[q] = 0u;
    for (int i = 0; i < n; i++) {
      int x, y;
      rotate_px(pat[2 * i], pat[2 * i + 1], c, s, x, y);
#pragma unroll
      for (int q = 0; q < kPlainCPT; q++) isum[q] += px(cx[q] + x, cy[q] + y);
    }
    double m1[kPlainCPT], c01[kPlainCPT], c11[kPlainCPT];
#pragma unroll
    for (int q = 0; q < kPlainCPT; q++) {
      m1[q] = (double)isum[q] / dn;
      c01[q] = 0.0;
      c11[q] = 0.0;
    }
    for (int i = 0; i < n; i++) {
      int x, y;
      rotate_px(pat[2 * i], pat[2 * i + 1], c, s, x, y);
      const double dd = d0[i];
#pragma unroll
      for (int q = 0; q < kPlainCPT; q++) {
        const double diffC1 = (double)px(cx[q] + x, cy[q] + y) - m1[q];
        c01[q] = c01[q] + dd * diffC1;
        c11[q] = c11[q] + diffC1 * diffC1;
      }
    }
#pragma unroll
    for (int q = 0; q < kPlainCPT; q++) {
      if (!valid[q]) continue;
      const double ncc = c01[q] / sqrt(C0C0 * c11[q]);
      if (better(ncc, cidx[q], best, bidx)) {
        best = ncc;
        bidx = cidx[q];
      }
    }
  }
  const int bx = X0 - rho, by = Y0 - rho;
  if (threadIdx.x == 0 && (bx < 0 || by < 0 || bx + U + 2 * rho > fr.dimx || by + V + 2 * rho > fr.dimy))
    atomicOr(&a.grains[g].status, (int)PTG_STATUS_CLAMPED);
  plain_argmax_write(best, bidx, k, S, out);
}

__global__ __launch_bounds__(kPlainThreads) void k_search_plain_global(SearchArgs a) {
  const int li = blockIdx.x / a.nsteps;
  const int k = blockIdx.x - li * a.nsteps;
  if (li >= *a.count) return;
  plain_slot_step_global(a, li, k, &a.best[(int64_t)li * a.nsteps + k]);
}

__global__ __launch_bounds__(kPlainThreads) void k_search_plain(SearchArgs a) {
  const int li = blockIdx.x / a.nsteps;
  const int k = blockIdx.x - li * a.nsteps;
  if (li >= *a.count) return;
  plain_slot_step(a, li, k, &a.best[(int64_t)li * a.nsteps + k]);
}

// The slots sub[0 .. *sub_count - 1] of the list; workgroup w takes step w % nsteps of
// entries w / nsteps, w / nsteps + gridDim.x / nsteps, ...
__global__ __launch_bounds__(kPlainThreads) void k_search_plain_sub(SearchArgs a, const int32_t* sub,
                                                                    const int32_t* sub_count, int best_stride) {
  const int nsteps = a.nsteps;
  const int k = blockIdx.x % nsteps;
  const int n = *sub_count;
  for (int j = blockIdx.x / nsteps; j < n; j += gridDim.x / nsteps) {
    const int li = sub[j];
    plain_slot_step(a, li, k, &a.best[(int64_t)li * best_stride + k]);
    __syncthreads();  // the next entry reuses the window, offsets and reduction slots
  }
}

hipError_t launch_search_plain(Context& c, const SearchArgs& a, int list_cap, size_t lds_bytes) {
  if (list_cap == 0) return hipSuccess;
  const int64_t blocks = (int64_t)list_cap * a.nsteps;
  if (hipError_t e = ensure_kernel_lds(c.device, (const void*)k_search_plain, lds_bytes); e != hipSuccess) return e;
  k_search_plain<<<dim3((unsigned)blocks), kPlainThreads, lds_bytes, c.stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_search_plain_global(Context& c, const SearchArgs& a, int list_cap) {
  if (list_cap == 0) return hipSuccess;
  const int64_t blocks = (int64_t)list_cap * a.nsteps;
  k_search_plain_global<<<dim3((unsigned)blocks), kPlainThreads, 0, c.stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_search_plain_sub(Context& c, const SearchArgs& a, int list_cap, const int32_t* sub,
                                   const int32_t* sub_count, int best_stride, size_t lds_bytes) {
  if (list_cap == 0) return hipSuccess;
  // up to 256 entries in flight; a workgroup with no entry returns at once
  const int64_t blocks = (int64_t)std::min(list_cap, 256) * a.nsteps;
  if (hipError_t e = ensure_kernel_lds(c.device, (const void*)k_search_plain_sub, lds_bytes); e != hipSuccess)
    return e;
  k_search_plain_sub<<<dim3((unsigned)blocks), kPlainThreads, lds_bytes, c.stream>>>(a, sub, sub_count, best_stride);
  return hipGetLastError();
}

// Reduces the partial argmaxes of a slot (global candidate indices, the first maximum
// in the reference's (step, u, v) order wins) and applies the per-stage output rules:
// main src/p-tracker.cpp:729-742, rescue :868-880, super-rescue :1005-1015.
__global__ void k_search_commit(SearchArgs a, int stage, double NCC_min, double NCC_min_super,
                                int32_t* __restrict__ next_list, int32_t* __restrict__ next_count,
                                const StepBest* __restrict__ parts, int nparts,
                                const int32_t* __restrict__ abort_flag) {
  // one wave per slot: its lanes read the slot's partial picks side by side and reduce them;
  // `better` orders candidates by NCC, then by the reference's loop index, and admits no NaN
  // and nothing at or below zero, so the reduction in any order gives the in-order loop's pick
  const int li = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (li >= *a.count) return;
  if (a.act && !a.act[li]) return;  // a slot of the plan that is not in this stage's queue
  // a search that could not be served exactly (capacity flag set) changes nothing
  if (abort_flag && *abort_flag) return;
  double best = 0.0;
  int bidx = INT_MAX;
  for (int p = lane; p < nparts; p += 64) {
    const StepBest sb = parts[(int64_t)li * nparts + p];
    if (better(sb.ncc, sb.idx, best, bidx)) {
      best = sb.ncc;
      bidx = sb.idx;
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double ob = __shfl_xor(best, m, 64);
    const int oi = __shfl_xor(bidx, m, 64);
    if (better(ob, oi, best, bidx)) {
      best = ob;
      bidx = oi;
    }
  }
  if (lane != 0) return;
  const int g = a.list[li];
  ptg_grain& gr = a.grains[g];
  const int U = a.left + a.right + 1, V = a.up + a.down + 1, S = U * V;
  int best_upix = 0, best_vpix = 0;
  double best_drot = 0.0;
  if (bidx != INT_MAX) {
    const int bk = bidx / S, rem = bidx - bk * S;
    best_upix = rem / V - a.left;
    best_vpix = rem % V - a.up;
    best_drot = a.steps[bk];
  }
  const int grain_dx = (int)round(gr.dx), grain_dy = (int)round(gr.dy);
  const double rest_dx = gr.dx - grain_dx, rest_dy = gr.dy - grain_dy;
  bool queued = false, update = false;
  if (stage == PTG_STAGE_MAIN) {
    gr.upix = gr.vpix = gr.rot_inc = 0.0;
    gr.NCC = best;
    gr.NCC_rescue = best;
    if (best < NCC_min) queued = true; else update = true;
  } else if (stage == PTG_STAGE_RESCUE) {
    gr.NCC_rescue = best;
    if (best < NCC_min_super) queued = true; else update = true;
  } else {
    gr.NCC_rescue = best;
    if (best > NCC_min_super) update = true;
  }
  if (update) {
    gr.upix = best_upix - rest_dx;
    gr.vpix = best_vpix - rest_dy;
    gr.rot_inc = best_drot;
    gr.dx = grain_dx + best_upix;
    gr.dy = grain_dy + best_vpix;
    gr.drot += best_drot;
  }
  if (queued && next_list) {
    const int pos = atomicAdd(next_count, 1);
    next_list[pos] = g;
  }
}

hipError_t launch_search_commit(Context& c, const SearchArgs& a, int list_cap, int stage, double NCC_min,
                                double NCC_min_super, int32_t* next_list, int32_t* next_count,
                                const StepBest* parts, int nparts, const int32_t* abort_flag) {
  if (list_cap == 0) return hipSuccess;
  // four slots (waves) per workgroup
  k_search_commit<<<(list_cap + 3) / 4, 256, 0, c.stream>>>(a, stage, NCC_min, NCC_min_super, next_list,
                                                             next_count, parts, nparts, abort_flag);
  return hipGetLastError();
}

}  // namespace ptg
