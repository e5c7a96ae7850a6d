// ptg_search_exact.hip — the exact-integer integer-pixel search of libptgpu on the
// gfx950 matrix cores.
//
// For a grain, rotation step k and shift s the reference's ZNCC needs three sums over
// the rotated pattern (src/p-tracker.cpp:698-716):
//   A = sum a_i I1(s + p_k,i),  B = sum I1(s + p_k,i),  C = sum I1(s + p_k,i)^2
// with a_i = I0 at the reference pattern pixel. As exact integers,
//   NCC = (N A - Sa B) / sqrt((N Saa - Sa^2) (N C - B^2)),
// equal to the reference's two-pass FP64 value up to its rounding.
//
// A is a correlation of the rotated template T_k (pixel values scattered onto their
// rounded offsets, duplicates summed) with the window, computed as a GEMM on
// v_mfma_i32_16x16x64_i8: rows = (step, template digit), columns = vertical shifts of
// one shift column, K = template cells. The operands are int8 digits chosen by the frames'
// declared depth, so every product is exact in the int32 accumulators:
//  - 8-bit frames: I1 as one offset digit b = I1 - 128 (the byte with its top bit flipped),
//    T as two 7-bit digits: A = 128 P_1 + P_0 + 128 Sa (Sa = sum of T): 2 products;
//  - 12-bit frames: I1 and T as two 7-bit digits each: 4 products;
//  - 16-bit frames: I1 as two offset bytes, I1 = 256 (h - 128) + (l - 128) + 32896 (each byte
//    with its top bit flipped), T as three 7-bit digits (T < 2^21): 6 products, and num,
//    d1 formed in int64 (N A reaches 2^61).
//  - Template fragments are built once per search in global memory in fragment order
//    (one 1 KiB block per M-tile and K-chunk) and read coalesced; all shift columns of
//    a grain run on one XCD so its L2 keeps them.
//  - K-chunks are 4 segments of 16 template cells of one row, segment starts aligned to
//    16. For each shift column the window digit planes are copied into LDS rows whose
//    cell c is window column su + c, so every B-fragment is one aligned ds_read_b128.
// B and C use a base multiset mask M_mid: the middle rotation step's, or for a pattern that
// is a full digital disk the pattern itself (a rotation maps the disk onto itself up to a
// few rim cells, so the rounded rotations differ from it only by their own duplicate and
// empty cells: 30-40% fewer Delta cells than against the middle step at base rotations of
// 0.05-0.3 rad, the same at zero). B_mid(s), C_mid(s) come from a second GEMM whose rows
// are copies of M_mid moved by 0..15 cells (k_exact_mid), and each step adds the sparse
// difference M_k - M_mid (a few tens of cells) on the vector units.
#include <limits.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "ptg_device.hpp"
#include "ptg_internal.hpp"

namespace ptg {

typedef int v4i __attribute__((ext_vector_type(4)));

// Per-slot geometry of the exact search (host computed).
struct ExactSlot {
  int32_t rho;       // every rotated offset lies in [-rho, rho]^2
  int32_t cw;        // template row width in cells, 16 * ceil((2 rho + 1) / 16)
  int32_t cs;        // LDS staging row stride (bytes), cw rounded to an odd multiple of 16
  int32_t kc;        // K-chunks (4 segments each)
  int64_t seg_off;   // first segment in the segment table
  int64_t tmpl_off;  // first byte of this slot's template fragments
  int64_t dl_off;    // first Delta entry (nsteps * 2 * dcap entries)
  int64_t mid_off;   // first int64 of (B_mid, C_mid) pairs (S pairs)
  int64_t mseg_off;  // first segment of the base-mask GEMM's segment table (k_exact_mid)
  int64_t mmask_off; // first byte of the base mask's multiplicity grid (side x side)
  int32_t mkc;       // its K-chunks (4 segments each)
  int32_t dcap;      // Delta entries per (step, sign), at most 1023
  int32_t base_pat;  // the base mask is the pattern itself (a full digital disk), not the middle step
  int32_t disk_m2;   // for such a disk: the largest x^2 + y^2 of its cells
};

struct ExactArgs {
  SearchArgs s;
  const ExactSlot* slots;      // per list slot
  const int32_t* bucket;       // list slots of this launch
  int32_t nbucket;             // their number (with nbucket_dev: the most there can be)
  const int32_t* nbucket_dev;  // device count of the launch's slots, or null (nbucket of them)
  const int32_t* seg;          // segment table: (qy + rho) << 16 | c0
  uint8_t* tmpl;               // template fragments
  int32_t* dlist;              // Delta cell offsets: per step dcap cells gained, then dcap cells lost (per slot)
                               // (a cell whose multiplicity changes by 2 appears twice)
  int32_t* dlen;               // per (slot, step): cells gained, cells lost
  int64_t* mid;                // B_mid, C_mid per (slot, shift)
  int64_t* mom;                // per slot: N, Sa, Saa
  int32_t* flags;              // [0]: template or Delta overflow
  ExactBest* xout;             // per (slot, shift block): the block's near-tie list, part_cap entries per slot
  StepBest* out;               // per slot: the pick (entry 0 of part_cap)
  int32_t* cand;               // per slot: the candidates re-scored in the reference's arithmetic (2)
  double thr;                  // the stage's queueing threshold (NCC_min, or NCC_min_super)
  int32_t part_cap;
  int32_t dcap;                // largest slot dcap of the search
  int32_t nmt;                 // M-tiles (vm_spt steps each)
  int32_t nsub;                // shift blocks per grain
  int32_t nsu;                 // shift columns per block (= waves)
  int32_t vp;                  // vertical shifts padded to 16
  int32_t ww, wh_pad;          // plain window width/height of the largest slot of the launch
  int32_t cs_max;              // staging row stride of the largest slot of the launch
  int32_t dl_lds_cap;          // Delta entries staged in LDS per slot (the rest is read from global)
  int32_t seg_cap;             // K-chunks of the largest slot of the launch (segment table in LDS)
  const int32_t* mseg;         // base-mask GEMM segments: (qy << 16) | x0, 0x8000 = empty
  int32_t mid_cbw;             // 16-column blocks per k_exact_mid workgroup
  int32_t mid_stride;          // digit-plane row stride (bytes) of the launch
  int32_t mid_mkc_cap;         // largest mkc of the launch
  int32_t abuf_alias;          // epilogue buffer in the wave's staging area (one vertical pass)
  uint8_t* mmask;              // base-mask multiplicity grids (written by k_exact_mid / k_exact_base_disk)
  int32_t mid_npl;             // digit planes staged per pass of k_exact_mid (6 or 3)
  int32_t mid_skip_disk;       // k_exact_mid leaves the disk-based slots to k_exact_base_disk
  int32_t* ovf_slot;           // per slot: 1 when its template or Delta lists overflowed (or null:
                               // an overflow then raises flags[0] bit 0 and voids the stage)
  int32_t* ovf_list;           // overflowed slots, flags[1] of them, searched by the plain kernel
  int32_t base_vb;             // k_exact_base_disk: vertical shifts per workgroup
  int32_t vmode;               // value mode (ExactMode)
  int32_t vmax;                // its largest window value
  int32_t pad_b;
};

// ---- the near-tie window -------------------------------------------------------------
// The reference's NCC of a candidate (two-pass FP64 sums in pattern order, unfused,
// src/p-tracker.cpp:698-716) differs from its exact value rho = num / sqrt(d0 d1), with
// num = N A - Sa B, d0 = N Saa - Sa^2, d1 = N C - B^2, by at most
//   tau = 2 (2u (R0 + R1) + 7u + 2 g_N) + 8 u^2 (R0 + sqrt N)(R1 + sqrt N),
// u = 2^-53, g_N = N u / (1 - N u), R0 = |mean0| / sigma0 = Sa / sqrt(d0) and
// R1 = |mean1| / sigma1 = B / sqrt(d1): the roundings of the two means (u |mean|), of the
// differences, products and sequential sums (g_N relative to the sums of absolute terms,
// which Cauchy-Schwarz bounds by sqrt(C0C0 C1C1) and sqrt(N C) |mean|), and of the final
// product, square root and division (3u), to first order and doubled, plus a bound of the
// second-order terms. For windows of integers at most vmax, R1 <= vmax N / sqrt(N - 1)
// (the least non-zero N C - B^2 is N - 1). q, formed here in FP64 from the exact
// integers, is within 4u of rho. So the reference orders two candidates whose q differ by
// more than W = 2 (tau + 4u) as q does, and its pick lies among the candidates with
// q >= q* - W. The search keeps that set, per grain, and re-scores it in the reference's
// own arithmetic.
__host__ __device__ inline double ncc_error_bound(double N, double R0, double R1) {
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double gN = N * u / (1.0 - N * u);
  const double sn = sqrt(N);
  return 2.0 * (2.0 * u * (R0 + R1) + 7.0 * u + 2.0 * gN) + 8.0 * u * u * (R0 + sn) * (R1 + sn);
}

// Grains whose best q is below 2 kQmin go to the plain kernel: the window on s = num^2 / d1
// (= q^2 d0) below is relative and holds for q* >= kQmin.
constexpr double kQmin = 1.0 / 64.0;
// Patterns of at most this many pixels have N A, Sa B, N C and B^2 below 2^53 for 12-bit
// values (N^2 4095^2 < 2^53), so the MFMA kernel forms num and d1 exactly in FP64, and below
// 2^62 for 16-bit values (formed in int64); larger ones (radius above about 85 px) go to the
// plain kernel.
constexpr long long kDoubleExactN = 23170;

// Value modes of the exact search, from the frames' declared depth (the larger of the two
// frames'): window values above the mode's vmax send the grain to the plain kernel.
enum ExactMode : int { kVm8 = 0, kVm12 = 1, kVm16 = 2 };
__host__ __device__ constexpr int vm_vmax(int vm) { return vm == kVm8 ? 255 : vm == kVm12 ? 4095 : 65535; }
// template digit rows per rotation step, steps per 16-row M-tile, window digit planes
__host__ __device__ constexpr int vm_digits(int vm) { return vm == kVm16 ? 3 : 2; }
__host__ __device__ constexpr int vm_spt(int vm) { return vm == kVm16 ? 5 : 8; }
__host__ __device__ constexpr int vm_planes(int vm) { return vm == kVm8 ? 1 : 2; }
// the offset the window's digits leave out: A = (digit products) + vm_offset Sa
__host__ __device__ constexpr double vm_offset(int vm) { return vm == kVm8 ? 128.0 : vm == kVm12 ? 0.0 : 32896.0; }

// The relative window on s = num^2 / d1: a candidate with s < s* lo cannot be the reference's
// pick when q* >= kQmin, since (q* - W)^2 / q*^2 >= (1 - W / kQmin)^2; 16u more cover the
// roundings of s (5u each) and of the comparisons. vmax bounds the window values; q is within
// 6u of rho when num and d1 are rounded to FP64 from int64 (16-bit frames), 4u otherwise.
__host__ __device__ inline double near_window_q(long long N, long long Sa, long long d0, int vmax) {
  const double u = 1.1102230246251565e-16;
  const double Nd = (double)N;
  const double R0 = (double)Sa / sqrt((double)d0);
  const double R1 = N > 1 ? (double)vmax * Nd / sqrt(Nd - 1.0) : 0.0;
  return 2.0 * (ncc_error_bound(Nd, R0, R1) + 6.0 * u);
}

__host__ __device__ inline double near_window_lo(long long N, long long Sa, long long d0, int vmax) {
  const double u = 1.1102230246251565e-16;
  const double r = 1.0 - near_window_q(N, Sa, d0, vmax) / kQmin;
  return r * r * (1.0 - 16.0 * u);
}

// A grain's (or a block's, or a lane's) best candidate by s and, when it lies within the
// window of the best, the next one; ovf: a third candidate lay within the window of the
// best of its time. Ranking: larger s first, the earlier candidate (lower index) among
// equal s. An empty entry has s = -1 and index INT_MAX.
__device__ __forceinline__ bool rank_gt(double s, int i, double t, int j) { return s > t || (s == t && i < j); }

struct Near2 {
  double s1, s2;
  int i1, i2, ovf;
};

__device__ __forceinline__ Near2 near2_none() { return Near2{-1.0, -1.0, INT_MAX, INT_MAX, 0}; }

// One candidate into a lane's list.
__device__ __forceinline__ void near2_add(Near2& a, double s, int idx, double lo) {
  if (rank_gt(s, idx, a.s1, a.i1)) {
    if (a.s1 >= s * lo) {
      a.ovf |= a.s2 >= s * lo ? 1 : 0;
      a.s2 = a.s1;
      a.i2 = a.i1;
    } else {
      a.s2 = -1.0;
      a.i2 = INT_MAX;
    }
    a.s1 = s;
    a.i1 = idx;
  } else if (s >= a.s1 * lo) {
    if (a.s2 >= a.s1 * lo) a.ovf = 1;
    if (rank_gt(s, idx, a.s2, a.i2)) {
      a.s2 = s;
      a.i2 = idx;
    }
  }
}

// near2_add for a candidate with numerator nd and d1 = dd, kept out of line (few candidates
// per lane reach it); the list goes in and out by value, in registers.
__device__ __attribute__((noinline)) Near2 near2_insert(Near2 a, double nd, double dd, int idx, double lo) {
  near2_add(a, nd * nd / dd, idx, lo);
  return a;
}

// Two lists into one: the best two of the four entries; a third entry within the window of
// the merged best, or a list that overflowed with its best within that window, overflows.
__device__ __forceinline__ Near2 near2_merge(const Near2& a, const Near2& b, double lo) {
  const double cs[4] = {a.s1, a.s2, b.s1, b.s2};
  const int ci[4] = {a.i1, a.i2, b.i1, b.i2};
  int bi = 0;
#pragma unroll
  for (int q = 1; q < 4; q++)
    if (rank_gt(cs[q], ci[q], cs[bi], ci[bi])) bi = q;
  int si = bi == 0 ? 1 : 0;
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (q != bi && rank_gt(cs[q], ci[q], cs[si], ci[si])) si = q;
  const double thr = cs[bi] * lo;
  int near = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) near += (ci[q] != INT_MAX && cs[q] >= thr) ? 1 : 0;
  Near2 m;
  m.s1 = cs[bi];
  m.i1 = ci[bi];
  m.s2 = cs[si];
  m.i2 = ci[si];
  m.ovf = (near > 2 || (a.ovf && a.s1 >= thr) || (b.ovf && b.s1 >= thr)) ? 1 : 0;
  return m;
}

__device__ __forceinline__ Near2 near2_shfl_xor(const Near2& a, int m) {
  Near2 b;
  b.s1 = __shfl_xor(a.s1, m, 64);
  b.s2 = __shfl_xor(a.s2, m, 64);
  b.i1 = __shfl_xor(a.i1, m, 64);
  b.i2 = __shfl_xor(a.i2, m, 64);
  b.ovf = __shfl_xor(a.ovf, m, 64);
  return b;
}

// Plain-window row length in uint16 elements: U - 1 + cw rounded up to 2 (mod 4), so
// that a row is an odd number of dwords and lanes reading consecutive rows hit
// distinct LDS banks.
__host__ __device__ __forceinline__ int win_width(int U, int cw) {
  const int w = U - 1 + cw;
  return w + ((2 - (w & 3)) & 3);
}

__device__ __forceinline__ int win_x0(const ptg_grain& gr, int left, int rho) {
  return gr.refcoord_xpix + (int)round(gr.dx) - left - rho;
}
__device__ __forceinline__ int win_y0(const ptg_grain& gr, int up, int rho) {
  return gr.refcoord_ypix + (int)round(gr.dy) - up - rho;
}

// Hands list slot li to the plain kernel (its MFMA search returns at once); without that
// fallback (the plain kernel's LDS cannot take the list) the stage is void.
__device__ __forceinline__ void overflow_slot(const ExactArgs& e, int li) {
  if (!e.ovf_slot) atomicOr(&e.flags[0], 1);
  else if (atomicExch(&e.ovf_slot[li], 1) == 0) e.ovf_list[atomicAdd(&e.flags[1], 1)] = li;
}

// Slots of this launch: the device count when the launch serves a rescue stage's queue.
__device__ __forceinline__ int bucket_n(const ExactArgs& e) { return e.nbucket_dev ? *e.nbucket_dev : e.nbucket; }

// Workgroups per grain of the MFMA search when a radius bucket has n grains: enough for about
// 512 workgroups in all, at most one per shift column.
__host__ __device__ __forceinline__ int exact_nsub_of(int n, int U) {
  return n > 0 ? max(1, min(U, (512 + n - 1) / n)) : 1;
}

// ---- template fragments and Delta lists: one workgroup per (slot, step) -------------
// LDS: packed[(2 rho + 1)^2] = sum of a_i | count << 24 for step k; mid[(2 rho + 1)^2]
// bytes = count of the base mask (from k_exact_mid or k_exact_base_disk, which run first). Rows beyond the
// last step, and an M-tile's rows beyond its vm_spt steps, are written as zeros.
__global__ __launch_bounds__(256) void k_exact_template(ExactArgs e) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int bi = blockIdx.x;
  if (bi >= bucket_n(e)) return;
  const int li = e.bucket[bi];
  if (li >= *e.s.count) return;
  const int k = blockIdx.y;  // 0 .. vm_spt * nmt - 1
  const int g = e.s.list[li];
  const ExactSlot sl = e.slots[li];
  const int rho = sl.rho, side = 2 * rho + 1, ncell = side * side;
  const int nsteps = e.s.nsteps;
  int32_t* packed = reinterpret_cast<int32_t*>(smem);
  uint8_t* cmid = smem + ((4 * ncell + 15) & ~15);
  int32_t* dflag = reinterpret_cast<int32_t*>(cmid + ((ncell + 15) & ~15));
  for (int t = threadIdx.x; t < ncell; t += blockDim.x) packed[t] = 0;
  if (threadIdx.x == 0) *dflag = 0;
  const bool real = k < nsteps;
  // the base mask's multiplicities (k_exact_mid or k_exact_base_disk), 16 cells per load
  // (the grids start on 16-byte boundaries and are padded to 16 bytes)
  if (real)
    for (int t = threadIdx.x; t < (ncell + 15) >> 4; t += blockDim.x)
      reinterpret_cast<uint4*>(cmid)[t] = reinterpret_cast<const uint4*>(e.mmask + sl.mmask_off)[t];
  __syncthreads();
  const int64_t b = e.s.pbegin[g];
  const int n = (int)(e.s.pbegin[g + 1] - b);
  // differs: this step's rotated offsets differ, in pattern order, from the previous
  // step's (step 0: always). A step whose list equals the previous one's scores every
  // shift with the reference's NCC bit for bit as that step did, earlier in the reference's
  // loop order, so none of its candidates can win (`best_NCC < NCC_test`,
  // src/p-tracker.cpp:718): the search skips it.
  int differs = k == 0 ? 1 : 0;
  if (real) {
    const double c = e.s.cs[2 * ((int64_t)g * nsteps + k)], s = e.s.cs[2 * ((int64_t)g * nsteps + k) + 1];
    double cp = 0.0, sp = 0.0;
    if (k > 0) {
      cp = e.s.cs[2 * ((int64_t)g * nsteps + k - 1)];
      sp = e.s.cs[2 * ((int64_t)g * nsteps + k - 1) + 1];
    }
    auto scatter = [&](uint32_t w) {
      const int px = (int)(int8_t)(w & 0xffu), py = (int)(int8_t)((w >> 8) & 0xffu);
      int x, y;
      rotate_px(px, py, c, s, x, y);
      atomicAdd(&packed[(y + rho) * side + (x + rho)], (int)(w >> 16) + (1 << 24));
      if (k > 0 && !differs) {
        int xp, yp;
        rotate_px(px, py, cp, sp, xp, yp);
        differs = (xp != x || yp != y) ? 1 : 0;
      }
    };
    // four pattern words in flight per thread before their scatters, then the rest
    const int nthr = blockDim.x;
    int i = threadIdx.x;
    for (; i + 3 * nthr < n; i += 4 * nthr) {
      const uint32_t w0 = e.s.pwords[b + i], w1 = e.s.pwords[b + i + nthr];
      const uint32_t w2 = e.s.pwords[b + i + 2 * nthr], w3 = e.s.pwords[b + i + 3 * nthr];
      scatter(w0);
      scatter(w1);
      scatter(w2);
      scatter(w3);
    }
    for (; i < n; i += nthr) scatter(e.s.pwords[b + i]);
  }
  // the block-wide OR through a word after the two grids (__syncthreads_or would reserve
  // static LDS)
  if (differs) *dflag = 1;
  __syncthreads();
  differs = *dflag;
  // fragments: rows nd j + d of M-tile k / spt (j = k mod spt) hold digit nd - 1 - d of the
  // template (7-bit digits, most significant first); the M-tile's last step also writes its
  // padding rows (16 - nd spt of them) as zeros
  const int nd = vm_digits(e.vmode), spt = vm_spt(e.vmode);
  const int mt = k / spt, kj = k - mt * spt, m = nd * kj;
  const int nr = nd + (kj == spt - 1 ? 16 - nd * spt : 0);
  const int vlim = 1 << (7 * nd);
  const int kc = sl.kc;
  bool overflow = false;
  for (int t = threadIdx.x; t < kc * 4 * nr; t += blockDim.x) {
    const int d = t % nr, r4 = t / nr, gg = r4 & 3, ch = r4 >> 2;
    const int segv = e.seg[sl.seg_off + 4 * ch + gg];
    const int qyr = segv >> 16, c0 = segv & 0x7fff;  // qy + rho, first cell column
    const bool empty = (segv & 0x8000) != 0;          // padding segment: zero template bytes
    unsigned int w[4] = {0u, 0u, 0u, 0u};
    if (real && !empty && d < nd) {
      const int sh = 7 * (nd - 1 - d);
      for (int j = 0; j < 16; j++) {
        const int col = c0 + j;  // qx + rho
        if (col < side) {
          const int v = packed[qyr * side + col] & 0xffffff;
          if (v >= vlim) overflow = true;
          const unsigned int dig = (unsigned)(v >> sh) & 127u;
          w[j >> 2] |= dig << (8 * (j & 3));
        }
      }
    }
    const int lane = m + d + 16 * gg;
    uint4 val = make_uint4(w[0], w[1], w[2], w[3]);
    *reinterpret_cast<uint4*>(e.tmpl + sl.tmpl_off + (((int64_t)mt * kc + ch) * 64 + lane) * 16) = val;
  }
  if (real && !differs) {
    // a repeated step: its Delta lists are marked absent (-1) and the search skips it
    if (threadIdx.x == 0) {
      e.dlen[2 * ((int64_t)li * nsteps + k)] = -1;
      e.dlen[2 * ((int64_t)li * nsteps + k) + 1] = -1;
    }
  } else if (real) {
    // Delta lists: cells whose multiplicity differs from the base mask's, split by sign
    __shared__ int cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    int32_t* dl = e.dlist + sl.dl_off + (int64_t)k * 2 * sl.dcap;
    const int ww = win_width(e.s.left + e.s.right + 1, sl.cw);  // plain window width of this slot
    for (int t = threadIdx.x; t < ncell; t += blockDim.x) {
        const int d = ((packed[t] >> 24) & 0xff) - (int)cmid[t];
      if (d != 0) {
        const int sg = d > 0 ? 0 : 1, m = d > 0 ? d : -d;
        const int pos = atomicAdd(&cnt[sg], m);
        const int qy = t / side, qx = t - qy * side;  // both + rho
        for (int r = 0; r < m; r++)
          if (pos + r < sl.dcap) dl[sg * sl.dcap + pos + r] = qy * ww + qx;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (cnt[0] > sl.dcap || cnt[1] > sl.dcap) overflow = true;
      e.dlen[2 * ((int64_t)li * nsteps + k)] = min(cnt[0], sl.dcap);
      e.dlen[2 * ((int64_t)li * nsteps + k) + 1] = min(cnt[1], sl.dcap);
    }
  }
  if (overflow) overflow_slot(e, li);
}

// ---- B_mid, C_mid for every shift and the reference moments -------------------------
// For a block of 16 shift columns su0 + m (m < 16), B_mid and C_mid of every vertical shift
// come from one GEMM on v_mfma_i32_16x16x64_i8: row m of A is the base multiset
// mask moved right by m cells (A_m[qy][x] = M_mid(x - m, qy)), columns are vertical shifts,
// K runs over (qy, x) in 16-cell segments, and the B operands are the window's digit planes
// starting at column su0: I1 = 128 h + l and I1^2 = sum_d 128^d q_d (7-bit digits), so
//   B_mid = 128 acc_h + acc_l,  C_mid = sum_d 128^d acc_qd,
// every product at most 4 * 127 and every sum exact in int32. For 16-bit frames the planes
// are the bytes of I1 and I1^2 with their top bits flipped (byte - 128, an int8), so
//   B_mid = 256 (acc_1 + 128 n) + (acc_0 + 128 n),  C_mid = sum_d 256^d (acc_qd + 128 n),
// n = the sum of the mask's multiplicities (the pattern's pixel count).
// One workgroup per (slot, column block); wave w takes N-tiles w, w + 4, ...
// LDS: NPL digit planes [WHm rows x stride] | mask rows [side x MLR] bytes (16 zero bytes on
// the left) | segment table. Large patterns stage the six planes in two passes; each
// pass adds its planes' share of B_mid and C_mid to the values the same lane wrote before.
constexpr int kMidPlanes = 6;
// Dynamic LDS a workgroup may take: the CU's 160 KiB less room for the kernels' static
// arrays.
constexpr size_t kLdsDynMax = 159 * 1024;

// Row stride of the digit planes for cbw column blocks: an odd multiple of 16 bytes that
// holds columns su0 .. su0 + 16 cbw + 2 rho + 15.
__host__ __device__ __forceinline__ int mid_stride(int rho, int cbw) {
  const int w = (16 * cbw + 2 * rho + 16 + 15) / 16 * 16;
  return ((w / 16) % 2 == 1) ? w : w + 16;
}
__host__ __device__ __forceinline__ int mid_mask_row(int rho) { return (2 * rho + 1 + 50 + 15) / 16 * 16; }
__host__ __device__ __forceinline__ size_t mid_lds_bytes(int rho, int vp, int cbw, int mkc, int npl) {
  const size_t whm = vp + 2 * rho;
  return (size_t)npl * whm * mid_stride(rho, cbw) + (size_t)(2 * rho + 1) * mid_mask_row(rho) +
         (size_t)16 * mkc + 64;
}

template <int NPL>
__global__ __launch_bounds__(1024) void k_exact_mid(ExactArgs e) {
  const FrameDesc fr = *e.s.fr;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ long long red[2][16];
  const int bi = blockIdx.x;
  if (bi >= bucket_n(e)) return;
  const int li = e.bucket[bi];
  if (li >= *e.s.count) return;
  const int g = e.s.list[li];
  const ExactSlot sl = e.slots[li];
  if (sl.base_pat && e.mid_skip_disk) return;  // k_exact_base_disk's
  const int rho = sl.rho, side = 2 * rho + 1;
  const int U = e.s.left + e.s.right + 1, V = e.s.up + e.s.down + 1;
  const int vp = e.vp, WHm = vp + 2 * rho;
  const int stride = e.mid_stride, MLR = mid_mask_row(rho);
  const int cbw = e.mid_cbw;
  const int su0 = 16 * cbw * blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  uint8_t* planes = smem;
  uint8_t* mrow = smem + (size_t)NPL * WHm * stride;
  int32_t* segt = reinterpret_cast<int32_t*>(mrow + (size_t)side * MLR);
  const ptg_grain& gr = e.s.grains[g];
  const int64_t b = e.s.pbegin[g];
  const int n = (int)(e.s.pbegin[g + 1] - b);
  const int nsteps = e.s.nsteps, kmid = nsteps / 2;
  // the base mask (bytes) and, in the first column block, the reference moments
  for (int t = threadIdx.x; t < side * MLR / 4; t += blockDim.x) reinterpret_cast<uint32_t*>(mrow)[t] = 0u;
  __syncthreads();
  const double c = e.s.cs[2 * ((int64_t)g * nsteps + kmid)], s = e.s.cs[2 * ((int64_t)g * nsteps + kmid) + 1];
  long long Sa = 0, Saa = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t pw = e.s.pwords[b + i];
    const int px = (int)(int8_t)(pw & 0xffu), py = (int)(int8_t)((pw >> 8) & 0xffu);
    int x = px, y = py;
    if (!sl.base_pat) rotate_px(px, py, c, s, x, y);
    // byte counters in 32-bit words; a count above 127 (not an int8 operand) raises the
    // capacity flag, which voids the search before a byte could carry into its neighbour
    const int off = (y + rho) * MLR + 16 + (x + rho);
    const unsigned int old = atomicAdd(reinterpret_cast<unsigned int*>(mrow + (off & ~3)), 1u << (8 * (off & 3)));
    if (((old >> (8 * (off & 3))) & 0xffu) >= 127u) atomicOr(&e.flags[0], 8);
    if (blockIdx.y == 0) {
      const long long a = pw >> 16;
      Sa += a;
      Saa += a * a;
    }
  }
  if (blockIdx.y == 0) {
    Sa = wave_sum_i64(Sa);
    Saa = wave_sum_i64(Saa);
    if (lane == 0) {
      red[0][wave] = Sa;
      red[1][wave] = Saa;
    }
  }
  // this slot's segment table (host built)
  const int nch = sl.mkc;
  for (int q = threadIdx.x; q < 4 * nch; q += blockDim.x) segt[q] = e.mseg[sl.mseg_off + q];
  const int X0 = gr.refcoord_xpix + (int)round(gr.dx) - e.s.left - rho + su0;
  const int Y0 = gr.refcoord_ypix + (int)round(gr.dy) - e.s.up - rho;
  const size_t psz = (size_t)WHm * stride;
  const int m = lane & 15, gq = lane >> 4;
  const int ntile = vp >> 4;
  __shared__ int big;
  if (threadIdx.x == 0) big = 0;
  const bool bytes = e.vmode == kVm16;
  const uint32_t vmax = (uint32_t)e.vmax;
  for (int p0 = 0; p0 < kMidPlanes; p0 += NPL) {
    // the window's digit planes p0 .. p0 + NPL - 1 for columns su0 .. su0 + stride - 1
    if (p0 > 0) __syncthreads();  // every wave is done with the previous pass's planes
    bool over = false;
    for (int t = threadIdx.x; t < WHm * (stride / 4); t += blockDim.x) {
      const int y = t / (stride / 4), x4 = (t - y * (stride / 4)) * 4;
      const int Y = clampi(Y0 + y, 0, fr.dimy - 1);
      uint32_t w[kMidPlanes] = {0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t v = fr.I1[(int64_t)Y * fr.pitch1 + clampi(X0 + x4 + j, 0, fr.dimx - 1)];
        over |= v > vmax;
        const uint32_t q = v * v;
        uint32_t dig[kMidPlanes];
        if (bytes) {
          dig[0] = ((v >> 8) & 255u) ^ 128u;
          dig[1] = (v & 255u) ^ 128u;
#pragma unroll
          for (int d = 0; d < 4; d++) dig[2 + d] = ((q >> (8 * d)) & 255u) ^ 128u;
        } else {
          dig[0] = (v >> 7) & 127u;
          dig[1] = v & 127u;
#pragma unroll
          for (int d = 0; d < 4; d++) dig[2 + d] = (q >> (7 * d)) & 127u;
        }
#pragma unroll
        for (int p = 0; p < kMidPlanes; p++) w[p] |= dig[p] << (8 * j);
      }
#pragma unroll
      for (int p = 0; p < NPL; p++)
        *reinterpret_cast<uint32_t*>(planes + p * psz + (size_t)y * stride + x4) = w[p0 + p];
    }
    if (over) big = 1;
    __syncthreads();
    if (big) {
      // a window value beyond the declared depth (the frame's data is not checked against
      // it): the digits would drop bits, so the plain kernel takes this grain
      if (threadIdx.x == 0) overflow_slot(e, li);
      return;
    }
    if (p0 == 0 && blockIdx.y == 0) {
      // the base mask's multiplicities for the template kernel's Delta lists
      for (int t = threadIdx.x; t < side * side; t += blockDim.x) {
        const int qy = t / side, qx = t - qy * side;
        e.mmask[sl.mmask_off + t] = mrow[(size_t)qy * MLR + 16 + qx];
      }
      if (threadIdx.x == 0) {
        long long A = 0, AA = 0;
        for (int w = 0; w < nwave; w++) {
          A += red[0][w];
          AA += red[1][w];
        }
        e.mom[3 * li] = n;
        e.mom[3 * li + 1] = A;
        e.mom[3 * li + 2] = AA;
      }
    }
    for (int job = wave; job < ntile * cbw; job += nwave) {
      const int nt = job % ntile, cbl = job / ntile;  // N-tile, column block within the workgroup
      if (su0 + 16 * cbl >= U) continue;
      v4i acc[NPL];
#pragma unroll
      for (int p = 0; p < NPL; p++) acc[p] = v4i{0, 0, 0, 0};
      for (int ch = 0; ch < nch; ch++) {
        const int sv = segt[4 * ch + gq];
        v4i af = v4i{0, 0, 0, 0};
        int rowoff = 0;
        if (!(sv & 0x8000)) {
          const int qy = sv >> 16, c0 = sv & 0x7fff;
          // A: bytes M(c0 + j - m, qy) = mask row byte 16 + c0 - m + j: five aligned words,
          // realigned per lane
          const int bs = 16 + c0 - m;
          const uint32_t* wp = reinterpret_cast<const uint32_t*>(mrow + (size_t)qy * MLR + (bs & ~3));
          const uint32_t w0 = wp[0], w1 = wp[1], w2 = wp[2], w3 = wp[3], w4 = wp[4];
          const uint32_t sh = (uint32_t)(bs & 3);
          af[0] = (int)__builtin_amdgcn_alignbyte(w1, w0, sh);
          af[1] = (int)__builtin_amdgcn_alignbyte(w2, w1, sh);
          af[2] = (int)__builtin_amdgcn_alignbyte(w3, w2, sh);
          af[3] = (int)__builtin_amdgcn_alignbyte(w4, w3, sh);
          rowoff = qy * stride + 16 * cbl + c0;
        }
        // B: plane p, row (16 nt + col) + qy, bytes c0 .. c0 + 15 (empty segments read row 0)
        const size_t bo = (size_t)(16 * nt + m) * stride + rowoff;
#pragma unroll
        for (int p = 0; p < NPL; p++) {
          const v4i bf = *reinterpret_cast<const v4i*>(planes + p * psz + bo);
          acc[p] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf, acc[p], 0, 0, 0);
        }
      }
      // lane holds rows 4 gq + i (columns su0 + 16 cbl + 4 gq + i) at vertical shift 16 nt + m
      const int svv = 16 * nt + m;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int su = su0 + 16 * cbl + 4 * gq + i;
        if (su < U && svv < V) {
          long long Bv = 0, Cv = 0;
#pragma unroll
          for (int p = 0; p < NPL; p++) {
            const int pl = p0 + p;
            if (bytes) {
              const long long a = (long long)acc[p][i] + 128ll * n;
              if (pl == 0) Bv += a * 256;
              else if (pl == 1) Bv += a;
              else Cv += a << (8 * (pl - 2));
            } else {
              const long long a = acc[p][i];
              if (pl == 0) Bv += a * 128;
              else if (pl == 1) Bv += a;
              else Cv += a << (7 * (pl - 2));
            }
          }
          const int64_t t = (int64_t)su * V + svv;
          if (p0 > 0) {
            Bv += e.mid[sl.mid_off + 2 * t];
            Cv += e.mid[sl.mid_off + 2 * t + 1];
          }
          e.mid[sl.mid_off + 2 * t] = Bv;
          e.mid[sl.mid_off + 2 * t + 1] = Cv;
        }
      }
    }
  }
}

// ---- B_mid, C_mid of a full digital disk from row prefix sums ---------------------------
// Every row qy of a disk base mask is one run of cells, |qx| <= xm(qy) = isqrt(m2 - qy^2),
// so over a window row the run's sums of I1 and I1^2 are differences of two prefix sums:
//   B_mid(s) = sum_qy P(row, hi) - P(row, lo),  C_mid likewise on I1^2,
// 2 (2 r + 1) LDS reads per shift instead of a GEMM over the disk's cells. The prefix of I1^2
// is kept modulo 2^32: a run has at most 2 rho - 1 <= 255 cells of at most 4095^2, so its
// true sum is below 2^32 and the modular difference is exact (a window value above 4095, or
// above the frames' declared depth, sends the grain to the plain kernel; 16-bit frames take
// k_exact_mid instead). Also writes the
// disk's multiplicity grid and, in the first block, the reference moments. One workgroup per
// (slot, block of base_vb vertical shifts); LDS: the window rows of the block (base_vb +
// 2 rho) x (U + 2 rho + 1) prefix pairs, then the run half-widths.
__global__ __launch_bounds__(512) void k_exact_base_disk(ExactArgs e) {
  const FrameDesc fr = *e.s.fr;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ long long red[2][8];
  __shared__ int big;
  const int bi = blockIdx.x;
  if (bi >= bucket_n(e)) return;
  const int li = e.bucket[bi];
  if (li >= *e.s.count) return;
  const ExactSlot sl = e.slots[li];
  if (!sl.base_pat) return;
  const int g = e.s.list[li];
  const int rho = sl.rho, side = 2 * rho + 1;
  const int U = e.s.left + e.s.right + 1, V = e.s.up + e.s.down + 1;
  const int sv0 = blockIdx.y * e.base_vb;
  if (sv0 >= V) return;
  const int nv = min(e.base_vb, V - sv0);
  const int nrow = nv + 2 * rho, ncol = U + 2 * rho, pst = ncol + 1;
  uint2* pre = reinterpret_cast<uint2*>(smem);
  int32_t* xm = reinterpret_cast<int32_t*>(smem + sizeof(uint2) * (size_t)nrow * pst);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const ptg_grain& gr = e.s.grains[g];
  const int X0 = win_x0(gr, e.s.left, rho), Y0 = win_y0(gr, e.s.up, rho) + sv0;
  const int m2 = sl.disk_m2;
  int r = (int)sqrt((double)m2);
  while (r * r > m2) r--;
  while ((r + 1) * (r + 1) <= m2) r++;
  if (threadIdx.x == 0) big = 0;
  for (int q = threadIdx.x; q < 2 * r + 1; q += blockDim.x) {
    const int qy = q - r, t2 = m2 - qy * qy;
    int w = (int)sqrt((double)t2);
    while (w * w > t2) w--;
    while ((w + 1) * (w + 1) <= t2) w++;
    xm[q] = w;
  }
  __syncthreads();
  // exclusive row prefix sums of I1 and I1^2 (mod 2^32), one wave per row; the window
  // reads clamp to the frame as the MFMA kernels' do
  bool over = false;
  // kRC 64-column chunks of a row loaded together, two rows at a time, then scanned
  constexpr int kRC = 4;
  for (int row0 = wave; row0 < nrow; row0 += 2 * nwave) {
    unsigned int va[2][kRC];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = row0 + h * nwave;
      const int64_t rb = (int64_t)clampi(Y0 + row, 0, fr.dimy - 1) * fr.pitch1;
#pragma unroll
      for (int j = 0; j < kRC; j++) {
        const int x = 64 * j + lane;
        va[h][j] = row < nrow && x < ncol ? frame_px(fr.I1, rb + clampi(X0 + x, 0, fr.dimx - 1)) : 0u;
      }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = row0 + h * nwave;
      if (row >= nrow) break;
      const int64_t rb = (int64_t)clampi(Y0 + row, 0, fr.dimy - 1) * fr.pitch1;
      uint2* dst = pre + (size_t)row * pst;
      unsigned int c1 = 0u, c2 = 0u;
      if (lane == 0) dst[0] = make_uint2(0u, 0u);
      for (int x0 = 0; x0 < ncol; x0 += 64) {
        const int x = x0 + lane, j = x0 >> 6;
        unsigned int v = 0u;
        if (j < kRC) {
#pragma unroll
          for (int jj = 0; jj < kRC; jj++)
            if (jj == j) v = va[h][jj];
        } else if (x < ncol) {
          v = frame_px(fr.I1, rb + clampi(X0 + x, 0, fr.dimx - 1));  // (rows wider than kRC chunks)
        }
        over |= v > (unsigned int)min(e.vmax, 4095);
        unsigned int a = v, b = v * v;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
          const unsigned int oa = __shfl_up(a, m, 64), ob = __shfl_up(b, m, 64);
          if (lane >= m) {
            a += oa;
            b += ob;
          }
        }
        if (x < ncol) dst[x + 1] = make_uint2(c1 + a, c2 + b);
        c1 += __shfl(a, 63, 64);
        c2 += __shfl(b, 63, 64);
      }
    }
  }
  if (over) big = 1;
  __syncthreads();
  if (big) {
    // beyond the depth: this grain goes to the plain kernel (its MFMA search returns at once)
    if (threadIdx.x == 0) overflow_slot(e, li);
    return;
  }
  // kSh shifts per thread and pass, (su, sv0 + v) for t = threadIdx.x + j blockDim.x: one
  // half-width read serves all of them, and their 2 kSh prefix reads per row are in flight
  // together
  constexpr int kSh = 4;
  const int nsh = U * nv;
  for (int t0 = threadIdx.x; t0 < nsh; t0 += kSh * blockDim.x) {
    const uint2* rowp[kSh];
    unsigned int B[kSh];
    unsigned long long C[kSh];
#pragma unroll
    for (int j = 0; j < kSh; j++) {
      const int t = min(t0 + j * (int)blockDim.x, nsh - 1);
      const int v = t / U, su = t - v * U;
      rowp[j] = pre + (size_t)(v + rho - r) * pst + su + rho;
      B[j] = 0u;
      C[j] = 0ull;
    }
    for (int q = 0; q <= 2 * r; q++) {
      const int w = xm[q];
      const int ro = q * pst;  // row q's offset (the window's prefix table is below 2^31 entries)
#pragma unroll
      for (int j = 0; j < kSh; j++) {
        const uint2 lo = rowp[j][ro - w], hi = rowp[j][ro + w + 1];
        B[j] += hi.x - lo.x;
        C[j] += (unsigned long long)(hi.y - lo.y);
      }
    }
#pragma unroll
    for (int j = 0; j < kSh; j++) {
      const int t = t0 + j * (int)blockDim.x;
      if (t < nsh) {
        const int v = t / U, su = t - v * U;
        const int64_t idx = (int64_t)su * V + sv0 + v;
        e.mid[sl.mid_off + 2 * idx] = (long long)B[j];
        e.mid[sl.mid_off + 2 * idx + 1] = (long long)C[j];
      }
    }
  }
  if (blockIdx.y == 0) {
    // the disk's multiplicities (each cell once) for the template kernel's Delta lists
    for (int t = threadIdx.x; t < side * side; t += blockDim.x) {
      const int qy = t / side - rho, qx = t - (t / side) * side - rho;
      e.mmask[sl.mmask_off + t] = (uint8_t)(qx * qx + qy * qy <= m2 ? 1 : 0);
    }
    const int64_t b = e.s.pbegin[g];
    const int n = (int)(e.s.pbegin[g + 1] - b);
    long long Sa = 0, Saa = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const long long a = e.s.pwords[b + i] >> 16;
      Sa += a;
      Saa += a * a;
    }
    Sa = wave_sum_i64(Sa);
    Saa = wave_sum_i64(Saa);
    if (lane == 0) {
      red[0][wave] = Sa;
      red[1][wave] = Saa;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      long long A = 0, AA = 0;
      for (int w = 0; w < nwave; w++) {
        A += red[0][w];
        AA += red[1][w];
      }
      e.mom[3 * li] = n;
      e.mom[3 * li + 1] = A;
      e.mom[3 * li + 2] = AA;
    }
  }
}

__host__ __forceinline__ size_t base_disk_lds_bytes(int rho, int U, int vb) {
  return sizeof(uint2) * (size_t)(vb + 2 * rho) * (U + 2 * rho + 1) + sizeof(int32_t) * (size_t)(2 * rho + 1) + 16;
}

// ---- the MFMA search ---------------------------------------------------------------
// Sum of the squares of eight window values: in 32 bits for values of at most 12 bits
// (8 x 4095^2 < 2^32), term by term in 64 bits for 16-bit values.
template <bool W64>
__device__ __forceinline__ unsigned long long sq8(unsigned int w0, unsigned int w1, unsigned int w2, unsigned int w3,
                                                  unsigned int w4, unsigned int w5, unsigned int w6, unsigned int w7) {
  if constexpr (W64) {
    typedef unsigned long long u64;
    return (((u64)(w0 * w0) + (u64)(w1 * w1)) + ((u64)(w2 * w2) + (u64)(w3 * w3))) +
           (((u64)(w4 * w4) + (u64)(w5 * w5)) + ((u64)(w6 * w6) + (u64)(w7 * w7)));
  } else {
    return (unsigned int)(((w0 * w0 + w1 * w1) + (w2 * w2 + w3 * w3)) + ((w4 * w4 + w5 * w5) + (w6 * w6 + w7 * w7)));
  }
}

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

// Sums of I1 and I1^2 over pairs of window values of at most 12 bits: v_dot2_u32_u16 adds both
// values (against (1, 1)) and both squares in one instruction each. A square is below 2^24, so a
// 32-bit sum of up to 128 squares is exact; callers pass at most 16 values' squares through
// csq32 before widening it.
__device__ __forceinline__ void dot2_acc(unsigned int a, unsigned int b, unsigned int& bsum, unsigned int& csq32) {
  const us2 v = us2{(unsigned short)a, (unsigned short)b};
  bsum = __builtin_amdgcn_udot2(v, us2{1, 1}, bsum, false);
  csq32 = __builtin_amdgcn_udot2(v, v, csq32, false);
}

// Sums of I1 and I1^2 at this lane's shift over the entries [lo, lo + len) of one Delta list
// held in LDS from dent[o] (lo a multiple of 4; 4 entries per 16-byte read, 16 window reads
// in flight) and, with tail set, over the list's entries [nl, n) beyond the LDS capacity from
// global memory (dg). A list has at most 1023 cells, so the sum of I1 is exact in 32 bits
// (1023 x 65535 < 2^26); I1^2 is summed as sq8 sums it, then in 64.
template <bool W64>
__device__ __forceinline__ void delta_seg(const uint16_t* win, const uint32_t* dent, const int32_t* dg, int o,
                                          int lo, int len, bool tail, int nl, int n, uint32_t wbase, int base,
                                          unsigned int& bsum, unsigned long long& csum) {
  o += lo;
  const unsigned char* wb = reinterpret_cast<const unsigned char*>(win) + wbase;
  auto px = [&](uint32_t off) { return (unsigned int)*reinterpret_cast<const uint16_t*>(wb + off); };
  int q = 0;
  for (; q + 16 <= len; q += 16) {
    const uint4 e0 = *reinterpret_cast<const uint4*>(dent + o + q);
    const uint4 e1 = *reinterpret_cast<const uint4*>(dent + o + q + 4);
    const uint4 e2 = *reinterpret_cast<const uint4*>(dent + o + q + 8);
    const uint4 e3 = *reinterpret_cast<const uint4*>(dent + o + q + 12);
    const unsigned int w0 = px(e0.x), w1 = px(e0.y), w2 = px(e0.z), w3 = px(e0.w);
    const unsigned int w4 = px(e1.x), w5 = px(e1.y), w6 = px(e1.z), w7 = px(e1.w);
    const unsigned int w8 = px(e2.x), w9 = px(e2.y), w10 = px(e2.z), w11 = px(e2.w);
    const unsigned int w12 = px(e3.x), w13 = px(e3.y), w14 = px(e3.z), w15 = px(e3.w);
    if constexpr (W64) {
      bsum += (((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7))) + (((w8 + w9) + (w10 + w11)) + ((w12 + w13) + (w14 + w15)));
      csum += sq8<W64>(w0, w1, w2, w3, w4, w5, w6, w7);
      csum += sq8<W64>(w8, w9, w10, w11, w12, w13, w14, w15);
    } else {
      unsigned int c32 = 0u;
      dot2_acc(w0, w1, bsum, c32);
      dot2_acc(w2, w3, bsum, c32);
      dot2_acc(w4, w5, bsum, c32);
      dot2_acc(w6, w7, bsum, c32);
      dot2_acc(w8, w9, bsum, c32);
      dot2_acc(w10, w11, bsum, c32);
      dot2_acc(w12, w13, bsum, c32);
      dot2_acc(w14, w15, bsum, c32);
      csum += c32;
    }
  }
  for (; q + 8 <= len; q += 8) {
    const uint4 e0 = *reinterpret_cast<const uint4*>(dent + o + q);
    const uint4 e1 = *reinterpret_cast<const uint4*>(dent + o + q + 4);
    const unsigned int w0 = px(e0.x), w1 = px(e0.y), w2 = px(e0.z), w3 = px(e0.w);
    const unsigned int w4 = px(e1.x), w5 = px(e1.y), w6 = px(e1.z), w7 = px(e1.w);
    if constexpr (W64) {
      bsum += ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7));
      csum += sq8<W64>(w0, w1, w2, w3, w4, w5, w6, w7);
    } else {
      unsigned int c32 = 0u;
      dot2_acc(w0, w1, bsum, c32);
      dot2_acc(w2, w3, bsum, c32);
      dot2_acc(w4, w5, bsum, c32);
      dot2_acc(w6, w7, bsum, c32);
      csum += c32;
    }
  }
  for (; q < len; q += 4) {
    const uint4 e0 = *reinterpret_cast<const uint4*>(dent + o + q);
    const unsigned int w0 = px(e0.x);
    const unsigned int w1 = q + 1 < len ? px(e0.y) : 0u;
    const unsigned int w2 = q + 2 < len ? px(e0.z) : 0u;
    const unsigned int w3 = q + 3 < len ? px(e0.w) : 0u;
    bsum += (w0 + w1) + (w2 + w3);
    csum += sq8<W64>(w0, w1, w2, w3, 0u, 0u, 0u, 0u);
  }
  if (tail)
    for (int r = nl; r < n; r++) {
      const unsigned int w = win[base + dg[r]];
      bsum += w;
      csum += (unsigned long long)(w * w);
    }
}

// The whole list ks = 2 step + sign at this lane's shift; an inactive lane (act false) sums
// nothing.
template <bool W64>
__device__ __forceinline__ void delta_sums(const ExactArgs& e, const uint16_t* win, const int32_t* dcnt,
                                           const int32_t* dofs, const uint32_t* dent, int64_t dl_off, int dcap,
                                           bool act, int ks, uint32_t wbase, int base, unsigned int& bsum,
                                           unsigned long long& csum) {
  const int n = act ? dcnt[ks] : 0, o = act ? dofs[ks] : 0;
  const int nl = max(0, min(n, e.dl_lds_cap - o));
  delta_seg<W64>(win, dent, e.dlist + dl_off + (int64_t)ks * dcap, o, 0, nl, true, nl, n, wbase, base, bsum, csum);
}

// Both lists of step k (cells gained, cells lost) at this lane's shift, for a wave whose lanes
// share the step: their common length is walked side by side, 8 entries of each list per
// iteration, so that the two lists' LDS latencies overlap; then the rest of each.
template <bool W64>
__device__ __forceinline__ void delta_pair(const ExactArgs& e, const uint16_t* win, const int32_t* dcnt,
                                           const int32_t* dofs, const uint32_t* dent, int64_t dl_off, int dcap,
                                           int k, uint32_t wbase, int base, unsigned int& bp,
                                           unsigned long long& cp, unsigned int& bn, unsigned long long& cn) {
  const int n0 = dcnt[2 * k], n1 = dcnt[2 * k + 1], o0 = dofs[2 * k], o1 = dofs[2 * k + 1];
  const int nl0 = max(0, min(n0, e.dl_lds_cap - o0)), nl1 = max(0, min(n1, e.dl_lds_cap - o1));
  const unsigned char* wb = reinterpret_cast<const unsigned char*>(win) + wbase;
  auto px = [&](uint32_t off) { return (unsigned int)*reinterpret_cast<const uint16_t*>(wb + off); };
  const int nc = min(nl0, nl1) & ~7;
  for (int q = 0; q < nc; q += 8) {
    const uint4 a0 = *reinterpret_cast<const uint4*>(dent + o0 + q);
    const uint4 a1 = *reinterpret_cast<const uint4*>(dent + o0 + q + 4);
    const uint4 b0 = *reinterpret_cast<const uint4*>(dent + o1 + q);
    const uint4 b1 = *reinterpret_cast<const uint4*>(dent + o1 + q + 4);
    const unsigned int w0 = px(a0.x), w1 = px(a0.y), w2 = px(a0.z), w3 = px(a0.w);
    const unsigned int w4 = px(a1.x), w5 = px(a1.y), w6 = px(a1.z), w7 = px(a1.w);
    const unsigned int v0 = px(b0.x), v1 = px(b0.y), v2 = px(b0.z), v3 = px(b0.w);
    const unsigned int v4 = px(b1.x), v5 = px(b1.y), v6 = px(b1.z), v7 = px(b1.w);
    if constexpr (W64) {
      bp += ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7));
      cp += sq8<W64>(w0, w1, w2, w3, w4, w5, w6, w7);
      bn += ((v0 + v1) + (v2 + v3)) + ((v4 + v5) + (v6 + v7));
      cn += sq8<W64>(v0, v1, v2, v3, v4, v5, v6, v7);
    } else {
      unsigned int cp32 = 0u, cn32 = 0u;
      dot2_acc(w0, w1, bp, cp32);
      dot2_acc(v0, v1, bn, cn32);
      dot2_acc(w2, w3, bp, cp32);
      dot2_acc(v2, v3, bn, cn32);
      dot2_acc(w4, w5, bp, cp32);
      dot2_acc(v4, v5, bn, cn32);
      dot2_acc(w6, w7, bp, cp32);
      dot2_acc(v6, v7, bn, cn32);
      cp += cp32;
      cn += cn32;
    }
  }
  delta_seg<W64>(win, dent, e.dlist + dl_off + (int64_t)(2 * k) * dcap, o0, nc, nl0 - nc, true, nl0, n0, wbase,
                 base, bp, cp);
  delta_seg<W64>(win, dent, e.dlist + dl_off + (int64_t)(2 * k + 1) * dcap, o1, nc, nl1 - nc, true, nl1, n1,
                 wbase, base, bn, cn);
}

// Hand-over waits of the producer/consumer form: a wait that lasts this many polls (each a short
// sleep, about 0.1 s in all) gives up and voids the search (a hand-over never waits that long).
constexpr int kWsSpinLimit = 1 << 21;

// WS: producer and consumer waves (e.nsu producers, as many consumers: at most 4 + 4); else every
// wave takes its own shift columns through the MFMA loop and the epilogue in turn.
template <int NMT, int NTP, int VM, bool WS>
__global__ __launch_bounds__(512) void k_search_exact(ExactArgs e) {
  const FrameDesc fr = *e.s.fr;
  constexpr int SPT = vm_spt(VM), NP = vm_planes(VM);
  extern __shared__ __align__(16) unsigned char smem[];
  // XCD-aware job mapping: all shift blocks of a grain share blockIdx % 8
  const int bx = blockIdx.x;
  const int xcd = bx & 7, t = bx >> 3;
  const int nbk = bucket_n(e);
  // workgroups per grain: the launch's, or for a rescue stage's queue the heuristic's for the
  // queued count (the grid covers the largest count)
  const int nsub = e.nbucket_dev ? exact_nsub_of(nbk, e.s.left + e.s.right + 1) : e.nsub;
  const int grp = t / nsub, blk = t - grp * nsub;
  const int bi = grp * 8 + xcd;
  if (bi >= nbk) return;
  const int li = e.bucket[bi];
  if (li >= *e.s.count) return;
  if (e.ovf_slot && e.ovf_slot[li]) return;  // served by the plain kernel
  const int g = e.s.list[li];
  const ExactSlot sl = e.slots[li];
  const int rho = sl.rho, cw = sl.cw, cs = sl.cs, kc = sl.kc;
  const int U = e.s.left + e.s.right + 1, V = e.s.up + e.s.down + 1, S = U * V;
  const int vp = e.vp;
  const int WW = win_width(U, cw), WH = vp + 2 * rho;
  const int nsteps = e.s.nsteps;
  const ptg_grain& gr = e.s.grains[g];
  const int X0 = win_x0(gr, e.s.left, rho), Y0 = win_y0(gr, e.s.up, rho);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nthreads = blockDim.x;

  uint16_t* win = reinterpret_cast<uint16_t*>(smem);
  const size_t win_bytes = ((size_t)e.ww * e.wh_pad * 2 + 15) / 16 * 16;
  // (the same layout as the host's LDS size for the launch: NP digit planes per wave)
  const size_t stage_bytes = ((size_t)NP * e.wh_pad * e.cs_max + 15) / 16 * 16;
  uint8_t* stage = reinterpret_cast<uint8_t*>(smem + win_bytes) + (size_t)wave * stage_bytes;
  unsigned char* tail = smem + win_bytes + (size_t)e.nsu * stage_bytes;
  // the epilogue's exchange buffer: the wave's own staging area when the column takes one
  // vertical pass (the staged planes are dead once its MFMAs are issued), else its own area;
  // for producer and consumer waves, one slot per pair with every M-tile's rows
  constexpr int ABR = VM == kVm16 ? 16 : 8;  // its rows: one per M-tile row (16-bit) or per step
  double* abuf = nullptr;
  double* slots = nullptr;
  if constexpr (WS) {
    slots = reinterpret_cast<double*>(tail);
    tail += (size_t)e.nsu * NMT * ABR * (NTP * 16) * 8;
  } else {
    abuf = e.abuf_alias ? reinterpret_cast<double*>(stage)
                        : reinterpret_cast<double*>(tail) + (size_t)wave * ABR * (NTP * 16);
    if (!e.abuf_alias) tail += (size_t)e.nsu * ABR * (NTP * 16) * 8;
  }
  int32_t* bflag = reinterpret_cast<int32_t*>(tail);  // the block's hand-over flag
  tail += 16;
  int32_t* hflag = reinterpret_cast<int32_t*>(tail);  // producer/consumer slot flags, then the abort word
  if constexpr (WS) tail += 80;
  int32_t* segl = reinterpret_cast<int32_t*>(tail);
  for (int q = threadIdx.x; q < 4 * kc; q += nthreads) segl[q] = e.seg[sl.seg_off + q];
  // Delta lists of this slot in LDS: counts, offsets of the lists in a compact table
  // (each list starting on a 16-byte boundary) and the cells as byte offsets into the
  // window; entries beyond dl_lds_cap stay in global memory
  const int nls = 2 * nsteps;
  int32_t* dcnt = segl + ((4 * e.seg_cap + 3) & ~3);
  int32_t* dofs = dcnt + ((nls + 3) & ~3);
  uint32_t* dent = reinterpret_cast<uint32_t*>(dofs + ((nls + 4) & ~3));
  for (int q = threadIdx.x; q < nls; q += nthreads) dcnt[q] = e.dlen[(int64_t)li * nls + q];

  if (threadIdx.x == 0) *bflag = 0;
  if constexpr (WS)
    if (threadIdx.x < 17) hflag[threadIdx.x] = 0;
  __syncthreads();
  bool clamped = false, over = false;
  {
    // the window is a rectangle: some pixel of it lies outside the frame iff one corner does
    clamped = X0 < 0 || Y0 < 0 || X0 + WW > fr.dimx || Y0 + WH > fr.dimy;
    auto px_of = [&](int q) -> unsigned int {
      const int y = q / WW, x = q - y * WW;
      const int X = clampi(X0 + x, 0, fr.dimx - 1), Y = clampi(Y0 + y, 0, fr.dimy - 1);
      return fr.I1[(int64_t)Y * fr.pitch1 + X];
    };
    // kB loads in flight per thread before their LDS stores, then the rest one at a time; the
    // largest value is kept for the depth check
    constexpr int kB = 4;
    const int nwin = WW * WH;
    unsigned int vmx = 0u;
    int q0 = threadIdx.x;
    for (; q0 + (kB - 1) * nthreads < nwin; q0 += kB * nthreads) {
      unsigned int v[kB];
#pragma unroll
      for (int j = 0; j < kB; j++) v[j] = px_of(q0 + j * nthreads);
#pragma unroll
      for (int j = 0; j < kB; j++) {
        vmx = max(vmx, v[j]);
        win[q0 + j * nthreads] = (uint16_t)v[j];
      }
    }
    for (; q0 < nwin; q0 += nthreads) {
      const unsigned int v = px_of(q0);
      vmx = max(vmx, v);
      win[q0] = (uint16_t)v;
    }
    over = vmx > (unsigned int)e.vmax;
  }
  if (clamped) atomicOr(&e.s.grains[g].status, (int)PTG_STATUS_CLAMPED);
  // a window value beyond the declared depth, or a pattern beyond the exact integers: the plain
  // kernel takes this grain (the flag word was cleared before the staging's barrier)
  if (over || (threadIdx.x == 0 && e.mom[3 * li] > kDoubleExactN)) *bflag = 1;
  __syncthreads();
  if (*bflag) {
    if (threadIdx.x == 0) overflow_slot(e, li);
    return;
  }
  if (wave == 0) {
    // exclusive prefix of the list lengths rounded up to 4 entries
    int run = 0;
    for (int q0 = 0; q0 < nls; q0 += 64) {
      const int q = q0 + lane;
      const int len = q < nls ? ((max(dcnt[q], 0) + 3) & ~3) : 0;  // a repeated step has no lists
      int inc = len;
#pragma unroll
      for (int m = 1; m < 64; m <<= 1) {
        const int o = __shfl_up(inc, m, 64);
        if (lane >= m) inc += o;
      }
      if (q < nls) dofs[q] = run + inc - len;
      run += __shfl(inc, 63, 64);
    }
    if (lane == 0) dofs[nls] = run;
  }
  __syncthreads();
  // each wave copies whole lists (their actual length, up to the LDS capacity)
  for (int ks = wave; ks < nls; ks += nthreads >> 6) {
    const int n = min(max(dcnt[ks], 0), e.dl_lds_cap - dofs[ks]), o = dofs[ks];
    const int32_t* src = e.dlist + sl.dl_off + (int64_t)ks * sl.dcap;
    for (int q = lane; q < n; q += 64) dent[o + q] = 2u * (uint32_t)src[q];
  }
  __syncthreads();
  const int gg = lane >> 4, col = lane & 15;
  // repeated steps (no Delta lists) as a bit mask, the same in every lane
  uint64_t rep = 0;
  for (int k = 0; k < nsteps; k++)
    if (dcnt[2 * k] < 0) rep |= 1ull << k;
  const int64_t* mom = e.mom + 3 * li;
  const long long N = mom[0], Sa = mom[1], Saa = mom[2];
  const long long d0i = N * Saa - Sa * Sa;
  // the relative near-tie window (none for a flat pattern, whose reference NCC is NaN at
  // every candidate: the list stays empty)
  const double lo = d0i > 0 ? near_window_lo(N, Sa, d0i, e.vmax) : 2.0;
  Near2 nl = near2_none();  // this lane's list
  double s1lo = -1.0;       // its best's s times lo: candidates below are rejected at once
  const double Nd = (double)N, Sad = (double)Sa;
  const uint8_t* tmpl = e.tmpl + sl.tmpl_off;
  const int32_t* segs = segl;
  // this workgroup's shift columns [c_lo, c_hi); wave w takes c_lo + w, c_lo + w + nsu, ...
  const int cpb = (U + nsub - 1) / nsub;
  const int c_lo = blk * cpb, c_hi = min(U, c_lo + cpb);
  if constexpr (!WS) {
  for (int su = c_lo + wave; su < c_hi; su += e.nsu) {
    // stage this shift column: cell c of staged row y is window column su + c; plane 0 holds
    // the high digit (12-bit: v >> 7; 16-bit: the high byte, top bit flipped) or, for 8-bit
    // frames, the only one (v - 128), plane 1 the low digit
    const int nq = WH * (cw >> 2);
    for (int q = lane; q < nq; q += 64) {
      const int y = q / (cw >> 2), c4 = (q - y * (cw >> 2)) * 4;
      unsigned int hi = 0u, lo = 0u;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const unsigned int v = win[y * WW + su + c4 + j];
        if constexpr (VM == kVm8) {
          hi |= ((v & 255u) ^ 128u) << (8 * j);
        } else if constexpr (VM == kVm12) {
          hi |= ((v >> 7) & 127u) << (8 * j);
          lo |= (v & 127u) << (8 * j);
        } else {
          hi |= (((v >> 8) & 255u) ^ 128u) << (8 * j);
          lo |= ((v & 255u) ^ 128u) << (8 * j);
        }
      }
      *reinterpret_cast<unsigned int*>(stage + (size_t)y * cs + c4) = hi;
      if constexpr (NP == 2) *reinterpret_cast<unsigned int*>(stage + (size_t)e.wh_pad * cs + (size_t)y * cs + c4) = lo;
    }
    // the stage is this wave's alone: a wave-level barrier orders its writes and reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  {
    const int ntile = vp >> 4;
    for (int p0 = 0; p0 < ntile; p0 += NTP) {
      v4i acc[NMT][NTP][NP];
#pragma unroll
      for (int a = 0; a < NMT; a++)
#pragma unroll
        for (int j = 0; j < NTP; j++)
#pragma unroll
          for (int p = 0; p < NP; p++) acc[a][j][p] = v4i{0, 0, 0, 0};
      // Software pipeline: the template fragments (global memory) of chunk ch + 1 are
      // requested before chunk ch multiplies (two named buffers, scheduling barriers keep
      // the loads ahead of their use); the window fragments (LDS) are read per chunk.
      v4i fa0[NMT], fa1[NMT], fb[NTP][NP];
      auto fetch_a = [&](int chq, v4i* fa) {
        const int chc = chq < kc ? chq : kc - 1;
#pragma unroll
        for (int a = 0; a < NMT; a++)
          fa[a] = *reinterpret_cast<const v4i*>(tmpl + (((int64_t)a * kc + chc) * 64 + lane) * 16);
      };
      auto fetch_b = [&](int chq) {
        const int segv = segs[4 * chq + gg];
        const int rowoff = (segv >> 16) * cs + (segv & 0x7fff);  // padding segments read row 0
#pragma unroll
        for (int j = 0; j < NTP; j++) {
          const uint8_t* bp = stage + (size_t)((p0 + j) * 16 + col) * cs + rowoff;
          fb[j][0] = *reinterpret_cast<const v4i*>(bp);
          if constexpr (NP == 2) fb[j][NP - 1] = *reinterpret_cast<const v4i*>(bp + (size_t)e.wh_pad * cs);
        }
      };
      auto mult = [&](const v4i* fa) {
#pragma unroll
        for (int j = 0; j < NTP; j++)
#pragma unroll
          for (int a = 0; a < NMT; a++)
#pragma unroll
            for (int p = 0; p < NP; p++)
              acc[a][j][p] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb[j][p], acc[a][j][p], 0, 0, 0);
      };
      fetch_a(0, fa0);
      for (int ch = 0; ch < kc; ch += 2) {  // kc is even (host pads with empty chunks)
        fetch_a(ch + 1, fa1);
        fetch_b(ch);
        __builtin_amdgcn_sched_barrier(0);
        mult(fa0);
        __builtin_amdgcn_sched_barrier(0);
        fetch_a(ch + 2, fa0);
        fetch_b(ch + 1);
        __builtin_amdgcn_sched_barrier(0);
        mult(fa1);
        __builtin_amdgcn_sched_barrier(0);
      }
      // epilogue, one M-tile (8 steps) at a time: the cross terms go through LDS so that
      // lanes become (step, shift) pairs. Three N-tiles per pass (48 shifts): lane = shift,
      // one step per round, so the step and its correction lists are wave-uniform. One or
      // two N-tiles (few vertical shifts, or many steps): the M-tile's (step, shift) pairs
      // laid out 64 per round, several steps side by side, so the lanes stay busy (17
      // shifts x 8 steps take 3 rounds, not 8; measured 7.28 -> 5.27 ms for 40 steps x
      // 41 x 41 shifts, while the uniform form stays faster with 48 shifts per pass)
      constexpr int TW = NTP * 16;
      constexpr bool kUniform = NTP == 3;
      const int vpass = min(TW, V - p0 * 16);  // vertical shifts of this pass
      const int svl0 = p0 * 16 + lane;  // the uniform form's shift of this lane
      const bool lane_ok = kUniform && lane < TW && svl0 < V;
      long long Bm0 = 0, Cm0 = 0;
      int base0 = 0, sidx0 = 0;
      if (lane_ok) {
        sidx0 = su * V + svl0;
        Bm0 = e.mid[sl.mid_off + 2 * (int64_t)sidx0];
        Cm0 = e.mid[sl.mid_off + 2 * (int64_t)sidx0 + 1];
        base0 = svl0 * WW + su;
      }
#pragma unroll
      for (int a = 0; a < NMT; a++) {
        // lane holds rows 4 gg + i (i < 4) of the M-tile. 8- and 12-bit frames: rows 2 h, 2 h + 1
        // of the lane are step 8 a + 2 gg + h, template digits 1 and 0, and A = T I1 goes to
        // abuf row 2 gg + h. 16-bit frames: row r = 3 j + d is step 5 a + j, template digit
        // 2 - d; its two planes' products, weighted, go to abuf row r, and a step's A is the sum
        // of its three rows. Every value is an integer below 2^53, exact in double.
#pragma unroll
        for (int j = 0; j < NTP; j++) {
          if constexpr (VM == kVm16) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
              const int r = 4 * gg + i, d = r % 3;
              const double wgt = d == 0 ? 16384.0 : d == 1 ? 128.0 : 1.0;
              abuf[r * (NTP * 16) + j * 16 + col] = ((double)acc[a][j][0][i] * 256.0 + (double)acc[a][j][1][i]) * wgt;
            }
          } else {
#pragma unroll
            for (int h = 0; h < 2; h++) {
              double A;
              if constexpr (VM == kVm12)
                A = (double)acc[a][j][0][2 * h] * 16384.0 +
                    ((double)acc[a][j][1][2 * h] + (double)acc[a][j][0][2 * h + 1]) * 128.0 +
                    (double)acc[a][j][1][2 * h + 1];
              else
                A = (double)acc[a][j][0][2 * h] * 128.0 + (double)acc[a][j][0][2 * h + 1];
              abuf[(2 * gg + h) * (NTP * 16) + j * 16 + col] = A;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // T I1 of step 8 a + kl (5 a + kl) at position pos of the pass, the offset put back
        auto a_of = [&](int kl, int pos) -> double {
          if constexpr (VM == kVm16)
            return ((abuf[(3 * kl) * TW + pos] + abuf[(3 * kl + 1) * TW + pos]) + abuf[(3 * kl + 2) * TW + pos]) +
                   vm_offset(VM) * Sad;
          else if constexpr (VM == kVm8)
            return abuf[kl * TW + pos] + vm_offset(VM) * Sad;
          else
            return abuf[kl * TW + pos];
        };
        // one candidate (step k, shift sidx) with its Delta sums: num = N A - Sa B and
        // d1 = N C - B^2 exactly (every term an integer below 2^53, patterns of at most
        // kDoubleExactN pixels; for 16-bit frames in int64, then rounded to double), then
        // s = num^2 / d1 = q^2 d0 in FP64 (5 roundings). Only
        // num > 0 can win, as best_NCC starts at 0 (src/p-tracker.cpp:680); d1 = 0 is a flat
        // window, whose reference NCC is NaN. A candidate below the window of the lane's best
        // is dropped here; the rest (a few per lane) go through the out-of-line insertion
        auto consider = [&](int k, int sidx, double A, long long Bm, long long Cm, unsigned int bp, unsigned int bn,
                            unsigned long long cp, unsigned long long cn) {
          double nd, dd;
          if constexpr (VM == kVm16) {
            const long long Ai = (long long)A;
            const long long Bi = Bm + (long long)bp - (long long)bn;
            const long long Ci = Cm + (long long)cp - (long long)cn;
            nd = (double)(N * Ai - Sa * Bi);
            dd = (double)(N * Ci - Bi * Bi);
          } else {
            const double B = (double)Bm + ((double)bp - (double)bn);
            const double C = (double)Cm + ((double)cp - (double)cn);
            nd = Nd * A - Sad * B;
            dd = Nd * C - B * B;
          }
          if (nd > 0.0 && dd > 0.0 && nd * nd >= s1lo * dd) {
            nl = near2_insert(nl, nd, dd, k * S + sidx, lo);
            s1lo = nl.s1 * lo;
          }
        };
        if constexpr (kUniform) {
          const uint32_t wbase = 2u * (uint32_t)base0;  // byte offset of this lane's window origin
          for (int kl = 0; kl < SPT; kl++) {
            const int k = SPT * a + kl;
            if (k >= nsteps) break;
            if ((rep >> k) & 1) continue;  // a repeated step
            unsigned int bp = 0u, bn = 0u;
            unsigned long long cp = 0ull, cn = 0ull;
            delta_pair<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, k, wbase, base0, bp, cp, bn, cn);
            if (lane_ok) consider(k, sidx0, a_of(kl, lane), Bm0, Cm0, bp, bn, cp, cn);
          }
        } else {
          const int nk = min(SPT, nsteps - SPT * a);
          const int nitem = nk > 0 ? nk * vpass : 0;
          for (int it0 = 0; it0 < nitem; it0 += 64) {
            const int it = it0 + lane;
            const int kl = it < nitem ? it / vpass : 0, vl = it < nitem ? it - kl * vpass : 0;
            const int k = SPT * a + kl, svl = p0 * 16 + vl;
            const bool act = it < nitem && !((rep >> k) & 1);  // a repeated step is skipped
            const int sidx = su * V + svl, base = svl * WW + su;
            long long Bm = 0, Cm = 0;
            if (act) {
              Bm = e.mid[sl.mid_off + 2 * (int64_t)sidx];
              Cm = e.mid[sl.mid_off + 2 * (int64_t)sidx + 1];
            }
            const uint32_t wbase = 2u * (uint32_t)base;
            unsigned int bp = 0u, bn = 0u;
            unsigned long long cp = 0ull, cn = 0ull;
            delta_sums<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, act, 2 * k, wbase, base, bp, cp);
            delta_sums<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, act, 2 * k + 1, wbase, base, bn,
                                    cn);
            if (act) consider(k, sidx, a_of(kl, vl), Bm, Cm, bp, bn, cp, cn);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    }
  }
  } else {
    // Producer and consumer waves: producer w (w < P) stages its shift columns and runs each
    // vertical pass's MFMA GEMM, then hands the pass's cross terms A (every step, TW shifts)
    // to consumer P + w through its LDS slot; the consumer does the pass's epilogue (Delta
    // sums and the near-tie list) while the producer multiplies the next pass. A flag per slot
    // (0 free, 1 full) orders the hand-over (release store, acquire load, workgroup scope).
    // Waits are bounded: one that runs out marks the search void (flags[0] bit 2) and tells
    // every wave to skip the rest, so all reach the end.
    const int P = e.nsu;
    const bool producer = wave < P;
    const int pid = producer ? wave : wave - P;
    int32_t* hf = hflag + pid;
    int32_t* habort = hflag + 16;
    const int ntile = vp >> 4;
    constexpr int TW = NTP * 16;
    double* slot = slots + (size_t)pid * (NMT * ABR * TW);
    auto ws_wait = [&](int want) -> bool {
      for (int n = 0; n < kWsSpinLimit; n++) {
        if (__hip_atomic_load(hf, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == want) return true;
        if (__hip_atomic_load(habort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
        __builtin_amdgcn_s_sleep(1);
      }
      if (lane == 0) atomicOr(&e.flags[0], 4);
      __hip_atomic_store(habort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return false;
    };
    // the pass's cross term of step k at position pos, the offset put back
    auto a_at = [&](int k, int pos) -> double {
      if constexpr (VM == kVm16) {
        const int a = k / 5, j = k - 5 * a;
        const double* r = slot + (size_t)(a * 16 + 3 * j) * TW + pos;
        return ((r[0] + r[TW]) + r[2 * TW]) + vm_offset(VM) * Sad;
      } else if constexpr (VM == kVm8) {
        return slot[(size_t)k * TW + pos] + vm_offset(VM) * Sad;
      } else {
        return slot[(size_t)k * TW + pos];
      }
    };
    // one candidate with its Delta sums (as the single-role form's)
    auto consider = [&](int k, int sidx, double A, long long Bm, long long Cm, unsigned int bp, unsigned int bn,
                        unsigned long long cp, unsigned long long cn) {
      double nd, dd;
      if constexpr (VM == kVm16) {
        const long long Ai = (long long)A;
        const long long Bi = Bm + (long long)bp - (long long)bn;
        const long long Ci = Cm + (long long)cp - (long long)cn;
        nd = (double)(N * Ai - Sa * Bi);
        dd = (double)(N * Ci - Bi * Bi);
      } else {
        const double B = (double)Bm + ((double)bp - (double)bn);
        const double C = (double)Cm + ((double)cp - (double)cn);
        nd = Nd * A - Sad * B;
        dd = Nd * C - B * B;
      }
      if (nd > 0.0 && dd > 0.0 && nd * nd >= s1lo * dd) {
        nl = near2_insert(nl, nd, dd, k * S + sidx, lo);
        s1lo = nl.s1 * lo;
      }
    };
    bool alive = true;
    for (int su = c_lo + pid; su < c_hi && alive; su += P) {
      if (producer) {
        // stage this shift column: cell c of staged row y is window column su + c; plane 0 holds
        // the high digit (12-bit: v >> 7; 16-bit: the high byte, top bit flipped) or, for 8-bit
        // frames, the only one (v - 128), plane 1 the low digit
        const int nq = WH * (cw >> 2);
        for (int q = lane; q < nq; q += 64) {
          const int y = q / (cw >> 2), c4 = (q - y * (cw >> 2)) * 4;
          unsigned int hi = 0u, lo = 0u;
    #pragma unroll
          for (int j = 0; j < 4; j++) {
            const unsigned int v = win[y * WW + su + c4 + j];
            if constexpr (VM == kVm8) {
              hi |= ((v & 255u) ^ 128u) << (8 * j);
            } else if constexpr (VM == kVm12) {
              hi |= ((v >> 7) & 127u) << (8 * j);
              lo |= (v & 127u) << (8 * j);
            } else {
              hi |= (((v >> 8) & 255u) ^ 128u) << (8 * j);
              lo |= ((v & 255u) ^ 128u) << (8 * j);
            }
          }
          *reinterpret_cast<unsigned int*>(stage + (size_t)y * cs + c4) = hi;
          if constexpr (NP == 2) *reinterpret_cast<unsigned int*>(stage + (size_t)e.wh_pad * cs + (size_t)y * cs + c4) = lo;
        }
        // the stage is this wave's alone: a wave-level barrier orders its writes and reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      for (int p0 = 0; p0 < ntile && alive; p0 += NTP) {
        if (producer) {
          v4i acc[NMT][NTP][NP];
    #pragma unroll
          for (int a = 0; a < NMT; a++)
    #pragma unroll
            for (int j = 0; j < NTP; j++)
    #pragma unroll
              for (int p = 0; p < NP; p++) acc[a][j][p] = v4i{0, 0, 0, 0};
          // Software pipeline: the template fragments (global memory) of chunk ch + 1 are
          // requested before chunk ch multiplies (two named buffers, scheduling barriers keep
          // the loads ahead of their use); the window fragments (LDS) are read per chunk.
          v4i fa0[NMT], fa1[NMT], fb[NTP][NP];
          auto fetch_a = [&](int chq, v4i* fa) {
            const int chc = chq < kc ? chq : kc - 1;
    #pragma unroll
            for (int a = 0; a < NMT; a++)
              fa[a] = *reinterpret_cast<const v4i*>(tmpl + (((int64_t)a * kc + chc) * 64 + lane) * 16);
          };
          auto fetch_b = [&](int chq) {
            const int segv = segs[4 * chq + gg];
            const int rowoff = (segv >> 16) * cs + (segv & 0x7fff);  // padding segments read row 0
    #pragma unroll
            for (int j = 0; j < NTP; j++) {
              const uint8_t* bp = stage + (size_t)((p0 + j) * 16 + col) * cs + rowoff;
              fb[j][0] = *reinterpret_cast<const v4i*>(bp);
              if constexpr (NP == 2) fb[j][NP - 1] = *reinterpret_cast<const v4i*>(bp + (size_t)e.wh_pad * cs);
            }
          };
          auto mult = [&](const v4i* fa) {
    #pragma unroll
            for (int j = 0; j < NTP; j++)
    #pragma unroll
              for (int a = 0; a < NMT; a++)
    #pragma unroll
                for (int p = 0; p < NP; p++)
                  acc[a][j][p] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb[j][p], acc[a][j][p], 0, 0, 0);
          };
          fetch_a(0, fa0);
          for (int ch = 0; ch < kc; ch += 2) {  // kc is even (host pads with empty chunks)
            fetch_a(ch + 1, fa1);
            fetch_b(ch);
            __builtin_amdgcn_sched_barrier(0);
            mult(fa0);
            __builtin_amdgcn_sched_barrier(0);
            fetch_a(ch + 2, fa0);
            fetch_b(ch + 1);
            __builtin_amdgcn_sched_barrier(0);
            mult(fa1);
            __builtin_amdgcn_sched_barrier(0);
          }
          if (!ws_wait(0)) {
            alive = false;
            break;
          }
          // A of every step of this pass into the slot: 8- and 12-bit frames one row per step,
          // 16-bit frames the M-tiles' rows (three per step, summed by the consumer)
#pragma unroll
          for (int a = 0; a < NMT; a++)
#pragma unroll
            for (int j = 0; j < NTP; j++) {
              if constexpr (VM == kVm16) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                  const int r = 4 * gg + i, d = r % 3;
                  const double wgt = d == 0 ? 16384.0 : d == 1 ? 128.0 : 1.0;
                  slot[(size_t)(a * 16 + r) * TW + j * 16 + col] =
                      ((double)acc[a][j][0][i] * 256.0 + (double)acc[a][j][1][i]) * wgt;
                }
              } else {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                  double A;
                  if constexpr (VM == kVm12)
                    A = (double)acc[a][j][0][2 * h] * 16384.0 +
                        ((double)acc[a][j][1][2 * h] + (double)acc[a][j][0][2 * h + 1]) * 128.0 +
                        (double)acc[a][j][1][2 * h + 1];
                  else
                    A = (double)acc[a][j][0][2 * h] * 128.0 + (double)acc[a][j][0][2 * h + 1];
                  slot[(size_t)(a * 8 + 2 * gg + h) * TW + j * 16 + col] = A;
                }
              }
            }
          __hip_atomic_store(hf, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
          if (!ws_wait(1)) {
            alive = false;
            break;
          }
          const int vpass = min(TW, V - p0 * 16);  // vertical shifts of this pass
          if constexpr (NTP == 3) {
            // lane = vertical shift, one step per round: the step's lists are wave-uniform
            const int svl0 = p0 * 16 + lane;
            const bool lane_ok = lane < TW && svl0 < V;
            long long Bm0 = 0, Cm0 = 0;
            int base0 = 0, sidx0 = 0;
            if (lane_ok) {
              sidx0 = su * V + svl0;
              Bm0 = e.mid[sl.mid_off + 2 * (int64_t)sidx0];
              Cm0 = e.mid[sl.mid_off + 2 * (int64_t)sidx0 + 1];
              base0 = svl0 * WW + su;
            }
            const uint32_t wbase = 2u * (uint32_t)base0;
            for (int k = 0; k < nsteps; k++) {
              if ((rep >> k) & 1) continue;  // a repeated step
              unsigned int bp = 0u, bn = 0u;
              unsigned long long cp = 0ull, cn = 0ull;
              delta_pair<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, k, wbase, base0, bp, cp, bn, cn);
              if (lane_ok) consider(k, sidx0, a_at(k, lane), Bm0, Cm0, bp, bn, cp, cn);
            }
          } else {
            // the pass's (step, shift) pairs 64 per round over every step
            const int nitem = nsteps * vpass;
            for (int it0 = 0; it0 < nitem; it0 += 64) {
              const int it = it0 + lane;
              const int k = it < nitem ? it / vpass : 0, vl = it < nitem ? it - k * vpass : 0;
              const int svl = p0 * 16 + vl;
              const bool act = it < nitem && !((rep >> k) & 1);  // a repeated step is skipped
              const int sidx = su * V + svl, base = svl * WW + su;
              long long Bm = 0, Cm = 0;
              if (act) {
                Bm = e.mid[sl.mid_off + 2 * (int64_t)sidx];
                Cm = e.mid[sl.mid_off + 2 * (int64_t)sidx + 1];
              }
              const uint32_t wbase = 2u * (uint32_t)base;
              unsigned int bp = 0u, bn = 0u;
              unsigned long long cp = 0ull, cn = 0ull;
              delta_sums<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, act, 2 * k, wbase, base, bp, cp);
              delta_sums<VM == kVm16>(e, win, dcnt, dofs, dent, sl.dl_off, sl.dcap, act, 2 * k + 1, wbase, base, bn,
                                      cn);
              if (act) consider(k, sidx, a_at(k, vl), Bm, Cm, bp, bn, cp, cn);
            }
          }
          __hip_atomic_store(hf, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
  }
  // the block's list: its best (lanes, then waves, as (s, index) pairs), then the entries
  // within the window of that best other than itself, gathered in LDS (they are few: none
  // for most blocks); more than one such entry, or a lane that overflowed with its best in the
  // window, overflows the block. The LDS is the window's, once every wave is done with it
  // (static arrays would add to the LDS every workgroup reserves).
  double bs = nl.s1;
  int bix = nl.i1;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double os = __shfl_xor(bs, m, 64);
    const int oi = __shfl_xor(bix, m, 64);
    if (rank_gt(os, oi, bs, bix)) {
      bs = os;
      bix = oi;
    }
  }
  double* red_s = reinterpret_cast<double*>(smem);  // 8 waves' best s, then the near entry
  int32_t* red_i = reinterpret_cast<int32_t*>(red_s + 10);  // their indices, the near count, ovf
  __syncthreads();
  if (lane == 0) {
    red_s[wave] = bs;
    red_i[wave] = bix;
  }
  if (threadIdx.x == 0) red_i[9] = red_i[10] = 0;
  __syncthreads();
  for (int w = 0; w < (nthreads >> 6); w++)
    if (rank_gt(red_s[w], red_i[w], bs, bix)) {
      bs = red_s[w];
      bix = red_i[w];
    }
  const double thr = bs * lo;
  const bool n1 = nl.i1 != INT_MAX && nl.i1 != bix && nl.s1 >= thr;
  const bool n2 = nl.i2 != INT_MAX && nl.i2 != bix && nl.s2 >= thr;
  if (n1 || n2) {
    const int at = atomicAdd(&red_i[9], (n1 ? 1 : 0) + (n2 ? 1 : 0));
    if (at == 0) {
      red_s[8] = n1 ? nl.s1 : nl.s2;
      red_i[8] = n1 ? nl.i1 : nl.i2;
    }
  }
  if (nl.ovf && nl.s1 >= thr) red_i[10] = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nn = red_i[9];
    ExactBest xb;
    xb.s1 = bs;
    xb.i1 = bix;
    xb.s2 = nn == 1 ? red_s[8] : -1.0;
    xb.i2 = nn == 1 ? red_i[8] : INT_MAX;
    xb.ovf = (red_i[10] || nn > 1) ? 1 : 0;
    xb.pad = 0;
    e.xout[(int64_t)li * e.part_cap + blk] = xb;
  }
}

// Clears a search's outputs and counters in one launch: the partial picks and the blocks'
// near-tie lists (n of each), the overflow flags of nlist slots (when kept), the capacity flags
// and, for a rescue stage's queue, the slots' marks and the buckets' counts.
__global__ void k_exact_reset(StepBest* __restrict__ best, ExactBest* __restrict__ xout, int n,
                              int32_t* __restrict__ ovf_slot, int32_t* __restrict__ act, int nlist,
                              int32_t* __restrict__ flags, int32_t* __restrict__ bcount, int nbuckets) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    StepBest sb;
    sb.ncc = 0.0;
    sb.idx = INT_MAX;
    sb.pad = 0;
    best[i] = sb;
    ExactBest xb;
    xb.s1 = xb.s2 = -1.0;
    xb.i1 = xb.i2 = INT_MAX;
    xb.ovf = xb.pad = 0;
    xout[i] = xb;
  }
  if (i < nlist) {
    if (ovf_slot) ovf_slot[i] = 0;
    if (act) act[i] = 0;
  }
  if (i < 4) flags[i] = 0;
  if (bcount && i < nbuckets) bcount[i] = 0;
}

// ---- the grain's near-tie set and its re-scoring --------------------------------------
// One wave per list slot: the blocks' lists merged into the grain's. A grain whose set could
// hold more than two candidates (ovf), whose best q is below 2 kQmin, or which has no
// candidate with num > 0 at all (the reference could then pick a candidate whose exact NCC
// is 0 or below, by rounding) goes to the plain kernel; a flat pattern has none and keeps
// none (the reference's NCC is NaN everywhere). A set of one candidate is the reference's
// pick: its NCC is taken as q = sqrt(s / d0) (within a few ulps of the reference's) unless q
// lies within the window of the stage's threshold, where the reference's last bits decide
// the queueing. Such grains and sets of two go to cand[2 li], cand[2 li + 1] (INT_MAX:
// none) for re-scoring.
constexpr int kMergeWaves = 4;

__global__ __launch_bounds__(64 * kMergeWaves) void k_exact_merge(ExactArgs e) {
  const int lane = threadIdx.x & 63;
  const int bi = blockIdx.x * kMergeWaves + (threadIdx.x >> 6);
  if (bi >= e.nbucket) return;
  const int li = e.bucket[bi];
  if (li >= *e.s.count) return;
  if (e.s.act && !e.s.act[li]) return;  // not in this rescue stage's queue
  if (lane == 0) e.cand[2 * li] = e.cand[2 * li + 1] = INT_MAX;
  if (e.ovf_slot && e.ovf_slot[li]) return;  // already the plain kernel's
  const int64_t* mom = e.mom + 3 * li;
  const long long N = mom[0], Sa = mom[1], Saa = mom[2];
  const long long d0 = N * Saa - Sa * Sa;
  if (d0 <= 0) return;
  const double lo = near_window_lo(N, Sa, d0, e.vmax);
  Near2 m = near2_none();
  const ExactBest* xb = e.xout + (int64_t)li * e.part_cap;
  for (int b = lane; b < e.part_cap; b += 64) {
    const ExactBest x = xb[b];
    m = near2_merge(m, Near2{x.s1, x.s2, x.i1, x.i2, x.ovf}, lo);
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) m = near2_merge(m, near2_shfl_xor(m, k), lo);
  if (lane != 0) return;
  const double q = m.i1 == INT_MAX ? 0.0 : sqrt(m.s1 / (double)d0);
  if (m.ovf || !(q >= 2.0 * kQmin)) {
    overflow_slot(e, li);
    return;
  }
  const bool two = m.i2 != INT_MAX && m.s2 >= m.s1 * lo;
  if (!two && fabs(q - e.thr) > near_window_q(N, Sa, d0, e.vmax)) {
    StepBest sb;
    sb.ncc = q;
    sb.idx = m.i1;
    sb.pad = 0;
    e.out[(int64_t)li * e.part_cap] = sb;
    return;
  }
  e.cand[2 * li] = m.i1;
  if (two) e.cand[2 * li + 1] = m.i2;
}

// One workgroup per list slot of a radius bucket: each candidate of the grain's near-tie set
// scored with the reference's arithmetic (src/p-tracker.cpp:698-716; the plain kernel's):
// the integer sum of I1 over the rotated pattern, one division for the mean; then, chunk by
// chunk, every lane forms products (I0 - mean0)(I1 - m1) and squares (I1 - m1)^2, each
// rounded as the reference rounds it, and one lane per sum adds them in pattern order,
// unfused. The larger NCC wins, the earlier candidate among equal values. The result is the
// slot's only partial, so the commit sees the reference's pick and NCC.
constexpr int kRescoreThreads = 256;
constexpr int kRescoreChunk = 1024;

__global__ __launch_bounds__(kRescoreThreads) void k_exact_rescore(ExactArgs e, StepBest* out) {
  const FrameDesc fr = *e.s.fr;
  __shared__ __align__(16) double terms[4][kRescoreChunk];  // (candidate, C0C1 | C1C1) x chunk
  __shared__ unsigned long long part[2][kRescoreThreads / 64];
  __shared__ double sums[4];
  if ((int)blockIdx.x >= e.nbucket) return;
  const int li = e.bucket[blockIdx.x];
  if (li >= *e.s.count) return;
  if (e.s.act && !e.s.act[li]) return;
  if (e.ovf_slot && e.ovf_slot[li]) return;
  const int ca[2] = {e.cand[2 * li], e.cand[2 * li + 1]};
  if (ca[0] == INT_MAX) return;
  const int nc = ca[1] == INT_MAX ? 1 : 2;
  const int g = e.s.list[li];
  const ptg_grain& gr = e.s.grains[g];
  const int64_t b = e.s.pbegin[g];
  const int n = (int)(e.s.pbegin[g + 1] - b);
  const int nsteps = e.s.nsteps;
  const int V = e.s.up + e.s.down + 1, S = (e.s.left + e.s.right + 1) * V;
  const int gdx = (int)round(gr.dx), gdy = (int)round(gr.dy);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int bx[2], by[2];
  double c[2], sn[2];
  for (int q = 0; q < 2; q++) {
    const int idx = q < nc ? ca[q] : ca[0];
    const int k = idx / S, rem = idx - k * S;
    bx[q] = gr.refcoord_xpix + gdx + rem / V - e.s.left;
    by[q] = gr.refcoord_ypix + gdy + rem % V - e.s.up;
    c[q] = e.s.cs[2 * ((int64_t)g * nsteps + k)];
    sn[q] = e.s.cs[2 * ((int64_t)g * nsteps + k) + 1];
  }
  auto pixel = [&](int q, int i) -> unsigned int {
    int x, y;
    rotate_px(e.s.pattern[2 * (b + i)], e.s.pattern[2 * (b + i) + 1], c[q], sn[q], x, y);
    const int X = clampi(bx[q] + x, 0, fr.dimx - 1), Y = clampi(by[q] + y, 0, fr.dimy - 1);
    return fr.I1[(int64_t)Y * fr.pitch1 + X];
  };
  // mean1 (:698-705): the sum of integers is exact in any order, then one division
  unsigned long long s1[2] = {0ull, 0ull};
  for (int i = threadIdx.x; i < n; i += kRescoreThreads)
    for (int q = 0; q < nc; q++) s1[q] += pixel(q, i);
  for (int q = 0; q < nc; q++) {
    const unsigned long long t = (unsigned long long)wave_sum_i64((long long)s1[q]);
    if (lane == 0) part[q][wave] = t;
  }
  __syncthreads();
  double m1[2] = {0.0, 0.0};
  for (int q = 0; q < nc; q++) {
    unsigned long long t = 0ull;
    for (int w = 0; w < kRescoreThreads / 64; w++) t += part[q][w];
    m1[q] = (double)t / (double)n;
  }
  // C0C1 and C1C1 (:706-714): lane 2q + j of wave 0 adds the terms of sum j of candidate q
  const double* d0 = e.s.d0 + b;
  const int adder = threadIdx.x < 2 * nc ? threadIdx.x : -1;
  double acc = 0.0;
  for (int i0 = 0; i0 < n; i0 += kRescoreChunk) {
    const int m = min(kRescoreChunk, n - i0);
    __syncthreads();  // the previous chunk is summed
    for (int t = threadIdx.x; t < m; t += kRescoreThreads) {
      const double dd = d0[i0 + t];
      for (int q = 0; q < nc; q++) {
        const double diffC1 = (double)pixel(q, i0 + t) - m1[q];
        terms[2 * q][t] = dd * diffC1;
        terms[2 * q + 1][t] = diffC1 * diffC1;
      }
    }
    __syncthreads();
    if (adder >= 0) {
      // the next 8 terms are read while the current 8 are added, so the chain of dependent
      // additions does not wait on LDS
      const double2* tp = reinterpret_cast<const double2*>(terms[adder]);
      const int m8 = m & ~7;
      double2 a0, a1, a2, a3;
      if (m8 > 0) {
        a0 = tp[0];
        a1 = tp[1];
        a2 = tp[2];
        a3 = tp[3];
      }
      for (int t = 0; t < m8; t += 8) {
        const int nx = t + 8 < m8 ? (t + 8) / 2 : t / 2;
        const double2 b0 = tp[nx], b1 = tp[nx + 1], b2 = tp[nx + 2], b3 = tp[nx + 3];
        acc = acc + a0.x;
        acc = acc + a0.y;
        acc = acc + a1.x;
        acc = acc + a1.y;
        acc = acc + a2.x;
        acc = acc + a2.y;
        acc = acc + a3.x;
        acc = acc + a3.y;
        a0 = b0;
        a1 = b1;
        a2 = b2;
        a3 = b3;
      }
      for (int t = m8; t < m; t++) acc = acc + terms[adder][t];
    }
  }
  if (adder >= 0) sums[adder] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    // src/p-tracker.cpp:716-718: NCC = C0C1 / sqrt(C0C0 C1C1); strictly larger wins, from
    // best_NCC = 0; the earlier among equals
    double best = 0.0;
    int bidx = INT_MAX;
    for (int q = 0; q < nc; q++) {
      const double v = sums[2 * q] / sqrt(gr.C0C0 * sums[2 * q + 1]);
      if (v > best || (v == best && v > 0.0 && ca[q] < bidx)) {
        best = v;
        bidx = ca[q];
      }
    }
    StepBest sb;
    sb.ncc = best;
    sb.idx = bidx;
    sb.pad = 0;
    out[(int64_t)li * e.part_cap] = sb;
  }
}

// A rescue stage's queued grains as the plan's slots: act[slot] = 1 and the slot appended to its
// radius bucket's list (bucket_act[bucket_off[b] ...], bcount[b] of them). act and bcount start
// at zero.
__global__ void k_exact_activate(const int32_t* __restrict__ queue, const int32_t* __restrict__ qcount,
                                 const int32_t* __restrict__ slot_of, const int32_t* __restrict__ slot_bucket,
                                 const int32_t* __restrict__ bucket_off, int32_t* __restrict__ act,
                                 int32_t* __restrict__ bcount, int32_t* __restrict__ bucket_act) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *qcount) return;
  const int li = slot_of[queue[j]];
  if (li < 0) return;  // (a queue holds only the list's grains)
  act[li] = 1;
  const int b = slot_bucket[li];
  bucket_act[bucket_off[b] + atomicAdd(&bcount[b], 1)] = li;
}

// ---- host side -----------------------------------------------------------------------
template <int NMT, int NTP, int VM, bool WS>
static hipError_t launch_main(int dev, const ExactArgs& e, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  if (hipError_t r = ensure_kernel_lds(dev, (const void*)k_search_exact<NMT, NTP, VM, WS>, lds); r != hipSuccess)
    return r;
  k_search_exact<NMT, NTP, VM, WS><<<grid, block, lds, st>>>(e);
  return hipGetLastError();
}

// The instances: M-tiles x N-tiles per pass for each value mode. 8- and 12-bit frames take
// up to 6 M-tiles (48 steps), 16-bit frames up to 8 (40 steps, 5 per M-tile).
#define PTG_EXACT_SHAPES(X, VM)                                                                                \
  X(1, 3, VM) X(2, 3, VM) X(3, 3, VM) X(4, 2, VM) X(5, 2, VM) X(6, 2, VM) X(1, 2, VM) X(2, 2, VM) X(3, 2, VM) \
  X(1, 1, VM) X(2, 1, VM) X(3, 1, VM) X(4, 1, VM) X(5, 1, VM) X(6, 1, VM)
#define PTG_EXACT_INSTANCES(X)                                                                          \
  PTG_EXACT_SHAPES(X, kVm8) PTG_EXACT_SHAPES(X, kVm12) PTG_EXACT_SHAPES(X, kVm16) X(7, 1, kVm16) \
  X(8, 1, kVm16)
constexpr int kExactMaxMt = 8;

// Vector registers and static LDS bytes of a k_search_exact instance (from its code object),
// for occupancy.
struct MainAttr {
  int vgprs = 256;
  size_t static_lds = 1024;
};
static MainAttr main_attr(int vm, int nmt, int ntp, bool ws) {
  const void* fn = nullptr;
#define PTG_FN(M, T, VM)                                                                                \
  if (vm == VM && nmt == M && ntp == T)                                                                 \
    fn = ws ? (const void*)k_search_exact<M, T, VM, true> : (const void*)k_search_exact<M, T, VM, false>;
  PTG_EXACT_INSTANCES(PTG_FN)
#undef PTG_FN
  static MainAttr cache[2][3][kExactMaxMt + 1][4] = {};
  static bool known[2][3][kExactMaxMt + 1][4] = {};
  if (!fn) return MainAttr{};
  if (!known[ws][vm][nmt][ntp]) {
    hipFuncAttributes at;
    if (hipFuncGetAttributes(&at, fn) == hipSuccess && at.numRegs > 0) {
      cache[ws][vm][nmt][ntp].vgprs = at.numRegs;
      cache[ws][vm][nmt][ntp].static_lds = at.sharedSizeBytes;
    }
    known[ws][vm][nmt][ntp] = true;
  }
  return cache[ws][vm][nmt][ntp];
}

static hipError_t dispatch_main(int dev, int vm, int nmt, int ntp, bool ws, const ExactArgs& e, dim3 grid,
                                dim3 block, size_t lds, hipStream_t st) {
#define PTG_CASE(M, T, VM)                                                                   \
  if (vm == VM && nmt == M && ntp == T)                                                      \
    return ws ? launch_main<M, T, VM, true>(dev, e, grid, block, lds, st)                    \
              : launch_main<M, T, VM, false>(dev, e, grid, block, lds, st);
  PTG_EXACT_INSTANCES(PTG_CASE)
#undef PTG_CASE
  return hipErrorInvalidValue;
}

namespace {
// Launch shapes of one radius bucket.
struct BucketLaunch {
  int cbw = 0, npl = 0;          // k_exact_mid: column blocks per workgroup, planes per pass (6, 3)
  int vb = 0;                    // k_exact_base_disk: vertical shifts per workgroup (0: k_exact_mid's)
  size_t lds_d = 0;              // its LDS
  size_t lds_m = 0, lds_t = 0;   // LDS of k_exact_mid, k_exact_template
  int cs = 0, ww = 0, wh = 0;    // k_search_exact: staging stride, window width and height
  int dl_cap = 0, nsub = 0, nsu = 0, ntp = 0;
  bool alias = false;
  bool ws = false;               // producer and consumer waves (nsu producers)
  size_t lds = 0;
  uint64_t key = 0;              // the bucket's key in the autotuned shapes
};

// Whether k_search_exact has an instance for nmt M-tiles x ntp N-tiles in value mode vm.
static bool exact_instance(int vm, int nmt, int ntp) {
#define PTG_HAS(M, T, VM) \
  if (vm == VM && nmt == M && ntp == T) return true;
  PTG_EXACT_INSTANCES(PTG_HAS)
#undef PTG_HAS
  return false;
}

// The key of a radius bucket's launch shape: everything the shape's cost depends on.
static uint64_t bucket_key(int vm, int nsteps, int U, int V, int rho, int nb) {
  uint64_t h = 1469598103934665603ull;
  for (const int64_t v : {(int64_t)vm, (int64_t)nsteps, (int64_t)U, (int64_t)V, (int64_t)rho, (int64_t)nb}) {
    h ^= (uint64_t)v;
    h *= 1099511628211ull;
  }
  return h;
}

struct Bucket {
  int rho_max = 0;
  int ndisk = 0;  // slots whose base mask is their (disk) pattern
  int max_kc = 0;
  int max_mkc = 0;
  std::vector<int32_t> slots;
};

// Host tables of one search list: per-slot geometry, segment table, radius buckets and
// the layout of the device scratch. They depend on the list, the patterns' extents and
// the search zone's shape only, so a keyed list reuses them across calls.
struct ExactPlan {
  uint64_t key = 0;
  int nlist = 0, U = 0, V = 0, nsteps = 0, vm = 0;
  int nmt = 0, vp = 0, dcap = 0;
  int64_t dl_total = 0;
  std::vector<ExactSlot> slots;
  std::vector<int32_t> segs;
  std::vector<int32_t> msegs;
  std::vector<Bucket> buckets;
  std::vector<int> by_radius, by_work;  // launch orders: increasing radius, decreasing work
  std::vector<int32_t> bucket_cat;
  // a rescue stage's queue to slots: grain -> slot (-1: not in the list), slot -> bucket,
  // bucket -> first position of its slots in bucket_cat
  std::vector<int32_t> slot_of, slot_bucket, bucket_off;
  size_t b_slot_of = 0, b_slot_bucket = 0, b_bucket_off = 0, b_act = 0, b_bcount = 0, b_bact = 0;
  size_t b_mseg = 0, b_mmask = 0;
  size_t b_slots = 0, b_seg = 0, b_bucket = 0, b_tmpl = 0, b_dl = 0, b_dlen = 0, b_mid = 0, b_mom = 0, b_flags = 0;
  size_t b_ovf = 0, b_xout = 0, b_cand = 0;
  size_t total = 0;
  void* dev_base = nullptr;  // scratch the tables were last uploaded to
};
}  // namespace

// Streams the radius buckets' MFMA searches alternate over: the library stream and
// kSearchLanes - 1 more. A bucket's launch ends with a tail of its last workgroups; the
// next bucket, on another stream, fills the CUs it leaves (measured on 2,000 grains in five
// buckets: 3.10 ms for the main stage on one stream, 2.93 ms on two).
constexpr int kSearchLanes = 2;

int Context::ensure_aux(size_t nevents) {
  hipError_t e;
  if (!aux_stream && (e = hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking)) != hipSuccess)
    return check(e, "hipStreamCreateWithFlags");
  while (search_streams.size() < (size_t)(kSearchLanes - 1)) {
    hipStream_t st;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
      return check(e, "hipStreamCreateWithFlags");
    search_streams.push_back(st);
    hipEvent_t ev;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return check(e, "hipEventCreate");
    search_joins.push_back(ev);
  }
  while (aux_events.size() < nevents) {
    hipEvent_t ev;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return check(e, "hipEventCreate");
    aux_events.push_back(ev);
  }
  return PTG_OK;
}

void free_exact_plans(Context& c) {
  for (int st = 0; st < 3; st++) {
    delete static_cast<ExactPlan*>(c.exact_plan[st]);
    c.exact_plan[st] = nullptr;
    c.exact_dev_key[st] = 0;
  }
}

static void build_plan(const Context& c, ExactPlan& P, const std::vector<int32_t>& list, int nlist, int U, int V,
                       int nsteps, int vm) {
  P.nlist = nlist;
  P.U = U;
  P.V = V;
  P.nsteps = nsteps;
  P.vm = vm;
  P.nmt = (nsteps + vm_spt(vm) - 1) / vm_spt(vm);
  P.vp = (V + 15) / 16 * 16;
  const int S = U * V;
  P.slots.assign(nlist, ExactSlot{});
  P.segs.clear();
  int64_t tmpl_total = 0, mid_total = 0, mmask_total = 0;
  int rho_max = 0;
  for (int li = 0; li < nlist; li++) {
    const int g = list[li];
    const int rho = c.h_prad[g];
    rho_max = std::max(rho_max, rho);
    ExactSlot& s = P.slots[li];
    s.rho = rho;
    s.cw = (2 * rho + 1 + 15) / 16 * 16;
    s.cs = ((s.cw / 16) % 2 == 1) ? s.cw : s.cw + 16;
    s.seg_off = (int64_t)P.segs.size();
    for (int qy = -rho; qy <= rho; qy++) {
      const int xm = (int)floor(sqrt((double)(rho * rho - qy * qy)));
      const int c_lo = (rho - xm) / 16 * 16, c_hi = rho + xm;
      for (int c0 = c_lo; c0 <= c_hi; c0 += 16) P.segs.push_back(((qy + rho) << 16) | c0);
    }
    // padding segments (zero template): a whole number of chunk pairs
    while ((P.segs.size() - s.seg_off) % 8) P.segs.push_back(0x8000);
    s.kc = (int)((P.segs.size() - s.seg_off) / 4);
    s.tmpl_off = tmpl_total;
    tmpl_total += (int64_t)P.nmt * s.kc * 1024;
    s.mid_off = mid_total;
    mid_total += 2 * (int64_t)S;
    s.mmask_off = mmask_total;
    mmask_total += ((int64_t)(2 * rho + 1) * (2 * rho + 1) + 15) & ~(int64_t)15;  // (16-byte aligned grids)
    // base-mask GEMM: per mask row, 16-aligned segments covering x in
    // [rho - xm, rho + xm + 15] (the row's cells moved right by up to 15)
    s.mseg_off = (int64_t)P.msegs.size();
    for (int qy = -rho; qy <= rho; qy++) {
      const int xm = (int)floor(sqrt((double)(rho * rho - qy * qy)));
      const int c_lo = (rho - xm) & ~15, c_hi = rho + xm + 15;
      for (int c0 = c_lo; c0 <= c_hi; c0 += 16) P.msegs.push_back(((qy + rho) << 16) | c0);
    }
    while ((P.msegs.size() - s.mseg_off) % 4) P.msegs.push_back(0x8000);
    s.mkc = (int)((P.msegs.size() - s.mseg_off) / 4);
  }
  // Delta capacity per (step, sign): about 4 rho + 64 cells near a base rotation of zero;
  // away from zero the rounded rotated disk has duplicate and empty cells whose positions
  // move with the step, and the lists grow with the pattern and the rotation (a radius-60
  // disk, steps within 0.04 rad: 344 cells at 0.1 rad, 944 at 0.2 rad). At most 1023
  // cells; a longer list sends the grain to the plain kernel
  P.dcap = 0;
  int64_t dl_total = 0;
  for (int li = 0; li < nlist; li++) {
    ExactSlot& s = P.slots[li];
    const int g = list[li];
    const int64_t n = c.h_pbegin[g + 1] - c.h_pbegin[g];
    s.dcap = (int32_t)std::min<int64_t>(1023, std::max<int64_t>(std::min(255, 4 * s.rho + 64), n / 8 + 64));
    s.base_pat = c.h_disk[g] ? 1 : 0;
    s.disk_m2 = c.h_disk[g] ? c.h_diskm2[g] : 0;
    s.dl_off = dl_total;
    dl_total += (int64_t)nsteps * 2 * s.dcap;
    P.dcap = std::max(P.dcap, (int)s.dcap);
  }
  P.dl_total = dl_total;
  // buckets by rho so each launch reserves LDS for its own grains
  const int bounds[] = {13, 21, 29, 37, 45, 53, 61, 69, 1 << 20};
  P.buckets.assign(sizeof(bounds) / sizeof(bounds[0]), Bucket{});
  for (int li = 0; li < nlist; li++) {
    int bk = 0;
    while (P.slots[li].rho > bounds[bk]) bk++;
    Bucket& b = P.buckets[bk];
    b.slots.push_back(li);
    b.rho_max = std::max(b.rho_max, P.slots[li].rho);
    b.max_kc = std::max(b.max_kc, P.slots[li].kc);
    b.max_mkc = std::max(b.max_mkc, P.slots[li].mkc);
    if (P.slots[li].base_pat) b.ndisk++;
  }
  // within a bucket the largest patterns first: their workgroups start first and the launch
  // ends on small ones
  for (Bucket& b : P.buckets)
    std::stable_sort(b.slots.begin(), b.slots.end(), [&](int32_t x, int32_t y) {
      const int32_t gx = list[x], gy = list[y];
      return c.h_pbegin[gx + 1] - c.h_pbegin[gx] > c.h_pbegin[gy + 1] - c.h_pbegin[gy];
    });
  // launch orders: by increasing radius, and by decreasing work (pattern pixels in all: the
  // longest launch starts first and the shorter ones fill the CUs its tail leaves)
  {
    std::vector<int64_t> work(P.buckets.size(), 0);
    for (size_t b = 0; b < P.buckets.size(); b++)
      for (int32_t li : P.buckets[b].slots) work[b] += c.h_pbegin[list[li] + 1] - c.h_pbegin[list[li]];
    P.by_radius.resize(P.buckets.size());
    for (size_t b = 0; b < P.buckets.size(); b++) P.by_radius[b] = (int)b;
    P.by_work = P.by_radius;
    std::stable_sort(P.by_work.begin(), P.by_work.end(), [&](int x, int y) { return work[x] > work[y]; });
  }
  P.bucket_cat.clear();
  P.slot_bucket.assign(nlist, 0);
  P.bucket_off.assign(P.buckets.size(), 0);
  for (size_t b = 0; b < P.buckets.size(); b++) {
    P.bucket_off[b] = (int32_t)P.bucket_cat.size();
    for (int32_t li : P.buckets[b].slots) P.slot_bucket[li] = (int32_t)b;
    P.bucket_cat.insert(P.bucket_cat.end(), P.buckets[b].slots.begin(), P.buckets[b].slots.end());
  }
  P.slot_of.assign(c.G, -1);
  for (int li = 0; li < nlist; li++) P.slot_of[list[li]] = li;
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  P.b_slot_of = al(sizeof(int32_t) * (size_t)c.G);
  P.b_slot_bucket = al(sizeof(int32_t) * (size_t)nlist);
  P.b_bucket_off = al(sizeof(int32_t) * P.buckets.size());
  P.b_act = al(sizeof(int32_t) * (size_t)nlist);
  P.b_bcount = al(sizeof(int32_t) * P.buckets.size());
  P.b_bact = al(sizeof(int32_t) * (size_t)nlist);
  P.b_slots = al(sizeof(ExactSlot) * nlist);
  P.b_seg = al(sizeof(int32_t) * P.segs.size());
  P.b_mseg = al(sizeof(int32_t) * P.msegs.size());
  P.b_bucket = al(sizeof(int32_t) * nlist);
  P.b_tmpl = al((size_t)tmpl_total);
  P.b_dl = al(sizeof(int32_t) * (size_t)P.dl_total);
  P.b_dlen = al(sizeof(int32_t) * 2 * (size_t)nlist * nsteps);
  P.b_mid = al(sizeof(int64_t) * (size_t)mid_total);
  P.b_mom = al(sizeof(int64_t) * 3 * nlist);
  P.b_mmask = al((size_t)mmask_total);
  P.b_flags = al(sizeof(int32_t) * 4);
  P.b_ovf = al(sizeof(int32_t) * 2 * (size_t)nlist);  // per-slot overflow flags, then the overflow list
  P.b_xout = al(sizeof(ExactBest) * (size_t)nlist * std::max(U, nsteps));  // the blocks' near-tie lists
  P.b_cand = al(sizeof(int32_t) * 2 * (size_t)nlist);                   // the grains' sets to re-score
  P.total = P.b_mmask + P.b_slots + P.b_seg + P.b_mseg + P.b_bucket + P.b_tmpl + P.b_dl + P.b_dlen + P.b_mid + P.b_mom +
            P.b_flags + P.b_ovf + P.b_xout + P.b_cand + P.b_slot_of + P.b_slot_bucket + P.b_bucket_off + P.b_act +
            P.b_bcount + P.b_bact;
  P.dev_base = nullptr;
}

int run_search_exact(Context& c, const SearchArgs& sa, int nlist, const std::vector<int32_t>& list,
                     uint64_t plan_key, double stage_thr, StepBest* out_best, int* nparts_out,
                     const int32_t** flag_dev, bool time_kernel, int stage, const int32_t* queue,
                     const int32_t* qcount, const int32_t** act_out) {
  // the value mode from the frames' declared depth; at most 48 rotation steps (40 for 16-bit
  // frames)
  const Image &i0 = c.img[c.iref], &i1 = c.img[c.icur];
  const int bits = std::max(i0.bits, i1.bits);
  const int vm = bits <= 8 ? kVm8 : bits <= 12 ? kVm12 : kVm16;
  const int nsteps = sa.nsteps;
  const int nmt = (nsteps + vm_spt(vm) - 1) / vm_spt(vm);
  if (nmt > (vm == kVm16 ? kExactMaxMt : 6)) return kExactUnfit;
  const int U = sa.left + sa.right + 1, V = sa.up + sa.down + 1;
  if ((int64_t)nsteps * U * V >= INT_MAX / 2) return kExactUnfit;
  // N-tiles per pass: 3 (2 for more than 3 M-tiles, 1 for more than 6: the accumulators'
  // register budget), and no more than the vertical shifts need (a zone of +-8 has 17: 2
  // tiles, not 3)
  const int ntp = std::min(nmt <= 3 ? 3 : nmt <= 6 ? 2 : 1, (V + 15) / 16);
  // the host plan: cached for a keyed list, else built for this call
  ExactPlan local;
  ExactPlan* P = &local;
  if (plan_key != 0) {
    ExactPlan* cached = static_cast<ExactPlan*>(c.exact_plan[stage]);
    if (!cached || cached->key != plan_key || cached->nlist != nlist || cached->U != U || cached->V != V ||
        cached->nsteps != nsteps || cached->vm != vm) {
      if (!cached) {
        cached = new ExactPlan();
        c.exact_plan[stage] = cached;
      }
      build_plan(c, *cached, list, nlist, U, V, nsteps, vm);
      cached->key = plan_key;
      c.exact_dev_key[stage] = 0;
    }
    P = cached;
  } else {
    build_plan(c, local, list, nlist, U, V, nsteps, vm);
  }
  for (const Bucket& bk : P->buckets)
    if (!bk.slots.empty() && bk.rho_max > 120)  // pattern words hold offsets as int8
      return kExactUnfit;
  // launch shapes of every bucket, all checked against the CU's LDS before anything runs
  std::vector<BucketLaunch> L(P->buckets.size());
  for (size_t bi = 0; bi < P->buckets.size(); bi++) {
    const Bucket& bk = P->buckets[bi];
    const int nb = (int)bk.slots.size();
    if (nb == 0) continue;
    BucketLaunch& l = L[bi];
    const int rho = bk.rho_max, side = 2 * rho + 1;
    const int cw = (side + 15) / 16 * 16;
    l.cs = ((cw / 16) % 2 == 1) ? cw : cw + 16;
    // base-mask window sums B_mid, C_mid and the reference moments: as many 16-column
    // blocks per workgroup as keep two workgroups per CU (at least one block); all six
    // digit planes in one pass when they fit the CU's LDS, else two passes of 3 (a
    // pattern too large for that is too large for the MFMA kernel's window as well)
    const int ncb = (U + 15) / 16, ntile_m = P->vp / 16;
    l.cbw = ncb;
    l.npl = kMidPlanes;
    while (l.cbw > 1 && (mid_lds_bytes(rho, P->vp, l.cbw, bk.max_mkc, l.npl) > 80 * 1024 || ntile_m * l.cbw > 16))
      l.cbw--;
    if (mid_lds_bytes(rho, P->vp, l.cbw, bk.max_mkc, l.npl) > kLdsDynMax) l.npl = 3;
    l.lds_m = mid_lds_bytes(rho, P->vp, l.cbw, bk.max_mkc, l.npl);
    // disk base masks from row prefix sums: all vertical shifts per workgroup when that
    // keeps two workgroups per CU, else as many as the LDS holds (none: k_exact_mid's)
    l.vb = 0;
    if (bk.ndisk > 0 && vm != kVm16) {  // (its prefix sums of I1^2 hold 12-bit values)
      int vb = V;
      while (vb > 1 && base_disk_lds_bytes(rho, U, vb) > 80 * 1024) vb--;
      if (vb < 8 && vb < V) {
        vb = V;
        while (vb > 1 && base_disk_lds_bytes(rho, U, vb) > kLdsDynMax) vb--;
      }
      if (base_disk_lds_bytes(rho, U, vb) <= kLdsDynMax) {
        l.vb = vb;
        l.lds_d = base_disk_lds_bytes(rho, U, vb);
      }
    }
    // templates and Delta lists: the step's packed template, the base mask's grid and a flag word
    l.lds_t = (size_t)side * side * 5 + 32;
    // the MFMA search
    l.ww = win_width(U, cw);
    l.wh = P->vp + 2 * rho;
    const size_t win_bytes = ((size_t)l.ww * l.wh * 2 + 15) / 16 * 16;
    // the launch shape: the autotuned one for this bucket's key (or the shape a tuning run
    // probes), else the heuristic's below
    l.key = bucket_key(vm, nsteps, U, V, rho, nb);
    ExactShape want;
    if (auto it = c.exact_tuned.find(l.key); it != c.exact_tuned.end()) want = it->second;
    if (c.tune_run && c.tune_bucket == (int)bi) want = c.tune_shape;
    l.ntp = want.ntp > 0 && want.ntp <= 3 && exact_instance(vm, nmt, want.ntp) ? want.ntp : ntp;
    const size_t stage_bytes = ((size_t)vm_planes(vm) * l.wh * l.cs + 15) / 16 * 16;
    const size_t abuf_bytes = (size_t)(vm == kVm16 ? 16 : 8) * (l.ntp * 16) * 8;
    l.alias = P->vp / 16 <= l.ntp && abuf_bytes <= stage_bytes;
    // wave roles: producer/consumer pairs hand each pass's cross terms of every step over in a slot
    l.ws = (want.roles > 0 ? want.roles : c.exact_roles) == 2;
    const size_t slot_bytes = (size_t)nmt * abuf_bytes;
    const size_t seg_bytes = (size_t)((4 * bk.max_kc + 3) & ~3) * 4;
    // Delta tables: counts, offsets and up to dl_cap entries (+ 4 of slack for the last
    // 16-byte read); typical lists total about 1.5 rho entries per rotation step
    const int nls = 2 * nsteps;
    l.dl_cap = std::min(nls * P->dcap, 2 * rho * nsteps + 4 * nls);
    const size_t dl_bytes = (size_t)(((nls + 3) & ~3) + ((nls + 4) & ~3) + l.dl_cap + 4) * 4;
    auto need = [&](int w) {
      if (l.ws) return win_bytes + w * (stage_bytes + slot_bytes) + 16 + 80 + seg_bytes + dl_bytes;
      return win_bytes + w * (stage_bytes + (l.alias ? 0 : abuf_bytes)) + 16 + seg_bytes + dl_bytes;
    };
    // shift columns split over nsub workgroups when the bucket has few grains
    l.nsub = want.nsub > 0 ? std::min(U, want.nsub) : std::max(1, std::min(U, (512 + nb - 1) / nb));
    const int cpb = (U + l.nsub - 1) / l.nsub;
    // waves per workgroup: the most waves resident per CU under the register limit (a
    // SIMD holds 512 / VGPRs waves) and the LDS limit, fewer and larger workgroups on ties
    const MainAttr ma = main_attr(vm, nmt, l.ntp, l.ws);
    const int wave_cap = 4 * std::max(1, 512 / std::max(ma.vgprs, 1));
    // (producer/consumer: w producers and w consumers per workgroup, at most 4 + 4)
    const int wmax = std::min(l.ws ? 4 : 8, cpb), wpw = l.ws ? 2 : 1;
    int best_res = 0;
    l.nsu = 1;
    for (int w = 1; w <= wmax; w++) {
      if (need(w) > kLdsDynMax) break;
      // waves resident per CU: workgroups by LDS times their waves, under the register limit
      const int res = std::min(wave_cap, (int)(160 * 1024 / (need(w) + ma.static_lds)) * w * wpw);
      if (res >= best_res) {
        best_res = res;
        l.nsu = w;
      }
    }
    if (want.nsu > 0 && want.nsu <= wmax && need(want.nsu) <= kLdsDynMax) l.nsu = want.nsu;
    l.lds = need(l.nsu);
    if (l.lds_m > kLdsDynMax || l.lds_t > kLdsDynMax || l.lds > kLdsDynMax) return kExactUnfit;
  }
  if (c.tune_run) {
    c.tune_keys.assign(P->buckets.size(), 0);
    c.tune_used.assign(P->buckets.size(), ExactShape{});
    for (size_t bi = 0; bi < P->buckets.size(); bi++) {
      c.tune_keys[bi] = L[bi].key;
      c.tune_used[bi] = ExactShape{L[bi].ntp, L[bi].nsu, L[bi].nsub, L[bi].ws ? 2 : 1};
    }
  }
  int rc = c.ensure(&c.d_exact_scratch[stage], &c.exact_scratch_cap[stage], P->total);
  if (rc != PTG_OK) return rc;
  char* base = (char*)c.d_exact_scratch[stage];
  ExactArgs e;
  e.s = sa;
  e.slots = (const ExactSlot*)base;
  base += P->b_slots;
  e.seg = (const int32_t*)base;
  char* seg_dev = base;
  base += P->b_seg;
  e.mseg = (const int32_t*)base;
  char* mseg_dev = base;
  base += P->b_mseg;
  int32_t* bucket_dev = (int32_t*)base;
  base += P->b_bucket;
  e.tmpl = (uint8_t*)base;
  base += P->b_tmpl;
  e.dlist = (int32_t*)base;
  base += P->b_dl;
  e.dlen = (int32_t*)base;
  base += P->b_dlen;
  e.mid = (int64_t*)base;
  base += P->b_mid;
  e.mom = (int64_t*)base;
  base += P->b_mom;
  e.mmask = (uint8_t*)base;
  base += P->b_mmask;
  e.flags = (int32_t*)base;
  base += P->b_flags;
  e.ovf_slot = (int32_t*)base;
  e.ovf_list = e.ovf_slot + nlist;
  base += P->b_ovf;
  e.xout = (ExactBest*)base;
  base += P->b_xout;
  e.cand = (int32_t*)base;
  base += P->b_cand;
  char* slot_of_dev = base;
  base += P->b_slot_of;
  char* slot_bucket_dev = base;
  base += P->b_slot_bucket;
  char* bucket_off_dev = base;
  base += P->b_bucket_off;
  int32_t* act_dev = (int32_t*)base;
  base += P->b_act;
  int32_t* bcount_dev = (int32_t*)base;
  base += P->b_bcount;
  int32_t* bact_dev = (int32_t*)base;
  e.nbucket_dev = nullptr;
  e.s.act = queue ? act_dev : nullptr;
  if (act_out) *act_out = e.s.act;
  // grains whose lists overflow are searched by the plain kernel when its LDS takes the
  // list's largest window and pattern (else an overflow voids the stage)
  const size_t lds_plain = ((size_t)sa.win_cap * 2 + 15) / 16 * 16 + (size_t)sa.off_cap * 4;
  const bool plain_ok = lds_plain <= 159 * 1024;  // the plain kernel's static arrays take the rest
  if (!plain_ok) e.ovf_slot = e.ovf_list = nullptr;
  e.dcap = P->dcap;
  e.nmt = nmt;
  e.vmode = vm;
  e.vmax = vm_vmax(vm);
  e.vp = P->vp;
  hipError_t he;
  // the tables: uploaded unless the scratch already holds this keyed plan's
  const bool fresh = plan_key != 0 && c.exact_dev_key[stage] == plan_key && P->dev_base == c.d_exact_scratch[stage];
  if (!fresh) {
    if (c.capturing) return c.fail(PTG_ERR_STATE, "exact search tables changed during a pair's capture");
    c.table_uploads++;
    if ((he = hipMemcpyAsync((void*)e.slots, P->slots.data(), sizeof(ExactSlot) * nlist, hipMemcpyHostToDevice,
                             c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(slot_of_dev, P->slot_of.data(), sizeof(int32_t) * P->slot_of.size(),
                             hipMemcpyHostToDevice, c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(slot_bucket_dev, P->slot_bucket.data(), sizeof(int32_t) * P->slot_bucket.size(),
                             hipMemcpyHostToDevice, c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(bucket_off_dev, P->bucket_off.data(), sizeof(int32_t) * P->bucket_off.size(),
                             hipMemcpyHostToDevice, c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(seg_dev, P->segs.data(), sizeof(int32_t) * P->segs.size(), hipMemcpyHostToDevice,
                             c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(mseg_dev, P->msegs.data(), sizeof(int32_t) * P->msegs.size(), hipMemcpyHostToDevice,
                             c.stream)) != hipSuccess ||
        (he = hipMemcpyAsync(bucket_dev, P->bucket_cat.data(), sizeof(int32_t) * P->bucket_cat.size(),
                             hipMemcpyHostToDevice, c.stream)) != hipSuccess)
      return c.check(he, "exact search tables");
    if (plan_key != 0) {
      // the cached host tables outlive the copies
      P->dev_base = c.d_exact_scratch[stage];
      c.exact_dev_key[stage] = plan_key;
    } else {
      // this call's tables are freed on return: let the copies finish first
      if ((he = hipStreamSynchronize(c.stream)) != hipSuccess) return c.check(he, "hipStreamSynchronize");
      c.exact_dev_key[stage] = 0;
    }
  }
  // shift blocks: results of block b of slot li go to out_best[li * part_cap + b]; the
  // plain kernel writes step k of an overflowed slot to out_best[li * part_cap + k]; the
  // rest stay "none"
  const int part_cap = std::max(U, nsteps);
  e.part_cap = part_cap;
  e.out = out_best;
  e.thr = stage_thr;
  *nparts_out = part_cap;
  {
    const int nb_tab = (int)P->buckets.size();
    const int n_all = std::max(nlist * part_cap, std::max(nlist, nb_tab));
    k_exact_reset<<<(n_all + 255) / 256, 256, 0, c.stream>>>(out_best, e.xout, nlist * part_cap, e.ovf_slot,
                                                             queue ? act_dev : nullptr, nlist, e.flags,
                                                             queue ? bcount_dev : nullptr, nb_tab);
    if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_reset");
  }
  if (queue) {
    // the queued grains' slots, per radius bucket
    k_exact_activate<<<(nlist + 255) / 256, 256, 0, c.stream>>>(
        queue, qcount, (const int32_t*)slot_of_dev, (const int32_t*)slot_bucket_dev, (const int32_t*)bucket_off_dev,
        act_dev, bcount_dev, bact_dev);
    if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_activate");
  }
  // the template and base-mask kernels of every bucket run in order on the second
  // stream; each bucket's MFMA kernel runs after its bucket's event, on the library stream
  // and the search streams in turn, so the next bucket's preparation overlaps the current
  // search and consecutive searches overlap their tails
  // (with the serial option every kernel runs in order on the library stream instead, so
  // that per-launch timings do not overlap)
  const size_t nb_all = P->buckets.size();
  const bool serial = c.serial_search;
  if (c.tune_run) {
    // a tuning run: GPU events around the whole search on the library stream
    if (!c.tune_ev[0] && (!(c.tune_ev[0] = c.get_event()) || !(c.tune_ev[1] = c.get_event())))
      return c.fail(PTG_ERR_HIP, "hipEventCreate");
    if ((he = hipEventRecord(c.tune_ev[0], c.stream)) != hipSuccess) return c.check(he, "hipEventRecord");
  }
  const std::vector<int>& order = c.bucket_order == 1 ? P->by_work : P->by_radius;
  if (int rca = c.ensure_aux(nb_all + 1); rca != PTG_OK) return rca;
  hipStream_t prep = serial ? c.stream : c.aux_stream;
  hipEvent_t fork = c.aux_events[nb_all];
  if (!serial && ((he = hipEventRecord(fork, c.stream)) != hipSuccess ||
                  (he = hipStreamWaitEvent(c.aux_stream, fork, 0)) != hipSuccess))
    return c.check(he, "exact search fork");
  hipStream_t lanes[kSearchLanes];
  const int nlanes = serial ? 1 : kSearchLanes;
  lanes[0] = c.stream;
  for (int q = 1; q < nlanes; q++) {
    lanes[q] = c.search_streams[q - 1];
    // the tables and cleared outputs above come first
    if ((he = hipStreamWaitEvent(lanes[q], fork, 0)) != hipSuccess) return c.check(he, "hipStreamWaitEvent");
  }
  if (c.chain_pending) {
    // the precomputation's ordered sums beside the searches: on the last search lane ahead of its
    // searches, joined back to the library stream with them before k_exact_rescore reads C0C0
    c.chain_pending = false;
    if ((he = launch_precompute_chain(c, lanes[nlanes - 1])) != hipSuccess) return c.check(he, "k_precompute_chain");
  }
  // each bucket's preparation, then its search, enqueued in turn (the first search is queued as
  // soon as its own preparation, not after every bucket's)
  int nlaunch = 0;
  for (const int bi : order) {
    const Bucket& bk = P->buckets[bi];
    const int nb = (int)bk.slots.size();
    if (nb == 0) continue;
    const BucketLaunch& l = L[bi];
    e.bucket = (queue ? bact_dev : bucket_dev) + P->bucket_off[bi];
    e.nbucket = nb;
    e.nbucket_dev = queue ? bcount_dev + bi : nullptr;
    const int rho = bk.rho_max;
    if (l.vb > 0) {
      if ((he = ensure_kernel_lds(c.device, (const void*)k_exact_base_disk, l.lds_d)) != hipSuccess)
        return c.check(he, "hipFuncSetAttribute");
      ExactArgs ed = e;
      ed.base_vb = l.vb;
      k_exact_base_disk<<<dim3(nb, (V + l.vb - 1) / l.vb), 512, l.lds_d, prep>>>(ed);
      if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_base_disk");
    }
    if (l.vb == 0 || bk.ndisk < nb) {
      const void* fm = l.npl == 6 ? (const void*)k_exact_mid<6> : (const void*)k_exact_mid<3>;
      if ((he = ensure_kernel_lds(c.device, fm, l.lds_m)) != hipSuccess) return c.check(he, "hipFuncSetAttribute");
      ExactArgs em = e;
      em.mid_skip_disk = l.vb > 0 ? 1 : 0;
      em.mid_cbw = l.cbw;
      em.mid_stride = mid_stride(rho, l.cbw);
      em.mid_mkc_cap = bk.max_mkc;
      em.mid_npl = l.npl;
      const int ncb = (U + 15) / 16, ntile_m = P->vp / 16;
      const dim3 grid_m(nb, (ncb + l.cbw - 1) / l.cbw), block_m(64 * std::min(16, ntile_m * l.cbw));
      if (l.npl == 6) k_exact_mid<6><<<grid_m, block_m, l.lds_m, prep>>>(em);
      else k_exact_mid<3><<<grid_m, block_m, l.lds_m, prep>>>(em);
      if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_mid");
    }
    if ((he = ensure_kernel_lds(c.device, (const void*)k_exact_template, l.lds_t)) != hipSuccess)
      return c.check(he, "hipFuncSetAttribute");
    k_exact_template<<<dim3(nb, vm_spt(vm) * nmt), 256, l.lds_t, prep>>>(e);
    if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_template");
    if ((he = hipEventRecord(c.aux_events[bi], prep)) != hipSuccess) return c.check(he, "hipEventRecord");
    // the search
    hipStream_t sx = lanes[nlaunch++ % nlanes];
    if (!serial && (he = hipStreamWaitEvent(sx, c.aux_events[bi], 0)) != hipSuccess)
      return c.check(he, "hipStreamWaitEvent");
    ExactArgs eb = e;
    eb.ww = l.ww;
    eb.wh_pad = l.wh;
    eb.nsub = l.nsub;
    eb.nsu = l.nsu;
    eb.cs_max = l.cs;
    eb.dl_lds_cap = l.dl_cap;
    eb.seg_cap = bk.max_kc;
    eb.abuf_alias = l.alias ? 1 : 0;
    // the grid: nsub workgroups per grain, or for a queue the most any queued count needs
    int nwg = 8 * ((nb + 7) / 8) * eb.nsub;
    if (queue) {
      nwg = 0;
      for (int n = 1; n <= nb; n++) nwg = std::max(nwg, 8 * ((n + 7) / 8) * exact_nsub_of(n, U));
    }
    dim3 grid(nwg), block(64 * eb.nsu * (l.ws ? 2 : 1));
    hipEvent_t ev = nullptr;
    if (time_kernel) c.mark_begin(PTG_T_SEARCH_KERNEL, &ev, sx);
    if ((he = dispatch_main(c.device, vm, nmt, l.ntp, l.ws, eb, grid, block, l.lds, sx)) != hipSuccess)
      return c.check(he, "k_search_exact");
    if (time_kernel) c.mark_end(PTG_T_SEARCH_KERNEL, ev, sx);
  }
  for (int q = 1; q < nlanes; q++) {
    if ((he = hipEventRecord(c.search_joins[q - 1], lanes[q])) != hipSuccess ||
        (he = hipStreamWaitEvent(c.stream, c.search_joins[q - 1], 0)) != hipSuccess)
      return c.check(he, "exact search join");
  }
  // every grain's near-tie set, re-scored in the reference's arithmetic into out_best (the
  // grains the merge hands over join the plain kernel's list)
  {
    ExactArgs ea = e;
    ea.bucket = bucket_dev;  // the buckets' slot lists, concatenated (a queue's marked in act)
    ea.nbucket = nlist;
    ea.nbucket_dev = nullptr;
    k_exact_merge<<<(nlist + kMergeWaves - 1) / kMergeWaves, 64 * kMergeWaves, 0, c.stream>>>(ea);
    if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_merge");
    k_exact_rescore<<<nlist, kRescoreThreads, 0, c.stream>>>(ea, out_best);
    if ((he = hipGetLastError()) != hipSuccess) return c.check(he, "k_exact_rescore");
  }
  if (c.tune_run && (he = hipEventRecord(c.tune_ev[1], c.stream)) != hipSuccess) return c.check(he, "hipEventRecord");
  if (e.ovf_slot) {
    if ((he = launch_search_plain_sub(c, sa, nlist, e.ovf_list, e.flags + 1, part_cap, lds_plain)) != hipSuccess)
      return c.check(he, "k_search_plain_sub");
  }
  // a template digit or Delta list overflow leaves a flag: the commit then changes nothing
  // and the caller reports PTG_ERR_LIMIT at its next synchronisation
  *flag_dev = e.flags;
  return PTG_OK;
}

}  // namespace ptg
