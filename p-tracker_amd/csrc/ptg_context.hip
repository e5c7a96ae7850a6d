// ptg_context.hip — the C ABI of libptgpu: context, frames, grains, and the per-pair
// stage sequence of p-tracker's particle_tracking() (src/p-tracker.cpp:460-581).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <omp.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "ptg_internal.hpp"

using namespace ptg;

namespace ptg {

int Context::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  err = buf;
  return code;
}

int Context::check(hipError_t e, const char* what) {
  if (e == hipSuccess) return PTG_OK;
  return fail(PTG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int Context::set_frames() {
  const Image &i0 = img[iref], &i1 = img[icur];
  FrameDesc f{};
  f.I0 = i0.d;
  f.I1 = i1.d;
  f.pitch0 = i0.pitch;
  f.pitch1 = i1.pitch;
  f.dimx = i1.dimx;
  f.dimy = i1.dimy;
  f.dimx0 = i0.dimx;
  f.dimy0 = i0.dimy;
  if (frames_written && f.I0 == h_frames.I0 && f.I1 == h_frames.I1 && f.pitch0 == h_frames.pitch0 &&
      f.pitch1 == h_frames.pitch1 && f.dimx == h_frames.dimx && f.dimy == h_frames.dimy &&
      f.dimx0 == h_frames.dimx0 && f.dimy0 == h_frames.dimy0)
    return PTG_OK;
  hipError_t e = launch_set_frames(*this, f);
  if (e != hipSuccess) return check(e, "k_set_frames");
  h_frames = f;
  frames_written = true;
  return PTG_OK;
}

int Context::ensure(void** p, size_t* cap, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (*p && *cap >= bytes) return PTG_OK;
  if (*p) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return check(e, "hipStreamSynchronize");
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(PTG_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  *cap = bytes;
  return PTG_OK;
}

hipEvent_t Context::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Events around work on stream `on` (the library stream when null).
void Context::mark_begin(int stage, hipEvent_t* ev, hipStream_t on) {
  *ev = nullptr;
  if (!timing) return;
  *ev = get_event();
  if (*ev) (void)hipEventRecord(*ev, on ? on : stream);
  (void)stage;
}

void Context::mark_end(int stage, hipEvent_t ev, hipStream_t on) {
  if (!timing || !ev) return;
  hipEvent_t e2 = get_event();
  if (!e2) return;
  (void)hipEventRecord(e2, on ? on : stream);
  pending.push_back({stage, ev, e2});
}

// Fills n int32 with a value (used instead of hipMemsetAsync so a captured stream
// records a kernel node).
__global__ void k_fill_i32(int32_t* p, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static hipError_t fill_i32(Context& c, int32_t* p, int n, int32_t v) {
  k_fill_i32<<<(n + 255) / 256, 256, 0, c.stream>>>(p, n, v);
  return hipGetLastError();
}

__global__ void k_fill_i64(long long* p, int n, long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t ensure_kernel_lds(int dev, const void* fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> set;
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = set[{dev, fn}];
  if (bytes <= have) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) have = bytes;
  return e;
}

}  // namespace ptg

// ---- staging of the allocation-free objective entry point ---------------------------------
// ptg_ncc_objective_batch_noalloc stages its grain indices, points and values in this
// file-scope array instead of memory the context allocates: at most kNoAllocMaxEvals
// evaluations per call. Each device holds its own copy until the process ends; host code
// uses only the per-device address that hipGetSymbolAddress returns (the array's name is a
// host-side stand-in, not a device address).
constexpr int32_t kNoAllocMaxEvals = 16384;
constexpr size_t kNoAllocIgBytes = sizeof(int32_t) * kNoAllocMaxEvals;  // a multiple of 256
constexpr size_t kNoAllocXBytes = sizeof(double) * 3 * kNoAllocMaxEvals;
constexpr size_t kNoAllocFBytes = sizeof(double) * kNoAllocMaxEvals;
static __device__ __align__(256) unsigned char scratch[kNoAllocIgBytes + kNoAllocXBytes + kNoAllocFBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on `device` (current), looked up once per device.
hipError_t scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(Context& c, size_t bytes) {
    if (p && cap >= bytes) return PTG_OK;
    if (p) {
      (void)hipStreamSynchronize(c.stream);
      (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
    }
    if (hipHostMalloc(&p, bytes < 64 ? 64 : bytes, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return c.fail(PTG_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
    }
    cap = bytes;
    return PTG_OK;
  }
};

// Words of pin_flags: [2 stage, 2 stage + 1] a stage's capacity flags and plain-served grains,
// [6 + stage] a device-queue stage's count, [9] the rescue queue's count read by track_pair.
constexpr int kPinFlags = 10;

// The context as seen through the C ABI, with the host-side staging it owns.
struct Ctx : Context {
  PinnedBuf pin_cs, pin_steps, pin_cs0, pin_list, pin_flags;
  // capacity flags of exact searches and their count of grains served by the plain
  // kernel, copied to pin_flags[2 stage], pin_flags[2 stage + 1] and read at the next
  // synchronisation of the stream
  bool flags_pending[3] = {false, false, false};
  // a rescue stage whose queue stayed on the device ran the plain form in place of the exact
  // one: its queue count, copied to pin_flags[6 + stage], says whether a search took place
  bool unfit_pending[3] = {false, false, false};
  int check_flags() {
    for (int st = 0; st < 3; st++) {
      if (!unfit_pending[st]) continue;
      unfit_pending[st] = false;
      if (static_cast<const int32_t*>(pin_flags.p)[6 + st] > 0) exact_fallbacks++;
    }
    for (int st = 0; st < 3; st++) {
      if (!flags_pending[st]) continue;
      flags_pending[st] = false;
      const int32_t f = static_cast<const int32_t*>(pin_flags.p)[2 * st];
      exact_grain_fallbacks += static_cast<const int32_t*>(pin_flags.p)[2 * st + 1];
      if (f)
        return fail(PTG_ERR_LIMIT,
                    "exact search capacity exceeded in stage %d (flags %d; 4: a producer/consumer hand-over "
                    "wait expired): the stage changed nothing; use PTG_SEARCH_PLAIN",
                    st, f);
    }
    return PTG_OK;
  }
  int sync_and_check() {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return check(e, "hipStreamSynchronize");
    return check_flags();
  }
  bool any_flags_pending() const {
    return flags_pending[0] || flags_pending[1] || flags_pending[2] || unfit_pending[0] || unfit_pending[1] ||
           unfit_pending[2];
  }
  enum { SNAP_INVALID = 0, SNAP_VALID = 1, SNAP_QUEUES = 2 } snap = SNAP_INVALID;
  std::vector<double> h_steps;
  // recorded after the last upload from the pinned staging buffers, waited on before
  // the host rewrites them
  hipEvent_t staging_done = nullptr;
  // the rescue queue's count copied for the host while the subpixel stage runs (track_pair), and
  // the stream that subpixel stage runs on beside the rescue stages, with its fork and join
  hipEvent_t count_ev = nullptr, sp_fork = nullptr, sp_join = nullptr;
  hipStream_t sp_stream = nullptr;
  int wait_staging() {
    if (staging_done) {
      hipError_t e = hipEventSynchronize(staging_done);
      if (e != hipSuccess) return check(e, "hipEventSynchronize");
    }
    return PTG_OK;
  }
  int mark_staging() {
    if (!staging_done && hipEventCreateWithFlags(&staging_done, hipEventDisableTiming) != hipSuccess)
      return fail(PTG_ERR_HIP, "hipEventCreate failed");
    return check(hipEventRecord(staging_done, stream), "hipEventRecord");
  }
  // Frame copies that overlap the context's work (ptg_set_image_async): a copy stream (made
  // at the first such copy), and per image slot the event recorded after its last copy.
  // up_pending: the library stream has not yet been told to wait for that copy.
  hipStream_t up_stream = nullptr;
  hipEvent_t up_after = nullptr;  // the library stream's work so far, waited on by a copy
  hipEvent_t up_done[kMaxImages] = {};
  bool up_pending[kMaxImages] = {};
  int upload_setup() {
    if (up_stream) return PTG_OK;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return check(e, "hipStreamCreateWithFlags (frame copies)");
    hipEvent_t ev[kMaxImages + 1] = {};
    for (int i = 0; i <= kMaxImages && e == hipSuccess; i++) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
      for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
      (void)hipStreamDestroy(s);
      return check(e, "hipEventCreateWithFlags (frame copies)");
    }
    up_after = ev[kMaxImages];
    for (int i = 0; i < kMaxImages; i++) up_done[i] = ev[i];
    up_stream = s;
    return PTG_OK;
  }
  // Work queued from now on waits on the GPU for image `index`'s last copy.
  int order_after_upload(int index) {
    if (index < 0 || index >= kMaxImages || !up_pending[index]) return PTG_OK;
    up_pending[index] = false;
    return check(hipStreamWaitEvent(stream, up_done[index], 0), "hipStreamWaitEvent (frame copy)");
  }
  // The host waits until image `index`'s last copy is done.
  int finish_upload(int index) {
    if (index < 0 || index >= kMaxImages || !up_done[index]) return PTG_OK;
    return check(hipEventSynchronize(up_done[index]), "hipEventSynchronize (frame copy)");
  }
  int finish_uploads() {
    if (!up_stream) return PTG_OK;
    return check(hipStreamSynchronize(up_stream), "hipStreamSynchronize (frame copies)");
  }
};

Ctx* as_ctx(void* p) { return reinterpret_cast<Ctx*>(p); }


#define CK(c, expr)                                                  \
  do {                                                               \
    hipError_t e_ = (expr);                                          \
    if (e_ != hipSuccess) return (c)->check(e_, #expr);              \
  } while (0)

#define TRY(expr)             \
  do {                        \
    int r_ = (expr);          \
    if (r_ != PTG_OK) return r_; \
  } while (0)

int valid_index(Ctx* c, int index) {
  if (index < 0 || index >= kMaxImages) return c->fail(PTG_ERR_ARG, "image index %d outside [0, %d)", index, kMaxImages);
  return PTG_OK;
}

// Rotation steps of src/p-tracker.cpp:684-686 in double, as the reference.
int make_steps(Ctx* c, const ptg_search_zone* z, std::vector<double>& out) {
  out.clear();
  if (!(z->inc_rot > 0.0))
    return c->fail(PTG_ERR_ARG, "inc_rot must be > 0 (the reference loop does not end otherwise; "
                                "`rotations 0` sets it to 1.0)");
  if (z->num_rot < 0) return c->fail(PTG_ERR_ARG, "num_rot must be >= 0");
  double rotmax = fabs(z->inc_rot * z->num_rot);
  for (double drot = -rotmax; drot <= rotmax; drot += z->inc_rot) {
    out.push_back(drot);
    if (out.size() > 100000) return c->fail(PTG_ERR_LIMIT, "more than 100000 rotation steps");
  }
  return PTG_OK;
}

int sync_snapshot(Ctx* c) {
  if (c->G == 0) return PTG_OK;
  CK(c, hipMemcpyAsync(c->h_grains.data(), c->d_grains, sizeof(ptg_grain) * c->G, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  c->snap = Ctx::SNAP_VALID;
  return c->check_flags();
}

// Grains were set (ptg_set_grains); zero grains is a valid set, on which the stage calls
// do nothing, as the reference's loops over an empty grain list do.
int need_grains(Ctx* c) {
  if (!c->grains_set) return c->fail(PTG_ERR_STATE, "no grains set (ptg_set_grains)");
  return PTG_OK;
}

int need_images(Ctx* c, bool ref, bool cur) {
  if (ref && !c->img[c->iref].d) return c->fail(PTG_ERR_STATE, "image[%d] (reference) not set", c->iref);
  if (cur && !c->img[c->icur].d) return c->fail(PTG_ERR_STATE, "image[%d] (current) not set", c->icur);
  if (ref && cur) {
    const Image &a = c->img[c->iref], &b = c->img[c->icur];
    if (a.dimx != b.dimx || a.dimy != b.dimy)
      return c->fail(PTG_ERR_ARG, "image sizes differ: %dx%d vs %dx%d", a.dimx, a.dimy, b.dimx, b.dimy);
  }
  // the work the caller queues next reads these frames: after their copies
  if (ref) TRY(c->order_after_upload(c->iref));
  if (cur) TRY(c->order_after_upload(c->icur));
  return PTG_OK;
}

// The ordered sums of the precomputation when they are still to be launched: on the context's
// stream, before any reader.
int flush_chain(Ctx* c) {
  if (!c->chain_pending) return PTG_OK;
  c->chain_pending = false;
  CK(c, launch_precompute_chain(*c, c->stream));
  return PTG_OK;
}

// defer: leave the ordered sums (C0C0, the sum of d0, d0) to the main stage's exact search, which
// launches them beside its searches (chain_pending until then)
int do_precompute(Ctx* c, bool defer) {
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  TRY(need_images(c, true, false));
  TRY(c->set_frames());
  if (!c->rot_host) {
    // cos/sin(refrot) on the device, guarded (checked by the caller after its stages)
    hipEvent_t ev;
    c->mark_begin(PTG_T_PRECOMPUTE, &ev);
    CK(c, launch_precompute(*c, true, !defer));
    c->mark_end(PTG_T_PRECOMPUTE, ev);
    c->chain_pending = defer;
    c->require_precomputations = false;
    return PTG_OK;
  }
  // cos/sin(refrot) per grain from the host's libm, as do_precomputations (:636-637)
  if (c->snap == Ctx::SNAP_INVALID) TRY(sync_snapshot(c));
  TRY(c->wait_staging());
  TRY(c->pin_cs0.ensure(*c, sizeof(double) * 2 * c->G));
  double* cs0 = (double*)c->pin_cs0.p;
  for (int g = 0; g < c->G; g++) {
    cs0[2 * g] = cos(c->h_grains[g].refrot);
    cs0[2 * g + 1] = sin(c->h_grains[g].refrot);
  }
  hipEvent_t ev;
  c->mark_begin(PTG_T_PRECOMPUTE, &ev);
  CK(c, hipMemcpyAsync(c->d_cs0, cs0, sizeof(double) * 2 * c->G, hipMemcpyHostToDevice, c->stream));
  TRY(c->mark_staging());
  CK(c, launch_precompute(*c, false, true));
  c->mark_end(PTG_T_PRECOMPUTE, ev);
  c->require_precomputations = false;
  return PTG_OK;
}

constexpr size_t kMaxLds = 160 * 1024 - 1024;  // dynamic LDS; the rest for the kernels' static arrays

// Rotation tables of a stage from the host's libm: cos/sin(refrot + drot + step) per (grain,
// step) of the list, evaluated as the reference (:687-689), uploaded with the step values.
static int host_rot_tables(Ctx* c, const std::vector<int32_t>& list, int stage) {
  const int nlist = (int)list.size(), nsteps = (int)c->h_steps.size();
  TRY(c->pin_cs.ensure(*c, sizeof(double) * 2 * (size_t)c->G * nsteps));
  double* cs = (double*)c->pin_cs.p;
  const double* steps = c->h_steps.data();
  // the CPUs this process may use less four, left to the GPU runtime's threads and the
  // caller's (with every CPU taken, one preempted thread holds the whole loop for a
  // scheduler time slice); chunks of 8 grains taken in turn
  const int nthr = std::max(1, std::min(omp_get_max_threads(), omp_get_num_procs() - 4));
  (void)nthr;  // (the device compilation pass ignores the pragma)
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthr)
  for (int li = 0; li < nlist; li++) {
    const int g = list[li];
    const ptg_grain& gr = c->h_grains[g];
    double* row = cs + 2 * (size_t)g * nsteps;
    for (int k = 0; k < nsteps; k++) {
      double total_rot = gr.refrot + gr.drot + steps[k];
      row[2 * k] = cos(total_rot);
      row[2 * k + 1] = sin(total_rot);
    }
  }
  // upload only the rows in use: for the main stage that is nearly all of them
  if (stage == PTG_STAGE_MAIN) {
    CK(c, hipMemcpyAsync(c->d_cs, cs, sizeof(double) * 2 * (size_t)c->G * nsteps, hipMemcpyHostToDevice, c->stream));
  } else {
    for (int li = 0; li < nlist; li++) {
      const size_t off = 2 * (size_t)list[li] * nsteps;
      CK(c, hipMemcpyAsync(c->d_cs + off, cs + off, sizeof(double) * 2 * nsteps, hipMemcpyHostToDevice, c->stream));
    }
  }
  return c->mark_staging();
}

// The stage's rotation-step values on the device, uploaded when they change (each stage keeps
// its own, where a captured pair's kernels read them).
static int upload_steps(Ctx* c, int stage) {
  if (c->d_steps_st[stage] && c->h_steps_st[stage] == c->h_steps) return PTG_OK;
  if (c->capturing) return c->fail(PTG_ERR_STATE, "rotation steps changed during a pair's capture");
  const size_t bytes = sizeof(double) * c->h_steps.size();
  TRY(c->wait_staging());
  TRY(c->pin_steps.ensure(*c, bytes));
  memcpy(c->pin_steps.p, c->h_steps.data(), bytes);
  TRY(c->ensure((void**)&c->d_steps_st[stage], &c->steps_cap_st[stage], bytes));
  CK(c, hipMemcpyAsync(c->d_steps_st[stage], c->pin_steps.p, bytes, hipMemcpyHostToDevice, c->stream));
  c->h_steps_st[stage] = c->h_steps;
  c->table_uploads++;
  return c->mark_staging();
}

int do_search(Ctx* c, int stage, const ptg_search_zone* z, double NCC_min, double NCC_min_super, int method) {
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  TRY(need_images(c, true, true));
  if (!z) return c->fail(PTG_ERR_ARG, "null search zone");
  if (stage < PTG_STAGE_MAIN || stage > PTG_STAGE_SUPER) return c->fail(PTG_ERR_ARG, "bad stage %d", stage);
  if (method != PTG_SEARCH_PLAIN && method != PTG_SEARCH_EXACT) return c->fail(PTG_ERR_ARG, "bad method %d", method);
  if (z->left < 0 || z->right < 0 || z->up < 0 || z->down < 0)
    return c->fail(PTG_ERR_ARG, "search zone extents must be >= 0");
  if ((int64_t)(z->left + z->right + 1) * (z->up + z->down + 1) > (1 << 24))
    return c->fail(PTG_ERR_LIMIT, "search zone too large");
  if (c->require_precomputations) TRY(do_precompute(c, false));
  TRY(c->set_frames());
  TRY(make_steps(c, z, c->h_steps));
  const int nsteps = (int)c->h_steps.size();

  // the stage's grain list on the host
  std::vector<int32_t> list;
  int32_t* d_list = c->d_list[stage];
  int32_t* d_count = c->d_counts + stage;
  // next queue counter starts at zero (enqueued ahead of the wait for this stage's list)
  if (stage < PTG_STAGE_SUPER) CK(c, fill_i32(*c, c->d_counts + stage + 1, 1, 0));
  // a rescue stage on device tables keeps its queue on the device: its exact search is planned
  // over the unmasked list and serves the queued grains among them (no host round trip)
  const bool devq = stage != PTG_STAGE_MAIN && !c->rot_host && (c->device_queues || c->graph_pair);
  if (stage == PTG_STAGE_MAIN) {
    if (c->rot_host && c->snap != Ctx::SNAP_VALID) TRY(sync_snapshot(c));
    list = c->h_unmasked;  // built at set/put time
  } else if (devq) {
    list = c->h_unmasked;  // (the launches are sized for it; the queue is read on the device)
  } else {
    // the queue was filled on the device by the previous stage
    if (c->rot_host && c->snap == Ctx::SNAP_INVALID) TRY(sync_snapshot(c));
    int32_t n = 0;
    CK(c, hipMemcpyAsync(&n, d_count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    TRY(c->sync_and_check());
    list.resize(n);
    if (n) CK(c, hipMemcpyAsync(list.data(), d_list, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
  }
  const int nlist = (int)list.size();
  if (nlist == 0) return flush_chain(c);

  // rotation tables: cos/sin(refrot + drot + step) per (grain, step), evaluated as the
  // reference (:687-689), on the device (guarded) or from the host's libm
  TRY(c->ensure((void**)&c->d_cs, &c->cs_cap, sizeof(double) * 2 * (size_t)c->G * nsteps));
  const int nparts_cap = std::max(nsteps, z->left + z->right + 1);
  TRY(c->ensure((void**)&c->d_best, &c->best_cap, sizeof(StepBest) * (size_t)nlist * nparts_cap));
  TRY(upload_steps(c, stage));
  if (!c->rot_host) {
    RotArgs r;
    r.grains = c->d_grains;
    r.pattern = c->d_pattern;
    r.pbegin = c->d_pbegin;
    r.list = d_list;
    r.count = d_count;
    r.steps = c->d_steps_st[stage];
    r.nsteps = nsteps;
    r.all = c->rot_guard_test ? 1 : 0;
    r.cs = c->d_cs;
    r.guard = c->d_guard;
    r.guard_count = c->d_guard_count;
    CK(c, launch_rot_tables(*c, r, nlist));
  } else {
    TRY(c->wait_staging());
    TRY(host_rot_tables(c, list, stage));
  }

  // LDS of the plain kernel: the largest window and pattern of the list; a window beyond
  // the LDS is searched in square tiles of shifts
  int64_t win_cap = 0, off_cap = 0;
  int rho_max = 0;
  const int U = z->left + z->right + 1, V = z->up + z->down + 1;
  for (int li = 0; li < nlist; li++) {
    const int g = list[li];
    rho_max = std::max(rho_max, c->h_prad[g]);
    off_cap = std::max<int64_t>(off_cap, c->h_pbegin[g + 1] - c->h_pbegin[g]);
  }
  int tile_u = U, tile_v = V;
  auto plain_lds = [&](int tu, int tv) {
    return ((size_t)(tu + 2 * rho_max) * (tv + 2 * rho_max) * 2 + 15) / 16 * 16 + (size_t)off_cap * 4;
  };
  if (plain_lds(U, V) > kMaxLds) {
    int t = std::max(U, V);
    while (t > 1 && plain_lds(std::min(t, U), std::min(t, V)) > kMaxLds) t--;
    tile_u = std::min(t, U);
    tile_v = std::min(t, V);
  }
  win_cap = (int64_t)(tile_u + 2 * rho_max) * (tile_v + 2 * rho_max);

  SearchArgs a;
  a.fr = c->d_frames;
  a.grains = c->d_grains;
  a.pattern = c->d_pattern;
  a.pbegin = c->d_pbegin;
  a.p0 = c->d_p0;
  a.d0 = c->d_d0;
  a.pwords = c->d_pwords;
  a.prad = c->d_prad;
  a.list = d_list;
  a.count = d_count;
  a.act = nullptr;
  a.cs = c->d_cs;
  a.steps = c->d_steps_st[stage];
  a.nsteps = nsteps;
  a.left = z->left;
  a.right = z->right;
  a.up = z->up;
  a.down = z->down;
  a.win_cap = (int32_t)win_cap;
  a.off_cap = (int32_t)off_cap;
  a.tile_u = tile_u;
  a.tile_v = tile_v;
  a.best = c->d_best;

  const int tstage = PTG_T_SEARCH_MAIN + stage;
  hipEvent_t ev;
  c->mark_begin(tstage, &ev);
  int nparts = nsteps;
  const int32_t* flag_dev = nullptr;
  SearchArgs ca = a;  // the commit's list: the stage's, or the plan's with the queued slots marked
  if (method != PTG_SEARCH_PLAIN) {
    // the unmasked list is the same on every call until it is rebuilt
    const uint64_t key = stage == PTG_STAGE_MAIN || devq ? c->unmasked_version : 0;
    const double thr = stage == PTG_STAGE_MAIN ? NCC_min : NCC_min_super;
    SearchArgs ax = a;
    if (devq) {
      ax.list = c->d_list[0];
      ax.count = c->d_counts;
    }
    const int32_t* act = nullptr;
    const int rc = run_search_exact(*c, ax, nlist, list, key, thr, c->d_best, &nparts, &flag_dev,
                                    stage == PTG_STAGE_MAIN, stage, devq ? d_list : nullptr, devq ? d_count : nullptr,
                                    &act);
    if (rc == kExactUnfit) {
      method = PTG_SEARCH_PLAIN;
      nparts = nsteps;
      if (devq) {
        // counted at the next synchronisation if the queue held grains
        TRY(c->pin_flags.ensure(*c, kPinFlags * sizeof(int32_t)));
        CK(c, hipMemcpyAsync(static_cast<int32_t*>(c->pin_flags.p) + 6 + stage, d_count, sizeof(int32_t),
                             hipMemcpyDeviceToHost, c->stream));
        c->unfit_pending[stage] = true;
      } else {
        c->exact_fallbacks++;
      }
    } else {
      ca = ax;
      ca.act = act;
      TRY(rc);
      TRY(c->pin_flags.ensure(*c, kPinFlags * sizeof(int32_t)));
      CK(c, hipMemcpyAsync(static_cast<int32_t*>(c->pin_flags.p) + 2 * stage, flag_dev, 2 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c->stream));
      c->flags_pending[stage] = true;
    }
  }
  TRY(flush_chain(c));  // (the exact search did not take them: before any reader)
  if (method == PTG_SEARCH_PLAIN) {
    // a pattern whose offsets and smallest window tile exceed the LDS: the kernel that reads
    // the frame through the caches
    const size_t lds = (size_t)(win_cap * 2 + 15) / 16 * 16 + (size_t)off_cap * 4;
    hipEvent_t ek;
    if (stage == PTG_STAGE_MAIN) c->mark_begin(PTG_T_SEARCH_KERNEL, &ek);
    if (lds > kMaxLds) CK(c, launch_search_plain_global(*c, a, nlist));
    else CK(c, launch_search_plain(*c, a, nlist, lds));
    if (stage == PTG_STAGE_MAIN) c->mark_end(PTG_T_SEARCH_KERNEL, ek);
  }
  int32_t* next_list = stage < PTG_STAGE_SUPER ? c->d_list[stage + 1] : nullptr;
  int32_t* next_count = stage < PTG_STAGE_SUPER ? c->d_counts + stage + 1 : nullptr;
  CK(c, launch_search_commit(*c, ca, nlist, stage, NCC_min, NCC_min_super, next_list, next_count, c->d_best,
                             nparts, flag_dev));
  c->mark_end(tstage, ev);
  // successful grains changed their dx/drot; queued ones did not, so the snapshot
  // still serves the following rescue stages
  c->snap = Ctx::SNAP_QUEUES;
  return PTG_OK;
}

// The records at the start of work done on device rotation tables, for a redo.
int save_grains(Ctx* c) {
  if (c->rot_host || c->G == 0) return PTG_OK;
  TRY(c->ensure((void**)&c->d_grains_save, &c->grains_save_cap, sizeof(ptg_grain) * c->G));
  CK(c, hipMemcpyAsync(c->d_grains_save, c->d_grains, sizeof(ptg_grain) * c->G, hipMemcpyDeviceToDevice, c->stream));
  return PTG_OK;
}

// After work on device rotation tables: each guard entry's device cosine and sine against the
// host C library's for the same angle and, where they differ, the rotated offsets of the grain's
// pattern with both (src/p-tracker.cpp:615-621). *redo when an offset differs, when the list
// overflowed, or when the test option asks for it; the list is emptied.
int check_rot_guard(Ctx* c, bool* redo) {
  *redo = false;
  if (c->rot_host) return PTG_OK;
  int32_t n = 0;
  CK(c, hipMemcpyAsync(&n, c->d_guard_count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  if (n == 0) return PTG_OK;
  std::vector<RotGuardEntry> ent(std::min(n, kRotGuardCap));
  CK(c, hipMemcpyAsync(ent.data(), c->d_guard, sizeof(RotGuardEntry) * ent.size(), hipMemcpyDeviceToHost,
                       c->stream));
  CK(c, fill_i32(*c, c->d_guard_count, 1, 0));
  CK(c, hipStreamSynchronize(c->stream));
  c->rot_checked += (int64_t)ent.size();
  if (n > kRotGuardCap) {
    *redo = true;
    return PTG_OK;
  }
  if (c->rot_guard_test == 2) *redo = true;
  for (const RotGuardEntry& e : ent) {
    const double ch = cos(e.theta), sh = sin(e.theta);
    c->rot_max_diff = std::max(c->rot_max_diff, std::max(fabs(ch - e.c), fabs(sh - e.s)));
    if (ch == e.c && sh == e.s) continue;
    for (int64_t i = c->h_pbegin[e.g]; i < c->h_pbegin[e.g + 1]; i++) {
      const double dx = (double)c->h_pattern[2 * i], dy = (double)c->h_pattern[2 * i + 1];
      const double xd = e.c * dx + e.s * dy, yd = -e.s * dx + e.c * dy;
      const double xh = ch * dx + sh * dy, yh = -sh * dx + ch * dy;
      if (round(xd) != round(xh) || round(yd) != round(yh)) {
        *redo = true;
        break;
      }
    }
  }
  return PTG_OK;
}

// Puts back the records saved by save_grains and marks the precomputations stale, before a
// redo on the host's tables.
int restore_grains(Ctx* c) {
  CK(c, hipMemcpyAsync(c->d_grains, c->d_grains_save, sizeof(ptg_grain) * c->G, hipMemcpyDeviceToDevice, c->stream));
  c->snap = Ctx::SNAP_INVALID;
  c->require_precomputations = true;
  return PTG_OK;
}

// Runs fn() on device rotation tables and checks the guard; when an entry's offsets differ,
// puts the records back and runs fn() again on the host's tables.
template <class F>
int with_rot_guard(Ctx* c, F fn) {
  if (c->rot_host || c->G == 0) return fn();
  TRY(save_grains(c));
  TRY(fn());
  bool redo = false;
  TRY(check_rot_guard(c, &redo));
  if (!redo) return PTG_OK;
  TRY(restore_grains(c));
  c->rot_host = true;
  const int rc = fn();
  c->rot_host = false;
  c->rot_redos++;
  return rc;
}

}  // namespace

extern "C" {

int ptg_abi_version(void) { return PTG_ABI_VERSION; }

// The message of the last ptg_create that failed on this thread (no context exists to
// hold it); ptg_last_error(NULL) returns it.
static thread_local std::string g_create_err;

static int create_fail(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "ptg_create: %s: %s", what, e == hipSuccess ? "invalid argument" : hipGetErrorString(e));
  g_create_err = buf;
  return code;
}

int ptg_create(void** out, int device) {
  if (!out) return create_fail(PTG_ERR_ARG, "null output pointer", hipSuccess);
  *out = nullptr;
  g_create_err.clear();
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return create_fail(PTG_ERR_HIP, "hipGetDeviceCount", e == hipSuccess ? hipErrorNoDevice : e);
  if (device < 0 || device >= n) return create_fail(PTG_ERR_ARG, "device index out of range", hipSuccess);
  if ((e = hipSetDevice(device)) != hipSuccess) return create_fail(PTG_ERR_HIP, "hipSetDevice", e);
  Ctx* c = new Ctx();
  c->device = device;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete c;
    return create_fail(PTG_ERR_HIP, "hipStreamCreateWithFlags", e);
  }
  // every failure below releases what was made before it
  auto undo = [c]() {
    if (c->d_counts) (void)hipFree(c->d_counts);
    if (c->d_evals) (void)hipFree(c->d_evals);
    if (c->d_guard_count) (void)hipFree(c->d_guard_count);
    if (c->d_frames) (void)hipFree(c->d_frames);
    (void)hipStreamDestroy(c->stream);
    delete c;
  };
  if ((e = hipMalloc((void**)&c->d_counts, sizeof(int32_t) * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&c->d_evals, sizeof(int64_t))) != hipSuccess ||
      (e = hipMalloc((void**)&c->d_guard_count, 64 + sizeof(RotGuardEntry) * kRotGuardCap)) != hipSuccess ||
      (e = hipMalloc((void**)&c->d_frames, sizeof(FrameDesc))) != hipSuccess) {
    undo();
    return create_fail(PTG_ERR_NOMEM, "hipMalloc", e);
  }
  c->d_guard = reinterpret_cast<RotGuardEntry*>(reinterpret_cast<char*>(c->d_guard_count) + 64);
  k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_counts, 4, 0);
  k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_guard_count, 1, 0);
  k_fill_i64<<<1, 64, 0, c->stream>>>((long long*)c->d_evals, 1, 0);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    undo();
    return create_fail(PTG_ERR_HIP, "first kernels on the new stream", e);
  }
  // the allocation-free entry point's staging array on this device (device current)
  if ((e = scratch_address(device, &c->d_noalloc)) != hipSuccess) {
    undo();
    return create_fail(PTG_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
  }
  *out = c;
  return PTG_OK;
}

static void free_grain_buffers(Ctx* c) {
  void** bufs[] = {(void**)&c->d_grains, (void**)&c->d_pattern, (void**)&c->d_pbegin, (void**)&c->d_p0,
                   (void**)&c->d_d0,     (void**)&c->d_pwords, (void**)&c->d_prad,    (void**)&c->d_list[0], (void**)&c->d_list[1],
                   (void**)&c->d_list[2], (void**)&c->d_cs0,    (void**)&c->d_sd0,  (void**)&c->d_list_lpt,
                   (void**)&c->d_spwords, (void**)&c->d_m0};
  for (void** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
}

static void drop_graph(Ctx* c);

int ptg_destroy(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_OK;
  (void)hipSetDevice(c->device);
  (void)c->finish_uploads();
  if (c->sp_stream) (void)hipStreamSynchronize(c->sp_stream);
  (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < kMaxImages; i++)
    if (c->img[i].owned && c->img[i].d) (void)hipFree(c->img[i].d);
  if (c->up_stream) {
    for (hipEvent_t ev : c->up_done) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(c->up_after);
    (void)hipStreamDestroy(c->up_stream);
  }
  free_grain_buffers(c);
  void* more[] = {c->d_counts,   c->d_cs,      c->d_steps_st[0], c->d_steps_st[1], c->d_steps_st[2], c->d_best,       c->d_evals,  c->d_evals_g,
                  c->d_exact_scratch[0], c->d_exact_scratch[1], c->d_exact_scratch[2], c->d_tmp, c->d_joint, c->d_tune_save, c->d_eq_pairs, c->d_eq_ref, c->d_eq_disp, c->d_eq_ncc,
                  c->d_eq_partial, c->d_eq_out, c->d_eq_res, c->d_eq_jac, c->d_eq_orig, c->d_guard_count,
                  c->d_grains_save, c->d_frames};
  for (void* b : more)
    if (b) (void)hipFree(b);
  free_exact_plans(*c);
  drop_graph(c);
  PinnedBuf* pins[] = {&c->pin_cs, &c->pin_steps, &c->pin_cs0, &c->pin_list, &c->pin_flags};
  for (PinnedBuf* b : pins)
    if (b->p) (void)hipHostFree(b->p);
  for (auto& pe : c->pending) {
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->tune_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->aux_events) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->search_joins) (void)hipEventDestroy(e);
  for (hipStream_t st : c->search_streams) (void)hipStreamDestroy(st);
  if (c->staging_done) (void)hipEventDestroy(c->staging_done);
  for (hipEvent_t e : {c->count_ev, c->sp_fork, c->sp_join})
    if (e) (void)hipEventDestroy(e);
  if (c->sp_stream) (void)hipStreamDestroy(c->sp_stream);
  if (c->aux_stream) (void)hipStreamDestroy(c->aux_stream);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return PTG_OK;
}

const char* ptg_last_error(const void* p) {
  const Ctx* c = reinterpret_cast<const Ctx*>(p);
  if (!c) return g_create_err.empty() ? "null context" : g_create_err.c_str();
  return c->err.c_str();
}

static int set_image_common(Ctx* c, int index, int32_t dimx, int32_t dimy, int64_t pitch, int32_t bits) {
  TRY(valid_index(c, index));
  if (dimx < 4 || dimy < 4) return c->fail(PTG_ERR_ARG, "image %dx%d too small", dimx, dimy);
  if (pitch < dimx) return c->fail(PTG_ERR_ARG, "pitch %lld < dimx %d", (long long)pitch, dimx);
  if (bits < 1 || bits > 16) return c->fail(PTG_ERR_ARG, "bits %d outside [1, 16]", bits);
  if (index == c->iref) c->require_precomputations = true;
  return PTG_OK;
}

// A frame copy between host and device rows: one linear copy when both sides' rows are
// contiguous (the runtime's pitched copy of pageable memory moves a row at a time).
static hipError_t copy_frame(void* dst, int64_t dpitch, const void* src, int64_t spitch, int32_t dimx, int32_t dimy,
                             hipMemcpyKind kind, hipStream_t st) {
  if (dpitch == dimx && spitch == dimx)
    return hipMemcpyAsync(dst, src, sizeof(uint16_t) * (size_t)dimx * dimy, kind, st);
  return hipMemcpy2DAsync(dst, sizeof(uint16_t) * dpitch, src, sizeof(uint16_t) * spitch, sizeof(uint16_t) * dimx,
                          dimy, kind, st);
}

// image[index] as a context-owned buffer of at least dimx x dimy pixels (pitch dimx), the
// old one released once no copy or queued work uses it.
static int image_buffer(Ctx* c, int index, int32_t dimx, int32_t dimy, int32_t bits) {
  Image& im = c->img[index];
  if (!im.owned) {
    im.d = nullptr;
    im.cap = 0;
  }
  const size_t bytes = sizeof(uint16_t) * (size_t)dimx * dimy;
  if (!im.owned || im.cap < bytes) {
    if (im.owned && im.d) {
      TRY(c->finish_upload(index));
      c->up_pending[index] = false;
      CK(c, hipStreamSynchronize(c->stream));
      (void)hipFree(im.d);
    }
    im.d = nullptr;
    if (hipMalloc((void**)&im.d, bytes) != hipSuccess) {
      im.owned = false;
      return c->fail(PTG_ERR_NOMEM, "hipMalloc(%zu) for image[%d]", bytes, index);
    }
    im.cap = bytes;
    im.owned = true;
  }
  im.dimx = dimx;
  im.dimy = dimy;
  im.pitch = dimx;
  im.bits = bits;
  return PTG_OK;
}

int ptg_set_image(void* p, int index, const uint16_t* host, int32_t dimx, int32_t dimy, int64_t pitch,
                  int32_t bits) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  CK(c, hipSetDevice(c->device));
  if (!host) return c->fail(PTG_ERR_ARG, "null host frame");
  TRY(set_image_common(c, index, dimx, dimy, pitch, bits));
  TRY(image_buffer(c, index, dimx, dimy, bits));
  TRY(c->order_after_upload(index));  // (after an earlier overlapped copy into the slot)
  Image& im = c->img[index];
  CK(c, copy_frame(im.d, dimx, host, pitch, dimx, dimy, hipMemcpyHostToDevice, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

// Whether host memory is page-locked memory of the HIP runtime (hipHostMalloc). For other
// memory the query may fail; its error is cleared so that no later launch check reports it.
static bool pinned_host(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

int ptg_set_image_async(void* p, int index, const uint16_t* host, int32_t dimx, int32_t dimy, int64_t pitch,
                        int32_t bits) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  CK(c, hipSetDevice(c->device));
  if (!host) return c->fail(PTG_ERR_ARG, "null host frame");
  TRY(set_image_common(c, index, dimx, dimy, pitch, bits));
  TRY(c->upload_setup());
  TRY(image_buffer(c, index, dimx, dimy, bits));
  Image& im = c->img[index];
  // after every call queued so far (they may read the slot's old contents)
  CK(c, hipEventRecord(c->up_after, c->stream));
  CK(c, hipStreamWaitEvent(c->up_stream, c->up_after, 0));
  CK(c, copy_frame(im.d, dimx, host, pitch, dimx, dimy, hipMemcpyHostToDevice, c->up_stream));
  CK(c, hipEventRecord(c->up_done[index], c->up_stream));
  c->up_pending[index] = true;
  // pageable memory: returns once the copy has read it, whatever the runtime does with the
  // source of an asynchronous copy from such memory
  if (!pinned_host(host)) TRY(c->finish_upload(index));
  return PTG_OK;
}

int ptg_wait_image(void* p, int index) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(valid_index(c, index));
  return c->finish_upload(index);
}

int ptg_host_alloc(void* p, size_t bytes, void** out) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!out) return c->fail(PTG_ERR_ARG, "null output pointer");
  *out = nullptr;
  CK(c, hipSetDevice(c->device));
  if (hipHostMalloc(out, bytes < 64 ? 64 : bytes, hipHostMallocDefault) != hipSuccess) {
    *out = nullptr;
    return c->fail(PTG_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
  }
  return PTG_OK;
}

int ptg_host_free(void* p, void* mem) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!mem) return PTG_OK;
  CK(c, hipHostFree(mem));
  return PTG_OK;
}

int ptg_set_image_device(void* p, int index, const uint16_t* dev, int32_t dimx, int32_t dimy, int64_t pitch,
                         int32_t bits) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!dev) return c->fail(PTG_ERR_ARG, "null device frame");
  TRY(set_image_common(c, index, dimx, dimy, pitch, bits));
  Image& im = c->img[index];
  if (im.owned && im.d) {
    TRY(c->finish_upload(index));
    CK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(im.d);
  }
  c->up_pending[index] = false;
  im.d = const_cast<uint16_t*>(dev);
  im.owned = false;
  im.cap = 0;
  im.dimx = dimx;
  im.dimy = dimy;
  im.pitch = pitch;
  im.bits = bits;
  return PTG_OK;
}

int ptg_get_image(void* p, int index, uint16_t* host, int64_t pitch) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(valid_index(c, index));
  const Image& im = c->img[index];
  if (!im.d) return c->fail(PTG_ERR_STATE, "image[%d] not set", index);
  if (!host) return c->fail(PTG_ERR_ARG, "null host buffer");
  if (pitch < im.dimx) return c->fail(PTG_ERR_ARG, "pitch %lld < dimx %d", (long long)pitch, im.dimx);
  TRY(c->order_after_upload(index));
  CK(c, copy_frame(host, pitch, im.d, im.pitch, im.dimx, im.dimy, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_set_image_indices(void* p, int im_index_ref, int im_index_current) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(valid_index(c, im_index_ref));
  TRY(valid_index(c, im_index_current));
  if (im_index_ref != c->iref) c->require_precomputations = true;
  c->iref = im_index_ref;
  c->icur = im_index_current;
  return PTG_OK;
}

// force: the device lists were reallocated or the patterns changed (ptg_set_grains)
static int upload_unmasked(Ctx* c, bool force) {
  std::vector<int32_t> um;
  um.reserve(c->G);
  for (int g = 0; g < c->G; g++)
    if (!c->h_grains[g].masked) um.push_back(g);
  // a new list (not only new records) invalidates the plans keyed on it
  const bool changed = force || um != c->h_unmasked || c->unmasked_version == 0;
  c->h_unmasked.swap(um);
  const std::vector<int32_t>& ul = c->h_unmasked;
  c->n_unmasked = (int32_t)ul.size();
  if (changed) {
    c->unmasked_version++;
    std::vector<int32_t>& lpt = c->h_list_lpt;
    lpt = ul;
    auto size_of = [c](int32_t g) { return c->h_pbegin[g + 1] - c->h_pbegin[g]; };
    std::stable_sort(lpt.begin(), lpt.end(), [&](int32_t a, int32_t b) { return size_of(a) > size_of(b); });
    c->max_prad_unmasked = 0;
    for (int32_t g : ul) c->max_prad_unmasked = std::max(c->max_prad_unmasked, c->h_prad[g]);
    if (!ul.empty()) {
      CK(c, hipMemcpyAsync(c->d_list[0], ul.data(), sizeof(int32_t) * ul.size(), hipMemcpyHostToDevice, c->stream));
      CK(c, hipMemcpyAsync(c->d_list_lpt, lpt.data(), sizeof(int32_t) * lpt.size(), hipMemcpyHostToDevice,
                           c->stream));
    }
  }
  CK(c, fill_i32(*c, c->d_counts, 1, c->n_unmasked));
  if (changed) CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_set_grains(void* p, int32_t G, const ptg_grain* grains, const int32_t* pattern_xy,
                   const int64_t* pattern_begin) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  CK(c, hipSetDevice(c->device));
  if (G < 0 || (G > 0 && (!grains || !pattern_begin))) return c->fail(PTG_ERR_ARG, "bad grain arguments");
  if (G > 0 && pattern_begin[0] != 0) return c->fail(PTG_ERR_ARG, "pattern_begin[0] must be 0");
  for (int g = 0; g < G; g++) {
    const int64_t n = pattern_begin[g + 1] - pattern_begin[g];
    if (n <= 0) return c->fail(PTG_ERR_ARG, "grain %d has an empty pattern", g);
    if (n > (1 << 20)) return c->fail(PTG_ERR_LIMIT, "grain %d pattern has %lld pixels", g, (long long)n);
  }
  const int64_t total = G ? pattern_begin[G] : 0;
  if (total > 0 && !pattern_xy) return c->fail(PTG_ERR_ARG, "null pattern_xy");
  CK(c, hipStreamSynchronize(c->stream));
  free_grain_buffers(c);
  c->G = G;
  c->grains_set = true;
  c->total_px = total;
  c->h_grains.assign(grains, grains + G);
  c->h_pattern.assign(pattern_xy, pattern_xy + 2 * total);
  c->h_pbegin.assign(pattern_begin, pattern_begin + G + 1);
  c->h_prad.resize(G);
  c->h_disk.assign(G, 0);
  c->h_diskm2.assign(G, 0);
  std::vector<uint8_t> seen;
  for (int g = 0; g < G; g++) {
    int64_t m2 = 0;
    for (int64_t i = pattern_begin[g]; i < pattern_begin[g + 1]; i++) {
      const int64_t dx = pattern_xy[2 * i], dy = pattern_xy[2 * i + 1];
      m2 = std::max(m2, dx * dx + dy * dy);
    }
    // every rotated offset round(c dx + s dy) lies within ceil(|(dx, dy)|) + 1
    c->h_prad[g] = (int32_t)ceil(sqrt((double)m2)) + 1;
    // a full digital disk: every cell with x^2 + y^2 <= m2 exactly once
    const int64_t n = pattern_begin[g + 1] - pattern_begin[g];
    int64_t r = (int64_t)sqrt((double)m2);
    while (r * r > m2) r--;
    while ((r + 1) * (r + 1) <= m2) r++;
    int64_t cells = 0;
    for (int64_t y = -r; y <= r; y++) {
      int64_t xm = (int64_t)sqrt((double)(m2 - y * y));
      while (xm * xm > m2 - y * y) xm--;
      while ((xm + 1) * (xm + 1) <= m2 - y * y) xm++;
      cells += 2 * xm + 1;
    }
    if (n > 0 && cells == n && r <= 1024) {
      c->h_diskm2[g] = (int32_t)m2;
      const int64_t side = 2 * r + 1;
      seen.assign((size_t)(side * side), 0);
      bool dup = false;
      for (int64_t i = pattern_begin[g]; i < pattern_begin[g + 1] && !dup; i++) {
        uint8_t& f = seen[(size_t)((pattern_xy[2 * i + 1] + r) * side + pattern_xy[2 * i] + r)];
        dup = f != 0;
        f = 1;
      }
      c->h_disk[g] = dup ? 0 : 1;
    }
  }
  if (G == 0) return PTG_OK;
  const size_t Gs = (size_t)G;
  if (hipMalloc((void**)&c->d_grains, sizeof(ptg_grain) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_pattern, sizeof(int32_t) * 2 * total) != hipSuccess ||
      hipMalloc((void**)&c->d_pbegin, sizeof(int64_t) * (Gs + 1)) != hipSuccess ||
      hipMalloc((void**)&c->d_p0, sizeof(int32_t) * 2 * total) != hipSuccess ||
      hipMalloc((void**)&c->d_d0, sizeof(double) * total) != hipSuccess ||
      hipMalloc((void**)&c->d_pwords, sizeof(uint32_t) * total) != hipSuccess ||
      hipMalloc((void**)&c->d_spwords, sizeof(uint32_t) * total) != hipSuccess ||
      hipMalloc((void**)&c->d_prad, sizeof(int32_t) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_list[0], sizeof(int32_t) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_list[1], sizeof(int32_t) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_list[2], sizeof(int32_t) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_list_lpt, sizeof(int32_t) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_cs0, sizeof(double) * 2 * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_sd0, sizeof(double) * Gs) != hipSuccess ||
      hipMalloc((void**)&c->d_m0, sizeof(double) * Gs) != hipSuccess) {
    free_grain_buffers(c);
    c->G = 0;
    return c->fail(PTG_ERR_NOMEM, "grain buffers for %d grains / %lld pixels", G, (long long)total);
  }
  CK(c, hipMemcpyAsync(c->d_grains, grains, sizeof(ptg_grain) * Gs, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(c->d_pattern, pattern_xy, sizeof(int32_t) * 2 * total, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(c->d_pbegin, pattern_begin, sizeof(int64_t) * (Gs + 1), hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(c->d_prad, c->h_prad.data(), sizeof(int32_t) * Gs, hipMemcpyHostToDevice, c->stream));
  CK(c, fill_i32(*c, c->d_counts, 4, 0));
  TRY(upload_unmasked(c, true));
  c->snap = Ctx::SNAP_VALID;
  c->require_precomputations = true;
  return PTG_OK;
}

int ptg_put_grains(void* p, const ptg_grain* grains) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  if (!grains) return c->fail(PTG_ERR_ARG, "null grains");
  CK(c, hipStreamSynchronize(c->stream));
  c->h_grains.assign(grains, grains + c->G);
  CK(c, hipMemcpyAsync(c->d_grains, c->h_grains.data(), sizeof(ptg_grain) * c->G, hipMemcpyHostToDevice, c->stream));
  TRY(upload_unmasked(c, false));
  c->snap = Ctx::SNAP_VALID;
  c->require_precomputations = true;
  return PTG_OK;
}

int ptg_get_grains(void* p, ptg_grain* out) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!out) return c->fail(PTG_ERR_ARG, "null output");
  if (c->G == 0) return PTG_OK;
  TRY(sync_snapshot(c));
  memcpy(out, c->h_grains.data(), sizeof(ptg_grain) * c->G);
  return PTG_OK;
}

int ptg_begin_pair(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  CK(c, launch_begin_pair(*c));
  CK(c, fill_i32(*c, c->d_counts + 1, 2, 0));
  // NCC fields and backups changed; dx/drot did not
  return PTG_OK;
}

int ptg_precompute(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  return with_rot_guard(c, [c]() -> int { return do_precompute(c, false); });
}

int ptg_search(void* p, int stage, const ptg_search_zone* zone, double NCC_min, double NCC_min_super, int method) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  return with_rot_guard(c, [&]() -> int {
    TRY(do_search(c, stage, zone, NCC_min, NCC_min_super, method));
    // an exact search reports its capacity flag before returning
    if (c->any_flags_pending()) TRY(c->sync_and_check());
    return PTG_OK;
  });
}

// Zeroes the subpixel stage's evaluation counters (the total and the per-grain counts).
__global__ void k_subpixel_reset(long long* evals, int32_t* evals_g, int G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *evals = 0;
  if (i < G) evals_g[i] = 0;
}

static int subpixel_reset(Ctx* c) {
  TRY(c->ensure((void**)&c->d_evals_g, &c->evals_g_cap, sizeof(int32_t) * c->G));
  k_subpixel_reset<<<std::max(1, (c->G + 255) / 256), 256, 0, c->stream>>>((long long*)c->d_evals, c->d_evals_g,
                                                                           c->G);
  CK(c, hipGetLastError());
  return PTG_OK;
}

static int subpixel_check(Ctx* c, const ptg_subpix_params* prm) {
  if (!prm) return c->fail(PTG_ERR_ARG, "null subpixel params");
  if (prm->method != PTG_SUBPIX_POWELL && prm->method != PTG_SUBPIX_LM)
    return c->fail(PTG_ERR_ARG, "bad subpixel method %d", prm->method);
  return PTG_OK;
}

// reset: the counters were already zeroed in this pair (track_pair does it before the
// rescue stages' host round trip, so that only the kernel launch follows it)
// sel, thr: every unmasked grain (0), or only those whose main-stage NCC reached thr (1) or the
// rest (2)
static int subpixel_run(Ctx* c, const ptg_subpix_params* prm, bool reset, int sel = 0, double thr = 0.0,
                        hipStream_t st = nullptr) {
  if (c->require_precomputations) TRY(do_precompute(c, false));
  TRY(c->set_frames());
  if (reset) TRY(subpixel_reset(c));
  hipEvent_t ev;
  c->mark_begin(PTG_T_SUBPIXEL, &ev, st ? st : c->stream);
  CK(c, launch_subpixel(*c, *prm, sel, thr, st));
  c->mark_end(PTG_T_SUBPIXEL, ev, st ? st : c->stream);
  c->snap = Ctx::SNAP_INVALID;
  return PTG_OK;
}

int ptg_subpixel(void* p, const ptg_subpix_params* prm) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(subpixel_check(c, prm));
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  TRY(need_images(c, true, true));
  return with_rot_guard(c, [&]() -> int { return subpixel_run(c, prm, true); });
}

int ptg_end_pair(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  if (c->G == 0) return PTG_OK;
  CK(c, launch_end_pair(*c));
  if (c->snap == Ctx::SNAP_VALID) {
    for (auto& g : c->h_grains) {
      g.upix = g.dx - g.dx_prev;
      g.vpix = g.dy - g.dy_prev;
      g.rot_inc = g.drot - g.drot_prev;
    }
  }
  return PTG_OK;
}

int ptg_change_reference(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  if (!c->rot_host) {
    // nothing on the host needs the records: the device tables read them where they are
    CK(c, launch_change_reference(*c));
    c->snap = Ctx::SNAP_INVALID;
    std::swap(c->iref, c->icur);
    c->require_precomputations = true;
    return PTG_OK;
  }
  if (c->snap != Ctx::SNAP_VALID) TRY(sync_snapshot(c));
  CK(c, launch_change_reference(*c));
  // the same update on the host snapshot keeps it valid (src/p-tracker.cpp:598-609)
  for (auto& r : c->h_grains) {
    double actual_pos = (double)(r.refcoord_xpix) + r.dx;
    r.refcoord_xpix = (int32_t)round(actual_pos);
    r.dx = actual_pos - r.refcoord_xpix;
    actual_pos = (double)(r.refcoord_ypix) + r.dy;
    r.refcoord_ypix = (int32_t)round(actual_pos);
    r.dy = actual_pos - r.refcoord_ypix;
    r.refrot += r.drot;
    r.drot = 0.0;
  }
  std::swap(c->iref, c->icur);
  c->require_precomputations = true;
  return PTG_OK;
}

// The precomputations, when stale, for work outside a guarded stage: checked at once.
static int ensure_precomputed(Ctx* c) {
  if (!c->require_precomputations) return PTG_OK;
  return with_rot_guard(c, [c]() -> int { return do_precompute(c, false); });
}

static int track_pair_run(Ctx* c, const ptg_track_params* prm) {
  void* p = c;
  // the ordered sums of the precomputation overlap the main stage's exact search; they are
  // launched on the context's stream before any other reader, whichever way the pair ends
  struct Flush {
    Ctx* c;
    ~Flush() { (void)flush_chain(c); }
  } flush{c};
  const bool defer = prm->rescue_level >= 0 && prm->search_method == PTG_SEARCH_EXACT && !c->rot_host;
  if (c->require_precomputations) TRY(do_precompute(c, defer));
  if (prm->rescue_level >= 0) {
    TRY(ptg_begin_pair(p));
    TRY(do_search(c, PTG_STAGE_MAIN, &prm->search_zone, prm->NCC_min, prm->NCC_min_super, prm->search_method));
  }
  if (prm->subpixel) TRY(subpixel_reset(c));
  auto rescue_stages = [&]() -> int {
    if (prm->rescue_level >= 1)
      TRY(do_search(c, PTG_STAGE_RESCUE, &prm->search_zone_rescue, prm->NCC_min, prm->NCC_min_super,
                    prm->search_method));
    if (prm->rescue_level >= 2)
      TRY(do_search(c, PTG_STAGE_SUPER, &prm->search_zone_super_rescue, prm->NCC_min, prm->NCC_min_super,
                    prm->search_method));
    return PTG_OK;
  };
  if (prm->subpixel && prm->rescue_level >= 1 && !c->rot_host && !c->device_queues && !c->graph_pair) {
    // The rescue queue is read back by the host. The subpixel stage of the grains the main stage
    // settled (NCC at least NCC_min, which no rescue stage moves) runs on a stream of its own,
    // beside the rescue stages, while the host reads the queue's count: an empty queue then
    // needs no more host round trips, else the rescue stages run and the subpixel stage of the
    // other grains follows them. Each grain's work is the same as in one subpixel stage.
    if (!c->sp_stream) {
      hipStream_t st = nullptr;
      CK(c, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      c->sp_stream = st;
    }
    for (hipEvent_t* e : {&c->count_ev, &c->sp_fork, &c->sp_join})
      if (!*e) CK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    TRY(c->pin_flags.ensure(*c, kPinFlags * sizeof(int32_t)));
    int32_t* pc = static_cast<int32_t*>(c->pin_flags.p) + 9;
    CK(c, hipMemcpyAsync(pc, c->d_counts + PTG_STAGE_RESCUE, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipEventRecord(c->count_ev, c->stream));
    CK(c, hipEventRecord(c->sp_fork, c->stream));
    CK(c, hipStreamWaitEvent(c->sp_stream, c->sp_fork, 0));
    // the library stream waits for that stage at the end of this block, however it ends
    struct Join {
      Ctx* c;
      ~Join() {
        (void)hipEventRecord(c->sp_join, c->sp_stream);
        (void)hipStreamWaitEvent(c->stream, c->sp_join, 0);
      }
    } join{c};
    TRY(subpixel_run(c, &prm->subpix, false, 1, prm->NCC_min, c->sp_stream));
    CK(c, hipEventSynchronize(c->count_ev));
    if (*pc > 0) {
      TRY(rescue_stages());
      TRY(subpixel_run(c, &prm->subpix, false, 2, prm->NCC_min));
    } else {
      // nothing queued, so nothing reaches the super-rescue either: its queue count is zero
      CK(c, fill_i32(*c, c->d_counts + PTG_STAGE_SUPER, 1, 0));
    }
  } else {
    TRY(rescue_stages());
    if (prm->subpixel) TRY(subpixel_run(c, &prm->subpix, false));
  }
  TRY(ptg_end_pair(p));
  // a stage that could not be served exactly is reported here, after the pair's work
  if (!c->capturing && c->any_flags_pending()) TRY(c->sync_and_check());
  return PTG_OK;
}

// The key of a captured pair: everything its launches and their arguments depend on besides the
// frames (read through the frame record) and the records' values.
static uint64_t pair_key(const Ctx* c, const ptg_track_params* p) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  auto mixd = [&mix](double d) {
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    mix(u);
  };
  for (const ptg_search_zone* z : {&p->search_zone, &p->search_zone_rescue, &p->search_zone_super_rescue}) {
    mix((uint64_t)(uint32_t)z->left | (uint64_t)(uint32_t)z->right << 32);
    mix((uint64_t)(uint32_t)z->up | (uint64_t)(uint32_t)z->down << 32);
    mix((uint64_t)(uint32_t)z->num_rot);
    mixd(z->inc_rot);
  }
  mixd(p->NCC_min);
  mixd(p->NCC_min_super);
  mix((uint64_t)(uint32_t)p->rescue_level | (uint64_t)(uint32_t)p->subpixel << 32);
  mix((uint64_t)(uint32_t)p->search_method);
  if (p->subpixel) {
    const ptg_subpix_params& q = p->subpix;
    mixd(q.subpix_tol);
    for (double d : q.initial_direction) mixd(d);
    mixd(q.subpix_displacement_max);
    mixd(q.overflow_stiffness);
    mix((uint64_t)(uint32_t)q.method | (uint64_t)(uint32_t)q.max_iter << 32);
  }
  const Image &i0 = c->img[c->iref], &i1 = c->img[c->icur];
  mix((uint64_t)std::max(i0.bits, i1.bits));
  mix(c->unmasked_version);
  mix(c->tune_gen);
  mix(c->table_uploads);  // (device tables another call replaced)
  // the buffers the launches address (each grows, and moves, with the work it serves)
  for (const void* q : {(const void*)c->d_cs, (const void*)c->d_best, (const void*)c->d_steps_st[0],
                        (const void*)c->d_steps_st[1], (const void*)c->d_steps_st[2],
                        (const void*)c->d_exact_scratch[0], (const void*)c->d_exact_scratch[1],
                        (const void*)c->d_exact_scratch[2], (const void*)c->d_evals_g, (const void*)c->pin_flags.p})
    mix((uint64_t)(uintptr_t)q);
  mix((uint64_t)c->G);
  mix((uint64_t)c->require_precomputations | (uint64_t)c->serial_search << 1);
  return h;
}

static void drop_graph(Ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key = 0;
}

// A pair as one HIP graph: the first pair with a new key runs as launches (building every plan,
// table and buffer the pair needs), the second is captured from the same launch sequence and
// replayed, and so is every later pair with that key, its frames given by the frame record.
// Rescue queues stay on the device; the rotation guard and the exact search's flags are checked
// after each replay as after a pair of launches.
static int track_pair_graph(Ctx* c, const ptg_track_params* prm) {
  const uint64_t key = pair_key(c, prm);
  c->graph_pair = true;
  struct Reset {
    Ctx* c;
    ~Reset() { c->graph_pair = false; }
  } reset{c};
  if (key == c->graph_bad_key || (key != c->graph_key && key != c->graph_warm_key)) {
    // a new key: a pair of launches first
    if (key != c->graph_bad_key) c->graph_warm_key = key;
    return with_rot_guard(c, [&]() -> int { return track_pair_run(c, prm); });
  }
  TRY(c->set_frames());
  TRY(save_grains(c));
  if (key != c->graph_key) {
    // capture the pair's launches
    drop_graph(c);
    const bool pre = c->require_precomputations;
    for (int st = 0; st < 3; st++) c->flags_pending[st] = c->unfit_pending[st] = false;
    hipError_t he = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (he != hipSuccess) return c->check(he, "hipStreamBeginCapture");
    c->capturing = true;
    const int rc = track_pair_run(c, prm);
    c->capturing = false;
    hipGraph_t g = nullptr;
    he = hipStreamEndCapture(c->stream, &g);
    hipGraphExec_t ge = nullptr;
    if (rc == PTG_OK && he == hipSuccess && g) he = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (rc != PTG_OK || he != hipSuccess || !ge) {
      // this key stays on launches
      if (ge) (void)hipGraphExecDestroy(ge);
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      c->graph_bad_key = key;
      c->require_precomputations = pre;
      for (int st = 0; st < 3; st++) c->flags_pending[st] = c->unfit_pending[st] = false;
      return with_rot_guard(c, [&]() -> int { return track_pair_run(c, prm); });
    }
    c->graph = g;
    c->graph_exec = ge;
    c->graph_key = key;
    for (int st = 0; st < 3; st++) {
      c->graph_flags[st] = c->flags_pending[st];
      c->graph_unfit[st] = c->unfit_pending[st];
    }
    c->graph_captures++;
  }
  CK(c, hipGraphLaunch(c->graph_exec, c->stream));
  c->graph_launches++;
  c->require_precomputations = false;
  c->snap = Ctx::SNAP_INVALID;
  for (int st = 0; st < 3; st++) {
    c->flags_pending[st] = c->graph_flags[st];
    c->unfit_pending[st] = c->graph_unfit[st];
  }
  if (c->any_flags_pending()) TRY(c->sync_and_check());
  bool redo = false;
  TRY(check_rot_guard(c, &redo));
  if (!redo) return PTG_OK;
  TRY(restore_grains(c));
  c->rot_host = true;
  const int rc = track_pair_run(c, prm);
  c->rot_host = false;
  c->rot_redos++;
  return rc;
}

int ptg_track_pair(void* p, const ptg_track_params* prm) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!prm) return c->fail(PTG_ERR_ARG, "null track params");
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  if (prm->subpixel) TRY(subpixel_check(c, &prm->subpix));
  if (c->G == 0) return PTG_OK;
  if (c->use_graph && !c->rot_host && !c->timing && c->rot_guard_test == 0) return track_pair_graph(c, prm);
  return with_rot_guard(c, [&]() -> int { return track_pair_run(c, prm); });
}

// The exact search's autotuner: for each radius bucket of the main stage, every launch shape
// of a candidate set (N-tiles per pass 1-3, waves per workgroup 1-8, workgroups per grain 1, 2
// or 4) is tried on the current frames and grains, and the whole search's GPU span (events on
// the library stream, the radius buckets overlapped over their streams as in any search) is
// timed; a shape replaces the heuristic's only when the median of three spans is shorter. The
// grains and the fallback counters are left as they were.
int ptg_autotune_search(void* p, const ptg_search_zone* z, int32_t* ntried, double* ms_before, double* ms_after) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!z) return c->fail(PTG_ERR_ARG, "null search zone");
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  int32_t tried = 0;
  double before = 0.0, after = 0.0;
  if (c->G > 0) {
    TRY(ensure_precomputed(c));
    const size_t gb = sizeof(ptg_grain) * (size_t)c->G;
    TRY(c->ensure(&c->d_tune_save, &c->tune_save_cap, gb));
    CK(c, hipMemcpyAsync(c->d_tune_save, c->d_grains, gb, hipMemcpyDeviceToDevice, c->stream));
    const int64_t fb0 = c->exact_fallbacks, fg0 = c->exact_grain_fallbacks;
    c->tune_run = true;
    // one main-stage search from the saved grains: its GPU span in ms
    auto run = [&](float* ms) -> int {
      CK(c, hipMemcpyAsync(c->d_grains, c->d_tune_save, gb, hipMemcpyDeviceToDevice, c->stream));
      c->snap = Ctx::SNAP_INVALID;
      TRY(ptg_begin_pair(p));
      c->tune_keys.clear();
      TRY(do_search(c, PTG_STAGE_MAIN, z, 0.7, 0.5, PTG_SEARCH_EXACT));
      TRY(c->sync_and_check());
      tried++;
      *ms = 0.0f;
      if (!c->tune_keys.empty()) CK(c, hipEventElapsedTime(ms, c->tune_ev[0], c->tune_ev[1]));
      return PTG_OK;
    };
    auto median3 = [&](float* ms) -> int {
      float t[3];
      for (int r = 0; r < 3; r++) TRY(run(&t[r]));
      std::sort(t, t + 3);
      *ms = t[1];
      return PTG_OK;
    };
    auto body = [&]() -> int {
      c->tune_bucket = -1;
      float ms = 0.0f;
      TRY(run(&ms));  // the buckets and their keys
      if (c->tune_keys.empty()) return PTG_OK;  // the search ran the plain form: nothing to tune
      for (uint64_t k : c->tune_keys) c->exact_tuned.erase(k);
      const std::vector<uint64_t> keys = c->tune_keys;
      float best_ms = 0.0f;
      TRY(median3(&best_ms));  // the heuristic's shapes
      before = best_ms;
      // the radius buckets' launch order (PTG_OPT_BUCKET_ORDER): the faster of the two
      {
        const int o0 = c->bucket_order;
        c->bucket_order = 1 - o0;
        float ms_other = 0.0f;
        TRY(median3(&ms_other));
        if (ms_other < best_ms) best_ms = ms_other;
        else c->bucket_order = o0;
      }
      for (size_t bi = 0; bi < keys.size(); bi++) {
        if (keys[bi] == 0) continue;  // an empty bucket
        c->tune_bucket = -1;
        TRY(run(&ms));
        const ExactShape base_shape = c->tune_used[bi];
        ExactShape best = base_shape;
        for (int roles = 1; roles <= 2; roles++)
          for (int ntp = 1; ntp <= 3; ntp++)
            for (int nsub : {1, 2, 4})
              for (int nsu = 1; nsu <= (roles == 2 ? 4 : 8); nsu++) {
              const ExactShape want{ntp, nsu, nsub, roles};
              if (want.ntp == base_shape.ntp && want.nsu == base_shape.nsu && want.nsub == base_shape.nsub &&
                  want.roles == base_shape.roles)
                continue;
              c->tune_bucket = (int)bi;
              c->tune_shape = want;
              TRY(run(&ms));
              const ExactShape& got = c->tune_used[bi];
              if (got.ntp != want.ntp || got.nsu != want.nsu || got.nsub != want.nsub || got.roles != want.roles)
                continue;  // did not fit
              if (ms < best_ms) {
                TRY(median3(&ms));
                if (ms < best_ms) {
                  best_ms = ms;
                  best = want;
                }
              }
            }
        // later buckets are tried with this one's choice in place
        c->exact_tuned[keys[bi]] = best;
      }
      c->tune_bucket = -1;
      TRY(median3(&best_ms));
      after = best_ms;
      return PTG_OK;
    };
    const int rc = body();
    c->tune_run = false;
    c->tune_gen++;
    c->tune_bucket = -1;
    c->exact_fallbacks = fb0;
    c->exact_grain_fallbacks = fg0;
    CK(c, hipMemcpyAsync(c->d_grains, c->d_tune_save, gb, hipMemcpyDeviceToDevice, c->stream));
    c->snap = Ctx::SNAP_INVALID;
    // the tuning searches' guard entries concern records that were put back
    CK(c, fill_i32(*c, c->d_guard_count, 1, 0));
    CK(c, hipStreamSynchronize(c->stream));
    TRY(rc);
  }
  if (ntried) *ntried = tried;
  if (ms_before) *ms_before = before;
  if (ms_after) *ms_after = after;
  return PTG_OK;
}

int ptg_clear_tuning(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  c->exact_tuned.clear();
  c->tune_gen++;
  return PTG_OK;
}

int ptg_get_queue_counts(void* p, int32_t* nres, int32_t* nsup) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  int32_t cnt[4] = {0, 0, 0, 0};
  CK(c, hipMemcpyAsync(cnt, c->d_counts, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  TRY(c->sync_and_check());
  if (nres) *nres = cnt[1];
  if (nsup) *nsup = cnt[2];
  return PTG_OK;
}

int ptg_sync(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(c->finish_uploads());
  return c->sync_and_check();
}

static int tmp_upload(Ctx* c, size_t bytes) { return c->ensure(&c->d_tmp, &c->tmp_cap, bytes); }

static int objective_common(Ctx* c, int32_t n, const int32_t* igrain, const double* X, double dmax, double stiff,
                            double* f, double* grad) {
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  if (n < 0 || (n > 0 && (!igrain || !X || !f))) return c->fail(PTG_ERR_ARG, "bad objective arguments");
  for (int k = 0; k < n; k++)
    if (igrain[k] < 0 || igrain[k] >= c->G) return c->fail(PTG_ERR_ARG, "grain index %d out of range", igrain[k]);
  if (n == 0) return PTG_OK;
  TRY(ensure_precomputed(c));
  const size_t bi = ((size_t)n * 4 + 255) / 256 * 256, bx = (size_t)n * 24, bf = (size_t)n * 8, bg = (size_t)n * 24;
  TRY(tmp_upload(c, bi + bx + bf + bg));
  char* base = (char*)c->d_tmp;
  int32_t* d_ig = (int32_t*)base;
  double* d_X = (double*)(base + bi);
  double* d_f = (double*)(base + bi + bx);
  double* d_g = (double*)(base + bi + bx + bf);
  CK(c, hipMemcpyAsync(d_ig, igrain, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(d_X, X, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_objective_batch(*c, n, d_ig, d_X, dmax, stiff, d_f, grad ? d_g : nullptr));
  CK(c, hipMemcpyAsync(f, d_f, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (grad) CK(c, hipMemcpyAsync(grad, d_g, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_ncc_objective(void* p, int32_t igrain, const double* X, double dmax, double stiff, double* f) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  return objective_common(c, 1, &igrain, X, dmax, stiff, f, nullptr);
}

int ptg_ncc_objective_batch(void* p, int32_t n, const int32_t* igrain, const double* X, double dmax, double stiff,
                            double* f) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  return objective_common(c, n, igrain, X, dmax, stiff, f, nullptr);
}

int ptg_ncc_objective_batch_noalloc(void* p, int32_t n, const int32_t* igrain, const double* X, double dmax,
                                    double stiff, double* f) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  if (n < 0 || (n > 0 && (!igrain || !X || !f))) return c->fail(PTG_ERR_ARG, "bad objective arguments");
  if (n > kNoAllocMaxEvals)
    return c->fail(PTG_ERR_LIMIT, "%d evaluations in one allocation-free call (at most %d)", n, kNoAllocMaxEvals);
  for (int k = 0; k < n; k++)
    if (igrain[k] < 0 || igrain[k] >= c->G) return c->fail(PTG_ERR_ARG, "grain index %d out of range", igrain[k]);
  if (n == 0) return PTG_OK;
  // the reference moments must exist already: computing them stages host tables
  if (c->require_precomputations)
    return c->fail(PTG_ERR_STATE, "reference moments not computed for the current frames (call ptg_precompute)");
  if (!c->d_noalloc) return c->fail(PTG_ERR_STATE, "no staging array on this device");
  char* base = (char*)c->d_noalloc;
  int32_t* d_ig = (int32_t*)base;
  double* d_X = (double*)(base + kNoAllocIgBytes);
  double* d_f = (double*)(base + kNoAllocIgBytes + kNoAllocXBytes);
  CK(c, hipMemcpyAsync(d_ig, igrain, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(d_X, X, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_objective_batch(*c, n, d_ig, d_X, dmax, stiff, d_f, nullptr));
  CK(c, hipMemcpyAsync(f, d_f, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_ncc_objective_grad_batch(void* p, int32_t n, const int32_t* igrain, const double* X, double* f,
                                 double* grad) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!grad) return c->fail(PTG_ERR_ARG, "null gradient output");
  return objective_common(c, n, igrain, X, 3.0, 0.0, f, grad);
}

int ptg_interp_cubic(void* p, int index, int32_t n, const double* xy, double* out) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(valid_index(c, index));
  if (!c->img[index].d) return c->fail(PTG_ERR_STATE, "image[%d] not set", index);
  if (n < 0 || (n > 0 && (!xy || !out))) return c->fail(PTG_ERR_ARG, "bad sampler arguments");
  if (n == 0) return PTG_OK;
  const size_t bx = ((size_t)n * 16 + 255) / 256 * 256;
  TRY(tmp_upload(c, bx + (size_t)n * 8));
  double* d_xy = (double*)c->d_tmp;
  double* d_o = (double*)((char*)c->d_tmp + bx);
  CK(c, hipMemcpyAsync(d_xy, xy, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_interp_batch(*c, index, n, d_xy, d_o));
  CK(c, hipMemcpyAsync(out, d_o, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_undistort_points(void* p, const double* params14, int32_t n, const double* xy_d, double* xy_u) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!params14 || n < 0 || (n > 0 && (!xy_d || !xy_u))) return c->fail(PTG_ERR_ARG, "bad undistort arguments");
  if (n == 0) return PTG_OK;
  const size_t bx = ((size_t)n * 16 + 255) / 256 * 256;
  TRY(tmp_upload(c, 2 * bx));
  double* d_in = (double*)c->d_tmp;
  double* d_out = (double*)((char*)c->d_tmp + bx);
  CK(c, hipMemcpyAsync(d_in, xy_d, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_undistort_points(*c, params14, n, d_in, d_out));
  CK(c, hipMemcpyAsync(xy_u, d_out, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

// ---- distortion calibration ----------------------------------------------------------
int ptg_equiproj_set_data(void* p, int32_t G, const double* ref_xy, int32_t nimg, const double* disp,
                          const double* ncc, int32_t npairs, const int32_t* pairs, double NCC_min) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (G < 0 || nimg < 0 || npairs < 0 || (G > 0 && !ref_xy) || (G > 0 && nimg > 0 && (!disp || !ncc)) ||
      (npairs > 0 && !pairs))
    return c->fail(PTG_ERR_ARG, "bad equiprojectivity data");
  if ((int64_t)nimg * npairs > ((int64_t)1 << 31)) return c->fail(PTG_ERR_LIMIT, "too many residuals");
  for (int32_t q = 0; q < 2 * npairs; q++)
    if (pairs[q] < 0 || pairs[q] >= G) return c->fail(PTG_ERR_ARG, "pair %d names grain %d of %d", q / 2, pairs[q], G);
  CK(c, hipStreamSynchronize(c->stream));
  TRY(c->ensure((void**)&c->d_eq_pairs, &c->eq_pairs_cap, sizeof(int32_t) * 2 * (size_t)std::max(npairs, 1)));
  TRY(c->ensure((void**)&c->d_eq_ref, &c->eq_ref_cap, sizeof(double) * 2 * (size_t)std::max(G, 1)));
  TRY(c->ensure((void**)&c->d_eq_disp, &c->eq_disp_cap, sizeof(double) * 2 * (size_t)std::max<int64_t>((int64_t)nimg * G, 1)));
  TRY(c->ensure((void**)&c->d_eq_ncc, &c->eq_ncc_cap, sizeof(double) * (size_t)std::max<int64_t>((int64_t)nimg * G, 1)));
  const int64_t m = (int64_t)nimg * npairs;
  TRY(c->ensure((void**)&c->d_eq_partial, &c->eq_partial_cap, sizeof(double) * 120 * (size_t)equiproj_partial_blocks(m)));
  TRY(c->ensure((void**)&c->d_eq_out, &c->eq_out_cap, sizeof(double) * 120));
  if (npairs) CK(c, hipMemcpyAsync(c->d_eq_pairs, pairs, sizeof(int32_t) * 2 * npairs, hipMemcpyHostToDevice, c->stream));
  if (G) CK(c, hipMemcpyAsync(c->d_eq_ref, ref_xy, sizeof(double) * 2 * G, hipMemcpyHostToDevice, c->stream));
  if (G && nimg) {
    CK(c, hipMemcpyAsync(c->d_eq_disp, disp, sizeof(double) * 2 * (size_t)nimg * G, hipMemcpyHostToDevice, c->stream));
    CK(c, hipMemcpyAsync(c->d_eq_ncc, ncc, sizeof(double) * (size_t)nimg * G, hipMemcpyHostToDevice, c->stream));
  }
  CK(c, hipStreamSynchronize(c->stream));
  c->eq_G = G;
  c->eq_nimg = nimg;
  c->eq_npairs = npairs;
  c->eq_NCC_min = NCC_min;
  c->eq_orig_set = false;  // new data: the first positions are undistorted again
  return PTG_OK;
}

int ptg_equiproj_set_origins(void* p, const double* orig_xy) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!c->d_eq_out) return c->fail(PTG_ERR_STATE, "ptg_equiproj_set_data not called");
  if (!orig_xy) {
    c->eq_orig_set = false;
    return PTG_OK;
  }
  if (c->eq_G > 0) {
    TRY(c->ensure((void**)&c->d_eq_orig, &c->eq_orig_cap, sizeof(double) * 2 * (size_t)c->eq_G));
    CK(c, hipMemcpyAsync(c->d_eq_orig, orig_xy, sizeof(double) * 2 * (size_t)c->eq_G, hipMemcpyHostToDevice,
                         c->stream));
    CK(c, hipStreamSynchronize(c->stream));
  }
  c->eq_orig_set = true;
  return PTG_OK;
}

// f and, when wanted, J^T r and J^T J (full 14 x 14) at X.
static int equiproj_eval(Ctx* c, const double* X, double* f, double* jtr, double* jtj, double* res, double* jac) {
  const int64_t m = (int64_t)c->eq_nimg * c->eq_npairs;
  EquiprojArgs a;
  a.G = c->eq_G;
  a.nimg = c->eq_nimg;
  a.npairs = c->eq_npairs;
  a.pad = 0;
  a.pairs = c->d_eq_pairs;
  a.ref = c->d_eq_ref;
  a.orig = c->eq_orig_set && c->eq_G > 0 ? c->d_eq_orig : nullptr;
  a.disp = c->d_eq_disp;
  a.ncc = c->d_eq_ncc;
  a.NCC_min = c->eq_NCC_min;
  a.partial = c->d_eq_partial;
  a.res = a.jac = nullptr;
  if (res && m) {
    TRY(c->ensure((void**)&c->d_eq_res, &c->eq_res_cap, sizeof(double) * (size_t)m));
    a.res = c->d_eq_res;
  }
  if (jac && m) {
    TRY(c->ensure((void**)&c->d_eq_jac, &c->eq_jac_cap, sizeof(double) * 14 * (size_t)m));
    a.jac = c->d_eq_jac;
  }
  CK(c, launch_equiproj(*c, a, X, c->d_eq_out));
  double out[120];
  CK(c, hipMemcpyAsync(out, c->d_eq_out, sizeof(out), hipMemcpyDeviceToHost, c->stream));
  if (a.res) CK(c, hipMemcpyAsync(res, a.res, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, c->stream));
  if (a.jac) CK(c, hipMemcpyAsync(jac, a.jac, sizeof(double) * 14 * (size_t)m, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  if (f) *f = out[0];
  if (jtr)
    for (int q = 0; q < 14; q++) jtr[q] = out[1 + q];
  if (jtj) {
    int q = 15;
    for (int a1 = 0; a1 < 14; a1++)
      for (int a2 = a1; a2 < 14; a2++, q++) jtj[a1 * 14 + a2] = jtj[a2 * 14 + a1] = out[q];
  }
  return PTG_OK;
}

int ptg_equiproj_eval(void* p, const double* params14, double* f, double* jtr, double* jtj, double* res,
                      double* jac) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!params14) return c->fail(PTG_ERR_ARG, "null parameters");
  if (!c->d_eq_out) return c->fail(PTG_ERR_STATE, "ptg_equiproj_set_data not called");
  return equiproj_eval(c, params14, f, jtr, jtj, res, jac);
}

// Solves A x = b for symmetric positive definite A (n x n, row major) by Cholesky; false
// when A is not positive definite.
static bool cholesky_solve(int n, std::vector<double> A, std::vector<double>& x, const std::vector<double>& b) {
  for (int j = 0; j < n; j++) {
    double d = A[j * n + j];
    for (int k = 0; k < j; k++) d -= A[j * n + k] * A[j * n + k];
    if (!(d > 0.0)) return false;
    A[j * n + j] = sqrt(d);
    for (int i = j + 1; i < n; i++) {
      double v = A[i * n + j];
      for (int k = 0; k < j; k++) v -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = v / A[j * n + j];
    }
  }
  x = b;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < i; k++) x[i] -= A[i * n + k] * x[k];
    x[i] /= A[i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    for (int k = i + 1; k < n; k++) x[i] -= A[k * n + i] * x[k];
    x[i] /= A[i * n + i];
  }
  return true;
}

int ptg_calibrate_distortion(void* p, double* params14, const int32_t* free14, int32_t max_iter, double* f_ini,
                             double* f_end, int32_t* iters) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!params14 || !free14 || max_iter < 0) return c->fail(PTG_ERR_ARG, "bad calibration arguments");
  if (!c->d_eq_out) return c->fail(PTG_ERR_STATE, "ptg_equiproj_set_data not called");
  std::vector<int> F;
  for (int q = 0; q < 14; q++)
    if (free14[q]) F.push_back(q);
  const int n = (int)F.size();
  double X[14], f = 0.0, jtr[14], jtj[196];
  for (int q = 0; q < 14; q++) X[q] = params14[q];
  TRY(equiproj_eval(c, X, &f, jtr, jtj, nullptr, nullptr));
  if (f_ini) *f_ini = f;
  // Levenberg-Marquardt on the free parameters, scaled by the square roots of the
  // diagonal of J^T J (K1 ~ 1e-9 and K2 ~ 1e-16 px^-2k sit beside pixel centres)
  double lambda = 1e-3;
  int it = 0;
  for (; it < max_iter && n > 0; it++) {
    std::vector<double> D(n), A(n * n), b(n), y;
    for (int a = 0; a < n; a++) {
      const double d = jtj[F[a] * 14 + F[a]];
      D[a] = d > 0.0 ? sqrt(d) : 1.0;
    }
    for (int a = 0; a < n; a++) {
      b[a] = -jtr[F[a]] / D[a];
      for (int a2 = 0; a2 < n; a2++) A[a * n + a2] = jtj[F[a] * 14 + F[a2]] / (D[a] * D[a2]);
    }
    bool improved = false, converged = false;
    while (lambda < 1e16) {
      std::vector<double> M = A;
      for (int a = 0; a < n; a++) M[a * n + a] += lambda * (A[a * n + a] > 0.0 ? A[a * n + a] : 1.0);
      if (!cholesky_solve(n, M, y, b)) {
        lambda *= 10.0;
        continue;
      }
      double Xn[14];
      for (int q = 0; q < 14; q++) Xn[q] = X[q];
      double step = 0.0;
      for (int a = 0; a < n; a++) {
        const double dq = y[a] / D[a];
        Xn[F[a]] = X[F[a]] + dq;
        step = std::max(step, fabs(dq) / (fabs(X[F[a]]) + 1e-300));
      }
      double fn = 0.0, jtrn[14], jtjn[196];
      TRY(equiproj_eval(c, Xn, &fn, jtrn, jtjn, nullptr, nullptr));
      if (fn < f) {
        converged = (f - fn) <= 1e-15 * f || step < 1e-15;
        for (int q = 0; q < 14; q++) X[q] = Xn[q];
        for (int q = 0; q < 14; q++) jtr[q] = jtrn[q];
        for (int q = 0; q < 196; q++) jtj[q] = jtjn[q];
        f = fn;
        lambda = std::max(lambda * 0.1, 1e-12);
        improved = true;
        break;
      }
      if (step < 1e-15) {
        converged = true;
        break;
      }
      lambda *= 10.0;
    }
    if (!improved || converged) {
      it++;
      break;
    }
  }
  for (int q = 0; q < 14; q++) params14[q] = X[q];
  if (f_end) *f_end = f;
  if (iters) *iters = it;
  return PTG_OK;
}

// ---- joint refinement of the grains' motion and a radial lens --------------------------
// Device layout of one call: the grain list (n ints, 256-aligned), the parameters (3n
// doubles), the per-grain outputs (joint_out_stride() n doubles).
static int joint_buffers(Ctx* c, int32_t n, int32_t** d_list, double** d_P, double** d_out) {
  const size_t bl = ((size_t)n * 4 + 255) / 256 * 256, bp = ((size_t)n * 24 + 255) / 256 * 256;
  const size_t bo = (size_t)n * 8 * joint_out_stride();
  TRY(c->ensure(&c->d_joint, &c->joint_cap, bl + bp + bo));
  *d_list = (int32_t*)c->d_joint;
  *d_P = (double*)((char*)c->d_joint + bl);
  *d_out = (double*)((char*)c->d_joint + bl + bp);
  return PTG_OK;
}

int ptg_joint_eval(void* p, int32_t n, const int32_t* igrain, const double* P, const double* lens, double* f,
                   double* grad, double* jtj) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  if (n < 0 || (n > 0 && (!igrain || !P || !lens || !f))) return c->fail(PTG_ERR_ARG, "bad joint arguments");
  for (int k = 0; k < n; k++)
    if (igrain[k] < 0 || igrain[k] >= c->G) return c->fail(PTG_ERR_ARG, "grain index %d out of range", igrain[k]);
  if (n == 0) return PTG_OK;
  TRY(ensure_precomputed(c));
  int32_t* d_list;
  double *d_P, *d_out;
  TRY(joint_buffers(c, n, &d_list, &d_P, &d_out));
  CK(c, hipMemcpyAsync(d_list, igrain, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(d_P, P, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_joint(*c, d_list, n, d_P, lens, grad || jtj, d_out));
  const int st = joint_out_stride();
  std::vector<double> out((size_t)n * st);
  CK(c, hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < n; k++) {
    const double* o = &out[(size_t)k * st];
    f[k] = o[0];
    if (grad)
      for (int q = 0; q < 5; q++) grad[5 * k + q] = o[2 + q];
    if (jtj)
      for (int q = 0; q < 15; q++) jtj[15 * k + q] = o[7 + q];
  }
  return PTG_OK;
}

namespace {
// A symmetric positive definite 3 x 3 system A X = B (B with m columns) by Cholesky; false
// when A is not positive definite.
bool chol3_solve(const double A[9], double* B, int m) {
  double L[9] = {0};
  for (int j = 0; j < 3; j++) {
    double d = A[3 * j + j];
    for (int k = 0; k < j; k++) d -= L[3 * j + k] * L[3 * j + k];
    if (!(d > 0.0)) return false;
    L[3 * j + j] = sqrt(d);
    for (int i = j + 1; i < 3; i++) {
      double v = A[3 * i + j];
      for (int k = 0; k < j; k++) v -= L[3 * i + k] * L[3 * j + k];
      L[3 * i + j] = v / L[3 * j + j];
    }
  }
  for (int col = 0; col < m; col++) {
    double y[3];
    for (int i = 0; i < 3; i++) {
      double v = B[3 * col + i];
      for (int k = 0; k < i; k++) v -= L[3 * i + k] * y[k];
      y[i] = v / L[3 * i + i];
    }
    for (int i = 2; i >= 0; i--) {
      double v = y[i];
      for (int k = i + 1; k < 3; k++) v -= L[3 * k + i] * B[3 * col + k];
      B[3 * col + i] = v / L[3 * i + i];
    }
  }
  return true;
}
}  // namespace

int ptg_refine_joint_distortion(void* p, double* lens, double ncc_min, int32_t max_iter, double* f_ini,
                                double* f_end, int32_t* iters, int32_t* ngrains) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!lens || max_iter < 0) return c->fail(PTG_ERR_ARG, "bad joint refinement arguments");
  TRY(need_grains(c));
  TRY(need_images(c, true, true));
  TRY(ensure_precomputed(c));
  if (c->snap != Ctx::SNAP_VALID) TRY(sync_snapshot(c));
  // the grains that take part: unmasked, with a subpixel NCC of at least ncc_min
  std::vector<int32_t> list;
  for (int g = 0; g < c->G; g++)
    if (!c->h_grains[g].masked && c->h_grains[g].NCC_subpix >= ncc_min) list.push_back(g);
  const int n = (int)list.size();
  if (ngrains) *ngrains = n;
  if (n == 0) {
    if (f_ini) *f_ini = 0.0;
    if (f_end) *f_end = 0.0;
    if (iters) *iters = 0;
    return PTG_OK;
  }
  int32_t* d_list;
  double *d_P, *d_out;
  TRY(joint_buffers(c, n, &d_list, &d_P, &d_out));
  CK(c, hipMemcpyAsync(d_list, list.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  std::vector<double> P(3 * (size_t)n), Pt(3 * (size_t)n);
  for (int k = 0; k < n; k++) {
    const ptg_grain& r = c->h_grains[list[k]];
    P[3 * k] = r.dx;
    P[3 * k + 1] = r.dy;
    P[3 * k + 2] = r.drot;
  }
  // the lens in relative units: kappa1 = k1 R^2, kappa2 = k2 R^4, R the distance from the
  // centre to the farthest frame corner (|kappa| is then the lens's relative displacement there)
  const Image& im = c->img[c->icur];
  const double xc = lens[2], yc = lens[3];
  const double R2 = std::max(std::max(xc * xc, (im.dimx - 1 - xc) * (im.dimx - 1 - xc)), 0.0) +
                    std::max(yc * yc, (im.dimy - 1 - yc) * (im.dimy - 1 - yc));
  const double s1 = R2, s2 = R2 * R2;
  double kap[2] = {lens[0] * s1, lens[1] * s2};
  const int st = joint_out_stride();
  std::vector<double> out((size_t)n * st);
  auto eval = [&](const std::vector<double>& Pv, const double* kp, bool derivs, double* F) -> int {
    const double L[4] = {kp[0] / s1, kp[1] / s2, xc, yc};
    CK(c, hipMemcpyAsync(d_P, Pv.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    CK(c, launch_joint(*c, d_list, n, d_P, L, derivs, d_out));
    CK(c, hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
    double t = 0.0;
    for (int k = 0; k < n; k++) t += out[(size_t)k * st];
    *F = t;
    return PTG_OK;
  };
  double F = 0.0;
  TRY(eval(P, kap, true, &F));
  if (f_ini) *f_ini = F;
  if (!std::isfinite(F)) return c->fail(PTG_ERR_ARG, "the joint objective is not finite at the start");
  std::vector<double> H = out;  // derivatives at the current point
  double lambda = 1e-3;
  int it = 0;
  std::vector<double> Y(9 * (size_t)n);  // per grain: A^-1 [B | b] (3 x 3 column-major)
  for (; it < max_iter; it++) {
    bool accepted = false, done = false;
    while (!accepted) {
      double S[4] = {0, 0, 0, 0}, sv[2] = {0, 0}, C[4] = {0, 0, 0, 0}, rk[2] = {0, 0};
      bool ok = true;
      for (int k = 0; k < n && ok; k++) {
        const double* o = &H[(size_t)k * st];
        const double* g = o + 2;
        const double* h = o + 7;  // upper triangle of (Tx, Ty, th, k1, k2)
        // h index of (p, q), p <= q: rows 0..4 hold 5, 4, 3, 2, 1 entries
        auto hp = [&](int a, int b2) {
          if (a > b2) std::swap(a, b2);
          static const int row0[5] = {0, 5, 9, 12, 14};
          return h[row0[a] + (b2 - a)];
        };
        const double sc[5] = {1.0, 1.0, 1.0, 1.0 / s1, 1.0 / s2};
        double A[9];
        for (int a = 0; a < 3; a++)
          for (int b2 = 0; b2 < 3; b2++) A[3 * a + b2] = hp(a, b2);
        for (int a = 0; a < 3; a++) A[4 * a] *= 1.0 + lambda;
        double* y = &Y[9 * (size_t)k];
        for (int a = 0; a < 3; a++) {
          y[a] = hp(a, 3) * sc[3];
          y[3 + a] = hp(a, 4) * sc[4];
          y[6 + a] = g[a];
        }
        double Bk[6] = {y[0], y[1], y[2], y[3], y[4], y[5]}, bk[3] = {g[0], g[1], g[2]};
        if (!chol3_solve(A, y, 3)) {
          ok = false;
          break;
        }
        for (int a = 0; a < 2; a++) {
          for (int b2 = 0; b2 < 2; b2++) {
            double t = 0.0;
            for (int q = 0; q < 3; q++) t += Bk[3 * a + q] * y[3 * b2 + q];
            S[2 * a + b2] += t;
          }
          double t = 0.0;
          for (int q = 0; q < 3; q++) t += Bk[3 * a + q] * y[6 + q];
          sv[a] += t;
          rk[a] += g[3 + a] * sc[3 + a];
          for (int b2 = 0; b2 < 2; b2++) C[2 * a + b2] += hp(3 + a, 3 + b2) * sc[3 + a] * sc[3 + b2];
        }
        (void)bk;
      }
      double dk[2] = {0.0, 0.0};
      if (ok) {
        double M[4] = {C[0] * (1.0 + lambda) - S[0], C[1] - S[1], C[2] - S[2], C[3] * (1.0 + lambda) - S[3]};
        const double r0 = -(rk[0] - sv[0]), r1 = -(rk[1] - sv[1]);
        const double det = M[0] * M[3] - M[1] * M[2];
        if (!(det > 0.0) || !(M[0] > 0.0)) ok = false;
        else {
          dk[0] = (M[3] * r0 - M[1] * r1) / det;
          dk[1] = (M[0] * r1 - M[2] * r0) / det;
        }
      }
      if (!ok) {
        lambda *= 4.0;
        if (lambda > 1e16) {
          done = true;
          break;
        }
        continue;
      }
      double step = std::fabs(dk[0]) + std::fabs(dk[1]);
      for (int k = 0; k < n; k++) {
        const double* y = &Y[9 * (size_t)k];
        for (int a = 0; a < 3; a++) {
          const double d = -(y[6 + a] + y[a] * dk[0] + y[3 + a] * dk[1]);
          Pt[3 * k + a] = P[3 * k + a] + d;
          step = std::max(step, std::fabs(d));
        }
      }
      const double kt[2] = {kap[0] + dk[0], kap[1] + dk[1]};
      double Ft = 0.0;
      TRY(eval(Pt, kt, true, &Ft));
      if (std::isfinite(Ft) && Ft < F) {
        const bool small = (F - Ft) <= 1e-13 * F || step < 1e-10;
        P.swap(Pt);
        kap[0] = kt[0];
        kap[1] = kt[1];
        F = Ft;
        H = out;
        lambda = std::max(lambda / 3.0, 1e-12);
        accepted = true;
        done = small;
      } else {
        lambda *= 4.0;
        if (lambda > 1e16 || step < 1e-12) {
          done = true;
          break;
        }
      }
    }
    if (done) {
      it++;
      break;
    }
  }
  // the refined motion into the grain records (the last accepted point's outputs are in H)
  for (int k = 0; k < n; k++) {
    ptg_grain& r = c->h_grains[list[k]];
    r.dx = P[3 * k];
    r.dy = P[3 * k + 1];
    r.drot = P[3 * k + 2];
    r.NCC_subpix = H[(size_t)k * st + 1];
  }
  CK(c, hipMemcpyAsync(c->d_grains, c->h_grains.data(), sizeof(ptg_grain) * c->G, hipMemcpyHostToDevice,
                       c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  lens[0] = kap[0] / s1;
  lens[1] = kap[1] / s2;
  if (f_end) *f_end = F;
  if (iters) *iters = it;
  return PTG_OK;
}

int ptg_remap(void* p, int src, int dst, const double* params14) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(valid_index(c, src));
  TRY(valid_index(c, dst));
  if (src == dst) return c->fail(PTG_ERR_ARG, "remap source and destination must differ");
  if (!params14) return c->fail(PTG_ERR_ARG, "null distortion parameters");
  const Image& s = c->img[src];
  if (!s.d) return c->fail(PTG_ERR_STATE, "image[%d] not set", src);
  Image& d = c->img[dst];
  TRY(c->order_after_upload(src));
  TRY(c->order_after_upload(dst));
  const size_t bytes = sizeof(uint16_t) * (size_t)s.dimx * s.dimy;
  if (!d.owned || d.cap < bytes) {
    TRY(c->finish_upload(dst));
    CK(c, hipStreamSynchronize(c->stream));
    if (d.owned && d.d) (void)hipFree(d.d);
    d.d = nullptr;
    if (hipMalloc((void**)&d.d, bytes) != hipSuccess) {
      d.owned = false;
      return c->fail(PTG_ERR_NOMEM, "hipMalloc(%zu) for image[%d]", bytes, dst);
    }
    d.owned = true;
    d.cap = bytes;
  }
  d.dimx = s.dimx;
  d.dimy = s.dimy;
  d.pitch = s.dimx;
  d.bits = s.bits;
  if (dst == c->iref) c->require_precomputations = true;
  hipEvent_t ev;
  c->mark_begin(PTG_T_REMAP, &ev);
  CK(c, launch_remap(*c, src, dst, params14));
  c->mark_end(PTG_T_REMAP, ev);
  return PTG_OK;
}

int ptg_set_option(void* p, int option, int64_t value) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (option == PTG_OPT_SEARCH_STREAMS) {
    if (value != 1 && value != 2) return c->fail(PTG_ERR_ARG, "PTG_OPT_SEARCH_STREAMS takes 1 or 2, not %lld",
                                                 (long long)value);
    c->serial_search = value == 1;
    return PTG_OK;
  }
  if (option == PTG_OPT_ROT_TABLES) {
    if (value != 0 && value != 1) return c->fail(PTG_ERR_ARG, "PTG_OPT_ROT_TABLES takes 0 or 1, not %lld",
                                                  (long long)value);
    TRY(c->sync_and_check());
    bool redo = false;
    TRY(check_rot_guard(c, &redo));  // (entries of unguarded work are dropped)
    c->rot_host = value == 1;
    c->snap = Ctx::SNAP_INVALID;
    return PTG_OK;
  }
  if (option == PTG_OPT_SEARCH_ROLES) {
    if (value != 1 && value != 2) return c->fail(PTG_ERR_ARG, "PTG_OPT_SEARCH_ROLES takes 1 or 2");
    c->exact_roles = (int)value;
    c->tune_gen++;  // (a captured pair's launches change)
    return PTG_OK;
  }
  if (option == PTG_OPT_BUCKET_ORDER) {
    if (value != 0 && value != 1) return c->fail(PTG_ERR_ARG, "PTG_OPT_BUCKET_ORDER takes 0 or 1");
    c->bucket_order = (int)value;
    c->tune_gen++;  // (a captured pair's launches change)
    return PTG_OK;
  }
  if (option == PTG_OPT_DEVICE_QUEUES) {
    if (value != 0 && value != 1) return c->fail(PTG_ERR_ARG, "PTG_OPT_DEVICE_QUEUES takes 0 or 1");
    c->device_queues = value == 1;
    return PTG_OK;
  }
  if (option == PTG_OPT_GRAPH) {
    if (value != 0 && value != 1) return c->fail(PTG_ERR_ARG, "PTG_OPT_GRAPH takes 0 or 1");
    c->use_graph = value == 1;
    if (!c->use_graph) {
      TRY(c->sync_and_check());
      drop_graph(c);
      c->graph_warm_key = c->graph_bad_key = 0;
    }
    return PTG_OK;
  }
  if (option == PTG_OPT_ROT_GUARD_TEST) {
    if (value < 0 || value > 2) return c->fail(PTG_ERR_ARG, "PTG_OPT_ROT_GUARD_TEST takes 0, 1 or 2");
    c->rot_guard_test = (int)value;
    return PTG_OK;
  }
  return c->fail(PTG_ERR_ARG, "unknown option %d", option);
}

int ptg_get_graph_stats(void* p, int64_t* captures, int64_t* launches, int64_t* nodes) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (captures) *captures = c->graph_captures;
  if (launches) *launches = c->graph_launches;
  if (nodes) {
    size_t n = 0;
    if (c->graph) CK(c, hipGraphGetNodes(c->graph, nullptr, &n));
    *nodes = (int64_t)n;
  }
  return PTG_OK;
}

int ptg_get_rot_guard_stats(void* p, int64_t* checked, int64_t* redone, double* max_diff) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(c->sync_and_check());
  if (checked) *checked = c->rot_checked;
  if (redone) *redone = c->rot_redos;
  if (max_diff) *max_diff = c->rot_max_diff;
  return PTG_OK;
}

int ptg_set_timing(void* p, int enable) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  c->timing = enable != 0;
  return PTG_OK;
}

static int drain_timing(Ctx* c) {
  CK(c, hipStreamSynchronize(c->stream));
  for (auto& pe : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      c->stage_ms[pe.stage] += ms;
      c->stage_launches[pe.stage] += 1;
    }
    c->event_pool.push_back(pe.a);
    c->event_pool.push_back(pe.b);
  }
  c->pending.clear();
  return PTG_OK;
}

int ptg_get_stage_times(void* p, double* ms, int64_t* launches, int n) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(drain_timing(c));
  for (int i = 0; i < n && i < PTG_T_COUNT; i++) {
    if (ms) ms[i] = c->stage_ms[i];
    if (launches) launches[i] = c->stage_launches[i];
  }
  return PTG_OK;
}

int ptg_reset_stage_times(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  TRY(drain_timing(c));
  for (int i = 0; i < PTG_T_COUNT; i++) {
    c->stage_ms[i] = 0.0;
    c->stage_launches[i] = 0;
  }
  return PTG_OK;
}

int ptg_get_subpix_evals_per_grain(void* p, int32_t* per_grain) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!per_grain) return c->fail(PTG_ERR_ARG, "null output");
  if (!c->d_evals_g || c->G == 0) {
    for (int g = 0; g < c->G; g++) per_grain[g] = 0;
    return PTG_OK;
  }
  CK(c, hipMemcpyAsync(per_grain, c->d_evals_g, sizeof(int32_t) * c->G, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

int ptg_get_exact_fallbacks(void* p, int64_t* searches, int64_t* grains) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!searches || !grains) return c->fail(PTG_ERR_ARG, "null output");
  if (c->any_flags_pending()) TRY(c->sync_and_check());
  *searches = c->exact_fallbacks;
  *grains = c->exact_grain_fallbacks;
  return PTG_OK;
}

int ptg_get_subpix_evals(void* p, int64_t* evals) {
  Ctx* c = as_ctx(p);
  if (!c) return PTG_ERR_ARG;
  if (!evals) return c->fail(PTG_ERR_ARG, "null output");
  CK(c, hipMemcpyAsync(evals, c->d_evals, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return PTG_OK;
}

}  // extern "C"

