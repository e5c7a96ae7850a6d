// ptg_internal.hpp — the context behind the C ABI of libptgpu and the launchers each
// kernel translation unit provides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ptgpu.h"

namespace ptg {

constexpr int kMaxImages = 4;

struct Image {
  uint16_t* d = nullptr;  // device frame, row-major
  int32_t dimx = 0, dimy = 0;
  int64_t pitch = 0;      // elements
  int32_t bits = 16;
  bool owned = false;
  size_t cap = 0;         // bytes allocated when owned
};

// Per-candidate-step argmax of one (grain, rotation step) work item.
struct StepBest {
  double ncc;
  int32_t idx;  // u-major candidate index within the step, INT32_MAX when none > 0
  int32_t pad;
};

// A block's near-tie list of the exact search: its best candidate by s = num^2 / d1 and the
// next one within the window, index INT32_MAX when none; ovf when a third lay within it.
struct ExactBest {
  double s1, s2;
  int32_t i1, i2;
  int32_t ovf, pad;
};

// A (grain, rotation step) whose device rotation might round differently from the host C
// library's: the angle and the device's cosine and sine, for the host to compare with its own
// (k = -1: the reference rotation refrot of the precomputation).
struct RotGuardEntry {
  int32_t g, k;
  double theta, c, s;
};
// Entries the guard list holds; more than that sends the stage back to the host tables.
constexpr int kRotGuardCap = 65536;

// Arguments of the device rotation tables (k_rot_tables).
struct RotArgs {
  const ptg_grain* grains;
  const int32_t* pattern;   // (dx, dy) pairs
  const int64_t* pbegin;    // G + 1
  const int32_t* list;      // grains of the stage
  const int32_t* count;     // device count of list entries
  const double* steps;      // rotation-step values
  int32_t nsteps;
  int32_t all;              // list every (grain, step) in the guard (a test of the host check)
  double* cs;               // out: per grain, steps x (cos, sin)
  RotGuardEntry* guard;
  int32_t* guard_count;
};

// The frames of a pair as the kernels of the tracking path read them: a small device record the
// host rewrites (k_set_frames, in stream order) when the images or their roles change, so that a
// captured pair replays on whatever frames the context holds.
struct FrameDesc {
  const uint16_t* I0;  // reference frame
  const uint16_t* I1;  // current frame
  int64_t pitch0, pitch1;
  int32_t dimx, dimy;    // the current frame's size (the same as the reference's for a search)
  int32_t dimx0, dimy0;  // the reference frame's size
};

// Arguments of the integer-search kernels (plain and exact forms).
struct SearchArgs {
  const FrameDesc* fr;      // the pair's frames (device)
  ptg_grain* grains;
  const int32_t* pattern;   // (dx, dy) pairs
  const int64_t* pbegin;    // G + 1
  const int32_t* p0;        // pattern0_rotated pairs, same indexing as pattern
  const double* d0;         // I0 at p0 minus mean0, same indexing
  const uint32_t* pwords;   // per pattern pixel: dx, dy as int8 and I0 at p0 (u16), same indexing
  const int32_t* prad;      // per grain: every rotated offset lies in [-prad, prad]^2
  const int32_t* list;      // grains of this stage
  const int32_t* count;     // device count of list entries
  const int32_t* act;       // per list slot: searched in this stage (null: every slot)
  const double* cs;         // per list slot g: steps x (cos, sin) of refrot + drot + step
  const double* steps;      // rotation-step values
  int32_t nsteps;
  int32_t left, right, up, down;
  int32_t win_cap;          // LDS elements reserved for the window (of one shift tile)
  int32_t off_cap;          // LDS elements reserved for rotated offsets
  int32_t tile_u, tile_v;   // plain kernel: shifts per tile (U, V when the window fits the LDS)
  StepBest* best;           // list slot x nsteps
};

// A launch shape of the exact search's MFMA kernel for one radius bucket: N-tiles of 16
// vertical shifts per pass, waves (shift columns in flight) per workgroup, workgroups per
// grain, and the wave roles (1: every wave multiplies and scores its own columns; 2: producer
// waves multiply, as many consumer waves score, nsu being the producers). Zero fields take the
// built-in heuristic's choice.
struct ExactShape {
  int32_t ntp = 0, nsu = 0, nsub = 0, roles = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of the exact search (each radius bucket's template and base-mask
  // kernels run ahead of the MFMA kernels of the previous buckets), default priority, and
  // its fork/join events (created on first use)
  hipStream_t aux_stream = nullptr;
  std::vector<hipEvent_t> aux_events;
  // further streams the radius buckets' MFMA searches alternate over (with the library
  // stream), and the events that join them back into it
  std::vector<hipStream_t> search_streams;
  std::vector<hipEvent_t> search_joins;
  int ensure_aux(size_t nevents);
  // PTG_OPT_SEARCH_STREAMS = 1: every kernel of an exact search in order on `stream`
  bool serial_search = false;
  // PTG_OPT_SEARCH_ROLES: the MFMA kernel's wave roles where no tuned shape says otherwise
  // (1: each wave multiplies and scores its own columns; 2: producer and consumer waves)
  int exact_roles = 1;
  // PTG_OPT_BUCKET_ORDER: the radius buckets' launches by increasing radius (0) or by
  // decreasing work (1)
  int bucket_order = 0;
  // k_precompute ran without its ordered sums: k_precompute_chain is still to be launched
  // (by the main stage's exact search beside its searches, else on `stream` before any reader)
  bool chain_pending = false;
  std::string err;

  Image img[kMaxImages];
  int iref = 0, icur = 1;
  FrameDesc* d_frames = nullptr;  // the frames' device record, and what was last written to it
  FrameDesc h_frames{};
  bool frames_written = false;
  int set_frames();  // writes the record when the images or their roles changed

  // grains
  int32_t G = 0;
  bool grains_set = false;  // ptg_set_grains was called (possibly with zero grains)
  ptg_grain* d_grains = nullptr;
  int32_t* d_pattern = nullptr;
  int64_t* d_pbegin = nullptr;
  int32_t* d_p0 = nullptr;
  double* d_d0 = nullptr;
  uint32_t* d_pwords = nullptr;  // per pattern pixel: (dx, dy) as int8, I0(p0) as u16
  uint32_t* d_spwords = nullptr; // per pattern pixel: rotated p0 offsets as int8, I0(p0) as u16
  int32_t* d_prad = nullptr;
  int64_t total_px = 0;
  std::vector<int32_t> h_pattern;
  std::vector<int64_t> h_pbegin;
  std::vector<int32_t> h_prad;
  std::vector<uint8_t> h_disk;  // per grain: the pattern is a full digital disk (each cell of x^2 + y^2 <= max once)
  std::vector<int32_t> h_diskm2; // for such a pattern: the largest x^2 + y^2
  std::vector<ptg_grain> h_grains;  // host snapshot, valid when h_valid
  bool h_valid = false;
  bool require_precomputations = true;

  // stage lists: [0] unmasked grains, [1] rescue queue, [2] super-rescue queue
  int32_t* d_list[3] = {nullptr, nullptr, nullptr};
  int32_t* d_counts = nullptr;  // 3 counters
  int32_t n_unmasked = 0;
  std::vector<int32_t> h_unmasked;  // host copy of list [0]
  // list [0] ordered by decreasing pattern size (ties by index): the subpixel stage
  // schedules its longest grains first
  int32_t* d_list_lpt = nullptr;
  std::vector<int32_t> h_list_lpt;
  int32_t max_prad_unmasked = 0;
  const std::vector<int32_t>& h_unmasked_ref() const { return h_unmasked; }

  // search scratch
  double* d_cs = nullptr;
  size_t cs_cap = 0;
  double* d_steps_st[3] = {nullptr, nullptr, nullptr};  // each stage's rotation-step values
  size_t steps_cap_st[3] = {0, 0, 0};
  std::vector<double> h_steps_st[3];  // what each holds
  StepBest* d_best = nullptr;
  size_t best_cap = 0;
  double* d_cs0 = nullptr;  // cos, sin of refrot per grain (precompute)
  // rotation tables: computed on the device (default) with a guard list checked by the host
  // after each stage, or on the host with its C library (rot_host, also the guard's fallback)
  bool rot_host = false;
  int rot_guard_test = 0;  // 1: list every (grain, step); 2: also treat every entry as differing
  // rescue stages' queues stay on the device (with device tables; else they are read back)
  bool device_queues = false;
  // a pair captured as one HIP graph and replayed (PTG_OPT_GRAPH): the capture's state
  bool use_graph = false;
  bool capturing = false;   // inside a capture: no host synchronisation and no upload
  bool graph_pair = false;  // a pair run for a graph or in one: rescue queues stay on the device
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_key = 0, graph_warm_key = 0, graph_bad_key = 0;
  bool graph_flags[3] = {false, false, false};  // exact-search flags the captured pair copies out
  bool graph_unfit[3] = {false, false, false};
  int64_t graph_captures = 0, graph_launches = 0;
  uint64_t tune_gen = 0;   // bumped when the tuned launch shapes change
  uint64_t table_uploads = 0;  // bumped whenever an exact search's tables or a stage's steps are uploaded
  int32_t* d_guard_count = nullptr;  // the guard list's count, then its entries
  RotGuardEntry* d_guard = nullptr;
  int64_t rot_redos = 0;   // stages (or pairs) redone with the host tables
  int64_t rot_checked = 0; // guard entries the host compared
  double rot_max_diff = 0.0;  // largest |device - host| cosine or sine among them
  ptg_grain* d_grains_save = nullptr;  // the records at the start of the work a redo repeats
  size_t grains_save_cap = 0;

  // subpixel
  int64_t* d_evals = nullptr;
  int32_t* d_evals_g = nullptr;  // per grain, last subpixel stage
  size_t evals_g_cap = 0;
  double* d_sd0 = nullptr;  // per grain: sum of d0 (used by the one-pass objective)
  double* d_m0 = nullptr;   // per grain: the mean0 d0 was formed with (d0 = I0(p0) - m0)

  // exact-form scratch and the cached host plan of the last keyed list
  // (one of each per search stage: the stages' zones differ, and their device tables stay
  // where a captured pair's kernels find them)
  void* d_exact_scratch[3] = {nullptr, nullptr, nullptr};
  size_t exact_scratch_cap[3] = {0, 0, 0};
  void* exact_plan[3] = {nullptr, nullptr, nullptr};
  uint64_t exact_dev_key[3] = {0, 0, 0};  // plan whose tables the scratch holds (0: none)
  uint64_t unmasked_version = 0;    // bumped whenever the unmasked list is rebuilt
  // launch shapes chosen by ptg_autotune_search, by bucket key (value mode, steps, zone,
  // largest pattern radius and grain count of the bucket); a tuning run probes tune_shape on
  // bucket tune_bucket, records per bucket its key and the shape used, and brackets the
  // whole search with two events on the library stream
  std::unordered_map<uint64_t, ExactShape> exact_tuned;
  int tune_bucket = -1;
  ExactShape tune_shape;
  bool tune_run = false;
  std::vector<uint64_t> tune_keys;
  std::vector<ExactShape> tune_used;
  hipEvent_t tune_ev[2] = {nullptr, nullptr};
  void* d_tune_save = nullptr;  // the grains, restored before each tuning run
  size_t tune_save_cap = 0;
  int64_t exact_fallbacks = 0;      // exact-form searches that ran the plain form instead
  int64_t exact_grain_fallbacks = 0;  // grains of exact-form searches served by the plain kernel

  // distortion calibration data (ptg_equiproj_set_data)
  int32_t eq_G = 0, eq_nimg = 0, eq_npairs = 0;
  double eq_NCC_min = 0.0;
  int32_t* d_eq_pairs = nullptr;
  double* d_eq_ref = nullptr;
  double* d_eq_disp = nullptr;
  double* d_eq_ncc = nullptr;
  double* d_eq_partial = nullptr;
  double* d_eq_out = nullptr;
  double* d_eq_res = nullptr;
  double* d_eq_jac = nullptr;
  double* d_eq_orig = nullptr;  // given first positions (ptg_equiproj_set_origins)
  bool eq_orig_set = false;
  size_t eq_orig_cap = 0;
  size_t eq_pairs_cap = 0, eq_ref_cap = 0, eq_disp_cap = 0, eq_ncc_cap = 0, eq_partial_cap = 0, eq_out_cap = 0,
         eq_res_cap = 0, eq_jac_cap = 0;

  // objective / sampler batches
  void* d_tmp = nullptr;
  size_t tmp_cap = 0;
  // this device's staging array of the allocation-free objective entry point (file-scope
  // device memory, not owned by the context)
  void* d_noalloc = nullptr;

  // joint refinement: grain list, parameters, per-grain outputs
  void* d_joint = nullptr;
  size_t joint_cap = 0;

  // timing
  bool timing = false;
  struct Pending { int stage; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double stage_ms[PTG_T_COUNT] = {0};
  int64_t stage_launches[PTG_T_COUNT] = {0};

  int fail(int code, const char* fmt, ...);
  int check(hipError_t e, const char* what);
  int ensure(void** p, size_t* cap, size_t bytes);
  hipEvent_t get_event();
  void mark_begin(int stage, hipEvent_t* ev, hipStream_t on = nullptr);
  void mark_end(int stage, hipEvent_t ev, hipStream_t on = nullptr);
};

// Raises kernel fn's dynamic-LDS limit on device dev to at least `bytes` when that exceeds
// the 64 KiB default: the runtime is called only when the size grows past the largest set
// before for (dev, fn), since a call per launch costs enough host time to show in a pair's.
hipError_t ensure_kernel_lds(int dev, const void* fn, size_t bytes);

// launchers (each returns hipError_t of the launch)
hipError_t launch_begin_pair(Context& c);
// dev_trig: cos/sin(refrot) on the device, with the guard of k_rot_tables (k = -1); else the
// host's from d_cs0.
// dev_trig: cos/sin(refrot) on the device (guarded); chain false leaves C0C0, the sum of d0 and
// d0 itself to launch_precompute_chain (on any stream ordered after it)
hipError_t launch_precompute(Context& c, bool dev_trig, bool chain);
hipError_t launch_precompute_chain(Context& c, hipStream_t st);
hipError_t launch_set_frames(Context& c, const FrameDesc& f);
// The rotation tables of a stage's grains on the device (grid: list_cap x step blocks).
hipError_t launch_rot_tables(Context& c, const RotArgs& r, int list_cap);
hipError_t launch_search_plain(Context& c, const SearchArgs& a, int list_cap, size_t lds_bytes);
// The plain kernel for patterns whose offsets and window exceed the LDS: the same candidates
// and arithmetic, the frame read through the caches.
hipError_t launch_search_plain_global(Context& c, const SearchArgs& a, int list_cap);
// The plain kernel over a device list of *sub_count slots of a's list (sub[j] < list_cap),
// step k of slot li written to a.best[li * best_stride + k].
hipError_t launch_search_plain_sub(Context& c, const SearchArgs& a, int list_cap, const int32_t* sub,
                                   const int32_t* sub_count, int best_stride, size_t lds_bytes);
// Reduces nparts partial argmaxes per list slot (global candidate index k * S + u * V + v,
// first maximum wins) and applies the stage's update rules.
// abort_flag (device, may be null): when non-zero the commit changes nothing.
hipError_t launch_search_commit(Context& c, const SearchArgs& a, int list_cap, int stage, double NCC_min,
                                double NCC_min_super, int32_t* next_list, int32_t* next_count,
                                const StepBest* parts, int nparts, const int32_t* abort_flag);
hipError_t launch_end_pair(Context& c);
hipError_t launch_change_reference(Context& c);
hipError_t launch_build_unmasked(Context& c);

// Returned by run_search_exact, before it launches anything, when the search is beyond
// the exact form's limits (frames above 12 bits, more than 48 rotation steps, a pattern
// radius above 120, or a bucket whose window and pattern exceed the CU's LDS): the caller
// then runs the plain form, whose picks are the same.
constexpr int kExactUnfit = 1;

// exact (MFMA) form: returns an error code of the C ABI (or kExactUnfit), writes message
// into c.err.
// Writes nparts partial argmaxes per list slot into out_best. The capacity flag (template
// digit or Delta list overflow) is left in *flag_dev on the device; the caller passes it
// to the commit and checks it at its next synchronisation.
// plan_key != 0 names a list that is the same on every call with that key (the unmasked
// list of one grain set): its host tables are then built once and reused. stage_thr is the
// threshold the commit compares the stage's NCC with (NCC_min for the main stage, else
// NCC_min_super): a grain whose NCC lies within rounding of it is re-scored in the
// reference's arithmetic.
// time_kernel: bracket each MFMA kernel launch with the PTG_T_SEARCH_KERNEL timer.
// stage: the plan and scratch slot (each stage keeps its own). queue/qcount (device, or null):
// the stage's grains are the queued ones among the list's (a rescue stage searched without a
// host read-back of its queue): the slots of the queued grains are marked in a.act (which the
// call sets) and listed per radius bucket on the device, and every launch is sized for the whole
// list, its surplus workgroups returning at once.
int run_search_exact(Context& c, const SearchArgs& a, int nlist, const std::vector<int32_t>& list,
                     uint64_t plan_key, double stage_thr, StepBest* out_best, int* nparts_out,
                     const int32_t** flag_dev, bool time_kernel, int stage, const int32_t* queue,
                     const int32_t* qcount, const int32_t** act_out);
void free_exact_plans(Context& c);

// Equiprojectivity criterion of the distortion calibration (device data, see ptgpu.h).
struct EquiprojArgs {
  int32_t G, nimg, npairs, pad;
  const int32_t* pairs;  // npairs x (i, j)
  const double* ref;     // G x 2: reference pixels
  const double* orig;    // G x 2: given first positions (the POWELL_ITER form), or null: undistor(ref)
  const double* disp;    // nimg x G x 2: tracked displacements
  const double* ncc;     // nimg x G: NCC_subpix
  double NCC_min;
  double* res;           // nimg * npairs residuals, or null
  double* jac;           // nimg * npairs x 14, or null
  double* partial;       // per workgroup: sum r^2, J^T r (14), J^T J upper triangle (105)
};
hipError_t launch_equiproj(Context& c, const EquiprojArgs& a, const double* params14, double* d_out);
int equiproj_partial_blocks(int64_t m);

// subpixel
// sel 0: every unmasked grain; 1: those whose NCC (the main stage's) is at least thr; 2: the rest
hipError_t launch_subpixel(Context& c, const ptg_subpix_params& p, int sel = 0, double thr = 0.0,
                           hipStream_t st = nullptr);
hipError_t launch_objective_batch(Context& c, int32_t n, const int32_t* d_ig, const double* d_X, double dmax,
                                  double stiff, double* d_f, double* d_grad);
hipError_t launch_interp_batch(Context& c, int index, int32_t n, const double* d_xy, double* d_out);
hipError_t launch_undistort_points(Context& c, const double* params, int32_t n, const double* d_in, double* d_out);
hipError_t launch_remap(Context& c, int src, int dst, const double* params);

// joint refinement (ptg_joint.hip): per entry of d_list, from its (Tx, Ty, theta) in d_P and
// the radial lens (k1, k2, xc, yc), joint_out_stride() doubles: f, NCC, df/dp (5), the
// Gauss-Newton matrix's upper triangle (15), C1C1, a clamped flag (derivatives only when asked)
hipError_t launch_joint(Context& c, const int32_t* d_list, int32_t nlist, const double* d_P, const double* lens,
                        bool derivs, double* d_out);
int joint_out_stride();

}  // namespace ptg
