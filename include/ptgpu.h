/*
 * ptgpu.h — C ABI of the MI355X (gfx950) per-grain tracking library (libptgpu.so).
 *
 * The library replaces the per-frame-pair hot path of p-tracker's particle_tracking()
 * (src/p-tracker.cpp:460-581): precomputation, integer ZNCC search over shifts x rotation
 * steps (main, rescue and super-rescue zones), bicubic sampling, the subpixel NCC
 * objective and its optimizer, and a distortion-correction remap built on undistor().
 *
 * Conventions
 *  - Every function returns int: PTG_OK (0) on success, a negative PTG_ERR_* code on
 *    failure; ptg_last_error(ctx) then holds a one-line message. No exception crosses
 *    this boundary.
 *  - Frames are row-major uint16, element (x, y) at base[y * pitch + x]. This is the
 *    reference's image[i][x][y] (src/p-tracker.hpp:80) relabelled, not transposed data.
 *  - Grain records are caller-owned host arrays copied in and out; the context owns
 *    all device memory except frames attached with ptg_set_image_device().
 *  - One context per host thread. All work of a context runs in order on one HIP
 *    stream; calls return before the GPU finishes, except those that hand results
 *    back to host memory (ptg_get_grains, ptg_get_queue_counts, ptg_ncc_objective*,
 *    ptg_interp_cubic, ptg_get_stage_times, ptg_sync).
 *  - Reads outside a frame are undefined behaviour in the reference
 *    (src/interpol.cpp:61, src/p-tracker.cpp:703). Here they are clamped to the frame
 *    edge and the grain's `status` gets PTG_STATUS_CLAMPED, so no call can fault.
 */
#ifndef PTGPU_H
#define PTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTG_ABI_VERSION 1

enum {
  PTG_OK = 0,
  PTG_ERR_ARG = -1,    /* bad argument (null pointer, size, index, zone) */
  PTG_ERR_HIP = -2,    /* HIP runtime error */
  PTG_ERR_STATE = -3,  /* call order: e.g. search before images/grains are set */
  PTG_ERR_NOMEM = -4,  /* device allocation failed */
  PTG_ERR_LIMIT = -5   /* a size above a documented kernel limit */
};

/* Per-grain status flags (ptg_grain.status). */
enum { PTG_STATUS_CLAMPED = 1 };

/* Search stages of one frame pair, src/p-tracker.cpp:472-537. */
enum { PTG_STAGE_MAIN = 0, PTG_STAGE_RESCUE = 1, PTG_STAGE_SUPER = 2 };

/* Integer-search arithmetic.
 *  PLAIN: per candidate, the reference's two-pass FP64 sums in the reference's pixel
 *         order, unfused; NCC is bit-identical to follow_pattern_pixel(). Patterns of any
 *         size: offsets and window staged in LDS, read from the frame through the caches
 *         when they exceed it (radius above about 85 px).
 *  EXACT: per candidate, exact integer moments (sum I1, sum I1^2, sum I0*I1) from an
 *         INT8 MFMA correlation, then the ratio in FP64. Every candidate whose exact NCC
 *         lies within a proven bound of the reference's FP64 rounding error of the best is
 *         re-scored in the reference's two-pass arithmetic, so the picks are the
 *         reference's (ties included); NCC is then the reference's value bit for bit, and
 *         elsewhere within a few units in the last place of it. The digits follow the frames'
 *         declared depth (the larger of the two): 8-bit windows as one offset-INT8 digit,
 *         12-bit as two 7-bit digits, 16-bit as two offset bytes (num and d1 in int64).
 *         Grains the exact kernels cannot settle (more than two such candidates, a best NCC
 *         below 1/32, window values above the declared depth, patterns above 23,170 pixels)
 *         are searched by the plain kernel in the same stage. */
enum { PTG_SEARCH_PLAIN = 0, PTG_SEARCH_EXACT = 1 };

/* Subpixel optimizer: POWELL is derivative-free, as the reference
 * (src/p-tracker.cpp:1114-1116); LM is Levenberg-Marquardt driven by the analytic
 * gradient of the ZNCC residual through the bicubic sampler. */
enum { PTG_SUBPIX_POWELL = 0, PTG_SUBPIX_LM = 1 };

/* Same members, order and layout as search_zone_type (src/search_zone_type.hpp:20-27). */
typedef struct ptg_search_zone {
  int32_t left;    /* towards decreasing x */
  int32_t right;   /* towards increasing x */
  int32_t up;      /* towards decreasing y */
  int32_t down;    /* towards increasing y */
  int32_t num_rot; /* rotation steps each side */
  double inc_rot;  /* rotation increment (rad) */
} ptg_search_zone;

/* The scalar members of grain_type_2D (src/grain_type_2D.hpp:23-56), in the order of
 * the 13 dic_out columns (src/p-tracker.cpp:1331-1343) followed by the backups,
 * the precomputed reference moments and the mask. The pattern offset list is passed
 * separately (ptg_set_grains). */
typedef struct ptg_grain {
  int32_t refcoord_xpix; /* col 1 */
  int32_t refcoord_ypix; /* col 2 */
  double refrot;         /* col 3 */
  double radius_pix;     /* col 4 */
  double dx, dy, drot;   /* col 5-7: cumulative motion */
  double upix, vpix, rot_inc;          /* col 8-10: increments of the last pair */
  double NCC, NCC_rescue, NCC_subpix;  /* col 11-13 */
  double dx_prev, dy_prev, drot_prev;  /* backup() */
  double mean0, C0C0;                  /* do_precomputations() */
  int32_t masked;
  int32_t status; /* PTG_STATUS_* flags, written by the library */
} ptg_grain;

/* Subpixel stage parameters; defaults of src/p-tracker.hpp:122-127 in brackets. */
typedef struct ptg_subpix_params {
  double subpix_tol;              /* [1e-8] Powell relative tolerance on the objective */
  double initial_direction[3];    /* [0.01, 0.01, 0.01] Powell initial step per axis */
  double subpix_displacement_max; /* [3.0] */
  double overflow_stiffness;      /* [0.0] */
  int32_t method;                 /* PTG_SUBPIX_* */
  int32_t max_iter;               /* Powell outer iterations / LM iterations cap, <=0: 200 */
} ptg_subpix_params;

/* One frame pair, src/p-tracker.cpp:472-581. */
typedef struct ptg_track_params {
  ptg_search_zone search_zone;
  ptg_search_zone search_zone_rescue;
  ptg_search_zone search_zone_super_rescue;
  double NCC_min;       /* [0.7] */
  double NCC_min_super; /* [0.5] */
  int32_t rescue_level; /* [2] <0: no integer search; >=1 rescue; >=2 super-rescue */
  int32_t subpixel;     /* [1] */
  int32_t search_method;/* PTG_SEARCH_* */
  int32_t reserved;
  ptg_subpix_params subpix;
} ptg_track_params;

/* ---- context ------------------------------------------------------------------ */
int ptg_create(void** ctx, int device);
int ptg_destroy(void* ctx);
/* The message of the context's last failed call; with ctx == NULL, that of the last
 * ptg_create that failed on the calling thread (which then released everything it made). */
const char* ptg_last_error(const void* ctx);
int ptg_abi_version(void);

/* ---- frames: image[index], index 0..3 (0/1 are the reference's image[0]/image[1]) -- */
/* Replaces read_image() storage (src/p-tracker.cpp:75-201): copies a host frame. */
int ptg_set_image(void* ctx, int index, const uint16_t* host, int32_t dimx, int32_t dimy,
                  int64_t pitch, int32_t bits);
/* Attaches a caller-owned device frame (no copy); it must outlive its use. */
int ptg_set_image_device(void* ctx, int index, const uint16_t* dev, int32_t dimx, int32_t dimy,
                         int64_t pitch, int32_t bits);
/* read_image()'s store into image[index] (src/p-tracker.cpp:75-201), which the reference
 * does between pairs (:426, :446), as a copy that overlaps the context's work: the copy runs
 * on a copy stream of the context, after every call made before this one (which may still
 * read the slot's old contents), and the first later call that reads or replaces
 * image[index] waits for it on the GPU. Returns once the copy is queued: from pinned memory
 * (ptg_host_alloc) the copy reads `host` while later calls run, so `host` stays unchanged
 * until ptg_wait_image(ctx, index) or ptg_sync returns; from any other memory the call
 * returns once the copy has read `host` (it then waits for the work queued before it). The
 * usual sequence: image k+1 into a slot the running pair does not use, then ptg_track_pair
 * on images k-1 and k. */
int ptg_set_image_async(void* ctx, int index, const uint16_t* host, int32_t dimx, int32_t dimy,
                        int64_t pitch, int32_t bits);
/* Waits on the host until image[index]'s last ptg_set_image_async copy has read its host
 * buffer (no-op without one). */
int ptg_wait_image(void* ctx, int index);
/* Pinned (page-locked) host memory for frames given to ptg_set_image_async, and its
 * release; memory from ptg_host_alloc may be freed only after the copies that read it
 * (ptg_wait_image, ptg_sync). No counterpart in the reference (its frames live in host
 * vectors). */
int ptg_host_alloc(void* ctx, size_t bytes, void** out);
int ptg_host_free(void* ctx, void* p);
/* Copies image[index] to a host buffer of row pitch `pitch` elements (synchronises): the
 * counterpart of the reference's direct access to image[i][x][y] (src/p-tracker.hpp:80),
 * e.g. for the frame ptg_remap wrote. */
int ptg_get_image(void* ctx, int index, uint16_t* host, int64_t pitch);
/* im_index_ref / im_index_current, src/p-tracker.hpp:82-83. */
int ptg_set_image_indices(void* ctx, int im_index_ref, int im_index_current);

/* ---- grains ------------------------------------------------------------------- */
/* grain[] and each grain's pattern (src/grain_type_2D.hpp:58): pattern_xy holds
 * (dx, dy) int pairs, grain g owning pairs [pattern_begin[g], pattern_begin[g+1]). */
int ptg_set_grains(void* ctx, int32_t num_grains, const ptg_grain* grains,
                   const int32_t* pattern_xy, const int64_t* pattern_begin);
/* Overwrites the scalar records only (patterns kept). */
int ptg_put_grains(void* ctx, const ptg_grain* grains);
/* Copies the records back (synchronises). */
int ptg_get_grains(void* ctx, ptg_grain* out);

/* ---- the per-pair stages -------------------------------------------------------- */
/* grain.reset(); grain.backup() for every grain, src/p-tracker.cpp:478-479. */
int ptg_begin_pair(void* ctx);
/* do_precomputations(), src/p-tracker.cpp:624-656. */
int ptg_precompute(void* ctx);
/* follow_pattern_pixel / _rescue_pixel / _super_rescue_pixel over the stage's grain
 * list, src/p-tracker.cpp:659-1016. MAIN runs every unmasked grain and queues those
 * with best NCC < NCC_min; RESCUE runs that queue and queues those < NCC_min_super;
 * SUPER runs the second queue. */
int ptg_search(void* ctx, int stage, const ptg_search_zone* zone, double NCC_min,
               double NCC_min_super, int method);
/* follow_pattern_subpixel_xyR for every unmasked grain, src/p-tracker.cpp:1090-1128. */
int ptg_subpixel(void* ctx, const ptg_subpix_params* params);
/* upix = dx - dx_prev, ..., src/p-tracker.cpp:577-581. */
int ptg_end_pair(void* ctx);
/* The reference swap at the end of a pair, src/p-tracker.cpp:587-611. */
int ptg_change_reference(void* ctx);
/* One whole frame pair, src/p-tracker.cpp:460-581 (precomputation when needed, the
 * stages gated by rescue_level and subpixel, then the increments). */
int ptg_track_pair(void* ctx, const ptg_track_params* params);
/* num_to_be_rescued / num_to_be_super_rescued of the last pair (synchronises). */
int ptg_get_queue_counts(void* ctx, int32_t* num_to_be_rescued, int32_t* num_to_be_super_rescued);
/* Waits until the context's work and frame copies are done. */
int ptg_sync(void* ctx);

/* ---- the subpixel API points ----------------------------------------------------- */
/* NCC_to_minimize_xyR(X) for grain igrain, src/p-tracker.cpp:1018-1088: the value the
 * NCC_optimizer functor returns (src/NCC_optimizer_functor.cpp:20). Batched form:
 * n evaluations, grain igrain[k] at X[3k..3k+2] -> f[k]. */
int ptg_ncc_objective(void* ctx, int32_t igrain, const double* X, double subpix_displacement_max,
                      double overflow_stiffness, double* f);
int ptg_ncc_objective_batch(void* ctx, int32_t n, const int32_t* igrain, const double* X,
                            double subpix_displacement_max, double overflow_stiffness, double* f);
/* ptg_ncc_objective_batch without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (host optimizers calling the objective in a
 * tight loop, code under HIP graph capture or in serving loops). The grain indices, points
 * and values are staged in a library-owned static device array of 576 KiB per GPU: at most
 * 16384 evaluations per call (PTG_ERR_LIMIT above). The reference moments must have been
 * computed for the current frames (PTG_ERR_STATE otherwise: call ptg_precompute first).
 * Every call on one GPU shares that array: calls that overlap on one GPU (several
 * contexts or threads at once) race. */
int ptg_ncc_objective_batch_noalloc(void* ctx, int32_t n, const int32_t* igrain, const double* X,
                                    double subpix_displacement_max, double overflow_stiffness,
                                    double* f);
/* Gradient of the same objective with respect to X (analytic, through the sampler). */
int ptg_ncc_objective_grad_batch(void* ctx, int32_t n, const int32_t* igrain, const double* X,
                                 double* f, double* grad);
/* image_interpCubic::getValue(image[index], x, y), src/interpol.cpp:55-86, for n points
 * xy[2k], xy[2k+1] -> out[k]. */
int ptg_interp_cubic(void* ctx, int index, int32_t n, const double* xy, double* out);

/* ---- distortion ------------------------------------------------------------------ */
/* undistor(X, xd, yd, xu, yu), src/distortion_correction.cpp:20-61, for n points. */
int ptg_undistort_points(void* ctx, const double* params14, int32_t n, const double* xy_d,
                         double* xy_u);
/* Distortion-correction remap: image[dst_index](x, y) = bicubic sample of
 * image[src_index] at the distorted point whose undistor() image is (x, y), rounded to
 * the nearest integer and clamped to [0, 2^bits - 1]. The inverse is found by Newton
 * iteration on undistor(). dst gets src's size and bits. */
int ptg_remap(void* ctx, int src_index, int dst_index, const double* params14);

/* ---- joint refinement of the grains' motion and a radial lens (differentiation form) --
 * No counterpart in the reference, which fits the lens only through the equiprojectivity
 * criterion (src/distortion_correction.cpp:135-479). undistor's radial part
 * U(p) = p + (p - c)(K1 r^2 + K2 r^4) (src/distortion_correction.cpp:20-39) maps frame pixels
 * to scene coordinates; grain g's pattern pixel q = ref + p0_i moves rigidly in the scene by
 * (Tx, Ty) and the rotation theta about U(ref), with NCC_to_minimize_xyR's rotation
 * convention, and is sampled in image[current] at U^-1 of the result. Its objective is
 * 1 - NCC (NCC_to_minimize_xyR with the lens; with K1 = K2 = 0 exactly that function at
 * X = (Tx - dx, Ty - dy, theta - drot)).
 * ptg_joint_eval: for grains igrain[k] at P[3k..3k+2] = (Tx, Ty, theta) and lens =
 * (K1, K2, xc, yc): f[k]; grad[5k..5k+4] = df/d(Tx, Ty, theta, K1, K2) (may be NULL); jtj[15k..]
 * = the upper triangle, row by row, of the Gauss-Newton matrix of the normalised ZNCC
 * residual over the same five (may be NULL). Analytic, through the sampler and the inverse
 * lens. */
int ptg_joint_eval(void* ctx, int32_t n, const int32_t* igrain, const double* P, const double* lens, double* f,
                   double* grad, double* jtj);
/* Minimises the sum of the objective over the unmasked grains with NCC_subpix >= ncc_min,
 * jointly over each grain's (Tx, Ty, theta), started from its (dx, dy, drot), and the lens's
 * (K1, K2) about the fixed centre (lens[2], lens[3]), started from lens[0..1]: Levenberg-
 * Marquardt on the Gauss-Newton matrix, the per-grain blocks eliminated (Schur complement)
 * for the lens step. On return lens[0..1] hold K1, K2, those grains' (dx, dy, drot) hold their
 * scene motion (Tx, Ty, theta) and NCC_subpix = 1 - f; f_ini, f_end (the sums), iters and
 * ngrains (the grains used) may be NULL. */
int ptg_refine_joint_distortion(void* ctx, double* lens, double ncc_min, int32_t max_iter, double* f_ini,
                                double* f_end, int32_t* iters, int32_t* ngrains);

/* ---- distortion calibration (procedure correction_distortion) --------------------- */
/* The equiprojectivity criterion of src/distortion_correction.cpp:135-479: G grains at
 * reference pixels ref_xy (G x 2, the reference's (double)refcoord_xpix / ypix), nimg
 * images after the reference with tracked displacements disp (nimg x G x 2: dx, dy) and
 * NCC_subpix ncc (nimg x G), and npairs grain pairs (npairs x 2: i, j; the reference's
 * precompute_paires, :62-85). Residual (k, l) is
 *   e . ((B2 - B1) - (A2 - A1)),  e = (B1 - A1) / |B1 - A1|,
 * with A1, B1 the undistor() images of the reference pixels of pair l and A2, B2 those of
 * its tracked positions in image k, or 0 when either grain has NCC_subpix < NCC_min
 * (the reference skips the pair). Copies the data to the device. */
int ptg_equiproj_set_data(void* ctx, int32_t G, const double* ref_xy, int32_t nimg, const double* disp,
                          const double* ncc, int32_t npairs, const int32_t* pairs, double NCC_min);
/* The POWELL_ITER form of the criterion, src/distortion_correction.cpp:316-354: A1, B1 are
 * the given positions orig_xy (G x 2; the reference's orig_x / orig_y: the reference pixels
 * at its first iteration, their undistor() images at the previous parameters afterwards)
 * instead of undistor() of the reference pixels, and carry no derivative. NULL returns to the
 * standard form, as does the next ptg_equiproj_set_data. */
int ptg_equiproj_set_origins(void* ctx, const double* orig_xy);
/* At params14 (undistor's 14 parameters, taken as given): f = the sum of squared residuals
 * (the reference's Powell functor, :277-303); jtr = J^T r (14) and jtj = J^T J (14 x 14,
 * row major) of the residuals' analytic Jacobian J; res = the residuals (nimg * npairs,
 * image major as the reference's fvec, :409-475) and jac = J (nimg * npairs x 14, row
 * major). Every output may be NULL. */
int ptg_equiproj_eval(void* ctx, const double* params14, double* f, double* jtr, double* jtj, double* res,
                      double* jac);
/* Minimises the criterion over the parameters with free14[k] != 0 (the others held) by
 * Levenberg-Marquardt on the analytic J^T J and J^T r, in place of the reference's Powell
 * or lmdif (forward-difference Jacobian) runs, :277-475. params14 is updated; f_ini and
 * f_end receive the criterion before and after, iters the iterations (each may be NULL). */
int ptg_calibrate_distortion(void* ctx, double* params14, const int32_t* free14, int32_t max_iter, double* f_ini,
                             double* f_end, int32_t* iters);

/* ---- timing ---------------------------------------------------------------------- */
/* Stage timers: HIP events around each stage on the context's stream. */
enum {
  PTG_T_PRECOMPUTE = 0,
  PTG_T_SEARCH_MAIN = 1,
  PTG_T_SEARCH_RESCUE = 2,
  PTG_T_SEARCH_SUPER = 3,
  PTG_T_SUBPIXEL = 4,
  PTG_T_REMAP = 5,
  PTG_T_SEARCH_KERNEL = 6, /* the candidate-scoring kernel(s) alone, main stage: the FP64
                              kernel (plain) or the MFMA kernel, one launch per radius
                              bucket (exact) */
  PTG_T_COUNT = 7
};
int ptg_set_timing(void* ctx, int enable);
/* Options of a context (defaults in brackets); no counterpart in the reference.
 *  PTG_OPT_SEARCH_STREAMS [2]: 2 overlaps an exact search's radius buckets over three
 *  streams (each bucket's preparation kernels ahead of the searches, consecutive searches
 *  on two streams); 1 runs every kernel of the search in order on the context's stream, so
 *  that per-launch timings (PTG_T_SEARCH_KERNEL, a profiler's trace) do not overlap.
 *  PTG_OPT_ROT_TABLES [0]: where the rotation tables cos/sin(refrot + drot + step)
 *  (src/p-tracker.cpp:636-637, 687-689) are formed. 0: on the device, from the device
 *  library's cosine and sine; every (grain, step) with a rotated pattern coordinate within
 *  (|dx| + |dy| + 2) 2^-44 of a half-integer (where a cosine or sine differing from the host C
 *  library's by up to 2^-50 could change a rounded offset) is listed, and after the call's
 *  stages the host compares those with its own C library's values and the offsets they give;
 *  if any offset differs, the call's records are put back and its work is redone on the host's
 *  tables (the rotated offsets, hence the picks, are the reference's either way). 1: on the
 *  host, as the reference (a host round trip per stage).
 *  PTG_OPT_ROT_GUARD_TEST [0]: a test of that check: 1 lists every (grain, step), 2 also
 *  redoes every call whose list is not empty.
 *  PTG_OPT_DEVICE_QUEUES [0]: 1, with device tables: a rescue stage's queue (the grains the
 *  previous stage left below its threshold, src/p-tracker.cpp:732-735, 865-866) stays on the
 *  device: the exact search is planned once over all unmasked grains and serves the queued
 *  ones, its launches sized for all, with no host round trip between the stages (what a captured
 *  pair needs; an empty queue still costs the stage's launches). 0 reads the queue back and
 *  plans the stage for it (an empty queue: no launch).
 *  PTG_OPT_GRAPH [0]: 1 runs ptg_track_pair as one HIP graph: the first pair with new
 *  parameters, grains or buffers runs as launches, the next is captured from the same launch
 *  sequence (rescue queues on the device) and it and every later pair with that key replay the
 *  graph, on the frames the context holds at the time (read through a device frame record), as
 *  a frame sequence does between reference swaps; stage timers are not kept for replayed pairs
 *  (with timing on, or host tables, pairs run as launches).
 */
/*  PTG_OPT_SEARCH_ROLES [1]: the wave roles of the exact search's MFMA kernel (where no tuned
 *  shape says otherwise). 1: every wave multiplies and scores its own shift columns in turn;
 *  2: producer waves stage shift columns and run the MFMA GEMM, each handing a vertical pass's
 *  cross terms to a consumer wave through LDS, which adds the window-sum corrections and keeps
 *  the near-tie candidates while its producer multiplies the next pass (at most 4 + 4 waves).
 *  The picks are the same; the autotuner tries both.
 *  PTG_OPT_BUCKET_ORDER [0]: the order of the exact search's radius-bucket launches. 0: by
 *  increasing pattern radius (the first launches' preparation is the cheapest); 1: by
 *  decreasing work (pattern pixels in all: the longest launch starts first and the last ones
 *  are short), which suits grain sets whose work grows with the radius. The picks are the same.
 */
enum {
  PTG_OPT_SEARCH_STREAMS = 0,
  PTG_OPT_ROT_TABLES = 1,
  PTG_OPT_ROT_GUARD_TEST = 2,
  PTG_OPT_DEVICE_QUEUES = 3,
  PTG_OPT_GRAPH = 4,
  PTG_OPT_SEARCH_ROLES = 5,
  PTG_OPT_BUCKET_ORDER = 6
};
int ptg_set_option(void* ctx, int option, int64_t value);
/* The rotation guard's counters since the context was created: guard entries the host
 * compared, calls redone on the host's tables, and the largest difference between a device
 * cosine or sine and the host C library's among the compared entries (the guard assumes at
 * most 2^-50). Synchronises. */
int ptg_get_rot_guard_stats(void* ctx, int64_t* checked, int64_t* redone, double* max_diff);
/* Pairs captured as HIP graphs and graph launches (PTG_OPT_GRAPH) since the context was
 * created, and the nodes of the current graph (the pair's kernel launches and copies). */
int ptg_get_graph_stats(void* ctx, int64_t* captures, int64_t* launches, int64_t* nodes);
/* The exact search's autotuner; no counterpart in the reference (it replaces the launch
 * shapes of the built-in heuristic, which picks them from the LDS and register budgets).
 * For each radius bucket of a main-stage search over `zone` on the current frames and
 * grains, tries each candidate shape: N-tiles of 16 vertical shifts per pass (1-3), waves per
 * workgroup (1-8) and workgroups per grain (1, 2, 4), timing the whole search's GPU span
 * (HIP events on the context's stream, the buckets overlapped as in any search); a shape
 * replaces the heuristic's when the median of three spans is shorter, and is kept under the
 * bucket's key (value mode, rotation steps, zone, largest pattern radius, grain count); the
 * buckets' launch order (PTG_OPT_BUCKET_ORDER) is set to the faster of the two first.
 * Later exact searches with a bucket of the same key use it; picks and NCC do not depend on
 * the shape. The grains are left as they were. ntried: searches run; ms_before, ms_after:
 * the search's span with the heuristic's shapes and with the chosen ones (medians of three;
 * any output may be NULL). */
int ptg_autotune_search(void* ctx, const ptg_search_zone* zone, int32_t* ntried, double* ms_before, double* ms_after);
/* Forgets every autotuned shape (the heuristic's shapes apply again). */
int ptg_clear_tuning(void* ctx);
/* Accumulated milliseconds and launch counts since the last reset (synchronises). */
int ptg_get_stage_times(void* ctx, double* ms, int64_t* launches, int n);
int ptg_reset_stage_times(void* ctx);
/* Work counters of the last subpixel stage: total objective evaluations, and per grain
 * (per_grain may be NULL; it receives G counts, 0 for masked grains). */
int ptg_get_subpix_evals(void* ctx, int64_t* evals);
int ptg_get_subpix_evals_per_grain(void* ctx, int32_t* per_grain);
/* Work the exact form handed to the plain form since the context was created (the picks
 * are the same; NCC comes from the two-pass formula). searches: PTG_SEARCH_EXACT searches
 * run wholly by the plain kernel (more than 48 rotation steps, 40 for 16-bit frames, a
 * pattern radius above 120 px, or a pattern and window beyond the exact kernels' LDS).
 * grains: grains of exact searches searched by the plain kernel within the same stage:
 * rotated templates that differ from the base mask in more cells than the exact kernel
 * holds (large patterns at large reference rotations), window values above the declared
 * depth,
 * patterns above 23,170 pixels, near-tie sets of more than two candidates, or a best NCC
 * below 1/32. Synchronises. */
int ptg_get_exact_fallbacks(void* ctx, int64_t* searches, int64_t* grains);

#ifdef __cplusplus
}
#endif

#endif /* PTGPU_H */
