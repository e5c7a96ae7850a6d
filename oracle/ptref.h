/*
 * ptref.h — CPU restatement ("oracle") of p-tracker's per-grain tracking path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for libptgpu: tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it. The product path
 * never links or calls it.
 *
 * Every function follows the reference's arithmetic in the reference's order, in
 * IEEE double without FMA contraction (built with -ffp-contract=off), using glibc
 * cos/sin/sqrt/round as the reference does. Frames are row-major uint16
 * (element (x, y) at f[y * pitch + x]); the reference stores image[x][y].
 */
#ifndef PTREF_H
#define PTREF_H

#include <stdint.h>

#include "../include/ptgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rotation-step values of `for (drot = -rotmax; drot <= rotmax; drot += inc_rot)`,
 * src/p-tracker.cpp:684-686. Writes at most cap values, returns the step count, or -1
 * when inc_rot <= 0 (the reference loop would not terminate). */
int ptref_rotation_steps(double inc_rot, int32_t num_rot, double* out, int32_t cap);

/* Pattern builders, src/p-tracker.cpp:1369-1433: (dx, dy) pairs, dy outer, dx inner.
 * Return the pixel count; out_xy may be NULL to count only. */
int ptref_make_circ_pattern(int32_t radius_pix, int32_t* out_xy);
int ptref_make_rect_pattern(int32_t half_width, int32_t half_height, int32_t* out_xy);
int ptref_make_ring_pattern(int32_t radius_in, int32_t radius_out, int32_t* out_xy);

/* rotate_pixel_pattern, src/p-tracker.cpp:615-621, applied to n offsets. */
void ptref_rotate_pattern(int32_t n, const int32_t* xy, double c, double s, int32_t* out_xy);

/* do_precomputations for one grain, src/p-tracker.cpp:624-656: writes
 * pattern0_rotated to p0_xy and g->mean0, g->C0C0. */
void ptref_precompute(const uint16_t* I0, int64_t pitch, ptg_grain* g, int32_t n,
                      const int32_t* pattern_xy, int32_t* p0_xy);

/* follow_pattern_pixel (stage 0, src/p-tracker.cpp:659-743),
 * follow_pattern_rescue_pixel (stage 1, :745-881) and
 * follow_pattern_super_rescue_pixel (stage 2, :883-1016) for one grain.
 * scratch_xy holds n int pairs. Returns 1 when the grain is queued for the next
 * stage, else 0. */
int ptref_follow_pattern(const uint16_t* I0, const uint16_t* I1, int64_t pitch, ptg_grain* g,
                         int32_t n, const int32_t* pattern_xy, const int32_t* p0_xy,
                         const ptg_search_zone* zone, int stage, double NCC_min,
                         double NCC_min_super, int32_t* scratch_xy);

/* The same candidate loop, returning every candidate's NCC in the reference loop
 * order (rotation step, then u, then v) into ncc_out (steps * S values). */
int ptref_candidate_ncc(const uint16_t* I0, const uint16_t* I1, int64_t pitch,
                        const ptg_grain* g, int32_t n, const int32_t* pattern_xy,
                        const int32_t* p0_xy, const ptg_search_zone* zone, double* ncc_out,
                        int64_t cap);

/* image_interpCubic::getValue, src/interpol.cpp:42-86 (no bounds check). */
double ptref_interp_cubic(const uint16_t* im, int64_t pitch, double x, double y);
void ptref_interp_cubic_batch(const uint16_t* im, int64_t pitch, int32_t n, const double* xy,
                              double* out);

/* NCC_to_minimize_xyR, src/p-tracker.cpp:1018-1088. */
double ptref_ncc_objective(const uint16_t* I0, const uint16_t* I1, int64_t pitch,
                           const ptg_grain* g, int32_t n, const int32_t* p0_xy, const double* X,
                           double subpix_displacement_max, double overflow_stiffness);

/* Powell's conjugate-direction minimizer with Brent line searches, written from the
 * published methods (Powell 1964; Brent 1973, ch. 5). x (3 values) is updated in
 * place; returns the final objective value; *nevals counts objective calls. */
typedef double (*ptref_fn3)(const double* x, void* user);
double ptref_powell(ptref_fn3 f, void* user, double* x, const double* dx, double ftol,
                    int32_t max_iter, int64_t* nevals);

/* The FP64 minimizer the LM form is checked against: Powell as above, then restarted from its
 * result with the initial axis directions until a restart no longer lowers the objective
 * (at most max_restarts). A finished Powell run can stall on a degenerate direction set in
 * a flat valley; a restart resets the set and continues downhill within the same basin. */
double ptref_minimize_fp64(ptref_fn3 f, void* user, double* x, const double* dx, double ftol,
                           int32_t max_iter, int32_t max_restarts, int64_t* nevals);

/* follow_pattern_subpixel_xyR for one grain, src/p-tracker.cpp:1090-1128. With
 * prm->method == PTG_SUBPIX_LM the oracle runs ptref_minimize_fp64 instead of one Powell
 * run: the converged minimum of the objective in Powell's basin, which the LM form is
 * checked against. */
void ptref_subpixel(const uint16_t* I0, const uint16_t* I1, int64_t pitch, ptg_grain* g,
                    int32_t n, const int32_t* p0_xy, const ptg_subpix_params* prm,
                    int64_t* nevals);

/* One frame pair, src/p-tracker.cpp:472-581, OpenMP schedule(dynamic) over grains as
 * the reference. Precomputes every grain first (the reference does so when the
 * reference image changed, which is every pair by default). counts[0..1] receive
 * num_to_be_rescued and num_to_be_super_rescued; nevals (may be NULL) receives the
 * total subpixel objective evaluations. nthreads <= 0 keeps the OpenMP default. */
int ptref_track_pair(const uint16_t* I0, const uint16_t* I1, int64_t pitch, int32_t G,
                     ptg_grain* grains, const int32_t* pattern_xy, const int64_t* pattern_begin,
                     const ptg_track_params* prm, int32_t nthreads, int32_t* counts,
                     int64_t* nevals);

/* The reference swap, src/p-tracker.cpp:598-609, on the grain records. */
void ptref_change_reference(int32_t G, ptg_grain* grains);

/* undistor, src/distortion_correction.cpp:20-61. */
void ptref_undistor(const double* X, double xd, double yd, double* xu, double* yu);

/* precompute_paires, src/distortion_correction.cpp:62-85: pairs (i < j) of reference
 * pixels whose distance lies in [dmin, dmax], i outer, j inner; writes up to cap pairs
 * and returns the number found. */
int64_t ptref_precompute_pairs(int32_t G, const double* ref_xy, double dmin, double dmax, int32_t* pairs, int64_t cap);
/* The equiprojectivity residuals of the LEVMAR functor, src/distortion_correction.cpp:409-475
 * (the Powell functor :277-303 sums their squares): res[k * npairs + l], 0 where either
 * grain of pair l has NCC_subpix < NCC_min in image k. */
void ptref_equiproj_residuals(const double* X, int32_t G, const double* ref_xy, int32_t nimg, const double* disp,
                              const double* ncc, int32_t npairs, const int32_t* pairs, double NCC_min, double* res);
/* The residuals of the POWELL_ITER functor, src/distortion_correction.cpp:316-354: the first
 * positions A1, B1 are given (orig_xy, G x 2: the reference pixels at the first iteration,
 * their undistorted positions at the previous parameters afterwards) and only the tracked
 * positions ref + disp are undistorted. */
void ptref_equiproj_residuals_orig(const double* X, int32_t G, const double* orig_xy, const double* ref_xy,
                                   int32_t nimg, const double* disp, const double* ncc, int32_t npairs,
                                   const int32_t* pairs, double NCC_min, double* res);

/* Remap of a whole frame (the library's ptg_remap): for every output pixel (x, y),
 * Newton iteration on undistor() for the distorted point, then a bicubic sample of src
 * with taps clamped to the frame, rounded and clamped to [0, 2^bits - 1]. */
void ptref_remap(const uint16_t* src, int32_t dimx, int32_t dimy, int64_t pitch, int32_t bits,
                 const double* X, uint16_t* dst, int64_t dst_pitch);
/* The distorted point for one undistorted point (the remap's inverse). */
void ptref_distort_point(const double* X, double xu, double yu, double* xd, double* yd);

/* The differentiation form's joint objective (no reference counterpart): 1 - NCC of grain g
 * whose pattern pixels move rigidly by P = (Tx, Ty, theta) in the undistorted coordinates of
 * the radial lens lens = (K1, K2, xc, yc), sampled in I1 through the inverse lens. */
double ptref_joint_objective(const uint16_t* I0, const uint16_t* I1, int64_t pitch, const ptg_grain* g,
                             int32_t n, const int32_t* p0_xy, const double* P, const double* lens);
void ptref_joint_objective_batch(const uint16_t* I0, const uint16_t* I1, int64_t pitch, const ptg_grain* grains,
                                 const int32_t* p0_xy, const int64_t* pattern_begin, int32_t n,
                                 const int32_t* list, const double* P, const double* lens, double* f,
                                 int32_t nthreads);

#ifdef __cplusplus
}
#endif

#endif
