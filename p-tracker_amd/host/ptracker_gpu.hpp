// ptracker_gpu.hpp — C++ host side over libptgpu with the reference's names.
//
// The state and member functions a caller of p-tracker's per-grain path uses
// (src/p-tracker.hpp:66-237): grain[], the three search zones, the thresholds, the
// subpixel settings, do_precomputations(), the per-pair body of particle_tracking()
// (track_pair), NCC_to_minimize_xyR() and the reference swap. The GPU work goes through
// the C ABI of include/ptgpu.h; errors raise std::runtime_error with the library's
// message.
#pragma once

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <fstream>
#include <future>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/ptgpu.h"
#include "powell_minimize.hpp"
#include "ptracker_io.hpp"
#include "tiff_reader.hpp"

namespace ptgpu {

static_assert(sizeof(search_zone_type) == sizeof(ptg_search_zone), "search zone layouts differ");

// Frames of a sequence decoded ahead of the tracking loop: up to `ahead` TIFF files decode at
// once on their own threads, each into a pinned buffer of a pool (ptg_host_alloc) from which
// ptg_set_image_async copies it; the loop hands a buffer back once its copy is done. The
// reader threads make no library call (the context belongs to the loop's thread).
class frame_reader {
 public:
  struct frame {
    const uint16_t* px = nullptr;
    int width = 0, height = 0;
    std::string dateTime;
    int buf = -1;
  };

  frame_reader(void* ctx, std::vector<std::string> names, int width, int height, int ahead)
      : ctx_(ctx), names_(std::move(names)), width_(width), height_(height) {
    ahead = std::max(1, ahead);
    const size_t bytes = sizeof(uint16_t) * (size_t)width * height;
    // every decode in flight holds a buffer, and so do the frames the loop has taken whose
    // copies may still run (the next frame's and the current one's)
    for (int b = 0; b < ahead + 2; b++) {
      void* p = nullptr;
      if (ptg_host_alloc(ctx_, bytes, &p) != PTG_OK) {
        free_buffers();
        throw std::runtime_error(std::string("ptg_host_alloc: ") + ptg_last_error(ctx_));
      }
      bufs_.push_back(static_cast<uint16_t*>(p));
      free_.push_back(b);
    }
    for (int k = 0; k < ahead; k++) launch();
  }
  ~frame_reader() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      closing_ = true;
    }
    cv_.notify_all();
    for (auto& f : pending_)
      if (f.valid()) f.wait();
    free_buffers();
  }
  frame_reader(const frame_reader&) = delete;
  frame_reader& operator=(const frame_reader&) = delete;

  // The next frame in order (waits for its decode); throws what the decode threw.
  frame next() {
    if (pending_.empty()) throw std::runtime_error("no frame left to read");
    frame f = pending_.front().get();
    pending_.pop_front();
    launch();
    return f;
  }
  // Hands buffer b back (its device copy is done).
  void release(int b) {
    if (b < 0) return;
    {
      std::lock_guard<std::mutex> lock(mu_);
      free_.push_back(b);
    }
    cv_.notify_one();
  }

 private:
  void launch() {
    if (next_ >= names_.size()) return;
    const std::string name = names_[next_++];
    pending_.push_back(std::async(std::launch::async, [this, name]() {
      tiff_frame t = read_tiff(name);
      if (t.width != width_ || t.height != height_)
        throw std::runtime_error("image " + name + " differs in size from the reference");
      int b = -1;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [this]() { return closing_ || !free_.empty(); });
        if (closing_) throw std::runtime_error("frame reader closed");
        b = free_.back();
        free_.pop_back();
      }
      std::memcpy(bufs_[b], t.gray.data(), sizeof(uint16_t) * t.gray.size());
      frame f;
      f.px = bufs_[b];
      f.width = t.width;
      f.height = t.height;
      f.dateTime = t.dateTime;
      f.buf = b;
      return f;
    }));
  }
  void free_buffers() {
    for (uint16_t* p : bufs_) (void)ptg_host_free(ctx_, p);
    bufs_.clear();
  }

  void* ctx_;
  std::vector<std::string> names_;
  int width_, height_;
  size_t next_ = 0;
  std::vector<uint16_t*> bufs_;
  std::deque<std::future<frame>> pending_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<int> free_;
  bool closing_ = false;
};

class ptracker_gpu {
 public:
  std::vector<grain_type_2D> grain;
  search_zone_type search_zone, search_zone_rescue, search_zone_super_rescue;
  int rescue_level{2};  // src/p-tracker.hpp:109-127 defaults
  double NCC_min{0.7};
  double NCC_min_super{0.5};
  int subpixel{1};
  double subpix_tol{1e-8};
  double initial_direction_dx{0.01}, initial_direction_dy{0.01}, initial_direction_dr{0.01};
  double subpix_displacement_max{3.0};
  double overflow_stiffness{0.0};
  int im_index_ref{0}, im_index_current{1};
  int search_method{PTG_SEARCH_EXACT};
  int subpix_method{PTG_SUBPIX_POWELL};
  // particle_tracking: TIFF decodes running ahead of the pairs, or the reference's serial order
  int frames_ahead{3};
  bool serial_frames{false};

  explicit ptracker_gpu(int device = 0) {
    if (ptg_create(&ctx_, device) != PTG_OK) throw std::runtime_error("ptg_create failed");
  }
  ~ptracker_gpu() { ptg_destroy(ctx_); }
  ptracker_gpu(const ptracker_gpu&) = delete;
  ptracker_gpu& operator=(const ptracker_gpu&) = delete;

  // image[index] from a row-major frame (element (x, y) at frame[y * dimx + x])
  void set_image(int index, const uint16_t* frame, int dimx, int dimy, int bits) {
    check(ptg_set_image(ctx_, index, frame, dimx, dimy, dimx, bits), "ptg_set_image");
  }

  // grain[] and patterns to the device; call after changing patterns or the grain count
  void upload_grains() {
    std::vector<int32_t> pxy;
    std::vector<int64_t> pbeg{0};
    for (auto& g : grain) {
      for (auto& p : g.pattern) {
        pxy.push_back(p.dx);
        pxy.push_back(p.dy);
      }
      pbeg.push_back((int64_t)pxy.size() / 2);
    }
    std::vector<ptg_grain> rec = records();
    check(ptg_set_grains(ctx_, (int32_t)grain.size(), rec.data(), pxy.data(), pbeg.data()), "ptg_set_grains");
  }

  void download_grains() {
    std::vector<ptg_grain> rec(grain.size());
    check(ptg_get_grains(ctx_, rec.data()), "ptg_get_grains");
    for (size_t g = 0; g < grain.size(); g++) from_record(rec[g], grain[g]);
  }

  // do_precomputations, src/p-tracker.cpp:624-656
  void do_precomputations() { check(ptg_precompute(ctx_), "ptg_precompute"); }

  // the per-pair body of particle_tracking, src/p-tracker.cpp:460-581; grain[] holds the
  // 13 columns afterwards
  void track_pair() {
    check(ptg_set_image_indices(ctx_, im_index_ref, im_index_current), "ptg_set_image_indices");
    ptg_track_params tp{};
    std::memcpy(&tp.search_zone, &search_zone, sizeof(search_zone));
    std::memcpy(&tp.search_zone_rescue, &search_zone_rescue, sizeof(search_zone));
    std::memcpy(&tp.search_zone_super_rescue, &search_zone_super_rescue, sizeof(search_zone));
    tp.NCC_min = NCC_min;
    tp.NCC_min_super = NCC_min_super;
    tp.rescue_level = rescue_level;
    tp.subpixel = subpixel;
    tp.search_method = search_method;
    tp.subpix = subpix_params();
    check(ptg_track_pair(ctx_, &tp), "ptg_track_pair");
    download_grains();
  }

  int num_to_be_rescued() {
    int32_t a = 0, b = 0;
    check(ptg_get_queue_counts(ctx_, &a, &b), "ptg_get_queue_counts");
    return a;
  }

  // NCC_to_minimize_xyR, src/p-tracker.cpp:1018-1088, for an explicit grain
  double NCC_to_minimize_xyR(int igrain, std::vector<double>& X) {
    double f = 0.0;
    check(ptg_ncc_objective(ctx_, igrain, X.data(), subpix_displacement_max, overflow_stiffness, &f),
          "ptg_ncc_objective");
    return f;
  }

  // the reference swap, src/p-tracker.cpp:587-611 (swaps the image indices too)
  void change_reference() {
    check(ptg_change_reference(ctx_), "ptg_change_reference");
    std::swap(im_index_ref, im_index_current);
    download_grains();
  }

  void* context() { return ctx_; }

  // The command-file state the per-pair body reads.
  void apply_settings(const tracker_settings& s) {
    search_zone = s.search_zone;
    search_zone_rescue = s.search_zone_rescue;
    search_zone_super_rescue = s.search_zone_super_rescue;
    rescue_level = s.rescue_level;
    NCC_min = s.NCC_min;
    NCC_min_super = s.NCC_min_super;
    subpixel = s.subpixel;
    subpix_tol = s.subpix_tol;
    initial_direction_dx = s.initial_direction_dx;
    initial_direction_dy = s.initial_direction_dy;
    initial_direction_dr = s.initial_direction_dr;
    subpix_displacement_max = s.subpix_displacement_max;
    overflow_stiffness = s.overflow_stiffness;
  }

  // grain[] without the pattern lists (what save_grains writes)
  std::vector<grain_type_2D> scalar_copy() {
    std::vector<std::vector<relative_coord_type>> pats(grain.size());
    for (size_t g = 0; g < grain.size(); g++) pats[g].swap(grain[g].pattern);
    std::vector<grain_type_2D> out = grain;
    for (size_t g = 0; g < grain.size(); g++) pats[g].swap(grain[g].pattern);
    return out;
  }

  // particle_tracking, src/p-tracker.cpp:423-611: images ibeg .. iend (step iinc) against
  // the reference image iref, frames read from the TIFF files named by image_name, one
  // dic_out_<num>.txt per image (in out_dir), and the reference swapped once
  // num_image - iref >= iraz. Frames hold the reference's 16-bit gray levels. Returns the
  // number of images tracked.
  //
  // The frames are read ahead (frames_ahead TIFF decodes at once, on their own threads), each
  // next frame is copied into the image slot the running pair does not use while that pair
  // runs (ptg_set_image_async from pinned memory), and each dic_out file is written while the
  // next pair runs; serial_frames runs the reference's order instead (read, copy, track, write),
  // with the same files.
  int particle_tracking(const tracker_settings& s, const std::string& out_dir = "") {
    if (s.RawImages) throw std::runtime_error("RAW images are not read (libraw is not part of this build)");
    if (s.image_interpolator != "cubic")
      throw std::runtime_error("only the cubic image interpolator is implemented (got " + s.image_interpolator + ")");
    if (s.iinc <= 0 && s.iend >= s.ibeg) throw std::runtime_error("iinc must be positive");
    apply_settings(s);
    if (serial_frames) return particle_tracking_serial(s, out_dir);
    int iref = s.iref;
    const tiff_frame f0 = read_tiff(numbered_name(s.image_name, iref));
    std::vector<int> nums;
    std::vector<std::string> names;
    for (int num = s.ibeg; num <= s.iend; num += s.iinc) {
      nums.push_back(num);
      names.push_back(numbered_name(s.image_name, num));
    }
    // three image slots: the reference frame, the current one and the next one
    int slot_ref = 0, slot_cur = 1, slot_next = 2;
    std::string datetime[3];
    set_image(slot_ref, f0.gray.data(), f0.width, f0.height, 16);
    datetime[slot_ref] = f0.dateTime;
    upload_grains();
    if (nums.empty()) return 0;
    frame_reader reader(ctx_, names, f0.width, f0.height, frames_ahead);
    auto copy_in = [&](int slot, const frame_reader::frame& f) {
      check(ptg_set_image_async(ctx_, slot, f.px, f.width, f.height, f.width, 16), "ptg_set_image_async");
      datetime[slot] = f.dateTime;
    };
    frame_reader::frame cur = reader.next();
    copy_in(slot_cur, cur);
    std::future<void> writing;
    int n = 0;
    for (size_t i = 0; i < nums.size(); i++) {
      const int num = nums[i];
      // the next frame goes in while this pair runs
      frame_reader::frame next;
      if (i + 1 < nums.size()) {
        next = reader.next();
        copy_in(slot_next, next);
      }
      im_index_ref = slot_ref;
      im_index_current = slot_cur;
      track_pair();
      // the current frame's copy is done (the pair waited for it): its buffer goes back
      check(ptg_wait_image(ctx_, slot_cur), "ptg_wait_image");
      reader.release(cur.buf);
      // dic_out of this pair, written while the next pair runs (from the scalar columns)
      if (writing.valid()) writing.get();
      writing = std::async(std::launch::async,
                           [snap = scalar_copy(), path = out_dir + numbered_name("dic_out_%d.txt", num),
                            from = numbered_name(s.image_name, iref), from_dt = datetime[slot_ref],
                            to = names[i], to_dt = datetime[slot_cur]]() {
                             save_grains(path, snap, false, from, from_dt, to, to_dt);
                           });
      n++;
      const int old_ref = slot_ref, old_cur = slot_cur;
      if (num - iref >= s.iraz) {
        change_reference();  // the current frame becomes the reference
        iref = num;
        slot_ref = old_cur;
        slot_cur = slot_next;
        slot_next = old_ref;
      } else {
        slot_cur = slot_next;
        slot_next = old_cur;
      }
      cur = next;
    }
    if (writing.valid()) writing.get();
    return n;
  }

  // The reference's order: each frame read and copied before its pair, each dic_out file
  // written after it.
  int particle_tracking_serial(const tracker_settings& s, const std::string& out_dir) {
    int iref = s.iref;
    std::string datetime[kFrames];
    const tiff_frame f0 = read_tiff(numbered_name(s.image_name, iref));
    set_image(im_index_ref, f0.gray.data(), f0.width, f0.height, 16);
    datetime[im_index_ref] = f0.dateTime;
    upload_grains();
    int n = 0;
    for (int num = s.ibeg; num <= s.iend; num += s.iinc) {
      const tiff_frame f = read_tiff(numbered_name(s.image_name, num));
      if (f.width != f0.width || f.height != f0.height)
        throw std::runtime_error("image " + numbered_name(s.image_name, num) + " differs in size from the reference");
      set_image(im_index_current, f.gray.data(), f.width, f.height, 16);
      datetime[im_index_current] = f.dateTime;
      track_pair();
      save_grains(out_dir + numbered_name("dic_out_%d.txt", num), grain, false, numbered_name(s.image_name, iref),
                  datetime[im_index_ref], numbered_name(s.image_name, num), datetime[im_index_current]);
      n++;
      if (num - iref >= s.iraz) {
        change_reference();
        iref = num;
      }
    }
    return n;
  }

  // correction_distortion, src/distortion_correction.cpp:135-608: the images
  // image_numbers_corrDisto[1..] tracked from image_numbers_corrDisto[0] (every grain
  // starting at the imposed displacement of the image; search and subpixel, no rescue),
  // the grain pairs between equiProj_dist_min and equiProj_dist_max apart, then the
  // equiprojectivity criterion (evaluated on the GPU) minimised over the parameters whose
  // perturbation (d*_corrDistor) is non-zero, by optim_algo: POWELL (Powell's method from
  // those perturbations, ftol 1e-12, restarted until a run no longer improves),
  // POWELL_ITER (powell_num_iter such minimisations with the first positions re-centred
  // between them) or LEVMAR (Levenberg-Marquardt on the analytic Jacobian). Writes dic_out_<num>.txt per image, corrDisto_log.txt (the reference's log
  // lines), errors.txt, distobox.txt and, for POWELL_ITER, convergence.txt into out_dir;
  // returns the 14 parameters.
  std::vector<double> correction_distortion(tracker_settings& s, const std::string& out_dir = "") {
    const std::vector<int>& im = s.image_numbers_corrDisto;
    if (im.size() < 2) throw std::runtime_error("image_numbers_corrDisto needs at least two images");
    if (s.RawImages) throw std::runtime_error("RAW images are not read (libraw is not part of this build)");
    apply_settings(s);
    const tiff_frame f0 = read_tiff(numbered_name(s.image_name, im[0]));
    set_image(im_index_ref, f0.gray.data(), f0.width, f0.height, 16);
    const int G = (int)grain.size();
    const int nimg = (int)im.size() - 1;
    std::vector<double> ref(2 * (size_t)G), disp(2 * (size_t)nimg * G), ncc((size_t)nimg * G);
    for (int g = 0; g < G; g++) {
      ref[2 * g] = grain[g].refcoord_xpix;
      ref[2 * g + 1] = grain[g].refcoord_ypix;
    }
    const int rescue = rescue_level;
    rescue_level = 0;  // no rescue is attempted (:193)
    for (int k = 1; k <= nimg; k++) {
      const tiff_frame f = read_tiff(numbered_name(s.image_name, im[k]));
      set_image(im_index_current, f.gray.data(), f.width, f.height, 16);
      for (auto& g : grain) {
        g.dx = k < (int)s.imposed_displ_x_pix.size() ? s.imposed_displ_x_pix[k] : 0.0;
        g.dy = k < (int)s.imposed_displ_y_pix.size() ? s.imposed_displ_y_pix[k] : 0.0;
        g.drot = 0.0;
        g.reset();
      }
      upload_grains();
      track_pair();
      for (int g = 0; g < G; g++) {
        disp[2 * ((size_t)(k - 1) * G + g)] = grain[g].dx;
        disp[2 * ((size_t)(k - 1) * G + g) + 1] = grain[g].dy;
        ncc[(size_t)(k - 1) * G + g] = grain[g].NCC_subpix;
      }
      save_grains(out_dir + numbered_name("dic_out_%d.txt", im[k]), grain, false, numbered_name(s.image_name, im[0]),
                  f0.dateTime, numbered_name(s.image_name, im[k]), f.dateTime);
    }
    rescue_level = rescue;
    // centres left negative default to the image centre (:230-233)
    std::vector<double> X(s.disto_parameters, s.disto_parameters + 14);
    if (X[0] < 0.0) X[0] = 0.5 * f0.width;
    if (X[1] < 0.0) X[1] = 0.5 * f0.height;
    if (X[8] < 0.0) X[8] = 0.5 * f0.width;
    if (X[9] < 0.0) X[9] = 0.5 * f0.height;
    const std::vector<int32_t> pairs = precompute_pairs(grain, s.equiProj_dist_min, s.equiProj_dist_max);
    check(ptg_equiproj_set_data(ctx_, G, ref.data(), nimg, disp.data(), ncc.data(), (int32_t)(pairs.size() / 2),
                                pairs.data(), s.NCC_min),
          "ptg_equiproj_set_data");
    int32_t free14[14];
    for (int q = 0; q < 14; q++) free14[q] = s.disto_parameters_perturb[q] != 0.0 ? 1 : 0;
    double f_ini = 0.0, f_end = 0.0;
    // the criterion (sum of squared residuals) on the GPU, and Powell's initial directions:
    // the perturbations of the parameters it moves (the reference's minimize(X, perturb))
    auto crit = [&](const std::vector<double>& P) {
      double f = 0.0;
      check(ptg_equiproj_eval(ctx_, P.data(), &f, nullptr, nullptr, nullptr, nullptr), "ptg_equiproj_eval");
      return f;
    };
    std::vector<std::vector<double>> dirs;
    for (int q = 0; q < 14; q++)
      if (s.disto_parameters_perturb[q] != 0.0) {
        dirs.emplace_back(14, 0.0);
        dirs.back()[q] = s.disto_parameters_perturb[q];
      }
    if (s.optim_algo == OPTIM_LEVMAR) {
      // Levenberg-Marquardt on the analytic Jacobian in place of lmdif (:401-475)
      int32_t iters = 0;
      check(ptg_calibrate_distortion(ctx_, X.data(), free14, 200, &f_ini, &f_end, &iters),
            "ptg_calibrate_distortion");
    } else if (s.optim_algo == OPTIM_POWELL) {
      // :241-300, Powell's method restarted until a run no longer improves (one run can stop
      // short of the minimum in the K1-K2 valley)
      f_ini = crit(X);
      powell_minimize_restarted(crit, X, dirs, 1e-12);
      f_end = crit(X);
    } else {
      // POWELL_ITER, :302-400: the first positions start at the reference pixels and, after
      // each Powell run, become their undistor() images at its result; the tracked
      // displacements move by the same correction; one line of convergence.txt per run
      std::vector<double> orig = ref, und(ref.size());
      const std::vector<double> disp0 = disp;
      check(ptg_equiproj_set_origins(ctx_, orig.data()), "ptg_equiproj_set_origins");
      f_ini = crit(X);
      std::ofstream conv(out_dir + "convergence.txt");
      conv << std::scientific << std::setprecision(15);
      for (int iter = 1; iter <= s.powell_num_iter; iter++) {
        powell_minimize_restarted(crit, X, dirs, 1e-12);
        const double idx = crit(X) / nimg;
        std::cout << std::scientific << std::setprecision(15) << "iteration " << iter << ", index = " << idx
                  << std::endl;
        conv << iter << ' ' << idx;
        for (int q = 0; q < 14; q++) conv << ' ' << X[q];
        conv << ' ' << std::endl;
        check(ptg_undistort_points(ctx_, X.data(), G, ref.data(), und.data()), "ptg_undistort_points");
        for (int g = 0; g < G; g++)
          for (int k = 0; k < nimg; k++) {
            const size_t o = 2 * ((size_t)k * G + g);
            disp[o] = disp0[o] + (und[2 * g] - ref[2 * g]);
            disp[o + 1] = disp0[o + 1] + (und[2 * g + 1] - ref[2 * g + 1]);
          }
        orig = und;
        check(ptg_equiproj_set_data(ctx_, G, ref.data(), nimg, disp.data(), ncc.data(), (int32_t)(pairs.size() / 2),
                                    pairs.data(), s.NCC_min),
              "ptg_equiproj_set_data");
        check(ptg_equiproj_set_origins(ctx_, orig.data()), "ptg_equiproj_set_origins");
      }
      f_end = crit(X);
    }
    // the reference reports the criterion divided by the number of image pairs (:234-236)
    const double index_ini = f_ini / nimg, index_end = f_end / nimg;
    std::ofstream log(out_dir + "corrDisto_log.txt");
    log << std::scientific << std::setprecision(15);
    log << "Initial distorsion index " << index_ini << '\n';
    log << "Final distorsion index " << index_end << '\n';
    log << "Index gain (>1 is good) " << index_ini / index_end << '\n';
    static const char* const names[14] = {"xc_corrDistor ", "yc_corrDistor ", "K1_corrDistor ", "K2_corrDistor ",
                                          "K3_corrDistor ", "P1_corrDistor ", "P2_corrDistor ", "P3_corrDistor ",
                                          "xc_corrParallax     ", "yc_corrParallax     ", "angle1_corrParallax ",
                                          "A1_corrParallax     ", "angle2_corrParallax ", "A2_corrParallax     "};
    for (int q = 0; q < 14; q++) log << names[q] << X[q] << (q == 13 ? "" : "\n");
    log << std::endl;
    write_correction_field(X, f0.width, f0.height, out_dir, log);
    for (int q = 0; q < 14; q++) s.disto_parameters[q] = X[q];
    return X;
  }

  // The correction field of the fitted parameters, :515-608: errors.txt, undistor(p) - p on
  // a grid of steps dimx / 50, dimy / 50 (at least 1) with a blank line after each column,
  // the largest components and length in the log, and distobox.txt along the frame's edge.
  void write_correction_field(const std::vector<double>& X, int dimx, int dimy, const std::string& out_dir,
                              std::ostream& log) {
    const double dxd = std::max(1.0, (double)dimx / 50.0), dyd = std::max(1.0, (double)dimy / 50.0);
    auto undistort = [&](const std::vector<double>& pts) {
      std::vector<double> u(pts.size());
      if (!pts.empty())
        check(ptg_undistort_points(ctx_, X.data(), (int32_t)(pts.size() / 2), pts.data(), u.data()),
              "ptg_undistort_points");
      return u;
    };
    std::vector<double> grid;
    std::vector<size_t> col_end;
    for (double xd = 0.0; xd < dimx; xd += dxd) {
      for (double yd = 0.0; yd < dimy; yd += dyd) {
        grid.push_back(xd);
        grid.push_back(yd);
      }
      col_end.push_back(grid.size() / 2);
    }
    const std::vector<double> gu = undistort(grid);
    std::ofstream ef(out_dir + "errors.txt");
    double xmax = -1e20, ymax = -1e20, emax = -1e20;
    size_t c = 0;
    for (size_t i = 0; i < grid.size() / 2; i++) {
      const double xe = gu[2 * i] - grid[2 * i], ye = gu[2 * i + 1] - grid[2 * i + 1];
      const double e = sqrt(xe * xe + ye * ye);
      xmax = std::max(xmax, std::fabs(xe));
      ymax = std::max(ymax, std::fabs(ye));
      emax = std::max(emax, e);
      ef << grid[2 * i] << ' ' << grid[2 * i + 1] << ' ' << xe << ' ' << ye << std::endl;
      if (i + 1 == col_end[c]) {
        ef << std::endl;
        c++;
      }
    }
    log << "max |xu - xd| " << xmax << '\n';
    log << "max |yu - yd| " << ymax << '\n';
    log << "errormax " << emax << std::endl;
    std::cout << std::scientific << std::setprecision(15) << "max |xu - xd| " << xmax << '\n'
              << "max |yu - yd| " << ymax << '\n'
              << "errormax      " << emax << std::endl;
    // the edge: top left to right, right side down, bottom right to left, left side up, and
    // the first corner again
    std::vector<double> box;
    auto add = [&](double x, double y) {
      box.push_back(x);
      box.push_back(y);
    };
    for (double xd = 0; xd < dimx; xd += dxd) add(xd, 0.0);
    for (double yd = 0; yd < dimy; yd += dyd) add(dimx - 1, yd);
    for (double xd = dimx - 1; xd >= 0; xd -= dxd) add(xd, dimy - 1);
    for (double yd = dimy - 1; yd >= 0; yd -= dyd) add(0.0, yd);
    add(0.0, 0.0);
    const std::vector<double> bu = undistort(box);
    std::ofstream bf(out_dir + "distobox.txt");
    for (size_t i = 0; i < box.size() / 2; i++)
      bf << box[2 * i] << " " << box[2 * i + 1] << " " << bu[2 * i] - box[2 * i] << " " << bu[2 * i + 1] - box[2 * i + 1]
         << std::endl;
  }

  // Searches (whole stages) and grains the exact form has handed to the plain FP64 form
  // since the tracker was created (the frames' depth, zone or pattern beyond its limits).
  std::pair<int64_t, int64_t> exact_fallbacks() {
    int64_t searches = 0, grains = 0;
    check(ptg_get_exact_fallbacks(ctx_, &searches, &grains), "ptg_get_exact_fallbacks");
    return {searches, grains};
  }

 private:
  static constexpr int kFrames = 2;  // image slots 0 and 1, as the reference's image[2]
  void* ctx_ = nullptr;

  void check(int rc, const char* what) {
    if (rc != PTG_OK) throw std::runtime_error(std::string(what) + ": " + ptg_last_error(ctx_));
  }

  ptg_subpix_params subpix_params() const {
    ptg_subpix_params p{};
    p.subpix_tol = subpix_tol;
    p.initial_direction[0] = initial_direction_dx;
    p.initial_direction[1] = initial_direction_dy;
    p.initial_direction[2] = initial_direction_dr;
    p.subpix_displacement_max = subpix_displacement_max;
    p.overflow_stiffness = overflow_stiffness;
    p.method = subpix_method;
    p.max_iter = 0;
    return p;
  }

  std::vector<ptg_grain> records() const {
    std::vector<ptg_grain> rec(grain.size());
    for (size_t i = 0; i < grain.size(); i++) {
      const grain_type_2D& g = grain[i];
      ptg_grain& r = rec[i];
      r = ptg_grain{};
      r.refcoord_xpix = g.refcoord_xpix;
      r.refcoord_ypix = g.refcoord_ypix;
      r.refrot = g.refrot;
      r.radius_pix = g.radius_pix;
      r.dx = g.dx;
      r.dy = g.dy;
      r.drot = g.drot;
      r.dx_prev = g.dx_prev;
      r.dy_prev = g.dy_prev;
      r.drot_prev = g.drot_prev;
      r.upix = g.upix;
      r.vpix = g.vpix;
      r.rot_inc = g.rot_inc;
      r.NCC = g.NCC;
      r.NCC_rescue = g.NCC_rescue;
      r.NCC_subpix = g.NCC_subpix;
      r.masked = g.masked ? 1 : 0;
    }
    return rec;
  }

  static void from_record(const ptg_grain& r, grain_type_2D& g) {
    g.refcoord_xpix = r.refcoord_xpix;
    g.refcoord_ypix = r.refcoord_ypix;
    g.refrot = r.refrot;
    g.dx = r.dx;
    g.dy = r.dy;
    g.drot = r.drot;
    g.dx_prev = r.dx_prev;
    g.dy_prev = r.dy_prev;
    g.drot_prev = r.drot_prev;
    g.upix = r.upix;
    g.vpix = r.vpix;
    g.rot_inc = r.rot_inc;
    g.NCC = r.NCC;
    g.NCC_rescue = r.NCC_rescue;
    g.NCC_subpix = r.NCC_subpix;
    g.mean0 = r.mean0;
    g.C0C0 = r.C0C0;
    g.status = r.status;
  }
};

// The NCC_optimizer functor (src/NCC_optimizer_functor.hpp:22-25) with an explicit grain.
struct NCC_optimizer {
  ptracker_gpu* ptrackerObject;
  int igrain;
  double operator()(std::vector<double>& X) { return ptrackerObject->NCC_to_minimize_xyR(igrain, X); }
};

}  // namespace ptgpu
