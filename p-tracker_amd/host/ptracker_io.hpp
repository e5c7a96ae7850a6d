// ptracker_io.hpp — the file side of p-tracker's tracking procedure, on the host (no GPU
// code): the command file with the reference's defaults, grain files in the simple and
// dic_out formats, the pattern builders, the point grid and the rectangular mask.
//
//   read_data            src/p-tracker.cpp:1487-1784 (token grammar and keywords)
//   tracker_settings     src/p-tracker.hpp:67-152, src/p-tracker.cpp:20-52 (defaults)
//   rotations 0          src/p-tracker.cpp:1191-1201 (applied by finish_settings)
//   read_grains          src/p-tracker.cpp:1213-1278
//   save_grains          src/p-tracker.cpp:1313-1367 (13 columns, scientific, precision
//                        15, widths 6 / 23, #_FROM_image trailer)
//   make_*_pattern       src/p-tracker.cpp:1369-1476
//   make_grid            src/p-tracker.cpp:1280-1311 (the C library's rand())
//   mask_rect            src/p-tracker.cpp:1478-1484
//
// Where the reference prints a message and calls std::exit, these functions throw
// std::runtime_error with the same message.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ptgpu {

// src/relative_coord_type.hpp:19-25
struct relative_coord_type {
  int dx{0};
  int dy{0};
};

// src/search_zone_type.hpp:20-27
struct search_zone_type {
  int left{0}, right{0}, up{0}, down{0}, num_rot{0};
  double inc_rot{1.0};
};

// src/grain_type_2D.hpp:23-71 (the neighbour list, dead in the reference, is left out)
struct grain_type_2D {
  int refcoord_xpix{0}, refcoord_ypix{0};
  double refrot{0.0}, radius_pix{0.0};
  double dx{0.0}, dy{0.0}, drot{0.0};
  double dx_prev{0.0}, dy_prev{0.0}, drot_prev{0.0};
  double upix{0.0}, vpix{0.0}, rot_inc{0.0};
  double NCC{0.0}, NCC_rescue{0.0}, NCC_subpix{0.0};
  double mean0{0.0}, C0C0{0.0};
  bool masked{false};
  int status{0};
  std::vector<relative_coord_type> pattern;
  void backup() { dx_prev = dx; dy_prev = dy; drot_prev = drot; }
  void reset() { NCC = NCC_rescue = NCC_subpix = 0.0; }
};

// optim_algo values, src/p-tracker.hpp:37-39
enum { OPTIM_POWELL = 0, OPTIM_POWELL_ITER = 1, OPTIM_LEVMAR = 2 };

// The command-file state of the reference's ptracker object with its defaults.
struct tracker_settings {
  std::string procedure{"particle_tracking"};
  int verbose_level{0};
  int rotations{1};
  int optim_algo{OPTIM_POWELL};
  int powell_num_iter{5};
  double targetRadiusPattern{-1.0};
  std::string image_name;  // printf style, e.g. "../ImageData/TEST%02d.tiff"
  std::string dic_name;    // read but unused by the reference (it writes dic_out_%d.txt)
  int RawImages{0};
  int DemosaicModel{0};
  int rescaleGrayLevels{0};
  int iref{0}, ibeg{0}, iend{0}, iinc{0}, iraz{0}, idelta{0};
  int rescue_level{2};
  double NCC_min{0.7};
  double NCC_min_super{0.5};
  int subpixel{1};
  double subpix_tol{1e-8};
  double initial_direction_dx{0.01}, initial_direction_dy{0.01}, initial_direction_dr{0.01};
  double subpix_displacement_max{3.0};
  double overflow_stiffness{0.0};
  std::string image_interpolator{"cubic"};
  int wanted_num_threads{4};
  int use_neighbour_list{1};
  int num_neighbour_max{40};
  double neighbour_dist_pix{250.0};
  int period_rebuild_neighbour_list{50};
  search_zone_type search_zone, search_zone_rescue, search_zone_super_rescue;
  double equiProj_dist_min{0.0}, equiProj_dist_max{1000.0};
  std::vector<int> image_numbers_corrDisto;
  std::vector<double> imposed_displ_x_pix, imposed_displ_y_pix;
  double disto_parameters[14]{-1.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, -1.0, 0.0, 0.0, 0.0, 0.0};
  double disto_parameters_perturb[14]{1e-6, 1e-6, 1e-6, 1e-6, 1e-6, 1e-6, 1e-6,
                                      1e-6, 1e-6, 1e-6, 1e-6, 1e-6, 1e-6, 1e-6};
};

// sprintf(out, pattern, num) as the reference names its image and dic_out files.
inline std::string numbered_name(const std::string& pattern, int num) {
  char buf[4096];
  std::snprintf(buf, sizeof(buf), pattern.c_str(), num);
  return std::string(buf);
}

// ---- patterns: offsets in row-major order (dy outer, dx inner) ---------------------------
inline void make_rect_pattern(grain_type_2D& g, int half_width_pix, int half_height_pix) {
  g.pattern.clear();
  for (int dy = -half_height_pix; dy <= half_height_pix; ++dy)
    for (int dx = -half_width_pix; dx <= half_width_pix; ++dx) g.pattern.push_back({dx, dy});
}

inline void make_circ_pattern(grain_type_2D& g, int radius_pix) {
  g.pattern.clear();
  for (int dy = -radius_pix; dy <= radius_pix; ++dy)
    for (int dx = -radius_pix; dx <= radius_pix; ++dx)
      if (std::sqrt((double)dx * (double)dx + (double)dy * (double)dy) <= (double)radius_pix)
        g.pattern.push_back({dx, dy});
}

inline void make_ring_pattern(grain_type_2D& g, int radius_IN_pix, int radius_OUT_pix) {
  g.pattern.clear();
  for (int dy = -radius_OUT_pix; dy <= radius_OUT_pix; ++dy)
    for (int dx = -radius_OUT_pix; dx <= radius_OUT_pix; ++dx) {
      const double dst = std::sqrt((double)dx * (double)dx + (double)dy * (double)dy);
      if (dst <= (double)radius_OUT_pix && dst >= (double)radius_IN_pix) g.pattern.push_back({dx, dy});
    }
}

// File: number of patterns (= number of grains), then per grain a count and that many
// "dx dy" pairs.
inline void make_custom_pattern(const std::string& name, std::vector<grain_type_2D>& grain) {
  std::ifstream f(name);
  if (!f) throw std::runtime_error("Cannot open file " + name);
  int num_patterns = 0;
  f >> num_patterns;
  if (num_patterns != (int)grain.size())
    throw std::runtime_error("The number of patterns is not equal to the number of grains! (num_patterns : " +
                             std::to_string(num_patterns) + ", num_grains : " + std::to_string(grain.size()) + ")");
  for (int i = 0; i < num_patterns; i++) {
    int n = 0;
    f >> n;
    grain[i].pattern.assign(n > 0 ? n : 0, relative_coord_type{});
    for (int k = 0; k < n; k++) f >> grain[i].pattern[k].dx >> grain[i].pattern[k].dy;
  }
}

// ---- grids and masks --------------------------------------------------------------------
inline void make_grid(std::vector<grain_type_2D>& grain, double xmin, double xmax, double ymin, double ymax, int nx,
                      int ny, int aleaMax) {
  const int dx = (int)std::floor((xmax - xmin) / (double)(nx - 1));
  const int dy = (int)std::floor((ymax - ymin) / (double)(ny - 1));
  grain.assign((size_t)nx * ny, grain_type_2D{});
  int i = 0;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      // angle in [0, 2 pi] and radius in [0, aleaMax] from the C library's generator
      const double angle = (double)std::rand() / (double)RAND_MAX * (2.0 * M_PI);
      const double alea = (double)std::rand() / (double)RAND_MAX * aleaMax;
      grain[i].refcoord_xpix = (int)(xmin + ix * dx + (int)std::round(alea * std::cos(angle)));
      grain[i].refcoord_ypix = (int)(ymin + iy * dy + (int)std::round(alea * std::sin(angle)));
      grain[i].refrot = 0.0;
      grain[i].radius_pix = 2.0;
      i++;
    }
}

inline void mask_rect(std::vector<grain_type_2D>& grain, int xmin, int xmax, int ymin, int ymax) {
  for (auto& g : grain)
    if (g.refcoord_xpix >= xmin && g.refcoord_xpix <= xmax && g.refcoord_ypix >= ymin && g.refcoord_ypix <= ymax)
      g.masked = true;
}

// ---- grain files -------------------------------------------------------------------------
// ptracker_format: the 13 dic_out columns per grain (a restart); otherwise the simple
// format "x y refrot radius" with real coordinates, whose fractional part becomes an
// initial displacement dx = upix (src/p-tracker.cpp:1265-1270).
inline void read_grains(const std::string& name, std::vector<grain_type_2D>& grain, bool ptracker_format) {
  std::ifstream f(name);
  if (!f) throw std::runtime_error("@ptracker::read_grains, cannot open file " + name);
  int num_grains = 0;
  f >> num_grains;
  grain.assign(num_grains > 0 ? num_grains : 0, grain_type_2D{});
  for (int i = 0; i < num_grains; i++) {
    grain_type_2D& g = grain[i];
    if (ptracker_format) {
      f >> g.refcoord_xpix >> g.refcoord_ypix >> g.refrot >> g.radius_pix >> g.dx >> g.dy >> g.drot >> g.upix >>
          g.vpix >> g.rot_inc >> g.NCC >> g.NCC_rescue >> g.NCC_subpix;
    } else {
      double x = 0.0, y = 0.0, r = 0.0;
      f >> x >> y >> g.refrot >> r;
      g.radius_pix = r;
      g.refcoord_xpix = (int)std::floor(x);
      g.dx = g.upix = x - std::floor(x);
      g.refcoord_ypix = (int)std::floor(y);
      g.dy = g.vpix = y - std::floor(y);
    }
  }
}

// The 13 columns of one grain as the reference writes them.
inline void write_dic_row(std::ostream& o, const grain_type_2D& g) {
  o << std::scientific << std::setprecision(15) << std::setw(6) << std::left << g.refcoord_xpix << ' '
    << std::setw(6) << std::left << g.refcoord_ypix << ' ' << std::setw(23) << std::right << g.refrot << ' '
    << std::setw(23) << std::right << g.radius_pix << ' ' << std::setw(23) << std::right << g.dx << ' '
    << std::setw(23) << std::right << g.dy << ' ' << std::setw(23) << std::right << g.drot << ' ' << std::setw(23)
    << std::right << g.upix << ' ' << std::setw(23) << std::right << g.vpix << ' ' << std::setw(23) << std::right
    << g.rot_inc << ' ' << std::setw(23) << std::right << g.NCC << ' ' << std::setw(23) << std::right
    << g.NCC_rescue << ' ' << std::setw(23) << std::right << g.NCC_subpix << '\n';
}

// dic_out file: the grain count, one row per grain, a blank line and the trailer naming
// the reference and current images with their TIFF date-times. simpleVersion writes
// "x y refrot radius" rows with the current positions (a start file).
inline void save_grains(const std::string& name, const std::vector<grain_type_2D>& grain, bool simpleVersion,
                        const std::string& from_image = "", const std::string& from_datetime = "",
                        const std::string& to_image = "", const std::string& to_datetime = "") {
  std::ofstream o(name);
  if (!o) throw std::runtime_error("Cannot open file " + name);
  if (simpleVersion) {
    o << grain.size() << std::endl;
    for (const auto& g : grain)
      o << g.refcoord_xpix + g.dx << ' ' << g.refcoord_ypix + g.dy << ' ' << g.refrot << ' ' << g.radius_pix << '\n';
    return;
  }
  o << grain.size() << std::endl;
  for (const auto& g : grain) write_dic_row(o, g);
  o << std::endl;
  o << "#_FROM_image: " << from_image << '\n';
  o << "#___dateTime: " << from_datetime << '\n';
  o << "#___TO_image: " << to_image << '\n';
  o << "#___dateTime: " << to_datetime << '\n';
  o << std::flush;
}

// ---- the command file --------------------------------------------------------------------
// Whitespace-separated tokens; a token starting with '!' or '#' skips the rest of its line.
// Grain files and patterns apply in file order, as in the reference (a pattern line acts on
// the grains read before it). Unknown tokens are reported on `log` and skipped.
inline void read_data(const std::string& name, tracker_settings& s, std::vector<grain_type_2D>& grain,
                      std::ostream& log = std::cout) {
  std::ifstream f(name);
  if (!f) throw std::runtime_error("Cannot open file " + name);
  std::string token;
  auto zone_field = [&](search_zone_type& z, const std::string& field) {
    if (field == "left") f >> z.left;
    else if (field == "right") f >> z.right;
    else if (field == "up") f >> z.up;
    else if (field == "down") f >> z.down;
    else if (field == "inc_rot") f >> z.inc_rot;
    else if (field == "num_rot") f >> z.num_rot;
    else return false;
    return true;
  };
  static const char* const kDisto[14] = {"xc_corrDistor",   "yc_corrDistor",       "K1_corrDistor",
                                         "K2_corrDistor",   "K3_corrDistor",       "P1_corrDistor",
                                         "P2_corrDistor",   "P3_corrDistor",       "xc_corrParallax",
                                         "yc_corrParallax", "angle1_corrParallax", "A1_corrParallax",
                                         "angle2_corrParallax", "A2_corrParallax"};
  while (f >> token) {
    bool known = true;
    if (token[0] == '!' || token[0] == '#') {
      std::getline(f, token);
    } else if (token == "image_name") {
      f >> s.image_name;
    } else if (token == "dic_name") {
      f >> s.dic_name;
    } else if (token == "RawImages") {
      f >> s.RawImages;
    } else if (token == "DemosaicModel") {
      f >> s.DemosaicModel;
    } else if (token == "rescaleGrayLevels") {
      f >> s.rescaleGrayLevels;
    } else if (token == "grains_to_follow" || token == "start_file") {
      std::string fn;
      f >> fn;
      read_grains(fn, grain, false);
    } else if (token == "restart_file") {
      std::string fn;
      f >> fn;
      read_grains(fn, grain, true);
    } else if (token == "make_grid") {
      double xmin, xmax, ymin, ymax;
      int nx, ny, alea;
      f >> xmin >> xmax >> ymin >> ymax >> nx >> ny >> alea;
      make_grid(grain, xmin, xmax, ymin, ymax, nx, ny, alea);
    } else if (token == "iref") {
      f >> s.iref;
    } else if (token == "ibeg") {
      f >> s.ibeg;
    } else if (token == "iend") {
      f >> s.iend;
    } else if (token == "iinc") {
      f >> s.iinc;
    } else if (token == "iraz") {
      f >> s.iraz;
    } else if (token == "idelta") {
      f >> s.idelta;
    } else if (token == "wanted_num_threads") {
      int n;
      f >> n;
      s.wanted_num_threads = std::abs(n);
    } else if (token == "procedure") {
      f >> s.procedure;
    } else if (token == "verbose_level") {
      f >> s.verbose_level;
    } else if (token == "rescue_level") {
      f >> s.rescue_level;
    } else if (token == "subpixel") {
      f >> s.subpixel;
    } else if (token == "rotations") {
      f >> s.rotations;
    } else if (token == "optim_algo") {
      std::string kw;
      f >> kw;
      if (kw == "POWELL") s.optim_algo = OPTIM_POWELL;
      else if (kw == "POWELL_ITER") s.optim_algo = OPTIM_POWELL_ITER;
      else if (kw == "LEVMAR") s.optim_algo = OPTIM_LEVMAR;
      else log << "WARNING, keyword " << kw << " is unknown for optim_algo\n";
    } else if (token == "powell_num_iter") {
      f >> s.powell_num_iter;
    } else if (token == "num_neighbour_max") {
      f >> s.num_neighbour_max;
    } else if (token == "neighbour_dist_pix") {
      f >> s.neighbour_dist_pix;
    } else if (token == "period_rebuild_neighbour_list") {
      f >> s.period_rebuild_neighbour_list;
    } else if (token == "image_interpolator") {
      std::string kw;
      f >> kw;
      if (kw == "linear" || kw == "cubic" || kw == "quintic") s.image_interpolator = kw;
    } else if (token == "subpix_tol") {
      f >> s.subpix_tol;
    } else if (token == "initial_direction_dx") {
      f >> s.initial_direction_dx;
    } else if (token == "initial_direction_dy") {
      f >> s.initial_direction_dy;
    } else if (token == "initial_direction_dr") {
      f >> s.initial_direction_dr;
    } else if (token == "overflow_stiffness") {
      f >> s.overflow_stiffness;
    } else if (token == "subpix_displacement_max") {
      f >> s.subpix_displacement_max;
    } else if (token.rfind("search_zone.", 0) == 0) {
      known = zone_field(s.search_zone, token.substr(12));
    } else if (token == "NCC_min") {
      f >> s.NCC_min;
    } else if (token.rfind("search_zone_rescue.", 0) == 0) {
      known = zone_field(s.search_zone_rescue, token.substr(19));
    } else if (token == "NCC_min_super") {
      f >> s.NCC_min_super;
    } else if (token.rfind("search_zone_super_rescue.", 0) == 0) {
      known = zone_field(s.search_zone_super_rescue, token.substr(25));
    } else if (token == "use_neighbour_list") {
      f >> s.use_neighbour_list;
    } else if (token == "targetRadiusPattern") {
      f >> s.targetRadiusPattern;
    } else if (token == "pattern") {
      std::string kind;
      f >> kind;
      auto selected = [&](const grain_type_2D& g) {
        return s.targetRadiusPattern < 0.0 || g.radius_pix == s.targetRadiusPattern;
      };
      if (kind == "circ") {
        int r;
        f >> r;
        for (auto& g : grain)
          if (selected(g)) make_circ_pattern(g, r);
      } else if (kind == "grain_diameter") {
        double reduction, min_radius;
        f >> reduction >> min_radius;
        for (auto& g : grain) {
          int r = (int)(g.radius_pix * reduction);
          if (r < min_radius) r = (int)min_radius;
          if (selected(g)) make_circ_pattern(g, r);
        }
      } else if (kind == "ring") {
        int rin, rout;
        f >> rin >> rout;
        for (auto& g : grain)
          if (selected(g)) make_ring_pattern(g, rin, rout);
      } else if (kind == "file") {
        std::string fn;
        f >> fn;
        make_custom_pattern(fn, grain);
      } else if (kind == "rect") {
        int hw, hh;
        f >> hw >> hh;
        for (auto& g : grain)
          if (selected(g)) make_rect_pattern(g, hw, hh);
      }
    } else if (token == "mask_rect") {
      int xmin, xmax, ymin, ymax;
      f >> xmin >> xmax >> ymin >> ymax;
      mask_rect(grain, xmin, xmax, ymin, ymax);
    } else if (token == "image_numbers_corrDisto") {
      size_t nb;
      f >> nb;
      s.image_numbers_corrDisto.assign(nb, 0);
      for (size_t i = 0; i < nb; i++) f >> s.image_numbers_corrDisto[i];
    } else if (token == "imposedDisplApprox_corrDisto") {
      const size_t n = s.image_numbers_corrDisto.size();
      s.imposed_displ_x_pix.assign(n, 0.0);
      s.imposed_displ_y_pix.assign(n, 0.0);
      for (size_t i = 1; i < n; i++) f >> s.imposed_displ_x_pix[i] >> s.imposed_displ_y_pix[i];
    } else if (token == "equiProj_dist_range") {
      f >> s.equiProj_dist_min >> s.equiProj_dist_max;
    } else {
      known = false;
      for (int k = 0; k < 14 && !known; k++) {
        if (token == kDisto[k]) {
          f >> s.disto_parameters[k];
          known = true;
        } else if (token == std::string("d") + kDisto[k]) {
          f >> s.disto_parameters_perturb[k];
          known = true;
        }
      }
    }
    if (!known) log << "Unknown token: " << token << "\n";
  }
}

// precompute_paires, src/distortion_correction.cpp:62-85: pairs (i < j) whose reference
// pixels lie between dmin and dmax apart, i outer, j inner.
inline std::vector<int32_t> precompute_pairs(const std::vector<grain_type_2D>& grain, double dmin, double dmax) {
  std::vector<int32_t> pairs;
  for (size_t i = 0; i < grain.size(); i++) {
    const double xi = grain[i].refcoord_xpix, yi = grain[i].refcoord_ypix;
    for (size_t j = i + 1; j < grain.size(); j++) {
      const double xij = grain[j].refcoord_xpix - xi, yij = grain[j].refcoord_ypix - yi;
      const double dij = std::sqrt(xij * xij + yij * yij);
      if (dij >= dmin && dij <= dmax) {
        pairs.push_back((int32_t)i);
        pairs.push_back((int32_t)j);
      }
    }
  }
  return pairs;
}

// What the reference's init does after reading the command file: without rotation
// tracking every zone searches one step (inc_rot 1, num_rot 0), src/p-tracker.cpp:1191-1201.
inline void finish_settings(tracker_settings& s) {
  if (s.rotations == 0) {
    for (search_zone_type* z : {&s.search_zone, &s.search_zone_rescue, &s.search_zone_super_rescue}) {
      z->inc_rot = 1.0;
      z->num_rot = 0;
    }
  }
}

}  // namespace ptgpu
