// ptracker_io_check.cpp — drives the host file side (no GPU) for the CPU tests.
//   parse <commands> <out>   settings and grains after read_data + finish_settings, one
//                            "key value" line each (grains: index, 13 columns, masked,
//                            pattern size and first / last offsets)
//   dic <in> <out> [ptracker_format 0|1] [simple 0|1]
//                            read_grains then save_grains (trailer names a, b, c, d)
//   tiff <in> <out>          the gray frame: "width height", the date-time line, then
//                            width x height uint16 (little endian) after a newline
//   grid xmin xmax ymin ymax nx ny alea <out>
//                            make_grid with the C library's default seed
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "ptracker_io.hpp"
#include "tiff_reader.hpp"

using namespace ptgpu;

static void zone(std::ostream& o, const char* n, const search_zone_type& z) {
  o << n << ' ' << z.left << ' ' << z.right << ' ' << z.up << ' ' << z.down << ' ' << z.num_rot << ' '
    << std::setprecision(17) << z.inc_rot << '\n';
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  try {
    if (mode == "parse" && argc == 4) {
      tracker_settings s;
      std::vector<grain_type_2D> g;
      std::ofstream o(argv[3]);
      read_data(argv[2], s, g, o);
      finish_settings(s);
      o << std::setprecision(17);
      o << "procedure " << s.procedure << "\nimage_name " << s.image_name << "\nrotations " << s.rotations
        << "\niref " << s.iref << "\nibeg " << s.ibeg << "\niend " << s.iend << "\niinc " << s.iinc << "\niraz "
        << s.iraz << "\nrescue_level " << s.rescue_level << "\nNCC_min " << s.NCC_min << "\nNCC_min_super "
        << s.NCC_min_super << "\nsubpixel " << s.subpixel << "\nsubpix_tol " << s.subpix_tol << "\ndirections "
        << s.initial_direction_dx << ' ' << s.initial_direction_dy << ' ' << s.initial_direction_dr
        << "\nsubpix_displacement_max " << s.subpix_displacement_max << "\noverflow_stiffness "
        << s.overflow_stiffness << "\nimage_interpolator " << s.image_interpolator << "\nwanted_num_threads "
        << s.wanted_num_threads << "\noptim_algo " << s.optim_algo << "\nuse_neighbour_list "
        << s.use_neighbour_list << "\n";
      zone(o, "search_zone", s.search_zone);
      zone(o, "search_zone_rescue", s.search_zone_rescue);
      zone(o, "search_zone_super_rescue", s.search_zone_super_rescue);
      o << "disto";
      for (double v : s.disto_parameters) o << ' ' << v;
      o << "\ngrains " << g.size() << '\n';
      for (size_t i = 0; i < g.size(); i++) {
        o << "grain " << i << ' ';
        o.flush();
        std::ostringstream row;
        write_dic_row(row, g[i]);
        std::string r = row.str();
        r.pop_back();
        o << r << ' ' << (g[i].masked ? 1 : 0) << ' ' << g[i].pattern.size();
        if (!g[i].pattern.empty())
          o << ' ' << g[i].pattern.front().dx << ' ' << g[i].pattern.front().dy << ' ' << g[i].pattern.back().dx
            << ' ' << g[i].pattern.back().dy;
        o << std::defaultfloat << std::setprecision(17) << '\n';
      }
    } else if (mode == "dic" && argc >= 4) {
      const bool pf = argc > 4 ? std::atoi(argv[4]) != 0 : true;
      const bool simple = argc > 5 ? std::atoi(argv[5]) != 0 : false;
      std::vector<grain_type_2D> g;
      read_grains(argv[2], g, pf);
      save_grains(argv[3], g, simple, "a", "b", "c", "d");
    } else if (mode == "tiff" && argc == 4) {
      const tiff_frame f = read_tiff(argv[2]);
      std::ofstream o(argv[3], std::ios::binary);
      o << f.width << ' ' << f.height << '\n' << f.dateTime << '\n';
      o.write(reinterpret_cast<const char*>(f.gray.data()), (std::streamsize)(f.gray.size() * 2));
    } else if (mode == "grid" && argc == 10) {
      std::vector<grain_type_2D> g;
      make_grid(g, std::atof(argv[2]), std::atof(argv[3]), std::atof(argv[4]), std::atof(argv[5]), std::atoi(argv[6]),
                std::atoi(argv[7]), std::atoi(argv[8]));
      std::ofstream o(argv[9]);
      for (const auto& x : g) o << x.refcoord_xpix << ' ' << x.refcoord_ypix << '\n';
    } else {
      return 2;
    }
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
