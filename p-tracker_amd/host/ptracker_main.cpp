// ptracker_main.cpp — the tracking program over libptgpu: `ptracker_gpu [options]
// [commands.txt]` reads a p-tracker command file (default commands.txt, as the
// reference's init, src/p-tracker.cpp:1170-1202) and runs its procedure.
//
// Options: --lm     subpixel by the LM form instead of Powell
//          --plain  integer search by the plain FP64 form instead of the exact MFMA form
//          --device N
//          --frames-ahead N  TIFF frames decoded at once ahead of the pairs (default 3)
//          --serial-frames   read, copy, track and write each image in turn, as the reference
// Procedures: particle_tracking (run once; the reference's run() runs it a second time
// for that procedure name, src/p-tracker.cpp:1802-1808) and correction_distortion (the
// equiprojectivity fit on the GPU by Levenberg-Marquardt in place of Powell / lmdif).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>

#include "ptracker_gpu.hpp"

int main(int argc, char** argv) {
  std::string cmd = "commands.txt";
  int device = 0;
  int search = PTG_SEARCH_EXACT, subpix = PTG_SUBPIX_POWELL;
  int ahead = 3;
  bool serial = false;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--lm")) subpix = PTG_SUBPIX_LM;
    else if (!std::strcmp(argv[i], "--serial-frames")) serial = true;
    else if (!std::strcmp(argv[i], "--frames-ahead") && i + 1 < argc) ahead = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--plain")) search = PTG_SEARCH_PLAIN;
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
    else cmd = argv[i];
  }
  try {
    ptgpu::tracker_settings s;
    std::vector<ptgpu::grain_type_2D> grains;
    ptgpu::read_data(cmd, s, grains);
    ptgpu::finish_settings(s);
    if (s.rotations == 0) std::cout << "Rotation are NOT TRACKED!" << std::endl;
    ptgpu::ptracker_gpu pt(device);
    pt.grain = grains;
    pt.search_method = search;
    pt.subpix_method = subpix;
    pt.frames_ahead = ahead;
    pt.serial_frames = serial;
    if (s.procedure == "correction_distortion") {
      const std::vector<double> X = pt.correction_distortion(s);
      std::cout << std::scientific << std::setprecision(15) << "K1_corrDistor " << X[2] << "\nK2_corrDistor " << X[3]
                << std::endl;
    } else if (s.procedure == "particle_tracking") {
      const int n = pt.particle_tracking(s);
      std::cout << "tracked " << pt.grain.size() << " grains over " << n << " images" << std::endl;
      const auto fb = pt.exact_fallbacks();
      std::cout << "plain-form fallbacks of the exact search: " << fb.first << " searches, " << fb.second
                << " grains" << std::endl;
    } else {
      std::cerr << "procedure " << s.procedure << " is not part of this program" << std::endl;
      return 1;
    }
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
