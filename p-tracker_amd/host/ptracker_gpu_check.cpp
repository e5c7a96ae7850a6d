// ptracker_gpu_check.cpp — drives ptgpu::ptracker_gpu through one frame pair.
//
// Input file (little endian): int32 dimx, dimy, bits, G, search_method, subpix_method,
// rescue_level, subpixel; 3 x search zone (5 int32 + 1 double, packed as
// ptg_search_zone); double NCC_min, NCC_min_super; uint16 I0[dimy][dimx], I1[dimy][dimx];
// per grain: int32 refx, refy, masked, npattern; double refrot, radius, dx, dy, drot;
// npattern x (int32 dx, int32 dy).
// Output file: per grain 9 doubles dx, dy, drot, upix, vpix, rot_inc, NCC, NCC_rescue,
// NCC_subpix; then int32 num_to_be_rescued and one objective value at X = (0.1, -0.1, 0.01)
// of grain 0.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ptracker_gpu.hpp"

template <class T>
static void rd(FILE* f, T* p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    std::exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t h[8];
  rd(f, h, 8);
  const int dimx = h[0], dimy = h[1], bits = h[2], G = h[3];
  ptgpu::ptracker_gpu pt(0);
  pt.search_method = h[4];
  pt.subpix_method = h[5];
  pt.rescue_level = h[6];
  pt.subpixel = h[7];
  rd(f, &pt.search_zone, 1);
  rd(f, &pt.search_zone_rescue, 1);
  rd(f, &pt.search_zone_super_rescue, 1);
  rd(f, &pt.NCC_min, 1);
  rd(f, &pt.NCC_min_super, 1);
  std::vector<uint16_t> I0((size_t)dimx * dimy), I1((size_t)dimx * dimy);
  rd(f, I0.data(), I0.size());
  rd(f, I1.data(), I1.size());
  pt.grain.resize(G);
  for (auto& g : pt.grain) {
    int32_t iv[4];
    double dv[5];
    rd(f, iv, 4);
    rd(f, dv, 5);
    g.refcoord_xpix = iv[0];
    g.refcoord_ypix = iv[1];
    g.masked = iv[2] != 0;
    g.refrot = dv[0];
    g.radius_pix = dv[1];
    g.dx = dv[2];
    g.dy = dv[3];
    g.drot = dv[4];
    g.pattern.resize(iv[3]);
    for (auto& p : g.pattern) {
      int32_t xy[2];
      rd(f, xy, 2);
      p.dx = xy[0];
      p.dy = xy[1];
    }
  }
  fclose(f);
  try {
    pt.set_image(0, I0.data(), dimx, dimy, bits);
    pt.set_image(1, I1.data(), dimx, dimy, bits);
    pt.upload_grains();
    pt.do_precomputations();
    pt.track_pair();
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    for (auto& g : pt.grain) {
      const double v[9] = {g.dx, g.dy, g.drot, g.upix, g.vpix, g.rot_inc, g.NCC, g.NCC_rescue, g.NCC_subpix};
      fwrite(v, sizeof(double), 9, o);
    }
    const int32_t nres = pt.num_to_be_rescued();
    fwrite(&nres, sizeof(nres), 1, o);
    std::vector<double> X{0.1, -0.1, 0.01};
    ptgpu::NCC_optimizer fn{&pt, 0};
    const double fx = fn(X);
    fwrite(&fx, sizeof(fx), 1, o);
    fclose(o);
  } catch (const std::exception& e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
