// ref_interp_driver.cpp — TEST INFRASTRUCTURE ONLY (oracle/_ref).
// A C entry point around the reference's own image_interpCubic (src/interpol.cpp:42-86),
// compiled unmodified from the reference tree, so the oracle's sampler and the GPU
// sampler can be checked against the reference's compiled code bit for bit.
#include <cstdint>
#include <vector>

#include "interpol.hpp"

extern "C" int refimpl_interp_cubic(const uint16_t* frame, int dimx, int dimy, int64_t pitch, int n,
                                    const double* xy, double* out) {
  // the reference's layout: image[x][y] (src/p-tracker.hpp:80)
  std::vector<std::vector<uint16_t>> im(dimx, std::vector<uint16_t>(dimy));
  for (int x = 0; x < dimx; x++)
    for (int y = 0; y < dimy; y++) im[x][y] = frame[(int64_t)y * pitch + x];
  image_interpCubic ic;
  for (int k = 0; k < n; k++) out[k] = ic.getValue(im, xy[2 * k], xy[2 * k + 1]);
  return 0;
}
