// Compiles against the C++ drop-in header the way reference code using
// RangeLib.h would: construct an OMap, a method, and cast a batch.
// Without a GPU the context cannot be created and a std::runtime_error is
// expected; with one, the program prints the ranges.
#include <cstdio>
#include <vector>

#include "RangeLibHIP.h"

int main() {
  ranges::OMap map(16, 16);
  for (int i = 0; i < 16; ++i) map.grid[i][15] = true;  // a wall at y = 15
  try {
    ranges::RayMarchingGPU rm(map, 50.0f);
    std::vector<float> ins = {3.5f, 2.5f, 1.5707964f, 8.0f, 8.0f, 1.5707964f};
    std::vector<float> outs(2);
    rm.calc_range_many(ins.data(), outs.data(), 2);
    ranges::BresenhamsLine bl(map, 50.0f);
    std::printf("rm %.6f %.6f bl %.6f\n", outs[0], outs[1], bl.calc_range(3.0f, 2.0f, 1.5707964f));
    return 0;
  } catch (const std::runtime_error& e) {
    std::printf("runtime_error: %s\n", e.what());
    return 3;
  }
}
