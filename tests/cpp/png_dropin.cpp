// Reference-style C++ use of the drop-in header with a map loaded from a PNG:
//   png_dropin <map.png> <threshold> info
//       prints width, height, error flag and the occupied cells (no GPU needed)
//   png_dropin <map.png> <threshold> cast <max_range> <theta_disc>
//       reads "x y theta" world-frame queries from stdin and prints, per query,
//       the bit patterns of the ranges of RayMarching, BresenhamsLine,
//       CDDTCast (pruned) and GiantLUTCast through numpy_calc_range
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "RangeLibHIP.h"

static unsigned bits(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  ranges::OMap map(std::string(argv[1]), (float)std::atof(argv[2]));
  if (std::strcmp(argv[3], "info") == 0) {
    long occupied = 0;
    for (unsigned x = 0; x < map.width; ++x)
      for (unsigned y = 0; y < map.height; ++y) occupied += map.isOccupied((int)x, (int)y) ? 1 : 0;
    std::printf("%u %u %d %ld\n", map.width, map.height, map.error() ? 1 : 0, occupied);
    return 0;
  }
  if (argc < 6) return 2;
  const float R = (float)std::atof(argv[4]);
  const unsigned td = (unsigned)std::atoi(argv[5]);
  std::vector<float> ins;
  float x, y, t;
  while (std::scanf("%f %f %f", &x, &y, &t) == 3) {
    ins.push_back(x);
    ins.push_back(y);
    ins.push_back(t);
  }
  const int n = (int)(ins.size() / 3);
  try {
    ranges::RayMarching rm(map, R);
    ranges::BresenhamsLine bl(map, R);
    ranges::CDDTCast cddt(map, R, td);
    cddt.prune(R);
    ranges::GiantLUTCast glt(map, R, td);
    std::vector<float> o[4];
    for (auto& v : o) v.resize(n);
    rm.numpy_calc_range(ins.data(), o[0].data(), n);
    bl.numpy_calc_range(ins.data(), o[1].data(), n);
    cddt.numpy_calc_range(ins.data(), o[2].data(), n);
    glt.numpy_calc_range(ins.data(), o[3].data(), n);
    for (int i = 0; i < n; ++i) std::printf("%08x %08x %08x %08x\n", bits(o[0][i]), bits(o[1][i]), bits(o[2][i]), bits(o[3][i]));
  } catch (const std::runtime_error& e) {
    std::printf("runtime_error: %s\n", e.what());
    return 3;
  }
  return 0;
}
