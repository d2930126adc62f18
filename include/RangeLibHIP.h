// RangeLibHIP.h -- C++ drop-in for the ray-casting classes of the reference's
// includes/RangeLib.h, implemented over the C ABI of rangelib_hip.h.
//
// Same namespace, class names, constructors and method signatures as
// RangeLib.h for the classes on the ray-casting path:
//   ranges::OMap            (RangeLib.h:121-322: grid, world fields, PNG
//                            constructors and save)
//   ranges::RayMarchingGPU  (RangeLib.h:774-920)
//   ranges::RayMarching     (RangeLib.h:922-969)
//   ranges::BresenhamsLine  (RangeLib.h:691-772)
//   ranges::CDDTCast        (RangeLib.h:971-1770: tables built and pruned on the GPU)
//   ranges::GiantLUTCast    (RangeLib.h:1772-1900: table built on the GPU)
// Every casting method runs on the GPU.  Differences from the reference:
//   * failures throw std::runtime_error carrying rl_last_error() (the
//     reference prints, or throws std::string when built without CUDA);
//   * OMap(filename) keeps no raw_grid of gray values (nothing on the
//     casting path reads it), and a PNG that fails to decode leaves a 1 x 1
//     free grid with has_error set (the reference leaves the grid empty);
//   * the GPU sensor-model call computes the weights (the reference's GPU
//     path prints "unimplemented", kernels.cu:281-315).
// Link with librangelib_hip.so.
#ifndef RANGE_LIB_HIP_H
#define RANGE_LIB_HIP_H

#include <stdint.h>
#include <stdio.h>

#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rangelib_hip.h"

namespace ranges {

namespace detail {
inline void check(int rc) {
  if (rc != RL_OK) throw std::runtime_error(std::string("rangelib_hip: ") + rl_last_error());
}

// One context per process on device 0, created on first use.
inline rl_ctx* context() {
  static std::once_flag once;
  static rl_ctx* ctx = nullptr;
  static int rc = RL_OK;
  static std::string err;
  std::call_once(once, [] {
    rc = rl_create(0, &ctx);
    if (rc != RL_OK) err = rl_last_error();
  });
  if (rc != RL_OK) throw std::runtime_error("rangelib_hip: " + err);
  return ctx;
}
}  // namespace detail

struct OMap {
  bool has_error;
  unsigned width;   // x axis
  unsigned height;  // y axis
  std::vector<std::vector<bool> > grid;  // grid[x][y]
  std::string fn;
  float world_scale, world_angle, world_origin_x, world_origin_y, world_sin_angle, world_cos_angle;

  OMap(int w, int h)
      : has_error(false), width(w), height(h), grid(w, std::vector<bool>(h, false)), fn(""),
        world_scale(1.0f), world_angle(0.0f), world_origin_x(0.0f), world_origin_y(0.0f),
        world_sin_angle(0.0f), world_cos_angle(1.0f) {}

  // RangeLib.h:158-201: occupied where the pixel's gray value is below the
  // threshold (the PNG is decoded by the library, rl_png_occupancy)
  OMap(std::string filename) : OMap(filename, 128) {}
  OMap(std::string filename, float threshold)
      : has_error(false), width(0), height(0), fn(filename), world_scale(1.0f), world_angle(0.0f),
        world_origin_x(0.0f), world_origin_y(0.0f), world_sin_angle(0.0f), world_cos_angle(1.0f) {
    int w = 0, h = 0;
    std::vector<uint8_t> occ;
    int rc = rl_png_occupancy(filename.c_str(), threshold, nullptr, &w, &h);
    if (rc == RL_OK) {
      occ.resize((size_t)w * h);
      rc = rl_png_occupancy(filename.c_str(), threshold, occ.data(), &w, &h);
    }
    if (rc != RL_OK) {
      std::printf("ERROR %d: %s\n", rc, rl_last_error());
      has_error = true;
      w = h = 1;
      occ.assign(1, 0);
    }
    width = (unsigned)w;
    height = (unsigned)h;
    grid.assign(width, std::vector<bool>(height, false));
    for (unsigned x = 0; x < width; ++x)
      for (unsigned y = 0; y < height; ++y) grid[x][y] = occ[(size_t)x * height + y] != 0;
  }

  // RangeLib.h:264-291: opaque RGBA PNG, black where occupied; returns true on error
  bool save(std::string filename) {
    std::vector<uint8_t> occ = xmajor();
    const int rc = rl_png_save_occupancy(filename.c_str(), occ.data(), (int)width, (int)height);
    if (rc != RL_OK) std::cout << "encoder error " << rc << ": " << rl_last_error() << std::endl;
    return rc != RL_OK;
  }

  bool get(int x, int y) { return grid[x][y]; }
  bool isOccupied(int x, int y) {
    if (x < 0 || x >= (int)width || y < 0 || y >= (int)height) return false;
    return grid[x][y];
  }
  bool isOccupiedNT(int x, int y) { return grid[x][y]; }
  bool error() { return has_error; }
  int memory() { return (int)(sizeof(bool) * width * height); }

  // x-major uint8 occupancy as the C ABI takes it
  std::vector<uint8_t> xmajor() const {
    std::vector<uint8_t> occ((size_t)width * height);
    for (unsigned x = 0; x < width; ++x)
      for (unsigned y = 0; y < height; ++y) occ[(size_t)x * height + y] = grid[x][y] ? 1 : 0;
    return occ;
  }
};

// Shared machinery of the three methods (RangeMethod, RangeLib.h:412-689).
class GpuRangeMethod {
 public:
  GpuRangeMethod(const OMap& m, float mr, int method) : method_(method), max_range_(mr) {
    std::vector<uint8_t> occ = m.xmajor();
    detail::check(rl_map_create(detail::context(), occ.data(), (int)m.width, (int)m.height, &map_));
    detail::check(rl_map_set_world(map_, m.world_scale, m.world_angle, m.world_origin_x, m.world_origin_y,
                                   m.world_sin_angle, m.world_cos_angle));
  }
  virtual ~GpuRangeMethod() { rl_map_destroy(map_); }
  GpuRangeMethod(const GpuRangeMethod&) = delete;
  GpuRangeMethod& operator=(const GpuRangeMethod&) = delete;

  float maxRange() { return max_range_; }

  // RangeLib.h:439-480
  void numpy_calc_range(float* ins, float* outs, int num_casts) {
    detail::check(rl_numpy_calc_range(map_, method_, max_range_, ins, outs, num_casts));
  }
  // RangeLib.h:482-520
  void numpy_calc_range_angles(float* ins, float* angles, float* outs, int num_particles, int num_angles) {
    detail::check(rl_numpy_calc_range_angles(map_, method_, max_range_, ins, angles, outs, num_particles,
                                             num_angles));
  }
  // RangeLib.h:523-532
  void set_sensor_model(double* table, int table_width) {
    detail::check(rl_map_set_sensor_table(map_, table, table_width));
  }
  // RangeLib.h:533-555
  void eval_sensor_model(float* obs, float* ranges, double* outs, int rays_per_particle, int particles) {
    detail::check(rl_eval_sensor_model(map_, obs, ranges, outs, rays_per_particle, particles));
  }
  // RangeLib.h:558-612
  void calc_range_repeat_angles_eval_sensor_model(float* ins, float* angles, float* obs, double* weights,
                                                  int num_particles, int num_angles) {
    detail::check(rl_calc_range_repeat_angles_eval_sensor_model(map_, method_, max_range_, ins, angles, obs,
                                                                weights, num_particles, num_angles));
  }
  void saveTrace(std::string) {
    std::cout << "WARNING: trace map not generated, must compile with trace support enabled." << std::endl;
  }
  rl_map* handle() { return map_; }

 protected:
  float single(float x, float y, float heading) {
    float out = 0.0f;
    detail::check(rl_calc_range(map_, method_, max_range_, x, y, heading, &out));
    return out;
  }
  int method_;
  float max_range_;
  rl_map* map_ = nullptr;
};

class RayMarchingGPU : public GpuRangeMethod {
 public:
  RayMarchingGPU(OMap m, float mr) : GpuRangeMethod(m, mr, RL_METHOD_RM) {}
  // RangeLib.h:800-807: the GPU class answers batched queries only
  float calc_range(float, float, float) {
    std::cout << "Do not call calc_range on RayMarchingGPU, requires batched queries" << std::endl;
    return -1.0f;
  }
  // RangeLib.h:819-831: grid frame, no world conversion
  void calc_range_many(float* ins, float* outs, int num_casts) {
    detail::check(rl_calc_range_many(map_, method_, max_range_, ins, outs, num_casts));
  }
};

class RayMarching : public GpuRangeMethod {
 public:
  RayMarching(OMap m, float mr) : GpuRangeMethod(m, mr, RL_METHOD_RM) {}
  float calc_range(float x, float y, float heading) { return single(x, y, heading); }  // RangeLib.h:927
};

class BresenhamsLine : public GpuRangeMethod {
 public:
  BresenhamsLine(OMap m, float mr) : GpuRangeMethod(m, mr, RL_METHOD_BL) {}
  float calc_range(float x, float y, float heading) { return single(x, y, heading); }  // RangeLib.h:696
  int memory() { return 0; }
};

class CDDTCast : public GpuRangeMethod {
 public:
  // RangeLib.h:974-1174: the compressed tables are built on the GPU
  CDDTCast(OMap m, float mr, unsigned int td) : GpuRangeMethod(m, mr, RL_METHOD_CDDT) {
    detail::check(rl_map_build_cddt(map_, (int)td));
  }
  float calc_range(float x, float y, float heading) { return single(x, y, heading); }  // RangeLib.h:1342
  // RangeLib.h:1176-1290
  void prune(float max_range) { detail::check(rl_map_prune_cddt(map_, max_range)); }
  // RangeLib.h:616-676
  void calc_range_many_radial_optimized(float* ins, float* outs, int num_particles, int num_rays, float min_angle,
                                        float max_angle) {
    detail::check(rl_calc_range_many_radial_optimized(map_, method_, max_range_, ins, outs, num_particles, num_rays,
                                                      min_angle, max_angle));
  }
};

class GiantLUTCast : public GpuRangeMethod {
 public:
  // RangeLib.h:1781-1823: the table is built on the GPU
  GiantLUTCast(OMap m, float mr, unsigned int td) : GpuRangeMethod(m, mr, RL_METHOD_GLT) {
    detail::check(rl_map_build_giant_lut(map_, mr, (int)td));
  }
  float calc_range(float x, float y, float heading) { return single(x, y, heading); }  // RangeLib.h:1869
};

}  // namespace ranges

#endif
