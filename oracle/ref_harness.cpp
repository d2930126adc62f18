// extern "C" harness around the reference's own CPU implementation
// (includes/RangeLib.h, compiled from the reference tree where it lies, CPU
// only: USE_CUDA is left undefined).  TEST INFRASTRUCTURE ONLY: it writes the
// golden vectors, cross-checks the oracle restatement and serves as the timed
// CPU baseline.  Output goes to oracle/_ref/ (see Makefile).
//
// RangeUtils.h holds non-inline definitions, so RangeLib.h is included in
// exactly this one translation unit.
#include "includes/RangeLib.h"

#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

namespace {

ranges::OMap make_map(const uint8_t* occ, int w, int h, const float* world) {
  ranges::OMap m(w, h);
  for (int x = 0; x < w; ++x)
    for (int y = 0; y < h; ++y) m.grid[x][y] = occ[(size_t)x * h + y] != 0;
  // PyOMap defaults (RangeLibc.pyx:174-180) unless parameters are given
  m.world_scale = world ? world[0] : 1.0f;
  m.world_angle = world ? world[1] : 0.0f;
  m.world_origin_x = world ? world[2] : 0.0f;
  m.world_origin_y = world ? world[3] : 0.0f;
  m.world_sin_angle = world ? world[4] : 0.0f;
  m.world_cos_angle = world ? world[5] : 1.0f;
  return m;
}

struct Handle {
  int kind;
  ranges::RangeMethod* method;
};

// Read access to GiantLUTCast's protected table and constants: pointers to
// members formed inside a derived class (never instantiated).
struct GltPeek : ranges::GiantLUTCast {
  typedef std::vector<std::vector<std::vector<uint16_t> > > table_t;
  static table_t ranges::GiantLUTCast::*lut() { return &GltPeek::giant_lut; }
  static float ranges::GiantLUTCast::*max_div_limits_() { return &GltPeek::max_div_limits; }
  static float ranges::GiantLUTCast::*limits_div_max_() { return &GltPeek::limits_div_max; }
};

}  // namespace

extern "C" {

// OMap(filename, threshold) (RangeLib.h:159-201).  Call with occ == NULL to
// get the size; returns 0 on success.
int ref_load_png(const char* path, float threshold, int* w, int* h, uint8_t* occ) {
  ranges::OMap m(std::string(path), threshold);
  if (m.error()) return 1;
  *w = (int)m.width;
  *h = (int)m.height;
  if (occ)
    for (unsigned x = 0; x < m.width; ++x)
      for (unsigned y = 0; y < m.height; ++y) occ[(size_t)x * m.height + y] = m.grid[x][y] ? 1 : 0;
  return 0;
}

// DistanceTransform(OMap*) (RangeLib.h:345-373), x-major output.
void ref_edt(const uint8_t* occ, int w, int h, float* out) {
  ranges::OMap m = make_map(occ, w, h, nullptr);
  ranges::DistanceTransform dt(&m);
  for (int x = 0; x < w; ++x)
    for (int y = 0; y < h; ++y) out[(size_t)x * h + y] = dt.grid[x][y];
}

// kind 0 = RayMarching, 1 = BresenhamsLine; world = 6 floats or NULL.
void* ref_create(int kind, const uint8_t* occ, int w, int h, float max_range, const float* world) {
  ranges::OMap m = make_map(occ, w, h, world);
  Handle* hd = new Handle;
  hd->kind = kind;
  if (kind == 1)
    hd->method = new ranges::BresenhamsLine(m, max_range);
  else
    hd->method = new ranges::RayMarching(m, max_range);
  return hd;
}

// GiantLUTCast(OMap, max_range, theta_discretization) (RangeLib.h:1781-1823):
// the LUT is built here, on the CPU, by the reference's own constructor.
void* ref_create_glt(const uint8_t* occ, int w, int h, float max_range, int theta_disc, const float* world) {
  ranges::OMap m = make_map(occ, w, h, world);
  Handle* hd = new Handle;
  hd->kind = 2;
  hd->method = new ranges::GiantLUTCast(m, max_range, theta_disc);
  return hd;
}

// The LUT as built (giant_lut[x][y][i], x-major, theta fastest) and its
// scale constants (max_div_limits, limits_div_max).
void ref_glt_lut(void* p, uint16_t* out, float* consts) {
  ranges::GiantLUTCast* g = (ranges::GiantLUTCast*)((Handle*)p)->method;
  size_t k = 0;
  for (auto& sx : g->*GltPeek::lut())
    for (auto& sy : sx)
      for (uint16_t v : sy) out[k++] = v;
  consts[0] = g->*GltPeek::max_div_limits_();
  consts[1] = g->*GltPeek::limits_div_max_();
}

// CDDTCast(OMap, max_range, theta_discretization) (RangeLib.h:974-1142).
void* ref_create_cddt(const uint8_t* occ, int w, int h, float max_range, int theta_disc, const float* world) {
  ranges::OMap m = make_map(occ, w, h, world);
  Handle* hd = new Handle;
  hd->kind = 3;
  hd->method = new ranges::CDDTCast(m, max_range, theta_disc);
  return hd;
}

void ref_cddt_prune(void* p, float max_range) {
  ((ranges::CDDTCast*)((Handle*)p)->method)->prune(max_range);
}

// Sizes of compressed_lut: number of bins over all angles, number of values.
void ref_cddt_sizes(void* p, long long* nbins, long long* nvals) {
  ranges::CDDTCast* c = (ranges::CDDTCast*)((Handle*)p)->method;
  long long b = 0, v = 0;
  for (auto& ang : c->compressed_lut)
    for (auto& bin : ang) {
      ++b;
      v += (long long)bin.size();
    }
  *nbins = b;
  *nvals = v;
}

// compressed_lut as CSR (bins of angle 0 first), per-angle widths and lut_translations.
void ref_cddt_export(void* p, int* widths, float* trans, long long* start, float* vals) {
  ranges::CDDTCast* c = (ranges::CDDTCast*)((Handle*)p)->method;
  long long b = 0, v = 0;
  start[0] = 0;
  for (size_t a = 0; a < c->compressed_lut.size(); ++a) {
    widths[a] = (int)c->compressed_lut[a].size();
    trans[a] = c->lut_translations[a];
    for (auto& bin : c->compressed_lut[a]) {
      for (float f : bin) vals[v++] = f;
      start[++b] = v;
    }
  }
}

// RangeMethod::calc_range_many_radial_optimized (RangeLib.h:616-676).
void ref_calc_range_many_radial_optimized(void* p, float* ins, float* outs, int np_, int nr, float mn, float mx) {
  ((Handle*)p)->method->calc_range_many_radial_optimized(ins, outs, np_, nr, mn, mx);
}

// CDDTCast::serializeJson (RangeLib.h:1690-1735) into out (cap bytes); returns the length.
long long ref_cddt_serialize_json(void* p, char* out, long long cap) {
  std::stringstream ss;
  ((ranges::CDDTCast*)((Handle*)p)->method)->serializeJson(&ss);
  std::string str = ss.str();
  if (out && cap > 0) {
    long long n = (long long)str.size() < cap - 1 ? (long long)str.size() : cap - 1;
    memcpy(out, str.data(), (size_t)n);
    out[n] = 0;
  }
  return (long long)str.size();
}

void ref_destroy(void* p) {
  Handle* hd = (Handle*)p;
  delete hd->method;
  delete hd;
}

void ref_set_sensor_model(void* p, double* table, int k) {
  ((Handle*)p)->method->set_sensor_model(table, k);
}

float ref_calc_range(void* p, float x, float y, float heading) {
  return ((Handle*)p)->method->calc_range(x, y, heading);
}

// Grid-frame batch (the loop RayMarchingGPU::calc_range_many performs).
void ref_calc_range_many(void* p, const float* ins, float* outs, int n) {
  ranges::RangeMethod* m = ((Handle*)p)->method;
  for (int i = 0; i < n; ++i) outs[i] = m->calc_range(ins[3 * i], ins[3 * i + 1], ins[3 * i + 2]);
}

void ref_numpy_calc_range(void* p, float* ins, float* outs, int n) {
  ((Handle*)p)->method->numpy_calc_range(ins, outs, n);
}

void ref_numpy_calc_range_angles(void* p, float* ins, float* angles, float* outs, int np_, int na) {
  ((Handle*)p)->method->numpy_calc_range_angles(ins, angles, outs, np_, na);
}

void ref_eval_sensor_model(void* p, float* obs, float* ranges_, double* outs, int rpp, int np_) {
  ((Handle*)p)->method->eval_sensor_model(obs, ranges_, outs, rpp, np_);
}

void ref_repeat_angles_eval_sensor_model(void* p, float* ins, float* angles, float* obs,
                                         double* weights, int np_, int na) {
  ((Handle*)p)->method->calc_range_repeat_angles_eval_sensor_model(ins, angles, obs, weights, np_, na);
}

// Particle-filter step split over `threads` std::threads (the methods are
// re-entrant with the trace map off): numpy_calc_range_angles followed by
// eval_sensor_model on each thread's slice of particles.
void ref_pf_step_mt(void* p, float* ins, float* angles, float* obs, float* outs, double* weights,
                    int np_, int na, int threads) {
  ranges::RangeMethod* m = ((Handle*)p)->method;
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  int per = (np_ + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int b = t * per, e = std::min(np_, b + per);
    if (b >= e) break;
    pool.emplace_back([=]() {
      m->numpy_calc_range_angles(ins + 3 * (size_t)b, angles, outs + (size_t)b * na, e - b, na);
      m->eval_sensor_model(obs, outs + (size_t)b * na, weights + b, na, e - b);
    });
  }
  for (auto& th : pool) th.join();
}

// numpy_calc_range split over threads.
void ref_numpy_calc_range_mt(void* p, float* ins, float* outs, int n, int threads) {
  ranges::RangeMethod* m = ((Handle*)p)->method;
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  int per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int b = t * per, e = std::min(n, b + per);
    if (b >= e) break;
    pool.emplace_back([=]() { m->numpy_calc_range(ins + 3 * (size_t)b, outs + b, e - b); });
  }
  for (auto& th : pool) th.join();
}

}  // extern "C"
