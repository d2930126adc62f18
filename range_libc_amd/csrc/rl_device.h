// Device-side building blocks of the ray-casting kernels (gfx950).
//
// Every float operation here mirrors the reference CPU code operation for
// operation; the file is compiled with contraction off so that no multiply-add
// is fused.  See the per-function citations.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "rl_trig.h"

#pragma clang fp contract(off)

namespace rl {

struct TrigDev {
  const uint32_t* off_s;
  const uint16_t* key_s;
  const uint32_t* off_c;
  const uint16_t* key_c;
  double delta_s, delta_c;
};

// World-to-grid constants, derived on the host exactly as RangeLib.h:442-450.
struct WorldDev {
  float inv_scale, scale, origin_x, origin_y, sin_a, cos_a, rotation_const;
};

struct MapDev {
  const float* edt;    // x-major W*H
  const uint8_t* occ;  // x-major W*H
  int width, height;
  WorldDev world;
};

__device__ __forceinline__ void sincos_libc(float h, const TrigDev& t, float* s, float* c) {
  rltrig::Table ts{t.off_s, t.key_s, t.delta_s};
  rltrig::Table tc{t.off_c, t.key_c, t.delta_c};
  rltrig::sincosf_libc(h, ts, tc, s, c);
}

// float -> int as x86 cvttss2si does it: truncation, and INT_MIN for NaN or
// out-of-range values (the reference's implicit conversions compile to it).
__device__ __forceinline__ int trunc_x86(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT_MIN;
}

// RayMarching::calc_range (RangeLib.h:927-962): distThreshold 0 (:967),
// step_coeff 0.999f (:968).  (dx, dy) = (cosf h, sinf h).
template <bool COUNT>
__device__ __forceinline__ float rm_march(const float* __restrict__ edt, int W, int H,
                                          float max_range, float x0, float y0, float dx, float dy,
                                          int* reads) {
  float t = 0.0f;
  int nr = 0;
  while (t < max_range) {
    int px = trunc_x86(x0 + dx * t);
    int py = trunc_x86(y0 + dy * t);
    if ((unsigned)px >= (unsigned)W || (unsigned)py >= (unsigned)H) {
      if (COUNT) *reads = nr;
      return max_range;
    }
    float d = edt[(size_t)px * (size_t)H + (size_t)py];
    if (COUNT) ++nr;
    if (d <= 0.0f) {
      float xd = (float)px - x0;
      float yd = (float)py - y0;
      if (COUNT) *reads = nr;
      return __builtin_sqrtf(xd * xd + yd * yd);
    }
    float st = d * 0.999f;
    t += (st < 1.0f) ? 1.0f : st;
  }
  if (COUNT) *reads = nr;
  return max_range;
}

// OMap::isOccupied (RangeLib.h:204-210): out of bounds reads as free.
__device__ __forceinline__ bool occupied(const uint8_t* __restrict__ occ, int W, int H, int x,
                                         int y) {
  if (x < 0 || (unsigned)x >= (unsigned)W || y < 0 || (unsigned)y >= (unsigned)H) return false;
  return occ[(size_t)x * (size_t)H + (size_t)y] != 0;
}

// BresenhamsLine::calc_range (RangeLib.h:696-769).  (s, c) = (sinf h, cosf h).
template <bool COUNT>
__device__ __forceinline__ float bl_trace(const uint8_t* __restrict__ occ, int W, int H,
                                          float max_range, float x, float y, float s, float c,
                                          int* visits) {
  int nv = 0;
  if (occupied(occ, W, H, trunc_x86(x), trunc_x86(y))) {
    if (COUNT) *visits = 0;
    return 0.0f;
  }
  float x0 = y;
  float y0 = x;
  float x1 = y + max_range * s;
  float y1 = x + max_range * c;
  bool steep = fabsf(y1 - y0) > fabsf(x1 - x0);
  if (steep) {
    float tmp = x0;
    x0 = y0;
    y0 = tmp;
    tmp = x1;
    x1 = y1;
    y1 = tmp;
  }
  float deltax = fabsf(x1 - x0);
  float deltay = fabsf(y1 - y0);
  float error = 0.0f;
  float _x = x0;
  float _y = y0;
  float xstep = (x0 < x1) ? 1.0f : -1.0f;
  float ystep = (y0 < y1) ? 1.0f : -1.0f;
  float fw = (float)(unsigned)W;
  float fh = (float)(unsigned)H;
  int xend = trunc_x86(x1 + xstep);
  // The reference loop only ends for finite end points below 2^23 (beyond,
  // unit steps stop being exact and it never terminates); such rays give NaN.
  if (!(fabsf(x0) < 8388608.0f && fabsf(x1) < 8388608.0f)) {
    if (COUNT) *visits = 0;
    return __builtin_nanf("");
  }
  // (int)_x moves by one cell per step, and may repeat once while crossing
  // zero, so the loop below ends within |xend - (int)x0| + 2 steps.
  int cap = abs(xend - trunc_x86(x0)) + 2;
  // When steep the roles of (x, y) swap back for the map lookup.
  float limx = steep ? fw : fh;  // bound on _x
  float limy = steep ? fh : fw;  // bound on _y
  for (int it = 0; trunc_x86(_x) != xend && it < cap; ++it) {
    _x += xstep;
    error += deltay;
    if (COUNT) ++nv;
    // error * 2.00 >= deltax is evaluated in double (RangeLib.h:749); doubling
    // a float is exact, so the float comparison is the same test.
    if (error * 2.0f >= deltax) {
      _y += ystep;
      error -= deltax;
    }
    if (0.0f <= _y && _y < limy && 0.0f <= _x && _x < limx) {
      int cx = steep ? (int)_x : (int)_y;
      int cy = steep ? (int)_y : (int)_x;
      if (occ[(size_t)cx * (size_t)H + (size_t)cy]) {
        float xd = _x - x0;
        float yd = _y - y0;
        if (COUNT) *visits = nv;
        return __builtin_sqrtf(xd * xd + yd * yd);
      }
    }
  }
  if (COUNT) *visits = nv;
  return max_range;
}

// World-to-grid conversion of one pose (RangeLib.h:468-473).
__device__ __forceinline__ void to_grid(const WorldDev& w, float xw, float yw, float* gx,
                                        float* gy) {
  float x = (xw - w.origin_x) * w.inv_scale;
  float y = (yw - w.origin_y) * w.inv_scale;
  float temp = x;
  x = w.cos_a * x - w.sin_a * y;
  y = w.sin_a * temp + w.cos_a * y;
  *gx = x;
  *gy = y;
}

// std::min<float>(std::max<float>(v, 0.0), (float)K - 1.0)  (RangeLib.h:547-550)
__device__ __forceinline__ float clamp_table(float v, float hi) {
  float lo = (v < 0.0f) ? 0.0f : v;
  return (hi < lo) ? hi : lo;
}

}  // namespace rl
