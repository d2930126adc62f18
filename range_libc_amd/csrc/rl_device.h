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

// Explicit address spaces for the hot loads: a load through a generic pointer
// is a flat access, which the compiler can only wait for together with every
// other outstanding load (vmcnt(0) lgkmcnt(0)); global and LDS pointers give
// global_load / ds_read with their own counters.
#define RL_GLB __attribute__((address_space(1)))
#define RL_LDS __attribute__((address_space(3)))

namespace rl {

template <class T>
__device__ __forceinline__ const RL_GLB T* glb(const T* p) {
  return (const RL_GLB T*)p;
}
template <class T>
__device__ __forceinline__ const RL_LDS T* lds(const T* p) {
  return (const RL_LDS T*)p;
}

struct TrigDev {
  rltrig::Bitmap sin_exc, cos_exc;  // exception bitmaps of the C library's sinf / cosf
};

// World-to-grid constants, derived on the host exactly as RangeLib.h:442-450.
struct WorldDev {
  float inv_scale, scale, origin_x, origin_y, sin_a, cos_a, rotation_const;
};

// Compact distance map: every exact distance-transform value is sqrtf(n) for
// an integer n that is a sum of two squares.  The 255 smallest such values
// (d <= 27.3) are tabulated once (lut); the map stores per cell the index of
// its value in that table, or RL_IDX8_ESCAPE when the value is not tabulated
// (the fp32 map is then read).  Cells are stored in 8x8 blocks of 64 bytes,
// blocks x-major, y-major inside a block: byte (((x >> 3) * nby + (y >> 3)) << 6) |
// ((y & 7) << 3) | (x & 7), which is (x >> 3) * 64 * nby + 8 * y + (x & 7): one
// shift, one mask, one shift-add and one multiply-add per cell address.
#define RL_IDX8_ESCAPE 255u
#define RL_LUT_SIZE 256

// CDDTCast tables (built by rl_map_build_cddt): angle a has width[a] bins;
// bin i of angle a holds vals[start[bin_off[a] + i] .. start[bin_off[a] + i + 1]),
// sorted and free of duplicates; cosv / sinv / trans are the reference's
// per-angle cosf, sinf and lut_translations.
struct CddtDev {
  const int* width;
  const float* trans;
  const float* cosv;
  const float* sinv;
  const long long* bin_off;
  const long long* start;
  const float* vals;
  int td;
  float td_div_2pi;
};

struct MapDev {
  const float* edt;     // x-major W*H
  const uint8_t* occ;   // x-major W*H
  const uint8_t* idx8;  // blocked index map (see above)
  const float* steps;   // RL_LUT_SIZE march steps max(d * 0.999f, 1) of the tabulated distances
  int width, height;
  int nby;              // blocks along y: (height + 7) / 8
  int use_box;          // stage the 64x64 cells around a particle in LDS
  WorldDev world;
  // GiantLUTCast table (null when not built): glt[(x*height + y)*glt_td + bin]
  const uint16_t* glt;
  int glt_td;
  float glt_R, glt_scale, glt_td_div_2pi;  // its max_range, max_div_limits, td / M_2PI
  CddtDev cddt;  // CDDTCast tables (null pointers when not built)
  // Occupancy bitmap in 8x8-cell tiles, one 64-bit word each, in two
  // orientations (null when not built): tile (tx, ty) = cells (8 tx + u,
  // 8 ty + v), u, v < 8.  Steep Bresenham walks (walking along x) read the
  // x-major tiles: word tx * bits_nty + ty, bit 8 u + v; the others (walking
  // along y) read the y-major tiles from byte bits_ybase: word ty * bits_ntx +
  // tx, bit 8 v + u.  The walking axis picks the tile row, the other axis
  // the word in the row and the byte.
  const uint8_t* bits;
  unsigned bits_bytes, bits_ybase;  // size of both orientations; start of the y-major one
  int bits_ntx, bits_nty;           // tiles along x ((W + 7) / 8) and y ((H + 7) / 8)
};

__device__ __forceinline__ void sincos_libc(float h, const TrigDev& t, float* s, float* c) {
  rltrig::sincosf_libc_bm(h, t.sin_exc, t.cos_exc, s, c);
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
    float d = glb(edt)[(size_t)px * (size_t)H + (size_t)py];
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

// LDS copy of the 64x64 cells (8x8 compact blocks) around a particle's origin
// cell, shared by the lanes of the wave that casts its beams.  Box block (i, j)
// is map block (bx0 + i, by0 + j), stored at lds[(i * 8 + j) * 64] in the
// compact layout; blocks outside the map are never read (the bounds test comes
// first).
#define RL_BOX_BLOCKS 8
#define RL_BOX_BYTES (RL_BOX_BLOCKS * RL_BOX_BLOCKS * 64)
struct Box {
  const uint8_t* lds;  // null: no box
  int bx0, by0;
};

// Loads the box around grid cell (x0, y0) into `lds` (RL_BOX_BYTES, one wave:
// lane l copies box block (l / 8, l % 8)).  Returns the box.
__device__ __forceinline__ Box stage_box(const MapDev& m, uint8_t* lds, float x0, float y0,
                                         bool active) {
  Box b{nullptr, 0, 0};
  if (!active || !(x0 >= 0.0f && x0 < (float)m.width && y0 >= 0.0f && y0 < (float)m.height)) return b;
  b.bx0 = ((int)x0 >> 3) - RL_BOX_BLOCKS / 2 + 1;
  b.by0 = ((int)y0 >> 3) - RL_BOX_BLOCKS / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int bx = b.bx0 + (lane >> 3), by = b.by0 + (lane & 7);
  const int nbx = (m.width + 7) >> 3;
  uint4* dst = reinterpret_cast<uint4*>(lds + lane * 64);
  if (bx >= 0 && bx < nbx && by >= 0 && by < m.nby) {
    const uint4* src = reinterpret_cast<const uint4*>(m.idx8 + (((size_t)bx * m.nby + by) << 6));
    uint4 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
    dst[0] = v0;
    dst[1] = v1;
    dst[2] = v2;
    dst[3] = v3;
  }
  // the wave reads what its other lanes wrote: complete the LDS stores first
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  b.lds = lds;
  return b;
}

// Speculation width and start.  Measured on a 2048^2 corridor map, 4000
// particles x 60 beams, with the first single-exit loops: plain 29.7 us per
// step; width 2 after 8 or 12 reads 25.1, 25.2 us; width 3 after 4, 8 or 12
// reads 25.7, 25.0, 24.9 us.  With the whole-wave finish and the current loop
// forms: width 2 after 6, 8 or 12 reads 20.45, 20.2, 20.38 us; width 3 after 6
// or 8 reads 21.2, 20.8 us.  Earlier speculation doubles the loads of the bulk
// of the launch; later speculation leaves the long rays' chains (runs of equal
// steps along corridor walls) unshortened.
#ifndef RL_SPEC
#define RL_SPEC 2  // positions loaded per round of the speculative march (1: plain)
#endif
#ifndef RL_SPEC_AFTER
#define RL_SPEC_AFTER 8  // reads of a ray before it starts to speculate
#endif

// Compact-map reads go through a buffer resource covering the map: a read
// past its end returns 0 instead of faulting, and a position off the map
// (negative or too large coordinates wrap in the 32-bit offset) reads either
// that 0 or some cell of the map.  Every march decides off-map positions from
// the position alone, before it looks at the code, so those reads need no
// clamp of the position (two selects per read less on the march's chain).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t idx8_rsrc(const MapDev& m) {
  // the compact map is built only below 4 GiB; 0x00020000: 32-bit data format
  const unsigned bytes = ((unsigned)((m.width + 7) >> 3) * (unsigned)m.nby) << 6;
  return __builtin_amdgcn_make_buffer_rsrc((void*)m.idx8, (short)0, (int)bytes, 0x00020000);
}

// Code of grid cell (ux, uy) of the compact map: from the LDS box when BOX and
// the box covers the cell, else from the map in HBM.
template <bool BOX>
__device__ __forceinline__ unsigned idx8_code(const MapDev& m, const Box& box, unsigned ux, unsigned uy) {
  if (BOX) {
    unsigned ib = (unsigned)((int)(ux >> 3) - box.bx0), jb = (unsigned)((int)(uy >> 3) - box.by0);
    if (ib < RL_BOX_BLOCKS && jb < RL_BOX_BLOCKS)
      return lds(box.lds)[((ib * RL_BOX_BLOCKS + jb) << 6) | ((uy & 7u) << 3) | (ux & 7u)];
  }
  // 32-bit offset from the map's base (the compact map is built only below
  // 4 GiB, so 64 * nby < 2^24): a multiply-add on the march's chain and a
  // load through the map's buffer resource (scalar base, range checked)
  const unsigned off = __umul24(ux >> 3, (unsigned)m.nby << 6) + ((uy << 3) | (ux & 7u));
  return __builtin_amdgcn_raw_buffer_load_b8(idx8_rsrc(m), off, 0, 0);
}

// March step max(d * 0.999f, 1.0f) from a cell with non-zero code c: from the
// step table staged in LDS (`lut`, the same float operations done once per
// code), or from the fp32 map for the escape code.
__device__ __forceinline__ float idx8_step(const MapDev& m, const float* lut, unsigned c, int px, int py) {
  float g = lds(lut)[c];  // lut[RL_IDX8_ESCAPE] exists: the read needs no branch
  if (c == RL_IDX8_ESCAPE) {
    const float st = glb(m.edt)[(size_t)px * (unsigned)m.height + py] * 0.999f;
    g = (st < 1.0f) ? 1.0f : st;
  }
  return g;
}

// The same march over the compact map.  `lut` is the step table staged in
// LDS.  Valid when |x0|, |y0| and max_range are below 1e9 (checked by the
// caller): every x0 + dx*t then lies in int range, so the conversion is a
// plain truncation.  Identical results to rm_march: a cell's code is 0 exactly
// when its distance is 0, and any other code maps back to the exact distance.
//
// Each loop has one way out: the ways a ray ends (past max_range, off the map,
// on an occupied cell) are decided with selects, and only the kind of end
// (1 = max_range, 2 = hit, 3 = handed to the whole wave) leaves the loop, with
// t at the ending position (a hit's cell is formed again from it); the
// exec-mask bookkeeping and register copies of several exits per round cost a
// quarter of the kernel.  Off-map positions are read through the map's buffer
// resource (0 or an ignored cell) and never used.
//
// Speculation (SPEC > 1, RL_SPEC unless a caller picks another width), after
// a ray's first RL_SPEC_AFTER reads: each round loads the cells at t, t + g,
// (t + g) + g, ... (SPEC positions, all loads issued before any is used),
// where g is the last step taken, and then
// walks them in order: position k + 1 is the true next position exactly when
// cell k's step equals g, i.e. when its code equals the code that gave g
// (distinct codes have distinct steps; escaped cells never match).  The walk
// stops at the first mismatch and the next round starts from the true next t.
// Positions, float operations and results are those of the plain loop, and
// `reads` counts resolved positions only.
// State a lane hands to solo_tail when it is the last of its wave still
// marching (see there): its ray and its position, last step and that step's
// code.  pending = false: the lane's range is final.
struct MarchTail {
  bool pending;
  float x0, y0, dx, dy, t, g;
  unsigned gcode;
};

template <bool COUNT, bool BOX, bool SOLO = false, int SPEC = RL_SPEC>
__device__ __forceinline__ float rm_march_idx8_t(const MapDev& m, const float* lut, const Box& box,
                                                 float max_range, float x0, float y0, float dx,
                                                 float dy, int* reads, float t, float g,
                                                 unsigned gcode, int nr, MarchTail* tail = nullptr) {
  // (t, g, gcode, nr): position, last step, the code that gave it (256 matches
  // no cell) and reads so far; (0, 1, 1, 0) for a march from its origin
  const unsigned W = (unsigned)m.width, H = (unsigned)m.height;
  // the loops leave t at the position that ended the ray; a hit's cell is
  // formed again from it after the loops (the same float operations)
  int end = 0;
  for (int i = nr; SPEC == 1 || i < RL_SPEC_AFTER; ++i) {
    const int px = (int)(x0 + dx * t), py = (int)(y0 + dy * t);
    const bool live = t < max_range;
    const bool inb = (unsigned)px < W && (unsigned)py < H;
    const unsigned c = idx8_code<BOX>(m, box, (unsigned)px, (unsigned)py);
    const int e = (!live || !inb) ? 1 : (c == 0u ? 2 : 0);
    if (COUNT) nr += (live && inb) ? 1 : 0;
    if (e) {
      end = e;
      break;
    }
    g = idx8_step(m, lut, c, px, py);
    gcode = (c != RL_IDX8_ESCAPE) ? c : 256u;
    t += g;
  }
  while (SPEC == 2 && !end) {  // width 2, written out
    if (SOLO && __popcll(__builtin_amdgcn_read_exec()) == 1) {
      // the wave's last marching lane: handed to the whole wave after the
      // loops (one exit block for every way out of the loop: a return from
      // here cost register copies of t, g and gcode in every round)
      end = 3;
      break;
    }
    const float t0 = t, t1 = t + g;
    const int px0 = (int)(x0 + dx * t0), py0 = (int)(y0 + dy * t0);
    const int px1 = (int)(x0 + dx * t1), py1 = (int)(y0 + dy * t1);
    const bool in0 = (unsigned)px0 < W && (unsigned)py0 < H;
    const bool in1 = (unsigned)px1 < W && (unsigned)py1 < H;
    const unsigned c0 = idx8_code<BOX>(m, box, (unsigned)px0, (unsigned)py0);
    const unsigned c1 = idx8_code<BOX>(m, box, (unsigned)px1, (unsigned)py1);
    // position k is readable when it is live and on the map; a readable cell
    // with g's code continues the run (g's code is never a hit's 0), so
    // position 1 is the true next position exactly then
    const bool ok0 = (t0 < max_range) & in0, ok1 = (t1 < max_range) & in1;
    const bool use1 = ok0 & (c0 == gcode);
    const bool oks = use1 ? ok1 : ok0;
    const unsigned cs = use1 ? c1 : c0;
    const float ts = use1 ? t1 : t0;
    const int es = !oks ? 1 : (cs == 0u ? 2 : 0);
    // positions before the resolved one were read and continued; the resolved
    // one was read unless it is past max_range or off the map (es == 1)
    if (COUNT) nr += (use1 ? 1 : 0) + (es != 1 ? 1 : 0);
    const int sx = use1 ? px1 : px0, sy = use1 ? py1 : py0;
    if (es) {
      end = es, t = ts;
      break;
    }
    g = idx8_step(m, lut, cs, sx, sy);
    gcode = (cs != RL_IDX8_ESCAPE) ? cs : 256u;
    t = ts + g;
  }
  while (SPEC == 3 && !end) {  // width 3, written out
    if (SOLO && __popcll(__builtin_amdgcn_read_exec()) == 1) {
      end = 3;  // the wave's last marching lane: handed to the whole wave after the loops
      break;
    }
    const float t0 = t, t1 = t + g, t2 = t1 + g;
    const int px0 = (int)(x0 + dx * t0), py0 = (int)(y0 + dy * t0);
    const int px1 = (int)(x0 + dx * t1), py1 = (int)(y0 + dy * t1);
    const int px2 = (int)(x0 + dx * t2), py2 = (int)(y0 + dy * t2);
    const bool in0 = (unsigned)px0 < W && (unsigned)py0 < H;
    const bool in1 = (unsigned)px1 < W && (unsigned)py1 < H;
    const bool in2 = (unsigned)px2 < W && (unsigned)py2 < H;
    const unsigned c0 = idx8_code<BOX>(m, box, (unsigned)px0, (unsigned)py0);
    const unsigned c1 = idx8_code<BOX>(m, box, (unsigned)px1, (unsigned)py1);
    const unsigned c2 = idx8_code<BOX>(m, box, (unsigned)px2, (unsigned)py2);
    const bool ok0 = (t0 < max_range) & in0, ok1 = (t1 < max_range) & in1, ok2 = (t2 < max_range) & in2;
    const bool u1 = ok0 & (c0 == gcode);
    const bool u2 = u1 & ok1 & (c1 == gcode);
    const bool oks = u2 ? ok2 : (u1 ? ok1 : ok0);
    const unsigned cs = u2 ? c2 : (u1 ? c1 : c0);
    const float ts = u2 ? t2 : (u1 ? t1 : t0);
    const int es = !oks ? 1 : (cs == 0u ? 2 : 0);
    if (COUNT) nr += (u1 ? 1 : 0) + (u2 ? 1 : 0) + (es != 1 ? 1 : 0);
    const int sx = u2 ? px2 : (u1 ? px1 : px0), sy = u2 ? py2 : (u1 ? py1 : py0);
    if (es) {
      end = es, t = ts;
      break;
    }
    g = idx8_step(m, lut, cs, sx, sy);
    gcode = (cs != RL_IDX8_ESCAPE) ? cs : 256u;
    t = ts + g;
  }
  while (SPEC > 3 && !end) {  // any width (slower selects than the written-out forms)
    float tk[SPEC];
    int px[SPEC], py[SPEC], e[SPEC];
    unsigned c[SPEC];
    bool in[SPEC];
#pragma unroll
    for (int k = 0; k < SPEC; ++k) {
      tk[k] = (k == 0) ? t : tk[k - 1] + g;
      px[k] = (int)(x0 + dx * tk[k]);
      py[k] = (int)(y0 + dy * tk[k]);
      in[k] = (unsigned)px[k] < W && (unsigned)py[k] < H;
    }
#pragma unroll
    for (int k = 0; k < SPEC; ++k)
      c[k] = idx8_code<BOX>(m, box, (unsigned)px[k], (unsigned)py[k]);
#pragma unroll
    for (int k = 0; k < SPEC; ++k) e[k] = (!(tk[k] < max_range) || !in[k]) ? 1 : (c[k] == 0u ? 2 : 0);
    int sel = SPEC - 1;
#pragma unroll
    for (int k = SPEC - 2; k >= 0; --k)
      if (e[k] != 0 || c[k] != gcode) sel = k;
    int es = e[0], sx = px[0], sy = py[0];
    unsigned cs = c[0];
    float ts = tk[0];
#pragma unroll
    for (int k = 1; k < SPEC; ++k)
      if (sel == k) es = e[k], sx = px[k], sy = py[k], cs = c[k], ts = tk[k];
    if (COUNT) nr += sel + (es != 1 ? 1 : 0);
    if (es) {
      end = es, t = ts;
      break;
    }
    g = idx8_step(m, lut, cs, sx, sy);
    gcode = (cs != RL_IDX8_ESCAPE) ? cs : 256u;
    t = ts + g;
  }
  if (SOLO && end == 3) {  // the whole wave takes over (solo_tail)
    *tail = MarchTail{true, x0, y0, dx, dy, t, g, gcode};
    return 0.0f;
  }
  if (COUNT) *reads = nr;
  if (end == 2) {
    const int hx = (int)(x0 + dx * t), hy = (int)(y0 + dy * t);
    const float xd = (float)hx - x0;
    const float yd = (float)hy - y0;
    return __builtin_sqrtf(xd * xd + yd * yd);
  }
  return max_range;
}

// SPEC: positions loaded per round once speculating (1: the plain march), as
// in rm_march_idx8_t.
template <bool COUNT, int SPEC = RL_SPEC>
__device__ __forceinline__ float rm_march_idx8(const MapDev& m, const float* lut, const Box& box,
                                               float max_range, float x0, float y0, float dx,
                                               float dy, int* reads) {
  if (dx != dx || dy != dy) {  // see rm_cast_idx8
    if (COUNT) *reads = 0;
    return max_range;
  }
  if (box.lds)
    return rm_march_idx8_t<COUNT, true, false, SPEC>(m, lut, box, max_range, x0, y0, dx, dy, reads, 0.0f, 1.0f, 1u, 0);
  return rm_march_idx8_t<COUNT, false, false, SPEC>(m, lut, box, max_range, x0, y0, dx, dy, reads, 0.0f, 1.0f, 1u, 0);
}

// RayMarching::calc_range (RangeLib.h:927-962) over the compact map, from the
// heading.  The first read is issued before the direction is known: at t = 0
// the position x0 + dx * 0 is x0 for every finite dx, so the origin cell's
// load overlaps the sinf/cosf evaluation.  A NaN direction (NaN or infinite
// heading) returns max_range, as in the reference, whose float-to-int
// conversion of a NaN position gives INT_MIN, a cell outside the map.
template <bool COUNT, bool BOX, bool SOLO = false>
__device__ __forceinline__ float rm_cast_idx8_t(const MapDev& m, const float* lut, const Box& box,
                                                const TrigDev& trig, float max_range, float x0, float y0,
                                                float heading, int* reads, MarchTail* tail = nullptr) {
  const unsigned W = (unsigned)m.width, H = (unsigned)m.height;
  const int px = (int)x0, py = (int)y0;
  const bool in = (unsigned)px < W && (unsigned)py < H;
  const unsigned c_any = idx8_code<BOX>(m, box, (unsigned)px, (unsigned)py);
  const unsigned c = in ? c_any : 0u;
  float dy, dx;
  sincos_libc(heading, trig, &dy, &dx);
  if (COUNT) *reads = 0;
  if (!(0.0f < max_range) || dx != dx || dy != dy || !in) return max_range;
  if (COUNT) *reads = 1;
  if (c == 0u) {
    const float xd = (float)px - x0;
    const float yd = (float)py - y0;
    return __builtin_sqrtf(xd * xd + yd * yd);
  }
  const float g = idx8_step(m, lut, c, px, py);
  return rm_march_idx8_t<COUNT, BOX, SOLO>(m, lut, box, max_range, x0, y0, dx, dy, reads, g, g,
                                           (c != RL_IDX8_ESCAPE) ? c : 256u, 1, tail);
}

// The end of a march whose lane was left alone in its wave (MarchTail): the
// whole wave, converged, follows that one ray with lane-parallel speculation.
// Lane k (of RL_SOLO_W) forms t_k = t + g + ... + g by k float additions in
// order -- the positions the march reaches if its next k steps all repeat the
// last step g -- and loads its cell.  The first lane whose cell ends the ray
// or carries another code than g's is the next true position: positions up
// to it were read and continued with step g, and the march goes on from its
// cell as rm_march_idx8_t does (same float operations, same bits).  When all
// RL_SOLO_W cells repeat g's code, the march continues from the last of them.
// Returns the ray's range on the pending lane and `d` on the others.  Must be
// reached by the whole wave.
#ifndef RL_SOLO_W
#define RL_SOLO_W 8
#endif
#ifndef RL_SOLO_TAIL
#define RL_SOLO_TAIL 1  // 0: every lane finishes its own march
#endif
__device__ __forceinline__ float solo_tail(const MapDev& m, const float* lut, float max_range,
                                           const MarchTail& tail, float d) {
  unsigned long long pm = __builtin_amdgcn_ballot_w64(tail.pending);
  auto rdf = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
  for (; pm; pm &= pm - 1ull) {  // one pending lane per wave in practice
  const int L = __builtin_ctzll(pm);
  const float x0 = rdf(tail.x0, L), y0 = rdf(tail.y0, L), dx = rdf(tail.dx, L), dy = rdf(tail.dy, L);
  float t = rdf(tail.t, L), g = rdf(tail.g, L);
  unsigned gcode = (unsigned)__builtin_amdgcn_readlane((int)tail.gcode, L);
  const unsigned W = (unsigned)m.width, H = (unsigned)m.height;
  const int k = (int)(threadIdx.x & (RL_SOLO_W - 1));  // lanes beyond RL_SOLO_W repeat the first ones
  int end, hx, hy;
  const unsigned long long lanes = (1ull << RL_SOLO_W) - 1ull;
  for (;;) {
    float tk = t;
#pragma unroll
    for (int j = 0; j < RL_SOLO_W - 1; ++j)
      if (j < k) tk = tk + g;
    const int px = (int)(x0 + dx * tk), py = (int)(y0 + dy * tk);
    const bool inb = (unsigned)px < W && (unsigned)py < H;
    // positions past max_range or off the map end the ray (decided before the load)
    const unsigned long long out = __builtin_amdgcn_ballot_w64(!(tk < max_range) || !inb) & lanes;
    const unsigned c = idx8_code<false>(m, Box{nullptr, 0, 0}, (unsigned)px, (unsigned)py);
    // a cell that is not g's code (a hit's code 0 never is) or an ending position stops the run
    const unsigned long long stop = (__builtin_amdgcn_ballot_w64(c != gcode) & lanes) | out;
    const int j = stop ? __builtin_ctzll(stop) : RL_SOLO_W - 1;
    const float tj = rdf(tk, j);
    // one way out of the loop (several cost register copies in every round):
    // the ray ends at position j when it is past max_range or off the map, or
    // on a hit; otherwise a stop takes the step of j's cell, and no stop keeps
    // g (every cell repeated g's code)
    const unsigned cj = (unsigned)__builtin_amdgcn_readlane((int)c, j);
    const int pxj = __builtin_amdgcn_readlane(px, j), pyj = __builtin_amdgcn_readlane(py, j);
    const bool is_out = (out >> j) & 1ull;
    if (is_out || (stop && cj == 0u)) {
      end = is_out ? 1 : 2, hx = pxj, hy = pyj;
      break;
    }
    if (stop) {
      g = idx8_step(m, lut, cj, pxj, pyj);
      gcode = (cj != RL_IDX8_ESCAPE) ? cj : 256u;
    }
    t = tj + g;
  }
  float r = max_range;
  if (end == 2) {
    const float xd = (float)hx - x0;
    const float yd = (float)hy - y0;
    r = __builtin_sqrtf(xd * xd + yd * yd);
  }
  if ((int)(threadIdx.x & 63) == L) d = r;
  }
  return d;
}

template <bool COUNT>
__device__ __forceinline__ float rm_cast_idx8(const MapDev& m, const float* lut, const Box& box,
                                              const TrigDev& trig, float max_range, float x0, float y0,
                                              float heading, int* reads) {
  if (box.lds) return rm_cast_idx8_t<COUNT, true>(m, lut, box, trig, max_range, x0, y0, heading, reads);
  return rm_cast_idx8_t<COUNT, false>(m, lut, box, trig, max_range, x0, y0, heading, reads);
}

// GiantLUTCast::discretize_theta (RangeLib.h:1832-1864) with
// _USE_ALTERNATE_MOD and _USE_CACHED_CONSTANTS (RangeLib.h:66-67): the wrap
// loops add or subtract the double M_2PI (RangeLib.h:61) and round back to
// float each time, then roundf(theta * (td / M_2PI)) % td.  -1 for a NaN
// heading (the reference indexes out of its table) and for a heading so large
// that a wrap step leaves it unchanged (the reference loops forever); every
// terminating case (|theta| below about 1.3e8) is followed exactly.
#define RL_M_2PI 6.28318530718
__device__ __forceinline__ int glt_bin(float theta, int td, float td_div_2pi) {
  if (theta != theta) return -1;
  if (theta < 0.0) {
    while (theta < 0.0) {
      const float next = (float)((double)theta + RL_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  } else if (theta > RL_M_2PI) {
    while (theta > RL_M_2PI) {
      const float next = (float)((double)theta - RL_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  }
  return (int)roundf(theta * td_div_2pi) % td;
}

// GiantLUTCast::calc_range (RangeLib.h:1866-1877).  NaN where glt_bin is -1
// and for a NaN origin (the reference converts it to an out-of-table index).
template <bool COUNT>
__device__ __forceinline__ float glt_calc_range(const MapDev& m, float x, float y, float heading, int* reads) {
  if (COUNT) *reads = 0;
  if (x < 0 || x >= (float)m.width || y < 0 || y >= (float)m.height) return m.glt_R;
  if (x != x || y != y) return __builtin_nanf("");
  const int b = glt_bin(heading, m.glt_td, m.glt_td_div_2pi);
  if (b < 0) return __builtin_nanf("");
  if (COUNT) *reads = 1;
  const size_t cell = (size_t)(int)x * (unsigned)m.height + (unsigned)(int)y;
  return (float)glb(m.glt)[cell * (unsigned)m.glt_td + (unsigned)b] * m.glt_scale;
}

// CDDTCast::discretize_theta (RangeLib.h:1279-1339): the GiantLUTCast wrap,
// then the half-turn symmetry (theta >= M_PI: flipped, theta -= M_PI in
// double), roundf(theta * (td / M_2PI)), and a bin that rounds up to td / 2
// wraps to 0 with the flip toggled.  -1 where the wrap never ends.
#define RL_M_PI 3.14159265358979323846
__device__ __forceinline__ int cddt_bin(const CddtDev& c, float theta, bool* flipped) {
  if (theta != theta) return -1;
  if (theta < 0.0) {
    while (theta < 0.0) {
      const float next = (float)((double)theta + RL_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  } else if (theta > RL_M_2PI) {
    while (theta > RL_M_2PI) {
      const float next = (float)((double)theta - RL_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  }
  bool fl = false;
  if (theta >= RL_M_PI) {
    fl = true;
    theta = (float)((double)theta - RL_M_PI);
  }
  int rounded = (int)roundf(theta * c.td_div_2pi);
  if (rounded == c.td >> 1) {
    rounded = 0;
    fl = !fl;
  }
  *flipped = fl;
  return rounded % c.td;
}

// CDDTCast::calc_range (RangeLib.h:1341-1514): the origin is projected onto
// the bin of the (discretized) heading and the nearest obstacle position
// ahead of it (behind it when flipped) is searched in that bin.  The
// reference uses upper_bound above 64 entries and a linear scan below; the
// scan's answer is the first entry >= x (lower_bound) or, flipped, the last
// entry <= x (upper_bound - 1), so both are binary searches here: a few
// dependent loads instead of up to 64 in sequence, same results.  NaN where the
// heading's wrap never ends and where the reference tests the occupancy of a
// cell outside the map (it indexes grid[x][y] out of bounds); 0 where its
// binary search would read past the bin (the query projects exactly onto the
// bin's last entry), which is what its linear scan returns.
template <bool COUNT>
__device__ __forceinline__ float cddt_calc_range(const MapDev& m, float max_range, float x, float y, float heading,
                                                 int* reads) {
  const CddtDev& c = m.cddt;
  if (COUNT) *reads = 0;
  bool fl;
  const int a = cddt_bin(c, -heading, &fl);
  if (a < 0) return __builtin_nanf("");
  const float cs = c.cosv[a], sn = c.sinv[a];
  const float lx = x * cs - y * sn;
  const float ly = (x * sn + y * cs) + c.trans[a];
  const unsigned li = (unsigned)trunc_x86(ly);
  if (li >= (unsigned)c.width[a]) return max_range;
  const long long b = c.bin_off[a] + li, s = c.start[b];
  const long long n = c.start[b + 1] - s;
  const RL_GLB float* v = glb(c.vals) + s;
  const long long high = n - 1;
  if (high == -1) return max_range;
  if (COUNT) *reads = 1;
  const bool inside = x > -1.0f && x < (float)m.width && y > -1.0f && y < (float)m.height;
  if (lx != lx) {  // a NaN projection (not reached from finite origins): the reference's searches as written
    if (!inside) return __builtin_nanf("");
    if (glb(m.occ)[(size_t)(int)x * (unsigned)m.height + (unsigned)(int)y]) return 0.0f;
    if (high > 64) return fl ? lx : 0.0f;
    return -1.0f;
  }
  if (fl) {
    if (v[0] > lx) return max_range;
    if (v[high] < lx) return lx - v[high];
    if (!inside) return __builtin_nanf("");
    if (glb(m.occ)[(size_t)(int)x * (unsigned)m.height + (unsigned)(int)y]) return 0.0f;
    long long lo = 0, hi = n;  // upper_bound (the linear scan's last entry <= x is its predecessor)
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (!(lx < v[mid])) lo = mid + 1; else hi = mid;
    }
    return lx - v[lo - 1];  // lo >= 1: v[0] <= lx
  } else {
    if (v[high] < lx) return max_range;
    if (v[0] > lx) return v[0] - lx;
    if (!inside) return __builtin_nanf("");
    if (glb(m.occ)[(size_t)(int)x * (unsigned)m.height + (unsigned)(int)y]) return 0.0f;
    long long lo = 0, hi = n;
    if (high > 64) {  // upper_bound; past the end only when the last entry equals x: 0, as the scan
      while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (!(lx < v[mid])) lo = mid + 1; else hi = mid;
      }
      return lo == n ? 0.0f : v[lo] - lx;
    }
    while (lo < hi) {  // lower_bound: the linear scan's first entry >= x
      const long long mid = (lo + hi) >> 1;
      if (v[mid] < lx) lo = mid + 1; else hi = mid;
    }
    return v[lo] - lx;  // lo < n: v[high] >= lx
  }
  return -1.0f;  // not reached: the checks above bracket lx within the bin
}

// CDDTCast::calc_range_pair (RangeLib.h:1521-1653): the ranges at the
// heading and at the heading + pi from one bin search.  The reference does not
// check the bin index (a max_range pair here), its unflipped occupancy test
// has no effect (no return: followed), and a flipped occupancy test outside
// the map is NaN here (out of bounds there).  The small-bin linear scans are
// the same binary searches as in cddt_calc_range.
__device__ __forceinline__ void cddt_calc_range_pair(const MapDev& m, float max_range, float x, float y,
                                                     float heading, float* r, float* r_inv) {
  const CddtDev& c = m.cddt;
  bool fl;
  const int a = cddt_bin(c, -heading, &fl);
  if (a < 0) {
    *r = *r_inv = __builtin_nanf("");
    return;
  }
  const float cs = c.cosv[a], sn = c.sinv[a];
  const float lx = x * cs - y * sn;
  const float ly = (x * sn + y * cs) + c.trans[a];
  const unsigned li = (unsigned)trunc_x86(ly);
  if (li >= (unsigned)c.width[a]) {
    *r = *r_inv = max_range;
    return;
  }
  const long long b = c.bin_off[a] + li, s = c.start[b];
  const long long n = c.start[b + 1] - s;
  const RL_GLB float* v = glb(c.vals) + s;
  const long long high = n - 1;
  if (high == -1) {
    *r = *r_inv = max_range;
    return;
  }
  const bool inside = x > -1.0f && x < (float)m.width && y > -1.0f && y < (float)m.height;
  if (fl) {
    if (v[0] > lx) {
      const float d = v[0] - lx;
      *r = max_range;
      *r_inv = (d < max_range) ? d : max_range;
      return;
    }
    if (v[high] < lx) {
      *r = lx - v[high];
      *r_inv = max_range;
      return;
    }
    if (!inside) {
      *r = *r_inv = __builtin_nanf("");
      return;
    }
    if (glb(m.occ)[(size_t)(int)x * (unsigned)m.height + (unsigned)(int)y]) {
      *r = *r_inv = 0.0f;
      return;
    }
    long long lo = 0, hi = n;  // upper_bound: the last entry <= x is lo - 1 (>= 0: v[0] <= x)
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (!(lx < v[mid])) lo = mid + 1; else hi = mid;
    }
    *r = lx - v[lo - 1];
    *r_inv = (lo == n) ? max_range : v[lo] - lx;
  } else {
    if (v[high] < lx) {
      const float d = lx - v[high];
      *r = max_range;
      *r_inv = (d < max_range) ? d : max_range;
      return;
    }
    long long lo = 0, hi = n;  // lower_bound: the first entry >= x (< n: v[high] >= x)
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (v[mid] < lx) lo = mid + 1; else hi = mid;
    }
    *r = v[lo] - lx;
    *r_inv = (lo == 0) ? max_range : lx - v[lo - 1];
  }
}

__device__ __forceinline__ bool fast_march_ok(float x0, float y0, float max_range) {
  return fabsf(x0) < 1.0e9f && fabsf(y0) < 1.0e9f && max_range < 1.0e9f;
}

// OMap::isOccupied (RangeLib.h:204-210): out of bounds reads as free.
__device__ __forceinline__ bool occupied(const uint8_t* __restrict__ occ, int W, int H, int x,
                                         int y) {
  if (x < 0 || (unsigned)x >= (unsigned)W || y < 0 || (unsigned)y >= (unsigned)H) return false;
  return glb(occ)[__umul24((unsigned)x, (unsigned)H) + (unsigned)y] != 0;  // W, H <= 65535: 32-bit
}

#ifndef RL_BL_UNROLL
#define RL_BL_UNROLL 8  // Bresenham steps (cell reads) per round of the walk
#endif

// A Bresenham walk in progress: BresenhamsLine::calc_range's loop state
// (RangeLib.h:727-768) with the walk's constants.
struct BlWalk {
  float x, y, err;           // _x, _y, error
  float deltax, deltay, xstep, ystep;
  float limx;                // bound of _x on the map
  unsigned limy;             // bound of the cell across, (int)floor(_y), on the map
  float x0, y0;              // the walk's start (axes as walked)
  unsigned sx, sy;           // byte offset of the cell at (_x, _y): (int)_x * sx + (int)_y * sy
  int xend, left;            // the loop's end value of (int)_x; steps still allowed
  unsigned bbase, brow;      // bitmap tiles of the walk's orientation: first byte, bytes per row
};

// Byte reads of the occupancy grid through a buffer resource covering it: a
// read past its end returns 0 instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t occ_rsrc(const uint8_t* occ, int W, int H) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)occ, (short)0, (int)((unsigned)W * (unsigned)H), 0x00020000);
}

// BresenhamsLine::calc_range (RangeLib.h:696-769) up to its loop, from
// (x, y) with (s, c) = (sinf h, cosf h): true with *w set when the loop runs,
// else false with the range in *r (an occupied start, a loop that does not
// run, or a walk the reference never ends).
__device__ __forceinline__ bool bl_begin(const MapDev& m, float max_range, float x, float y, float s, float c,
                                         BlWalk* w, float* r) {
  const uint8_t* __restrict__ occ = m.occ;
  const int W = m.width, H = m.height;
  if (occupied(occ, W, H, trunc_x86(x), trunc_x86(y))) {
    *r = 0.0f;
    return false;
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
  const float xstep = (x0 < x1) ? 1.0f : -1.0f;
  const int xend = trunc_x86(x1 + xstep);
  // The reference's loop runs while (int)_x != xend, both as x86 conversions
  // (INT_MIN for NaN or out-of-range values).  It does not run at all when
  // they agree at the start (a NaN origin, an origin beyond int range on the
  // walking axis): max_range.  For end points below 2^23, or an integer start
  // and end points below 2^24 (every float there with a fractional part is
  // beyond reach, and every integer is exact), each step moves (int)_x by one
  // cell, or by two when the step rounds a fractional _x into the next integer
  // at a power of two, and repeats it once while crossing zero; the loop ends
  // unless such a skipped cell is xend (the reference then never returns, and
  // the walk here stops at its step bound, below).  Past that a step can
  // round to no move and the reference never returns; such rays give NaN here.
  if (trunc_x86(x0) == xend) {
    *r = max_range;
    return false;
  }
  const bool exact = (fabsf(x0) < 8388608.0f && fabsf(x1) < 8388608.0f) ||
                     (x0 == __builtin_truncf(x0) && fabsf(x0) < 16777216.0f && fabsf(x1) < 16777214.0f);
  if (!exact) {
    *r = __builtin_nanf("");
    return false;
  }
  const float fw = (float)(unsigned)W;
  const float fh = (float)(unsigned)H;
  // A NaN start across (from a NaN origin) keeps _y NaN, off the map at every
  // step: the same walk as one that starts a cell before the map and never
  // moves across (y0 < y1 is false, so ystep is -1), with _y never NaN.
  const bool ynan = y0 != y0;
  w->x = x0;
  w->y = ynan ? -1.0f : y0;
  w->err = 0.0f;
  w->deltax = fabsf(x1 - x0);
  w->deltay = ynan ? 0.0f : fabsf(y1 - y0);
  w->xstep = xstep;
  w->ystep = (y0 < y1) ? 1.0f : -1.0f;
  // when steep the roles of (x, y) swap back for the map lookup
  w->limx = steep ? fw : fh;
  w->limy = steep ? (unsigned)H : (unsigned)W;
  w->x0 = x0;
  w->y0 = y0;
  w->sx = steep ? (unsigned)H : 1u;
  w->sy = steep ? 1u : (unsigned)H;
  // steep walks step along x: x-major tiles, rows of bits_nty words
  w->bbase = steep ? 0u : m.bits_ybase;
  w->brow = 8u * (unsigned)(steep ? m.bits_nty : m.bits_ntx);
  w->xend = xend;
  // (int)_x moves by at least one cell per step but once while crossing zero,
  // so the loop ends within |xend - (int)x0| + 2 steps, or has skipped xend.
  w->left = abs(xend - trunc_x86(x0)) + 2;
  return true;
}

// RL_BL_UNROLL steps of the walk's loop: true when one of them ends it (then
// *hit says whether on an occupied cell, and (w.x, w.y) is that step's
// position); *nv counts the steps the reference takes.  The cells a walk
// visits do not depend on what it reads, so the round takes its steps, issues
// their cell reads together, and then stops at the first step that ends the
// reference's loop (a hit, or the loop test on the updated _x); the steps
// after it are discarded.  No branch in the body: a step off the map, or past
// the walk's end, reads 0 or a cell that is then ignored.
template <bool COUNT>
__device__ __forceinline__ bool bl_round(__amdgpu_buffer_rsrc_t rsrc, BlWalk& w, bool* hit, int* nv) {
  float xs[RL_BL_UNROLL], ys[RL_BL_UNROLL];
  bool inb[RL_BL_UNROLL], fin[RL_BL_UNROLL];
  unsigned o[RL_BL_UNROLL];
#pragma unroll
  for (int k = 0; k < RL_BL_UNROLL; ++k) {
    w.x += w.xstep;
    w.err += w.deltay;
    // error * 2.00 >= deltax is evaluated in double (RangeLib.h:749); doubling
    // a float is exact, so the float comparison is the same test.
    const bool adv = w.err * 2.0f >= w.deltax;
    w.y = adv ? w.y + w.ystep : w.y;
    w.err = adv ? w.err - w.deltax : w.err;
    // _x stays below 2^24 in magnitude (checked at the start; a few steps past
    // the end only round); _y, off the map, may not (it is never NaN): the
    // cell across is floor(_y), clamped into int range first (the clamp
    // leaves every position on the map unchanged; floor is (int) there)
    const int b = (int)__builtin_floorf(__builtin_amdgcn_fmed3f(w.y, -1.0f, 1073741824.0f));
    inb[k] = ((unsigned)b < w.limy) & (0.0f <= w.x) & (w.x < w.limx);  // no short-circuit branch
    o[k] = __builtin_amdgcn_raw_buffer_load_b8(rsrc, __umul24((unsigned)(int)w.x, w.sx) + __umul24((unsigned)b, w.sy),
                                               0, 0);
    xs[k] = w.x;
    ys[k] = w.y;
    // the reference's loop test after this step, but for the hit
    fin[k] = ((int)w.x == w.xend) | (k + 1 >= w.left);
  }
  // the first step of the round that ends the walk, by selects from the last
  bool stop = false, hk = false;
  float ex = w.x, ey = w.y;
  int n = RL_BL_UNROLL;
#pragma unroll
  for (int k = RL_BL_UNROLL - 1; k >= 0; --k) {
    const bool h = inb[k] & (o[k] != 0u);
    const bool s = h | fin[k];
    ex = s ? xs[k] : ex;
    ey = s ? ys[k] : ey;
    hk = s ? h : hk;
    n = s ? k + 1 : n;
    stop |= s;
  }
  if (COUNT) *nv += n;
  w.x = ex;
  w.y = ey;
  w.left -= RL_BL_UNROLL;
  *hit = hk;
  return stop;
}

// The 2x2-tile window of the occupancy bitmap for the next RL_BL_UNROLL steps
// of walk w (bl_round_bits): tile rows ta, ta + 1 along the walking axis,
// tiles tb, tb + 1 across, starting at the tile of the walk's position and
// extending one tile along each of its step directions (across, not before
// tile 0).  A round moves at most 8 cells along each axis, so the window
// covers every cell of the round on the map.  A tile row holds its two tiles
// side by side: two 16-byte reads.
struct BlWindow {
  uint4 r0, r1;  // tile rows ta and ta + 1
  int a0, ta, tb;
};

__device__ __forceinline__ BlWindow bl_window(__amdgpu_buffer_rsrc_t bits, const BlWalk& w) {
  BlWindow v;
  // the walk stays below 2^24 on the walking axis (bl_begin): an exact floor
  v.a0 = (int)__builtin_floorf(w.x);
  const int b0 = (int)__builtin_floorf(__builtin_amdgcn_fmed3f(w.y, -1.0f, 1073741824.0f));
  v.ta = (v.a0 >> 3) - (w.xstep < 0.0f ? 1 : 0);
  // the first tile across is never before the map's first: a 16-byte read
  // that starts before the buffer is out of range as a whole, and would lose
  // tile 0 with tile -1 (tile 0 then sits in the window's first half; the
  // cells of tile -1 are off the map)
  v.tb = max((b0 >> 3) - (w.ystep < 0.0f ? 1 : 0), 0);
  // a window row before the map's first (each row is a read of its own, at a
  // 32-bit offset that wraps) reads past the buffer (0) or some other tile;
  // only off-map cells fall there
  const unsigned off = w.bbase + (unsigned)v.ta * w.brow + ((unsigned)v.tb << 3);
  v.r0 = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(bits, off, 0, 0));
  v.r1 = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(bits, off + w.brow, 0, 0));
  return v;
}

// bl_round over the occupancy bitmap: the same steps, stops and counts, with
// the round's cells read from one window (bl_window) instead of one byte per
// step.  The window loads are issued first and held there (a scheduling
// barrier: left to itself the compiler sinks them to their first use), so they
// are in flight while the round's steps are computed.  The window is regrouped
// into 16 columns of 16 bits along the walking axis (v_perm) in the lane's
// 32-byte slot of LDS, `win`; a step's column is its cell's (int)_x.
template <bool COUNT>
__device__ __forceinline__ bool bl_round_bits(__amdgpu_buffer_rsrc_t bits, __amdgpu_buffer_rsrc_t occ, BlWalk& w,
                                              uint4* win, bool* hit, int* nv) {
  const BlWindow cur = bl_window(bits, w);
  __builtin_amdgcn_sched_barrier(0);
  const float x_start = w.x, y_start = w.y, err_start = w.err;
  float xs[RL_BL_UNROLL], ys[RL_BL_UNROLL];
  bool inb[RL_BL_UNROLL], fin[RL_BL_UNROLL];
  int ix[RL_BL_UNROLL];
  unsigned nrow[RL_BL_UNROLL], o[RL_BL_UNROLL];
  const int nb = cur.tb << 3;
  // the steps first: they need no reads
#pragma unroll
  for (int k = 0; k < RL_BL_UNROLL; ++k) {
    w.x += w.xstep;
    w.err += w.deltay;
    // error * 2.00 >= deltax is evaluated in double (RangeLib.h:749); doubling
    // a float is exact, so the float comparison is the same test.
    const bool adv = w.err * 2.0f >= w.deltax;
    w.y = adv ? w.y + w.ystep : w.y;
    w.err = adv ? w.err - w.deltax : w.err;
    // the cell across, floor(_y) clamped into int range (the clamp leaves
    // every position on the map unchanged; _y is never NaN), and its bit in
    // the window's column: 0..15 on the map
    const int b = (int)__builtin_floorf(__builtin_amdgcn_fmed3f(w.y, -1.0f, 1073741824.0f));
    inb[k] = ((unsigned)b < w.limy) & (0.0f <= w.x) & (w.x < w.limx);  // no short-circuit branch
    nrow[k] = (unsigned)(b - nb);
    xs[k] = w.x;
    ys[k] = w.y;
    // (int)_x: the reference's loop test after this step (but for the hit),
    // and the step's column on the map, where it is floor(_x)
    ix[k] = (int)w.x;
    fin[k] = (ix[k] == w.xend) | (k + 1 >= w.left);
  }
  __builtin_amdgcn_sched_barrier(0);  // the steps before the window's first use
  // column c (0..15) = bits of walking-axis cell 8 ta + c: the low byte from
  // tile tb, the high byte from tile tb + 1 (bit n: cell 8 tb + n across)
  const uint4 q0 = cur.r0, q1 = cur.r1;
  win[0] = make_uint4(__builtin_amdgcn_perm(q0.z, q0.x, 0x05010400u), __builtin_amdgcn_perm(q0.z, q0.x, 0x07030602u),
                      __builtin_amdgcn_perm(q0.w, q0.y, 0x05010400u), __builtin_amdgcn_perm(q0.w, q0.y, 0x07030602u));
  win[1] = make_uint4(__builtin_amdgcn_perm(q1.z, q1.x, 0x05010400u), __builtin_amdgcn_perm(q1.z, q1.x, 0x07030602u),
                      __builtin_amdgcn_perm(q1.w, q1.y, 0x05010400u), __builtin_amdgcn_perm(q1.w, q1.y, 0x07030602u));
  const uint16_t* col = reinterpret_cast<const uint16_t*>(win);
  const int c8 = 8 * cur.ta;
  unsigned cv[RL_BL_UNROLL];
#pragma unroll
  for (int k = 0; k < RL_BL_UNROLL; ++k) cv[k] = col[(ix[k] - c8) & 15];  // an off-map step: some column
  __builtin_amdgcn_sched_barrier(0);  // the column reads issue together
#pragma unroll
  for (int k = 0; k < RL_BL_UNROLL; ++k) o[k] = __builtin_amdgcn_ubfe(cv[k], nrow[k], 1u);
  // On the map a step moves floor(_x) by one: a step that rounds (a fractional
  // _x crossing a power of two upward) does too, unless the rounding carries
  // into the next integer (a fraction within half a unit in the last place of
  // 1), which moves it by two.  Downward steps on the map never round.  So the
  // round's columns stay in the window (15 at most, the last step's largest)
  // but after such a carry; that round is taken again from its start with
  // byte reads.
  if ((unsigned)(ix[RL_BL_UNROLL - 1] - c8) > 15u) {
    w.x = x_start, w.y = y_start, w.err = err_start;
    return bl_round<COUNT>(occ, w, hit, nv);
  }
  // the first step of the round that ends the walk, by selects from the last
  bool stop = false, hk = false;
  float ex = w.x, ey = w.y;
  int n = RL_BL_UNROLL;
#pragma unroll
  for (int k = RL_BL_UNROLL - 1; k >= 0; --k) {
    const bool h = inb[k] & (o[k] != 0u);
    const bool s = h | fin[k];
    ex = s ? xs[k] : ex;
    ey = s ? ys[k] : ey;
    hk = s ? h : hk;
    n = s ? k + 1 : n;
    stop |= s;
  }
  if (COUNT) *nv += n;
  w.x = ex;
  w.y = ey;
  w.left -= RL_BL_UNROLL;
  *hit = hk;
  return stop;
}

// The range of a walk that ended (RangeLib.h:762-768).
__device__ __forceinline__ float bl_range(const BlWalk& w, bool hit, float max_range) {
  if (hit) {
    const float xd = w.x - w.x0;
    const float yd = w.y - w.y0;
    return __builtin_sqrtf(xd * xd + yd * yd);
  }
  return max_range;
}

// BresenhamsLine::calc_range (RangeLib.h:696-769).  (s, c) = (sinf h, cosf h).
template <bool COUNT>
__device__ __forceinline__ float bl_trace(const MapDev& m, float max_range, float x, float y, float s, float c,
                                          int* visits) {
  BlWalk w;
  float r;
  if (!bl_begin(m, max_range, x, y, s, c, &w, &r)) {
    if (COUNT) *visits = 0;
    return r;
  }
  const __amdgpu_buffer_rsrc_t rsrc = occ_rsrc(m.occ, m.width, m.height);
  int nv = 0;
  bool hit = false;
  for (;;)  // one way out
    if (bl_round<COUNT>(rsrc, w, &hit, &nv)) break;
  if (COUNT) *visits = nv;
  return bl_range(w, hit, max_range);
}

// bl_trace over the occupancy bitmap (m.bits built), with the lane's 32-byte
// window slot in LDS: the same ranges and counts.
template <bool COUNT>
__device__ __forceinline__ float bl_trace_bits(const MapDev& m, uint4* win, float max_range, float x, float y,
                                               float s, float c, int* visits) {
  BlWalk w;
  float r;
  if (!bl_begin(m, max_range, x, y, s, c, &w, &r)) {
    if (COUNT) *visits = 0;
    return r;
  }
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)m.bits, (short)0, (int)(m.bits_bytes + 16u), 0x00020000);
  const __amdgpu_buffer_rsrc_t occ = occ_rsrc(m.occ, m.width, m.height);
  int nv = 0;
  bool hit = false;
  for (;;)  // one way out
    if (bl_round_bits<COUNT>(rsrc, occ, w, win, &hit, &nv)) break;
  if (COUNT) *visits = nv;
  return bl_range(w, hit, max_range);
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

// std::min<float>(std::max<float>(v, 0.0), (float)K - 1.0)  (RangeLib.h:547-550).
// A NaN passes through the reference's min/max and its (int) conversion then
// indexes outside the table (undefined behaviour); here NaN gives K - 1, the
// max-range row/column, so the table read stays in bounds.
__device__ __forceinline__ float clamp_table(float v, float hi) {
  float lo = (v < 0.0f) ? 0.0f : v;
  lo = (hi < lo) ? hi : lo;
  return (v != v) ? hi : lo;
}

}  // namespace rl
