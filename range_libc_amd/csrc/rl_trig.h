// Single-precision sine/cosine that reproduce the host C library's sinf/cosf
// bit for bit, usable from host and device code.
//
// The reference casts rays with glibc's cosf/sinf (RangeLib.h:713-714 for
// Bresenham, RangeLib.h:931-932 for ray marching).  glibc's float routines are
// not correctly rounded: on about 0.2% of inputs with |x| <= 64 their result is
// the float on the far side of the rounding midpoint (the true value always lies within
// ~0.061 ulp of that midpoint).  A one-ulp change in a ray direction moves a
// sample across a cell edge now and then, so the GPU must reproduce glibc's
// choice exactly.
//
// Method (own derivation):
//   1. Evaluate sin/cos in double precision: Cody-Waite reduction by pi/2 with
//      a three-part constant (the first two parts carry 33 significant bits, so
//      n*P1 and n*P2 are exact for |n| < 2^20), then Taylor polynomials of
//      degree 17 (sin) and 18 (cos) on |r| <= pi/4 + eps, evaluated in Horner
//      form.  Truncation error < 1e-19, so the double result is within a few
//      ulp(double) of the true value.  Only IEEE +,-,*,/ and rint are used and
//      the file must be built with contraction off, so host and device produce
//      identical doubles.
//   2. Round to float.  If the double lies within `delta` of the midpoint
//      between the two candidate floats, consult an exception set listing
//      the inputs for which glibc returns the other candidate.  The set and
//      delta are produced by exhaustively enumerating every float x with
//      |x| <= 64 against the C library at build time (gen_trig_table.cpp),
//      stored compactly (sorted keys per 4096-value bucket) and expanded at
//      load time into a flat bitmap so that the device probes it with a
//      single load (exceptions cluster: up to ~350 per bucket).
//
// Outside |x| <= 64 the result is the double value rounded to nearest.  The C
// library (glibc 2.35) rounds the other way on 1.3% of such inputs for sinf and
// 1.3% for cosf (2.6% of inputs differ in one of the two; 2e7 samples, uniform
// over the float encodings of 64 < |x| < 1.6e6 and uniform in value over
// [64, 1000]); the result is never more than one float away while the
// reduction is exact, i.e. for |x| < 2^20 * pi/2 (about 1.6e6).  Beyond that
// n * PIO2_1 is no longer exact and the result can differ further.  Bit
// parity with the reference therefore holds for ray headings |h| <= 64.
#pragma once
#include <stdint.h>
#include <math.h>
#include <string.h>

#ifdef __HIPCC__
#define RL_HD __host__ __device__ __forceinline__
#else
#define RL_HD static inline
#endif

#pragma clang fp contract(off)

namespace rltrig {

// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (to ~2^-122)
#define RLT_PIO2_1 0x1.921fb54400000p+0
#define RLT_PIO2_2 0x1.0b4611a600000p-34
#define RLT_PIO2_3 0x1.3198a2e037073p-69
#define RLT_TWO_OVER_PI 0.6366197723675814

// Exception-table geometry: a bucket holds 4096 consecutive float encodings of
// one sign; buckets cover |x| <= 64.0f (encoding 0x42800000).
#define RLT_ABS_LIMIT 0x42800000u
#define RLT_BUCKET_SHIFT 12
#define RLT_SIGN_SHIFT 19
#define RLT_NBUCKETS (1u << 20)

struct Table {
  const uint32_t* off;   // RLT_NBUCKETS + 1 offsets into keys
  const uint16_t* keys;  // low 12 bits of |x|'s encoding, sorted per bucket
  double delta;          // probe only when |v - midpoint| / gap <= delta
};

RL_HD uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
RL_HD float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// sin(x) and cos(x) in double for float x.
RL_HD void sincos_d(float xf, double* s_out, double* c_out) {
  double x = (double)xf;
  if (!(fabs(x) < 1.0e9)) {  // inf, nan, or absurdly large headings
    if (x != x || fabs(x) == INFINITY) {
      *s_out = NAN;
      *c_out = NAN;
      return;
    }
  }
  if (x == 0.0) {  // keeps the sign of a zero input: sin(-0) = -0
    *s_out = x;
    *c_out = 1.0;
    return;
  }
  double n = rint(x * RLT_TWO_OVER_PI);
  double r = x;
  if (n != 0.0) {
    r = x - n * RLT_PIO2_1;
    r = r - n * RLT_PIO2_2;
    r = r - n * RLT_PIO2_3;
  }
  double r2 = r * r;
  // sin r = r + r^3 * P(r^2), P = -1/3! + r^2/5! - ...  (Taylor, to r^17: on
  // |r| <= pi/4 the first dropped term, r^19 / 19!, is below 1e-19)
  double ps = 2.8114572543455206e-15;
  ps = ps * r2 - 7.647163731819816e-13;
  ps = ps * r2 + 1.6059043836821613e-10;
  ps = ps * r2 - 2.505210838544172e-08;
  ps = ps * r2 + 2.7557319223985893e-06;
  ps = ps * r2 - 0.0001984126984126984;
  ps = ps * r2 + 0.008333333333333333;
  ps = ps * r2 - 0.16666666666666666;
  double sr = r + (r * r2) * ps;
  // cos r = 1 - r^2/2 + r^4 * Q(r^2), Q = 1/4! - r^2/6! + ...  (Taylor, to r^18:
  // the first dropped term, r^20 / 20!, is below 1e-20)
  double pc = -1.5619206968586225e-16;
  pc = pc * r2 + 4.779477332387385e-14;
  pc = pc * r2 - 1.1470745597729725e-11;
  pc = pc * r2 + 2.08767569878681e-09;
  pc = pc * r2 - 2.755731922398589e-07;
  pc = pc * r2 + 2.48015873015873e-05;
  pc = pc * r2 - 0.001388888888888889;
  pc = pc * r2 + 0.041666666666666664;
  double cr = (1.0 - 0.5 * r2) + (r2 * r2) * pc;
  int q = (int)(((long long)n) & 3);
  double s, c;
  if (q == 0) {
    s = sr;
    c = cr;
  } else if (q == 1) {
    s = cr;
    c = -sr;
  } else if (q == 2) {
    s = -sr;
    c = -cr;
  } else {
    s = -cr;
    c = sr;
  }
  *s_out = s;
  *c_out = c;
}

// Round v to float the way the C library does for input x: nearest unless x
// is a listed exception, in which case the far neighbour.
// `dist_out` (optional) receives |v - midpoint| / gap, 0.5 when exact.
RL_HD float neighbours(double v, float* other_out, double* dist_out) {
  float f = (float)v;
  double fd = (double)f;
  if (fd == v) {
    *other_out = f;
    *dist_out = 0.5;
    return f;
  }
  bool up = v > fd;
  uint32_t fb = f2u(f);
  uint32_t ob;
  if ((fb & 0x7fffffffu) == 0u) {
    ob = up ? 0x00000001u : 0x80000001u;
  } else if ((fb >> 31) == 0u) {
    ob = up ? fb + 1u : fb - 1u;
  } else {
    ob = up ? fb - 1u : fb + 1u;
  }
  float o = u2f(ob);
  double od = (double)o;
  double m = 0.5 * (fd + od);
  double gap = fabs(od - fd);
  *other_out = o;
  *dist_out = fabs(v - m) / gap;
  return f;
}

RL_HD bool table_has(const Table& t, float x) {
  uint32_t xb = f2u(x);
  uint32_t ab = xb & 0x7fffffffu;
  if (ab > RLT_ABS_LIMIT) return false;
  uint32_t b = (ab >> RLT_BUCKET_SHIFT) | ((xb >> 31) << RLT_SIGN_SHIFT);
  uint32_t lo = t.off[b], hi = t.off[b + 1];
  uint16_t key = (uint16_t)(ab & ((1u << RLT_BUCKET_SHIFT) - 1u));
  for (uint32_t i = lo; i < hi; ++i) {
    uint16_t k = t.keys[i];
    if (k == key) return true;
    if (k > key) break;
  }
  return false;
}

RL_HD float round_like_libc(double v, float x, const Table& t) {
  float o;
  double dist;
  float f = neighbours(v, &o, &dist);
  if (dist <= t.delta && table_has(t, x)) return o;
  return f;
}

// cosf(h) and sinf(h) as the C library computes them (bucketed table: a
// scan of sorted keys; used on the host).
RL_HD void sincosf_libc(float h, const Table& ts, const Table& tc, float* s, float* c) {
  double sd, cd;
  sincos_d(h, &sd, &cd);
  *s = round_like_libc(sd, h, ts);
  *c = round_like_libc(cd, h, tc);
}

// The same exception set as a flat bitmap: one bit per float encoding of
// each sign whose exponent field lies in [e_min, 133] (133 = exponent of
// 64.0f), so a lookup is a single byte load.  e_min is the smallest exponent
// field of any listed input.
#define RLT_EXP_MAX 133
struct Bitmap {
  const uint8_t* bits;
  int e_min;
  double delta;
};

RL_HD int bitmap_binades(int e_min) { return RLT_EXP_MAX - e_min + 1; }

RL_HD unsigned long long bitmap_bytes(int e_min) {
  return (2ull * (unsigned long long)bitmap_binades(e_min)) << 20;  // 2^23 bits per binade
}

// Bit index of x, or -1 when x is outside the covered encodings.
RL_HD long long bitmap_index(int e_min, float x) {
  uint32_t xb = f2u(x);
  uint32_t ab = xb & 0x7fffffffu;
  if (ab > RLT_ABS_LIMIT) return -1;
  int e = (int)(ab >> 23);
  if (e < e_min) return -1;
  unsigned long long slot = (unsigned long long)(xb >> 31) * (unsigned long long)bitmap_binades(e_min) +
                            (unsigned long long)(e - e_min);
  return (long long)((slot << 23) | (ab & 0x7fffffu));
}

// neighbours() without its division: returns the nearest float f and the
// other candidate, and whether |v - midpoint| / gap <= delta.  gap, the
// spacing of two adjacent floats, is a power of two, so |v - m| <= delta * gap
// is exactly that test.
RL_HD float neighbours_near(double v, double delta, float* other_out, bool* near_out) {
  float f = (float)v;
  double fd = (double)f;
  if (fd == v) {
    *other_out = f;
    *near_out = 0.5 <= delta;
    return f;
  }
  bool up = v > fd;
  uint32_t fb = f2u(f);
  uint32_t ob;
  if ((fb & 0x7fffffffu) == 0u) {
    ob = up ? 0x00000001u : 0x80000001u;
  } else if ((fb >> 31) == 0u) {
    ob = up ? fb + 1u : fb - 1u;
  } else {
    ob = up ? fb - 1u : fb + 1u;
  }
  float o = u2f(ob);
  double od = (double)o;
  double m = 0.5 * (fd + od);
  *other_out = o;
  *near_out = fabs(v - m) <= delta * fabs(od - fd);
  return f;
}

RL_HD void sincosf_libc_bm(float h, const Bitmap& bs, const Bitmap& bc, float* s, float* c) {
  double sd, cd;
  sincos_d(h, &sd, &cd);
  float os, oc;
  bool ns, nc;
  float fs = neighbours_near(sd, bs.delta, &os, &ns);
  float fc = neighbours_near(cd, bc.delta, &oc, &nc);
  long long is = ns ? bitmap_index(bs.e_min, h) : -1;
  long long ic = nc ? bitmap_index(bc.e_min, h) : -1;
  // both probes are issued before either is consumed
  uint32_t vs = is >= 0 ? (uint32_t)bs.bits[is >> 3] : 0u;
  uint32_t vc = ic >= 0 ? (uint32_t)bc.bits[ic >> 3] : 0u;
  *s = (is >= 0 && ((vs >> (is & 7)) & 1u)) ? os : fs;
  *c = (ic >= 0 && ((vc >> (ic & 7)) & 1u)) ? oc : fc;
}

// Host-side helpers (plain host functions under hipcc).
// Smallest exponent field among the listed inputs of a bucketed table (RLT_EXP_MAX + 1 if none).
static inline int table_min_exponent(const uint32_t* off, const uint16_t* keys) {
  int e_min = RLT_EXP_MAX + 1;
  for (uint32_t b = 0; b < RLT_NBUCKETS; ++b)
    for (uint32_t i = off[b]; i < off[b + 1]; ++i) {
      uint32_t ab = ((b & ((1u << RLT_SIGN_SHIFT) - 1u)) << RLT_BUCKET_SHIFT) | keys[i];
      int e = (int)(ab >> 23);
      if (e < e_min) e_min = e;
      break;  // keys ascend within a bucket: the first has the smallest exponent
    }
  return e_min;
}

// Sets the bits of every listed input; `bits` holds bitmap_bytes(e_min) zeroed bytes.
static inline void table_to_bitmap(const uint32_t* off, const uint16_t* keys, int e_min, uint8_t* bits) {
  for (uint32_t b = 0; b < RLT_NBUCKETS; ++b)
    for (uint32_t i = off[b]; i < off[b + 1]; ++i) {
      uint32_t xb = (((b & ((1u << RLT_SIGN_SHIFT) - 1u)) << RLT_BUCKET_SHIFT) | keys[i]) |
                    ((b >> RLT_SIGN_SHIFT) << 31);
      long long idx = bitmap_index(e_min, u2f(xb));
      if (idx >= 0) bits[idx >> 3] |= (uint8_t)(1u << (idx & 7));
    }
}

}  // namespace rltrig
