/* CPU oracle for the batched ray-casting path -- TEST INFRASTRUCTURE ONLY.
 * See rl_oracle.h for scope and pinning.  Restates, op for op:
 *   includes/RangeLib.h            345-373, 439-520, 533-612, 696-769, 927-968
 *   vendor/distance_transform.h    873-919, 1054-1096, 1117-1121
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math.
 */
#include "rl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------------
 * Distance transform.
 * distance_transform.h:1054-1096 (distanceL2): lower envelope of parabolas for
 * one line.  The types follow the reference: f values are float, the
 * parabola intersection `s` is double, `q*q` is an integer converted to float
 * when added to f[q], and D[q] is evaluated in float.
 */
static void dl2_line(const float* f, ptrdiff_t fs, float* D, ptrdiff_t ds, size_t n, size_t* v,
                     double* z) {
  if (n == 0) return;
  if (n == 1) { /* :1061-1063 */
    D[0] = f[0];
    return;
  }
  size_t k = 0;
  double s = 0.0;
  v[0] = 0;
  z[0] = -DBL_MAX;
  z[1] = DBL_MAX;
  for (size_t q = 1; q < n; ++q) { /* :1071-1083 */
    ++k;
    do {
      --k;
      float lhs = f[q * fs] + (float)(q * q);
      float rhs = f[v[k] * fs] + (float)(v[k] * v[k]);
      s = ((double)lhs - (double)rhs) / ((double)(2 * q) - (double)(2 * v[k]));
    } while (s <= z[k]);
    ++k;
    v[k] = q;
    z[k] = s;
    z[k + 1] = DBL_MAX;
  }
  k = 0;
  for (size_t q = 0; q < n; ++q) { /* :1085-1090 */
    while (z[k + 1] < (double)q) ++k;
    float dq = (float)q - (float)v[k];
    D[q * ds] = f[v[k] * fs] + dq * dq;
  }
}

/* RangeLib.h:345-373 -> distance_transform.h:873-919: f = 0 on occupied cells
 * and FLT_MAX elsewhere; pass 1 runs along y for every x, pass 2 along x for
 * every y; then an elementwise float square root (:1117-1121). */
void or_edt(const uint8_t* occ, int width, int height, float* out) {
  size_t W = (size_t)width, H = (size_t)height, N = W * H;
  float* f = (float*)malloc(N * sizeof(float));
  float* g = (float*)malloc(N * sizeof(float));
  size_t nmax = W > H ? W : H;
  size_t* v = (size_t*)malloc((nmax + 1) * sizeof(size_t));
  double* z = (double*)malloc((nmax + 2) * sizeof(double));
  for (size_t i = 0; i < N; ++i) f[i] = occ[i] ? 0.0f : FLT_MAX;
  for (size_t x = 0; x < W; ++x) dl2_line(f + x * H, 1, g + x * H, 1, H, v, z);
  for (size_t y = 0; y < H; ++y) dl2_line(g + y, (ptrdiff_t)H, out + y, (ptrdiff_t)H, W, v, z);
  for (size_t i = 0; i < N; ++i) out[i] = sqrtf(out[i]);
  free(f);
  free(g);
  free(v);
  free(z);
}

/* ---------------------------------------------------------------------------
 * RayMarching::calc_range (RangeLib.h:927-962) with distThreshold = 0
 * (:967) and step_coeff = 0.999f (:968). */
float or_rm_calc_range(const float* edt, int width, int height, float max_range, float x, float y,
                       float heading, int* reads) {
  const float step_coeff = 0.999f;
  const float dist_threshold = 0.0f;
  float x0 = x, y0 = y;
  float dx = cosf(heading);
  float dy = sinf(heading);
  int px = 0, py = 0;
  float t = 0.0f;
  int nr = 0;
  while (t < max_range) {
    px = (int)(x0 + dx * t);
    py = (int)(y0 + dy * t);
    if ((unsigned)px >= (unsigned)width || px < 0 || py < 0 || (unsigned)py >= (unsigned)height) {
      if (reads) *reads = nr;
      return max_range;
    }
    float d = edt[(size_t)px * (size_t)height + (size_t)py];
    ++nr;
    if (d <= dist_threshold) {
      float xd = (float)px - x0;
      float yd = (float)py - y0;
      if (reads) *reads = nr;
      return sqrtf(xd * xd + yd * yd);
    }
    float st = d * step_coeff;
    t += (st < 1.0f) ? 1.0f : st; /* std::max<float>(d*step_coeff, 1.0) */
  }
  if (reads) *reads = nr;
  return max_range;
}

/* ---------------------------------------------------------------------------
 * BresenhamsLine::calc_range (RangeLib.h:696-769). */
static int occupied(const uint8_t* occ, int width, int height, int x, int y) {
  /* OMap::isOccupied (RangeLib.h:204-210): out of bounds reads as free */
  if (x < 0 || (unsigned)x >= (unsigned)width || y < 0 || (unsigned)y >= (unsigned)height) return 0;
  return occ[(size_t)x * (size_t)height + (size_t)y] != 0;
}

float or_bl_calc_range(const uint8_t* occ, int width, int height, float max_range, float x,
                       float y, float heading, int* visits) {
  int nv = 0;
  if (occupied(occ, width, height, (int)x, (int)y)) {
    if (visits) *visits = 0;
    return 0.0f;
  }
  float x0 = y;
  float y0 = x;
  float x1 = y + max_range * sinf(heading);
  float y1 = x + max_range * cosf(heading);
  int steep = fabsf(y1 - y0) > fabsf(x1 - x0);
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
  float deltaerr = deltay;
  float _x = x0;
  float _y = y0;
  int xstep = (x0 < x1) ? 1 : -1;
  int ystep = (y0 < y1) ? 1 : -1;
  float fw = (float)(unsigned)width;  /* float < unsigned compares in float */
  float fh = (float)(unsigned)height;
  while ((int)_x != (int)(x1 + (float)xstep)) {
    _x += (float)xstep;
    error += deltaerr;
    ++nv;
    if ((double)error * 2.00 >= (double)deltax) {
      _y += (float)ystep;
      error -= deltax;
    }
    if (!steep) {
      if (0 <= _y && _y < fw && 0 <= _x && _x < fh && occupied(occ, width, height, (int)_y, (int)_x)) {
        float xd = _x - x0;
        float yd = _y - y0;
        if (visits) *visits = nv;
        return sqrtf(xd * xd + yd * yd);
      }
    } else {
      if (0 <= _x && _x < fw && 0 <= _y && _y < fh && occupied(occ, width, height, (int)_x, (int)_y)) {
        float xd = _x - x0;
        float yd = _y - y0;
        if (visits) *visits = nv;
        return sqrtf(xd * xd + yd * yd);
      }
    }
  }
  if (visits) *visits = nv;
  return max_range;
}

/* ---------------------------------------------------------------------------
 * GiantLUTCast (RangeLib.h:1772-1900) with _GIANT_LUT_SHORT_DATATYPE,
 * _USE_CACHED_CONSTANTS and _USE_ALTERNATE_MOD set, _USE_FAST_ROUND clear
 * (RangeLib.h:66-75).  M_2PI is the double literal of RangeLib.h:61. */
#define OR_M_2PI 6.28318530718

void or_glt_build(const float* edt, int width, int height, float max_range, int td, uint16_t* table,
                  float* consts) {
  /* RangeLib.h:1783-1786: int / double, double / float, float / int, int / float */
  const float td_div_2pi = (float)((double)td / OR_M_2PI);
  const float step = (float)(OR_M_2PI / (double)(float)td);
  const float max_div_limits = max_range / (float)65535;
  const float limits_div_max = (float)65535 / max_range;
  for (int x = 0; x < width; ++x)
    for (int y = 0; y < height; ++y)
      for (int i = 0; i < td; ++i) {
        const float angle = (float)i * step;                                   /* :1796 */
        float r = or_rm_calc_range(edt, width, height, max_range, (float)x, (float)y, angle, NULL);
        r = (r < max_range) ? r : max_range;                                   /* std::min, :1803 */
        table[((size_t)x * height + y) * td + i] = (uint16_t)(r * limits_div_max); /* :1805 */
      }
  if (consts) {
    consts[0] = max_div_limits;
    consts[1] = limits_div_max;
    consts[2] = td_div_2pi;
  }
}

int or_glt_discretize(float theta, int td, float td_div_2pi) {
  if (theta != theta) return -1;
  if (theta < 0.0) {                                                             /* :1834-1837 */
    while (theta < 0.0) {
      float next = (float)((double)theta + OR_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  } else if (theta > OR_M_2PI) {                                                 /* :1838-1841 */
    while (theta > OR_M_2PI) {
      float next = (float)((double)theta - OR_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  }
  int rounded = (int)roundf(theta * td_div_2pi);                                /* :1854 */
  return rounded % td;                                                           /* :1858 */
}

float or_glt_calc_range(const or_method* m, float x, float y, float heading) {
  /* RangeLib.h:1866-1870 */
  if (x < 0 || x >= (float)m->width || y < 0 || y >= (float)m->height) return m->max_range;
  if (x != x || y != y) return NAN; /* the reference indexes its table at (int)NaN */
  int b = or_glt_discretize(heading, m->glt_td, m->glt_td_div_2pi);
  if (b < 0) return NAN;
  return (float)m->glt[((size_t)(int)x * m->height + (int)y) * m->glt_td + b] * m->glt_max_div_limits;
}

float or_calc_range(const or_method* m, float x, float y, float heading) {
  if (m->kind == 1)
    return or_bl_calc_range(m->occ, m->width, m->height, m->max_range, x, y, heading, NULL);
  if (m->kind == 2) return or_glt_calc_range(m, x, y, heading);
  return or_rm_calc_range(m->edt, m->width, m->height, m->max_range, x, y, heading, NULL);
}

static float calc_range_counted(const or_method* m, float x, float y, float heading, int* cnt) {
  if (m->kind == 1) return or_bl_calc_range(m->occ, m->width, m->height, m->max_range, x, y, heading, cnt);
  if (m->kind == 2) {
    if (cnt) *cnt = (x < 0 || x >= (float)m->width || y < 0 || y >= (float)m->height) ? 0 : 1;
    return or_glt_calc_range(m, x, y, heading);
  }
  return or_rm_calc_range(m->edt, m->width, m->height, m->max_range, x, y, heading, cnt);
}

void or_calc_range_many(const or_method* m, const float* ins, float* outs, int n, int* reads) {
  for (int i = 0; i < n; ++i) {
    int c = 0;
    outs[i] = calc_range_counted(m, ins[3 * i], ins[3 * i + 1], ins[3 * i + 2], &c);
    if (reads) reads[i] = c;
  }
}

/* ---------------------------------------------------------------------------
 * World-to-grid batch drivers (ROS_WORLD_TO_GRID_CONVERSION == 1). */
typedef struct {
  float inv_world_scale, world_scale, origin_x, origin_y, sin_a, cos_a, rotation_const;
} world_consts;

static world_consts make_consts(const or_world* w) { /* RangeLib.h:442-450 */
  world_consts c;
  c.inv_world_scale = (float)(1.0 / (double)w->scale);
  c.world_scale = w->scale;
  c.origin_x = w->origin_x;
  c.origin_y = w->origin_y;
  c.sin_a = w->sin_angle;
  c.cos_a = w->cos_angle;
  c.rotation_const = (float)(-1.0 * (double)w->angle - 3.0 * M_PI / 2.0);
  return c;
}

static void to_grid(const world_consts* c, float xw, float yw, float* gx, float* gy) {
  float x = (xw - c->origin_x) * c->inv_world_scale; /* RangeLib.h:468-472 */
  float y = (yw - c->origin_y) * c->inv_world_scale;
  float temp = x;
  x = c->cos_a * x - c->sin_a * y;
  y = c->sin_a * temp + c->cos_a * y;
  *gx = x;
  *gy = y;
}

void or_numpy_calc_range(const or_method* m, const float* ins, float* outs, int n, int* reads) {
  world_consts c = make_consts(&m->world);
  for (int i = 0; i < n; ++i) { /* RangeLib.h:463-476 */
    float x, y;
    to_grid(&c, ins[i * 3], ins[i * 3 + 1], &x, &y);
    float theta = -ins[i * 3 + 2] + c.rotation_const;
    int cnt = 0;
    outs[i] = calc_range_counted(m, y, x, theta, &cnt) * c.world_scale;
    if (reads) reads[i] = cnt;
  }
}

void or_numpy_calc_range_angles(const or_method* m, const float* ins, const float* angles,
                                float* outs, int num_particles, int num_angles, int* reads) {
  world_consts c = make_consts(&m->world);
  for (int i = 0; i < num_particles; ++i) { /* RangeLib.h:506-519 */
    float theta = -ins[i * 3 + 2] + c.rotation_const;
    float x, y;
    to_grid(&c, ins[i * 3], ins[i * 3 + 1], &x, &y);
    for (int a = 0; a < num_angles; ++a) {
      int cnt = 0;
      outs[i * num_angles + a] = calc_range_counted(m, y, x, theta - angles[a], &cnt) * c.world_scale;
      if (reads) reads[i * num_angles + a] = cnt;
    }
  }
}

/* std::min<float>(std::max<float>(v, 0.0), (float)K - 1.0) */
static float clamp_table(float v, int K) {
  float lo = (v < 0.0f) ? 0.0f : v;
  float hi = (float)((double)(float)K - 1.0);
  return (hi < lo) ? hi : lo;
}

void or_eval_sensor_model(const or_method* m, const float* obs, const float* ranges, double* outs,
                          int rays_per_particle, int particles) {
  float inv_world_scale = (float)(1.0 / (double)m->world.scale); /* RangeLib.h:534 */
  int K = m->table_width;
  for (int i = 0; i < particles; ++i) { /* :542-554 */
    double weight = 1.0;
    for (int j = 0; j < rays_per_particle; ++j) {
      float r = clamp_table(obs[j] * inv_world_scale, K);
      float d = clamp_table(ranges[i * rays_per_particle + j] * inv_world_scale, K);
      weight *= m->table[(size_t)(int)r * (size_t)K + (size_t)(int)d];
    }
    outs[i] = weight;
  }
}

void or_calc_range_repeat_angles_eval_sensor_model(const or_method* m, const float* ins,
                                                   const float* angles, const float* obs,
                                                   double* weights, int num_particles,
                                                   int num_angles) {
  world_consts c = make_consts(&m->world);
  int K = m->table_width;
  for (int i = 0; i < num_particles; ++i) { /* RangeLib.h:587-610 */
    float theta = -ins[i * 3 + 2] + c.rotation_const;
    float x, y;
    to_grid(&c, ins[i * 3], ins[i * 3 + 1], &x, &y);
    double weight = 1.0;
    for (int a = 0; a < num_angles; ++a) {
      /* the range stays in grid units here: no world_scale (:602-603) */
      float d = clamp_table(or_calc_range(m, y, x, theta - angles[a]), K);
      float r = clamp_table(obs[a] * c.inv_world_scale, K);
      weight *= m->table[(size_t)(int)r * (size_t)K + (size_t)(int)d];
    }
    weights[i] = weight;
  }
}

void or_sinf_many(const float* x, float* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = sinf(x[i]);
}
void or_cosf_many(const float* x, float* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = cosf(x[i]);
}
