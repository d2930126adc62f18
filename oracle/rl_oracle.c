/* CPU oracle for the batched ray-casting path -- TEST INFRASTRUCTURE ONLY.
 * See rl_oracle.h for scope and pinning.  Restates, op for op:
 *   includes/RangeLib.h            345-373, 439-520, 533-612, 696-769, 927-968
 *   vendor/distance_transform.h    873-919, 1054-1096, 1117-1121
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math.
 */
#include "rl_oracle.h"

#include <float.h>
#include <limits.h>
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
  if (m->kind == 3) return or_cddt_calc_range(m->cddt, m->occ, m->width, m->height, m->max_range, x, y, heading);
  return or_rm_calc_range(m->edt, m->width, m->height, m->max_range, x, y, heading, NULL);
}

static float calc_range_counted(const or_method* m, float x, float y, float heading, int* cnt) {
  if (m->kind == 1) return or_bl_calc_range(m->occ, m->width, m->height, m->max_range, x, y, heading, cnt);
  if (m->kind == 2) {
    if (cnt) *cnt = (x < 0 || x >= (float)m->width || y < 0 || y >= (float)m->height) ? 0 : 1;
    return or_glt_calc_range(m, x, y, heading);
  }
  if (m->kind == 3) {
    if (cnt) *cnt = 0;
    return or_cddt_calc_range(m->cddt, m->occ, m->width, m->height, m->max_range, x, y, heading);
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

/* std::min<float>(std::max<float>(v, 0.0), (float)K - 1.0) (RangeLib.h:547-550).
 * A NaN passes through the reference's min/max and its (int) conversion then
 * indexes outside the table (undefined behaviour); here NaN gives K - 1, the
 * max-range row/column, as a no-return beam reads. */
static float clamp_table(float v, int K) {
  float hi = (float)((double)(float)K - 1.0);
  if (v != v) return hi;
  float lo = (v < 0.0f) ? 0.0f : v;
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

/* ---------------------------------------------------------------------------
 * CDDTCast (RangeLib.h:971-1770): compressed directional distance transform,
 * with _USE_CACHED_TRIG clear, _USE_CACHED_CONSTANTS and _USE_ALTERNATE_MOD
 * set, _EPSILON 1e-5 and _BINARY_SEARCH_THRESHOLD 64 (RangeLib.h:60-67).  The
 * bins of angle a are compressed_lut[a][i]; stored here as one CSR array. */
#define OR_EPS 0.00001
#define OR_BSEARCH 64
#define OR_M_PI 3.14159265358979323846

static int cmp_float(const void* a, const void* b) {
  float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}

static int edge_cell(const uint8_t* occ, int w, int h, int x, int y) {
  /* OMap::make_edge_map(true) (RangeLib.h:293-310), utils::outline (RangeUtils.h:36-52) */
  static const int dx[8] = {1, -1, 0, 0, 1, -1, 1, -1}, dy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
  if (!occ[(size_t)x * h + y]) return 0;
  for (int k = 0; k < 8; ++k) {
    int cx = x + dx[k], cy = y + dy[k];
    if (0 <= cx && 0 <= cy && cx < w && cy < h && !occ[(size_t)cx * h + cy]) return 1;
  }
  return 0;
}

or_cddt* or_cddt_build(const uint8_t* occ, int w, int h, int td) {
  or_cddt* c = (or_cddt*)calloc(1, sizeof(or_cddt));
  c->td = td;
  c->nang = 0;
  while (c->nang < td && c->nang < td / 2.0) c->nang++; /* a < theta_discretization / 2.0 */
  c->td_div_2pi = (float)((double)td / OR_M_2PI);              /* RangeLib.h:976 */
  c->m2pi_div_td = (float)(OR_M_2PI / (double)(float)td);      /* :977 */
  c->width = (int*)calloc(td, sizeof(int));
  c->trans = (float*)calloc(td, sizeof(float));
  c->cosv = (float*)calloc(td, sizeof(float));
  c->sinv = (float*)calloc(td, sizeof(float));
  c->bin_off = (long long*)calloc(td + 1, sizeof(long long));
  for (int i = 0; i < td; ++i) {                                /* :991-1058 */
    float angle = (float)i * c->m2pi_div_td;
    float cs = cosf(angle), sn = sinf(angle);
    c->cosv[i] = cs;
    c->sinv[i] = sn;
    float rotated_height = fabsf((float)w * sn) + fabsf((float)h * cs);
    c->width[i] = (int)(unsigned)ceil((double)rotated_height - OR_EPS);
    float lt = (float)h * cs, rt = (float)w * sn + (float)h * cs, rb = (float)w * sn;
    float m2 = (rb < rt) ? rb : rt;                              /* std::min(rt, rb) */
    float mn = (m2 < lt) ? m2 : lt;                              /* std::min(lt, ...) */
    double t = -1.0 * (double)mn - OR_EPS;
    c->trans[i] = (float)(t > 0.0 ? t : 0.0);                    /* std::max(0.0, t) */
    c->bin_off[i + 1] = c->bin_off[i] + c->width[i];
  }
  const long long nb = c->bin_off[td];
  long long* cnt = (long long*)calloc(nb + 1, sizeof(long long));
  /* two passes over the edge cells (x-major, as the reference): count, then fill */
  for (int pass = 0; pass < 2; ++pass) {
    for (int x = 0; x < w; ++x)
      for (int y = 0; y < h; ++y) {
        if (!edge_cell(occ, w, h, x, y)) continue;
        float px = (float)(x + 0.5), py = (float)(y + 0.5);     /* :1098 */
        for (int a = 0; a < c->nang; ++a) {
          float cs = c->cosv[a], sn = c->sinv[a];
          float half = (float)((double)(fabsf(sn) + fabsf(cs)) / 2.0);
          float cx = px * cs - py * sn;
          float cy = (px * sn + py * cs) + c->trans[a];
          int upper = (int)((double)(cy + half) - OR_EPS);
          int lower = (int)((double)(cy - half) + OR_EPS);
          for (int i = lower; i <= upper; ++i) {
            if (i < 0 || i >= c->width[a]) continue; /* outside the table: the reference writes out of bounds */
            long long b = c->bin_off[a] + i;
            if (pass == 0)
              cnt[b + 1]++;
            else
              c->vals[cnt[b]++] = cx;
          }
        }
      }
    if (pass == 0) {
      for (long long b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
      c->start = (long long*)malloc((nb + 1) * sizeof(long long));
      memcpy(c->start, cnt, (nb + 1) * sizeof(long long));
      c->vals = (float*)malloc((cnt[nb] > 0 ? cnt[nb] : 1) * sizeof(float));
    }
  }
  free(cnt);
  /* sort each bin, drop duplicates, compact (RangeLib.h:1131-1141) */
  long long out = 0;
  for (long long b = 0; b < nb; ++b) {
    long long s = c->start[b], e = c->start[b + 1];
    qsort(c->vals + s, (size_t)(e - s), sizeof(float), cmp_float);
    c->start[b] = out;
    for (long long k = s; k < e; ++k)
      if (k == s || c->vals[k] != c->vals[k - 1]) c->vals[out++] = c->vals[k];
  }
  c->start[nb] = out;
  return c;
}

void or_cddt_free(or_cddt* c) {
  if (!c) return;
  free(c->width);
  free(c->trans);
  free(c->cosv);
  free(c->sinv);
  free(c->bin_off);
  free(c->start);
  free(c->vals);
  free(c);
}

/* first index with v[i] >= x (lower_bound) / v[i] > x (upper_bound) */
static long long lower_b(const float* v, long long n, float x) {
  long long lo = 0, hi = n;
  while (lo < hi) {
    long long mid = (lo + hi) / 2;
    if (v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
static long long upper_b(const float* v, long long n, float x) {
  long long lo = 0, hi = n;
  while (lo < hi) {
    long long mid = (lo + hi) / 2;
    if (!(x < v[mid])) lo = mid + 1; else hi = mid;
  }
  return lo;
}

void or_cddt_prune(or_cddt* c, const uint8_t* occ, int w, int h, float max_range) {
  /* CDDTCast::prune (RangeLib.h:1172-1275).  Where the reference's linear
   * search leaves `index` unset (no entry >= the cell), the lower_bound value
   * (the bin size) is taken, as the binary search would. */
  const long long nb = c->bin_off[c->td];
  uint8_t* used = (uint8_t*)calloc(c->start[nb] + 1, 1);
  for (int a = 0; a < c->nang; ++a) {
    float cs = c->cosv[a], sn = c->sinv[a], tr = c->trans[a];
    for (int x = 0; x < w; ++x) {
      float fx = (float)(0.5 + x); /* float _x = 0.5 + x (RangeLib.h:1205) */
      for (int y = 0; y < h; ++y) {
        float fy = (float)(0.5 + y);
        float lx = fx * cs - fy * sn;
        float ly = (fx * sn + fy * cs) + tr;
        unsigned li = (unsigned)(int)ly;
        if (li >= (unsigned)c->width[a]) continue; /* outside the table (the reference reads out of bounds) */
        long long b = c->bin_off[a] + li, s = c->start[b];
        long long n = c->start[b + 1] - s;
        const float* v = c->vals + s;
        long long high = n - 1;
        if (high == -1) continue;
        if (occ[(size_t)x * h + y]) continue;
        if (v[high] < lx && lx - v[high] < max_range) {
          used[s + high] = 1;
          continue;
        }
        long long index = lower_b(v, n, lx); /* linear search (<= 64 entries) finds the same index */
        if (index < n) used[s + index] = 1;
        if (index - 1 >= 0) used[s + index - 1] = 1;
      }
    }
  }
  long long out = 0;
  for (long long b = 0; b < nb; ++b) {
    long long s = c->start[b], e = c->start[b + 1];
    c->start[b] = out;
    for (long long k = s; k < e; ++k)
      if (used[k]) c->vals[out++] = c->vals[k];
  }
  c->start[nb] = out;
  free(used);
}

static int cddt_discretize(const or_cddt* c, float theta, int* flipped) {
  /* CDDTCast::discretize_theta (RangeLib.h:1279-1339); -1 where the wrap never ends */
  if (theta != theta) return -1;
  if (theta < 0.0) {
    while (theta < 0.0) {
      float next = (float)((double)theta + OR_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  } else if (theta > OR_M_2PI) {
    while (theta > OR_M_2PI) {
      float next = (float)((double)theta - OR_M_2PI);
      if (next == theta) return -1;
      theta = next;
    }
  }
  int fl = 0;
  if (theta >= OR_M_PI) {
    fl = 1;
    theta = (float)((double)theta - OR_M_PI);
  }
  int rounded = (int)roundf(theta * c->td_div_2pi);
  if (rounded == c->td >> 1) {
    rounded = 0;
    fl = !fl;
  }
  *flipped = fl;
  return rounded % c->td;
}

float or_cddt_calc_range(const or_cddt* c, const uint8_t* occ, int w, int h, float max_range, float x, float y,
                         float heading) {
  /* CDDTCast::calc_range (RangeLib.h:1341-1514).  NaN where the reference's
   * heading wrap never ends, and where it tests the occupancy of a cell
   * outside the map (it indexes grid[x][y] with the raw coordinates). */
  int fl = 0;
  int a = cddt_discretize(c, (float)(-1.0 * (double)heading), &fl);
  if (a < 0) return NAN;
  float cs = c->cosv[a], sn = c->sinv[a];
  float lx = x * cs - y * sn;
  float ly = (x * sn + y * cs) + c->trans[a];
  int li_i = (ly >= -2147483648.0f && ly < 2147483648.0f) ? (int)ly : INT_MIN;
  unsigned li = (unsigned)li_i;
  if (li >= (unsigned)c->width[a]) return max_range;
  long long b = c->bin_off[a] + li, s = c->start[b];
  long long n = c->start[b + 1] - s;
  const float* v = c->vals + s;
  long long high = n - 1;
  if (high == -1) return max_range;
  int cx = (x > -1.0f && x < (float)w) ? (int)x : -1, cy = (y > -1.0f && y < (float)h) ? (int)y : -1;
  if (fl) {
    if (v[0] > lx) return max_range;
    if (v[high] < lx) return lx - v[high];
    if (cx < 0 || cy < 0) return NAN;
    if (occ[(size_t)cx * h + cy]) return 0.0f;
    if (high > OR_BSEARCH) return lx - v[upper_b(v, n, lx) - 1];
    for (long long i = high; i >= 0; --i)
      if (v[i] <= lx) return lx - v[i];
  } else {
    if (v[high] < lx) return max_range;
    if (v[0] > lx) return v[0] - lx;
    if (cx < 0 || cy < 0) return NAN;
    if (occ[(size_t)cx * h + cy]) return 0.0f;
    if (high > OR_BSEARCH) {
      long long idx = upper_b(v, n, lx);
      if (idx == n) return 0.0f; /* v[high] == lx: the reference reads past the bin; the linear search gives 0 */
      return v[idx] - lx;
    }
    for (long long i = 0; i < n; ++i)
      if (v[i] >= lx) return v[i] - lx;
  }
  return -1.0f; /* not reached (assert(0) in the reference) */
}

void or_cddt_calc_range_pair(const or_cddt* c, const uint8_t* occ, int w, int h, float max_range, float x, float y,
                             float heading, float* r, float* r_inv) {
  /* CDDTCast::calc_range_pair (RangeLib.h:1521-1653): the range at the heading
   * and at the heading + pi from one bin search.  The reference indexes its
   * table without checking the bin (max_range pair here), its unflipped
   * occupancy test has no effect (no return: followed), and the flipped one
   * outside the map is NaN here (out of bounds there). */
  int fl = 0;
  int a = cddt_discretize(c, (float)(-1.0 * (double)heading), &fl);
  if (a < 0) {
    *r = *r_inv = NAN;
    return;
  }
  float cs = c->cosv[a], sn = c->sinv[a];
  float lx = x * cs - y * sn;
  float ly = (x * sn + y * cs) + c->trans[a];
  int li_i = (ly >= -2147483648.0f && ly < 2147483648.0f) ? (int)ly : INT_MIN;
  unsigned li = (unsigned)li_i;
  if (li >= (unsigned)c->width[a]) {
    *r = *r_inv = max_range;
    return;
  }
  long long b = c->bin_off[a] + li, s = c->start[b];
  long long n = c->start[b + 1] - s;
  const float* v = c->vals + s;
  long long high = n - 1;
  if (high == -1) {
    *r = *r_inv = max_range;
    return;
  }
  int cx = (x > -1.0f && x < (float)w) ? (int)x : -1, cy = (y > -1.0f && y < (float)h) ? (int)y : -1;
  if (fl) {
    if (v[0] > lx) {
      float d = v[0] - lx;
      *r = max_range;
      *r_inv = (d < max_range) ? d : max_range;
      return;
    }
    if (v[high] < lx) {
      *r = lx - v[high];
      *r_inv = max_range;
      return;
    }
    if (cx < 0 || cy < 0) {
      *r = *r_inv = NAN;
      return;
    }
    if (occ[(size_t)cx * h + cy]) {
      *r = *r_inv = 0.0f;
      return;
    }
    long long index = upper_b(v, n, lx) - 1; /* the linear scan's last entry <= x */
    float val = v[index];
    *r = lx - val;
    *r_inv = (index + 1 == n) ? max_range : v[index + 1] - lx;
  } else {
    if (v[high] < lx) {
      float d = lx - v[high];
      *r = max_range;
      *r_inv = (d < max_range) ? d : max_range;
      return;
    }
    long long index = lower_b(v, n, lx); /* the linear scan's first entry >= x */
    float val = v[index];
    *r = val - lx;
    *r_inv = (index - 1 == -1) ? max_range : lx - v[index - 1];
  }
}

void or_calc_range_many_radial_optimized(const or_method* m, const float* ins, float* outs, int num_particles,
                                         int num_rays, float min_angle, float max_angle) {
  /* RangeMethod::calc_range_many_radial_optimized (RangeLib.h:616-676).  Pair
   * queries exist for CDDTCast only (other methods return (-1, -1),
   * RangeLib.h:419).  Slots at or past num_rays are not written (the reference
   * writes them into the next particle's, or past the array for the last). */
  world_consts c = make_consts(&m->world); /* the same constants, RangeLib.h:619-627 */
  float step = (max_angle - min_angle) / (float)(num_rays - 1);
  int max_pairwise_index = (int)((double)(float)num_rays / 3.0);
  float index_offset_float = (float)(((double)num_rays - 1.0) * OR_M_PI / (double)(max_angle - min_angle));
  int index_offset = (int)roundf(index_offset_float);
  for (int i = 0; i < num_particles; ++i) {
    float theta = -ins[3 * i + 2] + c.rotation_const, x, y;
    to_grid(&c, ins[3 * i], ins[3 * i + 1], &x, &y);
    float angle = min_angle;
    for (int a = 0; a <= max_pairwise_index; ++a) {
      float r, r_inv;
      if (m->kind == 3) {
        or_cddt_calc_range_pair(m->cddt, m->occ, m->width, m->height, m->max_range, y, x, theta - angle, &r, &r_inv);
      } else {
        r = r_inv = -1.0f;
      }
      if (a < num_rays) outs[(size_t)i * num_rays + a] = r * c.world_scale;
      if (a + index_offset < num_rays) outs[(size_t)i * num_rays + a + index_offset] = r_inv * c.world_scale;
      angle += step;
    }
    for (int a = max_pairwise_index + 1; a < index_offset; ++a) {
      if (a < num_rays) outs[(size_t)i * num_rays + a] = or_calc_range(m, y, x, theta - angle) * c.world_scale;
      angle += step;
    }
  }
}
