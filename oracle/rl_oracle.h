/* CPU oracle for the batched ray-casting path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithms, op for op, so that
 * the HIP library's results can be checked against it.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it; the
 * product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the range_libc source tree).  Build with -ffp-contract=off and without
 * -ffast-math: every float operation must round exactly as the reference's.
 *
 * Pinning (which independent check fixes each function):
 *   or_edt              golden vectors from the compiled reference (maps/*.png
 *                       and seeded maps), brute-force exact EDT on small maps,
 *                       all-free / all-occupied closed forms.
 *   or_rm_calc_range    golden vectors from the compiled reference; invariants
 *                       (0 <= r <= R, hit cell has EDT 0, r = |cell - origin|).
 *   or_bl_calc_range    golden vectors from the compiled reference; axis-aligned
 *                       rays reduce to a row/column scan (brute force).
 *   or_numpy_*          golden vectors (identity and non-identity world params);
 *                       identity-parameter closed form of the transform.
 *   or_eval_sensor_*    golden vectors; constant-table closed form c^M.
 *   or_glt_*            the compiled reference's own GiantLUTCast (live, and a
 *                       golden fixture of its table and queries); the table
 *                       entry equals the quantized or_rm_calc_range, and bins
 *                       of headings i * 2pi/td are i.
 */
#ifndef RL_ORACLE_H
#define RL_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* World-frame parameters of an OMap (RangeLib.h:134-141). */
typedef struct {
  float scale, angle, origin_x, origin_y, sin_angle, cos_angle;
} or_world;

/* CDDTCast tables (RangeLib.h:1744-1749): angle a has width[a] bins; bin i of
 * angle a holds vals[start[bin_off[a] + i] .. start[bin_off[a] + i + 1]). */
typedef struct {
  int td, nang;            /* theta discretization; angles that hold data */
  int* width;              /* [td] */
  float* trans;            /* [td] lut_translations */
  float* cosv;             /* [td] cosf of each bin's angle */
  float* sinv;             /* [td] */
  long long* bin_off;      /* [td + 1] */
  long long* start;        /* [bins + 1] */
  float* vals;             /* sorted, duplicate-free obstacle positions per bin */
  float td_div_2pi, m2pi_div_td;
} or_cddt;

/* One range method bound to a map (RangeMethod state, RangeLib.h:682-688). */
typedef struct {
  int kind;            /* 0 = RayMarching, 1 = BresenhamsLine, 2 = GiantLUTCast, 3 = CDDTCast */
  int width, height;   /* grid[x][y], x < width, y < height */
  const uint8_t* occ;  /* occupancy, x-major: occ[x*height + y] */
  const float* edt;    /* distance transform, x-major (RayMarching only) */
  float max_range;
  or_world world;
  const double* table; /* sensor model, row-major K x K (row = observed) */
  int table_width;
  const uint16_t* glt; /* GiantLUTCast table, glt[(x*height + y)*glt_td + bin] */
  int glt_td;          /* theta discretization */
  float glt_max_div_limits, glt_td_div_2pi;
  const or_cddt* cddt; /* CDDTCast tables */
} or_method;

/* DistanceTransform(OMap*) (RangeLib.h:345-373). */
void or_edt(const uint8_t* occ, int width, int height, float* out);

/* RayMarching::calc_range (RangeLib.h:927-962); *reads = EDT reads. */
float or_rm_calc_range(const float* edt, int width, int height, float max_range, float x, float y,
                       float heading, int* reads);

/* BresenhamsLine::calc_range (RangeLib.h:696-769); *visits = cells tested. */
float or_bl_calc_range(const uint8_t* occ, int width, int height, float max_range, float x,
                       float y, float heading, int* visits);

/* GiantLUTCast construction (RangeLib.h:1781-1816): table[(x*height + y)*td + i]
 * = (uint16) (min(R, RM(x, y, i * (2pi/td))) * (65535 / R)); consts receives
 * max_div_limits, limits_div_max, theta_discretization_div_M_2PI. */
void or_glt_build(const float* edt, int width, int height, float max_range, int td, uint16_t* table,
                  float* consts);

/* GiantLUTCast::discretize_theta (RangeLib.h:1832-1864, _USE_ALTERNATE_MOD,
 * _USE_CACHED_CONSTANTS, roundf).  Returns -1 where the reference does not
 * terminate (a heading so large that subtracting 2pi leaves it unchanged) or
 * reads outside its table (NaN heading). */
int or_glt_discretize(float theta, int td, float td_div_2pi);

/* GiantLUTCast::calc_range (RangeLib.h:1866-1877); NaN where or_glt_discretize is -1
 * and for a NaN origin (the reference indexes its table at (int)NaN). */
float or_glt_calc_range(const or_method* m, float x, float y, float heading);

/* CDDTCast(OMap, max_range, td) (RangeLib.h:974-1142); free with or_cddt_free. */
or_cddt* or_cddt_build(const uint8_t* occ, int width, int height, int td);
void or_cddt_free(or_cddt* c);
/* CDDTCast::prune (RangeLib.h:1172-1275). */
void or_cddt_prune(or_cddt* c, const uint8_t* occ, int width, int height, float max_range);
/* CDDTCast::calc_range (RangeLib.h:1341-1514); NaN for a heading whose wrap never
 * ends and for an occupancy test outside the map (out-of-bounds in the reference). */
float or_cddt_calc_range(const or_cddt* c, const uint8_t* occ, int width, int height, float max_range, float x,
                         float y, float heading);

/* CDDTCast::calc_range_pair (RangeLib.h:1521-1653): ranges at heading and heading + pi. */
void or_cddt_calc_range_pair(const or_cddt* c, const uint8_t* occ, int width, int height, float max_range, float x,
                             float y, float heading, float* r, float* r_inv);

/* RangeMethod::calc_range_many_radial_optimized (RangeLib.h:616-676); slots past
 * num_rays (written out of the particle's segment by the reference) are skipped. */
void or_calc_range_many_radial_optimized(const or_method* m, const float* ins, float* outs, int num_particles,
                                         int num_rays, float min_angle, float max_angle);

float or_calc_range(const or_method* m, float x, float y, float heading);

/* RayMarchingGPU::calc_range_many semantics (grid frame, no transform;
 * RangeLib.h:819-831 -> kernels.cu:19-57 with the CPU step constant). */
void or_calc_range_many(const or_method* m, const float* ins, float* outs, int n, int* reads);

/* RangeMethod::numpy_calc_range (RangeLib.h:439-480). */
void or_numpy_calc_range(const or_method* m, const float* ins, float* outs, int n, int* reads);

/* RangeMethod::numpy_calc_range_angles (RangeLib.h:482-520). */
void or_numpy_calc_range_angles(const or_method* m, const float* ins, const float* angles,
                                float* outs, int num_particles, int num_angles, int* reads);

/* RangeMethod::eval_sensor_model (RangeLib.h:533-555). */
void or_eval_sensor_model(const or_method* m, const float* obs, const float* ranges, double* outs,
                          int rays_per_particle, int particles);

/* RangeMethod::calc_range_repeat_angles_eval_sensor_model (RangeLib.h:558-612). */
void or_calc_range_repeat_angles_eval_sensor_model(const or_method* m, const float* ins,
                                                   const float* angles, const float* obs,
                                                   double* weights, int num_particles,
                                                   int num_angles);

/* The C library's sinf/cosf applied elementwise (for trig parity tests). */
void or_sinf_many(const float* x, float* out, long n);
void or_cosf_many(const float* x, float* out, long n);

#ifdef __cplusplus
}
#endif
#endif
