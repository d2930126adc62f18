/* rangelib_hip.h -- C ABI of the MI355X (gfx950) ray-casting library.
 *
 * Drop-in replacement for the reference's host/device seam (class
 * RayMarchingCUDA, includes/CudaRangeLib.h:11-61 and includes/kernels.cu:199-315)
 * and for the CPU-only batch methods of RangeMethod / BresenhamsLine
 * (includes/RangeLib.h:412-769) that the Cython module binds
 * (pywrapper/RangeLibc.pyx:25-92).  Plain pointers and sizes only; every
 * function returns RL_OK (0) or an error code and never throws; the message of
 * the last failure on the calling thread is rl_last_error().
 *
 * Conventions (identical to the reference's):
 *   occupancy / distance transform  x-major: a[x*height + y], x < width
 *   poses                           float32 triples (x, y, theta), C order
 *   sensor table                    float64 K x K row-major, row = observed range
 *   ranges                          float32; weights float64
 * Entry points without a _dev suffix take host pointers and return when the
 * results are in host memory (the reference's synchronous semantics).  The
 * _dev twins take device pointers, are ordered on the context's stream and
 * do not synchronise.
 *
 * Thread safety: calls on one context are serialised by a context mutex;
 * distinct contexts are independent.
 */
#ifndef RANGELIB_HIP_H
#define RANGELIB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RL_ABI_VERSION 1

#define RL_OK 0
#define RL_ERR_ARG 1     /* bad argument: null pointer, size, unknown method */
#define RL_ERR_HIP 2     /* HIP runtime error (message has the HIP error) */
#define RL_ERR_STATE 3   /* call not valid in the object's state */
#define RL_ERR_IO 4      /* missing or corrupt data file */
#define RL_ERR_NOMEM 5   /* device or host allocation failed */

/* Range methods: RayMarching / RayMarchingGPU (RangeLib.h:774-969) and
 * BresenhamsLine (RangeLib.h:691-772). */
#define RL_METHOD_RM 0
#define RL_METHOD_BL 1
#define RL_METHOD_GLT 2  /* GiantLUTCast: needs rl_map_build_giant_lut first */
#define RL_METHOD_CDDT 3 /* CDDTCast (PCDDT after rl_map_prune_cddt): needs rl_map_build_cddt first */

/* Coordinate frames of a batch. */
#define RL_FRAME_GRID 0    /* no conversion: RayMarchingGPU::calc_range_many (RangeLib.h:819-831) */
#define RL_FRAME_WORLD 1   /* RangeMethod::numpy_calc_range (RangeLib.h:439-480) */
#define RL_FRAME_ANGLES 2  /* RangeMethod::numpy_calc_range_angles (RangeLib.h:482-520) */

/* rl_map_create_ex flags */
#define RL_MAP_EDT_ENVELOPE_ONLY 1u  /* run the lower-envelope pass on columns too (test hook) */
#define RL_MAP_NO_COMPACT 2u         /* march the fp32 distance map only (no 8-bit index map) */

typedef struct rl_ctx rl_ctx;
typedef struct rl_map rl_map;

/* Message of the last failed call on this thread ("" if none). */
const char* rl_last_error(void);
int rl_abi_version(void);
/* Loads the sinf/cosf exception table (no GPU needed) and writes the C library
 * release it was generated with and the one this process runs (len bytes
 * each; len 0 writes nothing).  RL_ERR_STATE when they differ: the reference's
 * sinf/cosf (RangeLib.h:713-714,931-932) come from the running C library, so
 * bit parity needs the same release; rl_create refuses the same mismatch unless
 * RL_TRIG_ALLOW_LIBC_MISMATCH=1 is set.  No counterpart in the reference. */
int rl_trig_table_info(char* table_libc, char* running_libc, int len);

/* --- map ingestion (host only; no GPU needed) ------------------------------ */
/* OMap(filename, threshold) (RangeLib.h:158-201): decodes the PNG at `path` to
 * 8-bit RGBA and writes the occupancy x-major, occ[x*height + y] = 1 where the
 * pixel's gray value (int)(0.229 r + 0.587 g + 0.114 b) (RangeUtils.h:30-32, with
 * r, g, b the third, second and first byte of the pixel as the reference reads
 * them) is below `threshold` (128 in OMap(filename)).  With occ_xmajor NULL only
 * *width / *height are written (size query).  RL_ERR_IO when the file cannot be
 * read or decoded (the reference prints the decoder error and flags has_error). */
int rl_png_occupancy(const char* path, float threshold, uint8_t* occ_xmajor, int* width, int* height);
/* OMap::save (RangeLib.h:264-291): an opaque RGBA PNG, black where occupied,
 * white elsewhere.  RL_OK on success (the reference returns its encoder's
 * error code, 0 on success). */
int rl_png_save_occupancy(const char* path, const uint8_t* occ_xmajor, int width, int height);

/* --- context ------------------------------------------------------------ */
/* Binds a HIP device, creates a stream, uploads the sinf/cosf exception table.
 * Replaces the device set-up of RayMarchingCUDA::RayMarchingCUDA (kernels.cu:205-218). */
int rl_create(int device, rl_ctx** out);
/* Replaces RayMarchingCUDA::~RayMarchingCUDA (kernels.cu:220-222).  Handles
 * are reference counted: maps and batches keep their context alive, and
 * surrogates their map, so destroying them in any order is safe; the caller's
 * handle is invalid after its destroy call either way. */
int rl_destroy(rl_ctx* ctx);
/* Orders all later work of the context on `hip_stream` (a hipStream_t; NULL
 * restores the context's own stream).  No counterpart in the reference. */
int rl_set_stream(rl_ctx* ctx, void* hip_stream);
/* Waits for the context's stream (cudaDeviceSynchronize, kernels.cu:241). */
int rl_sync(rl_ctx* ctx);

/* Tuning options of a context (no counterpart in the reference; the
 * reference's only knobs are the compile-time CHUNK_SIZE / NUM_THREADS,
 * CMakeLists.txt:13-14). */
#define RL_OPT_XCD_SCHED 1          /* 1: bucket particles by map region per XCD (default 0) */
#define RL_OPT_LDS_BOX 2            /* 1: stage 64x64 cells around a particle in LDS (default 0) */
int rl_set_option(rl_ctx* ctx, int option, long long value);

/* --- maps ----------------------------------------------------------------- */
/* Uploads an occupancy grid and builds its distance transform on the GPU.
 * Replaces DistanceTransform(OMap*) (RangeLib.h:345-373) plus the H2D copy of
 * RayMarchingCUDA::RayMarchingCUDA (kernels.cu:210-216).  World parameters
 * default to the identity (RangeLibc.pyx:174-180). */
int rl_map_create(rl_ctx* ctx, const uint8_t* occ_xmajor, int width, int height, rl_map** out);
int rl_map_create_ex(rl_ctx* ctx, const uint8_t* occ_xmajor, int width, int height,
                     unsigned flags, rl_map** out);
/* A map whose distance transform the caller already has (x-major fp32, the
 * layout of DistanceTransform::grid flattened as in kernels.cu:213-216): no
 * transform is computed.  Ray marching then reads these values exactly as
 * given (the compact map keeps the exact ones, anything else is read from
 * them); Bresenham uses the occupancy. */
int rl_map_create_from_edt(rl_ctx* ctx, const uint8_t* occ_xmajor, const float* edt_xmajor, int width,
                           int height, unsigned flags, rl_map** out);
int rl_map_destroy(rl_map* map);
/* GiantLUTCast(OMap, max_range, theta_discretization) (RangeLib.h:1781-1816):
 * builds the map's W x H x td uint16 table on the device (the reference builds
 * it on the CPU, one marched ray per entry).  Afterwards RL_METHOD_GLT works in
 * every casting call at this max_range (RangeLib.h:1866-1877 lookups). */
int rl_map_build_giant_lut(rl_map* map, float max_range, int theta_discretization);
/* CDDTCast(OMap, max_range, theta_discretization) (RangeLib.h:974-1142): the
 * compressed tables built on the device; RL_METHOD_CDDT then works in every
 * casting call (RangeLib.h:1341-1514 lookups, any max_range). */
int rl_map_build_cddt(rl_map* map, int theta_discretization);
/* CDDTCast::prune(max_range) (RangeLib.h:1172-1275): PCDDT. */
int rl_map_prune_cddt(rl_map* map, float max_range);
/* The tables as CSR: widths[td], trans[td] (lut_translations), start[nbins + 1],
 * vals[nvals]; bins of angle 0 first.  Any pointer may be NULL. */
int rl_map_get_cddt(rl_map* map, int* theta_discretization, long long* nbins, long long* nvals, int* widths,
                    float* trans, long long* start, float* vals, float* build_ms);
/* RangeMethod::calc_range_many_radial_optimized (RangeLib.h:616-676): num_rays
 * beams evenly over [min_angle, max_angle] per world-frame pose, the beams one
 * half-turn apart answered by one CDDT pair query (calc_range_pair,
 * RangeLib.h:1521-1653; other methods give -1 for those slots, RangeLib.h:419).
 * Slots the reference leaves unwritten keep their contents; slots it would
 * write past a particle's num_rays are not written. */
int rl_calc_range_many_radial_optimized(rl_map* map, int method, float max_range, const float* ins, float* outs,
                                        int num_particles, int num_rays, float min_angle, float max_angle);
int rl_calc_range_many_radial_optimized_dev(rl_map* map, int method, float max_range, const float* d_ins,
                                            float* d_outs, int num_particles, int num_rays, float min_angle,
                                            float max_angle);
/* Copies the table (x-major, theta fastest: the reference's giant_lut[x][y][i])
 * when `out` is non-NULL; any other pointer may be NULL. */
int rl_map_get_giant_lut(rl_map* map, uint16_t* out, int* theta_discretization, float* max_range,
                         float* build_ms);
/* OMap world_* fields (RangeLib.h:134-141) as RayMarchingCUDA::set_conversion_params
 * consumes them (CudaRangeLib.h:24-35). */
int rl_map_set_world(rl_map* map, float scale, float angle, float origin_x, float origin_y,
                     float sin_angle, float cos_angle);
/* RangeMethod::set_sensor_model / RayMarchingCUDA::set_sensor_table
 * (RangeLib.h:523-532, 878-888; kernels.cu:224-229). */
int rl_map_set_sensor_table(rl_map* map, const double* table, int table_width);
/* Copies the distance transform to host memory, x-major (DistanceTransform::grid). */
int rl_map_get_edt(rl_map* map, float* out_xmajor);
/* Device pointers of the map's distance transform and occupancy (x-major). */
int rl_map_device_buffers(rl_map* map, const float** d_edt, const uint8_t** d_occ);
/* Milliseconds the last distance-transform build took on the device. */
int rl_map_edt_build_ms(rl_map* map, float* ms);

/* --- single query (host) ---------------------------------------------------- */
/* <Method>::calc_range (RangeLib.h:696, 927); grid frame. */
int rl_calc_range(rl_map* map, int method, float max_range, float x, float y, float heading,
                  float* out);

/* --- batches, host pointers, synchronous ----------------------------------- */
/* RayMarchingGPU::calc_range_many (RangeLib.h:819-831 -> kernels.cu:232-242). */
int rl_calc_range_many(rl_map* map, int method, float max_range, const float* ins, float* outs,
                       int n);
/* numpy_calc_range (RangeLib.h:439-480, 837-853 -> kernels.cu:245-259). */
int rl_numpy_calc_range(rl_map* map, int method, float max_range, const float* ins, float* outs,
                        int n);
/* numpy_calc_range_angles (RangeLib.h:482-520, 855-874 -> kernels.cu:263-279). */
int rl_numpy_calc_range_angles(rl_map* map, int method, float max_range, const float* ins,
                               const float* angles, float* outs, int num_particles,
                               int num_angles);
/* RangeMethod::eval_sensor_model (RangeLib.h:533-555); needs a sensor table. */
int rl_eval_sensor_model(rl_map* map, const float* obs, const float* ranges, double* outs,
                         int rays_per_particle, int particles);
/* calc_range_repeat_angles_eval_sensor_model (CPU RangeLib.h:558-612; the GPU
 * path RangeLib.h:892-910 -> kernels.cu:281-315 is a stub in the reference). */
int rl_calc_range_repeat_angles_eval_sensor_model(rl_map* map, int method, float max_range,
                                                  const float* ins, const float* angles,
                                                  const float* obs, double* weights,
                                                  int num_particles, int num_angles);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): the call above with
 * the same arguments, results and errors, except that the poses, angles,
 * observations and weights are staged in a static device array of
 * RL_NOALLOC_SCRATCH_BYTES per GPU (each part rounded up to 256 bytes:
 * 12*N + 8*M + 8*N bytes plus padding; about 200,000 particles at 60 angles)
 * instead of buffers the context allocates.  Larger calls fail with RL_ERR_ARG.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (from two contexts, or two threads with two contexts) race. */
#define RL_NOALLOC_SCRATCH_BYTES (4u << 20)
int rl_calc_range_repeat_angles_eval_sensor_model_noalloc(rl_map* map, int method, float max_range,
                                                          const float* ins, const float* angles,
                                                          const float* obs, double* weights,
                                                          int num_particles, int num_angles);

/* --- batches, device pointers, stream-ordered -------------------------------- */
int rl_calc_range_many_dev(rl_map* map, int method, float max_range, const float* d_ins,
                           float* d_outs, int n);
int rl_numpy_calc_range_dev(rl_map* map, int method, float max_range, const float* d_ins,
                            float* d_outs, int n);
int rl_numpy_calc_range_angles_dev(rl_map* map, int method, float max_range, const float* d_ins,
                                   const float* d_angles, float* d_outs, int num_particles,
                                   int num_angles);
int rl_eval_sensor_model_dev(rl_map* map, const float* d_obs, const float* d_ranges,
                             double* d_outs, int rays_per_particle, int particles);
int rl_calc_range_repeat_angles_eval_sensor_model_dev(rl_map* map, int method, float max_range,
                                                      const float* d_ins, const float* d_angles,
                                                      const float* d_obs, double* d_weights,
                                                      int num_particles, int num_angles);

/* numpy_calc_range_angles (RangeLib.h:482-520) then eval_sensor_model
 * (RangeLib.h:533-555) on its output -- the particle-filter update that
 * RangeLibc.pyx:328-329 + :334-335 perform as two calls -- in one launch.
 * Writes both outputs, each bit-identical to the two separate calls. */
int rl_pf_step_dev(rl_map* map, int method, float max_range, const float* d_ins,
                   const float* d_angles, const float* d_obs, float* d_ranges, double* d_weights,
                   int num_particles, int num_angles);

/* --- multi-map batches ----------------------------------------------------------
 * Many maps of uneven size, each with its own particles, beam count and
 * (optional) sensor model, cast in one launch.  No counterpart in the
 * reference, whose batches are one map per call in CHUNK_SIZE pieces
 * (RangeLib.h:819-874).  Each job is numpy_calc_range_angles (RangeLib.h:482-520)
 * of its map and, when d_obs is set, eval_sensor_model (RangeLib.h:533-555) of
 * the ranges; outputs are bit-identical to those calls.  All pointers are
 * device pointers; maps and buffers must outlive the launches. */
typedef struct {
  rl_map* map;
  float max_range;
  const float* d_ins;     /* num_particles x 3 poses (world frame) */
  const float* d_angles;  /* num_angles */
  const float* d_obs;     /* num_angles observed ranges, or NULL: no sensor model */
  float* d_ranges;        /* num_particles * num_angles */
  double* d_weights;      /* num_particles (when d_obs is set) */
  int num_particles;
  int num_angles;
} rl_batch_job;

typedef struct rl_batch rl_batch;
int rl_batch_create(rl_ctx* ctx, int method, rl_batch** out);
int rl_batch_destroy(rl_batch* batch);
/* Uploads the job descriptors and measures each map's mean reads per ray on a
 * sample of its rays (synchronous; not for graph capture). */
int rl_batch_set_jobs(rl_batch* batch, const rl_batch_job* jobs, int n_jobs);
/* Times every launch configuration (lanes per particle 64/32/16, LDS box,
 * work order, grid size) `reps` times on the current jobs and keeps the
 * fastest; *best_ms receives its time per launch. */
int rl_batch_autotune(rl_batch* batch, int reps, float* best_ms);
/* order: 0 job order, 1 estimated cost descending, 2 each map's work on one
 * XCD (balanced by estimated cost, longest first); grid: 0 one workgroup per
 * work item, else a persistent grid of that many workgroups. */
int rl_batch_set_config(rl_batch* batch, int rays_per_wave, int lds_box, int order, int grid);
int rl_batch_get_config(rl_batch* batch, int* rays_per_wave, int* lds_box, int* order, int* grid);
/* One launch over all jobs on the context's stream: no allocation, no copy,
 * so it may be captured in a HIP graph. */
int rl_batch_launch(rl_batch* batch);

/* --- learned range surrogate ------------------------------------------------- */
/* No counterpart in the reference.  A per-map Fourier-feature MLP (64 features
 * from 32 random frequencies of (x0/W, y0/H, sin h, cos h), three 256-wide ReLU
 * layers, two outputs: range / max_range and log of an error bound in cells),
 * bf16 weights with fp32 accumulation on the matrix cores, trained on the GPU on
 * rays cast by the exact marcher.  Inference is a checked start point for ray
 * marching: a ray whose predicted bound is <= bound_max cells starts at
 * max(0, range - bound - 1) if that cell is free and must hit by range + bound + 2;
 * every other ray is marched from its origin.  Results approximate the exact
 * march (a thin wall before the start point is missed); the error is reported,
 * not bounded.  A surrogate belongs to one map and must not outlive it. */
typedef struct rl_surrogate rl_surrogate;
int rl_surrogate_create(rl_map* map, unsigned seed, rl_surrogate** out);
int rl_surrogate_destroy(rl_surrogate* s);
/* Number of parameters (returned directly, not a status). */
int rl_surrogate_num_params(void);
/* Flat fp32 parameters W1 (256x64), b1, W2 (256x256), b2, W3, b3, W4 (2x256), b4
 * (row-major [out][in]) and the 32 x 4 frequencies; host pointers, either may be NULL. */
int rl_surrogate_get_params(rl_surrogate* s, float* params, float* freq);
int rl_surrogate_set_params(rl_surrogate* s, const float* params, const float* freq);
/* Network outputs (n x 2) for n grid-frame rays (x0, y0, sin h, cos h); host pointers. */
int rl_surrogate_forward(rl_surrogate* s, const float* rays4, int n, float* out);
/* Mean loss over n rays with target ranges (cells) and its gradient (num_params
 * floats); `features` (optional, n x 64 bf16 bit patterns) receives the network's input. */
int rl_surrogate_gradient(rl_surrogate* s, float max_range, const float* rays4, const float* targets, int n,
                          float* loss, float* grad, uint16_t* features);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the gradient of the last
 * rl_surrogate_gradient call, e.g. to train on caller-supplied rays. */
int rl_surrogate_step(rl_surrogate* s, float lr);
/* n_pairs / batch Adam steps on uniform free-space rays marched on the GPU;
 * loss_curve (host, optional) gets one loss per step.  Synchronous. */
int rl_surrogate_train(rl_surrogate* s, float max_range, long long n_pairs, int batch, float lr, unsigned seed,
                       float* loss_curve);
/* numpy_calc_range_angles through the checked surrogate (device pointers,
 * stream-ordered).  stats (host, 5 counters, optional; synchronises): rays
 * marched from the origin because the bound exceeded bound_max, because the
 * start check failed, rays finished from the start point, distance-map reads,
 * dependent read rounds (the path's latency in reads). */
int rl_surrogate_calc_range_angles_dev(rl_surrogate* s, float max_range, const float* d_ins,
                                       const float* d_angles, float* d_ranges, int num_particles,
                                       int num_angles, float bound_max, unsigned* stats);

/* --- instrumentation (device pointers) -------------------------------------- */
/* Per-ray work counts of a batch: distance-transform reads (RM) or cells
 * visited (BL).  `frame` is RL_FRAME_*; `angles`/`num_angles` are used by
 * RL_FRAME_ANGLES only, `num_particles` is the number of poses. */
int rl_count_steps_dev(rl_map* map, int method, int frame, float max_range, const float* d_ins,
                       const float* d_angles, int* d_counts, int num_particles, int num_angles);
/* The library's sinf/cosf (bit-identical to the host C library on |x| <= 64). */
int rl_debug_sincosf_dev(rl_ctx* ctx, const float* d_x, float* d_sin, float* d_cos, int n);
/* Device sqrtf as the kernels use it. */
int rl_debug_sqrtf_dev(rl_ctx* ctx, const float* d_x, float* d_out, int n);
/* Dependent-load latency probe (no reference counterpart): one lane loads
 * i = d_next[i] `steps` times from i = 0; d_out[0] = elapsed wall-clock ticks,
 * d_out[1] = the final index; *tick_ns (host, optional) = ns per tick.  Every
 * d_next entry must index inside the buffer. */
int rl_debug_chase_dev(rl_ctx* ctx, const unsigned* d_next, int steps, long long* d_out, double* tick_ns);
/* C = A * B through the surrogate's bf16 matrix-core GEMM (host row-major
 * matrices of bf16 bit patterns; fp32 result). */
int rl_debug_gemm_bf16(rl_ctx* ctx, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K);
/* rl_surrogate_forward through the per-layer kernels that training uses
 * (rl_surrogate_forward itself runs one fused kernel with the same outputs). */
int rl_debug_surrogate_forward_layers(rl_surrogate* s, const float* rays4, int n, float* out);
/* Number of float bit patterns (of all 2^32) on which the fused surrogate
 * kernel's bf16 conversion differs from the per-layer kernels' (expected 0). */
int rl_debug_bf16_check(rl_ctx* ctx, unsigned long long* mismatches);
/* C = A (M x K) * Bt (N x K)^T, the surrogate's forward-layer layout (the
 * 128 x 128 tile kernel when K % 32 == 0 and N % 128 == 0). */
int rl_debug_gemm_bf16_nt(rl_ctx* ctx, const uint16_t* A, const uint16_t* Bt, float* C, int M, int N, int K);
/* The CDDT build's device-wide exclusive sum (d_out[i] = d_in[0] + ... +
 * d_in[i-1], i < n) and its per-segment ascending float sort (segment s is
 * [d_start[s], d_start[s+1]), nseg + 1 offsets; -0 before +0), on device
 * buffers.  Both return after the work ran. */
int rl_debug_exclusive_sum_dev(rl_ctx* ctx, const unsigned long long* d_in, unsigned long long* d_out, long long n);
int rl_debug_segmented_sort_dev(rl_ctx* ctx, const float* d_in, float* d_out, const long long* d_start,
                                long long nseg);

#ifdef __cplusplus
}
#endif
#endif
