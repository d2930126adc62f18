// Multi-map batches: many maps of uneven size, each with its own particles,
// beam count and sensor model, cast in one persistent launch.
//
// Included at the end of rl_lib.hip (uses its internal types and
// particle_beams).  The work list holds one item per (map, particle chunk); a
// chunk is the particles one 256-thread workgroup casts together (4 waves x
// 64 / S particles, S = lanes per particle).  k_batch runs a fixed grid whose
// workgroups stride over the list, so long items placed first (order 1:
// estimated cost, descending) finish while later short items fill the gaps.
//
// Tuned per batch by rl_batch_autotune: lanes per particle S (16, 32 or 64,
// never above the beam count's power of two), LDS staging of the 64x64 cells
// around each particle (S = 64 only), item order, and grid size.  Order 2
// keeps each map on one XCD: workgroup b runs on XCD b mod X (X XCDs, each
// with its own L2), so the list is X interleaved queues, each map's items in
// one queue (a map costing more than a queue's share spills into the next
// least loaded one), queues balanced by estimated cost, longest items first
// within a queue, and padded with empty items to equal length.  The
// estimated cost of an item is its ray count times the mean number of
// distance-map reads per ray of its map, measured on a sample of the map's
// own rays when the jobs are set.

struct BatchMap {
  MapDev m;
  float max_range;
  const float* ins;
  const float* angles;
  const float* obs;
  const double* table;
  int K;
  float* ranges;
  double* weights;
  int N, M, S;
  int sensor;
};

struct BatchItem {
  int map;    // -1: an empty item (padding of the XCD-affine order)
  int first;  // first particle of the chunk
};

template <int METHOD>
__global__ __launch_bounds__(256) void k_batch(const BatchMap* __restrict__ maps,
                                               const BatchItem* __restrict__ items, int n_items,
                                               TrigDev trig, int use_box) {
  __shared__ float lut[RL_LUT_SIZE];
  __shared__ __align__(16) uint8_t boxes[4 * RL_BOX_BYTES];
  __shared__ double terms[4 * 64];
  if (METHOD == RL_METHOD_RM) {  // one distance table per context: the same for every map
    for (int i = threadIdx.x; i < RL_LUT_SIZE; i += blockDim.x) lut[i] = maps[0].m.steps[i];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const BatchItem item = items[it];
    if (item.map < 0) continue;
    const BatchMap& bm = maps[item.map];
    MapDev m = bm.m;
    m.use_box = use_box;
    const int S = bm.S;
    const int per_wave = 64 / S;
    const int seg = lane / S, j0 = lane - seg * S, base = seg * S;
    const int p = item.first + wib * per_wave + seg;
    const int pp = p < bm.N ? p : -1;
    const BeamIn none{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, false};
    // without the whole-wave finish of a wave's last ray: the batch is bound by
    // its total work (64-map batch 1.310 -> 1.285 ms)
    if (bm.sensor)
      particle_beams<METHOD, true, false>(m, lut, trig, bm.max_range, bm.ins, bm.angles, bm.obs, bm.table, bm.K,
                                          bm.ranges, bm.weights, pp, bm.M, S, j0, base,
                                          boxes + wib * RL_BOX_BYTES, terms + wib * 64, none);
    else
      particle_beams<METHOD, false, false>(m, lut, trig, bm.max_range, bm.ins, bm.angles, bm.obs, bm.table, bm.K,
                                           bm.ranges, bm.weights, pp, bm.M, S, j0, base,
                                           boxes + wib * RL_BOX_BYTES, terms + wib * 64, none);
  }
}

struct rl_batch {
  rl_ctx* ctx = nullptr;
  int method = RL_METHOD_RM;
  std::vector<rl_batch_job> jobs;     // each job's map holds one reference per job
  std::vector<double> reads_per_ray;  // measured per job
  std::vector<unsigned long long> gens;  // each job's map generation at the last upload
  // launch configuration
  int rays_per_wave = 64;  // upper bound on S
  int lds_box = 0;
  int order = 1;           // 0: job order, 1: estimated cost descending, 2: XCD-affine (see above)
  int grid = 0;            // 0: one workgroup per item
  int xcds = 8;            // XCDs of the device (workgroup b runs on XCD b mod xcds)
  // device descriptors of the current configuration
  BatchMap* d_maps = nullptr;
  BatchItem* d_items = nullptr;
  int n_items = 0;
  size_t cap_maps = 0, cap_items = 0;
};

namespace {

int batch_upload(rl_batch* b) {
  rl_ctx* c = b->ctx;
  const int nj = (int)b->jobs.size();
  std::vector<BatchMap> maps(nj);
  std::vector<std::pair<double, BatchItem>> items;
  b->gens.resize(nj);
  for (int i = 0; i < nj; ++i) b->gens[i] = b->jobs[i].map->gen.load();
  for (int i = 0; i < nj; ++i) {
    const rl_batch_job& j = b->jobs[i];
    BatchMap& bm = maps[i];
    bm.m = j.map->dev();
    bm.max_range = j.max_range;
    bm.ins = j.d_ins;
    bm.angles = j.d_angles;
    bm.obs = j.d_obs;
    bm.sensor = j.d_obs != nullptr;
    bm.table = bm.sensor ? j.map->d_table : nullptr;
    bm.K = bm.sensor ? j.map->table_width : 1;
    bm.ranges = j.d_ranges;
    bm.weights = j.d_weights;
    bm.N = j.num_particles;
    bm.M = j.num_angles;
    int S = segment_size(j.num_angles);
    if (S > b->rays_per_wave) S = b->rays_per_wave;
    bm.S = S;
    const int chunk = 4 * (64 / S);
    const double ray_cost = (double)j.num_angles * (1.0 + b->reads_per_ray[i]);
    for (int f = 0; f < j.num_particles; f += chunk) {
      int cnt = j.num_particles - f < chunk ? j.num_particles - f : chunk;
      items.push_back({cnt * ray_cost, BatchItem{i, f}});
    }
  }
  auto by_cost = [](const std::pair<double, BatchItem>& a, const std::pair<double, BatchItem>& z) {
    return a.first > z.first;
  };
  std::vector<BatchItem> flat;
  if (b->order == 2 && b->xcds > 1) {
    const int X = b->xcds;
    std::vector<double> jcost(nj, 0.0);
    std::vector<std::vector<std::pair<double, BatchItem>>> per_job(nj);
    double total = 0.0;
    for (const auto& it : items) {
      jcost[it.second.map] += it.first;
      per_job[it.second.map].push_back(it);
      total += it.first;
    }
    std::vector<int> jorder(nj);
    for (int i = 0; i < nj; ++i) jorder[i] = i;
    std::stable_sort(jorder.begin(), jorder.end(), [&](int a, int z) { return jcost[a] > jcost[z]; });
    const double share = total / X;
    std::vector<double> load(X, 0.0);
    std::vector<std::vector<std::pair<double, BatchItem>>> q(X);
    auto least = [&]() { return (int)(std::min_element(load.begin(), load.end()) - load.begin()); };
    for (int j : jorder) {
      int x = least();
      for (const auto& it : per_job[j]) {
        if (load[x] >= share) x = least();  // this map spills into the next queue
        q[x].push_back(it);
        load[x] += it.first;
      }
    }
    size_t len = 0;
    for (auto& v : q) {
      std::stable_sort(v.begin(), v.end(), by_cost);
      len = std::max(len, v.size());
    }
    flat.assign(len * X, BatchItem{-1, 0});
    for (int x = 0; x < X; ++x)
      for (size_t k = 0; k < q[x].size(); ++k) flat[k * X + x] = q[x][k].second;
  } else {
    if (b->order == 1) std::stable_sort(items.begin(), items.end(), by_cost);
    flat.resize(items.size());
    for (size_t k = 0; k < items.size(); ++k) flat[k] = items[k].second;
  }
  if (b->cap_maps < maps.size()) {
    if (b->d_maps) HIP_TRY(hipFree(b->d_maps));
    b->d_maps = nullptr;
    if (hipMalloc(&b->d_maps, maps.size() * sizeof(BatchMap)) != hipSuccess) {
      (void)hipGetLastError();
      b->cap_maps = 0;
      return fail(RL_ERR_NOMEM, "batch descriptor allocation failed");
    }
    b->cap_maps = maps.size();
  }
  if (b->cap_items < flat.size()) {
    if (b->d_items) HIP_TRY(hipFree(b->d_items));
    b->d_items = nullptr;
    if (hipMalloc(&b->d_items, flat.size() * sizeof(BatchItem)) != hipSuccess) {
      (void)hipGetLastError();
      b->cap_items = 0;
      return fail(RL_ERR_NOMEM, "batch work-list allocation failed");
    }
    b->cap_items = flat.size();
  }
  HIP_TRY(hipMemcpyAsync(b->d_maps, maps.data(), maps.size() * sizeof(BatchMap), hipMemcpyHostToDevice,
                         c->stream));
  if (!flat.empty())
    HIP_TRY(hipMemcpyAsync(b->d_items, flat.data(), flat.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  b->n_items = (int)flat.size();
  return RL_OK;
}

// Releases the map references of the current job list.
void batch_release_jobs(rl_batch* b) {
  for (const rl_batch_job& j : b->jobs) rl_map_destroy(j.map);
  b->jobs.clear();
  b->gens.clear();
}

// A job's map changed since the descriptors were uploaded (a setter replaced a
// buffer or the world frame): upload them again, which a graph capture cannot
// do.
int batch_refresh(rl_batch* b) {
  bool stale = false;
  for (size_t i = 0; i < b->jobs.size(); ++i) stale = stale || b->jobs[i].map->gen.load() != b->gens[i];
  if (!stale) return RL_OK;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(b->ctx->stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
    return fail(RL_ERR_STATE, "a map of this batch changed since its descriptors were uploaded; "
                              "launch once outside the capture first");
  return batch_upload(b);
}

int batch_launch_locked(rl_batch* b) {
  TRY(batch_refresh(b));
  if (b->n_items == 0) return RL_OK;
  rl_ctx* c = b->ctx;
  int grid = b->grid > 0 && b->grid < b->n_items ? b->grid : b->n_items;
  if (b->order == 2 && grid < b->n_items) {  // keep item i on workgroups of XCD i mod xcds
    grid -= grid % b->xcds;
    if (grid == 0) grid = b->n_items;
  }
  if (b->method == RL_METHOD_BL)
    hipLaunchKernelGGL(k_batch<RL_METHOD_BL>, dim3(grid), dim3(256), 0, c->stream, b->d_maps, b->d_items,
                       b->n_items, c->trig, b->lds_box);
  else if (b->method == RL_METHOD_GLT)
    hipLaunchKernelGGL(k_batch<RL_METHOD_GLT>, dim3(grid), dim3(256), 0, c->stream, b->d_maps, b->d_items,
                       b->n_items, c->trig, b->lds_box);
  else if (b->method == RL_METHOD_CDDT)
    hipLaunchKernelGGL(k_batch<RL_METHOD_CDDT>, dim3(grid), dim3(256), 0, c->stream, b->d_maps, b->d_items,
                       b->n_items, c->trig, b->lds_box);
  else
    hipLaunchKernelGGL(k_batch<RL_METHOD_RM>, dim3(grid), dim3(256), 0, c->stream, b->d_maps, b->d_items,
                       b->n_items, c->trig, b->lds_box);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

// Mean distance-map reads per ray of a job, on up to 64 of its particles.
int batch_probe_reads(rl_batch* b, int i, double* out) {
  const rl_batch_job& j = b->jobs[i];
  *out = 1.0;
  if (b->method != RL_METHOD_RM || j.num_particles == 0) return RL_OK;
  int np = j.num_particles < 64 ? j.num_particles : 64;
  size_t n = (size_t)np * j.num_angles;
  void* d = nullptr;
  TRY(b->ctx->staging(5, n * sizeof(int), &d));
  TRY(launch_cast(j.map, b->method, RL_FRAME_ANGLES, true, j.max_range, j.d_ins, j.d_angles, nullptr,
                  (int*)d, (int)n, j.num_angles));
  std::vector<int> h(n);
  HIP_TRY(hipMemcpyAsync(h.data(), d, n * sizeof(int), hipMemcpyDeviceToHost, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  double s = 0;
  for (int v : h) s += v;
  *out = s / (double)n;
  return RL_OK;
}

}  // namespace

extern "C" {

int rl_batch_create(rl_ctx* c, int method, rl_batch** out) {
  CHECK_ARG(c && out, "null argument");
  TRY(check_method(method));
  rl_batch* b = new rl_batch;
  b->ctx = c;
  {
    int x = 0;
    if (hipDeviceGetAttribute(&x, hipDeviceAttributeNumberOfXccs, c->device) == hipSuccess && x > 0) b->xcds = x;
    (void)hipGetLastError();
  }
  c->refs.fetch_add(1);
  b->method = method;
  *out = b;
  return RL_OK;
}

int rl_batch_destroy(rl_batch* b) {
  if (!b) return RL_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->d_maps) (void)hipFree(b->d_maps);
  if (b->d_items) (void)hipFree(b->d_items);
  rl_ctx* c = b->ctx;
  batch_release_jobs(b);
  delete b;
  return rl_destroy(c);  // drops the batch's reference
}

int rl_batch_set_jobs(rl_batch* b, const rl_batch_job* jobs, int n_jobs) {
  CHECK_ARG(b && (jobs || n_jobs == 0) && n_jobs >= 0, "bad argument");
  for (int i = 0; i < n_jobs; ++i) {
    const rl_batch_job& j = jobs[i];
    if (!j.map || !j.d_ins || !j.d_angles || !j.d_ranges)
      return fail(RL_ERR_ARG, "job %d: null map, poses, angles or ranges", i);
    if (j.map->ctx != b->ctx) return fail(RL_ERR_ARG, "job %d: map belongs to another context", i);
    if (j.num_particles < 0 || j.num_angles < 1) return fail(RL_ERR_ARG, "job %d: bad sizes", i);
    if ((long long)j.num_particles * j.num_angles > 0x7fffffffLL) return fail(RL_ERR_ARG, "job %d: too many rays", i);
    if (j.d_obs && (!j.d_weights || !j.map->d_table))
      return fail(RL_ERR_ARG, "job %d: the sensor model needs weights and a sensor table", i);
    if (check_map(j.map, b->method, j.max_range) != RL_OK) {
      std::string msg = g_err;
      return fail(RL_ERR_STATE, "job %d: %s", i, msg.c_str());
    }
  }
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  TRY(use_device(b->ctx));
  // the batch keeps its maps alive: a map destroyed by its caller is freed
  // when the batch lets it go (new jobs or rl_batch_destroy)
  for (int i = 0; i < n_jobs; ++i) jobs[i].map->refs.fetch_add(1);
  batch_release_jobs(b);
  b->jobs.assign(jobs, jobs + n_jobs);
  b->reads_per_ray.assign(n_jobs, 1.0);
  for (int i = 0; i < n_jobs; ++i) TRY(batch_probe_reads(b, i, &b->reads_per_ray[i]));
  return batch_upload(b);
}

int rl_batch_set_config(rl_batch* b, int rays_per_wave, int lds_box, int order, int grid) {
  CHECK_ARG(b, "batch is NULL");
  CHECK_ARG(rays_per_wave == 16 || rays_per_wave == 32 || rays_per_wave == 64, "rays_per_wave must be 16, 32 or 64");
  CHECK_ARG(order >= 0 && order <= 2, "order must be 0, 1 or 2");
  CHECK_ARG(grid >= 0, "grid must be >= 0");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  TRY(use_device(b->ctx));
  b->rays_per_wave = rays_per_wave;
  b->lds_box = lds_box != 0;
  b->order = order;
  b->grid = grid;
  return batch_upload(b);
}

int rl_batch_get_config(rl_batch* b, int* rays_per_wave, int* lds_box, int* order, int* grid) {
  CHECK_ARG(b, "batch is NULL");
  if (rays_per_wave) *rays_per_wave = b->rays_per_wave;
  if (lds_box) *lds_box = b->lds_box;
  if (order) *order = b->order;
  if (grid) *grid = b->grid;
  return RL_OK;
}

int rl_batch_launch(rl_batch* b) {
  CHECK_ARG(b, "batch is NULL");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  TRY(use_device(b->ctx));
  return batch_launch_locked(b);
}

int rl_batch_autotune(rl_batch* b, int reps, float* best_ms) {
  CHECK_ARG(b && reps >= 1, "bad argument");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  TRY(use_device(b->ctx));
  rl_ctx* c = b->ctx;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int cu = 256;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cu = prop.multiProcessorCount;
  }
  const int waves[3] = {64, 32, 16};
  const int grids[3] = {0, 8 * cu, 4 * cu};
  float best = 1e30f;
  int bw = b->rays_per_wave, bb = b->lds_box, bo = b->order, bg = b->grid;
  int rc = RL_OK;
  for (int w : waves) {
    for (int box = 0; box < 2 && rc == RL_OK; ++box) {
      if (box && w != 64) continue;  // the LDS box needs one particle per wave
      for (int order = 0; order < 3 && rc == RL_OK; ++order) {
        for (int g : grids) {
          b->rays_per_wave = w;
          b->lds_box = box;
          b->order = order;
          b->grid = g;
          if ((rc = batch_upload(b)) != RL_OK) break;
          if ((rc = batch_launch_locked(b)) != RL_OK) break;  // warm-up
          if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(RL_ERR_HIP, "event record"); break; }
          for (int r = 0; r < reps && rc == RL_OK; ++r) rc = batch_launch_locked(b);
          if (rc != RL_OK) break;
          if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(RL_ERR_HIP, "event synchronize");
            break;
          }
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          ms /= reps;
          if (ms < best) {
            best = ms;
            bw = w, bb = box, bo = order, bg = g;
          }
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != RL_OK) return rc;
  b->rays_per_wave = bw, b->lds_box = bb, b->order = bo, b->grid = bg;
  if (best_ms) *best_ms = best;
  return batch_upload(b);
}

}  // extern "C"
