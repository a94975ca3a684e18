// Batched query-vs-ref jobs (include/ppsk.h, ppsk_batch_*): plan building,
// graph capture of the persistent launch, and the tile-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "capi_common.h"

using namespace ppsk::capi;

// ---------------------------------------------------------------------------
// Batched jobs: one persistent launch over many query-vs-ref jobs.
// ---------------------------------------------------------------------------
struct ppsk_batch {
  struct Plan {  // one tile shape
    ppsk::BatchItem* items = nullptr;
    int64_t n_items = 0;
    int64_t queue_begin[ppsk::BATCH_QUEUES + 1] = {};
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  int device = 0;
  int32_t n_jobs = 0;
  int32_t max_k = 0;
  std::vector<ppsk_job> jobs;
  std::vector<ppsk::DistParams> params;
  std::vector<float*> tables;           // per-job device MC tables (or null)
  ppsk::DistParams* d_params = nullptr;
  unsigned long long* d_fail = nullptr; // [n_jobs][2]
  uint32_t* d_queue = nullptr;          // [BATCH_QUEUES]
  uint16_t* scratch = nullptr;
  size_t scratch_bytes = 0;
  Plan plans[ppsk::N_VARIANTS + 1];
  int variant = 0;                      // chosen tile shape (0: not chosen yet)
  double tune_ms[ppsk::N_VARIANTS + 1] = {};
  uint64_t signature = 0;
};

namespace {

// tile shapes the persistent kernel is built for, tried by the autotuner
const int kBatchVariants[] = {2, 4};

std::mutex g_tune_mu;
std::vector<std::pair<uint64_t, int>> g_tuned;  // job-shape signature -> tile shape

uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  return h;
}

// Shape signature of a batch: per job (rows, cols, n_k, s64) in order.
uint64_t batch_signature(const ppsk_batch* b) {
  uint64_t h = (uint64_t)b->n_jobs;
  for (const auto& j : b->jobs) {
    h = mix64(h, (uint64_t)j.query->n);
    h = mix64(h, (uint64_t)j.ref->n);
    h = mix64(h, (uint64_t)j.ref->n_k);
    h = mix64(h, (uint64_t)j.ref->s64);
  }
  return h;
}

// Items of one tile shape.  Every job's tiles (grouped order) are cut into
// BATCH_QUEUES contiguous runs, run q of every job going to queue q, so each
// queue (and the XCD that drains it) gets an equal share of every job and
// walks compact blocks of it.  Jobs enter the queues most expensive tile
// first, leaving the cheap tiles for the end of the launch.
int build_plan(ppsk_batch* b, int variant) {
  auto& pl = b->plans[variant];
  if (pl.exec) return PPSK_OK;
  const int tr = ppsk::variant_tile_rows(variant);
  const int tc = ppsk::variant_tile_cols(variant);
  std::vector<int> order(b->n_jobs);
  for (int j = 0; j < b->n_jobs; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const int64_t cx = (int64_t)b->jobs[x].ref->n_k * b->jobs[x].ref->s64;
    const int64_t cy = (int64_t)b->jobs[y].ref->n_k * b->jobs[y].ref->s64;
    return cx > cy;
  });
  std::vector<std::vector<ppsk::BatchItem>> q(ppsk::BATCH_QUEUES);
  std::vector<size_t> gs;
  for (int j : order) {
    const ppsk_job& jb = b->jobs[j];
    const bool self = jb.query == jb.ref;
    auto tiles = make_tiles(jb.query->n, jb.ref->n, tr, tc, self, 0, jb.query->n, tile_group(), &gs);
    const size_t n = tiles.size();
    for (int qi = 0; qi < ppsk::BATCH_QUEUES; ++qi) {
      const size_t a = n * qi / ppsk::BATCH_QUEUES, e = n * (qi + 1) / ppsk::BATCH_QUEUES;
      for (size_t x = a; x < e; ++x) q[qi].push_back({j, tiles[x].bi, tiles[x].ct, 0});
    }
  }
  std::vector<ppsk::BatchItem> all;
  pl.queue_begin[0] = 0;
  for (int qi = 0; qi < ppsk::BATCH_QUEUES; ++qi) {
    all.insert(all.end(), q[qi].begin(), q[qi].end());
    pl.queue_begin[qi + 1] = (int64_t)all.size();
  }
  pl.n_items = (int64_t)all.size();
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(b->device, &c)) return rc;
  HIP_TRY(ppsk::batch_grid(variant, &pl.grid));
  if (!b->scratch) {
    // sized once for every shape: captured graphs keep this address
    size_t need = 0;
    for (int v : kBatchVariants) {
      int g = 0;
      HIP_TRY(ppsk::batch_grid(v, &g));
      need = std::max(need, (size_t)g * b->max_k * ppsk::variant_tile_rows(v) *
                                ppsk::variant_tile_cols(v) * sizeof(uint16_t));
    }
    if (int rc = grow((void**)&b->scratch, &b->scratch_bytes, need)) return rc;
  }
  HIP_TRY(hipMalloc(&pl.items, std::max<size_t>(all.size(), 1) * sizeof(ppsk::BatchItem)));
  HIP_TRY(hipMemcpy(pl.items, all.data(), all.size() * sizeof(ppsk::BatchItem),
                    hipMemcpyHostToDevice));
  ppsk::BatchArgs a;
  std::memset(&a, 0, sizeof(a));
  a.jobs = b->d_params;
  a.items = pl.items;
  for (int qi = 0; qi <= ppsk::BATCH_QUEUES; ++qi) a.queue_begin[qi] = pl.queue_begin[qi];
  a.queue_next = b->d_queue;
  a.scratch = b->scratch;
  a.max_k = b->max_k;
  // both launches captured once, replayed by every run; the counters are
  // reset by a kernel (no memset node)
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  hipError_t e = ppsk::launch_batch_reset(b->d_queue, b->d_fail, b->n_jobs, c->stream);
  if (e == hipSuccess) e = ppsk::launch_batch(a, variant, pl.grid, c->stream);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (e == hipSuccess) e = ec;
  if (e == hipSuccess) e = hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0);
  pl.graph = g;
  if (e != hipSuccess) return fail(PPSK_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
  return PPSK_OK;
}

int run_plan(ppsk_batch* b, int variant, double* ms) {
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(b->device, &c)) return rc;
  if (int rc = build_plan(b, variant)) return rc;
  hipEvent_t e0 = pool_event(c, 0), e1 = pool_event(c, 1);
  if (!e0 || !e1) return fail(PPSK_ERR_HIP, "event creation failed");
  HIP_TRY(hipEventRecord(e0, c->stream));
  HIP_TRY(hipGraphLaunch(b->plans[variant].exec, c->stream));
  HIP_TRY(hipEventRecord(e1, c->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, e0, e1));
  *ms = f;
  return PPSK_OK;
}

}  // namespace

extern "C" {

int ppsk_batch_create(int device_id, const ppsk_job* jobs, int32_t n_jobs, ppsk_batch** out) {
  if (!out) return fail(PPSK_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_jobs < 1 || !jobs) return fail(PPSK_ERR_ARG, "a batch needs at least one job");
  for (int j = 0; j < n_jobs; ++j) {
    const ppsk_job& jb = jobs[j];
    const std::string at = " (job " + std::to_string(j) + ")";
    if (!jb.ref || !jb.query) return fail(PPSK_ERR_ARG, "job without sets" + at);
    if (jb.ref->device != device_id || jb.query->device != device_id) {
      return fail(PPSK_ERR_ARG, "job sets on another device" + at);
    }
    if (int rc = check_compatible(jb.ref, jb.query)) {
      g_last_error += at;
      return rc;
    }
    if (jb.out_kind < PPSK_OUT_CORE_ACC || jb.out_kind > PPSK_OUT_COUNTS) {
      return fail(PPSK_ERR_ARG, "unknown out_kind" + at);
    }
    if (jb.out_kind == PPSK_OUT_CORE_ACC && jb.ref->n_k < 2) {
      return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths" + at);
    }
    if (!jb.out) return fail(PPSK_ERR_ARG, "out is NULL" + at);
    if (int rc = check_random(jb.query, jb.ref, RandomIn{jb.random_mode, jb.n_clusters, jb.mc_table})) {
      g_last_error += at;
      return rc;
    }
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  auto* b = new ppsk_batch();
  b->device = device_id;
  b->n_jobs = n_jobs;
  b->jobs.assign(jobs, jobs + n_jobs);
  b->params.resize(n_jobs);
  b->tables.assign(n_jobs, nullptr);
  int rc = PPSK_OK;
  hipError_t e = hipMalloc(&b->d_fail, (size_t)n_jobs * 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&b->d_queue, ppsk::BATCH_QUEUES * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->d_params, (size_t)n_jobs * sizeof(ppsk::DistParams));
  if (e != hipSuccess) rc = fail(PPSK_ERR_MEMORY, "batch buffers");
  for (int j = 0; j < n_jobs && rc == PPSK_OK; ++j) {
    const ppsk_job& jb = jobs[j];
    const RandomIn rnd{jb.random_mode, jb.n_clusters, jb.mc_table};
    if ((rc = upload_table(c, jb.query, rnd, &b->tables[j]))) break;
    ppsk::DistParams& p = b->params[j];
    fill_common(p, jb.query, jb.ref, rnd, jb.tolerance);
    const bool self = jb.query == jb.ref;  // self job: condensed upper triangle
    p.self_mode = self ? 1 : 0;
    p.full_square = 0;
    p.out_kind = jb.out_kind;
    p.out = jb.out;
    p.mc_table = b->tables[j];
    p.mc_ref_is_col = self ? 0 : 1;  // query mode: random_match(r1 = ref, q)
    p.fail_count = b->d_fail + 2 * j;
    p.fail_first = b->d_fail + 2 * j + 1;
    b->max_k = std::max(b->max_k, jb.ref->n_k);
  }
  if (rc == PPSK_OK) {
    e = hipMemcpyAsync(b->d_params, b->params.data(), (size_t)n_jobs * sizeof(ppsk::DistParams),
                       hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(PPSK_ERR_HIP, std::string("batch upload: ") + hipGetErrorString(e));
  }
  if (rc != PPSK_OK) {
    ppsk_batch_free(b);
    return rc;
  }
  b->signature = batch_signature(b);
  *out = b;
  return PPSK_OK;
}

int ppsk_batch_run(ppsk_batch* b, int32_t variant, uint64_t* n_fail, int32_t* used_variant,
                   double* kernel_ms) {
  if (!b) return fail(PPSK_ERR_ARG, "batch is NULL");
  if (variant != 0) {
    bool ok = false;
    for (int v : kBatchVariants) ok |= (v == variant);
    if (!ok) return fail(PPSK_ERR_ARG, "tile shape " + std::to_string(variant) +
                                           " has no batch kernel (2 or 4, or 0 to autotune)");
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(b->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  int v = variant;
  double ms = 0.0;
  if (v == 0 && b->variant == 0) {
    if (const char* env = std::getenv("PPSK_BATCH_VARIANT")) {
      const int ev = std::atoi(env);
      for (int x : kBatchVariants) if (x == ev) b->variant = ev;
    }
  }
  if (v == 0 && b->variant == 0) {
    std::lock_guard<std::mutex> tl(g_tune_mu);
    for (auto& kv : g_tuned) if (kv.first == b->signature) b->variant = kv.second;
  }
  if (v == 0 && b->variant == 0) {
    // autotune: every candidate shape runs the whole batch twice (the first
    // run pays for graph set-up and cold caches); the fastest second run
    // wins.  Each run produces the complete, correct output.
    double best = 0.0;
    for (int cand : kBatchVariants) {
      double t0 = 0.0, t1 = 0.0;
      if (int rc = run_plan(b, cand, &t0)) return rc;
      if (int rc = run_plan(b, cand, &t1)) return rc;
      b->tune_ms[cand] = t1;
      if (b->variant == 0 || t1 < best) {
        best = t1;
        b->variant = cand;
      }
    }
    std::lock_guard<std::mutex> tl(g_tune_mu);
    g_tuned.emplace_back(b->signature, b->variant);
    v = b->variant;
    ms = b->tune_ms[v];
  } else {
    if (v == 0) v = b->variant;
    if (int rc = run_plan(b, v, &ms)) return rc;
  }
  std::vector<unsigned long long> f((size_t)b->n_jobs * 2);
  HIP_TRY(hipMemcpy(f.data(), b->d_fail, f.size() * sizeof(unsigned long long),
                    hipMemcpyDeviceToHost));
  uint64_t total = 0;
  for (int j = 0; j < b->n_jobs; ++j) {
    if (n_fail) n_fail[j] = f[2 * j];
    total += f[2 * j];
  }
  if (used_variant) *used_variant = v;
  if (kernel_ms) *kernel_ms = ms;
  g_last_ms = ms;
  g_last_launches = 1;
  g_last_items = b->plans[v].n_items;
  if (total) return fail(PPSK_ERR_FIT, "Fitting k-mer gradient failed for " + std::to_string(total) + " pair(s)");
  return PPSK_OK;
}

int ppsk_batch_tuning(const ppsk_batch* b, int32_t* variant, double* ms_per_variant) {
  if (!b) return fail(PPSK_ERR_ARG, "batch is NULL");
  if (variant) *variant = b->variant;
  if (ms_per_variant) {
    for (int v = 0; v <= ppsk::N_VARIANTS; ++v) ms_per_variant[v] = b->tune_ms[v];
  }
  return PPSK_OK;
}

int ppsk_batch_free(ppsk_batch* b) {
  if (!b) return PPSK_OK;
  hipSetDevice(b->device);
  for (auto& pl : b->plans) {
    if (pl.exec) hipGraphExecDestroy(pl.exec);
    if (pl.graph) hipGraphDestroy(pl.graph);
    if (pl.items) hipFree(pl.items);
  }
  for (float* t : b->tables) if (t) hipFree(t);
  if (b->d_params) hipFree(b->d_params);
  if (b->d_fail) hipFree(b->d_fail);
  if (b->d_queue) hipFree(b->d_queue);
  if (b->scratch) hipFree(b->scratch);
  delete b;
  return PPSK_OK;
}

}  // extern "C"
