// Batched query-vs-ref jobs (include/ppsk.h, ppsk_batch_*): plan building,
// graph capture of the persistent launch, and the tile-shape autotuner.
//
// Plans (one captured graph each), by index:
//   2, 4  every job in 64 x 256 / 32 x 256 tiles, on that shape's own
//         persistent kernel;
//   5     every job in 16 x 256 tiles, on the mixed-shape kernel;
//   6     each job's rows split over the three tile heights (row tiles laid
//         from its first row, e.g. 15 x 64 + 32 + 16 rows for 1,000
//         queries), on the mixed-shape kernel.
// The autotuner times plans 2, 4 and 5 on the batch itself, derives each
// shape's throughput on padded work from those runs, splits every job's rows
// for the least estimated time (padded work / throughput), times that plan
// too and keeps the fastest of the four.
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
namespace {
// Rows [row_begin, row_end) of a job in row tiles of one shape.
struct Seg {
  int shape;
  int64_t row_begin, row_end;
  bool operator==(const Seg& o) const {
    return shape == o.shape && row_begin == o.row_begin && row_end == o.row_end;
  }
};
using Split = std::vector<Seg>;
}  // namespace

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
  Plan plans[ppsk::BATCH_MIXED + 1];
  int variant = 0;                      // chosen plan (0: not chosen yet)
  double tune_ms[ppsk::BATCH_MIXED + 1] = {};
  std::vector<Split> job_split;         // plan 6: row split per job
  uint64_t signature = 0;
};

namespace {

// plans a batch can run (see the top of this file); the first three are the
// uniform shapes the autotuner times
const int kBatchPlans[] = {2, 4, 5, ppsk::BATCH_MIXED};
const int kUniformPlans[] = {2, 4, 5};

struct Tuned {
  uint64_t signature;
  int plan;
  std::vector<Split> job_split;
};
std::mutex g_tune_mu;
std::vector<Tuned> g_tuned;  // job-shape signature -> plan (and per-job splits)

// Kernel that runs a plan.
int plan_kernel(int plan) { return plan == 2 || plan == 4 ? plan : ppsk::BATCH_MIXED; }

// Row segments of job j in a plan.
Split plan_split(const ppsk_batch* b, int plan, int j) {
  if (plan == ppsk::BATCH_MIXED) return b->job_split[j];
  return {{plan, 0, b->jobs[j].query->n}};
}

// Column tiles kept for the row tile starting at row0 (self jobs: the upper
// triangle, as make_tiles keeps it: (ct + 1) * tc - 1 > row0).
int64_t kept_cols(const ppsk_job& jb, int v, int64_t row0) {
  const int tc = ppsk::variant_tile_cols(v);
  const int64_t nct = (jb.ref->n + tc - 1) / tc;
  if (jb.query != jb.ref) return nct;
  return nct - std::min<int64_t>(nct, (row0 + 1) / tc);
}

// Tiles of one segment (padding included).
int64_t seg_tiles(const ppsk_job& jb, const Seg& s) {
  const int tr = ppsk::variant_tile_rows(s.shape);
  const int64_t b0 = s.row_begin / tr, b1 = (s.row_end + tr - 1) / tr;
  if (jb.query != jb.ref) return (b1 - b0) * kept_cols(jb, s.shape, 0);
  int64_t t = 0;
  for (int64_t bi = b0; bi < b1; ++bi) t += kept_cols(jb, s.shape, bi * tr);
  return t;
}

int64_t plan_tiles(const ppsk_batch* b, int plan, int j) {
  int64_t t = 0;
  for (const Seg& s : plan_split(b, plan, j)) t += seg_tiles(b->jobs[j], s);
  return t;
}

// int32 matching operations of job j's padded tiles in a plan.
double plan_ops(const ppsk_batch* b, int plan, int j) {
  const ppsk_job& jb = b->jobs[j];
  double ops = 0.0;
  for (const Seg& s : plan_split(b, plan, j)) {
    ops += (double)seg_tiles(jb, s) * ppsk::variant_tile_rows(s.shape) *
           ppsk::variant_tile_cols(s.shape);
  }
  return ops * jb.ref->n_k * jb.ref->s64 * 30.0;
}

// Row split of one job for plan 6: a sequence of row tiles of the uniform
// shapes, each starting at a multiple of its own height, covering the job's
// rows with the least estimated time (tile work / shape throughput in ops
// per ms).  Dynamic programming over the rows in steps of the smallest tile
// height, from the last row up; ties go to the taller tile, so the split is
// taller tiles first.  With no measured throughputs (rate <= 0) every shape
// counts as equally fast: the split of least padding.
Split choose_split(const ppsk_job& jb, const double* rate) {
  const int unit = ppsk::variant_tile_rows(5);
  const int64_t n = jb.query->n, steps = (n + unit - 1) / unit;
  const double per_pair = (double)jb.ref->n_k * jb.ref->s64 * 30.0;
  std::vector<double> cost(steps + 1, 0.0);
  std::vector<int> pick(steps, 2);
  for (int64_t p = steps - 1; p >= 0; --p) {
    double best = -1.0;
    for (int v : kUniformPlans) {  // tallest first: it wins ties
      const int tr = ppsk::variant_tile_rows(v), tc = ppsk::variant_tile_cols(v);
      const int64_t len = tr / unit;
      if (p % len) continue;
      const double r = rate[v] > 0 ? rate[v] : 1.0;
      const double c = (double)kept_cols(jb, v, p * unit) * tr * tc * per_pair / r +
                       cost[std::min(steps, p + len)];
      if (best < 0.0 || c < best) {
        best = c;
        pick[p] = v;
      }
    }
    cost[p] = best;
  }
  Split out;
  for (int64_t p = 0; p < steps;) {
    const int v = pick[p];
    const int64_t len = ppsk::variant_tile_rows(v) / unit;
    const int64_t r0 = p * unit, r1 = std::min(n, (p + len) * unit);
    if (!out.empty() && out.back().shape == v) {
      out.back().row_end = r1;
    } else {
      out.push_back({v, r0, r1});
    }
    p += len;
  }
  return out;
}

std::vector<Split> choose_splits(const ppsk_batch* b, const double* rate) {
  std::vector<Split> s(b->n_jobs);
  for (int j = 0; j < b->n_jobs; ++j) s[j] = choose_split(b->jobs[j], rate);
  return s;
}

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

// Items of a plan.  Every (job, row segment) unit's tiles (grouped order) are
// cut into BATCH_QUEUES contiguous runs, run q of every unit going to queue
// q, so each queue (and the XCD that drains it) gets an equal share of every
// job and walks compact blocks of it.  Units enter the queues most expensive
// tile first, leaving the cheap tiles for the end of the launch.
int build_plan(ppsk_batch* b, int plan) {
  auto& pl = b->plans[plan];
  if (pl.exec) return PPSK_OK;
  const int kernel = plan_kernel(plan);
  struct Unit {
    int job;
    Seg seg;
    int64_t cost;  // matching work of one tile
  };
  std::vector<Unit> units;
  for (int j = 0; j < b->n_jobs; ++j) {
    for (const Seg& sg : plan_split(b, plan, j)) {
      units.push_back({j, sg, (int64_t)b->jobs[j].ref->n_k * b->jobs[j].ref->s64 *
                                  ppsk::variant_tile_rows(sg.shape)});
    }
  }
  std::stable_sort(units.begin(), units.end(),
                   [](const Unit& x, const Unit& y) { return x.cost > y.cost; });
  std::vector<std::vector<ppsk::BatchItem>> q(ppsk::BATCH_QUEUES);
  std::vector<size_t> gs;
  for (const Unit& u : units) {
    const ppsk_job& jb = b->jobs[u.job];
    const bool self = jb.query == jb.ref;
    const int shape = u.seg.shape;
    const int tr = ppsk::variant_tile_rows(shape), tc = ppsk::variant_tile_cols(shape);
    auto tiles = make_tiles(jb.query->n, jb.ref->n, tr, tc, self, u.seg.row_begin,
                            u.seg.row_end, tile_group(tr), &gs);
    const size_t n = tiles.size();
    for (int qi = 0; qi < ppsk::BATCH_QUEUES; ++qi) {
      const size_t a = n * qi / ppsk::BATCH_QUEUES, e = n * (qi + 1) / ppsk::BATCH_QUEUES;
      for (size_t x = a; x < e; ++x) q[qi].push_back({u.job, tiles[x].bi, tiles[x].ct, shape});
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
  HIP_TRY(ppsk::batch_grid(kernel, &pl.grid));
  if (!b->scratch) {
    // sized once for every kernel: captured graphs keep this address (the
    // mixed-shape kernel's blocks hold the largest shape's counts)
    size_t need = 0;
    for (int k : {2, 4, ppsk::BATCH_MIXED}) {
      int g = 0;
      HIP_TRY(ppsk::batch_grid(k, &g));
      const int v = k == ppsk::BATCH_MIXED ? 2 : k;
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
  a.ctl = c->ctl;
  // both launches captured once, replayed by every run; the counters are
  // reset by a kernel (no memset node)
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  hipError_t e = ppsk::launch_batch_reset(b->d_queue, b->d_fail, b->n_jobs, c->stream);
  if (e == hipSuccess) e = ppsk::launch_batch(a, kernel, pl.grid, c->stream);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (e == hipSuccess) e = ec;
  if (e == hipSuccess) e = hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0);
  pl.graph = g;
  if (e != hipSuccess) return fail(PPSK_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
  return PPSK_OK;
}

int run_plan(ppsk_batch* b, int plan, double* ms) {
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(b->device, &c)) return rc;
  if (int rc = build_plan(b, plan)) return rc;
  if (int rc = cancel_status(b->device)) return rc;
  uint64_t items = 0;
  for (int j = 0; j < b->n_jobs; ++j) {
    items += (uint64_t)plan_tiles(b, plan, j) * b->jobs[j].ref->n_k;
  }
  CallScope scope(c, items);
  hipEvent_t e0 = pool_event(c, 0), e1 = pool_event(c, 1);
  if (!e0 || !e1) return fail(PPSK_ERR_HIP, "event creation failed");
  HIP_TRY(hipEventRecord(e0, c->stream));
  HIP_TRY(hipGraphLaunch(b->plans[plan].exec, c->stream));
  HIP_TRY(hipEventRecord(e1, c->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, e0, e1));
  *ms = f;
  // the persistent kernel takes no more items once a cancellation is seen
  return cancel_status(b->device);
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
  const double none[ppsk::BATCH_MIXED + 1] = {};
  b->job_split = choose_splits(b, none);  // least padding until tuned
  *out = b;
  return PPSK_OK;
}

int ppsk_batch_run(ppsk_batch* b, int32_t variant, uint64_t* n_fail, int32_t* used_variant,
                   double* kernel_ms) {
  if (!b) return fail(PPSK_ERR_ARG, "batch is NULL");
  if (variant != 0) {
    bool ok = false;
    for (int v : kBatchPlans) ok |= (v == variant);
    if (!ok) return fail(PPSK_ERR_ARG, "no batch plan " + std::to_string(variant) +
                                           " (2, 4, 5, 6, or 0 to autotune)");
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(b->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  int v = variant;
  double ms = 0.0;
  if (v == 0 && b->variant == 0) {
    if (const char* env = std::getenv("PPSK_BATCH_VARIANT")) {
      const int ev = std::atoi(env);
      for (int x : kBatchPlans) if (x == ev) b->variant = ev;
    }
  }
  if (v == 0 && b->variant == 0) {
    std::lock_guard<std::mutex> tl(g_tune_mu);
    for (auto& t : g_tuned) {
      if (t.signature == b->signature) {
        b->variant = t.plan;
        if (t.plan == ppsk::BATCH_MIXED && b->plans[t.plan].exec == nullptr) {
          b->job_split = t.job_split;
        }
      }
    }
  }
  if (v == 0 && b->variant == 0) {
    // autotune: each uniform shape runs the whole batch twice (the first run
    // pays for graph set-up and cold caches) and its second run gives the
    // shape's throughput on padded work; the row splits built from those
    // throughputs are timed the same way; the fastest plan wins.  Each run
    // produces the complete, correct output.
    double rate[ppsk::BATCH_MIXED + 1] = {};
    auto time_plan = [&](int plan) -> int {
      double t0 = 0.0, t1 = 0.0;
      if (int rc = run_plan(b, plan, &t0)) return rc;
      if (int rc = run_plan(b, plan, &t1)) return rc;
      b->tune_ms[plan] = t1;
      if (b->variant == 0 || t1 < b->tune_ms[b->variant]) b->variant = plan;
      return PPSK_OK;
    };
    for (int cand : kUniformPlans) {
      if (int rc = time_plan(cand)) return rc;
      double ops = 0.0;
      for (int j = 0; j < b->n_jobs; ++j) ops += plan_ops(b, cand, j);
      rate[cand] = b->tune_ms[cand] > 0 ? ops / b->tune_ms[cand] : 0.0;
    }
    std::vector<Split> splits = choose_splits(b, rate);
    if (b->plans[ppsk::BATCH_MIXED].exec && splits != b->job_split) {
      // a per-job plan captured with other shapes: rebuilt on its next use
      auto& pl = b->plans[ppsk::BATCH_MIXED];
      hipGraphExecDestroy(pl.exec);
      hipGraphDestroy(pl.graph);
      hipFree(pl.items);
      pl = ppsk_batch::Plan();
    }
    b->job_split = splits;
    // every job in one segment of one shape: plan 6 is that uniform plan,
    // already timed
    const bool uniform = std::all_of(splits.begin(), splits.end(), [&](const Split& x) {
      return x.size() == 1 && x[0].shape == splits[0][0].shape;
    });
    if (!uniform) {
      if (int rc = time_plan(ppsk::BATCH_MIXED)) return rc;
    }
    std::lock_guard<std::mutex> tl(g_tune_mu);
    g_tuned.push_back({b->signature, b->variant, b->job_split});
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

int ppsk_batch_tuning(const ppsk_batch* b, int32_t* variant, double* ms_per_variant,
                      int32_t* row_tiles) {
  if (!b) return fail(PPSK_ERR_ARG, "batch is NULL");
  if (variant) *variant = b->variant;
  if (row_tiles) {
    for (int j = 0; j < b->n_jobs; ++j) {
      int32_t* t = row_tiles + 3 * (size_t)j;
      t[0] = t[1] = t[2] = 0;
      for (const Seg& sg : b->job_split[j]) {
        const int tr = ppsk::variant_tile_rows(sg.shape);
        const int x = sg.shape == 2 ? 0 : (sg.shape == 4 ? 1 : 2);
        t[x] += (int32_t)((sg.row_end + tr - 1) / tr - sg.row_begin / tr);
      }
    }
  }
  if (ms_per_variant) {
    for (int v = 0; v <= ppsk::BATCH_MIXED; ++v) ms_per_variant[v] = b->tune_ms[v];
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
