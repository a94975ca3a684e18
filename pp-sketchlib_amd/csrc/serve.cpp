// Serving (include/ppsk.h, ppsk_server_*): a reference set stays resident in
// HBM and query batches arrive at any time.  Continuous batching:
//
//  * every request waits in one FIFO queue;
//  * n_slots worker threads, each with its own stream, repeatedly fill one
//    step with query rows, oldest request first: at most max_rows rows and
//    the step's device memory, carved from one arena allocated at start-up.
//    A request that does not fit the step's remaining rows contributes a
//    piece (its next rows) and stays at the head of the queue, so a large
//    request spreads over several steps, which other slots run at the same
//    time, and no step grows beyond max_rows;
//  * requests that arrive while other steps are on the GPU are admitted as
//    soon as a slot and memory free up, and steps of different slots overlap
//    on the device;
//  * the query rows of a step share the row tiles of one query-vs-reference
//    distance pass (the dense band never leaves the device), followed by kNN
//    selection and threshold compaction straight into page-locked host
//    buffers;
//  * each piece is a contiguous slice of the step's rows, so its results are
//    that slice of the step's results, re-indexed to the request; a request
//    is complete when the step of its last piece has finished.
//
// The step budget bounds latency: the persistent distance kernel holds every
// CU for its whole pass, so work queued behind a step (another step's query
// striping, kNN, threshold kernels) waits for it.  Small steps keep that wait
// short; a few dozen query rows already keep the pass compute-bound.
//
// Results do not depend on the steps a request lands in: every output value
// is a function of its own (query, reference) pair only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "capi_common.h"
#include "select_kernels.h"

using namespace ppsk::capi;

namespace {

constexpr size_t ARENA_ALIGN = 256;

size_t align_up(size_t x) { return (x + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

// First-fit allocator over one device allocation (host-side bookkeeping
// only, so steps never call hipMalloc).
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length, coalesced
  size_t used = 0, peak = 0;

  void init(char* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    free_[0] = n;
  }
  bool fits(size_t n) const {
    for (auto& f : free_) if (f.second >= n) return true;
    return false;
  }
  bool alloc(size_t n, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n;
      const size_t at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used += n;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Threshold hits of one piece, query index relative to the request.
struct ThrPart {
  std::vector<int64_t> i, j;
  std::vector<float> d;
};

}  // namespace

struct ppsk_request {
  int64_t ticket = 0;
  int64_t n_q = 0;
  std::vector<uint64_t> raw;  // [n_q][n_k][s64 * 14]
  std::vector<double> jv;     // [n_q][n_k] equal-base terms
  std::vector<uint16_t> cl;   // [n_q] Monte Carlo cluster ids
  ppsk_result res;            // public view of the vectors below
  // kNN lists are sized at submit; each piece fills its own rows in place
  std::vector<int64_t> ki, kj, ti, tj;
  std::vector<float> kd, td;
  int64_t next_row = 0;   // first row not yet given to a step
  int64_t rows_done = 0;  // rows whose step has finished
  std::map<int64_t, ThrPart> thr_parts;  // by the piece's first row
  bool done = false;
  int status = PPSK_OK;
  std::string error;
};

struct ppsk_server {
  ppsk_set* ref = nullptr;
  int device = 0;
  int32_t random_mode = 0;
  int32_t n_clusters = 0;     // Monte Carlo table [n_k][C][C]
  float* d_table = nullptr;   // its device copy
  double tolerance = 0;
  int32_t kNN = 0, knn_col = 0, thr_col = 0;
  float cutoff = 0;
  int32_t max_rows = 0;
  int variant = 0;  // tile shape; 0: by step size (step_variant)
  int grid2 = 0, grid4 = 0;  // resident workgroups of the persistent kernel per shape
  bool force_persistent = false;  // PPSK_PERSISTENT=1: every step of shape 2 / 4
  size_t per_sample = 0;
  size_t scratch_budget = 0;
  Arena arena;
  void* arena_mem = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ppsk_request*> pending;
  std::map<int64_t, ppsk_request*> tickets;  // submitted, not yet collected
  int64_t next_ticket = 1;
  int32_t waiters = 0;  // threads inside ppsk_server_wait
  bool stop = false;
  std::vector<std::thread> workers;
  std::chrono::steady_clock::time_point t0;
  // statistics
  int64_t steps = 0, rows_done = 0, requests_done = 0;
  int32_t inflight = 0, peak_inflight = 0;
  double gpu_ms = 0;  // summed distance-kernel time of all steps
};

namespace {

// Tile shape of a step: the shortest row tile that holds the step's rows
// (a step of a few queries would otherwise compute mostly padding rows).
int step_variant(const ppsk_server* s, int64_t rows) {
  if (s->variant) return s->variant;
  if (rows <= 16) return 5;
  if (rows <= 32) return 4;
  return 2;
}

// Device memory of one step, carved from the arena.  A step with at least
// two tiles per resident workgroup runs the persistent tile kernel (whole
// tiles from work queues, per-workgroup count blocks), its last partial round
// as (tile, k-mer length) items when that round would leave most workgroup
// slots idle (tail tiles, per-tile count scratch); a smaller step is split
// per (tile, k-mer length) over more workgroups (per-tile count scratch,
// launched in chunks of chunk_tiles).
struct StepLayout {
  size_t staging, stripes, jv, clusters, band, tiles, scratch, tile_done, fail, fail_rows, params;
  size_t tail_tiles, tail_counts, tail_done, counts, knn_lists, total;
  int64_t nbq, n_tiles, chunk_tiles, tail;
  int segs;      // column segments per row of the threshold kernels
  int knn_segs;  // column segments per row of the kNN candidate lists
  bool persistent;
  int grid;
};

StepLayout layout(const ppsk_server* s, int64_t rows) {
  const ppsk_set* r = s->ref;
  StepLayout L{};
  const int v = step_variant(s, rows);
  const int tr = ppsk::variant_tile_rows(v);
  const int tc = ppsk::variant_tile_cols(v);
  L.nbq = (rows + ppsk::BLK - 1) / ppsk::BLK;
  L.n_tiles = ((rows + tr - 1) / tr) * ((r->n + tc - 1) / tc);
  const size_t per_tile = (size_t)r->n_k * tr * tc * sizeof(uint16_t);
  L.grid = v == 2 ? s->grid2 : (v == 4 ? s->grid4 : 0);
  L.persistent = L.grid > 0 && (s->force_persistent || L.n_tiles >= 2 * (int64_t)L.grid);
  L.chunk_tiles = std::max<int64_t>(1, std::min<int64_t>(L.n_tiles, s->scratch_budget / per_tile));
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align_up(std::max<size_t>(bytes, 1));
    return at;
  };
  L.staging = take((size_t)rows * s->per_sample);
  L.stripes = take((size_t)r->n_k * L.nbq * r->s64 * ppsk::HL * ppsk::BLK * sizeof(uint32_t));
  L.jv = take((size_t)rows * r->n_k * sizeof(double));
  L.clusters = take((size_t)rows * sizeof(uint16_t));
  L.band = take((size_t)rows * r->n * 2 * sizeof(float));
  if (L.persistent) {
    L.tail = (int64_t)tail_split_tiles((size_t)L.n_tiles, L.grid);
    L.tiles = take((size_t)(L.n_tiles - L.tail) * sizeof(ppsk::BatchItem));
    L.scratch = take((size_t)L.grid * per_tile);
    L.tile_done = take(ppsk::BATCH_QUEUES * sizeof(uint32_t));  // queue counters
    L.params = take(sizeof(ppsk::DistParams));
    if (L.tail > 0) {
      L.tail_tiles = take((size_t)L.tail * sizeof(ppsk::Tile));
      L.tail_counts = take((size_t)L.tail * per_tile);
      L.tail_done = take((size_t)L.tail * sizeof(uint32_t));
    }
  } else {
    L.tiles = take((size_t)L.n_tiles * sizeof(ppsk::Tile));
    L.scratch = take((size_t)L.chunk_tiles * per_tile);
    L.tile_done = take((size_t)L.chunk_tiles * sizeof(uint32_t));
    L.params = 0;
  }
  L.fail = take(2 * sizeof(unsigned long long));
  L.fail_rows = take((size_t)rows * sizeof(uint32_t));
  // about 2,048 threshold workgroups, segments of at least 4,096 entries
  L.segs = (int)std::max<int64_t>(1, std::min<int64_t>((2048 + rows - 1) / rows, r->n / 4096));
  L.counts = take((size_t)(rows * L.segs + 1) * sizeof(int64_t));
  // the same segments list kNN candidates (small kNN only: the lists grow
  // with it)
  L.knn_segs = s->kNN > 0 && s->kNN <= 64 ? L.segs : 1;
  L.knn_lists = take(ppsk::knn_rect_workspace(rows, L.knn_segs, std::max(1, s->kNN)));
  L.total = off;
  return L;
}

// Page-locked host buffer that grows on demand.
struct Pinned {
  void* p = nullptr;
  size_t n = 0;
  bool reserve(size_t bytes) {
    if (bytes <= n && p) return true;
    if (p) hipHostFree(p);
    p = nullptr;
    n = 0;
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    n = std::max<size_t>(bytes, 64);
    return true;
  }
  ~Pinned() {
    if (p) hipHostFree(p);
  }
};

struct Worker {
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Pinned knn_i, knn_j, knn_d, counts, thr_i, thr_j, thr_d, fail_rows;
};

// Rows [q0, q0 + n) of a request, carried by one step; thr and nf are the
// piece's results until they are merged into the request.
struct Piece {
  ppsk_request* r;
  int64_t q0, n;
  ThrPart thr;
  uint64_t nf = 0;
};

#define STEP_TRY(expr)                                                 \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess) {                                            \
      err = std::string(#expr) + ": " + hipGetErrorString(_e);         \
      return PPSK_ERR_HIP;                                             \
    }                                                                  \
  } while (0)

// One step over `batch` (the pieces' rows back to back in batch order).
int run_step(ppsk_server* s, Worker& w, std::vector<Piece>& batch, int64_t rows, char* mem,
             const StepLayout& L, std::string& err) {
  const ppsk_set* ref = s->ref;
  const int n_k = ref->n_k;
  uint64_t* staging = (uint64_t*)(mem + L.staging);
  uint32_t* stripes = (uint32_t*)(mem + L.stripes);
  double* jv = (double*)(mem + L.jv);
  uint16_t* qcl = (uint16_t*)(mem + L.clusters);
  float* band = (float*)(mem + L.band);
  ppsk::Tile* d_tiles = (ppsk::Tile*)(mem + L.tiles);
  uint16_t* scratch = (uint16_t*)(mem + L.scratch);
  uint32_t* tile_done = (uint32_t*)(mem + L.tile_done);
  unsigned long long* fail = (unsigned long long*)(mem + L.fail);
  uint32_t* fail_rows = (uint32_t*)(mem + L.fail_rows);

  // queries -> stripes (+ equal-base terms)
  const size_t words = s->per_sample / sizeof(uint64_t);
  int64_t o = 0;
  for (const Piece& pc : batch) {
    const ppsk_request* r = pc.r;
    STEP_TRY(hipMemcpyAsync(staging + (size_t)o * words, r->raw.data() + (size_t)pc.q0 * words,
                            (size_t)pc.n * s->per_sample, hipMemcpyHostToDevice, w.st));
    if (!r->jv.empty()) {
      STEP_TRY(hipMemcpyAsync(jv + (size_t)o * n_k, r->jv.data() + (size_t)pc.q0 * n_k,
                              (size_t)pc.n * n_k * sizeof(double), hipMemcpyHostToDevice, w.st));
    }
    if (!r->cl.empty()) {
      STEP_TRY(hipMemcpyAsync(qcl + o, r->cl.data() + pc.q0, (size_t)pc.n * sizeof(uint16_t),
                              hipMemcpyHostToDevice, w.st));
    }
    o += pc.n;
  }
  STEP_TRY(ppsk::launch_stripe(staging, rows, 0, n_k, ref->s64, (int32_t)L.nbq, stripes, w.st));
  // per-tile completion counters of the per-(tile, k) kernel (the queue
  // counters of a persistent step are reset by launch_batch_reset)
  if (!L.persistent) {
    STEP_TRY(hipMemsetAsync(tile_done, 0, (size_t)L.chunk_tiles * sizeof(uint32_t), w.st));
  }
  STEP_TRY(hipMemsetAsync(fail, 0, sizeof(unsigned long long), w.st));
  STEP_TRY(hipMemsetAsync(fail + 1, 0xFF, sizeof(unsigned long long), w.st));
  STEP_TRY(hipMemsetAsync(fail_rows, 0, (size_t)rows * sizeof(uint32_t), w.st));

  // distances: query rows x every reference, dense band in device memory
  ppsk::DistParams p;
  std::memset(&p, 0, sizeof(p));
  p.row_stripes = stripes;
  p.col_stripes = ref->stripes;
  p.n_rows = rows;
  p.n_cols = ref->n;
  p.nb_rows = (int32_t)L.nbq;
  p.nb_cols = ref->nb;
  p.n_k = n_k;
  p.s64 = ref->s64;
  for (int k = 0; k < n_k; ++k) p.kvals[k] = (float)ref->klist[k];
  p.tolerance = s->tolerance;
  p.random_mode = s->random_mode;
  p.jv_rows = s->random_mode == PPSK_RANDOM_EQUAL_BASE ? jv : nullptr;
  p.jv_cols = ref->jv;
  // Monte Carlo: table[k][cluster of the reference][cluster of the query],
  // the query's cluster being its nearest centroid (random_match(r1 = ref,
  // q_cluster, ...), src/random/random_match.cpp:185-197, 230-237)
  p.n_clusters = s->n_clusters;
  p.mc_table = s->d_table;
  p.cl_rows = s->random_mode == PPSK_RANDOM_MC ? qcl : nullptr;
  p.cl_cols = ref->clusters;
  p.mc_ref_is_col = 1;
  p.row_prefetch = row_prefetch();
  p.out_kind = PPSK_OUT_CORE_ACC;
  p.out = band;
  p.counts = scratch;
  p.tile_done = tile_done;
  p.fail_count = fail;
  p.fail_first = fail + 1;
  p.fail_rows = fail_rows;
  const int variant = step_variant(s, rows);
  const int tr = ppsk::variant_tile_rows(variant);
  const int tc = ppsk::variant_tile_cols(variant);
  std::vector<size_t> gs;
  const auto tiles = make_tiles(rows, ref->n, tr, tc, false, 0, rows, tile_group(tr), &gs);
  if (L.persistent) {
    // whole tiles from 8 queues, each a contiguous eighth of the tile list
    // less its tail
    std::vector<ppsk::BatchItem> items;
    items.reserve(tiles.size());
    ppsk::BatchArgs args;
    std::memset(&args, 0, sizeof(args));
    const size_t n = tiles.size() - (size_t)L.tail;
    for (int q = 0; q < ppsk::BATCH_QUEUES; ++q) {
      args.queue_begin[q] = (int64_t)items.size();
      for (size_t x = n * q / ppsk::BATCH_QUEUES; x < n * (q + 1) / ppsk::BATCH_QUEUES; ++x) {
        items.push_back({0, tiles[x].bi, tiles[x].ct, 0});
      }
    }
    args.queue_begin[ppsk::BATCH_QUEUES] = (int64_t)items.size();
    ppsk::DistParams* d_params = (ppsk::DistParams*)(mem + L.params);
    args.jobs = d_params;
    args.items = (const ppsk::BatchItem*)(mem + L.tiles);
    args.queue_next = tile_done;
    args.scratch = scratch;
    args.max_k = n_k;
    STEP_TRY(hipMemcpyAsync(mem + L.tiles, items.data(), items.size() * sizeof(ppsk::BatchItem),
                            hipMemcpyHostToDevice, w.st));
    STEP_TRY(hipMemcpyAsync(d_params, &p, sizeof(p), hipMemcpyHostToDevice, w.st));
    if (L.tail > 0) {
      STEP_TRY(hipMemcpyAsync(mem + L.tail_tiles, tiles.data() + n, (size_t)L.tail * sizeof(ppsk::Tile),
                              hipMemcpyHostToDevice, w.st));
      STEP_TRY(hipMemsetAsync(mem + L.tail_done, 0, (size_t)L.tail * sizeof(uint32_t), w.st));
    }
    STEP_TRY(ppsk::launch_batch_reset(tile_done, nullptr, 0, w.st));
    STEP_TRY(hipEventRecord(w.e0, w.st));
    STEP_TRY(ppsk::launch_batch(args, variant, L.grid, w.st));
    if (L.tail > 0) {
      // the last partial round as (tile, k) items, after the persistent
      // launch on the same stream
      ppsk::DistParams pt = p;
      pt.tiles = (const ppsk::Tile*)(mem + L.tail_tiles);
      pt.n_tiles = L.tail;
      pt.item_base = 0;
      pt.counts = (uint16_t*)(mem + L.tail_counts);
      pt.tile_done = (uint32_t*)(mem + L.tail_done);
      STEP_TRY(ppsk::launch_dists(pt, variant, L.tail * n_k, w.st));
    }
    STEP_TRY(hipEventRecord(w.e1, w.st));
  } else {
    STEP_TRY(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(ppsk::Tile),
                            hipMemcpyHostToDevice, w.st));
    STEP_TRY(hipEventRecord(w.e0, w.st));
    for (size_t a = 0; a < tiles.size(); a += (size_t)L.chunk_tiles) {
      const size_t b = std::min(tiles.size(), a + (size_t)L.chunk_tiles);
      p.tiles = d_tiles + a;
      p.n_tiles = (int64_t)(b - a);
      p.item_base = 0;
      STEP_TRY(ppsk::launch_dists(p, variant, (int64_t)(b - a) * n_k, w.st));
    }
    STEP_TRY(hipEventRecord(w.e1, w.st));
  }

  // kNN and threshold counts straight into page-locked host memory
  const size_t nk = (size_t)rows * s->kNN;
  if (s->kNN > 0) {
    if (!w.knn_i.reserve(nk * 8) || !w.knn_j.reserve(nk * 8) || !w.knn_d.reserve(nk * 4)) {
      err = "page-locked kNN buffers";
      return PPSK_ERR_MEMORY;
    }
    STEP_TRY(ppsk::launch_knn_select_rect(band, rows, ref->n, s->knn_col, s->kNN,
                                          (int64_t*)w.knn_i.p, (int64_t*)w.knn_j.p,
                                          (float*)w.knn_d.p, w.st, L.knn_segs,
                                          mem + L.knn_lists));
  }
  const int64_t nseg = rows * L.segs;  // (row, segment) counts, row-major
  if (!w.counts.reserve((size_t)(nseg + 1) * 8) || !w.fail_rows.reserve((size_t)rows * 4)) {
    err = "page-locked count buffers";
    return PPSK_ERR_MEMORY;
  }
  // threshold: per-row counts, offsets scanned on the device, hits written
  // straight to page-locked buffers of the capacity they already have, so the
  // whole step is queued at once and waited for once; only a step with more
  // hits than that capacity writes again into larger buffers
  int64_t* cnt = (int64_t*)w.counts.p;
  int64_t* dcnt = (int64_t*)(mem + L.counts);
  int64_t cap = 0;
  if (s->cutoff > 0) {
    const size_t want = std::max<size_t>((size_t)rows * 64, (size_t)1 << 16);
    if (w.thr_j.n / 8 < want &&
        (!w.thr_i.reserve(want * 8) || !w.thr_j.reserve(want * 8) || !w.thr_d.reserve(want * 4))) {
      err = "page-locked threshold buffers";
      return PPSK_ERR_MEMORY;
    }
    cap = (int64_t)std::min({w.thr_i.n / 8, w.thr_j.n / 8, w.thr_d.n / 4});
    STEP_TRY(ppsk::launch_threshold_count((const float2*)band, rows, ref->n, true, s->thr_col,
                                          s->cutoff, dcnt, w.st, L.segs));
    STEP_TRY(ppsk::launch_offsets_scan(dcnt, nseg, w.st));
    STEP_TRY(ppsk::launch_threshold_write((const float2*)band, rows, ref->n, true, s->thr_col,
                                          s->cutoff, dcnt, (int64_t*)w.thr_i.p,
                                          (int64_t*)w.thr_j.p, (float*)w.thr_d.p, w.st, cap,
                                          L.segs));
    STEP_TRY(hipMemcpyAsync(cnt, dcnt, (size_t)(nseg + 1) * 8, hipMemcpyDeviceToHost, w.st));
  }
  STEP_TRY(hipMemcpyAsync(w.fail_rows.p, fail_rows, (size_t)rows * 4, hipMemcpyDeviceToHost, w.st));
  STEP_TRY(hipStreamSynchronize(w.st));
  float ms = 0.f;
  STEP_TRY(hipEventElapsedTime(&ms, w.e0, w.e1));
  const int64_t total = s->cutoff > 0 ? cnt[nseg] : 0;
  if (total > cap) {
    if (!w.thr_i.reserve((size_t)total * 8) || !w.thr_j.reserve((size_t)total * 8) ||
        !w.thr_d.reserve((size_t)total * 4)) {
      err = "page-locked threshold buffers";
      return PPSK_ERR_MEMORY;
    }
    STEP_TRY(ppsk::launch_threshold_write((const float2*)band, rows, ref->n, true, s->thr_col,
                                          s->cutoff, dcnt, (int64_t*)w.thr_i.p,
                                          (int64_t*)w.thr_j.p, (float*)w.thr_d.p, w.st,
                                          INT64_MAX, L.segs));
    STEP_TRY(hipStreamSynchronize(w.st));
  }

  // each piece's slice of the step, re-indexed to its request's query
  // numbering; kNN rows go straight to the request's lists (pieces of one
  // request never overlap), threshold hits and failures stay with the piece
  const uint32_t* fr = (const uint32_t*)w.fail_rows.p;
  o = 0;
  for (Piece& pc : batch) {
    ppsk_request* r = pc.r;
    const int64_t s0 = o, s1 = o + pc.n, shift = pc.q0 - s0;
    if (s->kNN > 0) {
      const int64_t* si = (const int64_t*)w.knn_i.p;
      const int64_t* sj = (const int64_t*)w.knn_j.p;
      const float* sd = (const float*)w.knn_d.p;
      const size_t a = (size_t)s0 * s->kNN, b = (size_t)s1 * s->kNN;
      const size_t at = (size_t)pc.q0 * s->kNN;
      for (size_t x = a; x < b; ++x) r->ki[at + x - a] = si[x] + shift;
      std::copy(sj + a, sj + b, r->kj.begin() + at);
      std::copy(sd + a, sd + b, r->kd.begin() + at);
    }
    if (s->cutoff > 0 && total > 0) {
      const int64_t a = cnt[s0 * L.segs], b = cnt[s1 * L.segs];
      const int64_t* si = (const int64_t*)w.thr_i.p;
      pc.thr.i.resize(b - a);
      for (int64_t x = a; x < b; ++x) pc.thr.i[x - a] = si[x] + shift;
      pc.thr.j.assign((const int64_t*)w.thr_j.p + a, (const int64_t*)w.thr_j.p + b);
      pc.thr.d.assign((const float*)w.thr_d.p + a, (const float*)w.thr_d.p + b);
    }
    uint64_t nf = 0;
    for (int64_t x = s0; x < s1; ++x) nf += fr[x];
    pc.nf = nf;
    o = s1;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->gpu_ms += ms;
  return PPSK_OK;
}

// Rows of the next piece of the oldest request in a step of its own.
int64_t head_piece_rows(const ppsk_server* s) {
  const ppsk_request* r = s->pending.front();
  return std::min<int64_t>(r->n_q - r->next_row, s->max_rows);
}

// All pieces of r have finished: threshold hits in row order, done.
void finish_request(ppsk_request* r, double t) {
  size_t n = 0;
  for (auto& kv : r->thr_parts) n += kv.second.j.size();
  r->ti.reserve(n);
  r->tj.reserve(n);
  r->td.reserve(n);
  for (auto& kv : r->thr_parts) {
    r->ti.insert(r->ti.end(), kv.second.i.begin(), kv.second.i.end());
    r->tj.insert(r->tj.end(), kv.second.j.begin(), kv.second.j.end());
    r->td.insert(r->td.end(), kv.second.d.begin(), kv.second.d.end());
  }
  r->thr_parts.clear();
  r->res.t_done = t;
  r->done = true;
}

// One wave that sleeps `n` times 127 x 64 clock cycles (about 2 ms for
// n = 600 at 2.4 GHz).
__global__ void queue_probe_kernel(int n) {
  for (int x = 0; x < n; ++x) __builtin_amdgcn_s_sleep(127);
}

// Whether work on streams a and b runs at the same time: a wait on each, both
// queued at once, ends about one wait after the first starts when the streams
// feed different hardware queues and about two waits after it when they share
// one (the runtime maps streams onto a few queues, so two streams of one
// process may be served one after the other).
int streams_concurrent(hipStream_t a, hipStream_t b, int ticks, bool* concurrent) {
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  int rc = PPSK_OK;
  for (auto& x : e) {
    if (hipEventCreate(&x) != hipSuccess) rc = PPSK_ERR_HIP;
  }
  float ta = 0.f, tb = 0.f;
  if (rc == PPSK_OK) {
    queue_probe_kernel<<<1, 64, 0, a>>>(0);  // first launch on the stream
    if (hipStreamSynchronize(a) != hipSuccess || hipEventRecord(e[0], a) != hipSuccess) {
      rc = PPSK_ERR_HIP;
    }
  }
  if (rc == PPSK_OK) {
    // a's end event before b's wait: on a shared queue, anything queued after
    // b's wait would also wait for it
    queue_probe_kernel<<<1, 64, 0, a>>>(ticks);
    if (hipEventRecord(e[1], a) != hipSuccess) rc = PPSK_ERR_HIP;
    queue_probe_kernel<<<1, 64, 0, b>>>(ticks);
    if (rc != PPSK_OK || hipGetLastError() != hipSuccess ||
        hipEventRecord(e[2], b) != hipSuccess || hipStreamSynchronize(a) != hipSuccess ||
        hipStreamSynchronize(b) != hipSuccess || hipEventElapsedTime(&ta, e[0], e[1]) != hipSuccess ||
        hipEventElapsedTime(&tb, e[0], e[2]) != hipSuccess) {
      rc = PPSK_ERR_HIP;
    }
  }
  for (auto& x : e) {
    if (x) hipEventDestroy(x);
  }
  // ta: one wait; tb: one wait when concurrent, two when serialised
  *concurrent = rc == PPSK_OK && tb < 1.5f * ta;
  return rc;
}

// n streams for the serving slots, the first few placed on distinct hardware
// queues where the runtime allows: a new stream that runs serialised with an
// earlier one is kept aside (so that its queue counts as busier) and another
// is created, at most three times per slot; the streams kept aside are then
// destroyed.  Slots beyond the first four share queues in any case.
int create_slot_streams(int n, std::vector<hipStream_t>* out) {
  const int ticks = 600;
  std::vector<hipStream_t> aside;
  int rc = PPSK_OK;
  for (int i = 0; i < n && rc == PPSK_OK; ++i) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      rc = PPSK_ERR_HIP;
      break;
    }
    for (int tries = 0; i > 0 && i < 4 && tries < 3; ++tries) {
      bool all = true;
      for (hipStream_t o : *out) {
        bool c = false;
        if ((rc = streams_concurrent(o, st, ticks, &c)) != PPSK_OK) break;
        all = all && c;
      }
      if (rc != PPSK_OK || all) break;
      aside.push_back(st);
      st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        rc = PPSK_ERR_HIP;
        break;
      }
    }
    if (st) out->push_back(st);
  }
  for (hipStream_t st : aside) hipStreamDestroy(st);
  return rc;
}

void worker_main(ppsk_server* s, Worker* w) {
  hipSetDevice(s->device);
  for (;;) {
    std::vector<Piece> batch;
    int64_t rows = 0;
    size_t off = 0;
    StepLayout L{};
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // admission: a free slot (this worker) and device memory for at least
      // the next piece of the oldest request
      s->cv_work.wait(lk, [&] {
        return s->stop ||
               (!s->pending.empty() && s->arena.fits(layout(s, head_piece_rows(s)).total));
      });
      if (s->stop) return;
      // continuous batching: oldest first, while rows and memory allow; a
      // request larger than the room left gives a piece and keeps its place
      const int64_t step_id = s->steps;
      while (!s->pending.empty() && rows < s->max_rows) {
        ppsk_request* r = s->pending.front();
        const int64_t take = std::min<int64_t>(r->n_q - r->next_row, s->max_rows - rows);
        if (!s->arena.fits(layout(s, rows + take).total)) break;
        if (r->next_row == 0) {
          r->res.t_start = now_s(s->t0);
          r->res.step_id = step_id;
        }
        batch.push_back({r, r->next_row, take});
        r->next_row += take;
        r->res.n_steps += 1;
        rows += take;
        if (r->next_row == r->n_q) s->pending.pop_front();
      }
      for (const Piece& pc : batch) {
        if (pc.q0 == 0) pc.r->res.step_rows = rows;
      }
      L = layout(s, rows);
      s->arena.alloc(L.total, &off);  // fits() held under the same lock
      s->steps += 1;
      s->inflight += 1;
      s->peak_inflight = std::max(s->peak_inflight, s->inflight);
    }
    std::string err;
    const int rc = run_step(s, *w, batch, rows, s->arena.base + off, L, err);
    // a failed step may have queued kernels that still write into its arena
    // block: the block is released only once the stream has drained
    if (rc != PPSK_OK) hipStreamSynchronize(w->st);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->arena.release(off, L.total);
      s->inflight -= 1;
      const double t = now_s(s->t0);
      for (Piece& pc : batch) {
        ppsk_request* r = pc.r;
        if (rc != PPSK_OK) {
          if (r->status == PPSK_OK) {
            r->status = rc;
            r->error = err;
          }
        } else {
          r->res.n_fail += pc.nf;
          if (!pc.thr.j.empty()) r->thr_parts[pc.q0] = std::move(pc.thr);
        }
        r->rows_done += pc.n;
        if (r->rows_done == r->n_q) {
          finish_request(r, t);
          s->requests_done += 1;
        }
      }
      s->rows_done += rows;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

int ppsk_server_create(ppsk_set* ref, int32_t random_mode, int32_t n_clusters,
                       const float* mc_table, double tolerance, int32_t kNN, int32_t knn_col,
                       float cutoff, int32_t thr_col, int32_t max_rows, int32_t n_slots,
                       size_t arena_bytes, ppsk_server** out) {
  if (!out) return fail(PPSK_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (int rc = check_set(ref, "ref")) return rc;
  if (ref->n_k < 2) return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths");
  if (random_mode != PPSK_RANDOM_NONE && random_mode != PPSK_RANDOM_EQUAL_BASE &&
      random_mode != PPSK_RANDOM_MC) {
    return fail(PPSK_ERR_ARG, "unknown random_mode");
  }
  if (random_mode == PPSK_RANDOM_EQUAL_BASE && !ref->has_lengths) {
    return fail(PPSK_ERR_ARG, "equal-base random correction needs sample lengths (ppsk_set_meta)");
  }
  if (random_mode == PPSK_RANDOM_MC) {
    if (!ref->clusters) {
      return fail(PPSK_ERR_ARG, "Monte Carlo random correction needs cluster ids (ppsk_set_meta)");
    }
    if (n_clusters < 1 || n_clusters > 65536 || !mc_table) {
      return fail(PPSK_ERR_ARG, "Monte Carlo random correction needs a table");
    }
    if (ref->max_cluster >= n_clusters) {
      return fail(PPSK_ERR_ARG, "cluster id outside the Monte Carlo table");
    }
  }
  if (kNN < 0 || kNN > ref->n) return fail(PPSK_ERR_ARG, "kNN out of range");
  if ((knn_col != 0 && knn_col != 1) || (thr_col != 0 && thr_col != 1)) {
    return fail(PPSK_ERR_ARG, "dist_col out of range");
  }
  if (kNN == 0 && !(cutoff > 0)) return fail(PPSK_ERR_ARG, "nothing to return: kNN = 0 and cutoff <= 0");
  if (max_rows < 1 || n_slots < 1 || n_slots > 16) {
    return fail(PPSK_ERR_ARG, "max_rows >= 1 and n_slots in [1, 16] required");
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(ref->device, &c)) return rc;
  if (hipSetDevice(ref->device) != hipSuccess) return fail(PPSK_ERR_HIP, "hipSetDevice failed");
  auto* s = new ppsk_server();
  s->ref = ref;
  s->device = ref->device;
  s->random_mode = random_mode;
  s->n_clusters = random_mode == PPSK_RANDOM_MC ? n_clusters : 0;
  s->tolerance = tolerance;
  s->kNN = kNN;
  s->knn_col = knn_col;
  s->cutoff = cutoff;
  s->thr_col = thr_col;
  s->max_rows = max_rows;
  s->variant = std::getenv("PPSK_TILE_VARIANT") ? default_variant() : 0;
  s->per_sample = (size_t)ref->n_k * ref->s64 * ppsk::BBITS * sizeof(uint64_t);
  if (ppsk::batch_grid(2, &s->grid2) != hipSuccess) s->grid2 = 0;
  if (ppsk::batch_grid(4, &s->grid4) != hipSuccess) s->grid4 = 0;
  if (const char* e = std::getenv("PPSK_PERSISTENT")) {
    if (std::atoi(e) == 0) s->grid2 = s->grid4 = 0;
    s->force_persistent = std::atoi(e) == 1;
  }
  s->scratch_budget = (size_t)256 << 20;
  if (arena_bytes == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    arena_bytes = fr > ((size_t)4 << 30) ? fr - ((size_t)4 << 30) : fr / 2;
  }
  const size_t one = layout(s, 1).total;
  if (arena_bytes < one) {
    delete s;
    return fail(PPSK_ERR_MEMORY, "arena smaller than one single-query step (" +
                                     std::to_string(one) + " bytes)");
  }
  if (random_mode == PPSK_RANDOM_MC) {
    const size_t tb = (size_t)ref->n_k * n_clusters * n_clusters * sizeof(float);
    hipError_t et = hipMalloc(&s->d_table, tb);
    if (et == hipSuccess) et = hipMemcpy(s->d_table, mc_table, tb, hipMemcpyHostToDevice);
    if (et != hipSuccess) {
      if (s->d_table) hipFree(s->d_table);
      delete s;
      return fail(PPSK_ERR_MEMORY, std::string("serving Monte Carlo table: ") + hipGetErrorString(et));
    }
  }
  hipError_t e = hipMalloc(&s->arena_mem, arena_bytes);
  if (e != hipSuccess) {
    if (s->d_table) hipFree(s->d_table);
    delete s;
    return fail(PPSK_ERR_MEMORY, std::string("serving arena: ") + hipGetErrorString(e));
  }
  s->arena.init((char*)s->arena_mem, arena_bytes);
  s->t0 = std::chrono::steady_clock::now();
  std::vector<hipStream_t> streams;
  if (create_slot_streams(n_slots, &streams) != PPSK_OK) {
    for (hipStream_t st : streams) hipStreamDestroy(st);
    ppsk_server_free(s);
    return fail(PPSK_ERR_HIP, "serving stream creation");
  }
  for (int i = 0; i < n_slots; ++i) {
    auto* w = new Worker();
    w->st = streams[i];
    if (hipEventCreate(&w->e0) != hipSuccess || hipEventCreate(&w->e1) != hipSuccess) {
      for (int x = i; x < n_slots; ++x) hipStreamDestroy(streams[x]);
      if (w->e0) hipEventDestroy(w->e0);
      delete w;
      ppsk_server_free(s);
      return fail(PPSK_ERR_HIP, "serving stream creation");
    }
    s->workers.emplace_back([s, w] {
      worker_main(s, w);
      hipStreamDestroy(w->st);
      hipEventDestroy(w->e0);
      hipEventDestroy(w->e1);
      delete w;
    });
  }
  *out = s;
  return PPSK_OK;
}

int ppsk_server_submit(ppsk_server* s, int64_t n_queries, const uint64_t* raw,
                       const uint64_t* lengths, const uint16_t* clusters, int64_t* ticket) {
  if (!s || !ticket) return fail(PPSK_ERR_ARG, "server / ticket is NULL");
  if (n_queries < 1 || !raw) return fail(PPSK_ERR_ARG, "a request needs at least one query sketch");
  // a request larger than max_rows is served in pieces of at most max_rows
  // rows, so only a piece has to fit the arena
  if (layout(s, std::min<int64_t>(n_queries, s->max_rows)).total > s->arena.size) {
    return fail(PPSK_ERR_MEMORY, "request needs more device memory than the serving arena");
  }
  const ppsk_set* ref = s->ref;
  auto* r = new ppsk_request();
  r->n_q = n_queries;
  r->raw.assign(raw, raw + (size_t)n_queries * (s->per_sample / sizeof(uint64_t)));
  if (s->random_mode == PPSK_RANDOM_EQUAL_BASE) {
    if (!lengths) {
      delete r;
      return fail(PPSK_ERR_ARG, "equal-base correction needs query lengths");
    }
    // csrs terms per query and k (src/random/random_match.cpp:318-336)
    r->jv.resize((size_t)n_queries * ref->n_k);
    for (int kk = 0; kk < ref->n_k; ++kk) {
      const int64_t shift = (int64_t)(ref->klist[kk] - 1) * 2 + (ref->use_rc ? 1 : 0);
      const double base = 1 - (double)1 / (double)((size_t)1 << shift);
      for (int64_t q = 0; q < n_queries; ++q) {
        r->jv[(size_t)q * ref->n_k + kk] = 1 - std::pow(base, (double)lengths[q]);
      }
    }
  }
  if (s->random_mode == PPSK_RANDOM_MC) {
    if (!clusters) {
      delete r;
      return fail(PPSK_ERR_ARG, "Monte Carlo correction needs query clusters");
    }
    for (int64_t q = 0; q < n_queries; ++q) {
      if (clusters[q] >= s->n_clusters) {
        delete r;
        return fail(PPSK_ERR_ARG, "cluster id outside the Monte Carlo table");
      }
    }
    r->cl.assign(clusters, clusters + n_queries);
  }
  if (s->kNN > 0) {
    const size_t nk = (size_t)n_queries * s->kNN;
    r->ki.resize(nk);
    r->kj.resize(nk);
    r->kd.resize(nk);
  }
  std::memset(&r->res, 0, sizeof(r->res));
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete r;
      return fail(PPSK_ERR_STATE, "server stopped");
    }
    r->ticket = s->next_ticket++;
    r->res.n_queries = n_queries;
    r->res.t_submit = now_s(s->t0);
    s->tickets[r->ticket] = r;
    s->pending.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_work.notify_one();
  return PPSK_OK;
}

int ppsk_server_wait(ppsk_server* s, int64_t ticket, ppsk_result** out) {
  if (!s || !out) return fail(PPSK_ERR_ARG, "server / out is NULL");
  *out = nullptr;
  ppsk_request* r = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(PPSK_ERR_ARG, "unknown or already collected ticket");
    r = it->second;
    s->waiters += 1;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->tickets.erase(ticket);
    s->waiters -= 1;
    if (s->waiters == 0) s->cv_done.notify_all();  // ppsk_server_free may wait for this
  }
  if (r->status != PPSK_OK) {
    const int rc = r->status;
    const std::string msg = r->error;
    delete r;
    return fail(rc, "serving step failed: " + msg);
  }
  ppsk_result& res = r->res;
  res.knn_n = (int64_t)r->kj.size();
  res.knn_i = r->ki.data();
  res.knn_j = r->kj.data();
  res.knn_d = r->kd.data();
  res.thr_n = (int64_t)r->tj.size();
  res.thr_i = r->ti.data();
  res.thr_j = r->tj.data();
  res.thr_d = r->td.data();
  res.owner = r;
  *out = &res;
  return PPSK_OK;
}

void ppsk_result_free(ppsk_result* res) {
  if (res) delete (ppsk_request*)res->owner;
}

int ppsk_server_stats(ppsk_server* s, int64_t* steps, int64_t* rows, int64_t* requests,
                      int32_t* peak_inflight, size_t* peak_arena_bytes, double* gpu_ms) {
  if (!s) return fail(PPSK_ERR_ARG, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (steps) *steps = s->steps;
  if (rows) *rows = s->rows_done;
  if (requests) *requests = s->requests_done;
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  if (peak_arena_bytes) *peak_arena_bytes = s->arena.peak;
  if (gpu_ms) *gpu_ms = s->gpu_ms;
  return PPSK_OK;
}

int ppsk_server_free(ppsk_server* s) {
  if (!s) return PPSK_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // every request not yet served fails; waiters collect their failure and
    // leave before any request or the server is deleted
    for (auto& kv : s->tickets) {
      ppsk_request* r = kv.second;
      if (!r->done) {
        r->status = PPSK_ERR_STATE;
        r->error = "server stopped";
        r->done = true;
      }
    }
    s->pending.clear();
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
    for (auto& kv : s->tickets) delete kv.second;
    s->tickets.clear();
  }
  hipSetDevice(s->device);
  if (s->arena_mem) hipFree(s->arena_mem);
  if (s->d_table) hipFree(s->d_table);
  delete s;
  return PPSK_OK;
}

}  // extern "C"
