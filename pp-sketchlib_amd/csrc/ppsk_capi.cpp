// C-ABI implementation (include/ppsk.h): resident sketch sets, device
// workspaces, tile scheduling and the dense / kNN / threshold entry points.
// Compiled by hipcc together with dist_kernels.hip and select_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ppsk.h"
#include "ppsk_internal.h"
#include "select_kernels.h"
#include "capi_common.h"

namespace ppsk {
namespace capi {

thread_local std::string g_last_error;
thread_local double g_last_ms = 0.0;
thread_local int64_t g_last_launches = 0;
thread_local int64_t g_last_items = 0;
std::mutex g_ctx_mu;
std::vector<DeviceCtx*> g_ctx;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}



int get_ctx(int device, DeviceCtx** out) {
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n || device >= MAX_DEVICES) {
    return fail(PPSK_ERR_ARG, "device_id " + std::to_string(device) + " out of range");
  }
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if ((int)g_ctx.size() < n) g_ctx.resize(n, nullptr);
  if (!g_ctx[device]) {
    HIP_TRY(hipSetDevice(device));
    auto* c = new DeviceCtx();
    c->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&c->fail, 2 * sizeof(unsigned long long)));
    if (int rc = noalloc_init(&c->noalloc_scratch)) return rc;
    HIP_TRY(ppsk::upload_log_table());
    // progress / cancellation words: a device counter and two host-mapped,
    // coherent words (device_id < MAX_DEVICES here)
    HIP_TRY(hipMalloc(&c->ctl.progress, sizeof(unsigned long long)));
    HIP_TRY(hipMemset(c->ctl.progress, 0, sizeof(unsigned long long)));
    void* hw = nullptr;
    HIP_TRY(hipHostMalloc(&hw, 64, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(hw, 0, 64);
    void* dw = nullptr;
    HIP_TRY(hipHostGetDevicePointer(&dw, hw, 0));
    c->ctl.cancel = (const uint32_t*)dw;
    c->ctl.progress_host = (unsigned long long*)((char*)dw + 8);
    CallState& cs = call_state(device);
    cs.progress_host.store((volatile unsigned long long*)((char*)hw + 8));
    cs.cancel_host.store((volatile uint32_t*)hw);
    *(volatile uint32_t*)hw = cs.requested.load();  // a request made before the context existed
    g_ctx[device] = c;
  }
  *out = g_ctx[device];
  HIP_TRY(hipSetDevice(device));
  return PPSK_OK;
}

CallState g_calls[MAX_DEVICES];

CallState& call_state(int device) { return g_calls[device]; }

bool cancel_requested(int device) {
  return device >= 0 && device < MAX_DEVICES && g_calls[device].requested.load() != 0;
}

int cancel_status(int device) {
  if (!cancel_requested(device)) return PPSK_OK;
  return fail(PPSK_ERR_CANCELLED, "cancelled by ppsk_cancel");
}

// Largest progress value read so far (stores of different workgroups may
// arrive out of order).
uint64_t read_progress(CallState& cs) {
  uint64_t v = 0;
  if (!cs.running.load()) return cs.seen.load();  // final count of the last call
  if (volatile unsigned long long* p = cs.progress_host.load()) v = *p;
  uint64_t seen = cs.seen.load();
  while (v > seen && !cs.seen.compare_exchange_weak(seen, v)) {
  }
  return std::max(v, seen);
}

CallScope::CallScope(DeviceCtx* ctx, uint64_t total) : c(ctx), cs(g_calls[ctx->device]) {
  cs.total.store(total);
  cs.seen.store(0);
  if (volatile unsigned long long* p = cs.progress_host.load()) *p = 0;
  cs.running.store(1);
  // the device counter starts from zero in stream order
  hipMemsetAsync(c->ctl.progress, 0, sizeof(unsigned long long), c->stream);
}

CallScope::~CallScope() {
  unsigned long long v = 0;
  if (hipMemcpyAsync(&v, c->ctl.progress, sizeof(v), hipMemcpyDeviceToHost, c->stream) ==
          hipSuccess &&
      hipStreamSynchronize(c->stream) == hipSuccess) {
    cs.seen.store(v);
  } else {
    read_progress(cs);
  }
  cs.running.store(0);
}

uint64_t pass_items(int64_t n_rows, int64_t n_cols, int n_k, int variant, bool upper) {
  const int tr = ppsk::variant_tile_rows(variant), tc = ppsk::variant_tile_cols(variant);
  const int64_t nbr = (n_rows + tr - 1) / tr, nct = (n_cols + tc - 1) / tc;
  uint64_t tiles = 0;
  for (int64_t ct = 0; ct < nct; ++ct) {
    if (!upper) {
      tiles += (uint64_t)nbr;
      continue;
    }
    // make_tiles keeps (bi, ct) when bi * tr < (ct + 1) * tc - 1
    const int64_t lim = (ct + 1) * tc - 1;
    tiles += (uint64_t)std::min<int64_t>(nbr, (lim + tr - 1) / tr);
  }
  return tiles * (uint64_t)n_k;
}

DeviceCtx* find_ctx(int device) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (device < 0 || device >= (int)g_ctx.size() || !g_ctx[device]) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  return g_ctx[device];
}

int grow(void** ptr, size_t* have, size_t need) {
  if (*have >= need && *ptr) return PPSK_OK;
  if (*ptr) HIP_TRY(hipFree(*ptr));
  *ptr = nullptr;
  *have = 0;
  HIP_TRY(hipMalloc(ptr, need));
  *have = need;
  return PPSK_OK;
}

int grow_tile_done(DeviceCtx* c, size_t n) {
  if (c->tile_done_n >= n && c->tile_done) return PPSK_OK;
  if (c->tile_done) HIP_TRY(hipFree(c->tile_done));
  c->tile_done = nullptr;
  c->tile_done_n = 0;
  HIP_TRY(hipMalloc(&c->tile_done, n * sizeof(uint32_t)));
  // counters are left at zero by every tile's last workgroup afterwards
  HIP_TRY(hipMemset(c->tile_done, 0, n * sizeof(uint32_t)));
  c->tile_done_n = n;
  return PPSK_OK;
}

}  // namespace capi
}  // namespace ppsk


namespace ppsk {
namespace capi {

int check_set(const ppsk_set* s, const char* what) {
  if (!s) return fail(PPSK_ERR_ARG, std::string(what) + " is NULL");
  return PPSK_OK;
}

int check_compatible(const ppsk_set* a, const ppsk_set* b) {
  if (a->device != b->device) return fail(PPSK_ERR_ARG, "sketch sets on different devices");
  if (a->bbits != b->bbits) return fail(PPSK_ERR_ARG, "Mismatching bbits in sketches");
  if (a->s64 != b->s64) return fail(PPSK_ERR_ARG, "Mismatching sketchsize64 in sketches");
  if (a->klist != b->klist) return fail(PPSK_ERR_ARG, "Mismatching k-mer lengths in sketches");
  return PPSK_OK;
}

// Tile list for a tile shape of tr rows x tc columns over row tiles
// [row_begin / tr, ceil(row_end / tr)) (row_begin a multiple of tr).  Self
// mode keeps upper-triangle tiles only (at least one pair with i < j).
//
// Order: row tiles are taken in groups of `group`; inside a group the tiles
// run column by column, each column visiting every row tile of the group.
// Consecutive workgroups of a launch land on the same XCD eight apart, so the
// workgroups an XCD holds at once cover a compact block of `group` row tiles
// by a few column tiles: each column panel and row panel is fetched into that
// XCD's L2 once and reused by the whole block.  group_start receives the
// first tile of every group plus the end of the list.
std::vector<ppsk::Tile> make_tiles(int64_t n_rows, int64_t n_cols, int tr, int tc, bool upper,
                                   int64_t row_begin, int64_t row_end, int group,
                                   std::vector<size_t>* group_start) {
  std::vector<ppsk::Tile> t;
  const int64_t nbr = (n_rows + tr - 1) / tr;
  const int64_t nct = (n_cols + tc - 1) / tc;
  const int64_t b0 = row_begin / tr;
  const int64_t b1 = std::min(nbr, (row_end + tr - 1) / tr);
  group_start->clear();
  for (int64_t g0 = b0; g0 < b1; g0 += group) {
    const int64_t g1 = std::min(b1, g0 + group);
    const size_t before = t.size();
    for (int64_t ct = 0; ct < nct; ++ct) {
      for (int64_t bi = g0; bi < g1; ++bi) {
        if (upper && (ct + 1) * tc - 1 <= bi * tr) continue;
        t.push_back({(int32_t)bi, (int32_t)ct});
      }
    }
    if (t.size() > before) group_start->push_back(before);
  }
  group_start->push_back(t.size());
  return t;
}


int check_random(const ppsk_set* rows, const ppsk_set* cols, const RandomIn& rnd) {
  if (rnd.mode == PPSK_RANDOM_EQUAL_BASE) {
    if (!rows->has_lengths || !cols->has_lengths) {
      return fail(PPSK_ERR_ARG, "equal-base random correction needs sample lengths (ppsk_set_meta)");
    }
  } else if (rnd.mode == PPSK_RANDOM_MC) {
    if (!rows->clusters || !cols->clusters) {
      return fail(PPSK_ERR_ARG, "Monte Carlo random correction needs cluster ids (ppsk_set_meta)");
    }
    if (rnd.n_clusters < 1 || !rnd.table) {
      return fail(PPSK_ERR_ARG, "Monte Carlo random correction needs a table");
    }
    if (rows->max_cluster >= rnd.n_clusters || cols->max_cluster >= rnd.n_clusters) {
      return fail(PPSK_ERR_ARG, "cluster id outside the Monte Carlo table");
    }
  } else if (rnd.mode != PPSK_RANDOM_NONE) {
    return fail(PPSK_ERR_ARG, "unknown random_mode");
  }
  return PPSK_OK;
}

// Row tiles per scheduling group (make_tiles); PPSK_TILE_GROUP overrides
// (read on every call, so a process can compare settings).
int tile_group(int tile_rows) {
  const char* e = std::getenv("PPSK_TILE_GROUP");
  const int g = e ? std::atoi(e) : 512 / (tile_rows > 0 ? tile_rows : 64);
  return g < 1 ? 1 : g;
}

// Tile kernel configuration: PPSK_TILE_VARIANT overrides the default.
int default_variant() {
  const char* e = std::getenv("PPSK_TILE_VARIANT");
  const int v = e ? std::atoi(e) : ppsk::DEFAULT_VARIANT;
  return ppsk::variant_valid(v) ? v : ppsk::DEFAULT_VARIANT;
}


hipEvent_t pool_event(DeviceCtx* c, size_t i) {
  while (c->events.size() <= i) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->events.push_back(e);
  }
  return c->events[i];
}

// First output row of sample row r: self mode -> condensed index of (r, r+1),
// query mode -> r * n_cols.
int64_t out_row_start(bool self, int64_t r, int64_t n_rows, int64_t n_cols) {
  if (!self) return std::min(r, n_rows) * n_cols;
  const int64_t n = n_rows;
  if (r >= n - 1) return n * (n - 1) / 2;
  return n * r - ((r * (r + 1)) >> 1);
}

// Runs the tile kernel over the row tiles covering sample rows
// [row_begin, row_end) (row_begin a multiple of the tile height) against all
// columns, writing dense output into the device buffer p.out (row offset 0).
// Launches are cut at group boundaries (whole row tiles) so that each fits
// the per-tile count scratch; when host_out is set, the finished output rows
// of each launch are copied to host_out on the copy stream while the
// following launches compute.  Returns after both streams have drained.
int run_dense_tiles(DeviceCtx* c, ppsk::DistParams p, int variant, bool upper, int64_t row_begin,
                    int64_t row_end, bool self, size_t elem, void* host_out, Timing* tm) {
  const int tr = ppsk::variant_tile_rows(variant);
  const int tc = ppsk::variant_tile_cols(variant);
  const size_t per_tile = (size_t)p.n_k * tr * tc * sizeof(uint16_t);
  // count scratch per launch; PPSK_CHUNK_BYTES lowers it (more, smaller
  // launches: finer-grained streaming of host output)
  size_t budget = (size_t)1 << 30;
  if (const char* env = std::getenv("PPSK_CHUNK_BYTES")) {
    const long long v = std::atoll(env);
    if (v > 0) budget = (size_t)v;
  }
  const size_t max_tiles = std::max<size_t>(1, budget / per_tile);
  const int64_t nct = (p.n_cols + tc - 1) / tc;
  int group = tile_group(tr);
  group = (int)std::max<int64_t>(1, std::min<int64_t>(group, (int64_t)max_tiles / nct));
  std::vector<size_t> gs;
  const std::vector<ppsk::Tile> tiles =
      make_tiles(p.n_rows, p.n_cols, tr, tc, upper, row_begin, row_end, group, &gs);
  if (tiles.empty()) return PPSK_OK;
  // launches: runs of whole groups within the budget (at least one group)
  std::vector<std::pair<size_t, size_t>> chunks;
  for (size_t g = 0; g + 1 < gs.size();) {
    size_t h = g + 1;
    while (h + 1 < gs.size() && gs[h + 1] - gs[g] <= max_tiles) ++h;
    chunks.emplace_back(gs[g], gs[h]);
    g = h;
  }
  size_t most = 0;
  for (auto& ch : chunks) most = std::max(most, ch.second - ch.first);
  if (int rc = grow(&c->counts, &c->counts_bytes, most * per_tile)) return rc;
  if (int rc = grow_tile_done(c, most)) return rc;
  void* tptr = c->tiles;
  size_t tbytes = c->tiles_n * sizeof(ppsk::Tile);
  if (int rc = grow(&tptr, &tbytes, tiles.size() * sizeof(ppsk::Tile))) return rc;
  c->tiles = (ppsk::Tile*)tptr;
  c->tiles_n = tbytes / sizeof(ppsk::Tile);
  HIP_TRY(hipMemcpyAsync(c->tiles, tiles.data(), tiles.size() * sizeof(ppsk::Tile),
                         hipMemcpyHostToDevice, c->stream));
  p.counts = (uint16_t*)c->counts;
  p.tile_done = c->tile_done;
  p.out_row_offset = 0;
  char* dev = (char*)p.out;
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    const size_t a = chunks[ci].first, b = chunks[ci].second;
    p.tiles = c->tiles + a;
    p.n_tiles = (int64_t)(b - a);
    p.item_base = 0;
    const int64_t n_items = (int64_t)(b - a) * p.n_k;
    hipEvent_t e0 = pool_event(c, 2 * ci), e1 = pool_event(c, 2 * ci + 1);
    if (!e0 || !e1) return fail(PPSK_ERR_HIP, "event creation failed");
    HIP_TRY(hipEventRecord(e0, c->stream));
    HIP_TRY(ppsk::launch_dists(p, variant, n_items, c->stream));
    HIP_TRY(hipEventRecord(e1, c->stream));
    tm->launches += 1;
    tm->items += n_items;
  }
  // Copies are queued after every launch: a copy into pageable memory may
  // block this thread until it completes, which must not hold back launches.
  for (size_t ci = 0; host_out && ci < chunks.size(); ++ci) {
    const size_t a = chunks[ci].first, b = chunks[ci].second;
    int32_t bmin = tiles[a].bi, bmax = tiles[a].bi;
    for (size_t x = a; x < b; ++x) {
      bmin = std::min(bmin, tiles[x].bi);
      bmax = std::max(bmax, tiles[x].bi);
    }
    const int64_t r0 = (int64_t)bmin * tr;
    const int64_t r1 = (int64_t)(bmax + 1) * tr;
    const int64_t o0 = out_row_start(self, r0, p.n_rows, p.n_cols);
    const int64_t o1 = out_row_start(self, r1, p.n_rows, p.n_cols);
    if (o1 > o0) {
      HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->events[2 * ci + 1], 0));
      HIP_TRY(hipMemcpyAsync((char*)host_out + o0 * elem, dev + o0 * elem, (o1 - o0) * elem,
                             hipMemcpyDeviceToHost, c->copy_stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamSynchronize(c->copy_stream));
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->events[2 * ci], c->events[2 * ci + 1]));
    tm->ms += ms;
  }
  return PPSK_OK;
}

// Copies the finished dense rows of tiles [a, b) to the host on the copy
// stream once `done` has been reached on the launch stream.
int copy_rows(DeviceCtx* c, const std::vector<ppsk::Tile>& tiles, size_t a, size_t b, int tr,
              const ppsk::DistParams& p, bool self, size_t elem, void* host_out, hipEvent_t done) {
  int32_t bmin = tiles[a].bi, bmax = tiles[a].bi;
  for (size_t x = a; x < b; ++x) {
    bmin = std::min(bmin, tiles[x].bi);
    bmax = std::max(bmax, tiles[x].bi);
  }
  const int64_t o0 = out_row_start(self, (int64_t)bmin * tr, p.n_rows, p.n_cols);
  const int64_t o1 = out_row_start(self, (int64_t)(bmax + 1) * tr, p.n_rows, p.n_cols);
  if (o1 > o0) {
    const char* dev = (const char*)p.out;
    HIP_TRY(hipStreamWaitEvent(c->copy_stream, done, 0));
    HIP_TRY(hipMemcpyAsync((char*)host_out + o0 * elem, dev + o0 * elem, (o1 - o0) * elem,
                           hipMemcpyDeviceToHost, c->copy_stream));
  }
  return PPSK_OK;
}

// The same computation as run_dense_tiles with the persistent kernel: whole
// tiles (every k-mer length, then the epilogue) are work items of one launch
// over every CU, counts kept in a per-workgroup block.  For host output the
// rows are cut into about eight launches (whole groups of row tiles) so that
// each launch's rows are copied out while the next computes.
int run_dense_persistent(DeviceCtx* c, ppsk::DistParams p, int variant, bool upper,
                         int64_t row_begin, int64_t row_end, bool self, size_t elem,
                         void* host_out, Timing* tm) {
  const int tr = ppsk::variant_tile_rows(variant);
  const int tc = ppsk::variant_tile_cols(variant);
  // PPSK_CHUNK_BYTES (host output): one row tile per group and per launch,
  // the finest streaming granularity
  const bool finest = host_out && std::getenv("PPSK_CHUNK_BYTES");
  std::vector<size_t> gs;
  std::vector<ppsk::Tile> tiles = make_tiles(p.n_rows, p.n_cols, tr, tc, upper, row_begin,
                                             row_end, finest ? 1 : tile_group(tr), &gs);
  if (tiles.empty()) return PPSK_OK;
  // 16-row tiles run on the mixed-shape kernel (every item in that shape)
  const int kernel = variant == 5 ? ppsk::BATCH_MIXED : variant;
  int grid = 0;
  HIP_TRY(ppsk::batch_grid(kernel, &grid));
  // launches: runs of whole groups of about tiles / 8 (host output) or one
  std::vector<std::pair<size_t, size_t>> chunks;
  size_t target = host_out ? std::max<size_t>((size_t)grid * 4, tiles.size() / 8) : tiles.size();
  if (finest) target = 1;
  for (size_t g = 0; g + 1 < gs.size();) {
    size_t h = g + 1;
    while (h + 1 < gs.size() && gs[h] - gs[g] < target) ++h;
    chunks.emplace_back(gs[g], gs[h]);
    g = h;
  }
  // A chunk's last, partial round of whole tiles would leave most workgroup
  // slots idle while a few run a whole tile each: when it fills at most half
  // of the slots, those tiles go to a second launch as (tile, k-mer length)
  // items of the per-(tile, k) kernel instead (PPSK_TAIL_SPLIT=0 turns this
  // off).  Same tile code, same epilogue, same output.
  //
  // In a self pass (upper triangle) of 64 x 256 tiles, a diagonal tile whose
  // first row lies at least 128 columns past its first column has its first
  // 2 or 3 column blocks below the diagonal: each chunk's such tiles are
  // moved to its end, into the tail, whose kernel skips those blocks
  // (dist_tiles_diagonal); at 8,192 samples they are exactly the tail.  The
  // tiles that keep 2 blocks come before those that keep 1, so that the
  // items dispatched last, into the slots freed last, are the shortest.
  const bool diagonal = variant == 2 && self && upper && p.self_mode && !p.full_square;
  if (diagonal) {
    auto kept = [&](const ppsk::Tile& t) {  // 0: whole tile; 1: 2 blocks; 2: 1 block
      const int64_t d = (int64_t)t.bi * tr - (int64_t)t.ct * tc;
      return d >= 3 * 64 ? 2 : (d >= 2 * 64 ? 1 : 0);
    };
    for (const auto& ch : chunks) {
      std::stable_sort(tiles.begin() + ch.first, tiles.begin() + ch.second,
                       [&](const ppsk::Tile& a, const ppsk::Tile& b) { return kept(a) < kept(b); });
    }
  }
  std::vector<size_t> tail(chunks.size(), 0);
  std::vector<ppsk::Tile> tail_tiles;
  std::vector<size_t> tail_at(chunks.size(), 0);
  size_t most_tail = 0;
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    tail[ci] = tail_split_tiles(chunks[ci].second - chunks[ci].first, grid);
    tail_at[ci] = tail_tiles.size();
    tail_tiles.insert(tail_tiles.end(), tiles.begin() + (chunks[ci].second - tail[ci]),
                      tiles.begin() + chunks[ci].second);
    most_tail = std::max(most_tail, tail[ci]);
  }
  // items: chunk by chunk, each chunk's tiles cut into one contiguous run per
  // queue (so each XCD walks a compact block)
  std::vector<ppsk::BatchItem> items;
  items.reserve(tiles.size());
  std::vector<std::array<int64_t, ppsk::BATCH_QUEUES + 1>> qb(chunks.size());
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    const size_t a = chunks[ci].first, n = chunks[ci].second - a - tail[ci];
    for (int q = 0; q < ppsk::BATCH_QUEUES; ++q) {
      qb[ci][q] = (int64_t)items.size();
      for (size_t x = a + n * q / ppsk::BATCH_QUEUES; x < a + n * (q + 1) / ppsk::BATCH_QUEUES; ++x) {
        items.push_back({0, tiles[x].bi, tiles[x].ct, variant});
      }
    }
    qb[ci][ppsk::BATCH_QUEUES] = (int64_t)items.size();
  }
  // per-workgroup count blocks (the mixed-shape kernel's hold 64 x 256 counts)
  const size_t scratch = (size_t)grid * p.n_k * (kernel == ppsk::BATCH_MIXED ? 64 * 256 : tr * tc) *
                         sizeof(uint16_t);
  if (int rc = grow(&c->pscratch, &c->pscratch_bytes, scratch)) return rc;
  if (int rc = grow(&c->pitems, &c->pitems_bytes, items.size() * sizeof(ppsk::BatchItem))) return rc;
  if (!c->pparams) HIP_TRY(hipMalloc(&c->pparams, sizeof(ppsk::DistParams)));
  if (!c->pqueue) HIP_TRY(hipMalloc(&c->pqueue, ppsk::BATCH_QUEUES * sizeof(uint32_t)));
  p.out_row_offset = 0;
  if (most_tail > 0) {
    const size_t per_tile = (size_t)p.n_k * tr * tc * sizeof(uint16_t);
    if (int rc = grow(&c->counts, &c->counts_bytes, most_tail * per_tile)) return rc;
    if (int rc = grow_tile_done(c, most_tail)) return rc;
    void* tptr = c->tiles;
    size_t tbytes = c->tiles_n * sizeof(ppsk::Tile);
    if (int rc = grow(&tptr, &tbytes, tail_tiles.size() * sizeof(ppsk::Tile))) return rc;
    c->tiles = (ppsk::Tile*)tptr;
    c->tiles_n = tbytes / sizeof(ppsk::Tile);
    HIP_TRY(hipMemcpyAsync(c->tiles, tail_tiles.data(), tail_tiles.size() * sizeof(ppsk::Tile),
                           hipMemcpyHostToDevice, c->stream));
    p.counts = (uint16_t*)c->counts;
    p.tile_done = c->tile_done;
  }
  HIP_TRY(hipMemcpyAsync(c->pitems, items.data(), items.size() * sizeof(ppsk::BatchItem),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->pparams, &p, sizeof(p), hipMemcpyHostToDevice, c->stream));
  ppsk::BatchArgs args;
  std::memset(&args, 0, sizeof(args));
  args.jobs = c->pparams;
  args.items = (const ppsk::BatchItem*)c->pitems;
  args.queue_next = c->pqueue;
  args.scratch = (uint16_t*)c->pscratch;
  args.max_k = p.n_k;
  args.ctl = p.ctl;
  // The tail's launch goes on a second stream, after the persistent launch
  // has been queued: its (tile, k) items take the CU slots that the
  // persistent workgroups leave as they run out of tiles, instead of waiting
  // for the slowest of them (PPSK_TAIL_STREAM=0: the same stream).  The two
  // launches share nothing but the output rows they write, which are
  // disjoint; the main stream waits for the tail before the pass's end event.
  const char* ts_env = std::getenv("PPSK_TAIL_STREAM");
  const bool tail_beside = most_tail > 0 && !(ts_env && std::atoi(ts_env) == 0);
  if (tail_beside && !c->tail_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&c->tail_stream, hipStreamNonBlocking));
  }
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    for (int q = 0; q <= ppsk::BATCH_QUEUES; ++q) args.queue_begin[q] = qb[ci][q];
    hipEvent_t e0 = pool_event(c, 2 * ci), e1 = pool_event(c, 2 * ci + 1);
    hipEvent_t et = pool_event(c, 2 * chunks.size() + ci);
    if (!e0 || !e1 || !et) return fail(PPSK_ERR_HIP, "event creation failed");
    HIP_TRY(ppsk::launch_batch_reset(c->pqueue, nullptr, 0, c->stream));
    HIP_TRY(hipEventRecord(e0, c->stream));
    HIP_TRY(ppsk::launch_batch(args, kernel, grid, c->stream));
    tm->launches += 1;
    if (tail[ci] > 0) {
      ppsk::DistParams pt = p;
      pt.tiles = c->tiles + tail_at[ci];
      pt.n_tiles = (int64_t)tail[ci];
      pt.item_base = 0;
      hipStream_t ts = tail_beside ? c->tail_stream : c->stream;
      if (tail_beside) HIP_TRY(hipStreamWaitEvent(ts, e0, 0));
      if (diagonal) {
        HIP_TRY(ppsk::launch_dists_diagonal(pt, (int64_t)tail[ci] * p.n_k, ts));
      } else {
        HIP_TRY(ppsk::launch_dists(pt, variant, (int64_t)tail[ci] * p.n_k, ts));
      }
      if (tail_beside) {
        HIP_TRY(hipEventRecord(et, ts));
        HIP_TRY(hipStreamWaitEvent(c->stream, et, 0));
      }
      tm->launches += 1;
    }
    HIP_TRY(hipEventRecord(e1, c->stream));
    tm->items += (int64_t)(chunks[ci].second - chunks[ci].first) * p.n_k;
  }
  for (size_t ci = 0; host_out && ci < chunks.size(); ++ci) {
    if (int rc = copy_rows(c, tiles, chunks[ci].first, chunks[ci].second, tr, p, self, elem,
                           host_out, c->events[2 * ci + 1])) {
      return rc;
    }
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamSynchronize(c->copy_stream));
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->events[2 * ci], c->events[2 * ci + 1]));
    tm->ms += ms;
  }
  return PPSK_OK;
}

// Dense pass over rows [row_begin, row_end): the persistent kernel for the
// shapes it is built for when the pass has at least two tiles per resident
// workgroup (a smaller pass fills the GPU better split per (tile, k-mer
// length)); PPSK_PERSISTENT=0 / 1 forces the choice.
int run_dense(DeviceCtx* c, ppsk::DistParams p, int variant, bool upper, int64_t row_begin,
              int64_t row_end, bool self, size_t elem, void* host_out, Timing* tm) {
  const char* e = std::getenv("PPSK_PERSISTENT");
  bool persistent = variant == 2 || variant == 4 || variant == 5;
  if (e) {
    persistent = persistent && std::atoi(e) != 0;
  } else if (persistent) {
    const int tr = ppsk::variant_tile_rows(variant), tc = ppsk::variant_tile_cols(variant);
    const int64_t rt = (row_end - row_begin + tr - 1) / tr, ct = (p.n_cols + tc - 1) / tc;
    const int64_t tiles = upper ? rt * ct / 2 : rt * ct;
    int grid = 0;
    if (ppsk::batch_grid(variant == 5 ? ppsk::BATCH_MIXED : variant, &grid) != hipSuccess) grid = 0;
    persistent = tiles >= 2 * (int64_t)grid;
  }
  if (persistent) {
    return run_dense_persistent(c, p, variant, upper, row_begin, row_end, self, elem, host_out,
                                tm);
  }
  return run_dense_tiles(c, p, variant, upper, row_begin, row_end, self, elem, host_out, tm);
}

void fill_common(ppsk::DistParams& p, const ppsk_set* rows, const ppsk_set* cols,
                 const RandomIn& rnd, double tol) {
  std::memset(&p, 0, sizeof(p));
  p.row_stripes = rows->stripes;
  p.col_stripes = cols->stripes;
  p.n_rows = rows->n;
  p.n_cols = cols->n;
  p.nb_rows = rows->nb;
  p.nb_cols = cols->nb;
  p.n_k = rows->n_k;
  p.s64 = rows->s64;
  for (int k = 0; k < rows->n_k; ++k) p.kvals[k] = (float)rows->klist[k];
  p.tolerance = tol;
  p.random_mode = rnd.mode;
  p.n_clusters = rnd.n_clusters;
  p.jv_rows = rows->jv;
  p.jv_cols = cols->jv;
  p.cl_rows = rows->clusters;
  p.cl_cols = cols->clusters;
  p.row_prefetch = row_prefetch();
}

// Row-panel L2 prefetch in the pipelined tile loop: on unless
// PPSK_ROW_PREFETCH=0.
int row_prefetch() {
  const char* e = std::getenv("PPSK_ROW_PREFETCH");
  return e ? (std::atoi(e) != 0) : 1;
}

size_t tail_split_tiles(size_t n, int grid) {
  const char* e = std::getenv("PPSK_TAIL_SPLIT");
  if (e && std::atoi(e) == 0) return 0;
  const size_t g = (size_t)grid, r = g ? n % g : 0;
  return (n > g && r > 0 && r <= g / 2) ? r : 0;
}

int upload_table(DeviceCtx* c, const ppsk_set* s, const RandomIn& rnd, float** dtab) {
  *dtab = nullptr;
  if (rnd.mode != PPSK_RANDOM_MC) return PPSK_OK;
  const size_t bytes = (size_t)s->n_k * rnd.n_clusters * rnd.n_clusters * sizeof(float);
  HIP_TRY(hipMalloc((void**)dtab, bytes));
  HIP_TRY(hipMemcpyAsync(*dtab, rnd.table, bytes, hipMemcpyHostToDevice, c->stream));
  return PPSK_OK;
}

int reset_fail(DeviceCtx* c) {
  unsigned long long init[2] = {0ull, ~0ull};
  HIP_TRY(hipMemcpyAsync(c->fail, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  return PPSK_OK;
}

int read_fail(DeviceCtx* c, unsigned long long* cnt, unsigned long long* first) {
  unsigned long long v[2];
  HIP_TRY(hipMemcpyAsync(v, c->fail, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *cnt = v[0];
  *first = v[1];
  return PPSK_OK;
}

bool table_symmetric(const RandomIn& rnd, int n_k) {
  if (rnd.mode != PPSK_RANDOM_MC) return true;
  const int C = rnd.n_clusters;
  for (int k = 0; k < n_k; ++k) {
    for (int a = 0; a < C; ++a) {
      for (int b = a + 1; b < C; ++b) {
        if (rnd.table[((size_t)k * C + a) * C + b] != rnd.table[((size_t)k * C + b) * C + a]) {
          return false;
        }
      }
    }
  }
  return true;
}

// One dense pass (self or query) into the device buffer dout, optionally
// streamed to host_out; reports fit failures.
int dense_pass(DeviceCtx* c, const ppsk_set* rows, const ppsk_set* cols, bool self,
               const RandomIn& rnd, double tol, int32_t out_kind, void* dout, void* host_out,
               size_t elem, Timing* tm, unsigned long long* nf, unsigned long long* ff) {
  float* dtab = nullptr;
  if (int rc = upload_table(c, rows, rnd, &dtab)) return rc;
  ppsk::DistParams p;
  fill_common(p, rows, cols, rnd, tol);
  p.self_mode = self ? 1 : 0;
  p.full_square = 0;
  p.out_kind = out_kind;
  p.out = dout;
  p.mc_table = dtab;
  p.mc_ref_is_col = self ? 0 : 1;  // query mode: random_match(r1 = ref, q)
  p.fail_count = c->fail;
  p.fail_first = c->fail + 1;
  p.ctl = c->ctl;
  int rc = reset_fail(c);
  if (rc == PPSK_OK) {
    const int variant = default_variant();
    rc = run_dense(c, p, variant, self, 0, rows->n, self, elem, host_out, tm);
  }
  if (dtab) hipFree(dtab);
  if (rc) return rc;
  // the kernels skip their remaining work once a cancellation is seen
  if ((rc = cancel_status(c->device))) return rc;
  return read_fail(c, nf, ff);
}

// kNN selection over all rows of a symmetric self distance held as a
// condensed triangle on the device; outputs copied to host.
int knn_from_triangle(DeviceCtx* c, const float* tri, int64_t n, int col, int kNN, int64_t* i_vec,
                      int64_t* j_vec, float* dists) {
  int64_t *d_i = nullptr, *d_j = nullptr;
  float* d_d = nullptr;
  void* heap_ws = nullptr;  // row heaps above the LDS limit
  const size_t m = (size_t)n * kNN;
  const size_t hw = ppsk::knn_heap_workspace(n, kNN);
  hipError_t e = hipMalloc(&d_i, m * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc(&d_j, m * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc(&d_d, m * sizeof(float));
  if (e == hipSuccess && hw) e = hipMalloc(&heap_ws, hw);
  if (e == hipSuccess) {
    e = ppsk::launch_knn_select_triangle(tri, n, col, kNN, d_i, d_j, d_d, c->stream, heap_ws);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(i_vec, d_i, m * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(j_vec, d_j, m * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dists, d_d, m * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) hipStreamSynchronize(c->stream);  // nothing queued may outlive its buffers
  if (d_i) hipFree(d_i);
  if (d_j) hipFree(d_j);
  if (d_d) hipFree(d_d);
  if (heap_ws) hipFree(heap_ws);
  if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? PPSK_ERR_MEMORY : PPSK_ERR_HIP,
                                   std::string("kNN selection: ") + hipGetErrorString(e));
  return PPSK_OK;
}

// Threshold compaction of a device triangle into library-allocated host arrays.
int threshold_from_triangle(DeviceCtx* c, const float2* tri, int64_t n, int col, float cutoff,
                            int64_t** i_vec, int64_t** j_vec, float** dists, int64_t* n_out) {
  int64_t total = 0;
  int64_t *d_i = nullptr, *d_j = nullptr;
  float* d_d = nullptr;
  hipError_t e = ppsk::threshold_compact(tri, n, col, cutoff, &total, &d_i, &d_j, &d_d, c->stream);
  if (e != hipSuccess) return fail(PPSK_ERR_HIP, std::string("threshold: ") + hipGetErrorString(e));
  auto* hi = (int64_t*)std::malloc(std::max<int64_t>(total, 1) * sizeof(int64_t));
  auto* hj = (int64_t*)std::malloc(std::max<int64_t>(total, 1) * sizeof(int64_t));
  auto* hd = (float*)std::malloc(std::max<int64_t>(total, 1) * sizeof(float));
  if (hi && hj && hd && total > 0) {
    e = hipMemcpy(hi, d_i, total * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hj, d_j, total * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hd, d_d, total * sizeof(float), hipMemcpyDeviceToHost);
  }
  if (d_i) hipFree(d_i);
  if (d_j) hipFree(d_j);
  if (d_d) hipFree(d_d);
  if (!hi || !hj || !hd || e != hipSuccess) {
    std::free(hi);
    std::free(hj);
    std::free(hd);
    if (e != hipSuccess) return fail(PPSK_ERR_HIP, std::string("threshold copy: ") + hipGetErrorString(e));
    return fail(PPSK_ERR_MEMORY, "threshold host output");
  }
  *i_vec = hi;
  *j_vec = hj;
  *dists = hd;
  *n_out = total;
  return PPSK_OK;
}

int fit_status(unsigned long long nf, unsigned long long ff) {
  if (!nf) return PPSK_OK;
  return fail(PPSK_ERR_FIT, "Fitting k-mer gradient failed for " + std::to_string(nf) +
                                " pair(s); first at row " + std::to_string(ff));
}

void set_timing(const Timing& tm) {
  g_last_ms = tm.ms;
  g_last_launches = tm.launches;
  g_last_items = tm.items;
}

}  // namespace capi
}  // namespace ppsk

using namespace ppsk::capi;

extern "C" {

const char* ppsk_last_error(void) { return g_last_error.c_str(); }

const char* ppsk_version(void) { return "1.0.0"; }

int ppsk_device_init(int device_id, size_t* mem_free, size_t* mem_total, size_t* shared_size) {
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  size_t fr = 0, tot = 0;
  HIP_TRY(hipMemGetInfo(&fr, &tot));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_id));
  if (mem_free) *mem_free = fr;
  if (mem_total) *mem_total = tot;
  if (shared_size) *shared_size = prop.sharedMemPerBlock;
  return PPSK_OK;
}

int ppsk_set_create(int device_id, int64_t n_samples, int32_t n_k, const int32_t* klist,
                    int32_t sketchsize64, int32_t bbits, ppsk_set** out) {
  if (!out) return fail(PPSK_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_samples < 1) return fail(PPSK_ERR_ARG, "Query with empty ref or query list!");
  if (n_k < 1 || n_k > ppsk::MAXK || !klist) {
    return fail(PPSK_ERR_ARG, "n_k must be in [1, " + std::to_string(ppsk::MAXK) + "]");
  }
  for (int i = 1; i < n_k; ++i) {
    if (klist[i] <= klist[i - 1]) return fail(PPSK_ERR_ARG, "klist must be strictly ascending");
  }
  if (bbits != ppsk::BBITS) {
    return fail(PPSK_ERR_ARG, "only bbits = 14 sketches are supported (got " +
                                  std::to_string(bbits) + ")");
  }
  if (sketchsize64 < 1 || sketchsize64 > 1023) {
    return fail(PPSK_ERR_ARG, "sketchsize64 must be in [1, 1023]");
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  auto* s = new ppsk_set();
  s->device = device_id;
  s->n = n_samples;
  s->n_k = n_k;
  s->klist.assign(klist, klist + n_k);
  s->s64 = sketchsize64;
  s->bbits = bbits;
  s->nb = (int32_t)((n_samples + ppsk::BLK - 1) / ppsk::BLK);
  const size_t bytes = ((size_t)n_k * s->nb * sketchsize64 + ppsk::STRIPE_PAD_WORDS) *
                       ppsk::HL * ppsk::BLK * sizeof(uint32_t);
  hipError_t e = hipMalloc(&s->stripes, bytes);
  if (e != hipSuccess) {
    delete s;
    return fail(PPSK_ERR_MEMORY, std::string("sketch set allocation: ") + hipGetErrorString(e));
  }
  e = hipMemsetAsync(s->stripes, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    hipFree(s->stripes);
    delete s;
    return fail(PPSK_ERR_HIP, std::string("sketch set clear: ") + hipGetErrorString(e));
  }
  *out = s;
  return PPSK_OK;
}

int ppsk_set_upload(ppsk_set* s, int64_t first, int64_t count, const uint64_t* raw) {
  if (int rc = check_set(s, "set")) return rc;
  if (s->view) return fail(PPSK_ERR_STATE, "a prefix view is read-only");
  if (first < 0 || count < 0 || first + count > s->n) {
    return fail(PPSK_ERR_ARG, "upload range outside the set");
  }
  if (count == 0) return PPSK_OK;
  if (!raw) return fail(PPSK_ERR_ARG, "raw is NULL");
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(s->device, &c)) return rc;
  const size_t per_sample = (size_t)s->n_k * s->s64 * s->bbits * sizeof(uint64_t);
  const int64_t chunk = std::max<int64_t>(1, (int64_t)((256u << 20) / per_sample));
  void* staging = nullptr;
  HIP_TRY(hipMalloc(&staging, (size_t)std::min(chunk, count) * per_sample));
  int rc = PPSK_OK;
  for (int64_t c0 = 0; c0 < count && rc == PPSK_OK; c0 += chunk) {
    const int64_t m = std::min(chunk, count - c0);
    hipError_t e = hipMemcpyAsync(staging, (const char*)raw + (size_t)c0 * per_sample,
                                  (size_t)m * per_sample, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      e = ppsk::launch_stripe((const uint64_t*)staging, m, first + c0, s->n_k, s->s64, s->nb,
                              s->stripes, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(PPSK_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
  }
  hipFree(staging);
  return rc;
}

int ppsk_set_meta(ppsk_set* s, const uint64_t* lengths, int32_t use_rc, const uint16_t* clusters) {
  if (int rc = check_set(s, "set")) return rc;
  if (s->view) return fail(PPSK_ERR_STATE, "a prefix view is read-only");
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(s->device, &c)) return rc;
  if (lengths) {
    // csrs terms per sample and k (src/random/random_match.cpp:318-336)
    std::vector<double> jv((size_t)s->n * s->n_k);
    for (int kk = 0; kk < s->n_k; ++kk) {
      const int64_t k = s->klist[kk];
      const int64_t shift = (k - 1) * 2 + (use_rc ? 1 : 0);
      if (k < 1 || shift >= 64) {
        return fail(PPSK_ERR_ARG, "equal-base random correction is undefined for k = " +
                                      std::to_string(k) + " (k must be 1..32)");
      }
      const size_t match_chance = (size_t)1 << shift;
      const double base = 1 - (double)1 / match_chance;
      for (int64_t i = 0; i < s->n; ++i) {
        jv[(size_t)i * s->n_k + kk] = 1 - std::pow(base, (double)lengths[i]);
      }
    }
    if (!s->jv) HIP_TRY(hipMalloc(&s->jv, jv.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(s->jv, jv.data(), jv.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    s->has_lengths = true;
    s->use_rc = use_rc;
  }
  if (clusters) {
    s->max_cluster = 0;
    for (int64_t i = 0; i < s->n; ++i) s->max_cluster = std::max<int32_t>(s->max_cluster, clusters[i]);
    if (!s->clusters) HIP_TRY(hipMalloc(&s->clusters, (size_t)s->n * sizeof(uint16_t)));
    HIP_TRY(hipMemcpyAsync(s->clusters, clusters, (size_t)s->n * sizeof(uint16_t),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return PPSK_OK;
}

int ppsk_set_prefix(ppsk_set* parent, int64_t count, ppsk_set** out) {
  if (!out) return fail(PPSK_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (int rc = check_set(parent, "parent")) return rc;
  if (count < 1 || count > parent->n) {
    return fail(PPSK_ERR_ARG, "prefix length must be in [1, set size]");
  }
  // same buffers and stripe-block stride; only the sample count shrinks, so
  // every kernel index stays inside the parent's allocation
  auto* v = new ppsk_set(*parent);
  v->n = count;
  v->view = true;
  *out = v;
  return PPSK_OK;
}

int ppsk_set_free(ppsk_set* s) {
  if (!s) return PPSK_OK;
  if (s->view) {
    delete s;
    return PPSK_OK;
  }
  hipSetDevice(s->device);
  if (s->stripes) hipFree(s->stripes);
  if (s->jv) hipFree(s->jv);
  if (s->clusters) hipFree(s->clusters);
  delete s;
  return PPSK_OK;
}

int64_t ppsk_set_size(const ppsk_set* s) { return s ? s->n : 0; }

int ppsk_dense(ppsk_set* ref, ppsk_set* query, int32_t random_mode, int32_t n_clusters,
               const float* mc_table, double tolerance, int32_t out_kind, void* out,
               int32_t out_on_device, uint64_t* n_fail, int64_t* first_fail_row) {
  set_timing(Timing());
  if (int rc = check_set(ref, "ref")) return rc;
  const bool self = (query == nullptr || query == ref);
  if (!self) {
    if (int rc = check_compatible(ref, query)) return rc;
  }
  if (out_kind < PPSK_OUT_CORE_ACC || out_kind > PPSK_OUT_COUNTS) {
    return fail(PPSK_ERR_ARG, "unknown out_kind");
  }
  if (out_kind == PPSK_OUT_CORE_ACC && ref->n_k < 2) {
    return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths");
  }
  if (!out) return fail(PPSK_ERR_ARG, "out is NULL");
  const ppsk_set* rows = self ? ref : query;
  const ppsk_set* cols = ref;
  RandomIn rnd{random_mode, n_clusters, mc_table};
  if (int rc = check_random(rows, cols, rnd)) return rc;
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(ref->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  if (int rc = cancel_status(ref->device)) return rc;
  CallScope scope(c, pass_items(rows->n, cols->n, ref->n_k, default_variant(), self));

  const int64_t n_rows_out = self ? ref->n * (ref->n - 1) / 2 : query->n * ref->n;
  const size_t elem = out_kind == PPSK_OUT_CORE_ACC ? 2 * sizeof(float)
                      : out_kind == PPSK_OUT_JACCARD ? ref->n_k * sizeof(float)
                                                     : ref->n_k * sizeof(uint16_t);
  if (n_fail) *n_fail = 0;
  if (first_fail_row) *first_fail_row = -1;
  if (n_rows_out == 0) return PPSK_OK;
  void* dout = out;
  if (!out_on_device) {
    if (int rc = grow(&c->out, &c->out_bytes, (size_t)n_rows_out * elem)) return rc;
    dout = c->out;
  }
  Timing tm;
  unsigned long long nf = 0, ff = 0;
  int rc = dense_pass(c, rows, cols, self, rnd, tolerance, out_kind, dout,
                      out_on_device ? nullptr : out, elem, &tm, &nf, &ff);
  set_timing(tm);
  if (rc) return rc;
  if (n_fail) *n_fail = nf;
  if (first_fail_row) *first_fail_row = nf ? (int64_t)ff : -1;
  return fit_status(nf, ff);
}

int ppsk_knn(ppsk_set* ref, int32_t random_mode, int32_t n_clusters, const float* mc_table,
             double tolerance, int32_t jaccard, int32_t kNN, int32_t dist_col, int64_t* i_vec,
             int64_t* j_vec, float* dists, uint64_t* n_fail) {
  set_timing(Timing());
  if (int rc = check_set(ref, "ref")) return rc;
  if (kNN < 1 || kNN > ref->n) {
    return fail(PPSK_ERR_ARG, "kNN must be in [1, number of samples]");
  }
  if ((jaccard && dist_col >= ref->n_k) || (dist_col != 0 && dist_col != 1)) {
    return fail(PPSK_ERR_ARG, "dist_col out of range");
  }
  if (!jaccard && ref->n_k < 2) {
    return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths");
  }
  if (!i_vec || !j_vec || !dists) return fail(PPSK_ERR_ARG, "output is NULL");
  RandomIn rnd{random_mode, n_clusters, mc_table};
  if (int rc = check_random(ref, ref, rnd)) return rc;
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(ref->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  if (int rc = cancel_status(ref->device)) return rc;
  const int64_t n = ref->n;
  const int32_t kind = jaccard ? ppsk::OUT_ONE_MINUS_J : PPSK_OUT_CORE_ACC;
  Timing tm;
  unsigned long long nf = 0, ff = 0;
  int rc = PPSK_OK;
  const size_t tri_bytes = (size_t)n * (n - 1) / 2 * sizeof(float2);
  const bool symmetric = table_symmetric(rnd, ref->n_k) && tri_bytes <= ((size_t)96 << 30);
  CallScope scope(c, pass_items(n, n, ref->n_k, default_variant(), symmetric));
  if (symmetric) {
    // distances are symmetric: compute the upper triangle once and read both
    // (i, j) and (j, i) from it
    if ((rc = grow(&c->out, &c->out_bytes, std::max<size_t>(tri_bytes, 8)))) return rc;
    if (n > 1) {
      rc = dense_pass(c, ref, ref, true, rnd, tolerance, kind, c->out, nullptr, sizeof(float2),
                      &tm, &nf, &ff);
    }
    if (rc == PPSK_OK) {
      rc = knn_from_triangle(c, (const float*)c->out, n, dist_col, kNN, i_vec, j_vec, dists);
    }
  } else {
    // asymmetric Monte Carlo table: full square rows in bands
    const int variant = default_variant();
    int64_t band = std::max<int64_t>(ppsk::BLK, ((int64_t)(2ll << 30) / (n * 8)) / ppsk::BLK * ppsk::BLK);
    band = std::min<int64_t>(band, (n + ppsk::BLK - 1) / ppsk::BLK * ppsk::BLK);
    if ((rc = grow(&c->out, &c->out_bytes, (size_t)band * n * 2 * sizeof(float)))) return rc;
    float* dtab = nullptr;
    if ((rc = upload_table(c, ref, rnd, &dtab))) return rc;
    int64_t *d_i = nullptr, *d_j = nullptr;
    float* d_d = nullptr;
    void* heap_ws = nullptr;  // row heaps above the LDS limit
    const size_t hw = ppsk::knn_heap_workspace(band, kNN);
    hipError_t e = hipMalloc(&d_i, (size_t)band * kNN * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&d_j, (size_t)band * kNN * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&d_d, (size_t)band * kNN * sizeof(float));
    if (e == hipSuccess && hw) e = hipMalloc(&heap_ws, hw);
    if (e != hipSuccess) rc = fail(PPSK_ERR_MEMORY, "kNN output buffers");
    if (rc == PPSK_OK) rc = reset_fail(c);
    for (int64_t r0 = 0; r0 < n && rc == PPSK_OK; r0 += band) {
      const int64_t r1 = std::min(n, r0 + band);
      ppsk::DistParams p;
      fill_common(p, ref, ref, rnd, tolerance);
      p.self_mode = 1;
      p.full_square = 1;
      p.out_kind = kind;
      p.out = c->out;
      p.out_stride_rows = n;
      p.square_row0 = r0;
      p.mc_table = dtab;
      p.mc_ref_is_col = 0;
      p.fail_count = c->fail;
      p.fail_first = c->fail + 1;
      p.ctl = c->ctl;
      rc = run_dense(c, p, variant, false, r0, r1, true, sizeof(float2), nullptr, &tm);
      if (rc == PPSK_OK) rc = cancel_status(c->device);
      if (rc) break;
      e = ppsk::launch_knn_select((const float*)c->out, r0, r1 - r0, n, dist_col, kNN, d_i, d_j,
                                  d_d, c->stream, heap_ws);
      const size_t m = (size_t)(r1 - r0) * kNN;
      if (e == hipSuccess)
        e = hipMemcpyAsync(i_vec + r0 * kNN, d_i, m * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(j_vec + r0 * kNN, d_j, m * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(dists + r0 * kNN, d_d, m * sizeof(float), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = fail(PPSK_ERR_HIP, std::string("kNN select: ") + hipGetErrorString(e));
    }
    hipStreamSynchronize(c->stream);  // nothing queued may outlive its buffers
    if (dtab) hipFree(dtab);
    if (d_i) hipFree(d_i);
    if (d_j) hipFree(d_j);
    if (d_d) hipFree(d_d);
    if (heap_ws) hipFree(heap_ws);
    if (rc == PPSK_OK) rc = read_fail(c, &nf, &ff);
  }
  set_timing(tm);
  if (rc) return rc;
  if (jaccard) {
    // ranked on 1 - J, reported as J (src/api.cpp:423-429)
    for (int64_t x = 0; x < n * (int64_t)kNN; ++x) dists[x] = 1.0f - dists[x];
  }
  if (n_fail) *n_fail = nf;
  return fit_status(nf, ff);
}

int ppsk_threshold(ppsk_set* ref, int32_t random_mode, int32_t n_clusters, const float* mc_table,
                   double tolerance, int32_t dist_col, float cutoff, int64_t** i_vec,
                   int64_t** j_vec, float** dists, int64_t* n_out, uint64_t* n_fail) {
  set_timing(Timing());
  if (int rc = check_set(ref, "ref")) return rc;
  if (!(cutoff >= 0)) return fail(PPSK_ERR_ARG, "Distance threshold must be at least 0");
  if (dist_col != 0 && dist_col != 1) return fail(PPSK_ERR_ARG, "dist_col out of range");
  if (!i_vec || !j_vec || !dists || !n_out) return fail(PPSK_ERR_ARG, "output is NULL");
  if (ref->n_k < 2) {
    return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths");
  }
  *i_vec = nullptr;
  *j_vec = nullptr;
  *dists = nullptr;
  *n_out = 0;
  RandomIn rnd{random_mode, n_clusters, mc_table};
  if (int rc = check_random(ref, ref, rnd)) return rc;
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(ref->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  if (int rc = cancel_status(ref->device)) return rc;
  const int64_t n = ref->n;
  CallScope scope(c, pass_items(n, n, ref->n_k, default_variant(), true));
  const size_t tri_bytes = std::max<size_t>((size_t)n * (n - 1) / 2 * sizeof(float2), 8);
  if (int rc = grow(&c->out, &c->out_bytes, tri_bytes)) return rc;
  Timing tm;
  unsigned long long nf = 0, ff = 0;
  int rc = PPSK_OK;
  if (n > 1) {
    rc = dense_pass(c, ref, ref, true, rnd, tolerance, PPSK_OUT_CORE_ACC, c->out, nullptr,
                    sizeof(float2), &tm, &nf, &ff);
  }
  set_timing(tm);
  if (n_fail) *n_fail = nf;
  if (rc) return rc;
  if (nf) return fit_status(nf, ff);
  return threshold_from_triangle(c, (const float2*)c->out, n, dist_col, cutoff, i_vec, j_vec,
                                 dists, n_out);
}

int ppsk_self_pass(ppsk_set* ref, int32_t random_mode, int32_t n_clusters, const float* mc_table,
                   double tolerance, float* dense_out, int32_t kNN, int32_t knn_col,
                   int64_t* knn_i, int64_t* knn_j, float* knn_d, float cutoff, int32_t thr_col,
                   int64_t** thr_i, int64_t** thr_j, float** thr_d, int64_t* thr_n,
                   uint64_t* n_fail) {
  set_timing(Timing());
  if (int rc = check_set(ref, "ref")) return rc;
  if (ref->n_k < 2) {
    return fail(PPSK_ERR_ARG, "core/accessory distances need at least two k-mer lengths");
  }
  if (kNN < 0 || kNN > ref->n) {
    return fail(PPSK_ERR_ARG, "kNN must be in [0, number of samples]");
  }
  if (kNN > 0 && ((knn_col != 0 && knn_col != 1) || !knn_i || !knn_j || !knn_d)) {
    return fail(PPSK_ERR_ARG, "bad kNN output arguments");
  }
  if (cutoff > 0 && ((thr_col != 0 && thr_col != 1) || !thr_i || !thr_j || !thr_d || !thr_n)) {
    return fail(PPSK_ERR_ARG, "bad threshold output arguments");
  }
  RandomIn rnd{random_mode, n_clusters, mc_table};
  if (int rc = check_random(ref, ref, rnd)) return rc;
  if (!table_symmetric(rnd, ref->n_k)) {
    return fail(PPSK_ERR_ARG, "a single self pass needs symmetric random-match tables");
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(ref->device, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  if (int rc = cancel_status(ref->device)) return rc;
  const int64_t n = ref->n;
  CallScope scope(c, pass_items(n, n, ref->n_k, default_variant(), true));
  const size_t tri_bytes = std::max<size_t>((size_t)n * (n - 1) / 2 * sizeof(float2), 8);
  if (int rc = grow(&c->out, &c->out_bytes, tri_bytes)) return rc;
  Timing tm;
  unsigned long long nf = 0, ff = 0;
  int rc = PPSK_OK;
  if (n > 1) {
    rc = dense_pass(c, ref, ref, true, rnd, tolerance, PPSK_OUT_CORE_ACC, c->out, dense_out,
                    sizeof(float2), &tm, &nf, &ff);
  }
  set_timing(tm);
  if (n_fail) *n_fail = nf;
  if (rc) return rc;
  if (nf) return fit_status(nf, ff);
  if (kNN > 0) {
    if ((rc = knn_from_triangle(c, (const float*)c->out, n, knn_col, kNN, knn_i, knn_j, knn_d))) return rc;
  }
  if (cutoff > 0) {
    rc = threshold_from_triangle(c, (const float2*)c->out, n, thr_col, cutoff, thr_i, thr_j,
                                 thr_d, thr_n);
  }
  return rc;
}

int ppsk_progress(int device_id, uint64_t* done, uint64_t* total, int32_t* running,
                  int32_t* pending) {
  if (device_id < 0 || device_id >= MAX_DEVICES) return fail(PPSK_ERR_ARG, "device_id out of range");
  CallState& cs = call_state(device_id);
  if (done) *done = read_progress(cs);
  if (total) *total = cs.total.load();
  if (running) *running = cs.running.load();
  if (pending) *pending = cs.requested.load() ? 1 : 0;
  return PPSK_OK;
}

int ppsk_cancel(int device_id) {
  if (device_id < 0 || device_id >= MAX_DEVICES) return fail(PPSK_ERR_ARG, "device_id out of range");
  CallState& cs = call_state(device_id);
  cs.requested.store(1);
  if (volatile uint32_t* w = cs.cancel_host.load()) *w = 1u;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  return PPSK_OK;
}

int ppsk_cancel_reset(int device_id) {
  if (device_id < 0 || device_id >= MAX_DEVICES) return fail(PPSK_ERR_ARG, "device_id out of range");
  CallState& cs = call_state(device_id);
  cs.requested.store(0);
  if (volatile uint32_t* w = cs.cancel_host.load()) *w = 0u;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  return PPSK_OK;
}

void ppsk_free_host(void* ptr) { std::free(ptr); }

int ppsk_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(PPSK_ERR_ARG, "ptr is NULL");
  HIP_TRY(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  return PPSK_OK;
}

int ppsk_host_free(void* ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return PPSK_OK;
}

int ppsk_device_alloc(int device_id, size_t bytes, void** ptr) {
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  HIP_TRY(hipMalloc(ptr, bytes));
  return PPSK_OK;
}

int ppsk_device_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return PPSK_OK;
}

int ppsk_copy_to_host(void* dst, const void* device_src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, device_src, bytes, hipMemcpyDeviceToHost));
  return PPSK_OK;
}

int ppsk_last_timing(double* kernel_ms, int64_t* n_launches, int64_t* n_items) {
  if (kernel_ms) *kernel_ms = g_last_ms;
  if (n_launches) *n_launches = g_last_launches;
  if (n_items) *n_items = g_last_items;
  return PPSK_OK;
}

}  // extern "C"
