// ppsk_dense_noalloc (include/ppsk.h): the dense distance pass for callers
// that cannot allocate device memory.  Its working memory is one fixed
// file-scope device array; every GPU holds its own copy of it, whose address
// is looked up once per device (noalloc_init, at device-context creation) and
// passed to the kernels as an ordinary pointer.
//
// Working memory layout (byte offsets into the array, 256-byte aligned):
//   [fail counters: 2 x u64][Monte Carlo table: n_k x C x C float]
//   [tile list of one launch: chunk x Tile][tile completion counters: chunk x u32]
//   [per-tile match counts of one launch: chunk x n_k x rows x cols u16]
// The tile list is cut into launches of `chunk` tiles of the per-(tile, k)
// kernel, so any number of samples fits; only the table has a size limit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "capi_common.h"

static __device__ __align__(256) unsigned char scratch[PPSK_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

namespace ppsk {
namespace capi {

int noalloc_init(void** out) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    return fail(PPSK_ERR_HIP, std::string("scratch array address: ") + hipGetErrorString(e));
  }
  *out = p;
  return PPSK_OK;
}

}  // namespace capi
}  // namespace ppsk

using namespace ppsk::capi;

extern "C" {

int ppsk_dense_noalloc(ppsk_set* ref, ppsk_set* query, int32_t random_mode, int32_t n_clusters,
                       const float* mc_table, double tolerance, int32_t out_kind,
                       void* out_device, uint64_t* n_fail, int64_t* first_fail_row) {
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
  if (!out_device) return fail(PPSK_ERR_ARG, "out_device is NULL");
  const ppsk_set* rows = self ? ref : query;
  const ppsk_set* cols = ref;
  const RandomIn rnd{random_mode, n_clusters, mc_table};
  if (int rc = check_random(rows, cols, rnd)) return rc;
  DeviceCtx* c = find_ctx(ref->device);
  if (!c || !c->noalloc_scratch) {
    return fail(PPSK_ERR_STATE, "device " + std::to_string(ref->device) +
                                    " not initialised: call ppsk_device_init first");
  }
  std::lock_guard<std::mutex> lk(c->mu);
  if (int rc = cancel_status(ref->device)) return rc;
  if (n_fail) *n_fail = 0;
  if (first_fail_row) *first_fail_row = -1;
  const int64_t n_rows_out = self ? ref->n * (ref->n - 1) / 2 : query->n * ref->n;
  if (n_rows_out == 0) return PPSK_OK;

  // working memory: fixed part, then as many tiles per launch as fit
  char* base = (char*)c->noalloc_scratch;
  const size_t fail_off = 0;
  const size_t table_off = align_up(2 * sizeof(unsigned long long));
  const size_t table_bytes = rnd.mode == PPSK_RANDOM_MC
                                 ? (size_t)ref->n_k * n_clusters * n_clusters * sizeof(float)
                                 : 0;
  const size_t fixed = align_up(table_off + table_bytes);
  const int variant = default_variant();
  CallScope scope(c, pass_items(rows->n, cols->n, ref->n_k, variant, self));
  const int tr = ppsk::variant_tile_rows(variant);
  const int tc = ppsk::variant_tile_cols(variant);
  const size_t per_tile = (size_t)ref->n_k * tr * tc * sizeof(uint16_t);
  // per tile: its list entry, its completion counter and its counts, plus
  // two alignment paddings per launch
  const size_t room = PPSK_NOALLOC_SCRATCH_BYTES > fixed + 2 * kAlign
                          ? PPSK_NOALLOC_SCRATCH_BYTES - fixed - 2 * kAlign
                          : 0;
  const size_t chunk = room / (per_tile + sizeof(ppsk::Tile) + sizeof(uint32_t));
  if (chunk < 1) {
    return fail(PPSK_ERR_ARG, "Monte Carlo table of " + std::to_string(table_bytes) +
                                  " bytes does not fit the " +
                                  std::to_string((size_t)PPSK_NOALLOC_SCRATCH_BYTES) +
                                  "-byte scratch array of ppsk_dense_noalloc");
  }
  const size_t tiles_off = fixed;
  const size_t done_off = align_up(tiles_off + chunk * sizeof(ppsk::Tile));
  const size_t counts_off = align_up(done_off + chunk * sizeof(uint32_t));
  if (counts_off + chunk * per_tile > PPSK_NOALLOC_SCRATCH_BYTES) {
    return fail(PPSK_ERR_STATE, "scratch layout overflow");  // excluded by the sizes above
  }
  unsigned long long* d_fail = (unsigned long long*)(base + fail_off);
  ppsk::Tile* d_tiles = (ppsk::Tile*)(base + tiles_off);
  uint32_t* d_done = (uint32_t*)(base + done_off);

  ppsk::DistParams p;
  fill_common(p, rows, cols, rnd, tolerance);
  p.self_mode = self ? 1 : 0;
  p.full_square = 0;
  p.out_kind = out_kind;
  p.out = out_device;
  p.out_row_offset = 0;
  p.mc_ref_is_col = self ? 0 : 1;  // query mode: random_match(r1 = ref, q)
  p.fail_count = d_fail;
  p.fail_first = d_fail + 1;
  p.counts = (uint16_t*)(base + counts_off);
  p.tile_done = d_done;
  p.ctl = c->ctl;
  if (table_bytes) {
    HIP_TRY(hipMemcpyAsync(base + table_off, mc_table, table_bytes, hipMemcpyHostToDevice,
                           c->stream));
    p.mc_table = (const float*)(base + table_off);
  }
  // the array keeps the last call's contents: counters start from zero
  HIP_TRY(ppsk::launch_tile_reset(d_done, (int64_t)chunk, d_fail, c->stream));
  std::vector<size_t> gs;
  const std::vector<ppsk::Tile> tiles =
      make_tiles(rows->n, cols->n, tr, tc, self, 0, rows->n, tile_group(tr), &gs);
  Timing tm;
  for (size_t a = 0; a < tiles.size(); a += chunk) {
    const size_t b = std::min(tiles.size(), a + chunk);
    // stream-ordered after the previous launch, which read the same list slots
    HIP_TRY(hipMemcpyAsync(d_tiles, tiles.data() + a, (b - a) * sizeof(ppsk::Tile),
                           hipMemcpyHostToDevice, c->stream));
    p.tiles = d_tiles;
    p.n_tiles = (int64_t)(b - a);
    p.item_base = 0;
    HIP_TRY(ppsk::launch_dists(p, variant, (int64_t)(b - a) * p.n_k, c->stream));
    tm.launches += 1;
    tm.items += (int64_t)(b - a) * p.n_k;
  }
  unsigned long long f[2] = {0ull, 0ull};
  HIP_TRY(hipMemcpyAsync(f, d_fail, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  set_timing(tm);
  // the kernels skip their remaining work once a cancellation is seen
  if (int rc = cancel_status(ref->device)) return rc;
  if (n_fail) *n_fail = f[0];
  if (first_fail_row) *first_fail_row = f[0] ? (int64_t)f[1] : -1;
  return fit_status(f[0], f[1]);
}

}  // extern "C"
