// Shared internals of the C-ABI layer (ppsk_capi.cpp, batch.cpp, serve.cpp):
// the sketch-set handle, per-device contexts, argument checks and the tile
// scheduling helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ppsk.h"
#include "ppsk_internal.h"

struct ppsk_set {
  int device;
  int64_t n;
  int32_t n_k;
  std::vector<int32_t> klist;
  int32_t s64;
  int32_t bbits;
  int32_t nb;
  uint32_t* stripes = nullptr;  // [n_k][nb][s64][28][128]
  double* jv = nullptr;         // [n][n_k]
  uint16_t* clusters = nullptr; // [n]
  bool has_lengths = false;
  int32_t use_rc = 1;
  int32_t max_cluster = -1;     // largest cluster id uploaded
  bool view = false;            // prefix view: shares the parent's buffers
};

namespace ppsk {
namespace capi {

extern thread_local std::string g_last_error;
extern thread_local double g_last_ms;
extern thread_local int64_t g_last_launches;
extern thread_local int64_t g_last_items;

// Records msg as this thread's last error and returns code.
int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      return fail(_e == hipErrorOutOfMemory ? PPSK_ERR_MEMORY : PPSK_ERR_HIP,      \
                  std::string(#expr) + ": " + hipGetErrorString(_e));              \
    }                                                                              \
  } while (0)

// Per-device state: a launch stream, a copy stream, cached workspaces grown on
// demand.
struct DeviceCtx {
  int device = -1;
  hipStream_t stream = nullptr;
  void* counts = nullptr;
  size_t counts_bytes = 0;
  uint32_t* tile_done = nullptr;
  size_t tile_done_n = 0;
  ppsk::Tile* tiles = nullptr;
  size_t tiles_n = 0;
  unsigned long long* fail = nullptr;  // [0] count, [1] first row
  void* out = nullptr;
  size_t out_bytes = 0;
  hipStream_t copy_stream = nullptr;  // device -> host copies of finished rows
  std::vector<hipEvent_t> events;     // per-chunk timing / copy dependencies
  // persistent launches (dist_batch over the tiles of one call)
  void* pscratch = nullptr;           // per-workgroup count blocks
  size_t pscratch_bytes = 0;
  void* pitems = nullptr;             // BatchItem list
  size_t pitems_bytes = 0;
  ppsk::DistParams* pparams = nullptr;  // the call's parameters in device memory
  uint32_t* pqueue = nullptr;         // queue counters
  hipStream_t tail_stream = nullptr;  // the tail's (tile, k) launch beside the persistent one
  void* noalloc_scratch = nullptr;    // this device's copy of the fixed scratch array
  ppsk::Control ctl{};                // progress / cancellation pointers of the calls
  std::mutex mu;
};

// Progress and cancellation state of one device, kept in host memory so
// that ppsk_progress / ppsk_cancel work from any thread, with or without a
// device context.  The kernels see the request through a host-mapped word
// (cancel_host) and report progress through a device counter mirrored into
// progress_host; both words are allocated with the device context.
constexpr int MAX_DEVICES = 64;
struct CallState {
  std::atomic<uint32_t> requested{0};
  std::atomic<int32_t> running{0};
  std::atomic<uint64_t> total{0};
  std::atomic<uint64_t> seen{0};  // largest progress value read
  std::atomic<volatile uint32_t*> cancel_host{nullptr};
  std::atomic<volatile unsigned long long*> progress_host{nullptr};
};
CallState& call_state(int device);
bool cancel_requested(int device);
// Marks a call on the device as running with `total` work items and clears
// its progress (the device counter in stream order on `stream`); the
// destructor marks it finished.
// The destructor takes the final count from the device counter (the mirror
// may hold an earlier value when workgroups' stores landed out of order).
struct CallScope {
  DeviceCtx* c;
  CallState& cs;
  CallScope(DeviceCtx* c, uint64_t total);
  ~CallScope();
};
// PPSK_ERR_CANCELLED with a message when a cancellation stands, else PPSK_OK.
int cancel_status(int device);
// Work items ((tile, k-mer length) pairs) of one pass.
uint64_t pass_items(int64_t n_rows, int64_t n_cols, int n_k, int variant, bool upper);

// Context of a device (created on first use) made current.
int get_ctx(int device, DeviceCtx** out);
// Context of a device if it exists (made current), else null; never creates.
DeviceCtx* find_ctx(int device);
// Device address of the fixed scratch array of ppsk_dense_noalloc on the
// current device (hipGetSymbolAddress), called once per device at context
// creation.
int noalloc_init(void** scratch);
// Device buffer of at least `need` bytes (reallocated, contents lost).
int grow(void** ptr, size_t* have, size_t need);
// Tile completion counters, zero on return.
int grow_tile_done(DeviceCtx* c, size_t n);

int check_set(const ppsk_set* s, const char* what);
int check_compatible(const ppsk_set* a, const ppsk_set* b);

std::vector<ppsk::Tile> make_tiles(int64_t n_rows, int64_t n_cols, int tr, int tc, bool upper,
                                   int64_t row_begin, int64_t row_end, int group,
                                   std::vector<size_t>* group_start);

struct RandomIn {
  int32_t mode;
  int32_t n_clusters;
  const float* table;
};

int check_random(const ppsk_set* rows, const ppsk_set* cols, const RandomIn& rnd);
// Row tiles per scheduling group for tiles of tile_rows rows: groups of 512
// sample rows (8 tiles of 64 rows, 32 of 16), so that a column panel fetched
// into an XCD's L2 serves the same number of rows whatever the tile shape;
// PPSK_TILE_GROUP overrides the tile count.
int tile_group(int tile_rows);
int default_variant();

// Timing of the tile-kernel launches of one call: per-chunk events on the
// launch stream, summed once the call has finished.
struct Timing {
  double ms = 0;
  int64_t launches = 0;
  int64_t items = 0;
};

hipEvent_t pool_event(DeviceCtx* c, size_t i);

// Distance parameters shared by every entry point (rows = r1 of the Jaccard).
void fill_common(ppsk::DistParams& p, const ppsk_set* rows, const ppsk_set* cols,
                 const RandomIn& rnd, double tol);
// 1 unless PPSK_ROW_PREFETCH=0 (row-panel L2 prefetch of the tile kernels).
int row_prefetch();
// Tiles of a persistent launch of n whole tiles over `grid` workgroups that
// run instead as (tile, k-mer length) items of the per-(tile, k) kernel: the
// last, partial round when it would fill at most half of the workgroup slots
// (0 otherwise, or with PPSK_TAIL_SPLIT=0).
size_t tail_split_tiles(size_t n, int grid);
// Device copy of a Monte Carlo table (null for the other modes); caller frees.
int upload_table(DeviceCtx* c, const ppsk_set* s, const RandomIn& rnd, float** dtab);
// PPSK_OK, or PPSK_ERR_FIT with the failure count and first row.
int fit_status(unsigned long long nf, unsigned long long ff);
void set_timing(const Timing& tm);

}  // namespace capi
}  // namespace ppsk
