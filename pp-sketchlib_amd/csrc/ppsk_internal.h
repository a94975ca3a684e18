// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Device layout of a sketch set ("stripes"), for one k-mer length index kk:
//   S[kk][nb][w][hl][BLK]  (uint32)
//     nb : block of BLK = 128 samples (sample n -> nb = n / 128, lane n % 128)
//     w  : 64-bin word, 0 .. s64-1
//     hl : 32-bit plane, hl = h * BBITS + l, h = 0 low / 1 high dword of the
//          uint64 layer word, l = bit layer 0 .. BBITS-1
// so every (kk, nb, w) is one contiguous 28 x 128 dword slab (14 KiB) and a
// run of consecutive samples for one plane is contiguous.  Padding samples
// (n >= n_samples inside the last block) are zero and are masked on output.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace ppsk {

constexpr int BLK = 128;       // samples per stripe block
constexpr int BBITS = 14;      // bit layers per bin value (reference def_bbits)
constexpr int HL = 2 * BBITS;  // 32-bit planes per 64-bin word
constexpr int WAVES = 8;       // waves per workgroup
constexpr int MAXK = 32;       // most k-mer lengths per query
// zero words appended to every stripe buffer: the tile kernel stages whole
// groups of up to 8 words and may read one group past the last word
constexpr int STRIPE_PAD_WORDS = 8;

enum OutKind : int32_t { OUT_CORE_ACC = 0, OUT_JACCARD = 1, OUT_COUNTS = 2, OUT_ONE_MINUS_J = 3 };
enum RandomMode : int32_t { RANDOM_NONE = 0, RANDOM_EQUAL_BASE = 1, RANDOM_MC = 2 };

// Progress and cancellation of a call.  cancel: host-mapped flag written by
// ppsk_cancel, read by the kernels between work items with system-scope
// loads; progress: device counter of finished (tile, k) items, each update
// mirrored with a plain system-scope store into the host-mapped word
// progress_host (stores of different workgroups may land out of order: the
// host keeps the largest value it has read).
struct Control {
  const uint32_t* cancel;
  unsigned long long* progress;
  unsigned long long* progress_host;
};

// Tile: rows [bi * rows, +rows) x cols [ct * cols, +cols) of the variant's
// tile shape (variant_tile_rows / variant_tile_cols).
struct Tile {
  int32_t bi;
  int32_t ct;
};

struct DistParams {
  const uint32_t* row_stripes;  // set giving the tile rows (r1 of the Jaccard)
  const uint32_t* col_stripes;  // set giving the tile columns
  int64_t n_rows, n_cols;       // real sample counts
  int32_t nb_rows, nb_cols;     // stripe blocks in each set
  int32_t n_k, s64;
  int32_t self_mode;            // rows and cols are the same set, upper triangle
  int32_t full_square;          // self set, every i != j (kNN); i == j -> +inf
  const Tile* tiles;
  int64_t n_tiles;
  int64_t item_base;            // first (tile * n_k + kk) item of this launch
  uint16_t* counts;             // scratch [n_tiles][n_k][tile rows][tile cols]
  uint32_t* tile_done;          // [n_tiles] completion counters, left at zero
  // epilogue
  int32_t out_kind;
  void* out;                    // dense output (see OutKind)
  int64_t out_row_offset;       // subtracted from the dense row index
  int64_t out_stride_rows;      // for full_square: row stride of the [rows][n] block
  int64_t square_row0;          // for full_square: first global row in out
  float kvals[MAXK];            // k as float (reference kmer2mat)
  double tolerance;             // regression validity threshold
  int32_t random_mode;
  int32_t n_clusters;
  int32_t mc_ref_is_col;        // MC table index order: [col][row] instead of [row][col]
  const double* jv_rows;        // [n_rows][n_k] equal-base j(len) per sample
  const double* jv_cols;
  const float* mc_table;        // [n_k][C][C]
  const uint16_t* cl_rows;      // [n_rows] cluster ids
  const uint16_t* cl_cols;
  unsigned long long* fail_count;
  unsigned long long* fail_first;  // smallest failing dense row
  uint32_t* fail_rows;             // optional [n_rows] failures per tile row sample
  int32_t row_prefetch;            // pull each stage's row panel into L2 a stage ahead
  Control ctl;                     // progress and cancellation (all null: neither)
};

// Persistent batch kernel (dist_batch): work queues and items.
constexpr int BATCH_QUEUES = 8;  // one per XCD of the MI355X

struct BatchItem {  // one tile of one job: every k-mer length, then the epilogue
  int32_t job;
  int32_t bi;
  int32_t ct;
  int32_t pad;  // tile shape (variant) of the item for the mixed-shape kernel
};

struct BatchArgs {
  const DistParams* jobs;   // [n_jobs] per-job parameters (device memory)
  const BatchItem* items;   // queue q holds items [queue_begin[q], queue_begin[q + 1])
  int64_t queue_begin[BATCH_QUEUES + 1];
  uint32_t* queue_next;     // [BATCH_QUEUES] next item of each queue (zeroed before launch)
  uint16_t* scratch;        // [grid][max_k][tile rows][tile cols] per-workgroup counts
  int32_t max_k;
  Control ctl;              // progress and cancellation of the launch (null: neither)
};

// Launchers (dist_kernels.hip).  All work goes on `stream`.
hipError_t launch_stripe(const uint64_t* raw, int64_t count, int64_t first,
                         int32_t n_k, int32_t s64, int32_t nb, uint32_t* stripes,
                         hipStream_t stream);

// Tile kernel configurations ("variants"): rows per wave RQ (row operands in
// SGPRs), columns per lane RR, row-operand pipeline PB (0 = left to the
// compiler, 2 = explicit double-buffered scalar loads of 2 planes); the 8
// waves of a workgroup form a grid of row groups x column groups.
//   1: RQ16 RR1 PB0, waves 8 x 1 (128 x 64)   3: RQ16 RR2 PB0, 8 x 1 (128 x 128)
//   2: RQ8  RR4 PB2, waves 8 x 1 (64 x 256)   4: RQ8  RR2 PB2, 4 x 2 (32 x 256)
//   5: RQ8  RR1 PB2, waves 2 x 4 (16 x 256)
// A pipelined (PB = 2) configuration is only correct if the compiler keeps
// the in-flight SGPR tuples in place; the build rejects any that spill.
constexpr int DEFAULT_VARIANT = 2;
constexpr int N_VARIANTS = 5;
bool variant_valid(int variant);
int variant_tile_rows(int variant);
int variant_tile_cols(int variant);
hipError_t launch_dists(const DistParams& p, int variant, int64_t n_items,
                        hipStream_t stream);

// Persistent launch with a tile shape per item (BatchItem.pad = 2, 4 or 5):
// the `variant` of batch_grid / launch_batch that selects it.  Its scratch
// holds [grid][max_k][64][256] counts whatever the shapes.
constexpr int BATCH_MIXED = 6;

// Persistent batch launch: `grid` workgroups (batch_grid: every workgroup
// resident at once).  batch_reset zeroes the queue counters and sets the
// per-job fail counters ([n_jobs][2]: count, first row) to (0, max) with a
// kernel, so that both launches can be captured in a graph.
hipError_t batch_grid(int variant, int* grid);
hipError_t launch_batch(const BatchArgs& a, int variant, int grid, hipStream_t stream);
hipError_t launch_batch_reset(uint32_t* queue_next, unsigned long long* fail, int n_jobs,
                              hipStream_t stream, unsigned long long* progress = nullptr);
// Zeroes n tile completion counters and sets fail ([2]: count, first row; may
// be null) to (0, max), with a kernel.
// launch_dists for 64 x 256 tiles (variant 2) of a self pass (upper
// triangle): diagonal tiles skip the column blocks that hold no output pair.
hipError_t launch_dists_diagonal(const DistParams& p, int64_t n_items, hipStream_t stream);
// Writes the table of the distance epilogue's logarithm into the current
// device's copy (once per device, before its first distance kernel).
hipError_t upload_log_table();
hipError_t launch_tile_reset(uint32_t* tile_done, int64_t n, unsigned long long* fail,
                             hipStream_t stream);

}  // namespace ppsk
