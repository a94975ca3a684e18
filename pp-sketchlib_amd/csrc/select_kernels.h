// Sparse-output kernels: exact kNN selection and threshold compaction.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace ppsk {

// A row's heap of kNN + 1 entries lives in LDS up to this kNN; larger heaps
// live in a global-memory workspace (knn_heap_workspace bytes, passed as
// heap_ws), which the launchers fill with at most KNN_HEAP_ROWS rows at a
// time.
constexpr int KNN_LDS_MAX = 2048;
constexpr int64_t KNN_HEAP_ROWS = 32768;
size_t knn_heap_workspace(int64_t n_rows, int kNN);

// band: float2 (core, acc) rows [r0, r0 + n_rows) x n columns, row stride n.
hipError_t launch_knn_select(const float* band, int64_t r0, int64_t n_rows, int64_t n,
                             int dist_col, int kNN, int64_t* i_vec, int64_t* j_vec,
                             float* dists, hipStream_t stream, void* heap_ws = nullptr);

// Same selection over all n rows, reading (i, j) and (j, i) from the
// condensed upper triangle (float2 per pair) of a symmetric distance.
hipError_t launch_knn_select_triangle(const float* tri, int64_t n, int dist_col, int kNN,
                                      int64_t* i_vec, int64_t* j_vec, float* dists,
                                      hipStream_t stream, void* heap_ws = nullptr);

// Query rows against a separate reference set: band [n_rows][n] float2,
// nothing excluded; i_vec holds the band row.  segs > 1: each row is cut into
// segs column segments that list their candidates in parallel before one wave
// per row replays the heap over the lists (same result), for steps of few
// rows; the lists live in `workspace` (knn_rect_workspace bytes), which also
// holds the row heaps when kNN > KNN_LDS_MAX (then segs must be 1).
hipError_t launch_knn_select_rect(const float* band, int64_t n_rows, int64_t n, int dist_col,
                                  int kNN, int64_t* i_vec, int64_t* j_vec, float* dists,
                                  hipStream_t stream, int segs = 1, void* workspace = nullptr);
// Candidate list capacity per (row, segment); a longer list falls back to a
// direct scan of its segment in the merge.
int knn_rect_capacity(int kNN);
size_t knn_rect_workspace(int64_t n_rows, int segs, int kNN);

// Threshold in two passes over `rows` rows: counts of d < cutoff, then (after
// an exclusive scan of the counts into offsets) the entries in row-major
// order.  rect: rows of n entries of a query-major band (j = column);
// otherwise the condensed self triangle of n samples (rows = n - 1, j > i).
// Each row is cut into `segs` equal column segments, one workgroup each, with
// one count / offset per (row, segment) in row-major order (rows * segs
// entries), so that a few long rows still fill the GPU.  Outputs may be
// device or page-locked host memory.
hipError_t launch_threshold_count(const float2* dd, int64_t rows, int64_t n, bool rect,
                                  int dist_col, float cutoff, int64_t* counts,
                                  hipStream_t stream, int segs = 1);
// Entries at offsets >= capacity are not written (the caller sizes the
// outputs from the total and writes again).
hipError_t launch_threshold_write(const float2* dd, int64_t rows, int64_t n, bool rect,
                                  int dist_col, float cutoff, const int64_t* offsets,
                                  int64_t* i_vec, int64_t* j_vec, float* dists,
                                  hipStream_t stream, int64_t capacity = INT64_MAX,
                                  int segs = 1);
// Per-row counts -> exclusive offsets in place, total in counts[rows]
// (device memory of rows + 1 entries; one workgroup).
hipError_t launch_offsets_scan(int64_t* counts, int64_t rows, hipStream_t stream);

// dd: condensed float2 (core, acc) of the self upper triangle of n samples.
// Allocates the device outputs (*d_i, *d_j, *d_d; caller frees with hipFree).
hipError_t threshold_compact(const float2* dd, int64_t n, int dist_col, float cutoff,
                             int64_t* total, int64_t** d_i, int64_t** d_j, float** d_d,
                             hipStream_t stream);

}  // namespace ppsk
