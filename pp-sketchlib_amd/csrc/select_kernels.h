// Sparse-output kernels: exact kNN selection and threshold compaction.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace ppsk {

constexpr int KNN_MAX = 2048;  // heap kept in LDS

// band: float2 (core, acc) rows [r0, r0 + n_rows) x n columns, row stride n.
hipError_t launch_knn_select(const float* band, int64_t r0, int64_t n_rows, int64_t n,
                             int dist_col, int kNN, int64_t* i_vec, int64_t* j_vec,
                             float* dists, hipStream_t stream);

// Same selection over all n rows, reading (i, j) and (j, i) from the
// condensed upper triangle (float2 per pair) of a symmetric distance.
hipError_t launch_knn_select_triangle(const float* tri, int64_t n, int dist_col, int kNN,
                                      int64_t* i_vec, int64_t* j_vec, float* dists,
                                      hipStream_t stream);

// dd: condensed float2 (core, acc) of the self upper triangle of n samples.
// Allocates the device outputs (*d_i, *d_j, *d_d; caller frees with hipFree).
hipError_t threshold_compact(const float2* dd, int64_t n, int dist_col, float cutoff,
                             int64_t* total, int64_t** d_i, int64_t** d_j, float** d_d,
                             hipStream_t stream);

}  // namespace ppsk
