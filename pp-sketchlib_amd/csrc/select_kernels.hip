// Sparse outputs on the device.
//
// kNN: the reference (src/api.cpp:360-414) scans j = 0..n-1 for every row i
// and keeps the kNN smallest distances in a std::priority_queue max-heap,
// admitting an element when the heap is not full or when it is strictly
// smaller than the heap top, popping the top when the size exceeds kNN, and
// draining the heap back to front.  Which tied elements survive and their
// order depend on that exact heap algorithm, so it is replayed here: one wave
// per row, the row scanned 64 columns at a time, a wave ballot picking the
// columns that can be admitted against the current top (the top only ever
// decreases, so every other column is rejected exactly as in the scan), and
// lane 0 applying the binary-heap push / pop of the C++ standard library's
// priority_queue (hole sifting, right child preferred unless strictly
// smaller) to the admitted elements in column order.
//
// Threshold: the reference converts the dense column to a square matrix and
// emits the upper triangle row-major where d < cutoff
// (src/dist/matrix_ops.cpp:38-69).  Condensed order is exactly row-major
// upper-triangle order, so this is a stream compaction of the condensed
// column: per-row counts, an exclusive scan, then an order-preserving write.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "select_kernels.h"

namespace ppsk {

struct HeapEnt {
  float d;
  int32_t j;
};

// std::push_heap on [0, len] after placing v at index len (max-heap on d).
__device__ __forceinline__ void heap_sift_up(HeapEnt* h, int hole, int top, HeapEnt v) {
  int parent = (hole - 1) / 2;
  while (hole > top && h[parent].d < v.d) {
    h[hole] = h[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  h[hole] = v;
}

__device__ __forceinline__ void heap_push(HeapEnt* h, int len, HeapEnt v) {
  heap_sift_up(h, len, 0, v);
}

// std::pop_heap on [0, len): moves the top to index len - 1.
__device__ __forceinline__ void heap_pop(HeapEnt* h, int len) {
  if (len <= 1) return;
  const int L = len - 1;
  const HeapEnt v = h[L];
  h[L] = h[0];
  int hole = 0;
  int second = 0;
  while (second < (L - 1) / 2) {
    second = 2 * (second + 1);
    if (h[second].d < h[second - 1].d) second--;
    h[hole] = h[second];
    hole = second;
  }
  if ((L & 1) == 0 && second == (L - 2) / 2) {
    second = 2 * (second + 1);
    h[hole] = h[second - 1];
    hole = second - 1;
  }
  heap_sift_up(h, hole, 0, v);
}

// Row accessors: distance of (row gi, column j != gi) in column `col` of
// float2 (core, acc) data.
struct BandRows {  // square rows [r0, r0 + rows) x n, row stride n
  const float* band;
  int64_t r0, n;
  int col;
  __device__ float operator()(int64_t gi, int64_t j) const {
    return band[((gi - r0) * n + j) * 2 + col];
  }
};

struct TriangleRows {  // condensed upper triangle of n samples (symmetric distances)
  const float* tri;
  int64_t n;
  int col;
  __device__ float operator()(int64_t gi, int64_t j) const {
    const int64_t a = gi < j ? gi : j, b = gi < j ? j : gi;
    return tri[(n * a - ((a * (a + 1)) >> 1) + b - 1 - a) * 2 + col];
  }
};

// excl: column gi is the row's own sample (+inf, self kNN); otherwise the
// rows are queries against a separate reference set and nothing is excluded.
// Workgroup b handles row first_row + b (sample r0 + row); GHEAP: its heap is
// entries [b (kNN + 1), +kNN + 1) of gheap instead of LDS (kNN > KNN_LDS_MAX).
template <typename Rows, bool GHEAP>
__global__ __launch_bounds__(64) void knn_select_kernel(Rows rows, int64_t r0, int64_t first_row,
                                                        int64_t n, int kNN, bool excl,
                                                        int64_t* __restrict__ iv,
                                                        int64_t* __restrict__ jv,
                                                        float* __restrict__ dv,
                                                        HeapEnt* __restrict__ gheap) {
  extern __shared__ HeapEnt lds_heap[];
  HeapEnt* heap = GHEAP ? gheap + (size_t)blockIdx.x * (kNN + 1) : lds_heap;
  const int64_t row = first_row + blockIdx.x;
  const int64_t gi = r0 + row;
  const int lane = threadIdx.x;
  int len = 0;        // wave-uniform heap size
  float top = 0.0f;   // heap top once len == kNN
  // the next chunk's distances are loaded before the current chunk is
  // scanned, so the row streams at memory throughput rather than latency
  auto load = [&](int64_t c) {
    const int64_t jj = c + lane;
    return (jj < n && !(excl && jj == gi)) ? rows(gi, jj) : INFINITY;
  };
  float d_next = load(0);
  for (int64_t c0 = 0; c0 < n; c0 += 64) {
    const int64_t j = c0 + lane;
    const float d = d_next;
    d_next = load(c0 + 64);
    const bool cand = (j < n) && (len < kNN || d < top);
    unsigned long long mask = __ballot(cand);
    while (mask) {
      const int b = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const float db = __shfl(d, b);
      if (len < kNN || db < top) {
        float new_top = 0.0f;
        if (lane == 0) {
          heap_push(heap, len, HeapEnt{db, (int32_t)(c0 + b)});
          if (len + 1 > kNN) heap_pop(heap, len + 1);
          new_top = heap[0].d;
        }
        len = len < kNN ? len + 1 : kNN;
        new_top = __shfl(new_top, 0);
        if (len == kNN) top = new_top;
      }
    }
  }
  if (lane == 0) {
    const int64_t off = row * kNN;
    int l = len;
    for (int k = kNN - 1; k >= 0; --k) {
      iv[off + k] = gi;
      jv[off + k] = heap[0].j;
      dv[off + k] = heap[0].d;
      heap_pop(heap, l);
      --l;
    }
  }
}

// ---------------------------------------------------------------------------
// kNN of a few long query rows (serving): two passes, so that a step of a
// hundred rows still fills the GPU.
//
// 1. knn_cand_kernel, one wave per (row, column segment): the same scan,
//    started empty at the segment's first column, lists every element it
//    admits (column order).  An element the full scan admits is below the
//    k-th smallest of all earlier columns of its row (or the heap is not yet
//    full), and that is at most the k-th smallest of the earlier columns of
//    its segment, so the segment scan admits it too: the lists hold every
//    element the full scan admits.
// 2. knn_merge_kernel, one wave per row: the full scan's heap, replayed over
//    the lists of segments 0, 1, ... in column order.  Elements missing from
//    the lists are ones the full scan rejects without touching the heap, so
//    the heap goes through the same pushes and pops and ends the same.  A
//    segment whose list overflowed its capacity is scanned directly.
// ---------------------------------------------------------------------------

// Offers one chunk (lane t holds d, j; valid lanes only) to the heap in lane
// order, admitting as the scan does; lane 0 calls on_admit(d, j, base + x)
// for the x-th admitted element.  Returns the number admitted (wave-uniform).
template <typename F>
__device__ __forceinline__ int offer_chunk(HeapEnt* heap, int kNN, int& len, float& top,
                                           bool valid, float d, int32_t j, int lane, int base,
                                           F on_admit) {
  unsigned long long mask = __ballot(valid && (len < kNN || d < top));
  int x = 0;
  while (mask) {
    const int b = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const float db = __shfl(d, b);
    const int32_t jb = __shfl(j, b);
    if (len < kNN || db < top) {
      float new_top = 0.0f;
      if (lane == 0) {
        heap_push(heap, len, HeapEnt{db, jb});
        if (len + 1 > kNN) heap_pop(heap, len + 1);
        new_top = heap[0].d;
        on_admit(db, jb, base + x);
      }
      ++x;
      len = len < kNN ? len + 1 : kNN;
      new_top = __shfl(new_top, 0);
      if (len == kNN) top = new_top;
    }
  }
  return x;
}

// Scan of columns [lo, hi) of row `row` of a float2 band (rows of n), the next
// chunk loaded before the current one is offered.  Returns the number of
// admitted elements; on_admit as in offer_chunk, numbered from `base`.
template <typename F>
__device__ __forceinline__ int scan_columns(const float* band, int64_t row, int64_t n, int col,
                                            int64_t lo, int64_t hi, HeapEnt* heap, int kNN,
                                            int& len, float& top, int lane, int base,
                                            F on_admit) {
  auto load = [&](int64_t c) {
    const int64_t jj = c + lane;
    return jj < hi ? band[(row * n + jj) * 2 + col] : INFINITY;
  };
  int count = 0;
  float d_next = load(lo);
  for (int64_t c0 = lo; c0 < hi; c0 += 64) {
    const float d = d_next;
    d_next = load(c0 + 64);
    count += offer_chunk(heap, kNN, len, top, c0 + lane < hi, d, (int32_t)(c0 + lane), lane,
                         base + count, on_admit);
  }
  return count;
}

// Writes the kNN of row `row` from its heap (len == kNN), largest last, as the
// reference drains its priority_queue.
__device__ __forceinline__ void drain_heap(HeapEnt* heap, int kNN, int len, int64_t row,
                                           int64_t gi, int64_t* iv, int64_t* jv, float* dv) {
  const int64_t off = row * kNN;
  int l = len;
  for (int k = kNN - 1; k >= 0; --k) {
    iv[off + k] = gi;
    jv[off + k] = heap[0].j;
    dv[off + k] = heap[0].d;
    heap_pop(heap, l);
    --l;
  }
}

__global__ __launch_bounds__(64) void knn_cand_kernel(const float* __restrict__ band, int64_t n,
                                                      int col, int kNN, int segs, int cap,
                                                      HeapEnt* __restrict__ cand,
                                                      int32_t* __restrict__ cand_n) {
  extern __shared__ HeapEnt heap[];
  const int64_t b = blockIdx.x;
  const int64_t row = b / segs, g = b % segs;
  const int64_t lo = n * g / segs, hi = n * (g + 1) / segs;
  const int lane = threadIdx.x;
  HeapEnt* out = cand + b * cap;
  int len = 0;
  float top = 0.0f;
  const int count = scan_columns(band, row, n, col, lo, hi, heap, kNN, len, top, lane, 0,
                                 [&](float d, int32_t j, int x) {
                                   if (x < cap) out[x] = HeapEnt{d, j};
                                 });
  if (lane == 0) cand_n[b] = count;  // > cap: the list overflowed
}

__global__ __launch_bounds__(64) void knn_merge_kernel(const float* __restrict__ band, int64_t n,
                                                       int col, int kNN, int segs, int cap,
                                                       const HeapEnt* __restrict__ cand,
                                                       const int32_t* __restrict__ cand_n,
                                                       int64_t* __restrict__ iv,
                                                       int64_t* __restrict__ jv,
                                                       float* __restrict__ dv) {
  extern __shared__ HeapEnt heap[];
  const int64_t row = blockIdx.x;
  const int lane = threadIdx.x;
  int len = 0;
  float top = 0.0f;
  auto none = [](float, int32_t, int) {};
  for (int g = 0; g < segs; ++g) {
    const int64_t b = row * segs + g;
    const int cnt = cand_n[b];
    if (cnt <= cap) {
      const HeapEnt* in = cand + b * cap;
      for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int t = c0 + lane;
        const HeapEnt e = t < cnt ? in[t] : HeapEnt{INFINITY, 0};
        offer_chunk(heap, kNN, len, top, t < cnt, e.d, e.j, lane, 0, none);
      }
    } else {
      const int64_t lo = n * g / segs, hi = n * (g + 1) / segs;
      scan_columns(band, row, n, col, lo, hi, heap, kNN, len, top, lane, 0, none);
    }
  }
  if (lane == 0) drain_heap(heap, kNN, len, row, row, iv, jv, dv);
}

int knn_rect_capacity(int kNN) { return 16 * kNN + 64; }

size_t knn_heap_workspace(int64_t n_rows, int kNN) {
  if (kNN <= KNN_LDS_MAX) return 0;
  const int64_t rows = n_rows < KNN_HEAP_ROWS ? n_rows : KNN_HEAP_ROWS;
  return (size_t)rows * (kNN + 1) * sizeof(HeapEnt);
}

size_t knn_rect_workspace(int64_t n_rows, int segs, int kNN) {
  if (segs <= 1) return knn_heap_workspace(n_rows, kNN);
  const size_t lists = (size_t)n_rows * segs;
  return lists * knn_rect_capacity(kNN) * sizeof(HeapEnt) + lists * sizeof(int32_t);
}

// Rows [0, n_rows) of one selection, in launches of at most KNN_HEAP_ROWS
// rows when the heaps live in the workspace (each launch reuses it); row
// outputs at row * kNN.  The push writes slot kNN before the pop trims the
// heap back to kNN entries, hence kNN + 1 entries per heap.
template <typename Rows>
hipError_t knn_rows_launch(Rows rows, int64_t r0, int64_t n_rows, int64_t n, int kNN, bool excl,
                           int64_t* i_vec, int64_t* j_vec, float* dists, hipStream_t stream,
                           void* heap_ws) {
  if (kNN <= KNN_LDS_MAX) {
    knn_select_kernel<Rows, false><<<(unsigned)n_rows, 64, (kNN + 1) * sizeof(HeapEnt), stream>>>(
        rows, r0, 0, n, kNN, excl, i_vec, j_vec, dists, nullptr);
    return hipGetLastError();
  }
  if (!heap_ws) return hipErrorInvalidValue;
  for (int64_t f = 0; f < n_rows; f += KNN_HEAP_ROWS) {
    const int64_t m = n_rows - f < KNN_HEAP_ROWS ? n_rows - f : KNN_HEAP_ROWS;
    knn_select_kernel<Rows, true><<<(unsigned)m, 64, 0, stream>>>(
        rows, r0, f, n, kNN, excl, i_vec, j_vec, dists, (HeapEnt*)heap_ws);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_knn_select(const float* band, int64_t r0, int64_t n_rows, int64_t n,
                             int dist_col, int kNN, int64_t* i_vec, int64_t* j_vec,
                             float* dists, hipStream_t stream, void* heap_ws) {
  if (kNN < 1 || kNN > n) return hipErrorInvalidValue;
  if (n_rows <= 0) return hipSuccess;
  return knn_rows_launch(BandRows{band, r0, n, dist_col}, r0, n_rows, n, kNN, true, i_vec, j_vec,
                         dists, stream, heap_ws);
}

hipError_t launch_knn_select_rect(const float* band, int64_t n_rows, int64_t n, int dist_col,
                                  int kNN, int64_t* i_vec, int64_t* j_vec, float* dists,
                                  hipStream_t stream, int segs, void* workspace) {
  if (kNN < 1 || kNN > n) return hipErrorInvalidValue;
  if (n_rows <= 0) return hipSuccess;
  const size_t heap_bytes = (kNN + 1) * sizeof(HeapEnt);
  if (segs <= 1) {
    return knn_rows_launch(BandRows{band, 0, n, dist_col}, 0, n_rows, n, kNN, false, i_vec, j_vec,
                           dists, stream, workspace);
  }
  if (kNN > KNN_LDS_MAX || !workspace || n_rows * segs > INT32_MAX || segs > n) {
    return hipErrorInvalidValue;
  }
  const int cap = knn_rect_capacity(kNN);
  HeapEnt* cand = (HeapEnt*)workspace;
  int32_t* cand_n = (int32_t*)(cand + (size_t)n_rows * segs * cap);
  knn_cand_kernel<<<(unsigned)(n_rows * segs), 64, heap_bytes, stream>>>(band, n, dist_col, kNN,
                                                                          segs, cap, cand, cand_n);
  knn_merge_kernel<<<(unsigned)n_rows, 64, heap_bytes, stream>>>(band, n, dist_col, kNN, segs, cap,
                                                                  cand, cand_n, i_vec, j_vec,
                                                                  dists);
  return hipGetLastError();
}

hipError_t launch_knn_select_triangle(const float* tri, int64_t n, int dist_col, int kNN,
                                      int64_t* i_vec, int64_t* j_vec, float* dists,
                                      hipStream_t stream, void* heap_ws) {
  if (kNN < 1 || kNN > n) return hipErrorInvalidValue;
  if (n <= 0) return hipSuccess;
  return knn_rows_launch(TriangleRows{tri, n, dist_col}, 0, n, n, kNN, true, i_vec, j_vec, dists,
                         stream, heap_ws);
}

// ---------------------------------------------------------------------------
// Threshold compaction, row by row.  Triangle: the condensed self triangle,
// row i covering positions [n*i - i(i+1)/2, + n-1-i) with j = i+1+t.
// Rectangle: rows of n_cols entries (query-major band), j = t.
// ---------------------------------------------------------------------------
constexpr int TB = 256;

__device__ __forceinline__ float pick(const float2 v, int col) { return col ? v.y : v.x; }

struct RowSpan {
  int64_t start, len, j0;
};

__device__ __forceinline__ RowSpan row_span(bool rect, int64_t i, int64_t n) {
  if (rect) return RowSpan{i * n, n, 0};
  return RowSpan{n * i - ((i * (i + 1)) >> 1), n - 1 - i, i + 1};
}

// Block b covers segment b % segs of row b / segs: entries [lo, hi) of the
// row, lo = len * g / segs, hi = len * (g + 1) / segs; counts and offsets are
// per (row, segment) in row-major order, so the output order is row-major.
__global__ __launch_bounds__(TB) void thr_count_kernel(const float2* __restrict__ dd, int64_t n,
                                                       bool rect, int col, float cutoff,
                                                       int64_t* __restrict__ counts, int segs) {
  const int64_t b = blockIdx.x;
  const int64_t i = b / segs, g = b % segs;
  const RowSpan sp = row_span(rect, i, n);
  const int64_t lo = sp.len * g / segs, hi = sp.len * (g + 1) / segs;
  int64_t c = 0;
  for (int64_t t = lo + threadIdx.x; t < hi; t += TB) {
    c += pick(dd[sp.start + t], col) < cutoff ? 1 : 0;
  }
  __shared__ int64_t red[TB / 64];
  // wave reduction then across waves
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < TB / 64; ++w) s += red[w];
    counts[b] = s;
  }
}

// Exclusive scan of counts[0 .. rows) in place and the total in
// counts[rows]: one workgroup walks the rows in chunks of TB (a serving step
// has a few hundred rows).
__global__ __launch_bounds__(TB) void offsets_scan_kernel(int64_t* __restrict__ counts,
                                                          int64_t rows) {
  __shared__ int64_t wave_sum[TB / 64];
  __shared__ int64_t carry_sh;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_sh = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < rows; c0 += TB) {
    const int64_t t = c0 + threadIdx.x;
    const int64_t v = t < rows ? counts[t] : 0;
    int64_t x = v;  // inclusive scan inside the wave
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wave_sum[wave] = x;
    __syncthreads();
    int64_t before = carry_sh, chunk = 0;
    for (int w = 0; w < TB / 64; ++w) {
      if (w < wave) before += wave_sum[w];
      chunk += wave_sum[w];
    }
    if (t < rows) counts[t] = before + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry_sh += chunk;
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[rows] = carry_sh;
}

__global__ __launch_bounds__(TB) void thr_write_kernel(const float2* __restrict__ dd, int64_t n,
                                                       bool rect, int col, float cutoff,
                                                       const int64_t* __restrict__ offsets,
                                                       int64_t* __restrict__ iv,
                                                       int64_t* __restrict__ jv,
                                                       float* __restrict__ dv, int64_t cap,
                                                       int segs) {
  const int64_t b = blockIdx.x;
  const int64_t i = b / segs, g = b % segs;
  const RowSpan sp = row_span(rect, i, n);
  const int64_t lo = sp.len * g / segs, hi = sp.len * (g + 1) / segs;
  __shared__ int wave_tot[TB / 64];
  __shared__ int64_t base_sh;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) base_sh = offsets[b];
  __syncthreads();
  for (int64_t c0 = lo; c0 < hi; c0 += TB) {
    const int64_t t = c0 + threadIdx.x;
    float d = 0.0f;
    bool keep = false;
    if (t < hi) {
      d = pick(dd[sp.start + t], col);
      keep = d < cutoff;
    }
    const unsigned long long m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int wave_off = 0;
    int chunk_tot = 0;
    for (int w = 0; w < TB / 64; ++w) {
      if (w < wave) wave_off += wave_tot[w];
      chunk_tot += wave_tot[w];
    }
    const int64_t base = base_sh;
    const int64_t o = base + wave_off + before;
    if (keep && o < cap) {
      iv[o] = i;
      jv[o] = sp.j0 + t;
      dv[o] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) base_sh = base + chunk_tot;
    __syncthreads();
  }
}

hipError_t launch_threshold_count(const float2* dd, int64_t rows, int64_t n, bool rect,
                                  int dist_col, float cutoff, int64_t* counts,
                                  hipStream_t stream, int segs) {
  if (rows <= 0) return hipSuccess;
  if (segs < 1 || rows * segs > INT32_MAX) return hipErrorInvalidValue;
  thr_count_kernel<<<(unsigned)(rows * segs), TB, 0, stream>>>(dd, n, rect, dist_col, cutoff,
                                                               counts, segs);
  return hipGetLastError();
}

hipError_t launch_threshold_write(const float2* dd, int64_t rows, int64_t n, bool rect,
                                  int dist_col, float cutoff, const int64_t* offsets,
                                  int64_t* i_vec, int64_t* j_vec, float* dists,
                                  hipStream_t stream, int64_t capacity, int segs) {
  if (rows <= 0) return hipSuccess;
  if (segs < 1 || rows * segs > INT32_MAX) return hipErrorInvalidValue;
  thr_write_kernel<<<(unsigned)(rows * segs), TB, 0, stream>>>(dd, n, rect, dist_col, cutoff,
                                                               offsets, i_vec, j_vec, dists,
                                                               capacity, segs);
  return hipGetLastError();
}

hipError_t launch_offsets_scan(int64_t* counts, int64_t rows, hipStream_t stream) {
  if (rows < 0) return hipErrorInvalidValue;
  offsets_scan_kernel<<<1, TB, 0, stream>>>(counts, rows);
  return hipGetLastError();
}

hipError_t threshold_compact(const float2* dd, int64_t n, int dist_col, float cutoff,
                             int64_t* total, int64_t** d_i, int64_t** d_j, float** d_d,
                             hipStream_t stream) {
  *total = 0;
  *d_i = nullptr;
  *d_j = nullptr;
  *d_d = nullptr;
  const int64_t rows = n > 1 ? n - 1 : 0;
  if (rows == 0) return hipSuccess;
  int64_t* counts = nullptr;
  hipError_t e = hipMalloc(&counts, rows * sizeof(int64_t));
  if (e != hipSuccess) return e;
  e = launch_threshold_count(dd, rows, n, false, dist_col, cutoff, counts, stream);
  std::vector<int64_t> h(rows);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), counts, rows * sizeof(int64_t),
                                          hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    hipFree(counts);
    return e;
  }
  int64_t run = 0;
  for (int64_t r = 0; r < rows; ++r) {
    const int64_t c = h[r];
    h[r] = run;
    run += c;
  }
  *total = run;
  e = hipMemcpyAsync(counts, h.data(), rows * sizeof(int64_t), hipMemcpyHostToDevice, stream);
  const size_t m = (size_t)(run > 0 ? run : 1);
  if (e == hipSuccess) e = hipMalloc(d_i, m * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc(d_j, m * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc(d_d, m * sizeof(float));
  if (e == hipSuccess && run > 0) {
    e = launch_threshold_write(dd, rows, n, false, dist_col, cutoff, counts, *d_i, *d_j, *d_d,
                               stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  hipFree(counts);
  return e;
}

}  // namespace ppsk
