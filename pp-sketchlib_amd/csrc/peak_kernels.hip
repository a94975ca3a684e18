// Measurement kernels (not part of the distance path): the sustained int32
// VALU issue rate of v_bitop3_b32 on this device, used as the roofline peak
// next to the vendor figure, and a plain copy kernel whose HBM counter
// readings calibrate FETCH_SIZE / WRITE_SIZE.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int CHAINS = 16;   // independent accumulators per lane
constexpr int UNROLL = 32;   // bitop3 per chain per loop iteration

__global__ __launch_bounds__(256) void bitop3_peak(uint32_t* out, uint32_t a, uint32_t b,
                                                   int iters) {
  uint32_t acc[CHAINS];
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) acc[c] = threadIdx.x * 2654435761u + c;
  const uint32_t x = a ^ threadIdx.x, y = b + blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) acc[c] = __builtin_amdgcn_bitop3_b32(acc[c], x, y, 0x96);
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) r ^= acc[c];
  if (r == 0x12345678u) out[blockIdx.x * 256 + threadIdx.x] = r;  // keep the chains live
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    dst[i] = src[i];
  }
}

}  // namespace

extern "C" {

// Runs the bitop3 chain kernel; returns int32 VALU ops per second.
int ppsk_bench_valu_peak(int device, int blocks, int iters, double* ops_per_s, double* ms) {
  if (hipSetDevice(device) != hipSuccess) return 1;
  uint32_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  bitop3_peak<<<blocks, 256>>>(out, 0x1234u, 0x9876u, 4);  // warm-up
  hipEventRecord(e0);
  bitop3_peak<<<blocks, 256>>>(out, 0x1234u, 0x9876u, iters);
  hipEventRecord(e1);
  hipEventSynchronize(e1);
  float t = 0;
  hipEventElapsedTime(&t, e0, e1);
  *ms = t;
  *ops_per_s = (double)blocks * 256 * (double)iters * UNROLL * CHAINS / (t * 1e-3);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Copies `bytes` (multiple of 16) device to device; returns the kernel time.
int ppsk_bench_copy(int device, int64_t bytes, int reps, double* ms) {
  if (hipSetDevice(device) != hipSuccess) return 1;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return 1;
  if (hipMalloc(&b, bytes) != hipSuccess) return 1;
  hipMemset(a, 1, bytes);
  const int64_t n = bytes / 16;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  copy_kernel<<<4096, 256>>>((const uint4*)a, (uint4*)b, n);
  hipEventRecord(e0);
  for (int r = 0; r < reps; ++r) copy_kernel<<<4096, 256>>>((const uint4*)a, (uint4*)b, n);
  hipEventRecord(e1);
  hipEventSynchronize(e1);
  float t = 0;
  hipEventElapsedTime(&t, e0, e1);
  *ms = t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
