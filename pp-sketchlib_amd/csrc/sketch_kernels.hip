// Sketching on the GPU: sequences -> bindash sketches in the reference's
// bit-sliced word layout (sketch/sketch.cpp:118-176).
//
// Hashing (sketch/nthash.hpp:17-127): canonical ntHash in its
// split-rotation form.  rot(x) rotates the high 31 bits and the low 33 bits
// of x left by one each, so for a k-mer s_0 .. s_{k-1}
//   fh = XOR_i rot^(k-1-i)(seed(s_i)),  rh = XOR_i rot^i(seed(comp(s_i))),
// and sliding one base right is
//   fh' = rot(fh) ^ seed(in) ^ rot^k(seed(out)),
//   rh' = rot^-1(rh ^ seed(comp(out)) ^ rot^k(seed(comp(in)))).
// Characters other than A, C, G, T, U (either case) have seed 0, so the
// rolling values stay exact across them; a k-mer is hashed only if its
// window holds none of them.  Starting from fh = rh = 0 and rolling in the
// first k bases with no outgoing base gives the first window's values, so
// one loop serves the whole run.
//
// Work: a decode pass turns bytes into codes 0-4 (A, C, G, T/U, other).  A
// unit is a run of up to UNIT k-mer start positions of one contig; one lane
// rolls through a unit, reading its incoming and outgoing bases as two
// streams of 16-byte loads issued a block ahead.  The two XOR terms of a
// step depend only on the (incoming, outgoing) code pair: a 25-entry table
// in LDS per workgroup.  A workgroup takes a run of units of one sample for
// one k-mer length and keeps that sample's bin minima in LDS (ds_min_u64),
// merged into the global minima with atomicMin at the end.  A second kernel
// densifies (in the reference's sequential order) and bit-slices each
// (sample, k) with one wave: layer word w, layer i is the wave ballot of bit
// i of bins 64 w .. 64 w + 63.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "sketch_kernels.h"

namespace ppsk {

namespace {

constexpr uint64_t SIGN_MOD = (1ULL << 61) - 1ULL;
constexpr uint64_t SEED_A = 0x3c8bfbb395c60474ULL;
constexpr uint64_t SEED_C = 0x3193c18562a02b4cULL;
constexpr uint64_t SEED_G = 0x20323ed082572324ULL;
constexpr uint64_t SEED_T = 0x295549f54be24456ULL;

// 3-bit code of the letters a .. u (index = letter - 'a'): a 0, c 1, g 2,
// t 3, u 3, every other letter 4.
constexpr uint64_t letter_codes() {
  uint64_t t = 0;
  for (int x = 0; x < 21; ++x) {
    const uint64_t c = x == 0 ? 0 : x == 2 ? 1 : x == 6 ? 2 : (x == 19 || x == 20) ? 3 : 4;
    t |= c << (3 * x);
  }
  return t;
}
constexpr uint64_t LETTER_CODES = letter_codes();

// code of one byte: 0 A, 1 C, 2 G, 3 T/U (either case), 4 anything else
__device__ __forceinline__ uint32_t code_of(uint32_t c) {
  const uint32_t x = (c | 0x20u) - 0x61u;
  return x < 21u ? (uint32_t)((LETTER_CODES >> (3u * x)) & 7u) : 4u;
}

__device__ __forceinline__ uint32_t code4(uint32_t w) {
  return code_of(w & 0xFFu) | (code_of((w >> 8) & 0xFFu) << 8) |
         (code_of((w >> 16) & 0xFFu) << 16) | (code_of(w >> 24) << 24);
}

__device__ __forceinline__ uint64_t seed_of(int code) {
  return code == 0 ? SEED_A : code == 1 ? SEED_C : code == 2 ? SEED_G : code == 3 ? SEED_T : 0;
}

__device__ __forceinline__ uint64_t rot1(uint64_t x) {
  const uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  return ((((hi << 1) | (hi >> 30)) & 0x7FFFFFFFULL) << 33) | (((lo << 1) | (lo >> 32)) & 0x1FFFFFFFFULL);
}

__device__ __forceinline__ uint64_t rot1_inv(uint64_t x) {
  const uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  return ((((hi >> 1) | (hi << 30)) & 0x7FFFFFFFULL) << 33) | (((lo >> 1) | (lo << 32)) & 0x1FFFFFFFFULL);
}

__device__ __forceinline__ uint64_t rotj(uint64_t x, unsigned j) {
  uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  const unsigned a = j % 31, b = j % 33;
  if (a) hi = ((hi << a) | (hi >> (31 - a))) & 0x7FFFFFFFULL;
  if (b) lo = ((lo << b) | (lo >> (33 - b))) & 0x1FFFFFFFFULL;
  return (hi << 33) | lo;
}

// h mod (2^61 - 1)
__device__ __forceinline__ uint64_t mod_sign(uint64_t h) {
  uint64_t r = (h & SIGN_MOD) + (h >> 61);
  r = r >= SIGN_MOD ? r - SIGN_MOD : r;
  return r;
}

__global__ void fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
       x += (int64_t)gridDim.x * blockDim.x) {
    p[x] = v;
  }
}

// XOR terms of one rolling step for an (incoming, outgoing) code pair.
struct StepTerms {
  uint64_t f, r;
};

}  // namespace

// In place: bytes -> codes over [0, n), code 4 in the pads.
__global__ void sketch_decode_kernel(uint8_t* codes, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n16 = n / 16;
  for (int64_t x = t0; x < n16; x += stride) {
    uint4 v = reinterpret_cast<uint4*>(codes)[x];
    v.x = code4(v.x);
    v.y = code4(v.y);
    v.z = code4(v.z);
    v.w = code4(v.w);
    reinterpret_cast<uint4*>(codes)[x] = v;
  }
  for (int64_t x = n16 * 16 + t0; x < n; x += stride) codes[x] = (uint8_t)code_of(codes[x]);
  for (int64_t x = t0; x < CODE_PAD_FRONT; x += stride) codes[x - CODE_PAD_FRONT] = 4;
  for (int64_t x = t0; x < CODE_PAD_BACK; x += stride) codes[n + x] = 4;
}

template <bool LDS_BINS>
__global__ __launch_bounds__(SKETCH_TB) void sketch_hash_kernel(SketchArgs a, int64_t n_groups,
                                                                int* next_group) {
  extern __shared__ uint64_t lbins[];
  __shared__ StepTerms terms[25];
  __shared__ int64_t cur_group;
  __shared__ int cur_k;
  const int64_t nbins = (int64_t)a.s64 * 64;
  // exact division by binsize: quotient from the reciprocal (< 2^32), one correction
  const uint64_t binsize = a.binsize;
  const uint32_t recip = (uint32_t)a.bin_recip;
  const bool use_rc = a.use_rc != 0;
  if (threadIdx.x == 0) cur_k = -1;
  // groups are taken from a counter until none is left: every workgroup
  // leaves the loop once the counter passes n_groups
  for (;;) {
    __syncthreads();  // the previous group's flush and reads of cur_group are done
    if (threadIdx.x == 0) cur_group = atomicAdd(next_group, 1);
    __syncthreads();
    const int64_t gi = cur_group;
    if (gi >= n_groups) break;
    const SketchGroup g = a.groups[gi];
    const int kk = g.kk;
    const int k = a.klist[kk];
    uint64_t* gbins = a.signs + ((int64_t)g.sample * a.n_k + kk) * nbins;
    if (k != cur_k) {
      if (threadIdx.x < 25) {
        const int cin = threadIdx.x / 5, cout = threadIdx.x % 5;
        StepTerms t;
        t.f = (cin < 4 ? seed_of(cin) : 0ULL) ^ (cout < 4 ? rotj(seed_of(cout), k) : 0ULL);
        t.r = (cout < 4 ? seed_of(3 - cout) : 0ULL) ^ (cin < 4 ? rotj(seed_of(3 - cin), k) : 0ULL);
        terms[threadIdx.x] = t;
      }
    }
    if (LDS_BINS) {
      for (int64_t b = threadIdx.x; b < nbins; b += SKETCH_TB) lbins[b] = ~0ULL;
    }
    __syncthreads();
    if (threadIdx.x == 0) cur_k = k;
    for (int64_t u = g.unit0 + threadIdx.x; u < g.unit1; u += SKETCH_TB) {
      const SketchUnit su = a.units[u];
      // step i rolls in base i of the unit and, from i = k on, rolls out base
      // i - k; step i ends the window starting at i - k + 1
      const int steps = su.n + k - 1;
      const uintptr_t ia = (uintptr_t)(a.codes + su.pos), oa = ia - (uintptr_t)k;
      const uint32_t* ip = reinterpret_cast<const uint32_t*>(ia & ~(uintptr_t)3);
      const uint32_t* op = reinterpret_cast<const uint32_t*>(oa & ~(uintptr_t)3);
      const uint32_t ish = (uint32_t)(ia & 3), osh = (uint32_t)(oa & 3);
      uint32_t ic[4], in_[4], oc[4], on_[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ic[q] = ip[q];
        in_[q] = ip[4 + q];
        oc[q] = op[q];
        on_[q] = op[4 + q];
      }
      ip += 8;
      op += 8;
      uint64_t fh = 0, rh = 0;
      int last_bad = -1;
      for (int i0 = 0; i0 < steps; i0 += 16) {
        uint32_t inn[4], onn[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          inn[q] = ip[q];
          onn[q] = op[q];
        }
        ip += 4;
        op += 4;
        uint32_t wi[4], wo[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          wi[q] = __builtin_amdgcn_alignbyte(q < 3 ? ic[q + 1] : in_[0], ic[q], ish);
          wo[q] = __builtin_amdgcn_alignbyte(q < 3 ? oc[q + 1] : on_[0], oc[q], osh);
        }
        // each half block's step terms are read before any of its bin updates
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          StepTerms t[8];
          uint32_t bad = 0;
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) {
            const int j = half * 8 + jj;
            const uint32_t cin = (wi[j >> 2] >> (8 * (j & 3))) & 0xFFu;
            uint32_t cout = (wo[j >> 2] >> (8 * (j & 3))) & 0xFFu;
            cout = i0 + j >= k ? cout : 4u;
            t[jj] = terms[cin * 5 + cout];
            bad |= (cin == 4u ? 1u : 0u) << jj;
          }
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) {
            const int i = i0 + half * 8 + jj;
            fh = rot1(fh) ^ t[jj].f;
            rh = rot1_inv(rh ^ t[jj].r);
            last_bad = (bad >> jj) & 1u ? i : last_bad;
            if (i - k >= last_bad && i < steps) {
              const uint64_t h = mod_sign((use_rc && rh < fh) ? rh : fh);
              const uint64_t lo = (uint64_t)(uint32_t)h * recip;
              uint64_t bin = ((h >> 32) * recip + (lo >> 32)) >> 32;
              if (h - bin * binsize >= binsize) ++bin;
              if (LDS_BINS) {
                atomicMin((unsigned long long*)&lbins[bin], (unsigned long long)h);
              } else if (h < gbins[bin]) {
                atomicMin((unsigned long long*)&gbins[bin], (unsigned long long)h);
              }
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ic[q] = in_[q];
          in_[q] = inn[q];
          oc[q] = on_[q];
          on_[q] = onn[q];
        }
      }
    }
    if (LDS_BINS) {
      __syncthreads();
      for (int64_t b = threadIdx.x; b < nbins; b += SKETCH_TB) {
        const uint64_t v = lbins[b];
        if (v != ~0ULL && v < gbins[b]) atomicMin((unsigned long long*)&gbins[b], (unsigned long long)v);
      }
    }
  }
}

// Reference densifybin (sketch/sketch.cpp:93-115) and fillusigs (80-91) for
// one (sample, k) per 64-lane workgroup.
__global__ __launch_bounds__(64) void sketch_finish_kernel(SketchArgs a) {
  const int64_t sk = blockIdx.x;  // sample * n_k + kk
  const int64_t nbins = (int64_t)a.s64 * 64;
  uint64_t* sg = a.signs + sk * nbins;
  const int lane = threadIdx.x;
  uint64_t mn = ~0ULL, mx = 0;
  for (int64_t b = lane; b < nbins; b += 64) {
    const uint64_t v = sg[b];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t omn = __shfl_xor(mn, off), omx = __shfl_xor(mx, off);
    mn = omn < mn ? omn : mn;
    mx = omx > mx ? omx : mx;
  }
  if (lane == 0) a.inv_minhash[sk] = (double)sg[0] / (double)SIGN_MOD;
  int dens = 0;
  if (mx == ~0ULL) {
    dens = mn == ~0ULL ? -1 : 1;
    if (dens == 1 && lane == 0) {
      // sequential, in place: a probe may land on a bin filled earlier
      for (int64_t i = 0; i < nbins; ++i) {
        int64_t j = i;
        uint64_t attempts = 0;
        while (sg[j] == ~0ULL) {
          const uint64_t x = 1009ULL * (uint64_t)i + (1000ULL * 1000ULL + 3ULL) * attempts;
          j = (int64_t)(((48271ULL * x + 11ULL) % ((1ULL << 31) - 1ULL)) % (uint64_t)nbins);
          attempts++;
        }
        sg[i] = sg[j];
      }
    }
    __syncthreads();
  }
  if (lane == 0) a.densified[sk] = dens;
  uint64_t* out = a.raw + sk * (int64_t)a.s64 * BBITS_SK;
  for (int w = 0; w < a.s64; ++w) {
    const uint64_t v = sg[(int64_t)w * 64 + lane];
    uint64_t word_i = 0;
#pragma unroll
    for (int i = 0; i < BBITS_SK; ++i) {
      const uint64_t m = __ballot((v >> i) & 1ULL);
      if (lane == i) word_i = m;
    }
    if (lane < BBITS_SK) out[(int64_t)w * BBITS_SK + lane] = word_i;
  }
}

hipError_t launch_sketch_decode(uint8_t* codes, int64_t n, hipStream_t stream) {
  const int64_t work = (n + 15) / 16;
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 16384));
  sketch_decode_kernel<<<(unsigned)blocks, 256, 0, stream>>>(codes, n);
  return hipGetLastError();
}

hipError_t launch_sketch(const SketchArgs& a, int64_t n_groups, bool lds_bins, int* next_group,
                         hipStream_t stream) {
  const int64_t nbins = (int64_t)a.s64 * 64;
  const int64_t total = (int64_t)a.n_samples * a.n_k * nbins;
  fill_u64<<<(unsigned)((total + 255) / 256 > 65536 ? 65536 : (total + 255) / 256), 256, 0,
             stream>>>(a.signs, total, ~0ULL);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_groups > 0) {
    e = hipMemsetAsync(next_group, 0, sizeof(int), stream);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t lds = lds_bins ? (size_t)nbins * sizeof(uint64_t) : 0;
    if (e == hipSuccess) {
      e = lds_bins ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sketch_hash_kernel<true>,
                                                                  SKETCH_TB, lds)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, sketch_hash_kernel<false>, SKETCH_TB, 0);
    }
    if (e != hipSuccess) return e;
    const int64_t grid = std::max<int64_t>(
        1, std::min<int64_t>(n_groups, (int64_t)std::max(per_cu, 1) * std::max(cus, 1)));
    if (lds_bins) {
      sketch_hash_kernel<true><<<(unsigned)grid, SKETCH_TB, lds, stream>>>(a, n_groups, next_group);
    } else {
      sketch_hash_kernel<false><<<(unsigned)grid, SKETCH_TB, 0, stream>>>(a, n_groups, next_group);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  sketch_finish_kernel<<<(unsigned)(a.n_samples * a.n_k), 64, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace ppsk
