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
// window holds none of them.
//
// Work: a unit is a run of up to UNIT k-mer start positions of one contig;
// one thread rolls through a unit.  A workgroup takes a run of units of one
// sample for one k-mer length and keeps that sample's bin minima in LDS
// (ds_min_u64), merged into the global minima with atomicMin at the end.
// A second kernel densifies (in the reference's sequential order) and
// bit-slices each (sample, k) with one wave: layer word w, layer i is the
// wave ballot of bit i of bins 64 w .. 64 w + 63.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sketch_kernels.h"

namespace ppsk {

namespace {

constexpr uint64_t SIGN_MOD = (1ULL << 61) - 1ULL;
constexpr uint64_t SEED_A = 0x3c8bfbb395c60474ULL;
constexpr uint64_t SEED_C = 0x3193c18562a02b4cULL;
constexpr uint64_t SEED_G = 0x20323ed082572324ULL;
constexpr uint64_t SEED_T = 0x295549f54be24456ULL;

// base code: 0 A, 1 C, 2 G, 3 T/U, 4 other
__device__ __forceinline__ int base_code(unsigned char c) {
  switch (c | 0x20) {  // lower case
    case 'a':
      return 0;
    case 'c':
      return 1;
    case 'g':
      return 2;
    case 't':
    case 'u':
      return 3;
    default:
      return 4;
  }
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

}  // namespace

template <bool LDS_BINS>
__global__ __launch_bounds__(SKETCH_TB) void sketch_hash_kernel(SketchArgs a) {
  extern __shared__ uint64_t lbins[];
  const SketchGroup g = a.groups[blockIdx.x];
  const int kk = g.kk;
  const unsigned k = (unsigned)a.klist[kk];
  const int64_t nbins = (int64_t)a.s64 * 64;
  uint64_t* gbins = a.signs + ((int64_t)g.sample * a.n_k + kk) * nbins;
  if (LDS_BINS) {
    for (int64_t b = threadIdx.x; b < nbins; b += SKETCH_TB) lbins[b] = ~0ULL;
    __syncthreads();
  }
  // exact division by binsize: quotient from the reciprocal, one correction
  const uint64_t binsize = a.binsize;
  const uint64_t recip = a.bin_recip;
  // rot^k of the four seeds and of their complements (out-going terms)
  uint64_t fk[4], rk[4];
  for (int c = 0; c < 4; ++c) {
    fk[c] = rotj(seed_of(c), k);
    rk[c] = rotj(seed_of(3 - c), k);
  }
  for (int64_t u = g.unit0 + threadIdx.x; u < g.unit1; u += SKETCH_TB) {
    const SketchUnit su = a.units[u];
    const unsigned char* s = reinterpret_cast<const unsigned char*>(a.seq) + su.pos;
    // first window
    uint64_t fh = 0, rh = 0;
    int last_bad = -1;  // window index of the last non-ACGTU base
    for (unsigned i = 0; i < k; ++i) {
      const int c = base_code(s[i]);
      if (c == 4) last_bad = (int)i;
      fh = rot1(fh) ^ (c < 4 ? seed_of(c) : 0ULL);
      rh ^= rotj(c < 4 ? seed_of(3 - c) : 0ULL, i);
    }
    for (int p = 0;; ++p) {
      if (last_bad < p) {
        const uint64_t h = mod_sign((a.use_rc && rh < fh) ? rh : fh);
        uint64_t bin = __umul64hi(h, recip);
        if (h - bin * binsize >= binsize) ++bin;
        if (LDS_BINS) {
          atomicMin((unsigned long long*)&lbins[bin], (unsigned long long)h);
        } else if (h < gbins[bin]) {
          atomicMin((unsigned long long*)&gbins[bin], (unsigned long long)h);
        }
      }
      if (p + 1 >= su.n) break;
      const int cout = base_code(s[p]);
      const int cin = base_code(s[p + k]);
      if (cin == 4) last_bad = p + 1 + (int)k - 1;
      fh = rot1(fh) ^ (cin < 4 ? seed_of(cin) : 0ULL) ^ (cout < 4 ? fk[cout] : 0ULL);
      rh = rot1_inv(rh ^ (cout < 4 ? seed_of(3 - cout) : 0ULL) ^ (cin < 4 ? rk[cin] : 0ULL));
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

hipError_t launch_sketch(const SketchArgs& a, int64_t n_groups, bool lds_bins,
                         hipStream_t stream) {
  const int64_t nbins = (int64_t)a.s64 * 64;
  const int64_t total = (int64_t)a.n_samples * a.n_k * nbins;
  fill_u64<<<(unsigned)((total + 255) / 256 > 65536 ? 65536 : (total + 255) / 256), 256, 0,
             stream>>>(a.signs, total, ~0ULL);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_groups > 0) {
    if (lds_bins) {
      sketch_hash_kernel<true><<<(unsigned)n_groups, SKETCH_TB, nbins * sizeof(uint64_t), stream>>>(a);
    } else {
      sketch_hash_kernel<false><<<(unsigned)n_groups, SKETCH_TB, 0, stream>>>(a);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  sketch_finish_kernel<<<(unsigned)(a.n_samples * a.n_k), 64, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace ppsk
