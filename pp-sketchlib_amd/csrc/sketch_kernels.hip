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

#include <hipcub/hipcub.hpp>

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

// rot^S and rot^-S (S < 31)
template <int S>
__device__ __forceinline__ uint64_t rot_by(uint64_t x) {
  const uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  return ((((hi << S) | (hi >> (31 - S))) & 0x7FFFFFFFULL) << 33) |
         (((lo << S) | (lo >> (33 - S))) & 0x1FFFFFFFFULL);
}

template <int S>
__device__ __forceinline__ uint64_t rot_inv_by(uint64_t x) {
  const uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  return ((((hi >> S) | (hi << (31 - S))) & 0x7FFFFFFFULL) << 33) |
         (((lo >> S) | (lo << (33 - S))) & 0x1FFFFFFFFULL);
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

// The 25 step terms for an outgoing-base lag: entry cin * 5 + cout (codes 4
// contribute 0).  The lag is k for dense seeds and 3k for codon-phased ones.
__device__ __forceinline__ StepTerms step_terms(int e, int lag) {
  const int cin = e / 5, cout = e % 5;
  StepTerms t;
  t.f = (cin < 4 ? seed_of(cin) : 0ULL) ^ (cout < 4 ? rotj(seed_of(cout), lag) : 0ULL);
  t.r = (cout < 4 ? seed_of(3 - cout) : 0ULL) ^ (cin < 4 ? rotj(seed_of(3 - cin), lag) : 0ULL);
  return t;
}

// Seed shape: dense (STRIDE 1, span k) or codon-phased (STRIDE 3: bases at
// offsets 0, 3, ..., 3(k-1) of a span of 3k - 2, sketch/sketch.cpp:43-54).
// The codon-phased hash is the dense recurrence over every third base, so
// three interleaved states roll with rot^3, each seeing every third step,
// and the outgoing base of a state lies 3k bases back.  A window is valid
// when its whole span is free of code-4 bases.
template <int STRIDE>
__device__ __forceinline__ int seed_span(int k) {
  return STRIDE == 1 ? k : 3 * k - 2;
}

// One lane rolls through one unit: emit(p, h) for every window p (0-based in
// the unit, ascending) whose span is free of code-4 bases, h its ntHash (the
// smaller of the two strands with use_rc).  Step i rolls in base i of the
// unit and, from i = lag on, rolls out base i - lag; step i ends the window
// starting at i - span + 1.  Bases come as two streams of 16-byte loads
// issued a block ahead; each half block's step terms are read from LDS
// before its emits.
template <int STRIDE, class Emit>
__device__ __forceinline__ void roll_unit(const uint8_t* codes, const SketchUnit su, int k,
                                          bool use_rc, const StepTerms* terms, Emit&& emit) {
  const int span = seed_span<STRIDE>(k), lag = STRIDE * k;
  const int steps = su.n + span - 1;
  const uintptr_t ia = (uintptr_t)(codes + su.pos), oa = ia - (uintptr_t)lag;
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
  // fh[s], rh[s]: the state that step j of the current block updates when
  // j % STRIDE == s
  uint64_t fh[STRIDE], rh[STRIDE];
#pragma unroll
  for (int q = 0; q < STRIDE; ++q) fh[q] = rh[q] = 0;
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
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      StepTerms t[8];
      uint32_t bad = 0;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = half * 8 + jj;
        const uint32_t cin = (wi[j >> 2] >> (8 * (j & 3))) & 0xFFu;
        uint32_t cout = (wo[j >> 2] >> (8 * (j & 3))) & 0xFFu;
        cout = i0 + j >= lag ? cout : 4u;
        t[jj] = terms[cin * 5 + cout];
        bad |= (cin == 4u ? 1u : 0u) << jj;
      }
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = half * 8 + jj, q = j % STRIDE;
        const int i = i0 + j;
        fh[q] = rot_by<STRIDE>(fh[q]) ^ t[jj].f;
        rh[q] = rot_inv_by<STRIDE>(rh[q] ^ t[jj].r);
        last_bad = (bad >> jj) & 1u ? i : last_bad;
        if (i - span >= last_bad && i < steps) {
          emit(i - span + 1, (use_rc && rh[q] < fh[q]) ? rh[q] : fh[q]);
        }
      }
    }
    if (STRIDE == 3) {
      // the next block starts 16 steps on: its step 0 updates this block's state 1
      const uint64_t f0 = fh[0], r0 = rh[0];
      fh[0] = fh[1 % STRIDE];
      rh[0] = rh[1 % STRIDE];
      fh[1 % STRIDE] = fh[2 % STRIDE];
      rh[1 % STRIDE] = rh[2 % STRIDE];
      fh[2 % STRIDE] = f0;
      rh[2 % STRIDE] = r0;
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

// bin of h (< 2^61 - 1): quotient from the reciprocal (< 2^32), one correction
__device__ __forceinline__ uint64_t bin_of(uint64_t h, uint64_t binsize, uint32_t recip) {
  const uint64_t lo = (uint64_t)(uint32_t)h * recip;
  uint64_t bin = ((h >> 32) * recip + (lo >> 32)) >> 32;
  if (h - bin * binsize >= binsize) ++bin;
  return bin;
}

template <bool LDS_BINS, int STRIDE>
__global__ __launch_bounds__(SKETCH_TB) void sketch_hash_kernel(SketchArgs a, int64_t n_groups,
                                                                int* next_group) {
  extern __shared__ uint64_t lbins[];
  __shared__ StepTerms terms[25];
  __shared__ int64_t cur_group;
  __shared__ int cur_k;
  const int64_t nbins = (int64_t)a.s64 * 64;
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
    if (k != cur_k && threadIdx.x < 25) terms[threadIdx.x] = step_terms(threadIdx.x, STRIDE * k);
    if (LDS_BINS) {
      for (int64_t b = threadIdx.x; b < nbins; b += SKETCH_TB) lbins[b] = ~0ULL;
    }
    __syncthreads();
    if (threadIdx.x == 0) cur_k = k;
    for (int64_t u = g.unit0 + threadIdx.x; u < g.unit1; u += SKETCH_TB) {
      roll_unit<STRIDE>(a.codes, a.units[u], k, use_rc, terms, [&](int, uint64_t h0) {
        const uint64_t h = mod_sign(h0);
        const uint64_t bin = bin_of(h, binsize, recip);
        if (LDS_BINS) {
          atomicMin((unsigned long long*)&lbins[bin], (unsigned long long)h);
        } else if (h < gbins[bin]) {
          atomicMin((unsigned long long*)&gbins[bin], (unsigned long long)h);
        }
      });
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

// ---------------------------------------------------------------------------
// Reads count filter (sketch/countmin.cpp:37-85, sketch/sketch.cpp:128-149):
// every valid k-mer occurrence is counted, in sequence order, before it may
// be binned, and is binned only if its count after that occurrence reaches
// min_count.  Order matters only through the counts, so one (sample, k) job
// runs as: hash every window (h0, validity) -> occurrence order by prefix
// sum -> per-occurrence keys -> stable radix sort by key -> an occurrence's
// count is its rank among equal keys + 1, so "count >= m" is "the key m - 1
// places back in sorted order is the same" -> bin the passing occurrences.
//  - exact (HashCounter): key = h0.
//  - count-min: 4 keys per occurrence, the flat cells of the reference's
//    4 x (2^27 - 1) table (row r, 27 bits of hash r / 2; an index of
//    2^27 - 1 lands on the next row's cell 0, as in the reference's nested
//    std::array); the estimate is the minimum over the rows, so an
//    occurrence passes when all four of its events pass.  Hash 1 is the mix
//    of hash 0 of the first k-mer of the current run of valid k-mers: the
//    reference's hash iterator computes extra hashes only when it restarts
//    (sketch/stHashIterator.hpp:64-112).
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t CM_MASK = 0x7FFFFFFULL;
constexpr uint64_t MULTI_SEED = 0x90b45d39fb6da1faULL;

__device__ __forceinline__ uint64_t multi_hash1(uint64_t h0, int k) {
  const uint64_t t = h0 * (1ULL ^ (uint64_t)k * MULTI_SEED);
  return t ^ (t >> 27);
}

template <int STRIDE>
__global__ __launch_bounds__(256) void reads_hash_kernel(ReadsJob j, uint64_t* h0, uint8_t* st) {
  __shared__ StepTerms terms[25];
  if (threadIdx.x < 25) terms[threadIdx.x] = step_terms(threadIdx.x, STRIDE * j.k);
  __syncthreads();
  const bool use_rc = j.use_rc != 0;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < j.n_units;
       u += (int64_t)gridDim.x * blockDim.x) {
    const SketchUnit su = j.units[u];
    const int64_t w0 = j.wbase[u];
    roll_unit<STRIDE>(j.codes, su, j.k, use_rc, terms, [&](int p, uint64_t h) {
      h0[w0 + p] = h;
      st[w0 + p] = (p == 0 && su.pad) ? 3 : 1;
    });
  }
}

// rs[w] = w if a run of valid windows starts at w, else -1
__global__ void reads_run_kernel(const uint8_t* st, int64_t n, int64_t* rs) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n;
       w += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t s = st[w];
    rs[w] = (s != 0 && (s == 3 || w == 0 || st[w - 1] == 0)) ? w : -1;
  }
}

// occurrence t = tix[w] of every valid window: H0[t], H1[t].  Hash 1 mixes
// hash 0 of the run's first window (dense seeds) or of the window itself
// (codon-phased seeds: the spaced-seed iterator recomputes it every step);
// span is the iterator's k, the seed's span.
__global__ void reads_compact_kernel(const uint8_t* st, const uint64_t* h0, const int64_t* run,
                                     const int64_t* tix, int64_t n, int span, int own_hash,
                                     uint64_t* H0, uint64_t* H1) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n;
       w += (int64_t)gridDim.x * blockDim.x) {
    if (st[w] != 0) {
      const int64_t t = tix[w];
      H0[t] = h0[w];
      H1[t] = multi_hash1(h0[own_hash ? w : run[w]], span);
    }
  }
}

__global__ void reads_cm_keys_kernel(const uint64_t* H0, const uint64_t* H1, int64_t n,
                                     uint32_t* key, uint32_t* val) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < 4 * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e >> 2;
    const uint32_t r = (uint32_t)(e & 3);
    const uint64_t hv = r < 2 ? H0[t] : H1[t];
    key[e] = (uint32_t)(r * CM_MASK + ((hv >> (27 * (r & 1))) & CM_MASK));
    val[e] = (uint32_t)e;
  }
}

__global__ void reads_iota_kernel(uint32_t* val, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    val[e] = (uint32_t)e;
  }
}

// after the stable sort: the count of element s's key up to and including
// it is (rank among equal keys) + 1; it reaches m iff the key m - 1 places
// back is the same.  pass starts all 1 and only the (few) failing events are
// scattered back to their occurrence order.
template <class Key>
__global__ void reads_pass_kernel(const Key* key, const uint32_t* val, int64_t n, int m,
                                  uint8_t* pass) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    if (!(s >= m - 1 && key[s - (m - 1)] == key[s])) pass[val[s]] = 0;
  }
}

__global__ void reads_bin_kernel(const uint64_t* H0, const uint8_t* pass, int64_t n, int events,
                                 uint64_t binsize, uint32_t recip, uint64_t* bins) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    bool ok = true;
    for (int r = 0; r < events; ++r) ok = ok && pass[t * events + r];
    if (ok) {
      const uint64_t h = mod_sign(H0[t]);
      const uint64_t bin = bin_of(h, binsize, recip);
      if (h < bins[bin]) atomicMin((unsigned long long*)&bins[bin], (unsigned long long)h);
    }
  }
}

struct Max64 {
  __device__ __forceinline__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};

struct ValidOne {
  __device__ __forceinline__ int64_t operator()(uint8_t s) const { return s != 0 ? 1 : 0; }
};

inline unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// Workspace carving: every region 256-byte aligned.
struct Carve {
  char* p;
  size_t used = 0;
  template <class T>
  T* take(int64_t count) {
    T* r = reinterpret_cast<T*>(p ? p + used : nullptr);
    used += ((size_t)std::max<int64_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    return r;
  }
};

size_t cub_temp_bytes(int64_t max_windows) {
  size_t a = 0, b = 0, c = 0, d = 0;
  const int64_t W = std::max<int64_t>(max_windows, 1);
  hipcub::TransformInputIterator<int64_t, ValidOne, const uint8_t*> it(nullptr, ValidOne());
  (void)hipcub::DeviceScan::InclusiveScan(nullptr, a, (const int64_t*)nullptr, (int64_t*)nullptr,
                                    Max64(), (int)W);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, it, (int64_t*)nullptr, (int)W);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, c, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)(4 * W),
                                     0, 29);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, d, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)W, 0, 64);
  return std::max(std::max(a, b), std::max(c, d));
}

}  // namespace

size_t reads_workspace_bytes(int64_t max_windows) {
  Carve c{nullptr};
  const int64_t W = std::max<int64_t>(max_windows, 1);
  c.take<uint64_t>(W);      // h0
  c.take<uint8_t>(W);       // st
  c.take<int64_t>(W);       // rs
  c.take<int64_t>(W);       // run
  c.take<int64_t>(W);       // tix
  c.take<uint64_t>(W);      // H0
  c.take<uint64_t>(W);      // H1
  c.take<uint32_t>(4 * W);  // keys in (count-min) / 2 W for exact
  c.take<uint32_t>(4 * W);  // keys out
  c.take<uint32_t>(4 * W);  // vals in
  c.take<uint32_t>(4 * W);  // vals out
  c.take<uint8_t>(4 * W);   // pass
  c.take<char>((int64_t)cub_temp_bytes(W));
  return c.used;
}

hipError_t run_reads_job(const ReadsJob& j, void* ws, size_t ws_bytes, hipStream_t stream) {
  const int64_t W = j.n_windows;
  if (W <= 0) return hipSuccess;
  if (W > READS_MAX_WINDOWS || reads_workspace_bytes(W) > ws_bytes) return hipErrorInvalidValue;
  Carve c{static_cast<char*>(ws)};
  uint64_t* h0 = c.take<uint64_t>(W);
  uint8_t* st = c.take<uint8_t>(W);
  int64_t* rs = c.take<int64_t>(W);
  int64_t* run = c.take<int64_t>(W);
  int64_t* tix = c.take<int64_t>(W);
  uint64_t* H0 = c.take<uint64_t>(W);
  uint64_t* H1 = c.take<uint64_t>(W);
  uint32_t* kin = c.take<uint32_t>(4 * W);
  uint32_t* kout = c.take<uint32_t>(4 * W);
  uint32_t* vin = c.take<uint32_t>(4 * W);
  uint32_t* vout = c.take<uint32_t>(4 * W);
  uint8_t* pass = c.take<uint8_t>(4 * W);
  const size_t temp_bytes = cub_temp_bytes(W);
  void* temp = c.take<char>((int64_t)temp_bytes);

  hipError_t e = hipMemsetAsync(st, 0, (size_t)W, stream);
  if (e != hipSuccess) return e;
  if (j.codon) {
    reads_hash_kernel<3><<<grid_for(j.n_units), 256, 0, stream>>>(j, h0, st);
  } else {
    reads_hash_kernel<1><<<grid_for(j.n_units), 256, 0, stream>>>(j, h0, st);
  }
  reads_run_kernel<<<grid_for(W), 256, 0, stream>>>(st, W, rs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t tb = temp_bytes;
  e = hipcub::DeviceScan::InclusiveScan(temp, tb, rs, run, Max64(), (int)W, stream);
  if (e != hipSuccess) return e;
  hipcub::TransformInputIterator<int64_t, ValidOne, const uint8_t*> valid(st, ValidOne());
  tb = temp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(temp, tb, valid, tix, (int)W, stream);
  if (e != hipSuccess) return e;
  const int span = j.codon ? 3 * j.k - 2 : j.k;
  reads_compact_kernel<<<grid_for(W), 256, 0, stream>>>(st, h0, run, tix, W, span, j.codon, H0,
                                                        H1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // occurrence count
  int64_t last_t = 0;
  uint8_t last_st = 0;
  if ((e = hipMemcpyAsync(&last_t, tix + (W - 1), sizeof(int64_t), hipMemcpyDeviceToHost,
                          stream)) != hipSuccess ||
      (e = hipMemcpyAsync(&last_st, st + (W - 1), 1, hipMemcpyDeviceToHost, stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(stream)) != hipSuccess) {
    return e;
  }
  const int64_t N = last_t + (last_st != 0 ? 1 : 0);
  if (N == 0) return hipSuccess;
  int events;
  if (j.exact) {
    events = 1;
    uint64_t* k64in = reinterpret_cast<uint64_t*>(kin);
    uint64_t* k64out = reinterpret_cast<uint64_t*>(kout);
    e = hipMemcpyAsync(k64in, H0, (size_t)N * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return e;
    reads_iota_kernel<<<grid_for(N), 256, 0, stream>>>(vin, N);
    tb = temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(temp, tb, k64in, k64out, vin, vout, (int)N, 0, 64,
                                           stream);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(pass, 1, (size_t)N, stream);
    if (e != hipSuccess) return e;
    reads_pass_kernel<uint64_t><<<grid_for(N), 256, 0, stream>>>(k64out, vout, N, j.min_count,
                                                                  pass);
  } else {
    events = 4;
    reads_cm_keys_kernel<<<grid_for(4 * N), 256, 0, stream>>>(H0, H1, N, kin, vin);
    tb = temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(temp, tb, kin, kout, vin, vout, (int)(4 * N), 0, 29,
                                           stream);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(pass, 1, (size_t)(4 * N), stream);
    if (e != hipSuccess) return e;
    reads_pass_kernel<uint32_t><<<grid_for(4 * N), 256, 0, stream>>>(kout, vout, 4 * N,
                                                                      j.min_count, pass);
  }
  reads_bin_kernel<<<grid_for(N), 256, 0, stream>>>(H0, pass, N, events, j.binsize, j.recip,
                                                    j.bins);
  return hipGetLastError();
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

template <bool LDS_BINS, int STRIDE>
hipError_t launch_hash(const SketchArgs& a, int64_t n_groups, size_t lds, int* next_group,
                       hipStream_t stream) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sketch_hash_kernel<LDS_BINS, STRIDE>,
                                                     SKETCH_TB, lds);
  }
  if (e != hipSuccess) return e;
  const int64_t grid = std::max<int64_t>(
      1, std::min<int64_t>(n_groups, (int64_t)std::max(per_cu, 1) * std::max(cus, 1)));
  sketch_hash_kernel<LDS_BINS, STRIDE><<<(unsigned)grid, SKETCH_TB, lds, stream>>>(a, n_groups,
                                                                                 next_group);
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
    const size_t lds = lds_bins ? (size_t)nbins * sizeof(uint64_t) : 0;
    if (lds_bins) {
      e = a.codon ? launch_hash<true, 3>(a, n_groups, lds, next_group, stream)
                  : launch_hash<true, 1>(a, n_groups, lds, next_group, stream);
    } else {
      e = a.codon ? launch_hash<false, 3>(a, n_groups, lds, next_group, stream)
                  : launch_hash<false, 1>(a, n_groups, lds, next_group, stream);
    }
    if (e != hipSuccess) return e;
  }
  sketch_finish_kernel<<<(unsigned)(a.n_samples * a.n_k), 64, 0, stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_sketch_reset(const SketchArgs& a, hipStream_t stream) {
  const int64_t total = (int64_t)a.n_samples * a.n_k * a.s64 * 64;
  fill_u64<<<(unsigned)((total + 255) / 256 > 65536 ? 65536 : (total + 255) / 256), 256, 0,
             stream>>>(a.signs, total, ~0ULL);
  return hipGetLastError();
}

hipError_t launch_sketch_finish(const SketchArgs& a, hipStream_t stream) {
  sketch_finish_kernel<<<(unsigned)(a.n_samples * a.n_k), 64, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace ppsk
