// HIP kernels for the core/accessory distance path on gfx950 (MI355X).
//
// Hot loop (reference semantics: calc_intersize, dist/dist.cpp:50-78): for a
// pair of samples and one k-mer length, the match count is
//   sum over 64-bin words w of popcount( AND_l NOT(a[w][l] XOR b[w][l]) )
// over the 14 bit layers l.  Here every 64-bit layer word is processed as two
// 32-bit planes; per plane and pair that is one v_xnor_b32 (first layer) plus
// 13 v_bitop3_b32 (acc & ~(a ^ b)) and one v_bcnt_u32_b32: 30 int32 VALU ops
// per 64-bin word per pair, which is the algorithmic op count.
//
// Tiling (GEMM-shaped, integer VALU, no MFMA):
//  * a workgroup owns a tile of pairs for ONE k-mer length: its 8 waves form
//    a grid of row groups x column groups, each wave RQ rows x 64 RR columns
//    (RR columns per lane);
//  * the RQ row samples of a wave are wave-uniform, so their planes are read
//    with scalar loads into SGPRs: each v_bitop3 takes its row operand from
//    an SGPR;
//  * column planes are staged per group of words straight from global memory
//    into LDS (global_load_lds_dwordx4, double buffered) and read per lane;
//  * per-pair counts live in VGPRs for the whole word loop, then go to a
//    per-tile count scratch;
//  * the last of the n_k workgroups of a tile to finish (device-scope atomic)
//    runs the fused epilogue for the tile: random-match correction, Jaccard,
//    log-Jaccard-vs-k regression, dense output.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bin_fraction.h"
#include "log_jaccard.h"
#include "ppsk_internal.h"

namespace ppsk {

// ---------------------------------------------------------------------------
// Stripe transform: raw reference layout -> stripes.
// raw: [count][n_k][s64 * 14] uint64 (word w, layer l at w*14 + l,
// sketch/sketch.cpp:82-92), samples first .. first+count-1 of the set.
// One workgroup per (kk, stripe block, w): the block's 128 sample records
// of 14 uint64 (112 B each, contiguous in raw) are read with 16-byte loads
// into LDS, then the (kk, block, w) slab of 28 planes x 128 samples is
// written as one contiguous 14 KiB run.  Samples of the block outside
// [first, first + count) are left untouched (an upload in several chunks).
// ---------------------------------------------------------------------------
constexpr int STRIPE_TB = 256;
constexpr int STRIPE_REC = BBITS + 1;  // uint64 per LDS record (one pad word)

__global__ __launch_bounds__(STRIPE_TB) void stripe_kernel(const uint64_t* __restrict__ raw,
                                                           int64_t count, int64_t first,
                                                           int32_t n_k, int32_t s64, int32_t nb,
                                                           uint32_t* __restrict__ stripes) {
  __shared__ uint64_t rec[BLK * STRIPE_REC];
  const int64_t blk = (int64_t)first / BLK + blockIdx.y;
  const int kk = (int)(blockIdx.z);
  for (int w = (int)blockIdx.x; w < s64; w += (int)gridDim.x) {
    // samples of this block in this upload: n in [n0, n1)
    const int64_t n0 = blk * BLK > first ? blk * BLK : first;
    const int64_t n1 = (blk + 1) * BLK < first + count ? (blk + 1) * BLK : first + count;
    // load: 7 uint4 per sample record
    for (int x = threadIdx.x; x < BLK * 7; x += STRIPE_TB) {
      const int lane = x / 7, part = x % 7;
      const int64_t n = blk * BLK + lane;
      if (n >= n0 && n < n1) {
        const uint4* src = reinterpret_cast<const uint4*>(
            raw + (((n - first) * n_k + kk) * (int64_t)s64 + w) * BBITS);
        const uint4 v = src[part];
        uint64_t* d = rec + lane * STRIPE_REC + part * 2;
        d[0] = (uint64_t)v.x | ((uint64_t)v.y << 32);
        d[1] = (uint64_t)v.z | ((uint64_t)v.w << 32);
      }
    }
    __syncthreads();
    uint32_t* dst = stripes + (((int64_t)kk * nb + blk) * s64 + w) * (HL * BLK);
    for (int o = threadIdx.x; o < HL * BLK; o += STRIPE_TB) {
      const int hl = o / BLK, lane = o % BLK;
      const int64_t n = blk * BLK + lane;
      if (n >= n0 && n < n1) {
        const int h = hl / BBITS, l = hl % BBITS;
        const uint64_t v = rec[lane * STRIPE_REC + l];
        dst[o] = h ? (uint32_t)(v >> 32) : (uint32_t)v;
      }
    }
    __syncthreads();
  }
}

hipError_t launch_stripe(const uint64_t* raw, int64_t count, int64_t first,
                         int32_t n_k, int32_t s64, int32_t nb, uint32_t* stripes,
                         hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const int64_t b0 = first / BLK, b1 = (first + count - 1) / BLK;
  const unsigned words = (unsigned)(s64 < 64 ? s64 : 64);  // grid x strides over words
  dim3 grid(words, (unsigned)(b1 - b0 + 1), (unsigned)n_k);
  stripe_kernel<<<grid, STRIPE_TB, 0, stream>>>(raw, count, first, n_k, s64, nb, stripes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------
typedef const __attribute__((address_space(4))) uint32_t CU32;
typedef const __attribute__((address_space(4))) DistParams CParams;

__device__ __forceinline__ uint32_t match_and(uint32_t acc, uint32_t a, uint32_t b) {
  // truth table over (acc, a, b) with src0 = 0xF0, src1 = 0xCC, src2 = 0xAA:
  // acc & ~(a ^ b) = 0x90
  return __builtin_amdgcn_bitop3_b32(acc, a, b, 0x90);
}

// v_bcnt_u32_b32 with accumulate: cnt + popcount(x) in one VALU op (the
// compiler otherwise splits paired popcounts into two bcnt and a v_add3).
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t cnt) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(cnt));
  return r;
}

// Per-pair epilogue pieces.  Reference semantics (CPU path):
//  J_obs = c / (64 * s64)                        reference.cpp:222-232
//  J     = max(J_obs - r, 0) * 1 / (1 - r)       dist/dist.hpp:13-24 (double)
//  dists[kk] = (float) J                         linear_regression.cpp:62-71
//  window [first J > tol, last J > tol], LS of log J on k, core/acc
//                                                linear_regression.cpp:16-60
template <typename P>
__device__ __forceinline__ double random_for(const P& p, int kk, int64_t gi, int64_t gj) {
  if (p.random_mode == RANDOM_EQUAL_BASE) {
    // csrs combination (random_match.cpp:333-334) of per-sample terms
    const double j1 = p.jv_rows[gi * p.n_k + kk];
    const double j2 = p.jv_cols[gj * p.n_k + kk];
    double r = 0.0;
    if (j1 > 0 && j2 > 0) r = (j1 * j2) / (j1 + j2 - j1 * j2);
    return r;
  }
  if (p.random_mode == RANDOM_MC) {
    const int C = p.n_clusters;
    const int cr = p.cl_rows[gi], cc = p.cl_cols[gj];
    const int a = p.mc_ref_is_col ? cc : cr;
    const int b = p.mc_ref_is_col ? cr : cc;
    return (double)p.mc_table[(kk * C + a) * C + b];
  }
  return 0.0;
}

// Cancellation request of the call (ppsk_cancel), read from host-mapped
// memory with a system-scope load so that no cache holds a stale value.
__device__ __forceinline__ bool cancel_seen(const uint32_t* flag) {
  return flag != nullptr &&
         __hip_atomic_load(const_cast<uint32_t*>(flag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
}

// n more (tile, k-mer length) items finished: the device counter, mirrored
// into the host-mapped progress word with a plain store.
__device__ __forceinline__ void add_progress(const Control& c, unsigned long long n) {
  if (c.progress == nullptr) return;
  const unsigned long long old = atomicAdd(c.progress, n);
  if (c.progress_host != nullptr) {
    __hip_atomic_store(c.progress_host, old + n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__device__ __forceinline__ double jaccard_d(uint32_t c, double bins, double rbins, double r) {
  const double obs = bin_fraction(c, bins, rbins);
  const double diff = obs > r ? (obs - r) : 0.0;
  return (diff * 1.0) / (1.0 - r);
}

// Table of the regression's logarithm (log_jaccard.h); every device holds a
// copy, written once when the library sets the device up.
__device__ double g_log_tab[LOG_TAB_DOUBLES];


// ---------------------------------------------------------------------------
// Tile shape and LDS staging
// ---------------------------------------------------------------------------
// A stage is WS words x 28 planes x TC dwords = 1792 uint4 (28 KiB), double
// buffered (56 KiB per workgroup, two workgroups per CU).  Stages are copied
// global -> LDS directly: round i of wave w moves uint4 [i*512 + w*64, +64)
// of the stage; the last round is a half round.
//
// The 8 waves form a grid of RG row groups x CG column groups: wave w takes
// rows (w / CG) * RQ .. +RQ and columns (w % CG) * 64 RR .. +64 RR of the
// tile (lane l: columns l + 64 r, r < RR, of its group).  Every column
// operand a lane reads from LDS meets RQ row operands, so the VALU work per
// LDS byte is RQ / 4 ops: at RQ = 8 that matches the CU's VALU issue rate
// (256 lane-ops per clock) to its LDS read rate (128 B per clock).  The
// pipelined shapes therefore all keep RQ = 8 and span 256 columns
// (CG = 4 / RR): 64 x 256 (8 x 1 waves), 32 x 256 (4 x 2), 16 x 256 (2 x 4).
// The compiler-scheduled shapes (RQ = 16) stack all waves on the rows.
template <int RQ, int RR>
struct TileCfg {
  static constexpr int CG = RQ == 8 ? 4 / RR : 1;
  static constexpr int RG = WAVES / CG;
  static constexpr int TROWS = RG * RQ;
  static constexpr int TC = CG * 64 * RR;
  static constexpr int WS = 4 / (CG * RR);
  static_assert(CG * RG == WAVES, "the wave grid covers the workgroup");
  static constexpr int STAGE_U4 = WS * HL * TC / 4;  // 1792
  static constexpr int ROUNDS = (STAGE_U4 + WAVES * 64 - 1) / (WAVES * 64);  // 4
  static_assert(STAGE_U4 % 64 == 0, "stage must be whole wave rounds");
  static_assert(BLK % TROWS == 0, "row tiles must not straddle stripe blocks");
  // row panel of one stage (WS words x 28 planes x TROWS samples) in uint4
  static constexpr int ROW_U4 = WS * HL * TROWS / 4;
  static constexpr int ROW_ROUNDS = (ROW_U4 + WAVES * 64 - 1) / (WAVES * 64);
  // LDS target of the row prefetch: whole waves of uint4 (never read)
  static constexpr int ROW_LDS_U4 = (ROW_U4 + 63) / 64 * 64;
};

// s_waitcnt vmcnt(n) with expcnt / lgkmcnt left unconstrained (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// LDS byte address of a __shared__ object.
__device__ __forceinline__ uint32_t lds_addr(const uint4* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// One 16-byte-per-lane global -> LDS copy: lane L reads saddr + voff (bytes)
// and writes LDS bytes [lds_base + OFF + 16 L, +16).  Written as inline asm
// so that the compiler's wait-count pass does not treat it as a pending write
// to every LDS buffer (it cannot tell the two stage buffers apart and would
// wait for the copy of the next stage before reading the current one);
// completion is awaited explicitly with vmcnt before the stage barrier.  The
// LDS address is base + OFF with OFF an immediate, so a stage needs one SGPR
// per buffer rather than one per copy.
template <int OFF>
__device__ __forceinline__ void dma16(const uint32_t* saddr, uint32_t voff, uint32_t lds_base) {
  asm volatile("s_add_u32 m0, %0, %3\n\tglobal_load_lds_dwordx4 %1, %2"
               :
               : "s"(lds_base), "v"(voff), "s"(saddr), "i"(OFF)
               : "memory", "m0", "scc");
}

// Workgroup barrier without the release/acquire fences of __syncthreads():
// those wait for every outstanding vector-memory operation, including the
// next stage's direct-to-LDS copy that is meant to overlap the matching.
// Ordering of the stage copies is handled explicitly with vmcnt waits.
__device__ __forceinline__ void stage_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Column operands of one stage: words [w0, w0 + WS) of the tile's TC columns,
// which may span two stripe blocks (TC = 256) or half of one (TC = 64).
// Columns past the last stripe block read the last block (masked later).
// Round I of wave w copies uint4 [I * 512 + w * 64, +64) of the stage.
//
// Addressing: a lane's source is (tile column base) + (w0 * plane slab) +
// (its byte offset within the tile's <= 2 stripe blocks).  The last term does
// not depend on the word, so it is computed once per tile (ColOffsets, one
// VGPR per round, < 2^32 since it spans at most two blocks of one word group)
// and every stage only moves the scalar base.
template <int RQ, int RR>
struct ColOffsets {
  uint32_t v[TileCfg<RQ, RR>::ROUNDS];
};

template <int RQ, int RR>
__device__ __forceinline__ ColOffsets<RQ, RR> col_offsets(int64_t col0, int nb_cols, int s64,
                                                          int wave, int lane) {
  using Cfg = TileCfg<RQ, RR>;
  constexpr int TC = Cfg::TC;
  ColOffsets<RQ, RR> o;
  const int blk0 = (int)(col0 / BLK);
#pragma unroll
  for (int i = 0; i < Cfg::ROUNDS; ++i) {
    const int x = i * (WAVES * 64) + wave * 64 + lane;
    const int row = x / (TC / 4);  // ws * HL + hl
    const int64_t col = col0 + (x % (TC / 4)) * 4;
    int blk = (int)(col / BLK);
    blk = blk < nb_cols ? blk : nb_cols - 1;
    const uint32_t e = (uint32_t)(blk - blk0) * (uint32_t)s64 * (HL * BLK) +
                       (uint32_t)row * BLK + (uint32_t)(col % BLK);
    o.v[i] = e * 4u;
  }
  return o;
}

template <int RQ, int RR, int I>
__device__ __forceinline__ void issue_round(const uint32_t* sbase, const ColOffsets<RQ, RR>& off,
                                            uint32_t wbase, int wave) {
  using Cfg = TileCfg<RQ, RR>;
  if constexpr (I < Cfg::ROUNDS) {
    if (I * (WAVES * 64) + wave * 64 < Cfg::STAGE_U4) {  // wave-uniform
      dma16<I * WAVES * 64 * 16>(sbase, off.v[I], wbase);
    }
    issue_round<RQ, RR, I + 1>(sbase, off, wbase, wave);
  }
}

// Row-panel prefetch: the stage's row planes (read later with scalar loads)
// are copied into a scratch LDS area one stage ahead, only so that the scalar
// loads find them in L2.  Lane offsets as for the columns.
template <int RQ, int RR>
struct RowOffsets {
  uint32_t v[TileCfg<RQ, RR>::ROW_ROUNDS];
};

template <int RQ, int RR>
__device__ __forceinline__ RowOffsets<RQ, RR> row_offsets(int wave, int lane) {
  using Cfg = TileCfg<RQ, RR>;
  constexpr int PER_PLANE = Cfg::TROWS / 4;  // uint4 per plane of the panel
  RowOffsets<RQ, RR> o;
#pragma unroll
  for (int i = 0; i < Cfg::ROW_ROUNDS; ++i) {
    int x = i * (WAVES * 64) + wave * 64 + lane;
    if (x >= Cfg::ROW_U4) x = 0;  // lanes past the panel repeat its first uint4
    const int plane = x / PER_PLANE;  // ws * HL + hl
    const int part = x % PER_PLANE;
    o.v[i] = (uint32_t)(plane * BLK + part * 4) * 4u;
  }
  return o;
}

template <int RQ, int RR, int I>
__device__ __forceinline__ void prefetch_round(const uint32_t* rbase, const RowOffsets<RQ, RR>& off,
                                               uint32_t wbase, int wave) {
  using Cfg = TileCfg<RQ, RR>;
  if constexpr (I < Cfg::ROW_ROUNDS) {
    if (I * (WAVES * 64) + wave * 64 < Cfg::ROW_U4) {  // wave-uniform
      dma16<I * WAVES * 64 * 16>(rbase, off.v[I], wbase);
    }
    prefetch_round<RQ, RR, I + 1>(rbase, off, wbase, wave);
  }
}

// sbase: the tile's first column block, word w0 (wave-uniform).
template <int RQ, int RR>
__device__ __forceinline__ void issue_stage(const uint32_t* sbase, const ColOffsets<RQ, RR>& off,
                                            uint4* buf, int wave) {
  const uint32_t wbase = lds_addr(buf) + (uint32_t)(wave * 64 * 16);
  issue_round<RQ, RR, 0>(sbase, off, wbase, wave);
}

// ---------------------------------------------------------------------------
// Matching, row operands left to the compiler (PB = 0): one plane at a time.
// ---------------------------------------------------------------------------
template <int RQ, int RR>
__device__ __forceinline__ void match_stage_plain(const CU32* arow, int w0, int nw,
                                                  const uint32_t* sB, int lane,
                                                  uint32_t (&cnt)[RQ][RR]) {
  constexpr int TC = TileCfg<RQ, RR>::TC;
  static_assert(TileCfg<RQ, RR>::CG == 1, "compiler-scheduled shapes stack waves on rows");
  for (int ws = 0; ws < nw; ++ws) {
    const CU32* aw = arow + (int64_t)(w0 + ws) * (HL * BLK);
    const uint32_t* bw = sB + ws * (HL * TC) + lane;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t acc[RQ][RR];
#pragma unroll
      for (int l = 0; l < BBITS; ++l) {
        const int hl = h * BBITS + l;
        uint32_t a[RQ], b[RR];
#pragma unroll
        for (int q = 0; q < RQ; ++q) a[q] = aw[hl * BLK + q];
#pragma unroll
        for (int r = 0; r < RR; ++r) b[r] = bw[hl * TC + 64 * r];
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
#pragma unroll
          for (int r = 0; r < RR; ++r) {
            acc[q][r] = (l == 0) ? ~(a[q] ^ b[r]) : match_and(acc[q][r], a[q], b[r]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < RQ; ++q)
#pragma unroll
        for (int r = 0; r < RR; ++r) cnt[q][r] = bcnt_acc(acc[q][r], cnt[q][r]);
    }
  }
}

// ---------------------------------------------------------------------------
// Matching with an explicit two-deep operand pipeline (PB = 1 or 2).
// Scalar loads return out of order, so the only possible wait for them is
// "all outstanding" (lgkmcnt(0)), which also covers every LDS read; left to
// itself the compiler issues one plane's s_load and waits for it right away,
// exposing the full L2 latency every RQ * RR VALU ops.  Here the row
// operands of PB planes (scalar loads into SGPR tuples) and the matching
// column operands (ds_read2st64 into VGPR pairs) form one batch: batch b+1
// is issued before batch b is matched, and a single wait that names the
// registers it completes follows the matching.  Each round trip therefore
// overlaps PB * RQ * RR bitop3 of the same wave plus the other waves of the
// SIMD.  All loads of the pipeline are issued from inline asm, so the
// compiler's own wait counting never sees (or waits on) them.
// ---------------------------------------------------------------------------
template <int RQ>
struct ATuple;
template <>
struct ATuple<16> {
  typedef uint32_t T __attribute__((ext_vector_type(16)));
};
template <>
struct ATuple<8> {
  typedef uint32_t T __attribute__((ext_vector_type(8)));
};
template <>
struct ATuple<4> {
  typedef uint32_t T __attribute__((ext_vector_type(4)));
};
template <>
struct ATuple<2> {
  typedef uint32_t T __attribute__((ext_vector_type(2)));
};
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Column operands come in pairs from one ds_read2st64_b32: two column blocks
// of one plane (RR >= 2), or one column of two consecutive planes (RR = 1).
template <int RR, int PB>
struct ColRegs {
  static constexpr int N = RR >= 2 ? PB * (RR / 2) : PB / 2;
};

template <int RQ, int RR, int PB>
struct Batch {
  static_assert(RR >= 2 ? RR % 2 == 0 : PB % 2 == 0, "column operands are read in pairs");
  typename ATuple<RQ>::T a[PB];     // row planes: RQ samples each
  u32x2 b[ColRegs<RR, PB>::N];      // column planes: RR per lane and plane
};

// Column operand (plane pb of the batch, column block r) of a batch.
template <int RQ, int RR, int PB>
__device__ __forceinline__ uint32_t col_operand(const Batch<RQ, RR, PB>& x, int pb, int r) {
  if constexpr (RR >= 2) {
    return x.b[pb * (RR / 2) + r / 2][r % 2];
  } else {
    return x.b[pb / 2][pb % 2];
  }
}

template <int RQ, int OFF>
__device__ __forceinline__ void a_load(const CU32* aw, typename ATuple<RQ>::T& x) {
  if constexpr (RQ == 16) {
    asm volatile("s_load_dwordx16 %0, %1, %2" : "=&s"(x) : "s"(aw), "i"(OFF));
  } else if constexpr (RQ == 8) {
    asm volatile("s_load_dwordx8 %0, %1, %2" : "=&s"(x) : "s"(aw), "i"(OFF));
  } else if constexpr (RQ == 4) {
    asm volatile("s_load_dwordx4 %0, %1, %2" : "=&s"(x) : "s"(aw), "i"(OFF));
  } else {
    static_assert(RQ == 2, "row operands come in SGPR tuples of 2, 4, 8 or 16");
    asm volatile("s_load_dwordx2 %0, %1, %2" : "=&s"(x) : "s"(aw), "i"(OFF));
  }
}

// Issue planes P0 .. P0+PB-1 of the word at (aw, bw); P0 = HL addresses the
// first planes of the next word.  bw is the LDS byte address of plane 0 at the
// lane's first column (a stage holds TC columns per plane); column block r of
// plane P sits at bw + (P TC + 64 r) 4, i.e. at offset P TC / 64 + r in the
// 256-byte units of ds_read2st64.
template <int RQ, int RR, int PB, int P0>
__device__ __forceinline__ void batch_issue(const CU32* aw, uint32_t bw, Batch<RQ, RR, PB>& x) {
  constexpr int TC = TileCfg<RQ, RR>::TC;  // plane stride of the LDS stage
#pragma unroll
  for (int pb = 0; pb < PB; ++pb) {
    if (pb == 0) a_load<RQ, P0 * BLK * 4>(aw, x.a[0]);
    if (pb == 1) a_load<RQ, (P0 + 1) * BLK * 4>(aw, x.a[PB - 1]);
  }
  if constexpr (RR >= 2) {
#pragma unroll
    for (int pb = 0; pb < PB; ++pb) {
#pragma unroll
      for (int rp = 0; rp < RR / 2; ++rp) {
        asm volatile("ds_read2st64_b32 %0, %1 offset0:%2 offset1:%3"
                     : "=&v"(x.b[pb * (RR / 2) + rp])
                     : "v"(bw), "i"((P0 + pb) * TC / 64 + 2 * rp),
                       "i"((P0 + pb) * TC / 64 + 2 * rp + 1));
      }
    }
  } else {
    // one column: planes P0 + pb and P0 + pb + 1 in one read
#pragma unroll
    for (int pb = 0; pb < PB; pb += 2) {
      asm volatile("ds_read2st64_b32 %0, %1 offset0:%2 offset1:%3"
                   : "=&v"(x.b[pb / 2])
                   : "v"(bw), "i"((P0 + pb) * TC / 64), "i"((P0 + pb + 1) * TC / 64));
    }
  }
}

template <int RQ, int RR, int PB>
__device__ __forceinline__ void batch_wait(Batch<RQ, RR, PB>& x) {
  if constexpr (PB == 2 && RR == 4) {
    asm volatile("s_waitcnt lgkmcnt(0)"
                 : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]), "+v"(x.b[1]), "+v"(x.b[2]),
                   "+v"(x.b[3]));
  } else if constexpr (PB == 2 && RR == 2) {
    asm volatile("s_waitcnt lgkmcnt(0)" : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]), "+v"(x.b[1]));
  } else if constexpr (PB == 2 && RR == 1) {
    asm volatile("s_waitcnt lgkmcnt(0)" : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]));
  } else if constexpr (PB == 1 && RR == 4) {
    asm volatile("s_waitcnt lgkmcnt(0)" : "+s"(x.a[0]), "+v"(x.b[0]), "+v"(x.b[1]));
  } else {
    static_assert(PB == 1 && RR == 2, "unsupported batch shape");
    asm volatile("s_waitcnt lgkmcnt(0)" : "+s"(x.a[0]), "+v"(x.b[0]));
  }
}

// Keeps a batch's registers alive (unchanged) at this point of the program.
// Placed right after the next batch's loads are issued, it stops the
// compiler from matching the current batch first and then reusing its
// registers for the next batch's loads, which would serialise every load
// round trip behind the matching.
template <int RQ, int RR, int PB>
__device__ __forceinline__ void batch_hold(Batch<RQ, RR, PB>& x) {
  if constexpr (PB == 2 && RR == 4) {
    asm volatile("" : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]), "+v"(x.b[1]), "+v"(x.b[2]),
                      "+v"(x.b[3]));
  } else if constexpr (PB == 2 && RR == 2) {
    asm volatile("" : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]), "+v"(x.b[1]));
  } else if constexpr (PB == 2 && RR == 1) {
    asm volatile("" : "+s"(x.a[0]), "+s"(x.a[1]), "+v"(x.b[0]));
  } else if constexpr (PB == 1 && RR == 4) {
    asm volatile("" : "+s"(x.a[0]), "+v"(x.b[0]), "+v"(x.b[1]));
  } else {
    static_assert(PB == 1 && RR == 2, "unsupported batch shape");
    asm volatile("" : "+s"(x.a[0]), "+v"(x.b[0]));
  }
}

// Planes P .. P+PB-1 of one word against the lane's RR columns.
template <int RQ, int RR, int PB, int P>
__device__ __forceinline__ void match_batch(const Batch<RQ, RR, PB>& x, uint32_t (&acc)[RQ][RR]) {
#pragma unroll
  for (int pb = 0; pb < PB; ++pb) {
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        const uint32_t a = x.a[pb][q];
        const uint32_t b = col_operand<RQ, RR, PB>(x, pb, r);
        if ((P + pb) % BBITS == 0) {
          acc[q][r] = ~(a ^ b);
        } else {
          acc[q][r] = match_and(acc[q][r], a, b);
        }
      }
    }
  }
}

// Batches B .. NB-1 of one word; cur holds batch B (loaded and complete).
// The last batch prefetches the first batch of the next word (if any) into
// the registers the next word starts from.
template <int RQ, int RR, int PB, int B>
__device__ __forceinline__ void match_batches(const CU32* aw, uint32_t bw, bool next_word,
                                              Batch<RQ, RR, PB>& cur, Batch<RQ, RR, PB>& nxt,
                                              uint32_t (&acc)[RQ][RR], uint32_t (&cnt)[RQ][RR]) {
  constexpr int NB = HL / PB;
  if constexpr (B + 1 < NB) {
    batch_issue<RQ, RR, PB, (B + 1) * PB>(aw, bw, nxt);
  } else {
    if (next_word) batch_issue<RQ, RR, PB, HL>(aw, bw, nxt);  // word w + 1
  }
  batch_hold<RQ, RR, PB>(cur);
  match_batch<RQ, RR, PB, B * PB>(cur, acc);
  if constexpr (((B + 1) * PB) % BBITS == 0) {
    // end of a 32-bit half: fold its matches into the counts
#pragma unroll
    for (int q = 0; q < RQ; ++q)
#pragma unroll
      for (int r = 0; r < RR; ++r) cnt[q][r] = bcnt_acc(acc[q][r], cnt[q][r]);
  }
  if constexpr (B + 1 < NB) {
    batch_wait<RQ, RR, PB>(nxt);
    match_batches<RQ, RR, PB, B + 1>(aw, bw, next_word, nxt, cur, acc, cnt);
  }
}

// col: the lane's first column within the tile ((wave % CG) * 64 RR + lane).
template <int RQ, int RR, int PB>
__device__ __forceinline__ void match_stage_pipe(const CU32* arow, int w0, int nw,
                                                 const uint4* sbuf, int col,
                                                 uint32_t (&cnt)[RQ][RR]) {
  constexpr int TC = TileCfg<RQ, RR>::TC;
  static_assert((HL / PB) % 2 == 0, "an even batch count keeps the register sets aligned");
  Batch<RQ, RR, PB> c, n;
  const CU32* aw = arow + (int64_t)w0 * (HL * BLK);
  uint32_t bw = lds_addr(sbuf) + col * 4;
  batch_issue<RQ, RR, PB, 0>(aw, bw, c);
  for (int ws = 0; ws < nw; ++ws) {
    batch_wait<RQ, RR, PB>(c);
    uint32_t acc[RQ][RR];
    // batches alternate (c, n); the last one (odd) leaves the next word's
    // first batch in c
    match_batches<RQ, RR, PB, 0>(aw, bw, ws + 1 < nw, c, n, acc, cnt);
    aw += HL * BLK;
    bw += HL * TC * 4;
  }
}

// ---------------------------------------------------------------------------
// Fused epilogue of one tile (run by the last of its n_k workgroups):
// per-k counts -> random-corrected Jaccard -> dense output.  Out of line so
// that none of its state occupies registers during the matching loop.
// ---------------------------------------------------------------------------
// tile_counts: the tile's count block [n_k][tile rows][tile cols].
template <int RQ, int RR>
__device__ __noinline__ void tile_epilogue(CParams* pp, const uint16_t* tile_counts, Tile tile,
                                           int wave, int lane) {
  CParams& p = *pp;
  using Cfg = TileCfg<RQ, RR>;
  constexpr int TC = Cfg::TC;
  constexpr int TROWS = Cfg::TROWS;
  const double bins = 64.0 * (double)p.s64;
  const double rbins = 1.0 / bins;
  // this wave's rows and this lane's first column in the tile (wave grid)
  const int wrow = (wave / Cfg::CG) * RQ;
  const int wcol = (wave % Cfg::CG) * 64 * RR + lane;
  const int64_t gi0 = (int64_t)tile.bi * TROWS + wrow;
  const int64_t col0 = (int64_t)tile.ct * TC + wcol - lane;
  const uint16_t* tbase = tile_counts + (int64_t)wrow * TC + wcol;
  const int64_t kstride = (int64_t)TROWS * TC;
  for (int q = 0; q < RQ; ++q) {
    const int64_t gi = gi0 + q;
    for (int r = 0; r < RR; ++r) {
      const int64_t gj = col0 + lane + 64 * r;
      bool valid;
      if (p.self_mode) {
        valid = p.full_square ? (gi < p.n_rows && gj < p.n_cols) : (gi < gj && gj < p.n_cols);
      } else {
        valid = gi < p.n_rows && gj < p.n_cols;
      }
      if (!valid) continue;
      int64_t pos;
      if (p.full_square) {
        pos = (gi - p.square_row0) * p.out_stride_rows + gj;
      } else if (p.self_mode) {
        pos = p.n_cols * gi - ((gi * (gi + 1)) >> 1) + gj - 1 - gi;  // square_to_condensed
      } else {
        pos = gi * p.n_cols + gj;  // query-major: q * n_ref + r
      }
      pos -= p.out_row_offset;
      const uint16_t* cp = tbase + q * TC + 64 * r;
      if (p.out_kind == OUT_COUNTS) {
        uint16_t* o = reinterpret_cast<uint16_t*>(p.out) + pos * p.n_k;
        for (int k = 0; k < p.n_k; ++k) o[k] = cp[k * kstride];
        continue;
      }
      if (p.out_kind == OUT_JACCARD) {
        float* o = reinterpret_cast<float*>(p.out) + pos * p.n_k;
        for (int k = 0; k < p.n_k; ++k) {
          o[k] = (float)jaccard_d(cp[k * kstride], bins, rbins, random_for(p, k, gi, gj));
        }
        continue;
      }
      if (p.out_kind == OUT_ONE_MINUS_J) {
        // kNN on Jaccards: (float)(1 - J) of the first two k-mer lengths
        // (src/api.cpp:368-371 narrows 1.0f - double J to float)
        float v[2] = {INFINITY, INFINITY};
        for (int k = 0; k < 2 && k < p.n_k; ++k) {
          v[k] = (float)(1.0 - jaccard_d(cp[k * kstride], bins, rbins, random_for(p, k, gi, gj)));
        }
        reinterpret_cast<float2*>(p.out)[pos] = make_float2(v[0], v[1]);
        continue;
      }
      // core/acc: streaming window fit.  Points before the first valid one are
      // dropped; invalid points after the last valid one are held as pending
      // sums and committed only if a later valid point arrives.
      int first = -1, last = -1;
      double sy = 0.0, sxy = 0.0, py = 0.0, pxy = 0.0;
      for (int k = 0; k < p.n_k; ++k) {
        const float jf = (float)jaccard_d(cp[k * kstride], bins, rbins, random_for(p, k, gi, gj));
        const bool ok = (double)jf > p.tolerance;
        if (!ok && first < 0) continue;
        const double x = (double)p.kvals[k];
        const double y = log_jaccard(jf, g_log_tab);
        if (ok) {
          if (first < 0) {
            first = k;
            sy = y;
            sxy = x * y;
          } else {
            sy += py + y;
            sxy += pxy + x * y;
          }
          py = 0.0;
          pxy = 0.0;
          last = k;
        } else {
          py += y;
          pxy += x * y;
        }
      }
      float core = 0.0f, accd = 0.0f;
      if (first < 0 || last - first + 1 < 2) {
        atomicAdd(p.fail_count, 1ull);
        atomicMin(p.fail_first, (unsigned long long)(pos + p.out_row_offset));
        if (p.fail_rows) atomicAdd(&p.fail_rows[gi], 1u);
      } else {
        const double n = (double)(last - first + 1);
        double sx = 0.0, sxx = 0.0;
        for (int k = first; k <= last; ++k) {
          const double x = (double)p.kvals[k];
          sx += x;
          sxx += x * x;
        }
        const double mx = sx / n;
        const double my = sy / n;
        // centred sums: Sxx = sum x^2 - n mx^2, Sxy = sum xy - n mx my
        const double cxx = sxx - n * mx * mx;
        const double cxy = sxy - n * mx * my;
        const double slope = cxy / cxx;
        const double intercept = my - slope * mx;
        if (slope < 0) core = (float)(1.0 - exp(slope));
        if (intercept < 0) accd = (float)(1.0 - exp(intercept));
      }
      reinterpret_cast<float2*>(p.out)[pos] = make_float2(core, accd);
    }
  }
}

// ---------------------------------------------------------------------------
// Match counts of one tile for k-mer length index kk into cnt (zeroed here).
// pp: the distance parameters, read with scalar loads (kernarg segment or a
// job table entry).  sB0 / sB1: the two LDS stage buffers.
// ---------------------------------------------------------------------------
// The stage loop of tile_match with a narrower matching, for the diagonal
// tiles of a self pass (tile_match_diagonal below).  Copies and the row
// prefetch follow the tile's shape <RQ, RR>; the matching covers MR column
// blocks of 64 per lane from tile column `col`, against the same column
// stage (same plane stride and words per stage).  tile_match keeps its own
// copy of this loop so that the code of the kernels without the diagonal
// path stays exactly as it is.
template <int RQ, int RR, int PB, int MR>
__device__ __forceinline__ void tile_stages(const CU32* arow, const uint32_t* cbase,
                                            const ColOffsets<RQ, RR>& off,
                                            const uint32_t* rbase, const RowOffsets<RQ, RR>& roff,
                                            uint32_t rdst, bool rpf, int s64, uint4* sB0,
                                            uint4* sB1, int wave, int lane, int col,
                                            uint32_t (&cnt)[RQ][MR]) {
  constexpr int WS = TileCfg<RQ, RR>::WS;
  static_assert(MR == RR || (PB != 0 && TileCfg<RQ, MR>::TC == TileCfg<RQ, RR>::TC &&
                             TileCfg<RQ, MR>::WS == WS),
                "a narrower matching must read the same column stage");
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int r = 0; r < MR; ++r) cnt[q][r] = 0;

  const int n_stages = (s64 + WS - 1) / WS;
  issue_stage<RQ, RR>(cbase, off, sB0, wave);
  if (rpf) prefetch_round<RQ, RR, 0>(rbase, roff, rdst, wave);
  // One barrier per stage: after it, stage s is in LDS for every wave and
  // every wave has finished matching stage s - 1, whose buffer the copy of
  // stage s + 1 (issued right after the barrier) then overwrites while stage
  // s is matched.  Stages alternate between the two buffers; each pass of
  // this loop runs an even stage from sB0 and the following odd stage from
  // sB1.
  for (int st = 0; st < n_stages; st += 2) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int s = st + half;
      if (s < n_stages) {
        uint4* cur = half ? sB1 : sB0;
        uint4* nxt = half ? sB0 : sB1;
        const int w0 = s * WS;
        wait_vmcnt<0>();
        stage_barrier();
        if (s + 1 < n_stages) {
          issue_stage<RQ, RR>(cbase + (int64_t)(w0 + WS) * (HL * BLK), off, nxt, wave);
          if (rpf) prefetch_round<RQ, RR, 0>(rbase + (int64_t)(w0 + WS) * (HL * BLK), roff, rdst, wave);
        }
        const int nw = (s64 - w0) < WS ? (s64 - w0) : WS;
        if constexpr (PB == 0) {
          match_stage_plain<RQ, MR>(arow, w0, nw, reinterpret_cast<const uint32_t*>(cur), lane,
                                    cnt);
        } else {
          match_stage_pipe<RQ, MR, PB>(arow, w0, nw, cur, col, cnt);
        }
      }
    }
  }
  // every wave is done with both buffers before they are reused
  stage_barrier();
}

// Diagonal tiles of a self pass (upper triangle, i < j): a 64-column block
// of a 64 x 256 tile whose last column precedes the tile's first row holds no
// output pair.  A tile whose first row lies 128 (192) columns past its first
// column matches only its last 2 (1) blocks, with the matching of the
// 32-row (16-row) shape, which reads the same column stage; the skipped
// counts are zero and the epilogue never reads them.  Returns false (and
// does nothing) for any other tile.  Tile-uniform.
template <int RQ, int RR, int PB>
__device__ __forceinline__ bool tile_match_diagonal(CParams* pp, Tile tile, int kk, uint4* sB0,
                                                    uint4* sB1, uint4* sR, int wave, int lane,
                                                    uint32_t (&cnt)[RQ][RR]) {
  static_assert(RQ == 8 && RR == 4 && PB != 0, "64 x 256 pipelined tiles only");
  using Cfg = TileCfg<RQ, RR>;
  if (!pp->self_mode || pp->full_square) return false;
  const int64_t row0 = (int64_t)tile.bi * Cfg::TROWS;
  const int64_t col0 = (int64_t)tile.ct * Cfg::TC;
  const int64_t d = row0 - col0;
  if (d < 2 * 64) return false;
  const int s64 = pp->s64;
  const CU32* arow = (CU32*)(pp->row_stripes +
                             ((int64_t)kk * pp->nb_rows + row0 / BLK) * (int64_t)s64 * (HL * BLK) +
                             row0 % BLK + (wave / Cfg::CG) * RQ);
  const int nb_cols = pp->nb_cols;
  const uint32_t* cbase = pp->col_stripes + (int64_t)kk * nb_cols * (int64_t)s64 * (HL * BLK) +
                          (int64_t)(col0 / BLK) * s64 * (HL * BLK);
  const ColOffsets<RQ, RR> off = col_offsets<RQ, RR>(col0, nb_cols, s64, wave, lane);
  const bool rpf = pp->row_prefetch != 0;
  const uint32_t* rbase = pp->row_stripes +
                          ((int64_t)kk * pp->nb_rows + row0 / BLK) * (int64_t)s64 * (HL * BLK) +
                          row0 % BLK;
  const RowOffsets<RQ, RR> roff = row_offsets<RQ, RR>(wave, lane);
  const uint32_t rdst = lds_addr(sR) + (uint32_t)(wave * 64 * 16);
  if (d >= 3 * 64) {
    uint32_t c1[RQ][1];
    tile_stages<RQ, RR, PB, 1>(arow, cbase, off, rbase, roff, rdst, rpf, s64, sB0, sB1, wave,
                               lane, lane + 3 * 64, c1);
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      cnt[q][0] = cnt[q][1] = cnt[q][2] = 0;
      cnt[q][3] = c1[q][0];
    }
  } else {
    uint32_t c2[RQ][2];
    tile_stages<RQ, RR, PB, 2>(arow, cbase, off, rbase, roff, rdst, rpf, s64, sB0, sB1, wave,
                               lane, lane + 2 * 64, c2);
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      cnt[q][0] = cnt[q][1] = 0;
      cnt[q][2] = c2[q][0];
      cnt[q][3] = c2[q][1];
    }
  }
  return true;
}

// SKIP: diagonal tiles of a self pass go through tile_match_diagonal (only
// the kernels instantiated with it carry that code).
template <int RQ, int RR, int PB, bool SKIP = false>
__device__ __forceinline__ void tile_match(CParams* pp, Tile tile, int kk, uint4* sB0,
                                           uint4* sB1, uint4* sR, int wave, int lane,
                                           uint32_t (&cnt)[RQ][RR]) {
  if constexpr (SKIP) {
    if (tile_match_diagonal<RQ, RR, PB>(pp, tile, kk, sB0, sB1, sR, wave, lane, cnt)) return;
  }
  using Cfg = TileCfg<RQ, RR>;
  constexpr int TC = Cfg::TC;
  constexpr int WS = Cfg::WS;
  constexpr int TROWS = Cfg::TROWS;
  const int s64 = pp->s64;

  // row operands: RQ consecutive samples of this wave, wave-uniform address,
  // read through the constant address space so they are scalar loads into
  // SGPRs (the stripes are never written while a distance kernel runs)
  const int64_t row0 = (int64_t)tile.bi * TROWS;
  const CU32* arow = (CU32*)(pp->row_stripes +
                             ((int64_t)kk * pp->nb_rows + row0 / BLK) * (int64_t)s64 * (HL * BLK) +
                             row0 % BLK + (wave / Cfg::CG) * RQ);
  // the lane's first column within the tile
  const int col_in_tile = (wave % Cfg::CG) * 64 * RR + lane;
  // column slabs of this k-mer length.  The last stage may run past word
  // s64 - 1: the stripe buffer carries zero padding at its end, and those
  // words are never computed on.
  const int64_t col0 = (int64_t)tile.ct * TC;
  const int nb_cols = pp->nb_cols;
  const uint32_t* __restrict__ col_kk =
      pp->col_stripes + (int64_t)kk * nb_cols * (int64_t)s64 * (HL * BLK);

#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int r = 0; r < RR; ++r) cnt[q][r] = 0;

  const int n_stages = (s64 + WS - 1) / WS;
  const ColOffsets<RQ, RR> off = col_offsets<RQ, RR>(col0, nb_cols, s64, wave, lane);
  const uint32_t* cbase = col_kk + (int64_t)(col0 / BLK) * s64 * (HL * BLK);
  const bool rpf = PB != 0 && pp->row_prefetch != 0;
  const uint32_t* rbase = pp->row_stripes +
                          ((int64_t)kk * pp->nb_rows + row0 / BLK) * (int64_t)s64 * (HL * BLK) +
                          row0 % BLK;
  const RowOffsets<RQ, RR> roff = row_offsets<RQ, RR>(wave, lane);
  const uint32_t rdst = lds_addr(sR) + (uint32_t)(wave * 64 * 16);
  issue_stage<RQ, RR>(cbase, off, sB0, wave);
  if (rpf) prefetch_round<RQ, RR, 0>(rbase, roff, rdst, wave);
  // One barrier per stage: after it, stage s is in LDS for every wave and
  // every wave has finished matching stage s - 1, whose buffer the copy of
  // stage s + 1 (issued right after the barrier) then overwrites while stage
  // s is matched.  Stages alternate between the two buffers; each pass of
  // this loop runs an even stage from sB0 and the following odd stage from
  // sB1.
  for (int st = 0; st < n_stages; st += 2) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int s = st + half;
      if (s < n_stages) {
        uint4* cur = half ? sB1 : sB0;
        uint4* nxt = half ? sB0 : sB1;
        const int w0 = s * WS;
        wait_vmcnt<0>();
        stage_barrier();
        if (s + 1 < n_stages) {
          issue_stage<RQ, RR>(cbase + (int64_t)(w0 + WS) * (HL * BLK), off, nxt, wave);
          if (rpf) prefetch_round<RQ, RR, 0>(rbase + (int64_t)(w0 + WS) * (HL * BLK), roff, rdst, wave);
        }
        const int nw = (s64 - w0) < WS ? (s64 - w0) : WS;
        const uint32_t* sB = reinterpret_cast<const uint32_t*>(cur);
        if constexpr (PB == 0) {
          match_stage_plain<RQ, RR>(arow, w0, nw, sB, lane, cnt);
        } else {
          match_stage_pipe<RQ, RR, PB>(arow, w0, nw, cur, col_in_tile, cnt);
        }
      }
    }
  }
  // every wave is done with both buffers before they are reused
  stage_barrier();
}

// Counts of this wave / lane into a tile count block [n_k][rows][cols].
template <int RQ, int RR>
__device__ __forceinline__ void store_counts(uint16_t* tile_counts, int kk, int wave, int lane,
                                             const uint32_t (&cnt)[RQ][RR]) {
  using Cfg = TileCfg<RQ, RR>;
  constexpr int TC = Cfg::TC;
  uint16_t* cbase = tile_counts + ((int64_t)kk * Cfg::TROWS + (wave / Cfg::CG) * RQ) * TC +
                    (wave % Cfg::CG) * 64 * RR + lane;
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int r = 0; r < RR; ++r) cbase[q * TC + 64 * r] = (uint16_t)cnt[q][r];
}

// ---------------------------------------------------------------------------
// The tile kernel: one (tile, k-mer length) item per workgroup.
// ---------------------------------------------------------------------------
template <int RQ, int RR, int PB>
__global__ __launch_bounds__(WAVES * 64, 4) void dist_tiles(DistParams p) {
  using Cfg = TileCfg<RQ, RR>;
  __shared__ uint4 sB0[Cfg::STAGE_U4];
  __shared__ uint4 sB1[Cfg::STAGE_U4];
  __shared__ uint4 sR[PB ? Cfg::ROW_LDS_U4 : 64];  // row prefetch: pipelined shapes only
  __shared__ int s_last;
  __shared__ int s_stop;
  // the parameters are the kernel's only argument: scalar loads straight
  // from the kernarg segment (no copy)
  CParams* pp = (CParams*)__builtin_amdgcn_kernarg_segment_ptr();

  const int64_t item = p.item_base + blockIdx.x;
  const int64_t t = item / p.n_k;
  const int kk = (int)(item - t * p.n_k);
  Tile tile = p.tiles[t];
  // wave-uniform values (kept in SGPRs)
  tile.bi = __builtin_amdgcn_readfirstlane(tile.bi);
  tile.ct = __builtin_amdgcn_readfirstlane(tile.ct);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // a cancelled call skips the matching; one read per workgroup, shared, so
  // that every wave takes the same branch (the matching has barriers)
  if (threadIdx.x == 0) s_stop = cancel_seen(p.ctl.cancel) ? 1 : 0;
  __syncthreads();
  const int stop = __builtin_amdgcn_readfirstlane(s_stop);

  // per-tile count scratch: [tile][kk][row][col]
  uint16_t* tile_counts = p.counts + t * p.n_k * (int64_t)(Cfg::TROWS * Cfg::TC);
  if (!stop) {
    uint32_t cnt[RQ][RR];
    tile_match<RQ, RR, PB>(pp, tile, kk, sB0, sB1, sR, wave, lane, cnt);
    store_counts<RQ, RR>(tile_counts, kk, wave, lane, cnt);
  }

  // completion: the last workgroup of this tile runs the epilogue (unless
  // the call was cancelled, which then holds for every later read)
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = atomicAdd(&p.tile_done[t], 1u);
    const bool last = prev == (unsigned)(p.n_k - 1);
    s_last = last ? (cancel_seen(p.ctl.cancel) ? 2 : 1) : 0;
    if (last) p.tile_done[t] = 0;  // ready for the next launch
  }
  __syncthreads();
  if (s_last != 1) return;
  __threadfence();
  tile_epilogue<RQ, RR>(pp, tile_counts, tile, wave, lane);
  if (threadIdx.x == 0) add_progress(p.ctl, (unsigned long long)p.n_k);
}

// dist_tiles<8, 4, 2> for the (tile, k) items of a self pass's diagonal tiles:
// the column blocks that hold no output pair are skipped (tile_match_diagonal);
// any other tile runs as in dist_tiles.  A kernel of its own, so that the code
// of dist_tiles itself does not change with it.
__global__ __launch_bounds__(WAVES * 64, 4) void dist_tiles_diagonal(DistParams p) {
  constexpr int RQ = 8, RR = 4, PB = 2;
  using Cfg = TileCfg<RQ, RR>;
  __shared__ uint4 sB0[Cfg::STAGE_U4];
  __shared__ uint4 sB1[Cfg::STAGE_U4];
  __shared__ uint4 sR[PB ? Cfg::ROW_LDS_U4 : 64];  // row prefetch: pipelined shapes only
  __shared__ int s_last;
  __shared__ int s_stop;
  // the parameters are the kernel's only argument: scalar loads straight
  // from the kernarg segment (no copy)
  CParams* pp = (CParams*)__builtin_amdgcn_kernarg_segment_ptr();

  const int64_t item = p.item_base + blockIdx.x;
  const int64_t t = item / p.n_k;
  const int kk = (int)(item - t * p.n_k);
  Tile tile = p.tiles[t];
  // wave-uniform values (kept in SGPRs)
  tile.bi = __builtin_amdgcn_readfirstlane(tile.bi);
  tile.ct = __builtin_amdgcn_readfirstlane(tile.ct);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // a cancelled call skips the matching; one read per workgroup, shared, so
  // that every wave takes the same branch (the matching has barriers)
  if (threadIdx.x == 0) s_stop = cancel_seen(p.ctl.cancel) ? 1 : 0;
  __syncthreads();
  const int stop = __builtin_amdgcn_readfirstlane(s_stop);

  // per-tile count scratch: [tile][kk][row][col]
  uint16_t* tile_counts = p.counts + t * p.n_k * (int64_t)(Cfg::TROWS * Cfg::TC);
  if (!stop) {
    uint32_t cnt[RQ][RR];
    tile_match<RQ, RR, PB, true>(pp, tile, kk, sB0, sB1, sR, wave, lane, cnt);
    store_counts<RQ, RR>(tile_counts, kk, wave, lane, cnt);
  }

  // completion: the last workgroup of this tile runs the epilogue (unless
  // the call was cancelled, which then holds for every later read)
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = atomicAdd(&p.tile_done[t], 1u);
    const bool last = prev == (unsigned)(p.n_k - 1);
    s_last = last ? (cancel_seen(p.ctl.cancel) ? 2 : 1) : 0;
    if (last) p.tile_done[t] = 0;  // ready for the next launch
  }
  __syncthreads();
  if (s_last != 1) return;
  __threadfence();
  tile_epilogue<RQ, RR>(pp, tile_counts, tile, wave, lane);
  if (threadIdx.x == 0) add_progress(p.ctl, (unsigned long long)p.n_k);
}

// ---------------------------------------------------------------------------
// Persistent batch kernel: many query-vs-ref jobs (each with its own sets,
// k-mer lengths, sketch size and output) in one launch.  Work items are
// whole tiles (every k-mer length of the tile, then its epilogue) taken from
// NQ queues; workgroup b starts on queue b % NQ (the XCD it is dispatched
// to) and, once that queue is drained, moves on to the next ones in turn,
// never returning to a drained queue; it exits after the last one.  Counts
// go to the workgroup's own scratch block and are read back by the same
// lanes in the epilogue, so no cross-workgroup synchronisation is needed.
// ---------------------------------------------------------------------------
// LDS of the persistent kernels' matching: one set, sized for the largest
// tile shape (64 x 256) and shared by every shape a kernel runs, so a kernel
// that mixes shapes holds it once.  Every batch shape stages the same
// 28 KiB column stage (4 / RR words of 64 RR columns); the row prefetch
// area grows with the tile rows.
__shared__ uint4 g_batch_sB0[TileCfg<8, 4>::STAGE_U4];
__shared__ uint4 g_batch_sB1[TileCfg<8, 4>::STAGE_U4];
__shared__ uint4 g_batch_sR[TileCfg<8, 4>::ROW_LDS_U4];

// One (tile, k-mer length) of a batch item: counts into the workgroup's
// scratch block.  Out of line so that the matching loop is register-allocated
// on its own, exactly as in dist_tiles, instead of around the persistent
// loop's state (which would push the counts out to scratch memory on every
// word).  Arguments arrive in VGPRs and are made wave-uniform again here.
template <int RQ, int RR, int PB>
__device__ __noinline__ void batch_tile_k(CParams* pp_in, int bi_in, int ct_in, int kk_in,
                                          uint16_t* scratch_in) {
  using Cfg = TileCfg<RQ, RR>;
  static_assert(Cfg::STAGE_U4 == TileCfg<8, 4>::STAGE_U4 &&
                    Cfg::ROW_LDS_U4 <= TileCfg<8, 4>::ROW_LDS_U4,
                "batch shapes share the LDS stage buffers");
  uint4* sB0 = g_batch_sB0;
  uint4* sB1 = g_batch_sB1;
  uint4* sR = g_batch_sR;
  const uint64_t pv = (uint64_t)(uintptr_t)pp_in;
  const uint32_t plo = __builtin_amdgcn_readfirstlane((uint32_t)pv);
  const uint32_t phi = __builtin_amdgcn_readfirstlane((uint32_t)(pv >> 32));
  CParams* pp = (CParams*)(uintptr_t)(((uint64_t)phi << 32) | plo);
  const uint64_t sv = (uint64_t)(uintptr_t)scratch_in;
  const uint32_t slo = __builtin_amdgcn_readfirstlane((uint32_t)sv);
  const uint32_t shi = __builtin_amdgcn_readfirstlane((uint32_t)(sv >> 32));
  uint16_t* scratch = (uint16_t*)(uintptr_t)(((uint64_t)shi << 32) | slo);
  Tile tile;
  tile.bi = __builtin_amdgcn_readfirstlane(bi_in);
  tile.ct = __builtin_amdgcn_readfirstlane(ct_in);
  const int kk = __builtin_amdgcn_readfirstlane(kk_in);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  uint32_t cnt[RQ][RR];
  tile_match<RQ, RR, PB>(pp, tile, kk, sB0, sB1, sR, wave, lane, cnt);
  store_counts<RQ, RR>(scratch, kk, wave, lane, cnt);
}

template <int RQ, int RR, int PB>
__global__ __launch_bounds__(WAVES * 64, 4) void dist_batch(BatchArgs a) {
  using Cfg = TileCfg<RQ, RR>;
  __shared__ int64_t s_item;
  __shared__ int s_tried;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int home = (int)(blockIdx.x % BATCH_QUEUES);
  uint16_t* scratch =
      a.scratch + (int64_t)blockIdx.x * a.max_k * (int64_t)(Cfg::TROWS * Cfg::TC);
  int tried = 0;  // queues found drained so far (workgroup-uniform)
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t got = -1;
      if (cancel_seen(a.ctl.cancel)) tried = BATCH_QUEUES;  // cancelled: take no more items
      while (tried < BATCH_QUEUES) {
        const int q = (home + tried) % BATCH_QUEUES;
        const int64_t size = a.queue_begin[q + 1] - a.queue_begin[q];
        const unsigned idx = atomicAdd(&a.queue_next[q], 1u);
        if ((int64_t)idx < size) {
          got = a.queue_begin[q] + idx;
          break;
        }
        ++tried;
      }
      s_item = got;
      s_tried = tried;
    }
    __syncthreads();
    const int64_t it = s_item;
    tried = s_tried;
    __syncthreads();  // s_item is rewritten only after every thread read it
    if (it < 0) break;
    const BatchItem bi = a.items[it];
    const int job = __builtin_amdgcn_readfirstlane(bi.job);
    Tile tile;
    tile.bi = __builtin_amdgcn_readfirstlane(bi.bi);
    tile.ct = __builtin_amdgcn_readfirstlane(bi.ct);
    CParams* pp = (CParams*)(a.jobs + job);
    const int n_k = pp->n_k;
    for (int kk = 0; kk < n_k; ++kk) batch_tile_k<RQ, RR, PB>(pp, tile.bi, tile.ct, kk, scratch);
    tile_epilogue<RQ, RR>(pp, scratch, tile, wave, lane);
    if (threadIdx.x == 0) add_progress(a.ctl, (unsigned long long)n_k);
  }
}

// One whole tile of a batch item in shape RQ x RR: every k-mer length into
// the workgroup's scratch block, then the epilogue, then progress.  Out of
// line: the mixed-shape kernel's frame then only dispatches (a few of its
// values are saved around the calls).
template <int RQ, int RR, int PB>
__device__ __noinline__ void batch_tile(CParams* pp, Tile tile, uint16_t* scratch, Control ctl,
                                        int wave, int lane) {
  const int n_k = pp->n_k;
  for (int kk = 0; kk < n_k; ++kk) batch_tile_k<RQ, RR, PB>(pp, tile.bi, tile.ct, kk, scratch);
  tile_epilogue<RQ, RR>(pp, scratch, tile, wave, lane);
  if (threadIdx.x == 0) add_progress(ctl, (unsigned long long)n_k);
}

// The persistent kernel of dist_batch with a tile shape per item (item.pad:
// 2 = 64 x 256, 4 = 32 x 256, 5 = 16 x 256), so that each job of a batch can
// cover its rows with the tile heights that waste the least of its work on
// padding rows.  The shapes share the LDS stage buffers; a workgroup's
// scratch block holds the counts of the largest shape.
__global__ __launch_bounds__(WAVES * 64, 4) void dist_batch_mixed(BatchArgs a) {
  __shared__ int64_t s_item;
  __shared__ int s_tried;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int home = (int)(blockIdx.x % BATCH_QUEUES);
  uint16_t* scratch = a.scratch + (int64_t)blockIdx.x * a.max_k *
                                      (int64_t)(TileCfg<8, 4>::TROWS * TileCfg<8, 4>::TC);
  int tried = 0;  // queues found drained so far (workgroup-uniform)
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t got = -1;
      if (cancel_seen(a.ctl.cancel)) tried = BATCH_QUEUES;  // cancelled: take no more items
      while (tried < BATCH_QUEUES) {
        const int q = (home + tried) % BATCH_QUEUES;
        const int64_t size = a.queue_begin[q + 1] - a.queue_begin[q];
        const unsigned idx = atomicAdd(&a.queue_next[q], 1u);
        if ((int64_t)idx < size) {
          got = a.queue_begin[q] + idx;
          break;
        }
        ++tried;
      }
      s_item = got;
      s_tried = tried;
    }
    __syncthreads();
    const int64_t it = s_item;
    tried = s_tried;
    __syncthreads();  // s_item is rewritten only after every thread read it
    if (it < 0) break;
    const BatchItem bi = a.items[it];
    const int job = __builtin_amdgcn_readfirstlane(bi.job);
    const int shape = __builtin_amdgcn_readfirstlane(bi.pad);
    Tile tile;
    tile.bi = __builtin_amdgcn_readfirstlane(bi.bi);
    tile.ct = __builtin_amdgcn_readfirstlane(bi.ct);
    CParams* pp = (CParams*)(a.jobs + job);
    if (shape == 5) {
      batch_tile<8, 1, 2>(pp, tile, scratch, a.ctl, wave, lane);
    } else if (shape == 4) {
      batch_tile<8, 2, 2>(pp, tile, scratch, a.ctl, wave, lane);
    } else {
      batch_tile<8, 4, 2>(pp, tile, scratch, a.ctl, wave, lane);
    }
  }
}

__global__ void batch_reset_kernel(uint32_t* queue_next, unsigned long long* fail, int n_jobs,
                                   unsigned long long* progress) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < BATCH_QUEUES) queue_next[x] = 0u;
  if (x == 0 && progress) *progress = 0ull;
  if (fail && x < n_jobs) {
    fail[2 * x] = 0ull;
    fail[2 * x + 1] = ~0ull;
  }
}

bool variant_valid(int v) { return v >= 1 && v <= N_VARIANTS; }

int variant_tile_rows(int v) {
  switch (v) {
    case 1:
    case 3:
      return 128;
    case 4:
      return 32;
    case 5:
      return 16;
    default:
      return 64;
  }
}

int variant_tile_cols(int v) {
  switch (v) {
    case 1:
      return 64;
    case 3:
      return 128;
    default:
      return 256;
  }
}

hipError_t launch_dists(const DistParams& p, int variant, int64_t n_items, hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  const unsigned g = (unsigned)n_items;
  switch (variant) {
    case 1:
      dist_tiles<16, 1, 0><<<g, WAVES * 64, 0, stream>>>(p);
      break;
    case 3:
      dist_tiles<16, 2, 0><<<g, WAVES * 64, 0, stream>>>(p);
      break;
    case 4:
      dist_tiles<8, 2, 2><<<g, WAVES * 64, 0, stream>>>(p);
      break;
    case 5:
      dist_tiles<8, 1, 2><<<g, WAVES * 64, 0, stream>>>(p);
      break;
    default:
      dist_tiles<8, 4, 2><<<g, WAVES * 64, 0, stream>>>(p);
      break;
  }
  return hipGetLastError();
}

// Batch kernels exist for the pipelined shapes only; BATCH_MIXED runs the
// shape each item names.
template <typename F>
static hipError_t with_batch_kernel(int variant, F f) {
  switch (variant) {
    case 4:
      return f(dist_batch<8, 2, 2>);
    case 2:
      return f(dist_batch<8, 4, 2>);
    case BATCH_MIXED:
      return f(dist_batch_mixed);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t batch_grid(int variant, int* grid) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  return with_batch_kernel(variant, [&](auto kern) {
    int per_cu = 0;
    hipError_t r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WAVES * 64, 0);
    if (r != hipSuccess) return r;
    *grid = (per_cu > 0 ? per_cu : 1) * cus;
    return hipSuccess;
  });
}

hipError_t launch_batch(const BatchArgs& a, int variant, int grid, hipStream_t stream) {
  if (grid <= 0) return hipErrorInvalidValue;
  return with_batch_kernel(variant, [&](auto kern) {
    kern<<<grid, WAVES * 64, 0, stream>>>(a);
    return hipGetLastError();
  });
}

hipError_t launch_batch_reset(uint32_t* queue_next, unsigned long long* fail, int n_jobs,
                              hipStream_t stream, unsigned long long* progress) {
  const int n = n_jobs > BATCH_QUEUES ? n_jobs : BATCH_QUEUES;
  batch_reset_kernel<<<(n + 255) / 256, 256, 0, stream>>>(queue_next, fail, n_jobs, progress);
  return hipGetLastError();
}

__global__ void tile_reset_kernel(uint32_t* tile_done, int64_t n, unsigned long long* fail) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) tile_done[x] = 0u;
  if (x == 0 && fail) {
    fail[0] = 0ull;
    fail[1] = ~0ull;
  }
}

hipError_t launch_dists_diagonal(const DistParams& p, int64_t n_items, hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  dist_tiles_diagonal<<<(unsigned)n_items, WAVES * 64, 0, stream>>>(p);
  return hipGetLastError();
}

hipError_t upload_log_table() {
  double tab[LOG_TAB_DOUBLES];
  log_table_fill(tab);
  return hipMemcpyToSymbol(HIP_SYMBOL(g_log_tab), tab, sizeof(tab));
}

hipError_t launch_tile_reset(uint32_t* tile_done, int64_t n, unsigned long long* fail,
                             hipStream_t stream) {
  const int64_t m = n > 1 ? n : 1;
  tile_reset_kernel<<<(unsigned)((m + 255) / 256), 256, 0, stream>>>(tile_done, n, fail);
  return hipGetLastError();
}

}  // namespace ppsk
