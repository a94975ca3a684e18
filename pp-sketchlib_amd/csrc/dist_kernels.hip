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
//  * a workgroup owns a 128 x TC tile of pairs (TC = 64 * RR) for ONE k-mer
//    length; 8 waves x 16 rows each.
//  * the 16 row samples of a wave are wave-uniform, so their planes are read
//    with scalar loads and sit in SGPRs: each v_bitop3 takes its row operand
//    from an SGPR for free;
//  * column samples are per lane (RR per lane); their planes are staged per
//    group of words through LDS by the whole workgroup (coalesced 16-byte
//    global loads of contiguous stripe slabs) and read back per lane;
//  * per-pair counts live in VGPRs for the whole word loop, then go to a
//    per-tile count scratch;
//  * the last of the n_k workgroups of a tile to finish (device-scope atomic)
//    runs the fused epilogue for the tile: random-match correction, Jaccard,
//    log-Jaccard-vs-k regression, dense output.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>

#include "ppsk_internal.h"

namespace ppsk {

// ---------------------------------------------------------------------------
// Stripe transform: raw reference layout -> stripes.
// raw: [count][n_k][s64 * 14] uint64 (word w, layer l at w*14 + l,
// sketch/sketch.cpp:82-92).  One thread per (sample, kk, w).
// ---------------------------------------------------------------------------
__global__ void stripe_kernel(const uint64_t* __restrict__ raw, int64_t count,
                              int64_t first, int32_t n_k, int32_t s64, int32_t nb,
                              uint32_t* __restrict__ stripes) {
  const int64_t total = count * n_k * s64;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
       x += (int64_t)gridDim.x * blockDim.x) {
    // sample-fastest so that stores of one plane are coalesced
    const int64_t s = x % count;
    const int64_t rest = x / count;
    const int32_t w = (int32_t)(rest % s64);
    const int32_t kk = (int32_t)(rest / s64);
    const uint64_t* src = raw + ((s * n_k + kk) * (int64_t)s64 + w) * BBITS;
    const int64_t n = first + s;
    const int64_t blk = n / BLK, lane = n % BLK;
    uint32_t* dst = stripes + (((int64_t)kk * nb + blk) * s64 + w) * (HL * BLK) + lane;
#pragma unroll
    for (int l = 0; l < BBITS; ++l) {
      const uint64_t v = src[l];
      dst[l * BLK] = (uint32_t)v;
      dst[(BBITS + l) * BLK] = (uint32_t)(v >> 32);
    }
  }
}

hipError_t launch_stripe(const uint64_t* raw, int64_t count, int64_t first,
                         int32_t n_k, int32_t s64, int32_t nb, uint32_t* stripes,
                         hipStream_t stream) {
  const int64_t total = count * n_k * s64;
  if (total == 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  stripe_kernel<<<(unsigned)blocks, 256, 0, stream>>>(raw, count, first, n_k, s64, nb,
                                                        stripes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Distance tiles
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t match_and(uint32_t acc, uint32_t a, uint32_t b) {
  // truth table over (acc, a, b) with src0 = 0xF0, src1 = 0xCC, src2 = 0xAA:
  // acc & ~(a ^ b) = 0x90
  return __builtin_amdgcn_bitop3_b32(acc, a, b, 0x90);
}

// Words staged through LDS per step: WS * 28 planes * TC dwords = 512 uint4
// per plane-row group x 7, i.e. 7 16-byte loads per thread per stage.
template <int RR>
struct TileCfg {
  static constexpr int TC = 64 * RR;
  static constexpr int WS = 8 / RR;
  static constexpr int STAGE_U4 = WS * HL * TC / 4;  // 3584
  static constexpr int U4_PER_THREAD = STAGE_U4 / (WAVES * 64);  // 7
  static_assert(STAGE_U4 % (WAVES * 64) == 0, "stage must split evenly");
};

// Per-pair epilogue: counts of the n_k lengths -> dense output.
// Reference semantics (CPU path):
//  J_obs = c / (64 * s64)                        reference.cpp:222-232
//  J     = max(J_obs - r, 0) * 1 / (1 - r)       dist/dist.hpp:13-24 (double)
//  dists[kk] = (float) J                         linear_regression.cpp:62-71
//  window [first J > tol, last J > tol], LS of log J on k, core/acc
//                                                linear_regression.cpp:16-60
struct PairOut {
  float core, acc;
  bool ok;
};

__device__ __forceinline__ double random_for(const DistParams& p, int kk, int64_t gi,
                                             int64_t gj) {
  if (p.random_mode == RANDOM_EQUAL_BASE) {
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

__device__ __forceinline__ float jaccard_of(uint32_t c, double bins, double r) {
  const double obs = (double)c / bins;
  const double diff = obs > r ? (obs - r) : 0.0;
  return (float)((diff * 1.0) / (1.0 - r));
}

template <int RR>
__global__ __launch_bounds__(WAVES * 64, 4) void dist_tiles(DistParams p) {
  using Cfg = TileCfg<RR>;
  constexpr int TC = Cfg::TC;
  constexpr int WS = Cfg::WS;
  __shared__ uint4 sB4[Cfg::STAGE_U4];
  __shared__ int s_last;
  const uint32_t* sB = reinterpret_cast<const uint32_t*>(sB4);

  const int64_t item = p.item_base + blockIdx.x;
  const int64_t t = item / p.n_k;
  const int kk = (int)(item - t * p.n_k);
  const Tile tile = p.tiles[t];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int s64 = p.s64;

  // row operands: 16 consecutive samples of this wave, wave-uniform address
  const uint32_t* __restrict__ arow =
      p.row_stripes + ((int64_t)kk * p.nb_rows + tile.bi) * (int64_t)s64 * (HL * BLK) +
      wave * RQ;
  // column slab: stripe block of the column tile and the offset inside it
  const int64_t col0 = (int64_t)tile.ct * TC;
  const int cb = (int)(col0 / BLK);
  const int c_in = (int)(col0 % BLK);
  const uint32_t* __restrict__ bsrc =
      p.col_stripes + ((int64_t)kk * p.nb_cols + cb) * (int64_t)s64 * (HL * BLK) + c_in;

  uint32_t cnt[RQ][RR];
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int r = 0; r < RR; ++r) cnt[q][r] = 0;

  // global -> register staging of one stage (WS words of the column slab)
  uint4 stage[Cfg::U4_PER_THREAD];
  auto load_stage = [&](int w0) {
#pragma unroll
    for (int i = 0; i < Cfg::U4_PER_THREAD; ++i) {
      const int x = threadIdx.x + i * (WAVES * 64);   // uint4 index in stage
      const int u = x % (TC / 4);
      const int hl = (x / (TC / 4)) % HL;
      const int ws = x / (HL * TC / 4);
      const int w = w0 + ws;
      if (w < s64) {
        stage[i] = *reinterpret_cast<const uint4*>(bsrc + ((int64_t)w * HL + hl) * BLK + u * 4);
      } else {
        stage[i] = make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (int i = 0; i < Cfg::U4_PER_THREAD; ++i) sB4[threadIdx.x + i * (WAVES * 64)] = stage[i];
  };

  const int n_stages = (s64 + WS - 1) / WS;
  load_stage(0);
  store_stage();
  __syncthreads();
  for (int st = 0; st < n_stages; ++st) {
    const int w0 = st * WS;
    if (st + 1 < n_stages) load_stage(w0 + WS);
    const int nw = (s64 - w0) < WS ? (s64 - w0) : WS;
    for (int ws = 0; ws < nw; ++ws) {
      const uint32_t* __restrict__ aw = arow + (int64_t)(w0 + ws) * (HL * BLK);
      const uint32_t* bw = sB + ws * (HL * TC) + lane;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint32_t acc[RQ][RR];
#pragma unroll
        for (int l = 0; l < BBITS; ++l) {
          const int hl = h * BBITS + l;
          uint32_t b[RR];
#pragma unroll
          for (int r = 0; r < RR; ++r) b[r] = bw[hl * TC + 64 * r];
#pragma unroll
          for (int q = 0; q < RQ; ++q) {
            const uint32_t a = aw[hl * BLK + q];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
              if (l == 0) {
                acc[q][r] = ~(a ^ b[r]);
              } else {
                acc[q][r] = match_and(acc[q][r], a, b[r]);
              }
            }
          }
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q)
#pragma unroll
          for (int r = 0; r < RR; ++r) cnt[q][r] += __builtin_popcount(acc[q][r]);
      }
    }
    __syncthreads();
    if (st + 1 < n_stages) {
      store_stage();
      __syncthreads();
    }
  }

  // per-tile count scratch: [tile][kk][row][col]
  uint16_t* cbase = p.counts + ((t * p.n_k + kk) * TROWS + wave * RQ) * (int64_t)TC + lane;
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int r = 0; r < RR; ++r) cbase[q * TC + 64 * r] = (uint16_t)cnt[q][r];

  // completion: the last workgroup of this tile runs the epilogue
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = atomicAdd(&p.tile_done[t], 1u);
    s_last = (prev == (unsigned)(p.n_k - 1)) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  if (threadIdx.x == 0) p.tile_done[t] = 0;  // ready for the next launch

  const double bins = 64.0 * (double)s64;
  const int64_t gi0 = (int64_t)tile.bi * BLK + wave * RQ;
  const uint16_t* tbase = p.counts + (t * p.n_k * TROWS + wave * RQ) * (int64_t)TC + lane;
  for (int q = 0; q < RQ; ++q) {
    const int64_t gi = gi0 + q;
#pragma unroll
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
        for (int k = 0; k < p.n_k; ++k) o[k] = cp[(int64_t)k * TROWS * TC];
        continue;
      }
      if (p.out_kind == OUT_JACCARD) {
        float* o = reinterpret_cast<float*>(p.out) + pos * p.n_k;
        for (int k = 0; k < p.n_k; ++k) {
          o[k] = jaccard_of(cp[(int64_t)k * TROWS * TC], bins, random_for(p, k, gi, gj));
        }
        continue;
      }
      // core/acc: streaming window fit.  Points before the first valid one are
      // dropped; invalid points after the last valid one are held as pending
      // sums and committed only if a later valid point arrives.
      int first = -1, last = -1;
      double sy = 0.0, sxy = 0.0, py = 0.0, pxy = 0.0;
      for (int k = 0; k < p.n_k; ++k) {
        const float jf = jaccard_of(cp[(int64_t)k * TROWS * TC], bins, random_for(p, k, gi, gj));
        const bool ok = (double)jf > p.tolerance;
        if (!ok && first < 0) continue;
        const double x = (double)p.kvals[k];
        const double y = log((double)jf);
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
      float2* o = reinterpret_cast<float2*>(p.out) + pos;
      *o = make_float2(core, accd);
    }
  }
}

hipError_t launch_dists(const DistParams& p, int rr, int64_t n_items, hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  if (rr == 2) {
    dist_tiles<2><<<(unsigned)n_items, WAVES * 64, 0, stream>>>(p);
  } else {
    dist_tiles<1><<<(unsigned)n_items, WAVES * 64, 0, stream>>>(p);
  }
  return hipGetLastError();
}

}  // namespace ppsk
