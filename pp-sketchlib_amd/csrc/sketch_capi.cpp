// Sketching entry point (include/ppsk.h, ppsk_sketch): work units, device
// buffers and the launch of sketch_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "capi_common.h"
#include "sketch_kernels.h"

using namespace ppsk::capi;

namespace {

constexpr int64_t UNIT = 1024;          // k-mer start positions per thread
constexpr int64_t GROUP_UNITS = 1024;   // units per workgroup

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

}  // namespace

extern "C" {

int ppsk_sketch(int device_id, int64_t n_samples, const char* seq, int64_t seq_bytes,
                const int64_t* contig_off, int64_t n_contigs, const int64_t* sample_contig,
                int32_t n_k, const int32_t* klist, int32_t sketchsize64, int32_t use_rc,
                uint64_t* raw_out, double* inv_minhash, int32_t* densified) {
  if (n_samples < 1 || !sample_contig || !raw_out) {
    return fail(PPSK_ERR_ARG, "sketching needs at least one sample and an output buffer");
  }
  if (n_contigs < 0 || (n_contigs > 0 && (!contig_off || !seq))) {
    return fail(PPSK_ERR_ARG, "contig offsets / sequence missing");
  }
  if (n_k < 1 || n_k > ppsk::MAXK || !klist) {
    return fail(PPSK_ERR_ARG, "n_k must be in [1, " + std::to_string(ppsk::MAXK) + "]");
  }
  for (int i = 0; i < n_k; ++i) {
    if (klist[i] < 3) return fail(PPSK_ERR_ARG, "Minimum k must be 3 or higher");
    if (klist[i] > ppsk::SKETCH_MAX_K) {
      return fail(PPSK_ERR_ARG, "k must be at most " + std::to_string(ppsk::SKETCH_MAX_K));
    }
    if (i && klist[i] <= klist[i - 1]) return fail(PPSK_ERR_ARG, "klist must be strictly ascending");
  }
  if (sketchsize64 < 1 || sketchsize64 > 1023) {
    return fail(PPSK_ERR_ARG, "sketchsize64 must be in [1, 1023]");
  }
  if (sample_contig[0] != 0 || sample_contig[n_samples] != n_contigs) {
    return fail(PPSK_ERR_ARG, "sample_contig must run from 0 to n_contigs");
  }
  for (int64_t s = 0; s < n_samples; ++s) {
    if (sample_contig[s + 1] < sample_contig[s]) return fail(PPSK_ERR_ARG, "sample_contig not ascending");
  }
  for (int64_t c = 0; c < n_contigs; ++c) {
    if (contig_off[c + 1] < contig_off[c] || contig_off[c + 1] > seq_bytes || contig_off[c] < 0) {
      return fail(PPSK_ERR_ARG, "contig offsets outside the sequence buffer");
    }
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);

  // units (per k: every contig cut into runs of UNIT windows) and groups
  // (runs of units of one sample and one k)
  std::vector<ppsk::SketchUnit> units;
  std::vector<ppsk::SketchGroup> groups;
  for (int kk = 0; kk < n_k; ++kk) {
    const int64_t k = klist[kk];
    for (int64_t s = 0; s < n_samples; ++s) {
      const int64_t u0 = (int64_t)units.size();
      for (int64_t ci = sample_contig[s]; ci < sample_contig[s + 1]; ++ci) {
        const int64_t windows = contig_off[ci + 1] - contig_off[ci] - k + 1;
        for (int64_t w = 0; w < windows; w += UNIT) {
          units.push_back({contig_off[ci] + w, (int32_t)std::min(UNIT, windows - w), 0});
        }
      }
      const int64_t u1 = (int64_t)units.size();
      for (int64_t g = u0; g < u1; g += GROUP_UNITS) {
        groups.push_back({g, std::min(u1, g + GROUP_UNITS), (int32_t)s, kk});
      }
    }
  }

  const int64_t nbins = (int64_t)sketchsize64 * 64;
  const uint64_t sign_mod = (1ULL << 61) - 1ULL;
  ppsk::SketchArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n_k = n_k;
  a.s64 = sketchsize64;
  a.use_rc = use_rc ? 1 : 0;
  a.n_samples = n_samples;
  a.binsize = (sign_mod + (uint64_t)nbins - 1ULL) / (uint64_t)nbins;
  a.bin_recip = ~0ULL / a.binsize;

  DevBuf d_seq, d_units, d_groups, d_k, d_signs, d_raw, d_mh, d_dens, d_next;
  const size_t raw_words = (size_t)n_samples * n_k * sketchsize64 * ppsk::BBITS_SK;
  // base codes, padded both sides for the kernel's stream loads
  hipError_t e =
      hipMalloc(&d_seq.p, (size_t)(ppsk::CODE_PAD_FRONT + seq_bytes + ppsk::CODE_PAD_BACK));
  if (e == hipSuccess) e = hipMalloc(&d_units.p, std::max<size_t>(units.size(), 1) * sizeof(ppsk::SketchUnit));
  if (e == hipSuccess) e = hipMalloc(&d_groups.p, std::max<size_t>(groups.size(), 1) * sizeof(ppsk::SketchGroup));
  if (e == hipSuccess) e = hipMalloc(&d_k.p, n_k * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&d_signs.p, (size_t)n_samples * n_k * nbins * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&d_raw.p, raw_words * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&d_mh.p, (size_t)n_samples * n_k * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&d_dens.p, (size_t)n_samples * n_k * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&d_next.p, sizeof(int));
  if (e != hipSuccess) {
    return fail(PPSK_ERR_MEMORY, std::string("sketch buffers: ") + hipGetErrorString(e));
  }
  uint8_t* codes = (uint8_t*)d_seq.p + ppsk::CODE_PAD_FRONT;
  if (seq_bytes > 0) {
    HIP_TRY(hipMemcpyAsync(codes, seq, seq_bytes, hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(d_units.p, units.data(), units.size() * sizeof(ppsk::SketchUnit),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_groups.p, groups.data(), groups.size() * sizeof(ppsk::SketchGroup),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_k.p, klist, n_k * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  a.codes = codes;
  a.units = (const ppsk::SketchUnit*)d_units.p;
  a.groups = (const ppsk::SketchGroup*)d_groups.p;
  a.klist = (const int32_t*)d_k.p;
  a.signs = (uint64_t*)d_signs.p;
  a.raw = (uint64_t*)d_raw.p;
  a.inv_minhash = (double*)d_mh.p;
  a.densified = (int32_t*)d_dens.p;

  // bin minima in LDS when one sample's bins fit a workgroup's LDS
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_id));
  // (the step table takes another 400 bytes)
  const bool lds_bins = (size_t)nbins * sizeof(uint64_t) + 512 <= prop.sharedMemPerBlock;
  if (a.bin_recip >> 32) return fail(PPSK_ERR_ARG, "bin reciprocal out of range");
  hipEvent_t e0 = pool_event(c, 0), e1 = pool_event(c, 1);
  if (!e0 || !e1) return fail(PPSK_ERR_HIP, "event creation failed");
  HIP_TRY(hipEventRecord(e0, c->stream));
  HIP_TRY(ppsk::launch_sketch_decode(codes, seq_bytes, c->stream));
  HIP_TRY(ppsk::launch_sketch(a, (int64_t)groups.size(), lds_bins, (int*)d_next.p, c->stream));
  HIP_TRY(hipEventRecord(e1, c->stream));
  HIP_TRY(hipMemcpyAsync(raw_out, d_raw.p, raw_words * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         c->stream));
  std::vector<double> mh((size_t)n_samples * n_k);
  std::vector<int32_t> dn((size_t)n_samples * n_k);
  HIP_TRY(hipMemcpyAsync(mh.data(), d_mh.p, mh.size() * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(dn.data(), d_dens.p, dn.size() * sizeof(int32_t),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  g_last_ms = ms;
  g_last_launches = 1;
  g_last_items = (int64_t)units.size();
  for (int64_t s = 0; s < n_samples; ++s) {
    double sum = 0.0;
    int32_t d = 0;
    for (int kk = 0; kk < n_k; ++kk) {
      sum += mh[(size_t)s * n_k + kk];
      d |= dn[(size_t)s * n_k + kk] != 0 ? 1 : 0;
    }
    if (inv_minhash) inv_minhash[s] = sum;
    if (densified) densified[s] = d;
  }
  return PPSK_OK;
}

}  // extern "C"
