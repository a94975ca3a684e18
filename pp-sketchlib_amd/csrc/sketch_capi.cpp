// Sketching entry points (include/ppsk.h, ppsk_sketch / ppsk_sketch_ex):
// work units, device buffers and the launches of sketch_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "capi_common.h"
#include "sketch_kernels.h"

using namespace ppsk::capi;

namespace {

constexpr int64_t UNIT = 1024;          // k-mer start positions per lane
constexpr int64_t GROUP_UNITS = 1024;   // units per workgroup task

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

// Units of one sample's contigs for one k: runs of up to UNIT window starts,
// pad = 1 on the first unit of a contig; wbase = index of the unit's first
// window among the sample's windows (sequence order).  Returns the window
// count.
int64_t sample_units(const int64_t* contig_off, int64_t c0, int64_t c1, int64_t k,
                     std::vector<ppsk::SketchUnit>& units, std::vector<int64_t>* wbase) {
  int64_t nw = 0;
  for (int64_t ci = c0; ci < c1; ++ci) {
    const int64_t windows = contig_off[ci + 1] - contig_off[ci] - k + 1;
    for (int64_t w = 0; w < windows; w += UNIT) {
      units.push_back({contig_off[ci] + w, (int32_t)std::min(UNIT, windows - w), w == 0 ? 1 : 0});
      if (wbase) wbase->push_back(nw + w);
    }
    if (windows > 0) nw += windows;
  }
  return nw;
}

int sketch_impl(int device_id, int64_t n_samples, const char* seq, int64_t seq_bytes,
                const int64_t* contig_off, int64_t n_contigs, const int64_t* sample_contig,
                int32_t n_k, const int32_t* klist, int32_t sketchsize64, int32_t use_rc,
                int32_t codon, int32_t min_count, int32_t exact, uint64_t* raw_out,
                double* inv_minhash, int32_t* densified) {
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
    if ((codon ? 3 * klist[i] : klist[i]) > ppsk::SKETCH_MAX_K) {
      return fail(PPSK_ERR_ARG, "k must be at most " +
                                    std::to_string(ppsk::SKETCH_MAX_K / (codon ? 3 : 1)));
    }
    if (i && klist[i] <= klist[i - 1]) return fail(PPSK_ERR_ARG, "klist must be strictly ascending");
  }
  if (sketchsize64 < 1 || sketchsize64 > 1023) {
    return fail(PPSK_ERR_ARG, "sketchsize64 must be in [1, 1023]");
  }
  if (min_count < 0 || min_count > 255) return fail(PPSK_ERR_ARG, "min_count must be in [0, 255]");
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
  // windows span k bases, or 3k - 2 with codon-phased seeds
  auto span = [codon](int32_t k) -> int64_t { return codon ? 3 * (int64_t)k - 2 : k; };
  const bool filtered = min_count > 0;
  if (filtered) {
    for (int64_t s = 0; s < n_samples; ++s) {
      int64_t w = 0;
      for (int64_t ci = sample_contig[s]; ci < sample_contig[s + 1]; ++ci) {
        w += std::max<int64_t>(0, contig_off[ci + 1] - contig_off[ci] - span(klist[0]) + 1);
      }
      if (w > ppsk::READS_MAX_WINDOWS) {
        return fail(PPSK_ERR_ARG, "a sample has more than " +
                                      std::to_string(ppsk::READS_MAX_WINDOWS) +
                                      " k-mer windows for count filtering");
      }
    }
  }
  DeviceCtx* c = nullptr;
  if (int rc = get_ctx(device_id, &c)) return rc;
  std::lock_guard<std::mutex> lk(c->mu);

  // unfiltered: units (per k, every contig cut into runs of UNIT windows) and
  // groups (runs of units of one sample and one k) for one hashing launch
  std::vector<ppsk::SketchUnit> units;
  std::vector<ppsk::SketchGroup> groups;
  if (!filtered) {
    for (int kk = 0; kk < n_k; ++kk) {
      for (int64_t s = 0; s < n_samples; ++s) {
        const int64_t u0 = (int64_t)units.size();
        sample_units(contig_off, sample_contig[s], sample_contig[s + 1], span(klist[kk]), units,
                     nullptr);
        const int64_t u1 = (int64_t)units.size();
        for (int64_t g = u0; g < u1; g += GROUP_UNITS) {
          groups.push_back({g, std::min(u1, g + GROUP_UNITS), (int32_t)s, kk});
        }
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
  a.codon = codon ? 1 : 0;
  a.n_samples = n_samples;
  a.binsize = (sign_mod + (uint64_t)nbins - 1ULL) / (uint64_t)nbins;
  a.bin_recip = ~0ULL / a.binsize;
  if (a.bin_recip >> 32) return fail(PPSK_ERR_ARG, "bin reciprocal out of range");

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
  if (!units.empty()) {
    HIP_TRY(hipMemcpyAsync(d_units.p, units.data(), units.size() * sizeof(ppsk::SketchUnit),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_groups.p, groups.data(), groups.size() * sizeof(ppsk::SketchGroup),
                           hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(d_k.p, klist, n_k * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  a.codes = codes;
  a.units = (const ppsk::SketchUnit*)d_units.p;
  a.groups = (const ppsk::SketchGroup*)d_groups.p;
  a.klist = (const int32_t*)d_k.p;
  a.signs = (uint64_t*)d_signs.p;
  a.raw = (uint64_t*)d_raw.p;
  a.inv_minhash = (double*)d_mh.p;
  a.densified = (int32_t*)d_dens.p;

  hipEvent_t e0 = pool_event(c, 0), e1 = pool_event(c, 1);
  if (!e0 || !e1) return fail(PPSK_ERR_HIP, "event creation failed");
  HIP_TRY(hipEventRecord(e0, c->stream));
  HIP_TRY(ppsk::launch_sketch_decode(codes, seq_bytes, c->stream));
  int64_t items = (int64_t)units.size();
  if (!filtered) {
    // bin minima in LDS when one sample's bins fit a workgroup's LDS (the
    // step table takes another 400 bytes)
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_id));
    const bool lds_bins = (size_t)nbins * sizeof(uint64_t) + 512 <= prop.sharedMemPerBlock;
    HIP_TRY(ppsk::launch_sketch(a, (int64_t)groups.size(), lds_bins, (int*)d_next.p, c->stream));
  } else {
    // one (sample, k) at a time through the counting pipeline
    HIP_TRY(ppsk::launch_sketch_reset(a, c->stream));
    int64_t max_w = 0, max_u = 0;
    for (int64_t s = 0; s < n_samples; ++s) {
      std::vector<ppsk::SketchUnit> u;
      const int64_t w = sample_units(contig_off, sample_contig[s], sample_contig[s + 1],
                                     span(klist[0]), u, nullptr);
      max_w = std::max(max_w, w);
      max_u = std::max<int64_t>(max_u, (int64_t)u.size());
    }
    DevBuf d_ws, d_ju, d_jw;
    const size_t ws_bytes = ppsk::reads_workspace_bytes(max_w);
    e = hipMalloc(&d_ws.p, ws_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_ju.p, (size_t)std::max<int64_t>(max_u, 1) * sizeof(ppsk::SketchUnit));
    if (e == hipSuccess) e = hipMalloc(&d_jw.p, (size_t)std::max<int64_t>(max_u, 1) * sizeof(int64_t));
    if (e != hipSuccess) {
      return fail(PPSK_ERR_MEMORY, std::string("read filter workspace: ") + hipGetErrorString(e));
    }
    std::vector<ppsk::SketchUnit> ju;
    std::vector<int64_t> jw;
    for (int64_t s = 0; s < n_samples; ++s) {
      for (int kk = 0; kk < n_k; ++kk) {
        ju.clear();
        jw.clear();
        // the smallest k has the most windows: every later job fits max_w / max_u
        const int64_t nw = sample_units(contig_off, sample_contig[s], sample_contig[s + 1],
                                        span(klist[kk]), ju, &jw);
        if (nw == 0) continue;
        HIP_TRY(hipMemcpyAsync(d_ju.p, ju.data(), ju.size() * sizeof(ppsk::SketchUnit),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(d_jw.p, jw.data(), jw.size() * sizeof(int64_t),
                               hipMemcpyHostToDevice, c->stream));
        ppsk::ReadsJob rj;
        rj.codes = codes;
        rj.units = (const ppsk::SketchUnit*)d_ju.p;
        rj.wbase = (const int64_t*)d_jw.p;
        rj.n_units = (int64_t)ju.size();
        rj.n_windows = nw;
        rj.k = klist[kk];
        rj.use_rc = use_rc ? 1 : 0;
        rj.min_count = min_count;
        rj.exact = exact ? 1 : 0;
        rj.codon = codon ? 1 : 0;
        rj.binsize = a.binsize;
        rj.recip = (uint32_t)a.bin_recip;
        rj.bins = a.signs + (s * n_k + kk) * nbins;
        HIP_TRY(ppsk::run_reads_job(rj, d_ws.p, ws_bytes, c->stream));
        // the unit lists are rewritten by the next job
        HIP_TRY(hipStreamSynchronize(c->stream));
        items += (int64_t)ju.size();
      }
    }
    HIP_TRY(ppsk::launch_sketch_finish(a, c->stream));
  }
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
  g_last_items = items;
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

}  // namespace

extern "C" {

int ppsk_sketch(int device_id, int64_t n_samples, const char* seq, int64_t seq_bytes,
                const int64_t* contig_off, int64_t n_contigs, const int64_t* sample_contig,
                int32_t n_k, const int32_t* klist, int32_t sketchsize64, int32_t use_rc,
                uint64_t* raw_out, double* inv_minhash, int32_t* densified) {
  return sketch_impl(device_id, n_samples, seq, seq_bytes, contig_off, n_contigs, sample_contig,
                     n_k, klist, sketchsize64, use_rc, 0, 0, 0, raw_out, inv_minhash, densified);
}

int ppsk_sketch_ex(int device_id, int64_t n_samples, const char* seq, int64_t seq_bytes,
                   const int64_t* contig_off, int64_t n_contigs, const int64_t* sample_contig,
                   int32_t n_k, const int32_t* klist, int32_t sketchsize64, int32_t use_rc,
                   int32_t codon_phased, int32_t min_count, int32_t exact, uint64_t* raw_out,
                   double* inv_minhash, int32_t* densified) {
  return sketch_impl(device_id, n_samples, seq, seq_bytes, contig_off, n_contigs, sample_contig,
                     n_k, klist, sketchsize64, use_rc, codon_phased, min_count, exact, raw_out,
                     inv_minhash, densified);
}

}  // extern "C"
