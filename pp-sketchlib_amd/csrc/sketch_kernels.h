// Sketching kernels (sketch_kernels.hip): sequences -> bit-sliced sketches.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace ppsk {

constexpr int SKETCH_TB = 512;  // threads per hashing workgroup
constexpr int BBITS_SK = 14;    // bit layers written per bin
constexpr int SKETCH_MAX_K = 255;  // dense k; codon-phased: 3k <= this
// Base codes (one byte per base, 0 A, 1 C, 2 G, 3 T/U, 4 other) are kept
// with this many bytes of code 4 before and after the sequence, so that the
// per-lane 16-byte stream loads never leave the buffer.
constexpr int64_t CODE_PAD_FRONT = SKETCH_MAX_K + 17;
constexpr int64_t CODE_PAD_BACK = 64;

// A run of n k-mer start positions of one contig: windows start at
// seq[pos], seq[pos + 1], ..., seq[pos + n - 1] and never cross the contig.
struct SketchUnit {
  int64_t pos;
  int32_t n;
  int32_t pad;
};

// A hashing workgroup's work: units [unit0, unit1) of one sample, one k.
struct SketchGroup {
  int64_t unit0, unit1;
  int32_t sample;
  int32_t kk;
};

struct SketchArgs {
  const uint8_t* codes;      // base codes; codes[0] is the first base, padded both sides
  const SketchUnit* units;   // per k-mer length index: its own unit list
  const SketchGroup* groups;
  const int32_t* klist;      // [n_k]
  int32_t n_k;
  int32_t s64;
  int32_t use_rc;
  int32_t codon;             // codon-phased seeds (span 3k - 2, every third base)
  int64_t n_samples;
  uint64_t binsize;          // ceil((2^61 - 1) / (64 s64))
  uint64_t bin_recip;        // floor((2^64 - 1) / binsize), < 2^32
  uint64_t* signs;           // [n_samples][n_k][64 s64] bin minima (work)
  uint64_t* raw;             // [n_samples][n_k][s64 * 14] output words
  double* inv_minhash;       // [n_samples][n_k] bin 0 / (2^61 - 1), before densifying
  int32_t* densified;        // [n_samples][n_k] 0, 1, or -1 (every bin empty)
};

// In place: the bytes codes[0, n) become their base codes; the
// CODE_PAD_FRONT bytes before and CODE_PAD_BACK bytes after become code 4.
// codes must be 16-byte aligned.
hipError_t launch_sketch_decode(uint8_t* codes, int64_t n, hipStream_t stream);

// Clears the bin minima, hashes every group, densifies and bit-slices.
// lds_bins: keep each workgroup's bin minima in LDS (64 s64 * 8 bytes of
// dynamic LDS); otherwise update the global minima directly.  The hashing
// grid is as many workgroups as fit the device at once; they take groups
// from the counter next_group (one int of device memory, zeroed here).
hipError_t launch_sketch(const SketchArgs& a, int64_t n_groups, bool lds_bins, int* next_group,
                         hipStream_t stream);

// Bin-minima reset and the densify / bit-slice pass of launch_sketch, for
// callers that fill the minima themselves (run_reads_job).
hipError_t launch_sketch_reset(const SketchArgs& a, hipStream_t stream);
hipError_t launch_sketch_finish(const SketchArgs& a, hipStream_t stream);

// One (sample, k) of read sketching with a count filter: units cover the
// sample's contigs for this k (pad = 1 for the first unit of a contig),
// wbase[u] is the index of unit u's first window among the n_windows
// windows of the sample, in sequence order.
constexpr int64_t READS_MAX_WINDOWS = ((int64_t)1 << 29) - 1;  // 4 events each fit an int count
struct ReadsJob {
  const uint8_t* codes;
  const SketchUnit* units;
  const int64_t* wbase;
  int64_t n_units;
  int64_t n_windows;
  int k;
  int use_rc;
  int min_count;  // 1 .. 255
  int exact;      // hash counter instead of the count-min sketch
  int codon;      // codon-phased seeds
  uint64_t binsize;
  uint32_t recip;
  uint64_t* bins;  // this (sample, k)'s bin minima
};
// Device workspace a job of up to max_windows windows needs.
size_t reads_workspace_bytes(int64_t max_windows);
// Hashes, counts and bins one job into j.bins (synchronises the stream once
// to learn the occurrence count).
hipError_t run_reads_job(const ReadsJob& j, void* ws, size_t ws_bytes, hipStream_t stream);

}  // namespace ppsk
