// Host orchestration on top of the C ABI (libppsk): the roles of query_db /
// query_db_sparse / query_db_cuda / query_db_sparse_cuda (src/api.cpp:140-432,
// src/gpu/gpu_api.cpp:352-531), of create_sketches (src/api.cpp:39-138) and of
// the RandomMC constructor.  Every distance and sketch is computed by the HIP
// kernels; there is no CPU path.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <array>
#include <string>
#include <tuple>
#include <vector>

#include "random_match.hpp"
#include "sketch_db.hpp"

namespace ppsk_host {

using sparse_coo = std::tuple<std::vector<long>, std::vector<long>, std::vector<float>>;

// Row-major float matrix (the reference's NumpyMatrix, src/dist/matrix.hpp:25).
struct DenseMatrix {
  size_t rows = 0, cols = 0;
  std::vector<float> data;
};

// Distances failed for some pair: the reference prints and exit(1)s the
// process (src/dist/linear_regression.cpp:51-57); here it is an exception.
struct FitError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Regression tolerance of the reference: a function-local static set on the
// first core/accessory fit of the process, 5 / (64 * sketchsize64 of that
// call's first sample) (src/dist/linear_regression.cpp:22).
double regression_tolerance(size_t sketchsize64);

// query_db: self mode when ref and query hold the same names in the same
// order (src/api.cpp:195); dense core/acc or per-k Jaccard.
DenseMatrix query_db(std::vector<Sketch>& ref, std::vector<Sketch>& query,
                     const std::vector<size_t>& kmer_lengths, RandomMC& random, bool jaccard,
                     int device_id);

// query_db_sparse (kNN, src/api.cpp:319-432).
sparse_coo query_db_sparse(std::vector<Sketch>& ref, const std::vector<size_t>& kmer_lengths,
                           RandomMC& random, bool jaccard, int kNN, size_t dist_col,
                           int device_id);

// Self threshold sparsification computed on the device (the
// dist_cutoff > 0 route of querySelfSparse, src/sketchlib_bindings.cpp:159-169).
sparse_coo query_db_threshold(std::vector<Sketch>& ref, const std::vector<size_t>& kmer_lengths,
                              RandomMC& random, float cutoff, size_t dist_col, int device_id);

// One random-corrected Jaccard (src/sketchlib_bindings.cpp:204-216): the
// match count from the device, the double-precision correction on the host.
double jaccard_dist(const Sketch& s1, const Sketch& s2, size_t k, const RandomMC& random,
                    int device_id);

// Matrix utilities (src/dist/matrix_ops.cpp:38-150, matrix_idx.hpp:13-50).
DenseMatrix long_to_square(const float* rr, size_t n_rr, const float* qr, size_t n_qr,
                           const float* qq, size_t n_qq, unsigned num_threads);
std::vector<float> square_to_long(const float* square, size_t n, unsigned num_threads);
sparse_coo sparsify_dists_by_threshold(const float* dense, size_t rows, size_t cols,
                                       float cutoff, unsigned num_threads);

// Synthetic database writer: seeded sketches in the database layout.
// sketch_version attribute written to new databases: the reference's
// SKETCH_VERSION for 2.1.5, so databases compare equal across the two builds.
constexpr const char* SKETCH_VERSION = "445afa87649c32acbf9749c77231b5b0fcb08179";

// xoshiro256** with the reference's seeding (each state word the splitmix64
// of the previous one, src/random/rng.cpp:18-30); a C++11 URNG.
class Xoshiro256 {
 public:
  using result_type = uint64_t;
  static constexpr size_t min() { return 0; }
  static constexpr size_t max() { return ~(size_t)0; }
  explicit Xoshiro256(uint64_t seed);
  uint64_t operator()();
  void jump();  // 2^128 steps ahead

 private:
  static uint64_t splitmix(uint64_t x);
  uint64_t s_[4];
};

// KMeansRex with kmeans++ seeding from a 32-bit Mersenne Twister
// (src/random/kmeans/KMeansRexCore.cpp).  x: n rows of d values.
struct KMeansResult {
  std::vector<double> centroids;  // k x d
  std::vector<int> assignment;    // n
  int iterations = 0;
};
KMeansResult kmeans_plusplus(const std::vector<double>& x, size_t n, size_t d, size_t k,
                             int max_iter, int seed);

// Random genomes of the Monte Carlo tables (generate_random_sequence and the
// generator streams of the RandomMC constructor's OpenMP loop,
// src/random/random_match.cpp:97-118, 400-422): n_mc per cluster, in
// cluster-major order; centroids is C x 4 (A, C, G, T).
std::vector<std::string> random_genomes(const std::vector<float>& centroids, size_t n_clusters,
                                        unsigned n_mc, size_t length, bool use_rc,
                                        int num_threads);

// The RandomMC constructor (src/random/random_match.cpp:62-163): random
// genomes sketched and compared on the GPU.
RandomTables build_random(const std::vector<Sketch>& sketches, unsigned n_clusters, unsigned n_mc,
                          bool codon_phased, bool use_rc, int num_threads, int device_id);
// calculate_random (src/api.cpp:518-531): build_random, saved under /random.
void calculate_random(const std::vector<Sketch>& sketches, const std::string& db_name,
                      unsigned n_clusters, unsigned n_mc, bool codon_phased, bool use_rc,
                      int num_threads, int device_id);
constexpr unsigned DEFAULT_N_CLUSTERS = 3;  // src/random/random_match.hpp:20-21
constexpr unsigned DEFAULT_N_MC = 5;

// constructDatabase (src/sketchlib_bindings.cpp:45-84): sketches on the GPU.
// What the sequence reader keeps of a sample's files (src/sketch/seqio.cpp:
// 55-111): contigs at least kmer_len long, upper-cased; composition of the
// first 999,999 bases; N count; whether any record had a quality string.
struct SequenceStats {
  std::array<double, 4> bases{};
  size_t total = 0;
  unsigned long missing_bases = 0;
  bool reads = false;
  std::vector<std::string> contigs;
};
SequenceStats read_sequence_stats(const std::vector<std::string>& files, size_t kmer_len);

void construct_database(const std::string& db_name, const std::vector<std::string>& names,
                        const std::vector<std::vector<std::string>>& files,
                        std::vector<size_t> kmer_lengths, size_t sketchsize64, bool codon_phased,
                        bool calc_random, bool use_rc, size_t min_count, bool exact,
                        size_t num_threads, int device_id);

void write_synthetic_db(const std::string& db_name, int64_t n_samples,
                        const std::vector<size_t>& kmer_lengths, size_t sketchsize64,
                        uint64_t seed, double dcore_max, double dacc_max, double length_mean,
                        const std::string& name_prefix);

}  // namespace ppsk_host
