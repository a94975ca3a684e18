// Random-match model on the host side: which correction applies, per-sample
// cluster ids and the table handed to the device.  Mirrors the lookup side of
// RandomMC (src/random/random_match.hpp:27-117); the tables themselves are
// built by build_random (random_build.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "sketch_db.hpp"

namespace ppsk_host {

enum class RandomType { NoAdjust, EqualBase, MonteCarlo };

// csrs (src/random/random_match.cpp:318-336); k >= 33 is rejected.
double csrs(size_t k, bool use_rc, size_t l1, size_t l2);

class RandomMC {
 public:
  RandomMC();                     // no adjustment (random_match.cpp:49-51)
  explicit RandomMC(bool use_rc); // equal base frequencies (random_match.cpp:53-55)
  explicit RandomMC(const RandomTables& tables);  // Monte Carlo, from a database

  RandomType mode() const;
  bool use_rc() const { return use_rc_; }
  size_t n_clusters() const { return n_clusters_; }

  // random_match.cpp:165-182
  size_t min_supported_k(size_t seq_length) const;
  // random_match.cpp:185-206
  double random_match(const Sketch& r1, uint16_t q_cluster, size_t q_length, size_t k) const;
  double random_match(const Sketch& r1, const Sketch& r2, size_t k) const;
  // random_match.cpp:230-237 with nearest_neighbour / apply_rc (338-359)
  uint16_t closest_cluster(const Sketch& s) const;
  // random_match.cpp:239-252: adds samples missing from the cluster table
  bool check_present(const std::vector<Sketch>& sketches, bool update);
  // cluster of a sample already in the table (self mode lookups)
  uint16_t table_cluster(const std::string& name) const;
  // random_match.cpp:254-273: table entry, else the nearest centroid
  std::vector<uint16_t> lookup_array(const std::vector<Sketch>& sketches) const;
  // [n_k][C][C] float table for the device, zero where k >= max_k; throws
  // std::out_of_range when a k < max_k has no table (matches.at, :195)
  std::vector<float> device_table(const std::vector<size_t>& klist) const;

 private:
  bool no_adjustment_;
  bool no_mc_;
  bool use_rc_;
  unsigned min_k_, max_k_;
  size_t n_clusters_ = 0;
  std::unordered_map<std::string, uint16_t> cluster_table_;
  std::unordered_map<size_t, std::vector<float>> matches_;  // k -> C x C row-major
  std::vector<float> centroids_;                            // C x 4
};

}  // namespace ppsk_host
