#include "random_match.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>

namespace ppsk_host {

namespace {
const unsigned kDefaultMaxK = 101;  // src/random/random_match.hpp:19
}

double csrs(size_t k, bool use_rc, size_t l1, size_t l2) {
  const size_t shift = (k - 1) * 2 + (use_rc ? 1 : 0);
  if (k < 1 || shift >= 64) {
    throw std::runtime_error("equal-base random match chance is undefined for k = " +
                             std::to_string(k));
  }
  double random_chance = 0;
  const size_t match_chance = (size_t)1 << shift;
  const double j1 = 1 - std::pow(1 - (double)1 / match_chance, (double)l1);
  const double j2 = 1 - std::pow(1 - (double)1 / match_chance, (double)l2);
  if (j1 > 0 && j2 > 0) {
    random_chance = (j1 * j2) / (j1 + j2 - j1 * j2);
  }
  return random_chance;
}

RandomMC::RandomMC()
    : no_adjustment_(true), no_mc_(false), use_rc_(false), min_k_(0), max_k_(kDefaultMaxK) {}

RandomMC::RandomMC(bool use_rc)
    : no_adjustment_(false), no_mc_(true), use_rc_(use_rc), min_k_(0), max_k_(kDefaultMaxK) {}

RandomMC::RandomMC(const RandomTables& t)
    : no_adjustment_(false), no_mc_(false), use_rc_(t.use_rc), min_k_(t.k_min), max_k_(t.k_max),
      n_clusters_(t.n_clusters), centroids_(t.centroids) {
  for (size_t i = 0; i < t.table_keys.size(); ++i) cluster_table_[t.table_keys[i]] = t.table_values[i];
  const size_t cc = n_clusters_ * n_clusters_;
  for (size_t i = 0; i < t.matches_keys.size(); ++i) {
    matches_[t.matches_keys[i]] =
        std::vector<float>(t.matches.begin() + i * cc, t.matches.begin() + (i + 1) * cc);
  }
}

RandomType RandomMC::mode() const {
  if (no_adjustment_) return RandomType::NoAdjust;
  if (no_mc_) return RandomType::EqualBase;
  return RandomType::MonteCarlo;
}

size_t RandomMC::min_supported_k(size_t seq_length) const {
  // Equal-base mode skips the search (the reference loop is guarded by
  // !_no_MC).  With no adjustment the reference starts its search at k = 0,
  // where its csrs shift is undefined but evaluates to a zero chance on
  // x86-64, so the first k is accepted: 0.
  (void)seq_length;
  if (mode() != RandomType::MonteCarlo) return min_k_;
  size_t min_k_signal = min_k_;
  for (size_t k = min_k_; k < max_k_; ++k) {
    const auto it = matches_.find(k);
    if (it == matches_.end()) throw std::out_of_range("random match table has no k = " + std::to_string(k));
    double sum = 0;
    for (float v : it->second) sum += v;
    if (sum / it->second.size() < 1) {
      min_k_signal = k;
      break;
    }
  }
  return min_k_signal;
}

double RandomMC::random_match(const Sketch& r1, uint16_t q_cluster, size_t q_length, size_t k) const {
  double r = 0;
  if (no_mc_) {
    r = csrs(k, use_rc_, r1.seq_length, q_length);
  } else if (!no_adjustment_ && k < max_k_) {
    const uint16_t rc = table_cluster(r1.name);
    const auto it = matches_.find(k);
    if (it == matches_.end()) throw std::out_of_range("random match table has no k = " + std::to_string(k));
    r = it->second[(size_t)rc * n_clusters_ + q_cluster];
  }
  return r;
}

double RandomMC::random_match(const Sketch& r1, const Sketch& r2, size_t k) const {
  uint16_t q = 0;
  if (mode() == RandomType::MonteCarlo) q = table_cluster(r2.name);
  return random_match(r1, q, r2.seq_length, k);
}

uint16_t RandomMC::closest_cluster(const Sketch& s) const {
  if (mode() != RandomType::MonteCarlo) return 0;
  std::array<double, 4> b = s.bases;
  if (s.use_rc) {
    b[0] = 0.5 * (b[0] + b[3]);
    b[1] = 0.5 * (b[1] + b[2]);
    b[2] = b[1];
    b[3] = b[0];
  }
  const float v[4] = {(float)b[0], (float)b[1], (float)b[2], (float)b[3]};
  uint16_t best = 0;
  float best_d = std::numeric_limits<float>::infinity();
  for (size_t c = 0; c < n_clusters_; ++c) {
    float d = 0.0f;
    for (int j = 0; j < 4; ++j) {
      const float e = centroids_[c * 4 + j] - v[j];
      d += e * e;
    }
    if (d < best_d) {
      best_d = d;
      best = (uint16_t)c;
    }
  }
  return best;
}

bool RandomMC::check_present(const std::vector<Sketch>& sketches, bool update) {
  bool missing = false;
  if (mode() == RandomType::MonteCarlo) {
    for (const auto& s : sketches) {
      if (cluster_table_.find(s.name) == cluster_table_.end()) {
        missing = true;
        if (update) {
          cluster_table_[s.name] = closest_cluster(s);
        } else {
          break;
        }
      }
    }
  }
  return missing;
}

uint16_t RandomMC::table_cluster(const std::string& name) const {
  const auto it = cluster_table_.find(name);
  if (it == cluster_table_.end()) throw std::out_of_range("sample " + name + " not in the random match table");
  return it->second;
}

std::vector<uint16_t> RandomMC::lookup_array(const std::vector<Sketch>& sketches) const {
  std::vector<uint16_t> out;
  out.reserve(sketches.size());
  for (const auto& s : sketches) {
    if (mode() != RandomType::MonteCarlo) {
      out.push_back(0);
    } else {
      const auto it = cluster_table_.find(s.name);
      out.push_back(it == cluster_table_.end() ? closest_cluster(s) : it->second);
    }
  }
  return out;
}

std::vector<float> RandomMC::device_table(const std::vector<size_t>& klist) const {
  const size_t cc = n_clusters_ * n_clusters_;
  std::vector<float> t(klist.size() * cc, 0.0f);
  if (mode() != RandomType::MonteCarlo) return t;
  for (size_t i = 0; i < klist.size(); ++i) {
    if (klist[i] >= max_k_) continue;  // random_match.cpp:191: k >= max_k -> 0
    const auto it = matches_.find(klist[i]);
    if (it == matches_.end()) {
      throw std::out_of_range("random match table has no k = " + std::to_string(klist[i]));
    }
    std::copy(it->second.begin(), it->second.end(), t.begin() + i * cc);
  }
  return t;
}

}  // namespace ppsk_host
