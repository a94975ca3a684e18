// Monte Carlo random-match tables: the RandomMC constructor
// (src/random/random_match.cpp:62-163) and calculate_random
// (src/api.cpp:518-531).
//
//  1. k-means (KMeansRex, src/random/kmeans/KMeansRexCore.cpp: kmeans++
//     seeding, Lloyd iterations) of the strand-symmetrised base
//     compositions into n_clusters clusters, seed 3019, at most 300
//     iterations; each sample joins its nearest (float) centroid.
//  2. k-mer lengths 3, 4, ... while the equal-base chance of a match exceeds
//     one bin, 1 / (64 s64), for sketch 0's length (random_match.cpp:82-95).
//  3. n_MC random genomes per cluster, each as long as sketch 0, bases drawn
//     i.i.d. from the centroid composition with xoshiro256** and
//     std::discrete_distribution; with use_rc the second half is written as
//     complements (random_match.cpp:400-422).  The generator streams follow
//     the reference's OpenMP loop: thread t of num_threads holds its own copy
//     of the seed-1 generator and jumps it t times before each genome it
//     makes, the genomes being split between threads as libgomp's static
//     schedule splits the iterations (random_match.cpp:97-118).
//  4. The genomes are sketched on the GPU (ppsk_sketch); per k, the mean
//     uncorrected Jaccard over all pairs of genomes of two clusters (self
//     pairs left out of a cluster's own entry) fills the C x C table
//     (random_match.cpp:120-156), with bin-match counts from ppsk_dense.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "api.hpp"
#include "ppsk.h"

namespace ppsk_host {

namespace {

constexpr size_t N_BASES = 4;
constexpr size_t MIN_KMER = 3;    // random_match.cpp:30-31
constexpr size_t MAX_KMER = 31;
constexpr int KMEANS_ITER = 300;  // random_match.cpp:34-36
constexpr int KMEANS_SEED = 3019;

void check(int rc) {
  if (rc != PPSK_OK) throw std::runtime_error(ppsk_last_error());
}

// genrand_double of the k-means generator: a 32-bit MT19937 output over
// 2^32 - 1, so both ends of [0, 1] occur.
double mt_unit(std::mt19937& mt) { return (double)mt() * (1.0 / 4294967295.0); }

// D^2-weighted choice (KMeansRex discrete_rand): the first index whose
// running sum exceeds total * u.
size_t weighted_choice(const std::vector<double>& p, std::mt19937& mt) {
  double total = 0;
  for (double v : p) total += v;
  const double r = total * mt_unit(mt);
  double run = p[0];
  size_t pick = 0;
  while (r >= run && pick + 1 < p.size()) {
    ++pick;
    run += p[pick];
  }
  return pick;
}

double sq_dist(const double* a, const double* b, size_t d) {
  double s = 0;
  for (size_t j = 0; j < d; ++j) {
    const double e = a[j] - b[j];
    s += e * e;
  }
  return s;
}

// The reference's strand-symmetrised composition (apply_rc,
// random_match.cpp:338-348).
std::array<double, 4> sym_bases(const Sketch& s) {
  std::array<double, 4> b = s.bases;
  if (s.use_rc) {
    b[0] = 0.5 * (b[0] + b[3]);
    b[1] = 0.5 * (b[1] + b[2]);
    b[2] = b[1];
    b[3] = b[0];
  }
  return b;
}

uint16_t nearest_centroid(const Sketch& s, const std::vector<float>& centroids, size_t c) {
  const auto b = sym_bases(s);
  const float v[4] = {(float)b[0], (float)b[1], (float)b[2], (float)b[3]};
  uint16_t best = 0;
  float best_d = std::numeric_limits<float>::infinity();
  for (size_t i = 0; i < c; ++i) {
    float d = 0.0f;
    for (int j = 0; j < 4; ++j) {
      const float e = centroids[i * 4 + j] - v[j];
      d += e * e;
    }
    if (d < best_d) {  // first minimum wins
      best_d = d;
      best = (uint16_t)i;
    }
  }
  return best;
}

}  // namespace

// ---------------------------------------------------------------------------
uint64_t Xoshiro256::splitmix(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

Xoshiro256::Xoshiro256(uint64_t seed) {
  // chained: each word is the mix of the previous one (src/random/rng.cpp:18-30)
  s_[0] = splitmix(seed);
  for (int i = 1; i < 4; ++i) s_[i] = splitmix(s_[i - 1]);
}

uint64_t Xoshiro256::operator()() {
  auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
  const uint64_t out = rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl(s_[3], 45);
  return out;
}

void Xoshiro256::jump() {
  // the published xoshiro256 jump polynomial: 2^128 steps ahead
  static const uint64_t poly[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                                   0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
  uint64_t acc[4] = {0, 0, 0, 0};
  for (uint64_t w : poly) {
    for (int b = 0; b < 64; ++b) {
      if ((w >> b) & 1ULL) {
        for (int i = 0; i < 4; ++i) acc[i] ^= s_[i];
      }
      (*this)();
    }
  }
  std::memcpy(s_, acc, sizeof(s_));
}

// ---------------------------------------------------------------------------
KMeansResult kmeans_plusplus(const std::vector<double>& x, size_t n, size_t d, size_t k,
                             int max_iter, int seed) {
  if (n == 0 || k == 0 || x.size() != n * d) throw std::runtime_error("k-means: bad input");
  std::mt19937 mt((uint32_t)seed);
  KMeansResult r;
  r.centroids.assign(k * d, 0.0);
  r.assignment.assign(n, 0);
  // kmeans++ seeding (KMeansRex sampleRowsPlusPlus): the first centre is a
  // uniform row, later ones D^2-weighted by the distance to the nearest
  // centre so far; with more centres than rows only n are seeded.
  const size_t seeded = std::min(k, n);
  size_t pick = (size_t)((double)n * mt_unit(mt));
  pick = std::min(pick, n - 1);  // u = 1 exactly would index past the end
  std::copy(&x[pick * d], &x[pick * d] + d, &r.centroids[0]);
  std::vector<double> nearest(n), cur(n);
  for (size_t c = 1; c < seeded; ++c) {
    for (size_t i = 0; i < n; ++i) {
      cur[i] = sq_dist(&x[i * d], &r.centroids[(c - 1) * d], d);
      nearest[i] = c == 1 ? cur[i] : std::min(cur[i], nearest[i]);
    }
    pick = weighted_choice(nearest, mt);
    std::copy(&x[pick * d], &x[pick * d] + d, &r.centroids[c * d]);
  }
  // Lloyd iterations (run_lloyd): assign to the first nearest centre, move
  // centres to their members' means (an empty cluster's centre becomes 0),
  // stop when the summed distance repeats exactly.
  double prev = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> members(k);
  for (int it = 0; it < max_iter; ++it) {
    double total = 0;
    for (size_t i = 0; i < n; ++i) {
      double best = sq_dist(&x[i * d], &r.centroids[0], d);
      size_t bc = 0;
      for (size_t c = 1; c < k; ++c) {
        const double dd = sq_dist(&x[i * d], &r.centroids[c * d], d);
        if (dd < best) {
          best = dd;
          bc = c;
        }
      }
      total += best;
      r.assignment[i] = (int)bc;
    }
    std::fill(r.centroids.begin(), r.centroids.end(), 0.0);
    std::fill(members.begin(), members.end(), 0.0);
    for (size_t i = 0; i < n; ++i) {
      const size_t c = (size_t)r.assignment[i];
      for (size_t j = 0; j < d; ++j) r.centroids[c * d + j] += x[i * d + j];
      members[c] += 1;
    }
    for (size_t c = 0; c < k; ++c) {
      const double m = members[c] + 1e-100;
      for (size_t j = 0; j < d; ++j) r.centroids[c * d + j] /= m;
    }
    r.iterations = it + 1;
    if (prev == total) break;
    prev = total;
  }
  return r;
}

// ---------------------------------------------------------------------------
std::vector<std::string> random_genomes(const std::vector<float>& centroids, size_t n_clusters,
                                        unsigned n_mc, size_t length, bool use_rc,
                                        int num_threads) {
  static const char fwd[N_BASES] = {'A', 'C', 'G', 'T'};
  static const char cmp[N_BASES] = {'T', 'G', 'C', 'A'};
  const size_t n_it = n_clusters * n_mc;
  std::vector<std::string> out(n_it);
  const size_t T = (size_t)std::max(num_threads, 1);
  // libgomp's static split of n_it iterations over T threads: the first
  // n_it % T threads take one more, in contiguous blocks
  const size_t q = n_it / T, rem = n_it % T;
#pragma omp parallel for schedule(static, 1) num_threads(std::max<int>(num_threads, 1))
  for (long t = 0; t < (long)T; ++t) {
    const size_t count = q + ((size_t)t < rem ? 1 : 0);
    const size_t first = (size_t)t * q + std::min((size_t)t, rem);
    Xoshiro256 gen(1);
    for (size_t it = first; it < first + count; ++it) {
      for (long j = 0; j < t; ++j) gen.jump();
      const size_t cl = it / n_mc;
      std::discrete_distribution<int> base{centroids[cl * 4 + 0], centroids[cl * 4 + 1],
                                           centroids[cl * 4 + 2], centroids[cl * 4 + 3]};
      std::string& s = out[it];
      s.resize(length);
      for (size_t i = 0; i < length; ++i) {
        const int b = base(gen);
        s[i] = (use_rc && i > length / 2) ? cmp[b] : fwd[b];
      }
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
RandomTables build_random(const std::vector<Sketch>& sketches, unsigned n_clusters, unsigned n_mc,
                          bool codon_phased, bool use_rc, int num_threads, int device_id) {
  std::cerr << "Calculating random match chances using Monte Carlo" << std::endl;
  if (sketches.empty()) throw std::runtime_error("No sketches to calculate random matches from");
  if (n_clusters >= sketches.size()) {
    std::cerr << "Cannot make more base frequency clusters than sketches" << std::endl;
    n_clusters = (unsigned)sketches.size() - 1;
  }
  if (n_clusters < 2) throw std::runtime_error("Cannot make this few base frequency clusters");
  if (n_mc < 1) throw std::runtime_error("Need at least one random genome per cluster");
  const size_t s64 = sketches[0].sketchsize64;
  const size_t C = n_clusters;

  // 1. k-means of the symmetrised compositions
  const size_t n = sketches.size();
  std::vector<double> x(n * N_BASES);
  for (size_t i = 0; i < n; ++i) {
    const auto b = sym_bases(sketches[i]);
    std::copy(b.begin(), b.end(), &x[i * N_BASES]);
  }
  const KMeansResult km = kmeans_plusplus(x, n, N_BASES, C, KMEANS_ITER, KMEANS_SEED);
  RandomTables t;
  t.present = true;
  t.use_rc = sketches[0].use_rc;
  t.n_clusters = C;
  t.centroids.resize(C * N_BASES);
  for (size_t i = 0; i < C * N_BASES; ++i) t.centroids[i] = (float)km.centroids[i];
  for (const auto& s : sketches) {
    t.table_keys.push_back(s.name);
    t.table_values.push_back(nearest_centroid(s, t.centroids, C));
  }

  // 2. k-mer lengths with signal above one bin
  const size_t len0 = sketches[0].seq_length;
  std::vector<size_t> kl = {MIN_KMER};
  const double min_random = 1.0 / (double)(s64 * 64);
  for (size_t k = MIN_KMER + 1; k <= MAX_KMER; ++k) {
    if (csrs(k, use_rc, len0, len0) > min_random) {
      kl.push_back(k);
    } else {
      break;
    }
  }
  t.k_min = (unsigned)kl.front();
  t.k_max = (unsigned)kl.back();

  // 3. random genomes, 4. sketched on the GPU
  const std::vector<std::string> genomes = random_genomes(t.centroids, C, n_mc, len0, use_rc,
                                                          num_threads);
  const size_t n_rand = genomes.size();
  std::string seq;
  seq.reserve(n_rand * len0);
  std::vector<int64_t> off{0}, sc{0};
  for (const auto& g : genomes) {
    seq += g;
    off.push_back((int64_t)seq.size());
    sc.push_back((int64_t)off.size() - 1);
  }
  std::vector<int32_t> k32(kl.begin(), kl.end());
  const size_t words = s64 * 14, K = kl.size();
  std::vector<uint64_t> raw(n_rand * K * words);
  check(ppsk_sketch_ex(device_id, (int64_t)n_rand, seq.data(), (int64_t)seq.size(), off.data(),
                       (int64_t)n_rand, sc.data(), (int32_t)K, k32.data(), (int32_t)s64,
                       use_rc ? 1 : 0, codon_phased ? 1 : 0, 0, 0, raw.data(), nullptr,
                       nullptr));
  std::string().swap(seq);

  // pairwise bin-match counts of the random genomes, every k
  std::vector<uint16_t> counts;
  if (n_rand > 1) {
    ppsk_set* set = nullptr;
    check(ppsk_set_create(device_id, (int64_t)n_rand, (int32_t)K, k32.data(), (int32_t)s64, 14,
                          &set));
    const int rc_up = ppsk_set_upload(set, 0, (int64_t)n_rand, raw.data());
    counts.resize(n_rand * (n_rand - 1) / 2 * K);
    const int rc_d = rc_up != PPSK_OK
                         ? rc_up
                         : ppsk_dense(set, nullptr, PPSK_RANDOM_NONE, 0, nullptr, 0.0,
                                      PPSK_OUT_COUNTS, counts.data(), 0, nullptr, nullptr);
    ppsk_set_free(set);
    check(rc_d);
  }
  auto jaccard = [&](size_t a, size_t b, size_t kk) -> double {
    if (a == b) return 1.0;  // a sketch matches itself in every bin
    const size_t i = std::min(a, b), j = std::max(a, b);
    const size_t row = n_rand * i - i * (i + 1) / 2 + j - 1 - i;
    return counts[row * K + kk] / (double)(64 * s64);
  };
  t.matches_keys.assign(kl.begin(), kl.end());
  t.matches.assign(K * C * C, 0.0f);
  for (size_t kk = 0; kk < K; ++kk) {
    for (size_t ci = 0; ci < C; ++ci) {
      for (size_t cj = ci; cj < C; ++cj) {
        double sum = 0;
        for (size_t a = 0; a < n_mc; ++a) {
          for (size_t b = 0; b < n_mc; ++b) sum += jaccard(ci * n_mc + a, cj * n_mc + b, kk);
        }
        double count = (double)n_mc * n_mc;
        if (ci == cj) {  // the n_mc self comparisons (== 1) are removed
          sum -= n_mc;
          count -= n_mc;
        }
        const float m = (float)(sum / count);
        t.matches[(kk * C + ci) * C + cj] = m;
        t.matches[(kk * C + cj) * C + ci] = m;
      }
    }
  }
  return t;
}

void calculate_random(const std::vector<Sketch>& sketches, const std::string& db_name,
                      unsigned n_clusters, unsigned n_mc, bool codon_phased, bool use_rc,
                      int num_threads, int device_id) {
  const RandomTables t =
      build_random(sketches, n_clusters, n_mc, codon_phased, use_rc, num_threads, device_id);
  write_random(db_name + ".h5", t);
}

}  // namespace ppsk_host
