#include "api.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <memory>

#include "ppsk.h"

namespace ppsk_host {

namespace {

struct SetHandle {
  ppsk_set* h = nullptr;
  SetHandle() = default;
  SetHandle(const SetHandle&) = delete;
  SetHandle& operator=(const SetHandle&) = delete;
  ~SetHandle() {
    if (h) ppsk_set_free(h);
  }
};

void check(int rc) {
  if (rc == PPSK_OK) return;
  if (rc == PPSK_ERR_FIT) {
    throw FitError(std::string("Fitting k-mer gradient failed: ") + ppsk_last_error() +
                   "\nCheck for low quality genomes");
  }
  throw std::runtime_error(ppsk_last_error());
}

// checkSketchParamsMatch (src/gpu/gpu_api.cpp:170-185).  The k-mer test is
// "every requested length present": load_sketches keeps stored lengths above
// the largest requested one (src/api.cpp:471-478) and the reference CPU path
// only ever reads the requested ones.
void check_params(const std::vector<Sketch>& s, const std::vector<size_t>& klist, size_t bbits,
                  size_t s64) {
  for (const auto& x : s) {
    if (x.bbits != bbits) throw std::runtime_error("Mismatching bbits in sketches");
    if (x.sketchsize64 != s64) throw std::runtime_error("Mismatching sketchsize64 in sketches");
    for (size_t k : klist) {
      if (std::find(x.kmers.begin(), x.kmers.end(), k) == x.kmers.end()) {
        throw std::runtime_error("Mismatching k-mer lengths in sketches");
      }
    }
  }
}

// Resident device copy of a sketch vector with lengths and cluster ids.
void upload(SetHandle& set, const std::vector<Sketch>& s, const std::vector<size_t>& klist,
            int device, const RandomMC& random, const std::vector<uint16_t>* clusters) {
  const size_t s64 = s[0].sketchsize64, bbits = s[0].bbits;
  std::vector<int32_t> kl(klist.begin(), klist.end());
  check(ppsk_set_create(device, (int64_t)s.size(), (int32_t)kl.size(), kl.data(), (int32_t)s64,
                        (int32_t)bbits, &set.h));
  const size_t words = s64 * bbits;
  const size_t per_sample = klist.size() * words;
  const size_t chunk = std::max<size_t>(1, (size_t)(256u << 20) / (per_sample * 8));
  std::vector<uint64_t> raw;
  for (size_t c0 = 0; c0 < s.size(); c0 += chunk) {
    const size_t m = std::min(chunk, s.size() - c0);
    raw.resize(m * per_sample);
    for (size_t i = 0; i < m; ++i) {
      for (size_t kk = 0; kk < klist.size(); ++kk) {
        const auto& v = s[c0 + i].get_sketch(klist[kk]);
        if (v.size() != words) throw std::runtime_error("sketch of unexpected size in " + s[c0 + i].name);
        std::copy(v.begin(), v.end(), raw.begin() + i * per_sample + kk * words);
      }
    }
    check(ppsk_set_upload(set.h, (int64_t)c0, (int64_t)m, raw.data()));
  }
  std::vector<uint64_t> lengths(s.size());
  for (size_t i = 0; i < s.size(); ++i) lengths[i] = s[i].seq_length;
  const bool need_len = random.mode() == RandomType::EqualBase;
  check(ppsk_set_meta(set.h, need_len ? lengths.data() : nullptr, random.use_rc() ? 1 : 0,
                      clusters ? clusters->data() : nullptr));
}

int device_mode(const RandomMC& r) {
  switch (r.mode()) {
    case RandomType::EqualBase:
      return PPSK_RANDOM_EQUAL_BASE;
    case RandomType::MonteCarlo:
      return PPSK_RANDOM_MC;
    default:
      return PPSK_RANDOM_NONE;
  }
}

void check_signal(const std::vector<Sketch>& ref, const std::vector<size_t>& klist,
                  const RandomMC& random) {
  if (klist.empty()) throw std::runtime_error("No k-mer lengths given");
  if (klist[0] < random.min_supported_k(ref[0].seq_length)) {
    throw std::runtime_error(
        "Smallest k-mer has no signal above random chance; increase minimum k-mer length");
  }
}

void warn_missing(RandomMC& random, const std::vector<Sketch>& ref) {
  if (random.check_present(ref, true)) {
    std::cerr << "Some members of the reference database were not found "
                 "in its random match chances. Consider refreshing with addRandom"
              << std::endl;
  }
}

std::vector<uint16_t> table_clusters(const RandomMC& random, const std::vector<Sketch>& s) {
  std::vector<uint16_t> c(s.size(), 0);
  if (random.mode() == RandomType::MonteCarlo) {
    for (size_t i = 0; i < s.size(); ++i) c[i] = random.table_cluster(s[i].name);
  }
  return c;
}

}  // namespace

double regression_tolerance(size_t sketchsize64) {
  static const double tolerance = 5.0 / (sketchsize64 * 64);
  return tolerance;
}

DenseMatrix query_db(std::vector<Sketch>& ref, std::vector<Sketch>& query,
                     const std::vector<size_t>& kmer_lengths, RandomMC& random, bool jaccard,
                     int device_id) {
  if (ref.empty() || query.empty()) throw std::runtime_error("Query with empty ref or query list!");
  check_signal(ref, kmer_lengths, random);
  warn_missing(random, ref);
  const bool self = same_names(ref, query);
  const size_t bbits = ref[0].bbits, s64 = ref[0].sketchsize64;
  check_params(ref, kmer_lengths, bbits, s64);
  if (!self) check_params(query, kmer_lengths, bbits, s64);
  if (!jaccard && kmer_lengths.size() < 2) {
    throw FitError("Fitting k-mer gradient failed: fewer than two k-mer lengths");
  }
  const bool mc = random.mode() == RandomType::MonteCarlo;
  std::vector<float> table;
  if (mc) table = random.device_table(kmer_lengths);

  DenseMatrix out;
  out.cols = jaccard ? kmer_lengths.size() : 2;
  out.rows = self ? ref.size() * (ref.size() - 1) / 2 : ref.size() * query.size();
  out.data.assign(out.rows * out.cols, 0.0f);

  SetHandle rset, qset;
  if (self) {
    std::vector<uint16_t> cl = table_clusters(random, ref);
    upload(rset, ref, kmer_lengths, device_id, random, mc ? &cl : nullptr);
  } else {
    std::vector<uint16_t> rcl = random.lookup_array(ref);
    std::vector<uint16_t> qcl(query.size(), 0);
    for (size_t i = 0; i < query.size(); ++i) qcl[i] = random.closest_cluster(query[i]);
    upload(rset, ref, kmer_lengths, device_id, random, mc ? &rcl : nullptr);
    upload(qset, query, kmer_lengths, device_id, random, mc ? &qcl : nullptr);
  }
  if (out.rows == 0) return out;
  const double tol = jaccard ? 0.0 : regression_tolerance(self ? ref[0].sketchsize64 : query[0].sketchsize64);
  uint64_t n_fail = 0;
  int64_t first_fail = -1;
  check(ppsk_dense(rset.h, self ? nullptr : qset.h, device_mode(random),
                   (int32_t)random.n_clusters(), mc ? table.data() : nullptr, tol,
                   jaccard ? PPSK_OUT_JACCARD : PPSK_OUT_CORE_ACC, out.data.data(), 0, &n_fail,
                   &first_fail));
  return out;
}

sparse_coo query_db_sparse(std::vector<Sketch>& ref, const std::vector<size_t>& kmer_lengths,
                           RandomMC& random, bool jaccard, int kNN, size_t dist_col,
                           int device_id) {
  // check_sparse_inputs (src/api.cpp:293-316)
  if (ref.empty()) throw std::runtime_error("Query with empty ref or query list!");
  check_signal(ref, kmer_lengths, random);
  if ((jaccard && dist_col >= kmer_lengths.size()) || (dist_col != 0 && dist_col != 1)) {
    throw std::runtime_error("dist_col out of range");
  }
  warn_missing(random, ref);
  const size_t n = ref.size();
  if (kNN <= 0) return sparse_coo();
  if ((size_t)kNN > n) {
    throw std::runtime_error("kNN must not exceed the number of samples");
  }
  check_params(ref, kmer_lengths, ref[0].bbits, ref[0].sketchsize64);
  if (!jaccard && kmer_lengths.size() < 2) {
    throw FitError("Fitting k-mer gradient failed: fewer than two k-mer lengths");
  }
  const bool mc = random.mode() == RandomType::MonteCarlo;
  std::vector<float> table;
  if (mc) table = random.device_table(kmer_lengths);
  std::vector<uint16_t> cl = table_clusters(random, ref);
  SetHandle set;
  upload(set, ref, kmer_lengths, device_id, random, mc ? &cl : nullptr);
  std::vector<int64_t> iv(n * kNN), jv(n * kNN);
  std::vector<float> dv(n * kNN);
  const double tol = jaccard ? 0.0 : regression_tolerance(ref[0].sketchsize64);
  uint64_t n_fail = 0;
  check(ppsk_knn(set.h, device_mode(random), (int32_t)random.n_clusters(),
                 mc ? table.data() : nullptr, tol, jaccard ? 1 : 0, kNN, (int32_t)dist_col,
                 iv.data(), jv.data(), dv.data(), &n_fail));
  return sparse_coo(std::vector<long>(iv.begin(), iv.end()), std::vector<long>(jv.begin(), jv.end()),
                    std::move(dv));
}

sparse_coo query_db_threshold(std::vector<Sketch>& ref, const std::vector<size_t>& kmer_lengths,
                              RandomMC& random, float cutoff, size_t dist_col, int device_id) {
  if (ref.empty()) throw std::runtime_error("Query with empty ref or query list!");
  check_signal(ref, kmer_lengths, random);
  warn_missing(random, ref);
  if (cutoff < 0) throw std::runtime_error("Distance threshold must be at least 0");
  check_params(ref, kmer_lengths, ref[0].bbits, ref[0].sketchsize64);
  if (kmer_lengths.size() < 2) {
    throw FitError("Fitting k-mer gradient failed: fewer than two k-mer lengths");
  }
  const bool mc = random.mode() == RandomType::MonteCarlo;
  std::vector<float> table;
  if (mc) table = random.device_table(kmer_lengths);
  std::vector<uint16_t> cl = table_clusters(random, ref);
  SetHandle set;
  upload(set, ref, kmer_lengths, device_id, random, mc ? &cl : nullptr);
  int64_t *pi = nullptr, *pj = nullptr;
  float* pd = nullptr;
  int64_t m = 0;
  uint64_t n_fail = 0;
  const int rc = ppsk_threshold(set.h, device_mode(random), (int32_t)random.n_clusters(),
                                mc ? table.data() : nullptr,
                                regression_tolerance(ref[0].sketchsize64), (int32_t)dist_col,
                                cutoff, &pi, &pj, &pd, &m, &n_fail);
  check(rc);
  sparse_coo out(std::vector<long>(pi, pi + m), std::vector<long>(pj, pj + m),
                 std::vector<float>(pd, pd + m));
  ppsk_free_host(pi);
  ppsk_free_host(pj);
  ppsk_free_host(pd);
  return out;
}

double jaccard_dist(const Sketch& s1, const Sketch& s2, size_t k, const RandomMC& random,
                    int device_id) {
  std::vector<Sketch> pair = {s1, s2};
  const std::vector<size_t> klist = {k};
  check_params(pair, klist, s1.bbits, s1.sketchsize64);
  SetHandle set;
  upload(set, pair, klist, device_id, RandomMC(), nullptr);
  uint16_t count = 0;
  check(ppsk_dense(set.h, nullptr, PPSK_RANDOM_NONE, 0, nullptr, 0.0, PPSK_OUT_COUNTS, &count, 0,
                   nullptr, nullptr));
  // Reference::jaccard_dist (src/reference.cpp:222-238): r1 = s1
  const double r = random.random_match(s1, s2, k);
  const size_t unionsize = 64 * s1.sketchsize64;
  const double jaccard_obs = count / (double)unionsize;
  const int max = 1;
  const double diff = jaccard_obs > r ? (jaccard_obs - r) : 0;
  return diff * max / (max - r);
}

// ---------------------------------------------------------------------------
// Matrix utilities
// ---------------------------------------------------------------------------
namespace {
inline size_t rows_to_samples(size_t rows) { return (size_t)(0.5 * (1 + std::sqrt(1 + 8 * (double)rows))); }
inline long calc_row_idx(long long k, long n) {
  return n - 2 - (long)std::floor(std::sqrt((double)(-8 * k + 4 * (long long)n * (n - 1) - 7)) / 2 - 0.5);
}
inline long calc_col_idx(long long k, long i, long n) {
  return (long)(k + i + 1 - (long long)n * (n - 1) / 2 + (long long)(n - i) * ((n - i) - 1) / 2);
}
inline long long square_to_condensed(long i, long j, long n) {
  return (n * i - ((i * (i + 1)) >> 1) + j - 1 - i);
}
}  // namespace

DenseMatrix long_to_square(const float* rr, size_t n_rr, const float* qr, size_t n_qr,
                           const float* qq, size_t n_qq, unsigned num_threads) {
  const size_t nrr = rows_to_samples(n_rr);
  size_t nqq = 0;
  if (n_qr > 0) {
    nqq = rows_to_samples(n_qq);
    if (nqq < 1) nqq = 1;
    if (n_qr != nrr * nqq) {
      throw std::runtime_error("Shape of reference, query and ref vs query matrices mismatches");
    }
  }
  DenseMatrix m;
  m.rows = m.cols = nrr + nqq;
  m.data.assign(m.rows * m.cols, 0.0f);
  const size_t side = m.cols;
  float* d = m.data.data();
#pragma omp parallel for schedule(static) num_threads(num_threads)
  for (long long x = 0; x < (long long)n_rr; ++x) {
    const long i = calc_row_idx(x, (long)nrr);
    const long j = calc_col_idx(x, i, (long)nrr);
    d[i * side + j] = rr[x];
    d[j * side + i] = rr[x];
  }
  if (n_qq > 0) {
#pragma omp parallel for schedule(static) num_threads(num_threads)
    for (long long x = 0; x < (long long)n_qq; ++x) {
      const long i = calc_row_idx(x, (long)nqq) + (long)nrr;
      const long j = calc_col_idx(x, i - (long)nrr, (long)nqq) + (long)nrr;
      d[i * side + j] = qq[x];
      d[j * side + i] = qq[x];
    }
  }
  if (n_qr > 0) {
#pragma omp parallel for schedule(static) num_threads(num_threads)
    for (long long x = 0; x < (long long)n_qr; ++x) {
      const size_t i = x % nrr;
      const size_t j = x / nrr + nrr;
      d[i * side + j] = qr[x];
      d[j * side + i] = qr[x];
    }
  }
  return m;
}

std::vector<float> square_to_long(const float* square, size_t n, unsigned num_threads) {
  std::vector<float> out(n * (n - 1) / 2);
#pragma omp parallel for schedule(guided, 1) num_threads(num_threads)
  for (long i = (long)n - 2; i >= 0; i--) {
    for (long j = i + 1; j < (long)n; j++) {
      out[square_to_condensed(i, j, (long)n)] = square[i * n + j];
    }
  }
  return out;
}

sparse_coo sparsify_dists_by_threshold(const float* dense, size_t rows, size_t cols, float cutoff,
                                       unsigned num_threads) {
  if (cutoff < 0) throw std::runtime_error("Distance threshold must be at least 0");
  std::vector<std::vector<long>> iv(rows), jv(rows);
  std::vector<std::vector<float>> dv(rows);
#pragma omp parallel for schedule(static) num_threads(num_threads)
  for (long i = 0; i < (long)rows; ++i) {
    for (long j = i + 1; j < (long)cols; ++j) {
      const float d = dense[i * cols + j];
      if (d < cutoff) {
        iv[i].push_back(i);
        jv[i].push_back(j);
        dv[i].push_back(d);
      }
    }
  }
  sparse_coo out;
  for (size_t i = 0; i < rows; ++i) {
    std::get<0>(out).insert(std::get<0>(out).end(), iv[i].begin(), iv[i].end());
    std::get<1>(out).insert(std::get<1>(out).end(), jv[i].begin(), jv[i].end());
    std::get<2>(out).insert(std::get<2>(out).end(), dv[i].begin(), dv[i].end());
  }
  return out;
}

void write_synthetic_db(const std::string& db_name, int64_t n_samples,
                        const std::vector<size_t>& kmer_lengths, size_t sketchsize64,
                        uint64_t seed, double dcore_max, double dacc_max, double length_mean,
                        const std::string& name_prefix) {
  const size_t K = kmer_lengths.size();
  std::vector<int32_t> kl(kmer_lengths.begin(), kmer_lengths.end());
  const size_t words = sketchsize64 * 14;
  std::vector<uint64_t> raw((size_t)n_samples * K * words);
  std::vector<uint64_t> lengths(n_samples);
  std::vector<double> bases((size_t)n_samples * 4);
  check(ppsk_synth_sketches(n_samples, (int32_t)K, kl.data(), (int32_t)sketchsize64, seed,
                            dcore_max, dacc_max, length_mean, raw.data(), lengths.data(),
                            bases.data(), 0));
  write_new_db(db_name + ".h5", true, false, "synthetic");
  const int64_t chunk = 1024;
  for (int64_t c0 = 0; c0 < n_samples; c0 += chunk) {
    std::vector<Sketch> batch;
    for (int64_t i = c0; i < std::min(n_samples, c0 + chunk); ++i) {
      Sketch s;
      s.name = name_prefix + std::to_string(i);
      s.sketchsize64 = sketchsize64;
      s.bbits = 14;
      s.seq_length = lengths[i];
      for (int b = 0; b < 4; ++b) s.bases[b] = bases[i * 4 + b];
      s.kmers = kmer_lengths;
      for (size_t kk = 0; kk < K; ++kk) {
        const uint64_t* src = raw.data() + ((size_t)i * K + kk) * words;
        s.usigs.emplace_back(src, src + words);
      }
      batch.push_back(std::move(s));
    }
    write_sketches(db_name + ".h5", batch);
  }
}

}  // namespace ppsk_host
