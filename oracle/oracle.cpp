/*
 * oracle.cpp -- CPU restatement of pp-sketchlib's core/accessory distance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product path
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function restates one piece of the reference CPU path and cites the
 * reference file:line it follows (paths relative to the pp-sketchlib source
 * tree).  The restatement is written from the reference's behaviour, in plain
 * C++ with OpenMP, without Eigen (absent here).
 *
 * Floating point: compile with -ffp-contract=off and without -ffast-math so the
 * double-precision Jaccard and random-match arithmetic is IEEE-exact and
 * reproducible (the reference's own CMake Release build adds -ffast-math,
 * CMakeLists.txt:41; its Makefile build does not, src/Makefile_fedora38:12).
 *
 * Deviation that is deliberate and bounded: fit_slope's Eigen float LDLT solve
 * (dist/linear_regression.cpp:37-42) is replaced by a double-precision centred
 * closed-form least-squares fit; on the reference's golden outputs
 * (test/ppsketch_ref*.npy) this reproduces core/accessory to ~1e-7.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <queue>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

/* ---------------------------------------------------------------------------
 * calc_intersize -- dist/dist.cpp:50-78 (bindash bin-match count).
 * Per 64-bin word: AND over the bbits layers of NOT(XOR), then popcount.
 * The trailing observed_excess with expected = (64*s64) >> bbits is kept as in
 * the reference: for expected >= 1 it returns the raw count, for expected == 0
 * observed_excess(samebits, 0, maxnbits) is the identity.
 * ------------------------------------------------------------------------- */
size_t orc_calc_intersize(const uint64_t *s1, const uint64_t *s2,
                          size_t sketchsize64, size_t bbits) {
  size_t samebits = 0;
  for (size_t w = 0; w < sketchsize64; w++) {
    uint64_t bits = ~(uint64_t)0;
    for (size_t l = 0; l < bbits; l++) {
      bits &= ~(s1[w * bbits + l] ^ s2[w * bbits + l]);
    }
    samebits += (size_t)__builtin_popcountll(bits);
  }
  const size_t maxnbits = sketchsize64 * 64;
  const size_t expected = maxnbits >> bbits;
  if (expected) {
    return samebits;
  }
  /* observed_excess<size_t,size_t>(samebits, 0, maxnbits), dist/dist.hpp:19-24 */
  size_t diff = samebits > expected ? samebits - expected : 0;
  return diff * maxnbits / (maxnbits - expected);
}

/* ---------------------------------------------------------------------------
 * Reference::jaccard_dist(query, k, random_jaccard) -- reference.cpp:86-96
 * (file lines 222-232): J_obs = intersize / (64 * s64 of r1), then
 * observed_excess<double,int>(J_obs, r, 1) -- dist/dist.hpp:13-24.
 * ------------------------------------------------------------------------- */
double orc_jaccard(size_t intersize, size_t sketchsize64, double random_jaccard) {
  const size_t unionsize = 64 * sketchsize64;
  const double jaccard_obs = intersize / (double)unionsize;
  const int max = 1;
  const double diff = jaccard_obs > random_jaccard ? (jaccard_obs - random_jaccard) : 0;
  return diff * max / (max - random_jaccard);
}

/* ---------------------------------------------------------------------------
 * csrs -- random/random_match.cpp:318-336 (equal-base random match chance).
 * m = 2^((k-1)*2 + rc); j(l) = 1 - (1 - 1/m)^l; r = j1 j2 / (j1 + j2 - j1 j2).
 * The reference's shift is undefined for k >= 33; the oracle returns NaN there
 * and callers reject such k.
 * ------------------------------------------------------------------------- */
double orc_csrs_j(size_t k, int use_rc, size_t length) {
  const size_t shift = (k - 1) * 2 + (use_rc ? 1 : 0);
  if (k < 1 || shift >= 64) {
    return std::numeric_limits<double>::quiet_NaN();
  }
  const size_t match_chance = (size_t)1 << shift;
  return 1 - std::pow(1 - (double)1 / match_chance, (double)length);
}

double orc_csrs_combine(double j1, double j2) {
  double random_chance = 0;
  if (j1 > 0 && j2 > 0) {
    random_chance = (j1 * j2) / (j1 + j2 - j1 * j2);
  }
  return random_chance;
}

double orc_csrs(size_t k, int use_rc, size_t l1, size_t l2) {
  return orc_csrs_combine(orc_csrs_j(k, use_rc, l1), orc_csrs_j(k, use_rc, l2));
}

/* ---------------------------------------------------------------------------
 * fit_slope -- dist/linear_regression.cpp:16-60.
 *  valid = { i : dists[i] > tol }; window = [valid.front(), valid.back()]
 *  (interior invalid points ARE included); window < 2 -> failure (the
 *  reference prints and exit(1)s the process; here: return -1).
 *  Least squares of log(J) on k over the window; core = 1 - exp(slope) if
 *  slope < 0 else 0, acc = 1 - exp(intercept) if intercept < 0 else 0.
 * ------------------------------------------------------------------------- */
int orc_fit_slope(const float *kmers, const float *dists, int n_k, double tol,
                  float *core_out, float *acc_out) {
  int first = -1, last = -1;
  for (int i = 0; i < n_k; ++i) {
    if ((double)dists[i] > tol) {
      if (first < 0) first = i;
      last = i;
    }
  }
  *core_out = 0.0f;
  *acc_out = 0.0f;
  if (first < 0 || last - first + 1 < 2) {
    return -1;
  }
  const int n = last - first + 1;
  double sx = 0, sy = 0;
  for (int i = first; i <= last; ++i) {
    sx += (double)kmers[i];
    sy += std::log((double)dists[i]);
  }
  const double mx = sx / n, my = sy / n;
  double sxx = 0, sxy = 0;
  for (int i = first; i <= last; ++i) {
    const double dx = (double)kmers[i] - mx;
    sxx += dx * dx;
    sxy += dx * (std::log((double)dists[i]) - my);
  }
  const double slope = sxy / sxx;
  const double intercept = my - slope * mx;
  if (slope < 0) {
    *core_out = (float)(1.0 - std::exp(slope));
  }
  if (intercept < 0) {
    *acc_out = (float)(1.0 - std::exp(intercept));
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * apply_rc + nearest_neighbour -- random/random_match.cpp:338-359;
 * RandomMC::closest_cluster -- random_match.cpp:230-237.
 * Base composition (A,C,G,T) is strand-symmetrised when use_rc, cast to float,
 * and the nearest centroid row (float squared distance, first minimum) wins.
 * ------------------------------------------------------------------------- */
int orc_closest_cluster(const double *bases, int use_rc, const float *centroids,
                        int n_clusters) {
  double b[4] = {bases[0], bases[1], bases[2], bases[3]};
  if (use_rc) {
    b[0] = 0.5 * (b[0] + b[3]);
    b[1] = 0.5 * (b[1] + b[2]);
    b[2] = b[1];
    b[3] = b[0];
  }
  float v[4] = {(float)b[0], (float)b[1], (float)b[2], (float)b[3]};
  int best = 0;
  float best_d = std::numeric_limits<float>::infinity();
  for (int c = 0; c < n_clusters; ++c) {
    float d = 0.0f;
    for (int j = 0; j < 4; ++j) {
      const float e = centroids[c * 4 + j] - v[j];
      d += e * e;
    }
    if (d < best_d) {
      best_d = d;
      best = c;
    }
  }
  return best;
}

/* ---------------------------------------------------------------------------
 * Random-match model shared by the query loops.
 *  mode 0 NoAdjust (RandomMC(), random_match.cpp:49-51): r = 0.
 *  mode 1 EqualBase (RandomMC(use_rc)): r = csrs(k, use_rc, len(r1), len(q)),
 *         random_match.cpp:185-197.
 *  mode 2 MonteCarlo: r = matches[k](cluster(r1), cluster(q)) when k < max_k,
 *         else 0 (random_match.cpp:191-195).  table is [n_k][C][C] aligned
 *         with klist; k_lt_max[kk] carries the "k < max_k" test.
 * In self mode r1 = row sample i (api.cpp:209-215); in query mode r1 = the
 * reference sample and q = the query (api.cpp:255-267).
 * ------------------------------------------------------------------------- */
struct orc_random {
  int mode;
  int use_rc;
  int n_clusters;
  const float *table;
  const int *k_lt_max;
};

static double random_for(const orc_random *rnd, int kk, size_t k, size_t len_r1,
                         size_t len_q, uint16_t c_r1, uint16_t c_q) {
  if (rnd->mode == 0) return 0.0;
  if (rnd->mode == 1) return orc_csrs(k, rnd->use_rc, len_r1, len_q);
  if (!rnd->k_lt_max[kk]) return 0.0;
  const int C = rnd->n_clusters;
  return (double)rnd->table[(size_t)kk * C * C + (size_t)c_r1 * C + c_q];
}

/* square_to_condensed -- dist/matrix_idx.hpp:43-50 */
static inline long long sq2cond(long i, long j, long n) {
  return (n * i - ((i * (i + 1)) >> 1) + j - 1 - i);
}

/*
 * Per-pair distance: Reference::core_acc_dist -> regress_kmers -> jaccard_dist
 * (reference.cpp:104-138, dist/linear_regression.cpp:62-83).  r1 supplies the
 * sketch-size denominator.  Writes core/acc (or the K Jaccards when jaccard).
 * Returns -1 when the regression window has fewer than two points.
 */
static int pair_dist(const uint64_t *s_r1, const uint64_t *s_q, int n_k,
                     const int *klist, size_t s64, size_t bbits,
                     const orc_random *rnd, size_t len_rnd_r1, size_t len_rnd_q,
                     uint16_t c_rnd_r1, uint16_t c_rnd_q, double tol, int jaccard,
                     float *out, uint16_t *counts) {
  const size_t words = s64 * bbits;
  float dists[64];
  float kf[64];
  for (int kk = 0; kk < n_k; ++kk) {
    const size_t inter = orc_calc_intersize(s_r1 + kk * words, s_q + kk * words, s64, bbits);
    if (counts) counts[kk] = (uint16_t)inter;
    const double r = random_for(rnd, kk, (size_t)klist[kk], len_rnd_r1, len_rnd_q,
                                c_rnd_r1, c_rnd_q);
    dists[kk] = (float)orc_jaccard(inter, s64, r);
    kf[kk] = (float)klist[kk];
  }
  if (jaccard) {
    for (int kk = 0; kk < n_k; ++kk) out[kk] = dists[kk];
    return 0;
  }
  return orc_fit_slope(kf, dists, n_k, tol, &out[0], &out[1]);
}

/* ---------------------------------------------------------------------------
 * query_db -- api.cpp:140-291 (dense output).
 * Sketch arrays are [n][n_k][s64*bbits] uint64 in the reference's own word
 * layout (sketch/sketch.cpp:82-92).  Self mode (query == NULL): rows are the
 * condensed upper triangle, row = square_to_condensed(i, j, n) for i < j,
 * r1 = sample i (api.cpp:195-229).  Query mode: row = q * n_ref + r,
 * r1 = query for the Jaccard, r1 = ref for the random match (api.cpp:247-268).
 * Output is float32 row-major [rows][2] (core, acc) or [rows][n_k] (jaccard).
 * counts (optional) receives the per-k match counts, [rows][n_k] uint16.
 * Returns the number of pairs whose fit failed (first failing row in
 * *fail_row, -1 if none).
 * ------------------------------------------------------------------------- */
long long orc_query_db(const uint64_t *ref, long long n_ref, const uint64_t *query,
                       long long n_query, int n_k, const int *klist, size_t s64,
                       size_t bbits, const uint64_t *ref_len,
                       const uint64_t *query_len, const uint16_t *ref_cluster,
                       const uint16_t *query_cluster, int rmode, int use_rc,
                       int n_clusters, const float *table, const int *k_lt_max,
                       double tol, int jaccard, int num_threads, float *out,
                       uint16_t *counts, long long *fail_row) {
  orc_random rnd = {rmode, use_rc, n_clusters, table, k_lt_max};
  const size_t per_sample = (size_t)n_k * s64 * bbits;
  const int cols = jaccard ? n_k : 2;
  long long n_fail = 0;
  long long first_fail = -1;
  if (query == nullptr) {
    const long long n = n_ref;
#pragma omp parallel for schedule(dynamic, 5) num_threads(num_threads) reduction(+ : n_fail)
    for (long long i = 0; i < n; i++) {
      for (long long j = i + 1; j < n; j++) {
        const long long pos = sq2cond(i, j, n);
        int rc = pair_dist(ref + i * per_sample, ref + j * per_sample, n_k, klist,
                           s64, bbits, &rnd, ref_len[i], ref_len[j],
                           ref_cluster ? ref_cluster[i] : 0,
                           ref_cluster ? ref_cluster[j] : 0, tol, jaccard,
                           out + pos * cols, counts ? counts + pos * n_k : nullptr);
        if (rc) {
          n_fail++;
#pragma omp critical
          if (first_fail < 0 || pos < first_fail) first_fail = pos;
        }
      }
    }
  } else {
#pragma omp parallel for collapse(2) schedule(static) num_threads(num_threads) reduction(+ : n_fail)
    for (long long q = 0; q < n_query; q++) {
      for (long long r = 0; r < n_ref; r++) {
        const long long row = q * n_ref + r;
        int rc = pair_dist(query + q * per_sample, ref + r * per_sample, n_k, klist,
                           s64, bbits, &rnd, ref_len[r], query_len[q],
                           ref_cluster ? ref_cluster[r] : 0,
                           query_cluster ? query_cluster[q] : 0, tol, jaccard,
                           out + row * cols, counts ? counts + row * n_k : nullptr);
        if (rc) {
          n_fail++;
#pragma omp critical
          if (first_fail < 0 || row < first_fail) first_fail = row;
        }
      }
    }
  }
  if (fail_row) *fail_row = first_fail;
  return n_fail;
}

/* ---------------------------------------------------------------------------
 * query_db_sparse -- api.cpp:319-432 (self kNN).
 * Row i scans every j (i == j -> +inf), keeping the kNN smallest in a
 * std::priority_queue<SparseDist> max-heap: push when size < kNN or
 * d < top (strict), pop when size > kNN.  Rows are drained back to front so
 * each row is ascending.  The kept set and the order inside ties are those of
 * this standard library's heap, as in the reference.
 * jaccard: the row distance is 1 - J(klist[dist_col]) and J is restored at
 * the end (api.cpp:368-371, 423-429).
 * ------------------------------------------------------------------------- */
struct OrcSparseDist {
  float dist;
  long j;
};
static bool operator<(OrcSparseDist const &a, OrcSparseDist const &b) {
  return a.dist < b.dist;
}

long long orc_query_db_sparse(const uint64_t *ref, long long n_ref, int n_k,
                              const int *klist, size_t s64, size_t bbits,
                              const uint64_t *ref_len, const uint16_t *ref_cluster,
                              int rmode, int use_rc, int n_clusters,
                              const float *table, const int *k_lt_max, double tol,
                              int jaccard, int kNN, int dist_col, int num_threads,
                              long *i_vec, long *j_vec, float *dists) {
  orc_random rnd = {rmode, use_rc, n_clusters, table, k_lt_max};
  const size_t per_sample = (size_t)n_k * s64 * bbits;
  long long n_fail = 0;
#pragma omp parallel for schedule(static) num_threads(num_threads) reduction(+ : n_fail)
  for (long long i = 0; i < n_ref; i++) {
    std::priority_queue<OrcSparseDist> min_dists;
    for (long long j = 0; j < n_ref; j++) {
      float row_dist = std::numeric_limits<float>::infinity();
      if (i != j) {
        float o[64];
        if (jaccard) {
          const size_t words = s64 * bbits;
          const size_t inter = orc_calc_intersize(
              ref + i * per_sample + (size_t)dist_col * words,
              ref + j * per_sample + (size_t)dist_col * words, s64, bbits);
          const double r = random_for(&rnd, dist_col, (size_t)klist[dist_col],
                                      ref_len[i], ref_len[j],
                                      ref_cluster ? ref_cluster[i] : 0,
                                      ref_cluster ? ref_cluster[j] : 0);
          /* 1.0f - double is evaluated in double, then narrowed (api.cpp:370) */
          row_dist = (float)(1.0f - orc_jaccard(inter, s64, r));
        } else {
          if (pair_dist(ref + i * per_sample, ref + j * per_sample, n_k, klist, s64,
                        bbits, &rnd, ref_len[i], ref_len[j],
                        ref_cluster ? ref_cluster[i] : 0,
                        ref_cluster ? ref_cluster[j] : 0, tol, 0, o, nullptr)) {
            n_fail++;
          }
          row_dist = dist_col == 0 ? o[0] : o[1];
        }
      }
      if ((long long)min_dists.size() < kNN || row_dist < min_dists.top().dist) {
        OrcSparseDist nm = {row_dist, (long)j};
        min_dists.push(nm);
        if ((long long)min_dists.size() > kNN) {
          min_dists.pop();
        }
      }
    }
    const long long offset = i * (long long)kNN;
    for (int k = 0; k < kNN; ++k) i_vec[offset + k] = (long)i;
    for (int k = kNN - 1; k >= 0; --k) {
      OrcSparseDist e = min_dists.top();
      j_vec[offset + k] = e.j;
      dists[offset + k] = e.dist;
      min_dists.pop();
    }
  }
  if (jaccard) {
    for (long long x = 0; x < n_ref * (long long)kNN; ++x) dists[x] = 1.0f - dists[x];
  }
  return n_fail;
}

/*
 * Heap replay used to check kNN selection from a precomputed dense row set:
 * the same admission rule and std::priority_queue as orc_query_db_sparse, fed
 * from a square [n][n] distance matrix (diagonal ignored, treated as +inf).
 */
void orc_knn_from_square(const float *square, long long n, int kNN, long *i_vec,
                         long *j_vec, float *dists, int num_threads) {
#pragma omp parallel for schedule(static) num_threads(num_threads)
  for (long long i = 0; i < n; i++) {
    std::priority_queue<OrcSparseDist> min_dists;
    for (long long j = 0; j < n; j++) {
      float row_dist = i == j ? std::numeric_limits<float>::infinity() : square[i * n + j];
      if ((long long)min_dists.size() < kNN || row_dist < min_dists.top().dist) {
        OrcSparseDist nm = {row_dist, (long)j};
        min_dists.push(nm);
        if ((long long)min_dists.size() > kNN) min_dists.pop();
      }
    }
    const long long offset = i * (long long)kNN;
    for (int k = 0; k < kNN; ++k) i_vec[offset + k] = (long)i;
    for (int k = kNN - 1; k >= 0; --k) {
      OrcSparseDist e = min_dists.top();
      j_vec[offset + k] = e.j;
      dists[offset + k] = e.dist;
      min_dists.pop();
    }
  }
}

/*
 * The same heap over an explicit set of m dense rows [m][n]: row r scans
 * columns 0..n-1 with column row_ids[r] treated as +inf (row_ids[r] < 0: no
 * column excluded, as for a query row against a separate reference set).
 * i_vec receives row_ids[r] (or r when it is negative).
 */
void orc_knn_rows(const float *rows, const long *row_ids, long long m, long long n, int kNN,
                  long *i_vec, long *j_vec, float *dists, int num_threads) {
#pragma omp parallel for schedule(static) num_threads(num_threads)
  for (long long r = 0; r < m; r++) {
    std::priority_queue<OrcSparseDist> min_dists;
    const long self = row_ids[r];
    for (long long j = 0; j < n; j++) {
      float row_dist = j == self ? std::numeric_limits<float>::infinity() : rows[r * n + j];
      if ((long long)min_dists.size() < kNN || row_dist < min_dists.top().dist) {
        OrcSparseDist nm = {row_dist, (long)j};
        min_dists.push(nm);
        if ((long long)min_dists.size() > kNN) min_dists.pop();
      }
    }
    const long long offset = r * (long long)kNN;
    for (int k = 0; k < kNN; ++k) i_vec[offset + k] = self >= 0 ? self : (long)r;
    for (int k = kNN - 1; k >= 0; --k) {
      OrcSparseDist e = min_dists.top();
      j_vec[offset + k] = e.j;
      dists[offset + k] = e.dist;
      min_dists.pop();
    }
  }
}

/* ---------------------------------------------------------------------------
 * Sketching -- sketch/sketch.cpp:22-176 (bindash one-permutation sketch),
 * sketch/nthash.hpp:17-127 and stHashIterator.hpp:56-116 (canonical ntHash).
 *
 * ntHash in its split-rotation form (Mohamadi et al. 2016): rot(x) rotates
 * the high 31 bits and the low 33 bits of x left by one position each.  For
 * a k-mer s_0 .. s_{k-1},
 *   fh = XOR_i rot^(k-1-i)(seed(s_i)),   rh = XOR_i rot^i(seed(comp(s_i))),
 * and the hash is min(fh, rh) (canonical) or fh.  k-mers containing any
 * character other than A, C, G, T, U (either case) are skipped.  The seeds
 * are the published ntHash constants (nthash_tables.hpp:22-25).
 * Each hash is reduced mod 2^61 - 1; bin = hash / ceil((2^61 - 1) / nbins)
 * keeps its minimum; empty bins are filled by universal-hash probing
 * (densifybin, sketch.cpp:93-115); bit i of each bin minimum goes to bit
 * (bin % 64) of layer word (bin / 64) * 14 + i (fillusigs, sketch.cpp:80-91).
 * ------------------------------------------------------------------------- */
static const uint64_t ORC_SIGN_MOD = (1ULL << 61) - 1ULL;

static uint64_t orc_seed(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0x3c8bfbb395c60474ULL;
    case 'C': case 'c': return 0x3193c18562a02b4cULL;
    case 'G': case 'g': return 0x20323ed082572324ULL;
    case 'T': case 't': case 'U': case 'u': return 0x295549f54be24456ULL;
    default: return 0;
  }
}

static uint64_t orc_seed_comp(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return orc_seed('T');
    case 'C': case 'c': return orc_seed('G');
    case 'G': case 'g': return orc_seed('C');
    case 'T': case 't': case 'U': case 'u': return orc_seed('A');
    default: return 0;
  }
}

/* rot^j: rotate the high 31 bits and the low 33 bits left by j each */
static uint64_t orc_rot(uint64_t x, unsigned j) {
  uint64_t hi = x >> 33, lo = x & 0x1FFFFFFFFULL;
  const unsigned a = j % 31, b = j % 33;
  hi = a ? (((hi << a) | (hi >> (31 - a))) & 0x7FFFFFFFULL) : hi;
  lo = b ? (((lo << b) | (lo >> (33 - b))) & 0x1FFFFFFFFULL) : lo;
  return (hi << 33) | lo;
}

static uint64_t orc_univhash(uint64_t s, uint64_t t) {
  uint64_t x = 1009ULL * s + (1000ULL * 1000ULL + 3ULL) * t;
  return (48271ULL * x + 11ULL) % ((1ULL << 31) - 1ULL);
}

/* One sample, one k: seq holds the contigs back to back, contig c at
 * [off[c], off[c+1]).  usigs: [s64 * 14].  Returns the densify code
 * (0 none, 1 densified, -1 every bin empty); *inv_minhash = bin 0 / (2^61-1)
 * before densification (sketch.cpp:72-75). */
int orc_sketch_filtered(const char *seq, const long long *off, int n_contigs, int k, int s64,
                        int use_rc, int codon, int min_count, int exact, uint64_t *usigs,
                        double *inv_minhash);

int orc_sketch(const char *seq, const long long *off, int n_contigs, int k, int s64,
               int use_rc, uint64_t *usigs, double *inv_minhash) {
  return orc_sketch_filtered(seq, off, n_contigs, k, s64, use_rc, 0, 0, 0, usigs, inv_minhash);
}

/* ---------------------------------------------------------------------------
 * Read filtering -- sketch/countmin.cpp:37-85 with countmin.hpp:19-23 and the
 * multi-hash of sketch/nthash.hpp:180-240.  With min_count > 0 every valid
 * k-mer occurrence, in sequence order, is counted before it may be binned,
 * and it is binned only if the count after this occurrence is >= min_count.
 *  - exact: counts per 64-bit hash value (HashCounter), saturating at 255.
 *  - count-min: 4 rows of 2^27 - 1 uint8 cells laid out back to back; row r
 *    takes 27 bits of hash r / 2 (bits 0-26 for even r, 27-53 for odd r),
 *    where hash 1 = mix(hash 0) = t ^ (t >> 27), t = hash0 * (1 ^ k * seed).
 *    The hash iterator computes hash 1 only when it (re)starts, at the first
 *    valid k-mer of a sequence and after each invalid base
 *    (sketch/stHashIterator.hpp:64-112: init() asks NTMC64 for h2 = 2 hashes,
 *    next() for h = 1): hash 1 is that of the first k-mer of the current run
 *    of valid k-mers.
 *    The cell index can equal 2^27 - 1, one past its row: the flat position
 *    then is the next row's cell 0 (the reference's std::array of arrays),
 *    and for the last row one past the table (modelled here as one more
 *    cell starting at 0).  A cell below 255 is incremented; the estimate is
 *    the minimum of the incremented values, 255 if every cell is saturated.
 * ------------------------------------------------------------------------- */
static const uint64_t ORC_CM_MASK = 0x7FFFFFFULL;
static const uint64_t ORC_MULTI_SEED = 0x90b45d39fb6da1faULL;

static uint64_t orc_multi_hash(uint64_t h0, unsigned k, unsigned i) {
  uint64_t t = h0 * ((uint64_t)i ^ (uint64_t)k * ORC_MULTI_SEED);
  return t ^ (t >> 27);
}

struct OrcCounter {
  int exact;
  unsigned k;
  std::unordered_map<uint64_t, uint8_t> cells;  // flat cell (count-min) or hash (exact)

  unsigned add(uint64_t h0, uint64_t h1) {
    if (exact) {
      uint8_t &c = cells[h0];
      if (c < 255) ++c;
      return c;
    }
    unsigned est = 255;
    for (unsigned r = 0; r < 4; ++r) {
      const uint64_t hv = r < 2 ? h0 : h1;
      const uint64_t cell = (hv >> (27 * (r % 2))) & ORC_CM_MASK;
      uint8_t &c = cells[r * ORC_CM_MASK + cell];
      if (c < 255) {
        ++c;
        if (c < est) est = c;
      }
    }
    return est;
  }
};

/* codon: the codon-phased seed of generate_seeds (sketch/sketch.cpp:43-54),
 * 1001001 then 001 repeated, a span of 3k - 2 bases whose every third base
 * (offsets 0, 3, 6, ...) counts; the hash is the span's ntHash with the
 * other positions' terms removed (sketch/nthash.hpp:634-722), and every base
 * of the span must be valid.  The spaced-seed iterator recomputes the extra
 * hashes at every k-mer (stHashIterator.hpp:96-99). */
int orc_sketch_filtered(const char *seq, const long long *off, int n_contigs, int k, int s64,
                        int use_rc, int codon, int min_count, int exact, uint64_t *usigs,
                        double *inv_minhash) {
  const uint64_t nbins = (uint64_t)s64 * 64;
  const uint64_t binsize = (ORC_SIGN_MOD + nbins - 1ULL) / nbins;
  std::vector<uint64_t> signs(nbins, UINT64_MAX);
  if (min_count > 255) min_count = 255;  // create_sketches truncates to uint8
  const int span = codon ? 3 * k - 2 : k;
  OrcCounter counter{exact, (unsigned)span, {}};
  for (int c = 0; c < n_contigs; ++c) {
    const long long a = off[c], len = off[c + 1] - off[c];
    bool prev_ok = false;
    uint64_t run_h1 = 0;
    for (long long p = 0; p + span <= len; ++p) {
      uint64_t fh = 0, rh = 0;
      bool ok = true;
      for (int i = 0; i < span; ++i) {
        const unsigned char ch = (unsigned char)seq[a + p + i];
        const uint64_t sd = orc_seed(ch);
        if (!sd) { ok = false; break; }
        if (codon && i % 3 != 0) continue;
        fh ^= orc_rot(sd, (unsigned)(span - 1 - i));
        rh ^= orc_rot(orc_seed_comp(ch), (unsigned)i);
      }
      if (!ok) {
        prev_ok = false;
        continue;
      }
      const uint64_t h0 = (use_rc && rh < fh) ? rh : fh;
      if (!prev_ok || codon) run_h1 = orc_multi_hash(h0, (unsigned)span, 1);
      prev_ok = true;
      if (min_count > 0 && (int)counter.add(h0, run_h1) < min_count) continue;
      const uint64_t h = h0 % ORC_SIGN_MOD;
      const uint64_t bin = h / binsize;
      if (h < signs[bin]) signs[bin] = h;
    }
  }
  *inv_minhash = signs[0] / (double)ORC_SIGN_MOD;
  uint64_t mn = UINT64_MAX, mx = 0;
  for (uint64_t v : signs) {
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  int dens = 0;
  if (mx == UINT64_MAX) {
    if (mn == UINT64_MAX) {
      dens = -1;
    } else {
      for (uint64_t i = 0; i < nbins; ++i) {
        uint64_t j = i, attempts = 0;
        while (signs[j] == UINT64_MAX) {
          j = orc_univhash(i, attempts) % nbins;
          attempts++;
        }
        signs[i] = signs[j];
      }
      dens = 1;
    }
  }
  std::fill(usigs, usigs + (size_t)s64 * 14, 0ULL);
  for (uint64_t b = 0; b < nbins; ++b) {
    for (int i = 0; i < 14; ++i) {
      usigs[(b / 64) * 14 + i] |= ((signs[b] >> i) & 1ULL) << (b % 64);
    }
  }
  return dens;
}

/* ---------------------------------------------------------------------------
 * Index math -- dist/matrix_idx.hpp:13-50.
 * ------------------------------------------------------------------------- */
long long orc_square_to_condensed(long i, long j, long n) { return sq2cond(i, j, n); }

long orc_calc_row_idx(long long k, long n) {
  return n - 2 - (long)std::floor(std::sqrt((double)(-8 * k + 4 * (long long)n * (n - 1) - 7)) / 2 - 0.5);
}

long orc_calc_col_idx(long long k, long i, long n) {
  return (long)(k + i + 1 - (long long)n * (n - 1) / 2 + (long long)(n - i) * ((n - i) - 1) / 2);
}

/* rows_to_samples -- dist/matrix_idx.hpp:13-17 */
size_t orc_rows_to_samples(long long rows) {
  return (size_t)(0.5 * (1 + std::sqrt(1 + 8 * (double)rows)));
}

/* ---------------------------------------------------------------------------
 * long_to_square -- dist/matrix_ops.cpp:71-128.  out is [(nrr+nqq)^2] float,
 * row-major; qr may be NULL/0-length, qq may be 0-length (then one query).
 * Returns the side length, or -1 on a shape mismatch.
 * ------------------------------------------------------------------------- */
long long orc_long_to_square_size(long long n_rr, long long n_qr, long long n_qq) {
  long long nrr = (long long)orc_rows_to_samples(n_rr);
  long long nqq = 0;
  if (n_qr > 0) {
    nqq = (long long)orc_rows_to_samples(n_qq);
    if (nqq < 1) nqq = 1;
    if (n_qr != nrr * nqq) return -1;
  }
  return nrr + nqq;
}

long long orc_long_to_square(const float *rr, long long n_rr, const float *qr,
                             long long n_qr, const float *qq, long long n_qq,
                             float *out) {
  const long long nrr = (long long)orc_rows_to_samples(n_rr);
  const long long side = orc_long_to_square_size(n_rr, n_qr, n_qq);
  if (side < 0) return -1;
  const long long nqq = side - nrr;
  std::fill(out, out + side * side, 0.0f);
  for (long long d = 0; d < n_rr; d++) {
    const long i = orc_calc_row_idx(d, (long)nrr);
    const long j = orc_calc_col_idx(d, i, (long)nrr);
    out[i * side + j] = rr[d];
    out[j * side + i] = rr[d];
  }
  if (n_qq > 0) {
    for (long long d = 0; d < n_qq; d++) {
      const long i = orc_calc_row_idx(d, (long)nqq) + (long)nrr;
      const long j = orc_calc_col_idx(d, i - (long)nrr, (long)nqq) + (long)nrr;
      out[i * side + j] = qq[d];
      out[j * side + i] = qq[d];
    }
  }
  if (n_qr > 0) {
    for (long long d = 0; d < n_qr; d++) {
      const long long i = d % nrr;
      const long long j = d / nrr + nrr;
      out[i * side + j] = qr[d];
      out[j * side + i] = qr[d];
    }
  }
  return side;
}

/* square_to_long -- dist/matrix_ops.cpp:130-150 */
void orc_square_to_long(const float *square, long long n, float *out) {
  for (long long i = n - 2; i >= 0; i--) {
    for (long long j = i + 1; j < n; j++) {
      out[sq2cond((long)i, (long)j, (long)n)] = square[i * n + j];
    }
  }
}

/*
 * sparsify_dists_by_threshold -- dist/matrix_ops.cpp:38-69.  Upper triangle
 * (j > i) of a [rows][cols] row-major matrix, row-major order, d < cutoff.
 * Call with i_vec == NULL to get the entry count.  Returns -1 if cutoff < 0.
 */
long long orc_sparsify_threshold(const float *dense, long long rows, long long cols,
                                 float cutoff, long *i_vec, long *j_vec,
                                 float *dists) {
  if (cutoff < 0) return -1;
  long long len = 0;
  for (long long i = 0; i < rows; i++) {
    for (long long j = i + 1; j < cols; j++) {
      const float d = dense[i * cols + j];
      if (d < cutoff) {
        if (i_vec) {
          i_vec[len] = (long)i;
          j_vec[len] = (long)j;
          dists[len] = d;
        }
        len++;
      }
    }
  }
  return len;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

} /* extern "C" */
