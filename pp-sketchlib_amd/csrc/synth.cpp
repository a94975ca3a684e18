// Seeded synthetic sketch generator: sketches written directly in the
// reference's bit-sliced word layout (src/sketch/sketch.cpp:82-92), evolved
// down a random binary tree.
//
// Model.  The root draws every bin of every k-mer length as a fresh uniform
// 14-bit value.  Along an edge with core divergence dc and accessory loss da,
// a child's bin keeps its parent's value with probability
//   (1 - dc)^k * (1 - da)
// and otherwise takes a fresh uniform 14-bit value.  Edge rates are
// dc ~ U[0, dcore_max / D], da ~ U[0, dacc_max / D] with D = ceil(log2 n), so
// root-to-leaf totals stay within dcore_max / dacc_max on average and every
// pair keeps a measurable Jaccard at the longest k.  Leaves are assigned to
// sample indices through a seeded permutation.  Each sample also gets a
// sequence length (length_mean +/- 5%) and a near-uniform base composition.
//
// Every random stream is a counter-based hash of (seed, node, k index), so the
// output is identical for any thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ppsk.h"

namespace {

inline uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z = splitmix(z);
      s[i] = z;
    }
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  inline uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  inline double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Node {
  int32_t left = -1, right = -1;  // children, -1 for a leaf
  int64_t leaf = -1;              // leaf order index
  double dc = 0, da = 0;          // edge from parent
};

void build_tree(std::vector<Node>& nodes, int64_t n, uint64_t seed, double dcore_max,
                double dacc_max) {
  const double D = std::max(1.0, std::ceil(std::log2((double)std::max<int64_t>(n, 2))));
  Rng rng(splitmix(seed ^ 0x7265657472656574ull));
  struct Job {
    int32_t node;
    int64_t lo, hi;
  };
  nodes.clear();
  nodes.push_back(Node());
  std::vector<Job> stack = {{0, 0, n}};
  while (!stack.empty()) {
    Job j = stack.back();
    stack.pop_back();
    const int64_t size = j.hi - j.lo;
    if (size == 1) {
      nodes[j.node].leaf = j.lo;
      continue;
    }
    int64_t split = size / 2;
    if (size >= 4) {
      const int64_t jitter = (int64_t)(rng.next() % (uint64_t)(size / 2 + 1)) - size / 4;
      split = std::min<int64_t>(size - 1, std::max<int64_t>(1, split + jitter));
    }
    const int32_t l = (int32_t)nodes.size();
    nodes.push_back(Node());
    const int32_t r = (int32_t)nodes.size();
    nodes.push_back(Node());
    for (int32_t c : {l, r}) {
      nodes[c].dc = rng.uniform() * dcore_max / D;
      nodes[c].da = rng.uniform() * dacc_max / D;
    }
    nodes[j.node].left = l;
    nodes[j.node].right = r;
    stack.push_back({r, j.lo + split, j.hi});
    stack.push_back({l, j.lo, j.lo + split});
  }
}

// Bit-slice 64 bin values into 14 layer words.
inline void pack_word(const uint16_t* v, uint64_t* out) {
  uint64_t layers[14] = {0};
  for (int b = 0; b < 64; ++b) {
    const uint64_t x = v[b];
    for (int l = 0; l < 14; ++l) layers[l] |= ((x >> l) & 1ull) << b;
  }
  std::memcpy(out, layers, sizeof(layers));
}

void gen_k(const std::vector<Node>& nodes, const std::vector<int64_t>& perm, int64_t n,
           int32_t n_k, int32_t kk, int32_t k, int32_t s64, uint64_t seed, uint64_t* raw) {
  const int64_t bins = (int64_t)s64 * 64;
  // depth-first over the tree with one value buffer per depth level
  std::vector<std::vector<uint16_t>> level;
  struct Frame {
    int32_t node;
    int32_t depth;
  };
  std::vector<Frame> stack = {{0, 0}};
  while (!stack.empty()) {
    const Frame f = stack.back();
    stack.pop_back();
    if ((int)level.size() <= f.depth) level.emplace_back(bins);
    std::vector<uint16_t>& cur = level[f.depth];
    const Node& nd = nodes[f.node];
    Rng rng(splitmix(seed ^ splitmix(((uint64_t)f.node << 8) ^ (uint64_t)kk)));
    if (f.depth == 0) {
      for (int64_t b = 0; b < bins; ++b) cur[b] = (uint16_t)(rng.next() & 0x3fff);
    } else {
      const std::vector<uint16_t>& par = level[f.depth - 1];
      const double keep = std::pow(1.0 - nd.dc, (double)k) * (1.0 - nd.da);
      for (int64_t b = 0; b < bins; ++b) {
        const bool kept = rng.uniform() < keep;
        cur[b] = kept ? par[b] : (uint16_t)(rng.next() & 0x3fff);
      }
    }
    if (nd.leaf >= 0) {
      const int64_t s = perm[nd.leaf];
      uint64_t* out = raw + ((size_t)s * n_k + kk) * (size_t)s64 * 14;
      for (int32_t w = 0; w < s64; ++w) pack_word(&cur[(size_t)w * 64], out + (size_t)w * 14);
    } else {
      stack.push_back({nd.right, f.depth + 1});
      stack.push_back({nd.left, f.depth + 1});
    }
  }
  (void)n;
}

}  // namespace

extern "C" int ppsk_synth_sketches(int64_t n, int32_t n_k, const int32_t* klist,
                                   int32_t sketchsize64, uint64_t seed, double dcore_max,
                                   double dacc_max, double length_mean, uint64_t* raw,
                                   uint64_t* lengths, double* base_freq, int32_t n_threads) {
  if (n < 1 || n_k < 1 || !klist || sketchsize64 < 1 || !raw) return PPSK_ERR_ARG;
  std::vector<Node> nodes;
  build_tree(nodes, n, seed, dcore_max, dacc_max);
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  Rng prng(splitmix(seed ^ 0x7065726d7574ull));
  for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)(prng.next() % (uint64_t)(i + 1))]);
  if (n_threads < 1) n_threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::thread> th;
  for (int32_t t = 0; t < std::min(n_threads, n_k); ++t) {
    th.emplace_back([&, t]() {
      for (int32_t kk = t; kk < n_k; kk += std::min(n_threads, n_k)) {
        gen_k(nodes, perm, n, n_k, kk, klist[kk], sketchsize64, seed, raw);
      }
    });
  }
  for (auto& x : th) x.join();
  Rng mrng(splitmix(seed ^ 0x6d657461ull));
  for (int64_t i = 0; i < n; ++i) {
    if (lengths) {
      lengths[i] = (uint64_t)std::llround(length_mean * (0.95 + 0.1 * mrng.uniform()));
    }
    if (base_freq) {
      const double gc = 0.5 + 0.04 * (mrng.uniform() - 0.5);
      const double at = 1.0 - gc;
      const double skew = 0.01 * (mrng.uniform() - 0.5);
      base_freq[i * 4 + 0] = at / 2 + skew;
      base_freq[i * 4 + 1] = gc / 2 - skew;
      base_freq[i * 4 + 2] = gc / 2 + skew;
      base_freq[i * 4 + 3] = at / 2 - skew;
    }
  }
  return PPSK_OK;
}
