// Sketch container and HDF5 sketch-database I/O (drop-in on-disk layout).
//
// Layout (src/database/database.cpp:48-112, 159-258; README.md:194-214):
//   /sketches                 attrs sketch_version (str), codon_phased (u8),
//                             reverse_complement (u8)
//   /sketches/<name>          attrs sketchsize64 (u64), bbits (u64), kmers
//                             (i32[K]), length (u64), missing_bases (u64),
//                             reads (u8), base_freq (f64[4], ACGT)
//   /sketches/<name>/<k>      u64[s64 * bbits], attr kmer-size (i32)
//   /random                   table_keys (str[]), table_values (u16[]),
//                             matches_keys (u64[]), matches_values (f32[],
//                             attr dims u64[3] = C, C, nkeys), centroids
//                             (f32[], attr dims u64[2] = C, 4), attrs k_min,
//                             k_max (u32), use_rc (u8)
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ppsk_host {

constexpr size_t DEFAULT_LENGTH = 3000000;  // src/database/database.hpp:22

// HDF5 failures (exported to Python as pp_sketchlib.HDF5Exception).
struct HDF5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// One sample's sketch: the accessors of Reference (src/reference.hpp:60-79).
struct Sketch {
  std::string name;
  size_t bbits = 14;
  size_t sketchsize64 = 156;
  size_t seq_length = 0;
  std::array<double, 4> bases{0.25, 0.25, 0.25, 0.25};
  unsigned long missing_bases = 0;
  bool reads = false;
  bool use_rc = true;  // always true for loaded sketches (src/reference.cpp:64)
  std::vector<size_t> kmers;                 // ascending
  std::vector<std::vector<uint64_t>> usigs;  // aligned with kmers

  const std::vector<uint64_t>& get_sketch(size_t k) const;
  void remove_kmer(size_t k);
};

// Reference's operator== on Reference compares names only
// (src/reference.hpp:86-89); vector equality is elementwise.
bool same_names(const std::vector<Sketch>& a, const std::vector<Sketch>& b);

// load_sketches (src/api.cpp:436-516): missing sample -> empty vector plus a
// message; requested k not in the DB -> empty vector; other read errors throw.
std::vector<Sketch> load_sketches(const std::string& db_name,
                                  const std::vector<std::string>& names,
                                  std::vector<size_t> kmer_lengths, bool messages);

// Names under /sketches in HDF5 link (name) order.
std::vector<std::string> list_sketch_names(const std::string& h5_path);

struct DbAttrs {
  std::string version;
  bool codon_phased = false;
};
DbAttrs read_db_attrs(const std::string& h5_path);

// Random-match tables as stored under /random.
struct RandomTables {
  bool present = false;
  bool use_rc = true;
  unsigned k_min = 0, k_max = 0;
  std::vector<std::string> table_keys;
  std::vector<uint16_t> table_values;
  std::vector<uint64_t> matches_keys;
  size_t n_clusters = 0;
  std::vector<float> matches;    // [nkeys][C][C]
  std::vector<float> centroids;  // [C][4]
};
RandomTables read_random(const std::string& h5_path);

// Writers (new_db / add_sketch / save_random, database.cpp:48-258).
void write_new_db(const std::string& h5_path, bool use_rc, bool codon_phased,
                  const std::string& version);
void write_sketches(const std::string& h5_path, const std::vector<Sketch>& sketches);
void write_random(const std::string& h5_path, const RandomTables& r);

}  // namespace ppsk_host
