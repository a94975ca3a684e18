/*
 * C entry points around the reference's own sketching core, built from its
 * sources with -DWEB_SKETCH (sketch/sketch.cpp, sketch/seqio.cpp,
 * sketch/countmin.cpp).  Test infrastructure: lets the oracle's restatement
 * and the HIP sketcher be cross-checked against the reference's compiled
 * bindash sketch and sequence reader.
 */
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sketch/sketch.hpp"

// Sketch of one sample given as n_seqs in-memory sequences (one per
// contig), one k-mer length, dense or codon-phased seeds.  usigs:
// [s64 * 14].  Returns 0, or -1 if the reference threw.
extern "C" int ref_sketch_seqs(const char *const *seqs, int n_seqs, int k, int s64,
                               int use_rc, int codon, uint64_t *usigs, double *inv_minhash,
                               int *densified) {
  try {
    std::vector<std::string> in;
    for (int i = 0; i < n_seqs; ++i) in.emplace_back(seqs[i]);
    SeqBuf buf(in);
    KmerSeeds seeds = generate_seeds({(size_t)k}, codon != 0);
    auto res = sketch(buf, (uint64_t)s64, seeds[(size_t)k], 14, codon != 0, use_rc != 0, 0,
                      false);
    const std::vector<uint64_t> &u = std::get<0>(res);
    std::memcpy(usigs, u.data(), u.size() * sizeof(uint64_t));
    *inv_minhash = std::get<1>(res);
    *densified = std::get<2>(res) ? 1 : 0;
    return 0;
  } catch (...) {
    return -1;
  }
}

// The reference's file reader (kseq over zlib): composition of the first
// 1M bases, total bases, N count, read flag and the number of sequences kept
// (those at least kmer_len long).
extern "C" int ref_read_stats(const char *path, int kmer_len, double *bases4, uint64_t *total,
                              uint64_t *n_count, int *reads, uint64_t *n_seqs) {
  try {
    SeqBuf buf(std::vector<std::string>{path}, (size_t)kmer_len);
    BaseComp<double> b = buf.get_composition();
    bases4[0] = b.a;
    bases4[1] = b.c;
    bases4[2] = b.g;
    bases4[3] = b.t;
    *total = (uint64_t)b.total;
    *n_count = buf.missing_bases();
    *reads = buf.is_reads() ? 1 : 0;
    *n_seqs = buf.nseqs();
    return 0;
  } catch (...) {
    return -1;
  }
}

// The sequences the reference's reader keeps from one file (upper-cased,
// at least kmer_len long), back to back in out[cap]; lens[i] per sequence.
// Returns the number of sequences, or -1 if the buffers are too small or
// the reference threw.
extern "C" int64_t ref_read_seqs(const char *path, int kmer_len, char *out, uint64_t cap,
                                 uint64_t *lens, uint64_t lens_cap) {
  try {
    SeqBuf buf(std::vector<std::string>{path}, (size_t)kmer_len);
    uint64_t used = 0, n = 0;
    // an empty buffer's iterator is unset (seqio.cpp reset): stop before it
    while (buf.nseqs() > 0 && !buf.eof()) {
      const std::string &s = *buf.getseq();
      if (n >= lens_cap || used + s.size() > cap) return -1;
      std::memcpy(out + used, s.data(), s.size());
      used += s.size();
      lens[n++] = s.size();
      buf.move_next_seq();
    }
    return (int64_t)n;
  } catch (...) {
    return -1;
  }
}

// The reference's Reference sketch of one file at one k-mer length
// (SeqBuf from the file keeping sequences at least keep_len long, then
// sketch(); min_count > 0 filters reads by count, exact picks the hash
// counter over the count-min sketch).  Returns 1 when the file holds no kept
// sequence (Reference throws there).
extern "C" int ref_sketch_file(const char *path, int keep_len, int k, int s64, int use_rc,
                               int codon, int min_count, int exact, uint64_t *usigs,
                               double *inv_minhash, int *densified) {
  try {
    SeqBuf buf(std::vector<std::string>{path}, (size_t)keep_len);
    if (buf.nseqs() == 0) return 1;
    KmerSeeds seeds = generate_seeds({(size_t)k}, codon != 0);
    if (min_count > 255) min_count = 255;
    auto res = sketch(buf, (uint64_t)s64, seeds[(size_t)k], 14, codon != 0, use_rc != 0,
                      (uint8_t)min_count, exact != 0);
    const std::vector<uint64_t> &u = std::get<0>(res);
    std::memcpy(usigs, u.data(), u.size() * sizeof(uint64_t));
    *inv_minhash = std::get<1>(res);
    *densified = std::get<2>(res) ? 1 : 0;
    return 0;
  } catch (...) {
    return -1;
  }
}
