// constructDatabase: read sequence files, sketch on the GPU (ppsk_sketch),
// write the reference's HDF5 layout.
//
// Reading follows the reference's SeqBuf (src/sketch/seqio.cpp:55-111) over
// kseq's record rules (src/sketch/kseq.h, __KSEQ_READ):
//  * a record starts at '>' or '@'; its sequence is every following line up
//    to a line starting with '>', '@' or '+' (empty lines skipped, a trailing
//    '\r' dropped);
//  * after '+', quality lines are read until they are as long as the
//    sequence; a length mismatch ends the file (kseq_read returns -2);
//  * records shorter than the largest k-mer length are dropped; kept ones are
//    upper-cased; the composition counts A, C, G, T/U among the first
//    999,999 bases (the counter is tested after its increment), divided by
//    min(total, 1,000,000); N is counted over all kept bases;
//  * any quality string marks the sample as reads.
// Sketching options follow create_sketches (src/api.cpp:39-138): an existing
// database holding every requested sample is loaded instead of re-sketched.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "api.hpp"
#include "sketch_db.hpp"
#include "../../../include/ppsk.h"

namespace ppsk_host {

namespace {

constexpr size_t MAX_COMPOSITION_SAMPLE = 1000000;  // seqio.cpp:24

struct SampleSeq {
  std::vector<std::string> contigs;
  size_t total = 0, a = 0, c = 0, g = 0, t = 0;
  unsigned long n_count = 0;
  bool reads = false;
};

// The whole (possibly gzipped) file, read straight into a buffer sized from
// the file (doubled as a compressed file inflates).  A plain file (no gzip
// magic bytes) is read with one fread.
std::string read_all(const std::string& path) {
  if (FILE* fh = std::fopen(path.c_str(), "rb")) {
    unsigned char magic[2] = {0, 0};
    const size_t m = std::fread(magic, 1, 2, fh);
    long size = -1;
    if (!(m == 2 && magic[0] == 0x1f && magic[1] == 0x8b) && std::fseek(fh, 0, SEEK_END) == 0) {
      size = std::ftell(fh);
    }
    if (size >= 0 && std::fseek(fh, 0, SEEK_SET) == 0) {
      std::string buf((size_t)size, '\0');
      const size_t got = size > 0 ? std::fread(&buf[0], 1, (size_t)size, fh) : 0;
      const bool err = std::ferror(fh) != 0;
      std::fclose(fh);
      if (err) throw std::runtime_error("Could not read " + path);
      buf.resize(got);
      return buf;
    }
    std::fclose(fh);
  }
  gzFile fp = gzopen(path.c_str(), "r");
  if (!fp) throw std::runtime_error("Could not open " + path);
  gzbuffer(fp, 1 << 20);
  struct stat st;
  size_t cap = stat(path.c_str(), &st) == 0 ? (size_t)st.st_size + 1 : (size_t)1 << 20;
  std::string buf(std::max<size_t>(cap, 1 << 16), '\0');
  size_t len = 0;
  int n = 0;
  for (;;) {
    if (len == buf.size()) buf.resize(buf.size() * 2);
    const size_t want = std::min<size_t>(buf.size() - len, (size_t)1 << 30);
    n = gzread(fp, &buf[len], (unsigned)want);
    if (n <= 0) break;
    len += (size_t)n;
  }
  gzclose(fp);
  if (n < 0) throw std::runtime_error("Could not read " + path);
  buf.resize(len);
  return buf;
}

// Upper-cases a kept contig in place and adds it to the sample's counts:
// A, C, G, T/U over the sample's first 999,999 bases (the reference's counter
// is incremented before its test against 1,000,000, so a base at index x of
// this contig counts while total + x + 1 < 1,000,000), N over every base.
void add_contig(std::string& seq, SampleSeq& out) {
  char* p = &seq[0];
  const size_t n = seq.size();
  for (size_t x = 0; x < n; ++x) {
    const char ch = p[x];
    p[x] = ('a' <= ch && ch <= 'z') ? (char)(ch ^ 0x20) : ch;
  }
  const size_t lim = out.total + 1 < MAX_COMPOSITION_SAMPLE
                         ? std::min(n, MAX_COMPOSITION_SAMPLE - 1 - out.total)
                         : 0;
  size_t a = 0, c = 0, g = 0, t = 0;
  for (size_t x = 0; x < lim; ++x) {
    const char ch = p[x];
    a += ch == 'A';
    c += ch == 'C';
    g += ch == 'G';
    t += (ch == 'T') | (ch == 'U');
  }
  size_t nn = 0;
  for (size_t x = 0; x < n; ++x) nn += p[x] == 'N';
  out.a += a;
  out.c += c;
  out.g += g;
  out.t += t;
  out.n_count += nn;
  out.total += n;
}

// Appends the line starting at i (without its '\n') to `to`, dropping a
// trailing '\r' when `to` is then longer than one character (kseq's test is
// on the accumulated string); i moves past the line.
void append_line(const std::string& b, size_t& i, std::string& to) {
  const size_t e = b.find('\n', i);
  const size_t end = e == std::string::npos ? b.size() : e;
  to.append(b, i, end - i);
  if (to.size() > 1 && to.back() == '\r') to.pop_back();
  i = e == std::string::npos ? b.size() : e + 1;
}

void read_file(const std::string& path, size_t kmer_len, SampleSeq& out) {
  const std::string b = read_all(path);
  size_t i = 0;
  std::string skip;
  for (;;) {
    // to the next header character (anywhere, as kseq after a quality string)
    while (i < b.size() && b[i] != '>' && b[i] != '@') ++i;
    if (i >= b.size()) break;
    ++i;                         // header character
    skip.clear();
    append_line(b, i, skip);     // name and comment
    std::string seq;
    while (i < b.size() && b[i] != '>' && b[i] != '+' && b[i] != '@') {
      if (b[i] == '\n') {
        ++i;
        continue;
      }
      append_line(b, i, seq);
    }
    bool qual = false;
    if (i < b.size() && b[i] == '+') {
      skip.clear();
      append_line(b, i, skip);   // '+' line
      // kseq reads at least one quality line, then more while it is short
      std::string q;
      do {
        if (i >= b.size()) break;
        append_line(b, i, q);
      } while (q.size() < seq.size());
      if (q.size() != seq.size()) return;  // kseq: a quality of another length ends the file
      qual = !q.empty();
    }
    // the reader keeps the C string (seqio.cpp:68-73): up to a NUL
    if (const void* z = std::memchr(seq.data(), 0, seq.size())) {
      seq.resize((size_t)((const char*)z - seq.data()));
    }
    if (seq.size() >= kmer_len) {
      add_contig(seq, out);
      out.contigs.push_back(std::move(seq));
    }
    if (qual) out.reads = true;
  }
}

}  // namespace

SequenceStats read_sequence_stats(const std::vector<std::string>& files, size_t kmer_len) {
  SampleSeq q;
  for (const auto& f : files) read_file(f, kmer_len, q);
  SequenceStats st;
  const double tot = (double)std::min(q.total, MAX_COMPOSITION_SAMPLE);
  st.bases = {q.a / tot, q.c / tot, q.g / tot, q.t / tot};
  st.total = q.total;
  st.missing_bases = q.n_count;
  st.reads = q.reads;
  st.contigs = std::move(q.contigs);
  return st;
}

namespace {

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

}  // namespace

// Reads and sketches every sample, writes a new database; returns the
// sketches as written.
std::vector<Sketch> sketch_files(const std::string& h5, const std::vector<std::string>& names,
                                 const std::vector<std::vector<std::string>>& files,
                                 const std::vector<size_t>& kmer_lengths, size_t sketchsize64,
                                 bool codon_phased, bool use_rc, size_t min_count, bool exact,
                                 size_t num_threads, int device_id);

void construct_database(const std::string& db_name, const std::vector<std::string>& names,
                        const std::vector<std::vector<std::string>>& files,
                        std::vector<size_t> kmer_lengths, size_t sketchsize64, bool codon_phased,
                        bool calc_random, bool use_rc, size_t min_count, bool exact,
                        size_t num_threads, int device_id) {
  if (names.size() != files.size()) throw std::runtime_error("names and files differ in length");
  std::sort(kmer_lengths.begin(), kmer_lengths.end());
  if (kmer_lengths.empty() || kmer_lengths.front() < 3) {
    throw std::runtime_error("Minimum k must be 3 or higher");
  }
  const std::string h5 = db_name + ".h5";
  std::vector<Sketch> all;
  bool resketch = true;
  if (file_exists(h5)) {
    // create_sketches reuses a database that already holds every sample
    all = load_sketches(db_name, names, kmer_lengths, false);
    resketch = all.size() != names.size();
  }
  if (resketch) {
    all = sketch_files(h5, names, files, kmer_lengths, sketchsize64, codon_phased, use_rc,
                       min_count, exact, num_threads, device_id);
  }
  // constructDatabase (src/sketchlib_bindings.cpp:77-86)
  if (calc_random) {
    if (all.size() >= DEFAULT_N_CLUSTERS) {
      calculate_random(all, db_name, DEFAULT_N_CLUSTERS, DEFAULT_N_MC, codon_phased, use_rc,
                       (int)num_threads, device_id);
    } else {
      std::cerr << "Too few input genomes to calculate random match chances" << std::endl;
    }
  }
}

namespace {

// Sketch words of samples [s0, s1) from one ppsk_sketch_ex call.
struct SketchedBatch {
  size_t s0 = 0, s1 = 0;
  std::vector<uint64_t> raw;  // [n][K][words]
  std::vector<double> mh;     // mean bin-0 minhash (reads)
  std::vector<int32_t> dn;    // densified
};

// Samples [c0, c1) (already read) on the device, in batches of one kind
// (count-filtered or not) bounded by their bases.  The contigs of each
// sample are released once copied into the batch buffer.
int sketch_range(std::vector<SampleSeq>& seqs, size_t c0, size_t c1,
                 const std::vector<int32_t>& kl, size_t sketchsize64, bool use_rc,
                 bool codon_phased, int32_t mc, bool exact, size_t num_threads, int device_id,
                 std::vector<SketchedBatch>& out, std::string& err) {
  const size_t K = kl.size(), words = sketchsize64 * 14;
  auto filtered = [&](size_t s) { return seqs[s].reads && mc > 0; };
  const size_t batch_bases = (size_t)1 << 31;
  size_t s0 = c0;
  while (s0 < c1) {
    size_t s1 = s0, bases = 0;
    while (s1 < c1 && (s1 == s0 || (bases + seqs[s1].total <= batch_bases &&
                                    filtered(s1) == filtered(s0)))) {
      bases += seqs[s1].total;
      ++s1;
    }
    // the batch's contigs back to back: offsets first, then the copies in
    // parallel into an uninitialised buffer
    std::vector<int64_t> off{0}, sc{0};
    std::vector<size_t> first_contig;
    for (size_t s = s0; s < s1; ++s) {
      first_contig.push_back(off.size() - 1);
      for (const auto& c : seqs[s].contigs) off.push_back(off.back() + (int64_t)c.size());
      sc.push_back((int64_t)off.size() - 1);
    }
    const size_t seq_len = (size_t)off.back();
    std::unique_ptr<char[]> seq(new char[std::max<size_t>(seq_len, 1)]);
#pragma omp parallel for schedule(dynamic, 1) num_threads(std::max<size_t>(num_threads, 1))
    for (long x = 0; x < (long)(s1 - s0); ++x) {
      size_t ci = first_contig[x];
      for (const auto& c : seqs[s0 + x].contigs) std::memcpy(seq.get() + off[ci++], c.data(), c.size());
      std::vector<std::string>().swap(seqs[s0 + x].contigs);
    }
    SketchedBatch b;
    b.s0 = s0;
    b.s1 = s1;
    const int64_t n = (int64_t)(s1 - s0);
    b.raw.resize((size_t)n * K * words);
    b.mh.resize(n);
    b.dn.resize(n);
    const int rc = ppsk_sketch_ex(device_id, n, seq.get(), (int64_t)seq_len, off.data(),
                                  (int64_t)off.size() - 1, sc.data(), (int32_t)K, kl.data(),
                                  (int32_t)sketchsize64, use_rc ? 1 : 0, codon_phased ? 1 : 0,
                                  filtered(s0) ? mc : 0, exact ? 1 : 0, b.raw.data(),
                                  b.mh.data(), b.dn.data());
    if (rc != PPSK_OK) {
      err = ppsk_last_error();  // thread-local to this thread
      return rc;
    }
    out.push_back(std::move(b));
    s0 = s1;
  }
  return PPSK_OK;
}

}  // namespace

// As create_sketches (src/api.cpp:39-138): every sample is read in parallel
// (errors collected for all of them, then reported and thrown), sketched on
// the device, and only then the new database is created and written.
std::vector<Sketch> sketch_files(const std::string& h5, const std::vector<std::string>& names,
                                 const std::vector<std::vector<std::string>>& files,
                                 const std::vector<size_t>& kmer_lengths, size_t sketchsize64,
                                 bool codon_phased, bool use_rc, size_t min_count, bool exact,
                                 size_t num_threads, int device_id) {
  std::cerr << "Sketching " << names.size() << " genomes on GPU device " << device_id << std::endl;
  const size_t n = names.size();
  const size_t threads = std::max<size_t>(num_threads, 1);
  std::vector<SampleSeq> seqs(n);
  std::vector<std::string> errors(n);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (long i = 0; i < (long)n; ++i) {
    try {
      for (const auto& f : files[i]) read_file(f, kmer_lengths.back(), seqs[i]);
      if (seqs[i].contigs.empty()) errors[i] = names[i] + " contains no sequence";
    } catch (const std::exception& e) {
      errors[i] = e.what();
    }
  }
  // messages on stdout, then throw (api.cpp:100-112)
  bool any_error = false;
  for (const auto& e : errors) {
    if (!e.empty()) {
      std::cout << e << std::endl;
      any_error = true;
    }
  }
  if (any_error) throw std::runtime_error("Errors during sketching");

  std::vector<int32_t> kl(kmer_lengths.begin(), kmer_lengths.end());
  // reads are counted before binning when min_count > 0 (sketch.cpp:128-149;
  // create_sketches truncates min_count to uint8, api.cpp:66-69)
  const int32_t mc = (int32_t)std::min<size_t>(min_count, 255);
  std::vector<SketchedBatch> batches;
  std::string device_err;
  if (sketch_range(seqs, 0, n, kl, sketchsize64, use_rc, codon_phased, mc, exact, threads,
                   device_id, batches, device_err) != PPSK_OK) {
    throw std::runtime_error(device_err);
  }

  if (codon_phased) std::cerr << "NB: codon phased seeds are ON" << std::endl;
  const size_t K = kl.size(), words = sketchsize64 * 14;
  std::vector<Sketch> all;
  all.reserve(n);
  for (const SketchedBatch& b : batches) {
    for (size_t x = 0; x < b.s1 - b.s0; ++x) {
      const SampleSeq& q = seqs[b.s0 + x];
      Sketch s;
      s.name = names[b.s0 + x];
      s.sketchsize64 = sketchsize64;
      s.bbits = 14;
      s.reads = q.reads;
      s.use_rc = use_rc;
      s.missing_bases = q.n_count;
      const double tot = (double)std::min(q.total, MAX_COMPOSITION_SAMPLE);
      s.bases = {q.a / tot, q.c / tot, q.g / tot, q.t / tot};
      // reads: size from the mean bin-0 minhash over k (reference.cpp:49-55)
      s.seq_length = q.reads ? (size_t)((double)K / b.mh[x]) : q.total;
      s.kmers = kmer_lengths;
      for (size_t kk = 0; kk < K; ++kk) {
        const uint64_t* src = b.raw.data() + (x * K + kk) * words;
        s.usigs.emplace_back(src, src + words);
      }
      if (b.dn[x]) std::cerr << "NOTE: " << s.name << " required densification" << std::endl;
      all.push_back(std::move(s));
    }
  }
  write_new_db(h5, use_rc, codon_phased, SKETCH_VERSION);
  write_sketches(h5, all);
  return all;
}

}  // namespace ppsk_host
