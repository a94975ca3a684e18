// pp_sketchlib: Python module with the reference's API
// (src/sketchlib_bindings.cpp:238-312): same function names, keyword
// arguments, defaults, return types and error messages.  Distances and
// sketches are computed by the HIP library (libppsk.so); `use_gpu` and
// `device_id` are accepted as in the reference, and a distance call that the
// reference rejects on its GPU path is rejected here in the same way.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <future>
#include <iostream>
#include <thread>

#include "api.hpp"
#include "ppsk.h"

namespace py = pybind11;
using namespace ppsk_host;

#ifndef PPSK_VERSION_INFO
#define PPSK_VERSION_INFO "2.1.5"
#endif
// Sketch-format identifier of pp-sketchlib 2.1.5 (sha1 of its sketcher
// sources, src/Makefile:103), so databases from that version compare equal.

namespace {

using farray = py::array_t<float, py::array::c_style>;

farray to_numpy(DenseMatrix&& m) {
  farray a({(py::ssize_t)m.rows, (py::ssize_t)m.cols});
  if (!m.data.empty()) std::memcpy(a.mutable_data(), m.data.data(), m.data.size() * sizeof(float));
  return a;
}

void require_ascending(const std::vector<size_t>& k) {
  for (size_t i = 1; i < k.size(); ++i) {
    if (k[i] <= k[i - 1]) throw std::runtime_error("Mismatching k-mer lengths in sketches");
  }
}

// get_random (src/api.cpp:533-537, src/database/database.cpp:195-226)
RandomMC get_random(const std::string& db_name, bool use_rc_default) {
  RandomTables t = read_random(db_name + ".h5");
  if (t.present) return RandomMC(t);
  std::cerr << "Could not find random match chances in database, "
               "calculating assuming equal base frequencies"
            << std::endl;
  return RandomMC(use_rc_default);
}

std::vector<Sketch> load_or_throw(const std::string& db, const std::vector<std::string>& names,
                                  const std::vector<size_t>& klist) {
  std::vector<Sketch> s = load_sketches(db, names, klist, false);
  if (s.empty()) throw std::runtime_error("Query with empty ref or query list!");
  return s;
}

py::tuple coo_to_py(sparse_coo&& c) {
  return py::make_tuple(py::cast(std::get<0>(c)), py::cast(std::get<1>(c)),
                        py::cast(std::get<2>(c)));
}

// Runs fn without the GIL on a worker thread while the calling thread
// watches for signals, as the reference's distance loops poll
// PyErr_CheckSignals (src/api.cpp:217-226, 285-287) and its GPU path polls
// progress and sets a kill flag (src/gpu/dist.cu:351-370): on a pending
// signal the device's calls are cancelled (ppsk_cancel), the worker returns
// early, the request is withdrawn (ppsk_cancel_reset) and the Python
// exception (KeyboardInterrupt for Ctrl-C) is raised here.  A call running
// for more than a second prints its progress to stderr, as the reference's
// GPU path does.
template <typename F>
auto interruptible(int device_id, F fn) -> decltype(fn()) {
  using R = decltype(fn());
  std::packaged_task<R()> task(std::move(fn));
  std::future<R> fut = task.get_future();
  bool interrupted = false;
  {
    py::gil_scoped_release nogil;
    std::thread worker(std::move(task));
    const auto t0 = std::chrono::steady_clock::now();
    auto shown = t0;
    bool printed = false;
    while (fut.wait_for(std::chrono::milliseconds(20)) != std::future_status::ready) {
      if (!interrupted) {
        py::gil_scoped_acquire gil;
        if (PyErr_CheckSignals() != 0) {
          interrupted = true;
          ppsk_cancel(device_id);
        }
      }
      const auto now = std::chrono::steady_clock::now();
      if (!interrupted && now - t0 > std::chrono::seconds(1) &&
          now - shown > std::chrono::milliseconds(250)) {
        uint64_t done = 0, total = 0;
        int32_t running = 0;
        if (ppsk_progress(device_id, &done, &total, &running, nullptr) == PPSK_OK && running &&
            total > 0) {
          std::fprintf(stderr, "%cProgress (GPU): %.1lf%%", 13,
                       100.0 * (double)std::min(done, total) / (double)total);
          printed = true;
          shown = now;
        }
      }
    }
    worker.join();
    if (interrupted) ppsk_cancel_reset(device_id);
    if (printed) std::fprintf(stderr, "%cProgress (GPU): 100.0%%\n", 13);
  }
  if (interrupted) throw py::error_already_set();
  return fut.get();
}

farray queryDatabase(const std::string& ref_db_name, const std::string& query_db_name,
                     const std::vector<std::string>& ref_names,
                     const std::vector<std::string>& query_names,
                     std::vector<size_t> kmer_lengths, bool random_correct, bool jaccard,
                     size_t num_threads, bool use_gpu, int device_id) {
  (void)num_threads;
  if (use_gpu && (jaccard || kmer_lengths.size() < 2)) {
    throw std::runtime_error("Extracting Jaccard distances not supported on GPU");
  }
  require_ascending(kmer_lengths);
  std::vector<Sketch> ref = load_or_throw(ref_db_name, ref_names, kmer_lengths);
  std::vector<Sketch> query;
  if (ref_db_name == query_db_name && ref_names == query_names) {
    query = ref;
  } else {
    query = load_or_throw(query_db_name, query_names, kmer_lengths);
  }
  RandomMC random = random_correct ? get_random(ref_db_name, ref[0].use_rc) : RandomMC();
  DenseMatrix d = interruptible(device_id, [&] {
    return query_db(ref, query, kmer_lengths, random, jaccard, device_id);
  });
  return to_numpy(std::move(d));
}

py::tuple querySelfSparse(const std::string& ref_db_name, const std::vector<std::string>& ref_names,
                          std::vector<size_t> kmer_lengths, bool random_correct, bool jaccard,
                          unsigned long kNN, float dist_cutoff, size_t dist_col,
                          size_t num_threads, bool use_gpu, int device_id) {
  (void)num_threads;
  if (use_gpu && (jaccard || kmer_lengths.size() < 2 || dist_col > 1)) {
    throw std::runtime_error("Extracting Jaccard distances not supported on GPU");
  }
  require_ascending(kmer_lengths);
  std::vector<Sketch> ref = load_or_throw(ref_db_name, ref_names, kmer_lengths);
  RandomMC random = random_correct ? get_random(ref_db_name, ref[0].use_rc) : RandomMC();
  sparse_coo out = interruptible(device_id, [&] {
    if (dist_cutoff > 0) {
      return query_db_threshold(ref, kmer_lengths, random, dist_cutoff, dist_col, device_id);
    }
    return query_db_sparse(ref, kmer_lengths, random, jaccard, (int)kNN, dist_col, device_id);
  });
  return coo_to_py(std::move(out));
}

double jaccardDist(const std::string& db_name, const std::string& sample1,
                   const std::string& sample2, size_t kmer_size, bool random_correct) {
  std::vector<Sketch> s = load_sketches(db_name, {sample1, sample2}, {kmer_size}, false);
  if (s.size() != 2) throw std::out_of_range("vector::_M_range_check");
  RandomMC random = random_correct ? get_random(db_name, s[0].use_rc) : RandomMC();
  random.check_present(s, true);
  return jaccard_dist(s[0], s[1], kmer_size, random, 0);
}

std::vector<float> flat(const farray& a) {
  return std::vector<float>(a.data(), a.data() + a.size());
}

farray longToSquare(const farray& distVec, unsigned num_threads) {
  return to_numpy(long_to_square(distVec.data(), distVec.size(), nullptr, 0, nullptr, 0, num_threads));
}

farray longToSquareMulti(const farray& distVec, const farray& qr, const farray& qq,
                         unsigned num_threads) {
  return to_numpy(long_to_square(distVec.data(), distVec.size(), qr.data(), qr.size(), qq.data(),
                                 qq.size(), num_threads));
}

farray squareToLong(const farray& distMat, unsigned num_threads) {
  if (distMat.ndim() != 2 || distMat.shape(0) != distMat.shape(1)) {
    throw std::runtime_error("square_to_long input must be a square matrix");
  }
  std::vector<float> v = square_to_long(distMat.data(), distMat.shape(0), num_threads);
  farray a(v.size());
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(float));
  return a;
}

py::tuple sparsifyDistsByThreshold(const farray& denseDists, float distCutoff, size_t num_threads) {
  if (denseDists.ndim() != 2) throw std::runtime_error("distMat must be a 2D matrix");
  return coo_to_py(sparsify_dists_by_threshold(denseDists.data(), denseDists.shape(0),
                                               denseDists.shape(1), distCutoff,
                                               (unsigned)num_threads));
}

py::tuple sparsifyDists(const farray&, float, int, size_t) {
  throw std::runtime_error(
      "pp_sketchlib.sparsifyDists was deprecated in v2.0.1\n"
      "Downgrade to v2.0.0, or update the calling package");
}

// Sketches on the GPU whatever use_gpu says (there is no CPU path).
void constructDatabase(const std::string& db_name, const std::vector<std::string>& names,
                       const std::vector<std::vector<std::string>>& files,
                       std::vector<size_t> kmer_lengths, size_t sketch_size, bool codon_phased,
                       bool calc_random, bool use_rc, size_t min_count, bool exact,
                       size_t num_threads, bool /*use_gpu*/, int device_id) {
  py::gil_scoped_release release;
  ppsk_host::construct_database(db_name, names, files, kmer_lengths, sketch_size, codon_phased,
                                calc_random, use_rc, min_count, exact, num_threads, device_id);
}

// addRandomToDb (src/sketchlib_bindings.cpp:189-202): Monte Carlo tables from
// the stored sketches, written under /random.
void addRandom(const std::string& db_name, const std::vector<std::string>& names,
               std::vector<size_t> klist, bool use_rc, size_t num_threads) {
  const DbAttrs attrs = read_db_attrs(db_name + ".h5");
  py::gil_scoped_release release;
  auto sketches = load_sketches(db_name, names, klist, false);
  if (sketches.size() != names.size()) {
    throw std::runtime_error("addRandom: some samples were not found in " + db_name);
  }
  ppsk_host::calculate_random(sketches, db_name, ppsk_host::DEFAULT_N_CLUSTERS,
                              ppsk_host::DEFAULT_N_MC, attrs.codon_phased, use_rc,
                              (int)num_threads, 0);
}

void writeSyntheticDatabase(const std::string& db_name, int64_t n_samples,
                            std::vector<size_t> klist, size_t sketch_size, uint64_t seed,
                            double dcore_max, double dacc_max, double length_mean,
                            const std::string& name_prefix) {
  require_ascending(klist);
  // sketch_size in bins, rounded to 64-bin words as the reference CLI does
  const size_t s64 = (size_t)std::llround(sketch_size / 64.0);
  write_synthetic_db(db_name, n_samples, klist, s64, seed, dcore_max, dacc_max, length_mean,
                     name_prefix);
}

void writeRandomMatches(const std::string& db_name, const std::vector<std::string>& table_keys,
                        const std::vector<uint16_t>& table_values,
                        const std::vector<uint64_t>& matches_keys,
                        py::array_t<float, py::array::c_style | py::array::forcecast> matches,
                        py::array_t<float, py::array::c_style | py::array::forcecast> centroids,
                        unsigned k_min, unsigned k_max, bool use_rc) {
  RandomTables t;
  t.present = true;
  t.use_rc = use_rc;
  t.k_min = k_min;
  t.k_max = k_max;
  t.table_keys = table_keys;
  t.table_values = table_values;
  t.matches_keys = matches_keys;
  if (matches.ndim() != 3 || (size_t)matches.shape(0) != matches_keys.size() ||
      matches.shape(1) != matches.shape(2)) {
    throw std::runtime_error("matches must be [n_keys][C][C]");
  }
  t.n_clusters = matches.shape(1);
  if (centroids.ndim() != 2 || (size_t)centroids.shape(0) != t.n_clusters || centroids.shape(1) != 4) {
    throw std::runtime_error("centroids must be [C][4]");
  }
  t.matches.assign(matches.data(), matches.data() + matches.size());
  t.centroids.assign(centroids.data(), centroids.data() + centroids.size());
  if (table_keys.size() != table_values.size()) throw std::runtime_error("table keys / values differ in length");
  write_random(db_name + ".h5", t);
}

std::vector<std::string> getSampleNames(const std::string& db_name) {
  return list_sketch_names(db_name + ".h5");
}

// One stored sample: attributes and the words of every k-mer length, as the
// reference's Reference accessors expose them (src/reference.hpp:60-79).
py::dict getSketch(const std::string& db_name, const std::string& name) {
  auto v = load_sketches(db_name, {name}, {}, false);
  if (v.empty()) throw std::runtime_error("sample " + name + " not found in " + db_name);
  const Sketch& s = v[0];
  py::dict d, words;
  for (size_t i = 0; i < s.kmers.size(); ++i) {
    py::array_t<uint64_t> a(s.usigs[i].size());
    std::memcpy(a.mutable_data(), s.usigs[i].data(), s.usigs[i].size() * sizeof(uint64_t));
    words[py::int_(s.kmers[i])] = a;
  }
  d["name"] = s.name;
  d["kmers"] = s.kmers;
  d["sketchsize64"] = s.sketchsize64;
  d["bbits"] = s.bbits;
  d["length"] = s.seq_length;
  d["missing_bases"] = s.missing_bases;
  d["reads"] = s.reads;
  d["base_freq"] = std::vector<double>(s.bases.begin(), s.bases.end());
  d["sketch"] = words;
  return d;
}

}  // namespace

PYBIND11_MODULE(pp_sketchlib, m) {
  m.doc() = "Sketch distances for PopPUNK (MI355X)";

  m.def("constructDatabase", &constructDatabase, "Create and save sketches", py::arg("db_name"),
        py::arg("samples"), py::arg("files"), py::arg("klist"), py::arg("sketch_size"),
        py::arg("codon_phased") = false, py::arg("calc_random") = true, py::arg("use_rc") = true,
        py::arg("min_count") = 0, py::arg("exact") = false, py::arg("num_threads") = 1,
        py::arg("use_gpu") = false, py::arg("device_id") = 0);

  m.def("queryDatabase", &queryDatabase, "Find distances between sketches; return all distances",
        py::arg("ref_db_name"), py::arg("query_db_name"), py::arg("rList"), py::arg("qList"),
        py::arg("klist"), py::arg("random_correct") = true, py::arg("jaccard") = false,
        py::arg("num_threads") = 1, py::arg("use_gpu") = false, py::arg("device_id") = 0);

  m.def("querySelfSparse", &querySelfSparse,
        "Find self-self distances between sketches; return a sparse matrix",
        py::arg("ref_db_name"), py::arg("rList"), py::arg("klist"),
        py::arg("random_correct") = true, py::arg("jaccard") = false, py::arg("kNN") = 0,
        py::arg("dist_cutoff") = 0.0f, py::arg("dist_col") = 0, py::arg("num_threads") = 1,
        py::arg("use_gpu") = false, py::arg("device_id") = 0);

  m.def("addRandom", &addRandom, "Add random match chances into older databases",
        py::arg("db_name"), py::arg("samples"), py::arg("klist"), py::arg("use_rc") = true,
        py::arg("num_threads") = 1);

  m.def("jaccardDist", &jaccardDist, "Calculate a raw Jaccard distance", py::arg("db_name"),
        py::arg("ref_name"), py::arg("query_name"), py::arg("kmer_length"),
        py::arg("random_correct") = false);

  m.def("squareToLong", &squareToLong, "Convert dense square matrices to long form",
        py::arg("distMat").noconvert(), py::arg("num_threads") = 1);

  m.def("longToSquare", &longToSquare, "Convert dense long form distance matrices to square form",
        py::arg("distVec").noconvert(), py::arg("num_threads") = 1);

  m.def("longToSquareMulti", &longToSquareMulti,
        "Convert dense long form distance matrices to square form (in three blocks)",
        py::arg("distVec").noconvert(), py::arg("query_ref_distVec").noconvert(),
        py::arg("query_query_distVec").noconvert(), py::arg("num_threads") = 1);

  m.def("sparsifyDistsByThreshold", &sparsifyDistsByThreshold,
        "Transform all distances into a sparse matrix", py::arg("distMat").noconvert(),
        py::arg("distCutoff") = 0, py::arg("num_threads") = 1);

  m.def("sparsifyDists", &sparsifyDists, "Transform all distances into a sparse matrix",
        py::arg("distMat").noconvert(), py::arg("distCutoff") = 0, py::arg("kNN") = 0,
        py::arg("num_threads") = 1);

  // additions of this library
  m.def("writeSyntheticDatabase", &writeSyntheticDatabase,
        "Write a database of seeded synthetic sketches (tree-evolved bins)", py::arg("db_name"),
        py::arg("n_samples"), py::arg("klist"), py::arg("sketch_size") = 10000,
        py::arg("seed") = 1, py::arg("dcore_max") = 0.05, py::arg("dacc_max") = 0.30,
        py::arg("length_mean") = 2000000.0, py::arg("name_prefix") = "sample");
  m.def("getSampleNames", &getSampleNames, "Sample names of a database, in HDF5 name order",
        py::arg("db_name"));
  m.def("getSketch", &getSketch, "Attributes and sketch words of one stored sample",
        py::arg("db_name"), py::arg("name"));
  m.def(
      "_xoshiro",
      [](uint64_t seed, size_t n, int jumps) {
        ppsk_host::Xoshiro256 g(seed);
        for (int j = 0; j < jumps; ++j) g.jump();
        py::array_t<uint64_t> a(n);
        for (size_t i = 0; i < n; ++i) a.mutable_data()[i] = g();
        return a;
      },
      "xoshiro256** stream of the random genome generator", py::arg("seed"), py::arg("n"),
      py::arg("jumps") = 0);
  m.def(
      "_kmeans",
      [](py::array_t<double, py::array::c_style | py::array::forcecast> x, size_t k, int seed,
         int max_iter) {
        if (x.ndim() != 2) throw std::runtime_error("x must be 2-d");
        const size_t n = x.shape(0), d = x.shape(1);
        std::vector<double> v(x.data(), x.data() + n * d);
        auto r = ppsk_host::kmeans_plusplus(v, n, d, k, max_iter, seed);
        py::array_t<double> c({k, d});
        std::copy(r.centroids.begin(), r.centroids.end(), c.mutable_data());
        return py::make_tuple(c, r.assignment, r.iterations);
      },
      "k-means (kmeans++ seeding, Lloyd) as the random-match clustering runs it", py::arg("x"),
      py::arg("k"), py::arg("seed") = 3019, py::arg("max_iter") = 300);
  m.def(
      "_randomGenomes",
      [](py::array_t<float, py::array::c_style | py::array::forcecast> centroids, unsigned n_mc,
         size_t length, bool use_rc, int num_threads) {
        if (centroids.ndim() != 2 || centroids.shape(1) != 4) {
          throw std::runtime_error("centroids must be [C][4]");
        }
        std::vector<float> c(centroids.data(), centroids.data() + centroids.size());
        std::vector<std::string> g;
        {
          py::gil_scoped_release release;
          g = ppsk_host::random_genomes(c, centroids.shape(0), n_mc, length, use_rc, num_threads);
        }
        return g;
      },
      "Random genomes of the Monte Carlo tables", py::arg("centroids"), py::arg("n_mc"),
      py::arg("length"), py::arg("use_rc"), py::arg("num_threads"));
  m.def(
      "_readRandom",
      [](const std::string& db_name) {
        auto r = read_random(db_name + ".h5");
        py::dict d;
        d["present"] = r.present;
        if (!r.present) return d;
        const size_t C = r.n_clusters;
        py::array_t<float> m({r.matches_keys.size(), C, C});
        std::copy(r.matches.begin(), r.matches.end(), m.mutable_data());
        py::array_t<float> c({C, (size_t)4});
        std::copy(r.centroids.begin(), r.centroids.end(), c.mutable_data());
        d["use_rc"] = r.use_rc;
        d["k_min"] = r.k_min;
        d["k_max"] = r.k_max;
        d["table_keys"] = r.table_keys;
        d["table_values"] = r.table_values;
        d["matches_keys"] = r.matches_keys;
        d["matches"] = m;
        d["centroids"] = c;
        return d;
      },
      "The /random group of a database", py::arg("db_name"));
  m.def(
      "_readSequences",
      [](const std::vector<std::string>& files, size_t kmer_len) {
        auto st = ppsk_host::read_sequence_stats(files, kmer_len);
        return py::make_tuple(std::vector<double>(st.bases.begin(), st.bases.end()), st.total,
                              st.missing_bases, st.reads, st.contigs);
      },
      "What the sequence reader keeps of one sample's files", py::arg("files"),
      py::arg("kmer_len"));
  m.def(
      "_interruptProbe",
      [](double seconds, int device_id) {
        // stands in for a long distance call: waits until the device's calls
        // are cancelled (True) or `seconds` pass (False), under the same
        // signal watch as queryDatabase / querySelfSparse
        return interruptible(device_id, [=] {
          const auto t0 = std::chrono::steady_clock::now();
          while (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() <
                 seconds) {
            int32_t pending = 0;
            if (ppsk_progress(device_id, nullptr, nullptr, nullptr, &pending) != PPSK_OK) {
              throw std::runtime_error(ppsk_last_error());
            }
            if (pending) return true;
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
          }
          return false;
        });
      },
      "Signal handling of the distance calls, without device work", py::arg("seconds"),
      py::arg("device_id") = 0);
  m.def("writeRandomMatches", &writeRandomMatches,
        "Store Monte Carlo random-match tables in a database's /random group",
        py::arg("db_name"), py::arg("table_keys"), py::arg("table_values"),
        py::arg("matches_keys"), py::arg("matches"), py::arg("centroids"), py::arg("k_min"),
        py::arg("k_max"), py::arg("use_rc") = true);

  m.attr("version") = PPSK_VERSION_INFO;
  m.attr("sketchVersion") = ppsk_host::SKETCH_VERSION;

  py::register_exception<HDF5Error>(m, "HDF5Exception");
}
