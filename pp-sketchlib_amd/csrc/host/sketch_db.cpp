// HDF5 sketch database reader / writer on the HDF5 C API.
#include "sketch_db.hpp"

#include <hdf5.h>

#include <algorithm>
#include <iostream>
#include <sstream>

namespace ppsk_host {

namespace {

// RAII wrapper for an hid_t with its close function.
struct H5Handle {
  hid_t id = -1;
  herr_t (*closer)(hid_t) = nullptr;
  H5Handle() = default;
  H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
  H5Handle(const H5Handle&) = delete;
  H5Handle& operator=(const H5Handle&) = delete;
  H5Handle(H5Handle&& o) noexcept : id(o.id), closer(o.closer) { o.id = -1; }
  ~H5Handle() {
    if (id >= 0 && closer) closer(id);
  }
  bool ok() const { return id >= 0; }
  operator hid_t() const { return id; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// File access for writing sketches: a 4 MiB sieve buffer gathers the many
// small dataset writes, and metadata and small raw data are allocated from
// 1 MiB blocks; this changes where objects sit in the file, not what it holds
// (64 samples of 8 k-mer lengths: about 15 ms instead of 18-26 ms).
H5Handle write_access() {
  H5Handle fapl(H5Pcreate(H5P_FILE_ACCESS), H5Pclose);
  H5Pset_sieve_buf_size(fapl, (size_t)4 << 20);
  H5Pset_meta_block_size(fapl, (hsize_t)1 << 20);
  H5Pset_small_data_block_size(fapl, (hsize_t)1 << 20);
  return fapl;
}

H5Handle open_file(const std::string& path, bool write) {
  quiet();
  H5Handle fapl = write ? write_access() : H5Handle(H5P_DEFAULT, nullptr);
  hid_t f = H5Fopen(path.c_str(), write ? H5F_ACC_RDWR : H5F_ACC_RDONLY,
                    write ? (hid_t)fapl : H5P_DEFAULT);
  if (f < 0) throw HDF5Error("Unable to open file " + path);
  return H5Handle(f, H5Fclose);
}

H5Handle open_group(hid_t loc, const std::string& name) {
  hid_t g = H5Gopen2(loc, name.c_str(), H5P_DEFAULT);
  if (g < 0) throw HDF5Error("Unable to open the group \"" + name + "\"");
  return H5Handle(g, H5Gclose);
}

bool has_attr(hid_t obj, const char* name) { return H5Aexists(obj, name) > 0; }

template <typename T>
std::vector<T> read_attr_vec(hid_t obj, const char* name, hid_t memtype) {
  H5Handle a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) throw HDF5Error(std::string("Unable to open the attribute \"") + name + "\"");
  H5Handle sp(H5Aget_space(a), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<T> v((size_t)std::max<hssize_t>(n, 0));
  if (!v.empty() && H5Aread(a, memtype, v.data()) < 0) {
    throw HDF5Error(std::string("Unable to read the attribute \"") + name + "\"");
  }
  return v;
}

template <typename T>
T read_attr(hid_t obj, const char* name, hid_t memtype) {
  auto v = read_attr_vec<T>(obj, name, memtype);
  if (v.empty()) throw HDF5Error(std::string("Empty attribute \"") + name + "\"");
  return v[0];
}

std::string read_str_attr(hid_t obj, const char* name) {
  H5Handle a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) throw HDF5Error(std::string("Unable to open the attribute \"") + name + "\"");
  H5Handle ft(H5Aget_type(a), H5Tclose);
  if (H5Tis_variable_str(ft) > 0) {
    H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, H5T_VARIABLE);
    char* s = nullptr;
    if (H5Aread(a, mt, &s) < 0) throw HDF5Error("Unable to read a string attribute");
    std::string out = s ? s : "";
    H5free_memory(s);
    return out;
  }
  const size_t sz = H5Tget_size(ft);
  std::string out(sz, '\0');
  H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mt, sz);
  if (H5Aread(a, mt, &out[0]) < 0) throw HDF5Error("Unable to read a string attribute");
  out.erase(std::find(out.begin(), out.end(), '\0'), out.end());
  return out;
}

template <typename T>
std::vector<T> read_dataset(hid_t loc, const std::string& name, hid_t memtype) {
  H5Handle d(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (!d.ok()) throw HDF5Error("Unable to open the dataset \"" + name + "\"");
  H5Handle sp(H5Dget_space(d), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<T> v((size_t)std::max<hssize_t>(n, 0));
  if (!v.empty() && H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0) {
    throw HDF5Error("Unable to read the dataset \"" + name + "\"");
  }
  return v;
}

std::vector<std::string> read_str_dataset(hid_t loc, const std::string& name) {
  H5Handle d(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (!d.ok()) throw HDF5Error("Unable to open the dataset \"" + name + "\"");
  H5Handle sp(H5Dget_space(d), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<std::string> out;
  if (n <= 0) return out;
  H5Handle ft(H5Dget_type(d), H5Tclose);
  if (H5Tis_variable_str(ft) > 0) {
    H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, H5T_VARIABLE);
    std::vector<char*> buf((size_t)n, nullptr);
    if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) {
      throw HDF5Error("Unable to read the dataset \"" + name + "\"");
    }
    for (auto* s : buf) out.emplace_back(s ? s : "");
    H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, buf.data());
  } else {
    const size_t sz = H5Tget_size(ft);
    std::vector<char> buf((size_t)n * sz);
    H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, sz);
    if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) {
      throw HDF5Error("Unable to read the dataset \"" + name + "\"");
    }
    for (hssize_t i = 0; i < n; ++i) {
      std::string s(&buf[(size_t)i * sz], sz);
      s.erase(std::find(s.begin(), s.end(), '\0'), s.end());
      out.push_back(s);
    }
  }
  return out;
}

// Database::load_sketch (src/database/database.cpp:114-156)
Sketch load_sketch(hid_t file, const std::string& name) {
  H5Handle g = open_group(file, "/sketches/" + name);
  Sketch s;
  s.name = name;
  auto kmers = read_attr_vec<int32_t>(g, "kmers", H5T_NATIVE_INT32);
  s.sketchsize64 = read_attr<uint64_t>(g, "sketchsize64", H5T_NATIVE_UINT64);
  s.bbits = read_attr<uint64_t>(g, "bbits", H5T_NATIVE_UINT64);
  s.seq_length = DEFAULT_LENGTH;
  if (has_attr(g, "length")) s.seq_length = read_attr<uint64_t>(g, "length", H5T_NATIVE_UINT64);
  if (has_attr(g, "base_freq")) {
    auto b = read_attr_vec<double>(g, "base_freq", H5T_NATIVE_DOUBLE);
    for (size_t i = 0; i < 4 && i < b.size(); ++i) s.bases[i] = b[i];
  }
  if (has_attr(g, "missing_bases")) {
    s.missing_bases = read_attr<uint64_t>(g, "missing_bases", H5T_NATIVE_UINT64);
  }
  if (has_attr(g, "reads")) s.reads = read_attr<uint8_t>(g, "reads", H5T_NATIVE_UINT8) != 0;
  std::vector<std::pair<size_t, std::vector<uint64_t>>> by_k;
  for (int32_t k : kmers) {
    by_k.emplace_back((size_t)k, read_dataset<uint64_t>(g, std::to_string(k), H5T_NATIVE_UINT64));
  }
  std::sort(by_k.begin(), by_k.end(),
            [](const auto& a, const auto& b) { return a.first < b.first; });
  for (auto& kv : by_k) {
    s.kmers.push_back(kv.first);
    s.usigs.push_back(std::move(kv.second));
  }
  return s;
}

herr_t collect_name(hid_t, const char* name, const H5L_info_t*, void* op) {
  static_cast<std::vector<std::string>*>(op)->emplace_back(name);
  return 0;
}

H5Handle create_group(hid_t loc, const std::string& name) {
  hid_t g = H5Gcreate2(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) throw HDF5Error("Unable to create the group \"" + name + "\"");
  return H5Handle(g, H5Gclose);
}

void write_attr(hid_t obj, const char* name, hid_t filetype, hid_t memtype, const void* data,
                hsize_t n, bool scalar) {
  H5Handle sp(scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &n, nullptr), H5Sclose);
  H5Handle a(H5Acreate2(obj, name, filetype, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a, memtype, data) < 0) {
    throw HDF5Error(std::string("Unable to write the attribute \"") + name + "\"");
  }
}

void write_str_attr(hid_t obj, const char* name, const std::string& v) {
  H5Handle t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t, H5T_VARIABLE);
  H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
  H5Handle a(H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  const char* p = v.c_str();
  if (!a.ok() || H5Awrite(a, t, &p) < 0) throw HDF5Error("Unable to write a string attribute");
}

H5Handle write_dataset(hid_t loc, const std::string& name, hid_t filetype, hid_t memtype,
                       const void* data, hsize_t n) {
  H5Handle sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  hid_t d = H5Dcreate2(loc, name.c_str(), filetype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (d < 0) throw HDF5Error("Unable to create the dataset \"" + name + "\"");
  H5Handle dh(d, H5Dclose);
  if (n > 0 && H5Dwrite(dh, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
    throw HDF5Error("Unable to write the dataset \"" + name + "\"");
  }
  return dh;
}

}  // namespace

const std::vector<uint64_t>& Sketch::get_sketch(size_t k) const {
  for (size_t i = 0; i < kmers.size(); ++i) {
    if (kmers[i] == k) return usigs[i];
  }
  throw std::runtime_error("Kmer length " + std::to_string(k) + " not found in sketch " + name);
}

void Sketch::remove_kmer(size_t k) {
  for (size_t i = 0; i < kmers.size(); ++i) {
    if (kmers[i] == k) {
      kmers.erase(kmers.begin() + i);
      usigs.erase(usigs.begin() + i);
      return;
    }
  }
}

bool same_names(const std::vector<Sketch>& a, const std::vector<Sketch>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i) {
    if (a[i].name != b[i].name) return false;
  }
  return true;
}

std::vector<Sketch> load_sketches(const std::string& db_name,
                                  const std::vector<std::string>& names,
                                  std::vector<size_t> kmer_lengths, bool messages) {
  std::vector<Sketch> sketches;
  std::sort(kmer_lengths.begin(), kmer_lengths.end());
  try {
    H5Handle f = open_file(db_name + ".h5", false);
    if (messages) std::cerr << "Looking for existing sketches in " + db_name + ".h5" << std::endl;
    sketches.reserve(names.size());
    for (const auto& nm : names) {
      Sketch s = load_sketch(f, nm);
      // drop unrequested k, fail on requested-but-missing k (api.cpp:466-491)
      std::vector<size_t> loaded = s.kmers;
      auto want = kmer_lengths.begin();
      auto have = loaded.begin();
      while (want != kmer_lengths.end() && have != loaded.end()) {
        if (*want == *have) {
          ++want;
        } else {
          s.remove_kmer(*have);
        }
        ++have;
      }
      if (want != kmer_lengths.end()) {
        std::stringstream old_k, new_k;
        for (auto k : loaded) old_k << k << ",";
        for (auto k : kmer_lengths) new_k << k << ",";
        throw std::runtime_error("k-mer lengths in old database (" + old_k.str() +
                                 ") do not match those requested (" + new_k.str() + ")");
      }
      sketches.push_back(std::move(s));
    }
  } catch (const HDF5Error& e) {
    std::cerr << "Missing sketch: " << e.what() << std::endl;
    sketches.clear();
  } catch (const std::exception& e) {
    std::cerr << "Mismatched data: " << e.what() << std::endl;
    sketches.clear();
  }
  return sketches;
}

std::vector<std::string> list_sketch_names(const std::string& h5_path) {
  H5Handle f = open_file(h5_path, false);
  H5Handle g = open_group(f, "/sketches");
  std::vector<std::string> names;
  hsize_t idx = 0;
  H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_name, &names);
  return names;
}

DbAttrs read_db_attrs(const std::string& h5_path) {
  H5Handle f = open_file(h5_path, false);
  H5Handle g = open_group(f, "/sketches");
  DbAttrs a;
  if (has_attr(g, "sketch_version")) a.version = read_str_attr(g, "sketch_version");
  if (has_attr(g, "codon_phased")) {
    a.codon_phased = read_attr<uint8_t>(g, "codon_phased", H5T_NATIVE_UINT8) != 0;
  }
  return a;
}

// Database::load_random (src/database/database.cpp:195-226)
RandomTables read_random(const std::string& h5_path) {
  RandomTables r;
  H5Handle f = open_file(h5_path, false);
  if (H5Lexists(f, "random", H5P_DEFAULT) <= 0) return r;
  H5Handle g = open_group(f, "/random");
  r.present = true;
  r.table_keys = read_str_dataset(g, "table_keys");
  r.table_values = read_dataset<uint16_t>(g, "table_values", H5T_NATIVE_UINT16);
  r.matches_keys = read_dataset<uint64_t>(g, "matches_keys", H5T_NATIVE_UINT64);
  r.matches = read_dataset<float>(g, "matches_values", H5T_NATIVE_FLOAT);
  {
    H5Handle d(H5Dopen2(g, "matches_values", H5P_DEFAULT), H5Dclose);
    auto dims = read_attr_vec<uint64_t>(d, "dims", H5T_NATIVE_UINT64);
    if (dims.size() < 2 || dims[0] != dims[1]) throw HDF5Error("bad random match dims");
    r.n_clusters = dims[0];
  }
  r.centroids = read_dataset<float>(g, "centroids", H5T_NATIVE_FLOAT);
  r.k_min = read_attr<uint32_t>(g, "k_min", H5T_NATIVE_UINT32);
  r.k_max = read_attr<uint32_t>(g, "k_max", H5T_NATIVE_UINT32);
  r.use_rc = read_attr<uint8_t>(g, "use_rc", H5T_NATIVE_UINT8) != 0;
  if (r.table_keys.size() != r.table_values.size() ||
      r.matches.size() != r.matches_keys.size() * r.n_clusters * r.n_clusters ||
      r.centroids.size() != r.n_clusters * 4) {
    throw HDF5Error("inconsistent random match tables");
  }
  return r;
}

void write_new_db(const std::string& h5_path, bool use_rc, bool codon_phased,
                  const std::string& version) {
  quiet();
  H5Handle fapl = write_access();
  hid_t fid = H5Fcreate(h5_path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
  if (fid < 0) throw HDF5Error("Unable to create file " + h5_path);
  H5Handle f(fid, H5Fclose);
  H5Handle g = create_group(f, "sketches");
  write_str_attr(g, "sketch_version", version);
  const uint8_t cp = codon_phased ? 1 : 0, rc = use_rc ? 1 : 0;
  write_attr(g, "codon_phased", H5T_STD_U8LE, H5T_NATIVE_UINT8, &cp, 1, true);
  write_attr(g, "reverse_complement", H5T_STD_U8LE, H5T_NATIVE_UINT8, &rc, 1, true);
}

void write_sketches(const std::string& h5_path, const std::vector<Sketch>& sketches) {
  H5Handle f = open_file(h5_path, true);
  for (const auto& s : sketches) {
    H5Handle g = create_group(f, "/sketches/" + s.name);
    const uint64_t s64 = s.sketchsize64, bb = s.bbits, len = s.seq_length, miss = s.missing_bases;
    const uint8_t rd = s.reads ? 1 : 0;
    write_attr(g, "sketchsize64", H5T_STD_U64LE, H5T_NATIVE_UINT64, &s64, 1, true);
    write_attr(g, "bbits", H5T_STD_U64LE, H5T_NATIVE_UINT64, &bb, 1, true);
    write_attr(g, "length", H5T_STD_U64LE, H5T_NATIVE_UINT64, &len, 1, true);
    write_attr(g, "missing_bases", H5T_STD_U64LE, H5T_NATIVE_UINT64, &miss, 1, true);
    write_attr(g, "reads", H5T_STD_U8LE, H5T_NATIVE_UINT8, &rd, 1, true);
    write_attr(g, "base_freq", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, s.bases.data(), 4, false);
    std::vector<int32_t> ks(s.kmers.begin(), s.kmers.end());
    write_attr(g, "kmers", H5T_STD_I32LE, H5T_NATIVE_INT32, ks.data(), ks.size(), false);
    for (size_t i = 0; i < s.kmers.size(); ++i) {
      H5Handle d = write_dataset(g, std::to_string(s.kmers[i]), H5T_STD_U64LE, H5T_NATIVE_UINT64,
                                 s.usigs[i].data(), s.usigs[i].size());
      const int32_t k = (int32_t)s.kmers[i];
      write_attr(d, "kmer-size", H5T_STD_I32LE, H5T_NATIVE_INT32, &k, 1, true);
    }
  }
}

void write_random(const std::string& h5_path, const RandomTables& r) {
  H5Handle f = open_file(h5_path, true);
  if (H5Lexists(f, "random", H5P_DEFAULT) > 0) {
    throw std::runtime_error("Random matches already exist in " + h5_path);
  }
  H5Handle g = create_group(f, "random");
  {
    H5Handle t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t, H5T_VARIABLE);
    hsize_t n = r.table_keys.size();
    std::vector<const char*> ptrs;
    for (auto& s : r.table_keys) ptrs.push_back(s.c_str());
    write_dataset(g, "table_keys", t, t, ptrs.data(), n);
  }
  write_dataset(g, "table_values", H5T_STD_U16LE, H5T_NATIVE_UINT16, r.table_values.data(),
                r.table_values.size());
  write_dataset(g, "matches_keys", H5T_STD_U64LE, H5T_NATIVE_UINT64, r.matches_keys.data(),
                r.matches_keys.size());
  {
    H5Handle d = write_dataset(g, "matches_values", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT,
                               r.matches.data(), r.matches.size());
    const uint64_t dims[3] = {r.n_clusters, r.n_clusters, r.matches_keys.size()};
    write_attr(d, "dims", H5T_STD_U64LE, H5T_NATIVE_UINT64, dims, 3, false);
  }
  {
    H5Handle d = write_dataset(g, "centroids", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT,
                               r.centroids.data(), r.centroids.size());
    const uint64_t dims[2] = {r.n_clusters, 4};
    write_attr(d, "dims", H5T_STD_U64LE, H5T_NATIVE_UINT64, dims, 2, false);
  }
  write_attr(g, "k_min", H5T_STD_U32LE, H5T_NATIVE_UINT32, &r.k_min, 1, true);
  write_attr(g, "k_max", H5T_STD_U32LE, H5T_NATIVE_UINT32, &r.k_max, 1, true);
  const uint8_t rc = r.use_rc ? 1 : 0;
  write_attr(g, "use_rc", H5T_STD_U8LE, H5T_NATIVE_UINT8, &rc, 1, true);
}

}  // namespace ppsk_host
