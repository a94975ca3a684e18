"""ctypes front end of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Each wrapper names
the reference function (file:line, relative to the pp-sketchlib source tree)
that the underlying C++ restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PPSK_ORACLE_LIB: another build of the same source (the CPU baseline of
# bench.py compiles one for its host's CPU)
_LIB_PATH = os.environ.get("PPSK_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_dist.so")

_c = ctypes
_p = _c.c_void_p
_ll = _c.c_longlong
_sz = _c.c_size_t


def build(ref=False):
    """Compile the oracle (and the reference's calc_intersize into _ref/)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-C", _HERE] + targets, check=True,
                   stdout=subprocess.DEVNULL)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = _c.CDLL(_LIB_PATH)
    lib.orc_calc_intersize.restype = _sz
    lib.orc_calc_intersize.argtypes = [_p, _p, _sz, _sz]
    lib.orc_jaccard.restype = _c.c_double
    lib.orc_jaccard.argtypes = [_sz, _sz, _c.c_double]
    lib.orc_csrs.restype = _c.c_double
    lib.orc_csrs.argtypes = [_sz, _c.c_int, _sz, _sz]
    lib.orc_csrs_j.restype = _c.c_double
    lib.orc_csrs_j.argtypes = [_sz, _c.c_int, _sz]
    lib.orc_csrs_combine.restype = _c.c_double
    lib.orc_csrs_combine.argtypes = [_c.c_double, _c.c_double]
    lib.orc_fit_slope.restype = _c.c_int
    lib.orc_fit_slope.argtypes = [_p, _p, _c.c_int, _c.c_double, _p, _p]
    lib.orc_closest_cluster.restype = _c.c_int
    lib.orc_closest_cluster.argtypes = [_p, _c.c_int, _p, _c.c_int]
    lib.orc_query_db.restype = _ll
    lib.orc_query_db.argtypes = [_p, _ll, _p, _ll, _c.c_int, _p, _sz, _sz, _p, _p,
                                 _p, _p, _c.c_int, _c.c_int, _c.c_int, _p, _p,
                                 _c.c_double, _c.c_int, _c.c_int, _p, _p, _p]
    lib.orc_query_db_sparse.restype = _ll
    lib.orc_query_db_sparse.argtypes = [_p, _ll, _c.c_int, _p, _sz, _sz, _p, _p,
                                        _c.c_int, _c.c_int, _c.c_int, _p, _p,
                                        _c.c_double, _c.c_int, _c.c_int, _c.c_int,
                                        _c.c_int, _p, _p, _p]
    lib.orc_knn_from_square.restype = None
    lib.orc_knn_from_square.argtypes = [_p, _ll, _c.c_int, _p, _p, _p, _c.c_int]
    lib.orc_sketch.restype = _c.c_int
    lib.orc_sketch.argtypes = [_c.c_char_p, _p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _p,
                               _c.POINTER(_c.c_double)]
    lib.orc_sketch_filtered.restype = _c.c_int
    lib.orc_sketch_filtered.argtypes = [_c.c_char_p, _p, _c.c_int, _c.c_int, _c.c_int,
                                        _c.c_int, _c.c_int, _c.c_int, _c.c_int, _p,
                                        _c.POINTER(_c.c_double)]
    lib.orc_knn_rows.restype = None
    lib.orc_knn_rows.argtypes = [_p, _p, _ll, _ll, _c.c_int, _p, _p, _p, _c.c_int]
    lib.orc_square_to_condensed.restype = _ll
    lib.orc_square_to_condensed.argtypes = [_c.c_long, _c.c_long, _c.c_long]
    lib.orc_calc_row_idx.restype = _c.c_long
    lib.orc_calc_row_idx.argtypes = [_ll, _c.c_long]
    lib.orc_calc_col_idx.restype = _c.c_long
    lib.orc_calc_col_idx.argtypes = [_ll, _c.c_long, _c.c_long]
    lib.orc_long_to_square_size.restype = _ll
    lib.orc_long_to_square_size.argtypes = [_ll, _ll, _ll]
    lib.orc_long_to_square.restype = _ll
    lib.orc_long_to_square.argtypes = [_p, _ll, _p, _ll, _p, _ll, _p]
    lib.orc_square_to_long.restype = None
    lib.orc_square_to_long.argtypes = [_p, _ll, _p]
    lib.orc_sparsify_threshold.restype = _ll
    lib.orc_sparsify_threshold.argtypes = [_p, _ll, _ll, _c.c_float, _p, _p, _p]
    lib.orc_max_threads.restype = _c.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# --------------------------------------------------------------------------
# Scalar pieces
# --------------------------------------------------------------------------
def calc_intersize(s1, s2, sketchsize64, bbits):
    """dist/dist.cpp:50-78"""
    s1 = _u64(s1)
    s2 = _u64(s2)
    return lib().orc_calc_intersize(_ptr(s1), _ptr(s2), sketchsize64, bbits)


def ref_calc_intersize_available():
    return os.path.exists(_REF_PATH)


def ref_calc_intersize(s1, s2, sketchsize64, bbits):
    """The reference's own compiled calc_intersize (oracle/_ref)."""
    r = _c.CDLL(_REF_PATH)
    r.ref_calc_intersize.restype = _sz
    r.ref_calc_intersize.argtypes = [_p, _p, _sz, _sz]
    s1 = _u64(s1)
    s2 = _u64(s2)
    return r.ref_calc_intersize(_ptr(s1), _ptr(s2), sketchsize64, bbits)


def jaccard(intersize, sketchsize64, random_jaccard=0.0):
    """reference.cpp:222-232 + dist/dist.hpp:13-24 (double)."""
    return lib().orc_jaccard(int(intersize), int(sketchsize64), float(random_jaccard))


def csrs(k, use_rc, l1, l2):
    """random/random_match.cpp:318-336"""
    return lib().orc_csrs(int(k), int(bool(use_rc)), int(l1), int(l2))


def csrs_j(k, use_rc, length):
    return lib().orc_csrs_j(int(k), int(bool(use_rc)), int(length))


def csrs_combine(j1, j2):
    return lib().orc_csrs_combine(float(j1), float(j2))


def fit_slope(kmers, dists, tol):
    """dist/linear_regression.cpp:16-60.  Returns (core, acc, ok)."""
    k = np.ascontiguousarray(kmers, dtype=np.float32)
    d = np.ascontiguousarray(dists, dtype=np.float32)
    core = np.zeros(1, np.float32)
    acc = np.zeros(1, np.float32)
    rc = lib().orc_fit_slope(_ptr(k), _ptr(d), len(k), float(tol), _ptr(core), _ptr(acc))
    return float(core[0]), float(acc[0]), rc == 0


def closest_cluster(bases, use_rc, centroids):
    """random/random_match.cpp:338-359"""
    b = np.ascontiguousarray(bases, dtype=np.float64)
    c = np.ascontiguousarray(centroids, dtype=np.float32)
    return lib().orc_closest_cluster(_ptr(b), int(bool(use_rc)), _ptr(c), c.shape[0])


# --------------------------------------------------------------------------
# Random-match description shared by the loops
# --------------------------------------------------------------------------
class Random:
    """mode 0 NoAdjust, 1 EqualBase, 2 MonteCarlo (random_match.cpp:185-206).

    table: [n_k][C][C] float32 aligned with klist; k_lt_max: [n_k] bool
    (the k < max_k rule of random_match.cpp:191)."""

    def __init__(self, mode=0, use_rc=True, table=None, k_lt_max=None):
        self.mode = mode
        self.use_rc = bool(use_rc)
        self.table = None if table is None else np.ascontiguousarray(table, np.float32)
        self.k_lt_max = None if k_lt_max is None else np.ascontiguousarray(k_lt_max, np.int32)
        self.n_clusters = 0 if table is None else self.table.shape[-1]


def query_db(ref, klist, sketchsize64, bbits, ref_len, query=None, query_len=None,
             ref_cluster=None, query_cluster=None, random=None, tol=None,
             jaccard=False, num_threads=None, want_counts=False):
    """api.cpp:140-291.  ref/query: uint64 [n][n_k][s64*bbits].

    Returns (dists, counts_or_None, n_fail, first_fail_row)."""
    L = lib()
    random = random or Random(0)
    ref = _u64(ref)
    n_ref = ref.shape[0]
    n_k = len(klist)
    kl = np.ascontiguousarray(klist, np.int32)
    rl = _u64(ref_len)
    rc = None if ref_cluster is None else np.ascontiguousarray(ref_cluster, np.uint16)
    if query is not None:
        query = _u64(query)
        n_q = query.shape[0]
        ql = _u64(query_len)
        qc = None if query_cluster is None else np.ascontiguousarray(query_cluster, np.uint16)
        rows = n_ref * n_q
    else:
        n_q = 0
        ql = None
        qc = None
        rows = n_ref * (n_ref - 1) // 2
    if tol is None:
        tol = 5.0 / (sketchsize64 * 64)
    cols = n_k if jaccard else 2
    out = np.zeros((rows, cols), np.float32)
    counts = np.zeros((rows, n_k), np.uint16) if want_counts else None
    fail = np.zeros(1, np.int64)
    nthr = num_threads or L.orc_max_threads()
    n_fail = L.orc_query_db(_ptr(ref), n_ref, _ptr(query), n_q, n_k, _ptr(kl),
                            sketchsize64, bbits, _ptr(rl), _ptr(ql), _ptr(rc), _ptr(qc),
                            random.mode, int(random.use_rc), random.n_clusters,
                            _ptr(random.table), _ptr(random.k_lt_max), float(tol),
                            int(bool(jaccard)), nthr, _ptr(out), _ptr(counts),
                            _ptr(fail))
    return out, counts, int(n_fail), int(fail[0])


def query_db_sparse(ref, klist, sketchsize64, bbits, ref_len, kNN, dist_col=0,
                    ref_cluster=None, random=None, tol=None, jaccard=False,
                    num_threads=None):
    """api.cpp:319-432 (kNN).  Returns (i_vec, j_vec, dists, n_fail)."""
    L = lib()
    random = random or Random(0)
    ref = _u64(ref)
    n = ref.shape[0]
    kl = np.ascontiguousarray(klist, np.int32)
    rl = _u64(ref_len)
    rc = None if ref_cluster is None else np.ascontiguousarray(ref_cluster, np.uint16)
    if tol is None:
        tol = 5.0 / (sketchsize64 * 64)
    iv = np.zeros(n * kNN, np.int64)
    jv = np.zeros(n * kNN, np.int64)
    dv = np.zeros(n * kNN, np.float32)
    nthr = num_threads or L.orc_max_threads()
    n_fail = L.orc_query_db_sparse(_ptr(ref), n, len(klist), _ptr(kl), sketchsize64,
                                   bbits, _ptr(rl), _ptr(rc), random.mode,
                                   int(random.use_rc), random.n_clusters,
                                   _ptr(random.table), _ptr(random.k_lt_max),
                                   float(tol), int(bool(jaccard)), int(kNN),
                                   int(dist_col), nthr, _ptr(iv), _ptr(jv), _ptr(dv))
    return iv, jv, dv, int(n_fail)


def knn_from_square(square, kNN, num_threads=None):
    """The api.cpp:360-414 heap, fed from a dense square matrix."""
    L = lib()
    sq = np.ascontiguousarray(square, np.float32)
    n = sq.shape[0]
    iv = np.zeros(n * kNN, np.int64)
    jv = np.zeros(n * kNN, np.int64)
    dv = np.zeros(n * kNN, np.float32)
    L.orc_knn_from_square(_ptr(sq), n, int(kNN), _ptr(iv), _ptr(jv), _ptr(dv),
                          num_threads or L.orc_max_threads())
    return iv, jv, dv


def sketch_filtered(contigs, k, sketchsize64, use_rc=True, min_count=0, exact=False,
                    codon=False):
    """sketch() with the reads count filter (sketch/countmin.cpp): a k-mer
    occurrence is binned only if its count so far reaches min_count; codon:
    the codon-phased seeds (sketch/sketch.cpp:43-54)."""
    L = lib()
    bs = [c.encode() if isinstance(c, str) else bytes(c) for c in contigs]
    seq = b"".join(bs)
    off = np.zeros(len(bs) + 1, np.int64)
    off[1:] = np.cumsum([len(b) for b in bs])
    us = np.zeros(sketchsize64 * 14, np.uint64)
    mh = _c.c_double()
    dens = L.orc_sketch_filtered(seq, _ptr(off), len(bs), int(k), int(sketchsize64),
                                 int(bool(use_rc)), int(bool(codon)), int(min_count),
                                 int(bool(exact)), _ptr(us), _c.byref(mh))
    return us, mh.value, dens


def sketch(contigs, k, sketchsize64, use_rc=True):
    """sketch/sketch.cpp:118-176 for one sample and one k.  contigs: list of
    str/bytes.  Returns (usigs [s64*14] uint64, inv_minhash, densify code)."""
    L = lib()
    bs = [c.encode() if isinstance(c, str) else bytes(c) for c in contigs]
    seq = b"".join(bs)
    off = np.zeros(len(bs) + 1, np.int64)
    off[1:] = np.cumsum([len(b) for b in bs])
    us = np.zeros(sketchsize64 * 14, np.uint64)
    mh = _c.c_double()
    dens = L.orc_sketch(seq, _ptr(off), len(bs), int(k), int(sketchsize64), int(bool(use_rc)),
                        _ptr(us), _c.byref(mh))
    return us, mh.value, dens


_REF_SKETCH = None


def ref_sketch_lib():
    """The reference's own sketch core (oracle/_ref), or None if not built."""
    global _REF_SKETCH
    if _REF_SKETCH is None:
        path = os.path.join(_HERE, "_ref", "libref_sketch.so")
        if not os.path.exists(path):
            return None
        L = _c.CDLL(path)
        L.ref_sketch_seqs.argtypes = [_c.POINTER(_c.c_char_p), _c.c_int, _c.c_int, _c.c_int,
                                      _c.c_int, _c.c_int, _p, _c.POINTER(_c.c_double),
                                      _c.POINTER(_c.c_int)]
        L.ref_read_stats.argtypes = [_c.c_char_p, _c.c_int, _p, _c.POINTER(_c.c_uint64),
                                     _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int),
                                     _c.POINTER(_c.c_uint64)]
        L.ref_read_seqs.restype = _c.c_int64
        L.ref_read_seqs.argtypes = [_c.c_char_p, _c.c_int, _c.c_char_p, _c.c_uint64, _p,
                                    _c.c_uint64]
        L.ref_sketch_file.argtypes = [_c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
                                      _c.c_int, _c.c_int, _c.c_int, _p, _c.POINTER(_c.c_double),
                                      _c.POINTER(_c.c_int)]
        _REF_SKETCH = L
    return _REF_SKETCH


def ref_sketch(contigs, k, sketchsize64, use_rc=True, codon=False):
    """The reference's compiled sketch() on the same input."""
    L = ref_sketch_lib()
    bs = [c.encode() if isinstance(c, str) else bytes(c) for c in contigs]
    arr = (_c.c_char_p * len(bs))(*bs)
    us = np.zeros(sketchsize64 * 14, np.uint64)
    mh = _c.c_double()
    dn = _c.c_int()
    if L.ref_sketch_seqs(arr, len(bs), int(k), int(sketchsize64), int(bool(use_rc)),
                         int(bool(codon)), _ptr(us), _c.byref(mh), _c.byref(dn)):
        raise RuntimeError("reference sketch failed")
    return us, mh.value, dn.value


def ref_read_stats(path, kmer_len):
    L = ref_sketch_lib()
    b = np.zeros(4, np.float64)
    tot, nc, ns = _c.c_uint64(), _c.c_uint64(), _c.c_uint64()
    rd = _c.c_int()
    if L.ref_read_stats(path.encode(), int(kmer_len), _ptr(b), _c.byref(tot), _c.byref(nc),
                        _c.byref(rd), _c.byref(ns)):
        raise RuntimeError("reference reader failed")
    return b, tot.value, nc.value, bool(rd.value), ns.value


def ref_read_seqs(path, kmer_len, cap=1 << 26, max_seqs=1 << 20):
    """The sequences the reference's reader keeps from one file, as str."""
    L = ref_sketch_lib()
    buf = _c.create_string_buffer(cap)
    lens = np.zeros(max_seqs, np.uint64)
    n = L.ref_read_seqs(path.encode(), int(kmer_len), buf, cap, _ptr(lens), max_seqs)
    if n < 0:
        raise RuntimeError("reference reader failed")
    out, pos, raw = [], 0, buf.raw
    for ln in lens[:n].tolist():
        out.append(raw[pos:pos + ln].decode("latin-1"))
        pos += ln
    return out


def ref_sketch_file(path, keep_len, k, sketchsize64, use_rc=True, min_count=0, exact=False,
                    codon=False):
    """The reference's sketch of one file at one k (sequences kept if at
    least keep_len long; reads filtered by count when min_count > 0); None
    when it keeps no sequence."""
    L = ref_sketch_lib()
    us = np.zeros(sketchsize64 * 14, np.uint64)
    mh = _c.c_double()
    dn = _c.c_int()
    rc = L.ref_sketch_file(path.encode(), int(keep_len), int(k), int(sketchsize64),
                           int(bool(use_rc)), int(bool(codon)), int(min_count), int(bool(exact)),
                           _ptr(us), _c.byref(mh), _c.byref(dn))
    if rc < 0:
        raise RuntimeError("reference sketch failed")
    if rc == 1:
        return None
    return us, mh.value, dn.value


# --------------------------------------------------------------------------
# Monte Carlo random-match tables (random/random_match.cpp:62-163)
# --------------------------------------------------------------------------
_REF_RANDOM = None


def ref_random_lib():
    """The reference's random generators (oracle/_ref), or None if not built."""
    global _REF_RANDOM
    if _REF_RANDOM is None:
        path = os.path.join(_HERE, "_ref", "libref_random.so")
        if not os.path.exists(path):
            return None
        L = _c.CDLL(path)
        L.ref_xoshiro.argtypes = [_c.c_uint64, _c.c_int, _c.c_uint64, _p]
        L.ref_mt_double.argtypes = [_c.c_uint32, _c.c_uint64, _p]
        _REF_RANDOM = L
    return _REF_RANDOM


def ref_xoshiro(seed, n, jumps=0):
    out = np.zeros(n, np.uint64)
    ref_random_lib().ref_xoshiro(seed, jumps, n, _ptr(out))
    return out


def ref_mt_double(seed, n):
    out = np.zeros(n, np.float64)
    ref_random_lib().ref_mt_double(seed, n, _ptr(out))
    return out


def csrs_py(k, use_rc, l1, l2):
    """random_match.cpp:318-336 (k < 33)."""
    m = float(1 << ((k - 1) * 2 + (1 if use_rc else 0)))
    j1 = 1 - (1 - 1 / m) ** l1
    j2 = 1 - (1 - 1 / m) ** l2
    return (j1 * j2) / (j1 + j2 - j1 * j2) if (j1 > 0 and j2 > 0) else 0.0


def mc_kmer_lengths(length, sketchsize64, use_rc):
    """random_match.cpp:82-95: k = 3, 4, ... while the equal-base chance for
    two genomes of sketch 0's length exceeds 1 / (64 s64)."""
    kl = [3]
    for k in range(4, 32):
        if csrs_py(k, use_rc, length, length) > 1.0 / (sketchsize64 * 64):
            kl.append(k)
        else:
            break
    return kl


def mc_matches(genomes, klist, sketchsize64, n_clusters, n_mc, use_rc=True):
    """random_match.cpp:120-156: per k a C x C float table, the mean
    uncorrected Jaccard over genome pairs of two clusters (genomes in
    cluster-major order), a cluster's own entry without its n_mc self pairs.
    Sums run copy-major in double, self pairs counting 1 and removed after."""
    sk = [[sketch([g], k, sketchsize64, use_rc)[0] for k in klist] for g in genomes]
    out = np.zeros((len(klist), n_clusters, n_clusters), np.float32)
    for kk in range(len(klist)):
        for ci in range(n_clusters):
            for cj in range(ci, n_clusters):
                total = 0.0
                for a in range(n_mc):
                    for b in range(n_mc):
                        x, y = ci * n_mc + a, cj * n_mc + b
                        c = calc_intersize(sk[x][kk], sk[y][kk], sketchsize64, 14)
                        total += c / float(64 * sketchsize64)
                count = n_mc * n_mc
                if ci == cj:
                    total -= n_mc
                    count -= n_mc
                out[kk, ci, cj] = out[kk, cj, ci] = np.float32(total / count)
    return out


def knn_rows(rows, row_ids, kNN, num_threads=None):
    """The same heap over dense rows [m][n]; column row_ids[r] counts as +inf
    (row_ids[r] < 0: none excluded).  Returns (i_vec, j_vec, dists)."""
    L = lib()
    d = np.ascontiguousarray(rows, np.float32)
    m, n = d.shape
    ids = np.ascontiguousarray(row_ids, np.int64)
    iv = np.zeros(m * kNN, np.int64)
    jv = np.zeros(m * kNN, np.int64)
    dv = np.zeros(m * kNN, np.float32)
    L.orc_knn_rows(_ptr(d), _ptr(ids), m, n, int(kNN), _ptr(iv), _ptr(jv), _ptr(dv),
                   num_threads or L.orc_max_threads())
    return iv, jv, dv


# --------------------------------------------------------------------------
# Matrix utilities (dist/matrix_ops.cpp, dist/matrix_idx.hpp)
# --------------------------------------------------------------------------
def square_to_condensed(i, j, n):
    return lib().orc_square_to_condensed(i, j, n)


def calc_row_idx(k, n):
    return lib().orc_calc_row_idx(k, n)


def calc_col_idx(k, i, n):
    return lib().orc_calc_col_idx(k, i, n)


def long_to_square(rr, qr=None, qq=None):
    L = lib()
    rr = np.ascontiguousarray(rr, np.float32).ravel()
    qr = np.zeros(0, np.float32) if qr is None else np.ascontiguousarray(qr, np.float32).ravel()
    qq = np.zeros(0, np.float32) if qq is None else np.ascontiguousarray(qq, np.float32).ravel()
    side = L.orc_long_to_square_size(len(rr), len(qr), len(qq))
    if side < 0:
        raise RuntimeError("Shape of reference, query and ref vs query matrices mismatches")
    out = np.zeros((side, side), np.float32)
    L.orc_long_to_square(_ptr(rr), len(rr), _ptr(qr), len(qr), _ptr(qq), len(qq), _ptr(out))
    return out


def square_to_long(square):
    sq = np.ascontiguousarray(square, np.float32)
    n = sq.shape[0]
    out = np.zeros(n * (n - 1) // 2, np.float32)
    lib().orc_square_to_long(_ptr(sq), n, _ptr(out))
    return out


def sparsify_threshold(dense, cutoff):
    L = lib()
    d = np.ascontiguousarray(dense, np.float32)
    rows, cols = d.shape
    n = L.orc_sparsify_threshold(_ptr(d), rows, cols, float(cutoff), None, None, None)
    if n < 0:
        raise RuntimeError("Distance threshold must be at least 0")
    iv = np.zeros(n, np.int64)
    jv = np.zeros(n, np.int64)
    dv = np.zeros(n, np.float32)
    L.orc_sparsify_threshold(_ptr(d), rows, cols, float(cutoff), _ptr(iv), _ptr(jv), _ptr(dv))
    return iv, jv, dv


# --------------------------------------------------------------------------
# Layout helpers (sketch/sketch.cpp:82-92): bit-sliced words <-> bin values
# --------------------------------------------------------------------------
def unpack_bins(words, sketchsize64, bbits):
    """uint64 [..., s64*bbits] -> uint32 bin values [..., s64*64]."""
    w = np.asarray(words, dtype=np.uint64).reshape(words.shape[:-1] + (sketchsize64, bbits))
    bitpos = np.arange(64, dtype=np.uint64)
    bits = (w[..., None] >> bitpos) & np.uint64(1)  # [..., s64, bbits, 64]
    vals = np.zeros(words.shape[:-1] + (sketchsize64, 64), np.uint64)
    for l in range(bbits):
        vals |= bits[..., l, :] << np.uint64(l)
    return vals.reshape(words.shape[:-1] + (sketchsize64 * 64,)).astype(np.uint32)


def pack_bins(vals, sketchsize64, bbits):
    """uint bin values [..., s64*64] -> bit-sliced uint64 [..., s64*bbits]."""
    v = np.asarray(vals, dtype=np.uint64).reshape(vals.shape[:-1] + (sketchsize64, 64))
    bitpos = np.arange(64, dtype=np.uint64)
    out = np.zeros(vals.shape[:-1] + (sketchsize64, bbits), np.uint64)
    for l in range(bbits):
        out[..., l] = np.bitwise_or.reduce(((v >> np.uint64(l)) & np.uint64(1)) << bitpos, axis=-1)
    return out.reshape(vals.shape[:-1] + (sketchsize64 * bbits,))
