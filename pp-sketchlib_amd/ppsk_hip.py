"""ctypes binding of libppsk.so, the C ABI declared in include/ppsk.h.

This is the thin Python view of the HIP library used by the tests and the
benchmark (the drop-in `pp_sketchlib` module is the C++ host layer).  There is
no CPU fallback: if libppsk.so is missing or no HIP device is present, calls
raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppsk.so")

OK, ERR_ARG, ERR_HIP, ERR_MEMORY, ERR_FIT, ERR_STATE, ERR_CANCELLED = range(7)
RANDOM_NONE, RANDOM_EQUAL_BASE, RANDOM_MC = 0, 1, 2
OUT_KINDS = {"core_acc": 0, "jaccard": 1, "counts": 2}

_c = ctypes
_p = _c.c_void_p
_i32 = _c.c_int32
_i64 = _c.c_int64
_u64 = _c.c_uint64

# every symbol include/ppsk.h declares
EXPORTS = [
    "ppsk_last_error", "ppsk_version", "ppsk_device_init", "ppsk_set_create",
    "ppsk_set_upload", "ppsk_set_meta", "ppsk_set_free", "ppsk_set_size", "ppsk_dense",
    "ppsk_knn", "ppsk_threshold", "ppsk_self_pass", "ppsk_free_host", "ppsk_host_alloc",
    "ppsk_host_free", "ppsk_device_alloc", "ppsk_device_free", "ppsk_copy_to_host",
    "ppsk_last_timing", "ppsk_synth_sketches", "ppsk_set_prefix", "ppsk_batch_create",
    "ppsk_batch_run", "ppsk_batch_tuning", "ppsk_batch_free", "ppsk_server_create",
    "ppsk_server_submit", "ppsk_server_wait", "ppsk_result_free", "ppsk_server_stats",
    "ppsk_server_free", "ppsk_sketch", "ppsk_sketch_ex", "ppsk_dense_noalloc",
    "ppsk_progress", "ppsk_cancel", "ppsk_cancel_reset",
]


class Job(_c.Structure):
    """ppsk_job (include/ppsk.h)."""
    _fields_ = [("ref", _p), ("query", _p), ("random_mode", _i32), ("n_clusters", _i32),
                ("mc_table", _p), ("tolerance", _c.c_double), ("out_kind", _i32),
                ("out", _p)]


class Result(_c.Structure):
    """ppsk_result (include/ppsk.h)."""
    _fields_ = [("n_queries", _i64), ("knn_n", _i64), ("knn_i", _c.POINTER(_i64)),
                ("knn_j", _c.POINTER(_i64)), ("knn_d", _c.POINTER(_c.c_float)),
                ("thr_n", _i64), ("thr_i", _c.POINTER(_i64)), ("thr_j", _c.POINTER(_i64)),
                ("thr_d", _c.POINTER(_c.c_float)), ("n_fail", _u64), ("t_submit", _c.c_double),
                ("t_start", _c.c_double), ("t_done", _c.c_double), ("step_id", _i64),
                ("step_rows", _i64), ("n_steps", _i64), ("owner", _p)]


class PpskError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PpskError(ERR_STATE, "libppsk.so is not built (run pp-sketchlib_amd/build.py)")
    L = _c.CDLL(LIB_PATH)
    L.ppsk_last_error.restype = _c.c_char_p
    L.ppsk_version.restype = _c.c_char_p
    L.ppsk_device_init.argtypes = [_c.c_int, _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t),
                                   _c.POINTER(_c.c_size_t)]
    L.ppsk_set_create.argtypes = [_c.c_int, _i64, _i32, _p, _i32, _i32, _c.POINTER(_p)]
    L.ppsk_set_upload.argtypes = [_p, _i64, _i64, _p]
    L.ppsk_set_meta.argtypes = [_p, _p, _i32, _p]
    L.ppsk_set_free.argtypes = [_p]
    L.ppsk_set_size.argtypes = [_p]
    L.ppsk_set_size.restype = _i64
    L.ppsk_dense.argtypes = [_p, _p, _i32, _i32, _p, _c.c_double, _i32, _p, _i32,
                             _c.POINTER(_u64), _c.POINTER(_i64)]
    L.ppsk_progress.argtypes = [_c.c_int, _c.POINTER(_u64), _c.POINTER(_u64),
                                _c.POINTER(_i32), _c.POINTER(_i32)]
    L.ppsk_cancel.argtypes = [_c.c_int]
    L.ppsk_cancel_reset.argtypes = [_c.c_int]
    L.ppsk_dense_noalloc.argtypes = [_p, _p, _i32, _i32, _p, _c.c_double, _i32, _p,
                                     _c.POINTER(_u64), _c.POINTER(_i64)]
    L.ppsk_knn.argtypes = [_p, _i32, _i32, _p, _c.c_double, _i32, _i32, _i32, _p, _p, _p,
                           _c.POINTER(_u64)]
    L.ppsk_threshold.argtypes = [_p, _i32, _i32, _p, _c.c_double, _i32, _c.c_float,
                                 _c.POINTER(_p), _c.POINTER(_p), _c.POINTER(_p),
                                 _c.POINTER(_i64), _c.POINTER(_u64)]
    L.ppsk_self_pass.argtypes = [_p, _i32, _i32, _p, _c.c_double, _p, _i32, _i32, _p, _p, _p,
                                 _c.c_float, _i32, _c.POINTER(_p), _c.POINTER(_p),
                                 _c.POINTER(_p), _c.POINTER(_i64), _c.POINTER(_u64)]
    L.ppsk_free_host.argtypes = [_p]
    L.ppsk_free_host.restype = None
    L.ppsk_host_alloc.argtypes = [_c.c_size_t, _c.POINTER(_p)]
    L.ppsk_host_free.argtypes = [_p]
    L.ppsk_device_alloc.argtypes = [_c.c_int, _c.c_size_t, _c.POINTER(_p)]
    L.ppsk_device_free.argtypes = [_p]
    L.ppsk_copy_to_host.argtypes = [_p, _p, _c.c_size_t]
    L.ppsk_last_timing.argtypes = [_c.POINTER(_c.c_double), _c.POINTER(_i64), _c.POINTER(_i64)]
    L.ppsk_synth_sketches.argtypes = [_i64, _i32, _p, _i32, _u64, _c.c_double, _c.c_double,
                                      _c.c_double, _p, _p, _p, _i32]
    L.ppsk_set_prefix.argtypes = [_p, _i64, _c.POINTER(_p)]
    L.ppsk_batch_create.argtypes = [_c.c_int, _c.POINTER(Job), _i32, _c.POINTER(_p)]
    L.ppsk_batch_run.argtypes = [_p, _i32, _p, _c.POINTER(_i32), _c.POINTER(_c.c_double)]
    L.ppsk_batch_tuning.argtypes = [_p, _c.POINTER(_i32), _p, _p]
    L.ppsk_batch_free.argtypes = [_p]
    L.ppsk_server_create.argtypes = [_p, _i32, _i32, _p, _c.c_double, _i32, _i32, _c.c_float,
                                     _i32, _i32, _i32, _c.c_size_t, _c.POINTER(_p)]
    L.ppsk_server_submit.argtypes = [_p, _i64, _p, _p, _p, _c.POINTER(_i64)]
    L.ppsk_server_wait.argtypes = [_p, _i64, _c.POINTER(_c.POINTER(Result))]
    L.ppsk_result_free.argtypes = [_c.POINTER(Result)]
    L.ppsk_result_free.restype = None
    L.ppsk_server_stats.argtypes = [_p, _c.POINTER(_i64), _c.POINTER(_i64), _c.POINTER(_i64),
                                    _c.POINTER(_i32), _c.POINTER(_c.c_size_t),
                                    _c.POINTER(_c.c_double)]
    L.ppsk_server_free.argtypes = [_p]
    L.ppsk_sketch.argtypes = [_c.c_int, _i64, _c.c_char_p, _i64, _p, _i64, _p, _i32, _p, _i32,
                              _i32, _p, _p, _p]
    L.ppsk_sketch_ex.argtypes = [_c.c_int, _i64, _c.c_char_p, _i64, _p, _i64, _p, _i32, _p,
                                 _i32, _i32, _i32, _i32, _i32, _p, _p, _p]
    _lib = L
    return L


def _check(rc, allow=()):
    if rc != OK and rc not in allow:
        raise PpskError(rc, lib().ppsk_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def version():
    return lib().ppsk_version().decode()


def device_init(device=0):
    fr, tot, sh = _c.c_size_t(), _c.c_size_t(), _c.c_size_t()
    _check(lib().ppsk_device_init(device, _c.byref(fr), _c.byref(tot), _c.byref(sh)))
    return fr.value, tot.value, sh.value


def progress(device=0):
    """Progress of the distance call running on a device (or the last one):
    dict(done, total, running, pending) in (tile, k-mer length) work items;
    pending is True while a cancellation request stands."""
    d, t = _u64(), _u64()
    r, p = _i32(), _i32()
    _check(lib().ppsk_progress(device, _c.byref(d), _c.byref(t), _c.byref(r), _c.byref(p)))
    return dict(done=d.value, total=t.value, running=bool(r.value), pending=bool(p.value))


def cancel(device=0):
    """Ask the call running on `device` (and calls starting there until
    cancel_reset) to stop; they raise PpskError with code ERR_CANCELLED."""
    _check(lib().ppsk_cancel(device))


def cancel_reset(device=0):
    _check(lib().ppsk_cancel_reset(device))


def synth(n, klist, sketchsize64=156, seed=1, dcore_max=0.05, dacc_max=0.30,
          length_mean=2_000_000.0, threads=0):
    """Seeded synthetic sketches in the reference word layout.

    Returns (raw [n][n_k][s64*14] uint64, lengths [n] uint64, base_freq [n][4])."""
    kl = np.ascontiguousarray(klist, np.int32)
    raw = np.zeros((n, len(kl), sketchsize64 * 14), np.uint64)
    lengths = np.zeros(n, np.uint64)
    bf = np.zeros((n, 4), np.float64)
    _check(lib().ppsk_synth_sketches(n, len(kl), _ptr(kl), sketchsize64, seed, dcore_max,
                                     dacc_max, length_mean, _ptr(raw), _ptr(lengths), _ptr(bf),
                                     threads))
    return raw, lengths, bf


class SketchSet:
    """A resident sketch set in HBM (ppsk_set)."""

    def __init__(self, raw, klist, sketchsize64, device=0, lengths=None, use_rc=True,
                 clusters=None):
        raw = np.ascontiguousarray(raw, np.uint64)
        self.n = raw.shape[0]
        self.klist = [int(k) for k in klist]
        self.s64 = int(sketchsize64)
        self.device = device
        kl = np.ascontiguousarray(self.klist, np.int32)
        h = _p()
        _check(lib().ppsk_set_create(device, self.n, len(kl), _ptr(kl), self.s64, 14,
                                     _c.byref(h)))
        self._h = h
        _check(lib().ppsk_set_upload(self._h, 0, self.n, _ptr(raw)))
        if lengths is not None or clusters is not None:
            self.meta(lengths, use_rc, clusters)

    def meta(self, lengths=None, use_rc=True, clusters=None):
        le = None if lengths is None else np.ascontiguousarray(lengths, np.uint64)
        cl = None if clusters is None else np.ascontiguousarray(clusters, np.uint16)
        _check(lib().ppsk_set_meta(self._h, _ptr(le), int(bool(use_rc)), _ptr(cl)))

    @property
    def handle(self):
        if self._h is None:
            raise PpskError(ERR_STATE, "sketch set freed")
        return self._h

    def prefix(self, count):
        """Read-only view of the first `count` samples (no copy)."""
        return SetView(self, count)

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ppsk_set_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SetView(SketchSet):
    """ppsk_set_prefix view; keeps its parent alive."""

    def __init__(self, parent, count):
        h = _p()
        _check(lib().ppsk_set_prefix(parent.handle, int(count), _c.byref(h)))
        self._h = h
        self.parent = parent
        self.n = int(count)
        self.klist = parent.klist
        self.s64 = parent.s64
        self.device = parent.device


class Batch:
    """A batch of query-vs-ref jobs run by the persistent kernel.

    jobs: list of dicts with keys ref, query (SketchSet), and optionally
    random_mode, table, tol, out ("core_acc" | "jaccard" | "counts").  Each
    job's output lives in device memory; results() copies them back."""

    def __init__(self, jobs, device=0):
        self.device = device
        self._keep = []
        self.specs = []
        arr = (Job * len(jobs))()
        for x, j in enumerate(jobs):
            ref, query = j["ref"], j["query"]
            kind = OUT_KINDS[j.get("out", "core_acc")]
            mode, ncl, tab = _random_args(j.get("random_mode", RANDOM_NONE), j.get("table"))
            tol = j.get("tol")
            tol = default_tolerance(ref.s64) if tol is None else tol
            self_job = query is ref
            rows = ref.n * (ref.n - 1) // 2 if self_job else query.n * ref.n
            n_k = len(ref.klist)
            if kind == 0:
                shape, dt = (rows, 2), np.float32
            elif kind == 1:
                shape, dt = (rows, n_k), np.float32
            else:
                shape, dt = (rows, n_k), np.uint16
            nbytes = int(np.prod(shape)) * np.dtype(dt).itemsize
            dptr = device_alloc(nbytes, device)
            self.specs.append((dptr, shape, dt))
            self._keep += [ref, query, tab]
            arr[x] = Job(ref.handle, query.handle, mode, ncl, _ptr(tab), tol, kind, dptr)
        self._jobs = arr
        h = _p()
        try:
            _check(lib().ppsk_batch_create(device, arr, len(jobs), _c.byref(h)))
        except Exception:
            self._free_outputs()
            raise
        self._h = h

    def run(self, variant=0):
        """Run every job.  Returns (n_fail per job, variant used, kernel ms)."""
        nf = np.zeros(len(self.specs), np.uint64)
        used = _i32()
        ms = _c.c_double()
        _check(lib().ppsk_batch_run(self._h, variant, _ptr(nf), _c.byref(used), _c.byref(ms)),
               allow=(ERR_FIT,))
        return nf, used.value, ms.value

    def tuning(self):
        """(chosen plan, time per plan [0..6], row tiles of plan 6: an
        (n_jobs, 3) array of each job's row tiles of 64, 32 and 16 rows)."""
        v = _i32()
        ms = np.zeros(7, np.float64)
        tiles = np.zeros((len(self.specs), 3), np.int32)
        _check(lib().ppsk_batch_tuning(self._h, _c.byref(v), _ptr(ms), _ptr(tiles)))
        return v.value, ms, tiles

    def result(self, x):
        dptr, shape, dt = self.specs[x]
        out = np.empty(shape, dt)
        copy_to_host(out, dptr)
        return out

    def _free_outputs(self):
        for dptr, _, _ in self.specs:
            device_free(dptr)
        self.specs = []

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ppsk_batch_free(self._h)
            self._h = None
        if getattr(self, "specs", None):
            self._free_outputs()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


def sketch(samples, klist, sketchsize64, use_rc=True, device=0, min_count=0, exact=False,
           codon=False):
    """Sketch sequences on the GPU.  samples: list of samples, each a list of
    contigs (str or bytes); min_count > 0 applies the reads count filter and
    codon the codon-phased seeds (ppsk_sketch_ex).  Returns (raw
    [n][n_k][s64*14] uint64, inv_minhash_sum [n], densified [n])."""
    contigs = []
    sample_contig = [0]
    for smp in samples:
        for c in smp:
            contigs.append(c.encode() if isinstance(c, str) else bytes(c))
        sample_contig.append(len(contigs))
    seq = b"".join(contigs)
    off = np.zeros(len(contigs) + 1, np.int64)
    off[1:] = np.cumsum([len(c) for c in contigs])
    sc = np.asarray(sample_contig, np.int64)
    kl = np.ascontiguousarray(klist, np.int32)
    n = len(samples)
    raw = np.zeros((n, len(kl), sketchsize64 * 14), np.uint64)
    mh = np.zeros(n, np.float64)
    dn = np.zeros(n, np.int32)
    if min_count > 0 or codon:
        _check(lib().ppsk_sketch_ex(device, n, seq, len(seq), _ptr(off), len(contigs), _ptr(sc),
                                    len(kl), _ptr(kl), sketchsize64, int(bool(use_rc)),
                                    int(bool(codon)), int(min_count), int(bool(exact)),
                                    _ptr(raw), _ptr(mh), _ptr(dn)))
    else:
        _check(lib().ppsk_sketch(device, n, seq, len(seq), _ptr(off), len(contigs), _ptr(sc),
                                 len(kl), _ptr(kl), sketchsize64, int(bool(use_rc)), _ptr(raw),
                                 _ptr(mh), _ptr(dn)))
    return raw, mh, dn


def _copy_out(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, (n,)).copy()


class Server:
    """Serving loop over a resident reference set (ppsk_server_*).

    submit() queues a query batch and returns a ticket; wait(ticket) blocks
    (outside the GIL) and returns a dict with the kNN and threshold lists and
    the request's timestamps."""

    def __init__(self, ref, kNN=10, knn_col=0, cutoff=0.05, thr_col=0,
                 random_mode=RANDOM_NONE, tol=None, max_rows=128, n_slots=2, arena_bytes=0,
                 table=None):
        self.ref = ref
        self.random_mode = random_mode
        tol = default_tolerance(ref.s64) if tol is None else tol
        mode, ncl, tab = _random_args(random_mode, table)
        h = _p()
        _check(lib().ppsk_server_create(ref.handle, mode, ncl, _ptr(tab), tol, kNN, knn_col,
                                        cutoff, thr_col, max_rows, n_slots, arena_bytes,
                                        _c.byref(h)))
        self._h = h

    def submit(self, raw, lengths=None, clusters=None):
        """Queue a request; lengths for the equal-base correction, clusters
        (each query's nearest centroid) for a Monte Carlo table."""
        raw = np.ascontiguousarray(raw, np.uint64)
        le = None if lengths is None else np.ascontiguousarray(lengths, np.uint64)
        cl = None if clusters is None else np.ascontiguousarray(clusters, np.uint16)
        t = _i64()
        _check(lib().ppsk_server_submit(self._h, raw.shape[0], _ptr(raw), _ptr(le), _ptr(cl),
                                        _c.byref(t)))
        return t.value

    def wait(self, ticket):
        r = _c.POINTER(Result)()
        _check(lib().ppsk_server_wait(self._h, ticket, _c.byref(r)))
        try:
            x = r.contents
            out = dict(
                knn_i=_copy_out(x.knn_i, x.knn_n, np.int64),
                knn_j=_copy_out(x.knn_j, x.knn_n, np.int64),
                knn_d=_copy_out(x.knn_d, x.knn_n, np.float32),
                thr_i=_copy_out(x.thr_i, x.thr_n, np.int64),
                thr_j=_copy_out(x.thr_j, x.thr_n, np.int64),
                thr_d=_copy_out(x.thr_d, x.thr_n, np.float32),
                n_fail=x.n_fail, t_submit=x.t_submit, t_start=x.t_start, t_done=x.t_done,
                step_id=x.step_id, step_rows=x.step_rows, n_steps=x.n_steps,
                n_queries=x.n_queries)
        finally:
            lib().ppsk_result_free(r)
        return out

    def stats(self):
        st, rows, req = _i64(), _i64(), _i64()
        pk = _i32()
        pa = _c.c_size_t()
        ms = _c.c_double()
        _check(lib().ppsk_server_stats(self._h, _c.byref(st), _c.byref(rows), _c.byref(req),
                                       _c.byref(pk), _c.byref(pa), _c.byref(ms)))
        return dict(steps=st.value, rows=rows.value, requests=req.value,
                    peak_inflight=pk.value, peak_arena_bytes=pa.value, gpu_ms=ms.value)

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ppsk_server_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


def _random_args(random_mode, table):
    tab = None if table is None else np.ascontiguousarray(table, np.float32)
    ncl = 0 if tab is None else tab.shape[-1]
    return random_mode, ncl, tab


def default_tolerance(sketchsize64):
    return 5.0 / (sketchsize64 * 64)


def dense(ref, query=None, random_mode=RANDOM_NONE, table=None, tol=None, out="core_acc"):
    """Dense distances.  Returns (array, n_fail, first_fail_row)."""
    kind = OUT_KINDS[out]
    n_k = len(ref.klist)
    rows = ref.n * (ref.n - 1) // 2 if query is None else query.n * ref.n
    if kind == 0:
        res = np.zeros((rows, 2), np.float32)
    elif kind == 1:
        res = np.zeros((rows, n_k), np.float32)
    else:
        res = np.zeros((rows, n_k), np.uint16)
    mode, ncl, tab = _random_args(random_mode, table)
    nf = _u64()
    ff = _i64()
    tol = default_tolerance(ref.s64) if tol is None else tol
    _check(lib().ppsk_dense(ref.handle, None if query is None else query.handle, mode, ncl,
                            _ptr(tab), tol, kind, _ptr(res), 0, _c.byref(nf), _c.byref(ff)),
           allow=(ERR_FIT,))
    return res, nf.value, ff.value


def dense_device(ref, query=None, random_mode=RANDOM_NONE, table=None, tol=None,
                 out="core_acc", dptr=None):
    """Dense distances into device memory (dptr from device_alloc, or allocated
    here and returned).  Returns (dptr, n_fail)."""
    kind = OUT_KINDS[out]
    mode, ncl, tab = _random_args(random_mode, table)
    nf = _u64()
    ff = _i64()
    tol = default_tolerance(ref.s64) if tol is None else tol
    if dptr is None:
        rows = ref.n * (ref.n - 1) // 2 if query is None else query.n * ref.n
        per = 8 if kind == 0 else (4 if kind == 1 else 2) * len(ref.klist)
        dptr = device_alloc(rows * per, ref.device)
    _check(lib().ppsk_dense(ref.handle, None if query is None else query.handle, mode, ncl,
                            _ptr(tab), tol, kind, dptr, 1, _c.byref(nf), _c.byref(ff)),
           allow=(ERR_FIT,))
    return dptr, nf.value


def dense_noalloc(ref, dptr, query=None, random_mode=RANDOM_NONE, table=None, tol=None,
                  out="core_acc"):
    """ppsk_dense_noalloc into the device buffer dptr.  Returns (n_fail,
    first_fail_row)."""
    kind = OUT_KINDS[out]
    mode, ncl, tab = _random_args(random_mode, table)
    nf = _u64()
    ff = _i64()
    tol = default_tolerance(ref.s64) if tol is None else tol
    _check(lib().ppsk_dense_noalloc(ref.handle, None if query is None else query.handle, mode,
                                    ncl, _ptr(tab), tol, kind, dptr, _c.byref(nf),
                                    _c.byref(ff)), allow=(ERR_FIT,))
    return nf.value, ff.value


def knn(ref, kNN, dist_col=0, random_mode=RANDOM_NONE, table=None, tol=None, jaccard=False):
    """Self kNN: (i_vec, j_vec, dists, n_fail)."""
    iv = np.zeros(ref.n * kNN, np.int64)
    jv = np.zeros(ref.n * kNN, np.int64)
    dv = np.zeros(ref.n * kNN, np.float32)
    mode, ncl, tab = _random_args(random_mode, table)
    nf = _u64()
    tol = default_tolerance(ref.s64) if tol is None else tol
    _check(lib().ppsk_knn(ref.handle, mode, ncl, _ptr(tab), tol, int(bool(jaccard)), kNN,
                          dist_col, _ptr(iv), _ptr(jv), _ptr(dv), _c.byref(nf)),
           allow=(ERR_FIT,))
    return iv, jv, dv, nf.value


def _take_host_lists(pi, pj, pd, m):
    """Copy library-allocated (i, j, d) host arrays of m entries, then free them."""
    try:
        iv = np.ctypeslib.as_array(_c.cast(pi, _c.POINTER(_i64)), (max(m, 1),))[:m].copy()
        jv = np.ctypeslib.as_array(_c.cast(pj, _c.POINTER(_i64)), (max(m, 1),))[:m].copy()
        dv = np.ctypeslib.as_array(_c.cast(pd, _c.POINTER(_c.c_float)), (max(m, 1),))[:m].copy()
    finally:
        for x in (pi, pj, pd):
            if x:
                lib().ppsk_free_host(x)
    return iv, jv, dv


class PinnedBuffer:
    """Page-locked host memory from ppsk_host_alloc, viewed as a numpy array."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        self.ptr = _p()
        _check(lib().ppsk_host_alloc(max(nbytes, 1), _c.byref(self.ptr)))
        buf = (_c.c_char * max(nbytes, 1)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dtype).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            _check(lib().ppsk_host_free(self.ptr))
            self.ptr = _p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def self_pass(ref, dense_out=None, kNN=0, knn_col=0, cutoff=0.0, thr_col=0,
              random_mode=RANDOM_NONE, table=None, tol=None):
    """One self pass: condensed (core, acc) into dense_out (a float32
    [n(n-1)/2, 2] array, e.g. PinnedBuffer.array, or None), kNN lists and
    threshold lists.  Returns (knn (i, j, d) or None, thr (i, j, d) or None,
    n_fail)."""
    mode, ncl, tab = _random_args(random_mode, table)
    tol = default_tolerance(ref.s64) if tol is None else tol
    if dense_out is not None:
        rows = ref.n * (ref.n - 1) // 2
        if dense_out.dtype != np.float32 or dense_out.size != rows * 2 or \
                not dense_out.flags["C_CONTIGUOUS"]:
            raise PpskError(ERR_ARG, "dense_out must be contiguous float32 [n(n-1)/2, 2]")
    ki = kj = kd = None
    if kNN > 0:
        ki = np.zeros(ref.n * kNN, np.int64)
        kj = np.zeros(ref.n * kNN, np.int64)
        kd = np.zeros(ref.n * kNN, np.float32)
    pi, pj, pd = _p(), _p(), _p()
    n = _i64()
    nf = _u64()
    _check(lib().ppsk_self_pass(ref.handle, mode, ncl, _ptr(tab), tol, _ptr(dense_out), kNN,
                                knn_col, _ptr(ki), _ptr(kj), _ptr(kd), cutoff, thr_col,
                                _c.byref(pi), _c.byref(pj), _c.byref(pd), _c.byref(n),
                                _c.byref(nf)), allow=(ERR_FIT,))
    thr = _take_host_lists(pi, pj, pd, n.value) if cutoff > 0 else None
    return (None if kNN == 0 else (ki, kj, kd)), thr, nf.value


def threshold(ref, cutoff, dist_col=0, random_mode=RANDOM_NONE, table=None, tol=None):
    """Self threshold sparsification: (i_vec, j_vec, dists, n_fail)."""
    mode, ncl, tab = _random_args(random_mode, table)
    pi, pj, pd = _p(), _p(), _p()
    n = _i64()
    nf = _u64()
    tol = default_tolerance(ref.s64) if tol is None else tol
    _check(lib().ppsk_threshold(ref.handle, mode, ncl, _ptr(tab), tol, dist_col, cutoff,
                                _c.byref(pi), _c.byref(pj), _c.byref(pd), _c.byref(n),
                                _c.byref(nf)), allow=(ERR_FIT,))
    iv, jv, dv = _take_host_lists(pi, pj, pd, n.value)
    return iv, jv, dv, nf.value


def device_alloc(nbytes, device=0):
    p = _p()
    _check(lib().ppsk_device_alloc(device, max(int(nbytes), 1), _c.byref(p)))
    return p


def device_free(p):
    _check(lib().ppsk_device_free(p))


def copy_to_host(dst, dptr):
    _check(lib().ppsk_copy_to_host(_ptr(dst), dptr, dst.nbytes))
    return dst


def last_timing():
    ms = _c.c_double()
    nl = _i64()
    ni = _i64()
    lib().ppsk_last_timing(_c.byref(ms), _c.byref(nl), _c.byref(ni))
    return ms.value, nl.value, ni.value
