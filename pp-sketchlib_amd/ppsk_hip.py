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

OK, ERR_ARG, ERR_HIP, ERR_MEMORY, ERR_FIT, ERR_STATE = range(6)
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
    "ppsk_last_timing", "ppsk_synth_sketches",
]


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
