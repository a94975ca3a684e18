"""ppsk_dense_noalloc (fixed device scratch, no allocation) against ppsk_dense.

Collected last (file name), so that it also runs after every other GPU test
of the process has used the device.

Both entry points are called through the C ABI on the same resident sets;
their device outputs must be byte-identical (counts, Jaccards and
core/accessory are all deterministic functions of each pair).  Every case
runs two calls in a row, the second on a smaller input, since the scratch
array keeps the previous call's contents.  The oracle pins one case so that
a shared mistake of both entry points cannot pass.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import KLIST8, mc_table, synth

pytestmark = pytest.mark.gpu


def _out_spec(ref, query, out):
    rows = ref.n * (ref.n - 1) // 2 if query is None else query.n * ref.n
    n_k = len(ref.klist)
    if out == "core_acc":
        return (rows, 2), np.float32
    if out == "jaccard":
        return (rows, n_k), np.float32
    return (rows, n_k), np.uint16


def _both(ref, query=None, **kw):
    """(ppsk_dense result, ppsk_dense_noalloc result) as host arrays."""
    shape, dt = _out_spec(ref, query, kw.get("out", "core_acc"))
    nbytes = int(np.prod(shape)) * np.dtype(dt).itemsize
    a = H.device_alloc(nbytes, ref.device)
    b = H.device_alloc(nbytes, ref.device)
    try:
        _, nf_a = H.dense_device(ref, query, dptr=a, **kw)
        nf_b, _ = H.dense_noalloc(ref, b, query=query, **kw)
        assert nf_a == nf_b
        ha = H.copy_to_host(np.empty(shape, dt), a)
        hb = H.copy_to_host(np.empty(shape, dt), b)
    finally:
        H.device_free(a)
        H.device_free(b)
    return ha, hb


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("out", ["core_acc", "counts", "jaccard"])
def test_self_two_calls(out):
    H.device_init(0)
    raw, le, _ = synth(700)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        a, b = _both(s, random_mode=H.RANDOM_EQUAL_BASE, out=out)
        _same(a, b)
        with s.prefix(300) as v:  # second call, smaller input
            a2, b2 = _both(v, random_mode=H.RANDOM_EQUAL_BASE, out=out)
        _same(a2, b2)
    if out == "counts":
        ref_c = O.query_db(raw[:300], KLIST8, 156, 14, le[:300], random=O.Random(1, True),
                           want_counts=True)[1]
        assert np.array_equal(b2, ref_c)


def test_query_mc_two_calls():
    H.device_init(0)
    raw, le, _ = synth(900)
    table, _ = mc_table(KLIST8, symmetric=False, zero_from=27)
    rc = (np.arange(650) * 5 % 3).astype(np.uint16)
    qc = (np.arange(250) % 3).astype(np.uint16)
    with H.SketchSet(raw[250:], KLIST8, 156, lengths=le[250:], clusters=rc) as r, \
            H.SketchSet(raw[:250], KLIST8, 156, lengths=le[:250], clusters=qc) as q:
        a, b = _both(r, q, random_mode=H.RANDOM_MC, table=table)
        _same(a, b)
        with q.prefix(40) as qv:
            a2, b2 = _both(r, qv, random_mode=H.RANDOM_MC, table=table)
        _same(a2, b2)


def test_many_launches():
    """A pass larger than the scratch holds: several launches of the tile kernel."""
    H.device_init(0)
    raw, le, _ = synth(4000)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        a, b = _both(s, random_mode=H.RANDOM_EQUAL_BASE)
        _same(a, b)
        with s.prefix(1000) as v:
            a2, b2 = _both(v, random_mode=H.RANDOM_EQUAL_BASE)
        _same(a2, b2)


def test_table_beyond_scratch_rejected():
    H.device_init(0)
    raw, le, _ = synth(300)
    C = 1600  # 8 x 1600^2 floats = 82 MB > the 64 MiB array
    table = np.zeros((len(KLIST8), C, C), np.float32)
    cl = (np.arange(300) % 3).astype(np.uint16)
    with H.SketchSet(raw, KLIST8, 156, lengths=le, clusters=cl) as s:
        dptr = H.device_alloc(300 * 299 // 2 * 8)
        try:
            with pytest.raises(H.PpskError) as e:
                H.dense_noalloc(s, dptr, random_mode=H.RANDOM_MC, table=table)
            assert e.value.code == H.ERR_ARG
            assert "scratch" in str(e.value)
        finally:
            H.device_free(dptr)
