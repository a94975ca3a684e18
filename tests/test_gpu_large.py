"""Large all-vs-all on one GPU: 32,768 samples, 8 k-mer lengths, 10,000 bins.

One self pass over 536,854,528 pairs streams the dense core/accessory
condensed matrix (4.3 GB) to page-locked host memory in row blocks while
later blocks compute, and returns kNN = 10 lists from the same distances.

Checked on a 1% sample of the samples (328 rows; their pairs with every
other sample, about 2% of all pairs):
  * per-k match counts bit-exact against the oracle;
  * streamed core/accessory within 1e-6 of the oracle;
  * kNN lists equal to the reference heap replayed over the streamed rows.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import KLIST8

pytestmark = pytest.mark.gpu

N = 32768
S64 = 156
KNN = 10
TOL = 1e-6


def _rows_from_condensed(dense, rows, n):
    """Dense rows d[r, :] (with +inf at j == r) gathered from condensed form."""
    j = np.arange(n, dtype=np.int64)
    out = np.empty((len(rows), n), np.float32)
    for x, r in enumerate(rows.tolist()):
        a = np.minimum(j, r)
        b = np.maximum(j, r)
        idx = n * a - (a * (a + 1)) // 2 + b - 1 - a
        idx[r] = 0
        out[x] = dense[idx]
        out[x, r] = np.inf
    return out


def test_large_self_pass_streamed_dense_and_knn():
    raw, le, _ = H.synth(N, KLIST8, S64, seed=3)
    rng = np.random.default_rng(7)
    rows = np.sort(rng.choice(N, N // 100, replace=False))
    pairs = N * (N - 1) // 2
    with H.SketchSet(raw, KLIST8, S64, lengths=le) as s, \
            H.PinnedBuffer((pairs, 2), np.float32) as pin:
        kn, _, nf = H.self_pass(s, pin.array, kNN=KNN, knn_col=0,
                                random_mode=H.RANDOM_EQUAL_BASE)
        _, launches, _ = H.last_timing()
        with H.SketchSet(raw[rows], KLIST8, S64, lengths=le[rows]) as q:
            counts, _, _ = H.dense(s, q, out="counts")
        core = pin.array[:, 0].copy()
        acc = pin.array[:, 1].copy()
    assert nf == 0
    assert launches > 1  # the dense output was streamed in several row blocks
    want_d, want_c, rnf, _ = O.query_db(raw, KLIST8, S64, 14, le, query=raw[rows],
                                        query_len=le[rows], random=O.Random(1, True),
                                        want_counts=True)
    assert rnf == 0
    assert np.array_equal(counts, want_c)
    got_core = _rows_from_condensed(core, rows, N)
    got_acc = _rows_from_condensed(acc, rows, N)
    want_core = want_d[:, 0].reshape(len(rows), N).copy()
    want_acc = want_d[:, 1].reshape(len(rows), N).copy()
    for x, r in enumerate(rows.tolist()):
        want_core[x, r] = np.inf
        want_acc[x, r] = np.inf
    fin = np.isfinite(want_core)
    assert np.array_equal(fin, np.isfinite(got_core))
    assert np.abs(got_core[fin] - want_core[fin]).max() <= TOL
    assert np.abs(got_acc[fin] - want_acc[fin]).max() <= TOL
    ki, kj, kd = (a.reshape(N, KNN)[rows].ravel() for a in kn)
    ei, ej, ed = O.knn_rows(got_core, rows, KNN)
    assert np.array_equal(ki, ei)
    assert np.array_equal(kj, ej)
    assert np.array_equal(kd.view(np.uint32), ed.view(np.uint32))
