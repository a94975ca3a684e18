"""HIP path vs the CPU oracle on seeded synthetic sketch sets.

Bar: per-k match counts and per-k Jaccards bit-exact; core/accessory within
1e-6 absolute of the oracle (the contract with the reference CPU path is
1e-5; the oracle itself reproduces the reference's golden core/accessory to
~1.2e-7); kNN and threshold indices and distances exact.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import KLIST8, mc_table, synth, with_duplicates, max_abs_diff

pytestmark = pytest.mark.gpu

CORE_ACC_TOL = 1e-6


def _oracle_random(mode, table=None, lt=None):
    if mode == H.RANDOM_MC:
        return O.Random(2, True, table, lt)
    return O.Random(mode, True)


@pytest.mark.parametrize("n", [1, 2, 127, 300])
def test_self_counts_and_jaccard_bit_exact(n):
    raw, le, _ = synth(max(n, 2))
    raw, le = raw[:n], le[:n]
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        cnt, nf, _ = H.dense(s, out="counts")
        jac, _, _ = H.dense(s, out="jaccard", random_mode=H.RANDOM_EQUAL_BASE)
    ref_j, ref_c, _, _ = O.query_db(raw, KLIST8, 156, 14, le, random=O.Random(1, True),
                                    jaccard=True, want_counts=True)
    assert cnt.shape == ref_c.shape
    assert np.array_equal(cnt, ref_c)
    assert np.array_equal(jac.view(np.uint32), ref_j.view(np.uint32))


@pytest.mark.parametrize("mode", [H.RANDOM_NONE, H.RANDOM_EQUAL_BASE, H.RANDOM_MC])
def test_self_core_acc(mode):
    raw, le, _ = synth(300)
    table = lt = cl = None
    if mode == H.RANDOM_MC:
        table, lt = mc_table(KLIST8, zero_from=27)
        cl = (np.arange(300) * 7 % 3).astype(np.uint16)
    with H.SketchSet(raw, KLIST8, 156, lengths=le, clusters=cl) as s:
        d, nf, _ = H.dense(s, random_mode=mode, table=table)
    ref, _, rnf, _ = O.query_db(raw, KLIST8, 156, 14, le, ref_cluster=cl,
                                random=_oracle_random(mode, table, lt))
    assert nf == rnf == 0
    assert np.abs(d - ref).max() <= CORE_ACC_TOL


@pytest.mark.parametrize("mode", [H.RANDOM_NONE, H.RANDOM_EQUAL_BASE, H.RANDOM_MC])
def test_query_vs_ref(mode):
    raw, le, _ = synth(300)
    q_raw, q_le = raw[:77].copy(), le[:77].copy()
    r_raw, r_le = raw[77:].copy(), le[77:].copy()
    table = lt = None
    qc = rc = None
    if mode == H.RANDOM_MC:
        table, lt = mc_table(KLIST8, symmetric=False, zero_from=29)
        qc = (np.arange(77) % 3).astype(np.uint16)
        rc = (np.arange(223) * 5 % 3).astype(np.uint16)
    with H.SketchSet(r_raw, KLIST8, 156, lengths=r_le, clusters=rc) as r, \
            H.SketchSet(q_raw, KLIST8, 156, lengths=q_le, clusters=qc) as q:
        d, nf, _ = H.dense(r, q, random_mode=mode, table=table)
        c, _, _ = H.dense(r, q, out="counts")
        j, _, _ = H.dense(r, q, random_mode=mode, table=table, out="jaccard")
    ref, ref_c, rnf, _ = O.query_db(r_raw, KLIST8, 156, 14, r_le, query=q_raw, query_len=q_le,
                                    ref_cluster=rc, query_cluster=qc,
                                    random=_oracle_random(mode, table, lt), want_counts=True)
    ref_j, _, _, _ = O.query_db(r_raw, KLIST8, 156, 14, r_le, query=q_raw, query_len=q_le,
                                ref_cluster=rc, query_cluster=qc,
                                random=_oracle_random(mode, table, lt), jaccard=True)
    assert np.array_equal(c, ref_c)
    assert np.array_equal(j.view(np.uint32), ref_j.view(np.uint32))
    assert nf == rnf == 0
    assert np.abs(d - ref).max() <= CORE_ACC_TOL


@pytest.mark.parametrize("knn,col", [(1, 0), (5, 0), (5, 1), (40, 0)])
def test_knn_matches_heap_order(knn, col):
    raw, le, _ = synth(260)
    # exact duplicates create tied distances (including ties at 0)
    raw, le = with_duplicates(raw, le, [(3, 10), (11, 10), (200, 10), (57, 58), (90, 58)])
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        iv, jv, dv, nf = H.knn(s, knn, dist_col=col, random_mode=H.RANDOM_EQUAL_BASE)
    ri, rj, rd, rnf = O.query_db_sparse(raw, KLIST8, 156, 14, le, knn, dist_col=col,
                                        random=O.Random(1, True))
    assert nf == rnf == 0
    assert np.array_equal(iv, ri)
    # distances agree to the core/acc tolerance; indices exactly where the
    # distance is not within tolerance of a tie
    assert max_abs_diff(dv, rd) <= CORE_ACC_TOL
    assert np.array_equal(jv, rj)


@pytest.mark.parametrize("mode", [H.RANDOM_EQUAL_BASE, H.RANDOM_MC])
def test_knn_above_lds_heap(mode):
    # kNN = 3,000 > 2,048: each row's heap lives in device memory instead of
    # LDS; same heap algorithm, so the same lists and tie order as the
    # reference's priority_queue (src/api.cpp:360-413).  The asymmetric
    # Monte Carlo table takes the full-square band path.
    n, k = 3300, 3000
    raw, le, _ = synth(n)
    raw, le = with_duplicates(raw, le, [(3, 10), (11, 10), (2000, 10)])
    table = lt = cl = None
    if mode == H.RANDOM_MC:
        table, lt = mc_table(KLIST8, symmetric=False, zero_from=27)
        cl = (np.arange(n) * 7 % 3).astype(np.uint16)
    with H.SketchSet(raw, KLIST8, 156, lengths=le, clusters=cl) as s:
        iv, jv, dv, nf = H.knn(s, k, random_mode=mode, table=table)
    ri, rj, rd, rnf = O.query_db_sparse(raw, KLIST8, 156, 14, le, k, ref_cluster=cl,
                                        random=_oracle_random(mode, table, lt))
    assert nf == rnf == 0
    assert np.array_equal(iv, ri)
    assert max_abs_diff(dv, rd) <= CORE_ACC_TOL
    assert np.array_equal(jv, rj)


def test_knn_equals_n():
    raw, le, _ = synth(40)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        iv, jv, dv, _ = H.knn(s, 40)
    ri, rj, rd, _ = O.query_db_sparse(raw, KLIST8, 156, 14, le, 40)
    assert np.array_equal(jv, rj)
    assert np.isinf(dv.reshape(40, 40)[:, -1]).all()


@pytest.mark.parametrize("cutoff,col", [(0.01, 0), (0.05, 1), (0.0, 0), (1.0, 0)])
def test_threshold(cutoff, col):
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        if cutoff == 0.0:
            iv, jv, dv, _ = H.threshold(s, cutoff, dist_col=col)
            assert len(iv) == 0
            return
        iv, jv, dv, nf = H.threshold(s, cutoff, dist_col=col, random_mode=H.RANDOM_EQUAL_BASE)
    dense, _, _, _ = O.query_db(raw, KLIST8, 156, 14, le, random=O.Random(1, True))
    square = O.long_to_square(dense[:, col])
    ri, rj, rd = O.sparsify_threshold(square, cutoff)
    assert np.array_equal(iv, ri)
    assert np.array_equal(jv, rj)
    assert max_abs_diff(dv, rd) <= CORE_ACC_TOL


@pytest.mark.parametrize("chunk_bytes", [None, "1"])
def test_self_pass_equals_separate_calls(monkeypatch, chunk_bytes):
    # chunk_bytes "1": one row tile per launch, so the host copies of the
    # dense rows run block by block behind the computation
    if chunk_bytes:
        monkeypatch.setenv("PPSK_CHUNK_BYTES", chunk_bytes)
    raw, le, _ = synth(300)
    raw, le = with_duplicates(raw, le, [(3, 10), (57, 58)])
    n = raw.shape[0]
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        d_sep, _, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
        k_sep = H.knn(s, 10, dist_col=1, random_mode=H.RANDOM_EQUAL_BASE)
        t_sep = H.threshold(s, 0.02, dist_col=0, random_mode=H.RANDOM_EQUAL_BASE)
        _, launches, _ = H.last_timing()
        with H.PinnedBuffer((n * (n - 1) // 2, 2), np.float32) as pin:
            kn, th, nf = H.self_pass(s, pin.array, kNN=10, knn_col=1, cutoff=0.02, thr_col=0,
                                     random_mode=H.RANDOM_EQUAL_BASE)
            dense = pin.array.copy()
        _, launches_pass, _ = H.last_timing()
    if chunk_bytes:
        assert launches_pass > 1
    assert nf == 0
    assert np.array_equal(dense.view(np.uint32), d_sep.view(np.uint32))
    for a, b in zip(kn, k_sep[:3]):
        assert np.array_equal(a, b)
    for a, b in zip(th, t_sep[:3]):
        assert np.array_equal(a, b)
    ref, _, _, _ = O.query_db(raw, KLIST8, 156, 14, le, random=O.Random(1, True))
    assert np.abs(dense - ref).max() <= CORE_ACC_TOL


def test_dense_streamed_chunks_match_single_launch(monkeypatch):
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        one, _, _ = H.dense(s, out="counts")
        q = H.SketchSet(raw[:50].copy(), KLIST8, 156, lengths=le[:50].copy())
        with q:
            qone, _, _ = H.dense(s, q, random_mode=H.RANDOM_EQUAL_BASE)
            monkeypatch.setenv("PPSK_CHUNK_BYTES", "1")
            many, _, _ = H.dense(s, out="counts")
            _, launches, _ = H.last_timing()
            qmany, _, _ = H.dense(s, q, random_mode=H.RANDOM_EQUAL_BASE)
    assert launches > 1
    assert np.array_equal(one, many)
    assert np.array_equal(qone.view(np.uint32), qmany.view(np.uint32))


def test_knn_asymmetric_table_band_path():
    # an asymmetric Monte Carlo table makes d(i, j) != d(j, i): kNN must
    # come from full square rows, not from the triangle
    raw, le, _ = synth(200)
    table, lt = mc_table(KLIST8, symmetric=False, zero_from=27)
    cl = (np.arange(200) * 7 % 3).astype(np.uint16)
    with H.SketchSet(raw, KLIST8, 156, lengths=le, clusters=cl) as s:
        iv, jv, dv, nf = H.knn(s, 6, dist_col=0, random_mode=H.RANDOM_MC, table=table)
        with pytest.raises(H.PpskError):
            H.self_pass(s, None, kNN=6, random_mode=H.RANDOM_MC, table=table)
    ri, rj, rd, rnf = O.query_db_sparse(raw, KLIST8, 156, 14, le, 6, ref_cluster=cl,
                                        random=O.Random(2, True, table, lt))
    assert nf == rnf == 0
    assert np.array_equal(iv, ri) and np.array_equal(jv, rj)
    assert max_abs_diff(dv, rd) <= CORE_ACC_TOL


@pytest.mark.parametrize("persistent", ["1", "0"])
@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_every_tile_shape(monkeypatch, variant, persistent):
    # every compiled tile configuration, self and query mode, against the
    # oracle (sample counts not a multiple of any tile edge); shapes 2, 4 and
    # 5 run both as a persistent kernel (5: the mixed-shape kernel) and as one
    # launch per (tile, k)
    if persistent == "1" and variant not in (2, 4, 5):
        pytest.skip("the persistent kernels are built for shapes 2, 4 and 5 only")
    monkeypatch.setenv("PPSK_TILE_VARIANT", str(variant))
    monkeypatch.setenv("PPSK_PERSISTENT", persistent)
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        c, _, _ = H.dense(s, out="counts")
        d, nf, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
        with H.SketchSet(raw[:45].copy(), KLIST8, 156, lengths=le[:45].copy()) as q:
            qd, _, _ = H.dense(s, q, random_mode=H.RANDOM_EQUAL_BASE)
    ref, ref_c, _, _ = O.query_db(raw, KLIST8, 156, 14, le, random=O.Random(1, True),
                                  want_counts=True)
    qref, _, _, _ = O.query_db(raw, KLIST8, 156, 14, le, query=raw[:45], query_len=le[:45],
                               random=O.Random(1, True))
    assert nf == 0
    assert np.array_equal(c, ref_c)
    assert np.abs(d - ref).max() <= CORE_ACC_TOL
    assert np.abs(qd - qref).max() <= CORE_ACC_TOL


def test_row_prefetch_off_is_identical(monkeypatch):
    # the row-panel L2 prefetch only moves data earlier: outputs are unchanged
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        on, _, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
        monkeypatch.setenv("PPSK_ROW_PREFETCH", "0")
        off, _, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
    assert np.array_equal(on.view(np.uint32), off.view(np.uint32))


@pytest.mark.parametrize("variant", [2, 4, 5])
def test_tail_split_persistent_pass(monkeypatch, variant):
    # a pass just over one round of whole tiles: its partial last round runs
    # as (tile, k) items of the per-(tile, k) kernel; outputs are identical to
    # the whole-tile run and to the oracle.  With 64 x 256 tiles (variant 2)
    # the tail of this self pass is its diagonal tiles whose first row lies 128
    # or 192 columns past their first column (31 of them at 4,000 samples),
    # which dist_tiles_diagonal runs with their leading column blocks skipped
    monkeypatch.setenv("PPSK_TILE_VARIANT", str(variant))
    monkeypatch.setenv("PPSK_PERSISTENT", "1")
    # just over one round of the 512 resident workgroups
    n = {2: 4000, 4: 2900, 5: 2000}[variant]
    raw, le, _ = synth(n)
    with H.SketchSet(raw, KLIST8, 156, lengths=le) as s:
        c, _, _ = H.dense(s, out="counts")
        launches = H.last_timing()[1]
        d, nf, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
        monkeypatch.setenv("PPSK_TAIL_SPLIT", "0")
        c0, _, _ = H.dense(s, out="counts")
        launches0 = H.last_timing()[1]
        d0, _, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
    assert launches == launches0 + 1, (launches, launches0)  # the tail launch ran
    assert nf == 0
    assert np.array_equal(c, c0)
    assert np.array_equal(d.view(np.uint32), d0.view(np.uint32))
    ref, ref_c, _, _ = O.query_db(raw, KLIST8, 156, 14, le, random=O.Random(1, True),
                                  want_counts=True)
    assert np.array_equal(c, ref_c)
    assert np.abs(d - ref).max() <= CORE_ACC_TOL
