"""Serving loop: a resident reference set answers query batches with kNN and
threshold lists under continuous batching.

* A request's results are the same served alone and served in a mix with
  many concurrent requests (sharing steps with them), bit for bit.
* Results equal the plain path: ppsk_dense in query mode, then the
  reference heap (kNN, no sample excluded) and a row-major threshold scan
  over the same dense rows; the dense rows agree with the oracle.
"""
import threading

import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import KLIST8

pytestmark = pytest.mark.gpu

N_REF = 3000
N_Q = 700
S64 = 156
KNN = 10
CUTOFF = 0.05


@pytest.fixture(scope="module")
def data():
    raw, le, _ = H.synth(N_REF + N_Q, KLIST8, S64, seed=17)
    ref_raw, ref_le = raw[:N_REF].copy(), le[:N_REF].copy()
    q_raw, q_le = raw[N_REF:].copy(), le[N_REF:].copy()
    ref = H.SketchSet(ref_raw, KLIST8, S64, lengths=ref_le)
    yield dict(ref=ref, ref_raw=ref_raw, ref_le=ref_le, q_raw=q_raw, q_le=q_le)
    ref.free()


def _plain(ref, q_raw, q_le):
    """The plain path for one request: dense rows, heap kNN, threshold scan."""
    with H.SketchSet(q_raw, KLIST8, S64, lengths=q_le) as q:
        d, nf, _ = H.dense(ref, q, random_mode=H.RANDOM_EQUAL_BASE)
    nq = q_raw.shape[0]
    core = d[:, 0].reshape(nq, ref.n)
    ki, kj, kd = O.knn_rows(core, np.full(nq, -1), KNN)
    ti, tj = np.nonzero(core < CUTOFF)
    td = core[ti, tj]
    return d, nf, (ki, kj, kd), (ti.astype(np.int64), tj.astype(np.int64), td)


def _same(a, b):
    for k in ("knn_i", "knn_j", "thr_i", "thr_j"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("knn_d", "thr_d"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert a["n_fail"] == b["n_fail"]


@pytest.mark.parametrize("persistent", ["0", "1"])
def test_request_same_alone_and_in_mix(data, monkeypatch, persistent):
    # steps run either split per (tile, k-mer length) or on the persistent
    # kernel (large steps choose it by themselves; forced here)
    monkeypatch.setenv("PPSK_PERSISTENT", persistent)
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    a0, a1 = 100, 137  # the request under test: 37 queries
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=256, n_slots=3) as srv:
        alone = srv.wait(srv.submit(q_raw[a0:a1], q_le[a0:a1]))
        assert alone["step_rows"] == a1 - a0
        # the mix: 40 requests of 1..64 queries from 4 submitting threads,
        # the request under test among them
        rng = np.random.default_rng(5)
        spans = []
        for _ in range(40):
            s = int(rng.integers(0, N_Q - 64))
            spans.append((s, s + int(rng.integers(1, 65))))
        spans.insert(17, (a0, a1))
        tickets = [None] * len(spans)

        def submit(part):
            for x in part:
                s, e = spans[x]
                tickets[x] = srv.submit(q_raw[s:e], q_le[s:e])

        parts = [list(range(t, len(spans), 4)) for t in range(4)]
        th = [threading.Thread(target=submit, args=(p,)) for p in parts]
        for t in th:
            t.start()
        for t in th:
            t.join()
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    mixed = res[17]
    _same(alone, mixed)
    # continuous batching happened: some steps carried several requests
    assert st["steps"] < st["requests"]
    assert max(r["step_rows"] for r in res) > max(e - s for s, e in spans)
    # every request in the mix equals its own plain-path result
    for x in (0, 17, 39):
        s, e = spans[x]
        _, _, (ki, kj, kd), (ti, tj, td) = _plain(ref, q_raw[s:e], q_le[s:e])
        r = res[x]
        assert np.array_equal(r["knn_i"], ki) and np.array_equal(r["knn_j"], kj)
        assert np.array_equal(r["knn_d"].view(np.uint32), kd.view(np.uint32))
        assert np.array_equal(r["thr_i"], ti) and np.array_equal(r["thr_j"], tj)
        assert np.array_equal(r["thr_d"].view(np.uint32), td.view(np.uint32))


def test_server_against_oracle(data):
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    s, e = 300, 320
    with H.Server(ref, kNN=KNN, knn_col=1, cutoff=0.3, thr_col=1,
                  random_mode=H.RANDOM_EQUAL_BASE, max_rows=64, n_slots=2) as srv:
        r = srv.wait(srv.submit(q_raw[s:e], q_le[s:e]))
    want, _, nf, _ = O.query_db(data["ref_raw"], KLIST8, S64, 14, data["ref_le"],
                                query=q_raw[s:e], query_len=q_le[s:e],
                                random=O.Random(1, True))
    assert nf == 0 and r["n_fail"] == 0
    acc = want[:, 1].reshape(e - s, N_REF)
    ki, kj, kd = O.knn_rows(acc, np.full(e - s, -1), KNN)
    assert np.array_equal(r["knn_i"], ki)
    # distances within the core/accessory tolerance; neighbours exact unless
    # two candidates are within that tolerance of each other
    assert np.abs(r["knn_d"] - kd).max() <= 1e-6
    got_d = acc[r["knn_i"], r["knn_j"]]
    assert np.abs(got_d - r["knn_d"]).max() <= 1e-6
    _threshold_matches(r, acc, 0.3)


def _threshold_matches(r, dist, cutoff, tol=1e-6):
    """The server's threshold hits equal the oracle's on both sides, except
    pairs within the core/accessory tolerance of the cutoff."""
    got = set(zip(r["thr_i"].tolist(), r["thr_j"].tolist()))
    ti, tj = np.nonzero(dist < cutoff)
    want = set(zip(ti.tolist(), tj.tolist()))
    assert len(want) > 0
    for i, j in got ^ want:
        assert abs(float(dist[i, j]) - cutoff) <= tol, (i, j, float(dist[i, j]))
    # hits come row-major with their own distances
    order = np.lexsort((r["thr_j"], r["thr_i"]))
    assert np.array_equal(order, np.arange(len(order)))
    assert np.abs(dist[r["thr_i"], r["thr_j"]] - r["thr_d"]).max() <= tol


def test_server_monte_carlo_against_oracle(data):
    # references carry table clusters, each query its nearest centroid's
    # cluster: table[k][cluster(ref)][cluster(query)], zero from k_max = 27
    from helpers import mc_table
    ref_raw, ref_le = data["ref_raw"], data["ref_le"]
    q_raw, q_le = data["q_raw"], data["q_le"]
    rcl = (np.arange(N_REF) * 5 % 3).astype(np.uint16)
    qcl = (np.arange(N_Q) % 3).astype(np.uint16)
    table, lt = mc_table(KLIST8, symmetric=False, seed=4, zero_from=27)
    spans = [(5, 6), (100, 129), (400, 470)]
    with H.SketchSet(ref_raw, KLIST8, S64, lengths=ref_le, clusters=rcl) as ref:
        with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_MC, table=table,
                      max_rows=32, n_slots=2) as srv:
            tickets = [srv.submit(q_raw[s:e], clusters=qcl[s:e]) for s, e in spans]
            res = [srv.wait(t) for t in tickets]
        with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_MC, table=table,
                      max_rows=32, n_slots=1) as bad:
            with pytest.raises(H.PpskError, match="clusters"):
                bad.submit(q_raw[:2])
    for (s, e), r in zip(spans, res):
        want, _, nf, _ = O.query_db(ref_raw, KLIST8, S64, 14, ref_le, query=q_raw[s:e],
                                    query_len=q_le[s:e], ref_cluster=rcl,
                                    query_cluster=qcl[s:e], random=O.Random(2, True, table, lt))
        assert nf == 0 and r["n_fail"] == 0
        core = want[:, 0].reshape(e - s, N_REF)
        ki, kj, kd = O.knn_rows(core, np.full(e - s, -1), KNN)
        assert np.array_equal(r["knn_i"], ki)
        assert np.abs(r["knn_d"] - kd).max() <= 1e-6
        assert np.abs(core[r["knn_i"], r["knn_j"]] - r["knn_d"]).max() <= 1e-6
        _threshold_matches(r, core, CUTOFF)


@pytest.mark.parametrize("max_rows,n_slots", [(24, 1), (24, 3), (1, 2)])
def test_requests_split_over_steps(data, max_rows, n_slots):
    # requests larger than the room in a step are served in pieces of at most
    # max_rows rows (down to one row per step), pieces of different requests
    # sharing steps; each result equals the same request served in one step of
    # its own
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    spans = [(10, 60), (200, 207), (300, 333), (400, 420), (500, 501), (600, 664)]
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=64, n_slots=1) as whole:
        want = [whole.wait(whole.submit(q_raw[s:e], q_le[s:e])) for s, e in spans]
    assert all(w["n_steps"] == 1 for w in want)
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=max_rows, n_slots=n_slots) as srv:
        tickets = [srv.submit(q_raw[s:e], q_le[s:e]) for s, e in spans]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for (s, e), w, r in zip(spans, want, res):
        _same(w, r)
        assert r["n_queries"] == e - s
        assert r["n_steps"] >= -(-(e - s) // max_rows)
    assert st["rows"] == sum(e - s for s, e in spans)
    assert st["requests"] == len(spans)
    assert st["steps"] >= -(-st["rows"] // max_rows)  # no step above max_rows
    # the 64-query request against the plain path
    _, _, (ki, kj, kd), (ti, tj, td) = _plain(ref, q_raw[600:664], q_le[600:664])
    r = res[-1]
    assert np.array_equal(r["knn_i"], ki) and np.array_equal(r["knn_j"], kj)
    assert np.array_equal(r["knn_d"].view(np.uint32), kd.view(np.uint32))
    assert np.array_equal(r["thr_i"], ti) and np.array_equal(r["thr_j"], tj)
    assert np.array_equal(r["thr_d"].view(np.uint32), td.view(np.uint32))


def test_server_rejects_bad_requests(data):
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=16, n_slots=1) as srv:
        with pytest.raises(H.PpskError, match="lengths"):
            srv.submit(q_raw[:2])
        with pytest.raises(H.PpskError, match="ticket"):
            srv.wait(12345)


def test_admission_by_memory(data):
    # an arena that holds about one 64-query step: with 3 slots the workers
    # must wait for device memory, never exceed the arena, and still return
    # the same results as requests served alone
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    arena = 48 << 20
    spans = [(x * 64, x * 64 + 64) for x in range(8)]
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=64, n_slots=3) as solo:
        alone = [solo.wait(solo.submit(q_raw[s:e], q_le[s:e])) for s, e in spans]
    with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=64, n_slots=3, arena_bytes=arena) as srv:
        tickets = [srv.submit(q_raw[s:e], q_le[s:e]) for s, e in spans]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert 0 < st["peak_arena_bytes"] <= arena
    assert st["peak_inflight"] == 1  # two 64-query steps never fit together
    for a, b in zip(alone, res):
        _same(a, b)
    with pytest.raises(H.PpskError, match="arena"):
        H.Server(ref, kNN=KNN, cutoff=CUTOFF, max_rows=64, n_slots=1, arena_bytes=1 << 20)


def test_persistent_step_tail_split(monkeypatch):
    # 33,000 references: a 512-query step is 8 x 129 tiles of 64 x 256 on the
    # persistent kernel, two full rounds of its 512 resident workgroups and 8
    # tiles more, which run as (tile, k-mer length) items after it; results
    # equal the step without that split and the plain path.  The threshold
    # kernels cut each row into 4 column segments here (8 for the 40-query
    # request below)
    n_ref, n_q = 33_000, 512
    # core divergence up to 15% keeps the 0.05 threshold to close relatives
    raw, le, _ = H.synth(n_ref + n_q, KLIST8, S64, seed=23, dcore_max=0.15)
    q_raw, q_le = raw[n_ref:].copy(), le[n_ref:].copy()
    with H.SketchSet(raw[:n_ref], KLIST8, S64, lengths=le[:n_ref]) as ref:
        del raw
        res = {}
        for split in ("1", "0"):
            monkeypatch.setenv("PPSK_TAIL_SPLIT", split)
            with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                          max_rows=n_q, n_slots=1) as srv:
                res[split] = srv.wait(srv.submit(q_raw, q_le))
            assert res[split]["n_steps"] == 1
        _same(res["1"], res["0"])
        with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                      max_rows=n_q, n_slots=1) as srv:
            small = srv.wait(srv.submit(q_raw[100:140], q_le[100:140]))
        for r, (s, e) in ((res["1"], (0, n_q)), (small, (100, 140))):
            _, nf, (ki, kj, kd), (ti, tj, td) = _plain(ref, q_raw[s:e], q_le[s:e])
            assert r["n_fail"] == nf
            assert np.array_equal(r["knn_i"], ki) and np.array_equal(r["knn_j"], kj)
            assert np.array_equal(r["knn_d"].view(np.uint32), kd.view(np.uint32))
            assert len(ti) > 0
            assert np.array_equal(r["thr_i"], ti) and np.array_equal(r["thr_j"], tj)
            assert np.array_equal(r["thr_d"].view(np.uint32), td.view(np.uint32))


def test_threshold_hits_beyond_buffer_capacity(data):
    # cutoff 1.0 keeps nearly every pair: a 64-query step has ~192,000 hits,
    # more than the page-locked buffers first hold (64 per row, at least
    # 65,536), so the step writes its hits again into larger buffers; the next
    # step fits them at once
    ref, q_raw, q_le = data["ref"], data["q_raw"], data["q_le"]
    spans = [(0, 64), (64, 128)]
    with H.Server(ref, kNN=KNN, cutoff=1.0, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=64, n_slots=1) as srv:
        res = [srv.wait(srv.submit(q_raw[s:e], q_le[s:e])) for s, e in spans]
    for (s, e), r in zip(spans, res):
        with H.SketchSet(q_raw[s:e], KLIST8, S64, lengths=q_le[s:e]) as q:
            d, _, _ = H.dense(ref, q, random_mode=H.RANDOM_EQUAL_BASE)
        core = d[:, 0].reshape(e - s, N_REF)
        ti, tj = np.nonzero(core < 1.0)
        assert len(ti) > 65536
        assert np.array_equal(r["thr_i"], ti) and np.array_equal(r["thr_j"], tj)
        assert np.array_equal(r["thr_d"].view(np.uint32), core[ti, tj].view(np.uint32))


def test_knn_candidate_lists_overflow():
    # 33,000 references ordered by decreasing core distance to the query: a
    # one-query step cuts its row into 8 column segments, each of which admits
    # every one of its ~4,000 columns, far more than a candidate list holds,
    # so the merge scans those segments directly; the kNN still equals the
    # reference heap over the whole row (ties included)
    n_ref = 33_000
    raw, le, _ = H.synth(n_ref + 1, KLIST8, S64, seed=29, dcore_max=0.15)
    q_raw, q_le = raw[n_ref:].copy(), le[n_ref:].copy()
    with H.SketchSet(raw[:n_ref], KLIST8, S64, lengths=le[:n_ref]) as ref0, \
            H.SketchSet(q_raw, KLIST8, S64, lengths=q_le) as q:
        d0, _, _ = H.dense(ref0, q, random_mode=H.RANDOM_EQUAL_BASE)
    order = np.argsort(-d0[:, 0], kind="stable")
    with H.SketchSet(raw[:n_ref][order], KLIST8, S64, lengths=le[:n_ref][order]) as ref:
        del raw
        with H.Server(ref, kNN=KNN, cutoff=CUTOFF, random_mode=H.RANDOM_EQUAL_BASE,
                      max_rows=64, n_slots=1) as srv:
            r = srv.wait(srv.submit(q_raw, q_le))
        d, nf, (ki, kj, kd), _ = _plain(ref, q_raw, q_le)
    assert np.all(np.diff(d[:, 0]) <= 0)  # non-increasing along the row
    assert np.array_equal(r["knn_i"], ki) and np.array_equal(r["knn_j"], kj)
    assert np.array_equal(r["knn_d"].view(np.uint32), kd.view(np.uint32))
