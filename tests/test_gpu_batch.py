"""Persistent batch kernel: many uneven query-vs-ref jobs in one launch.

Every job's output must equal the single-job ppsk_dense call on the same
sets (bit for bit: both run the same epilogue on the same counts), and a
subset is checked against the oracle (counts and Jaccards bit-exact,
core/accessory within 1e-6).  Jobs mix k-sets (including k >= 33, where only
no correction or a Monte Carlo table applies), sketch sizes, sample counts
from 1 query upwards, random modes and output kinds, and query prefixes of
resident databases.  Self jobs (query set = reference set) give the
condensed upper triangle, as ppsk_dense with no query set does.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import mc_table
import functools

pytestmark = pytest.mark.gpu

TOL = 1e-6

# klist, s64, reference DB size, query DB size
CLASSES = [
    ([15, 17, 19, 21], 16, 800, 70),
    ([13, 16, 19, 22, 25, 28], 47, 1500, 200),
    ([15, 17, 19, 21, 23, 25, 27, 29], 156, 600, 130),
    ([21, 24, 27, 30, 33, 36, 39], 30, 900, 40),
]
# class, refs, queries (0: self job, the refs against themselves), random
# correction, output
JOBS = [
    (0, 700, 1, "none", "core_acc"),
    (0, 800, 33, "equal", "jaccard"),
    (1, 1500, 200, "equal", "core_acc"),
    (1, 513, 64, "mc", "core_acc"),
    (2, 600, 130, "equal", "core_acc"),
    (2, 257, 2, "none", "counts"),
    (3, 900, 40, "mc", "core_acc"),
    (3, 129, 1, "none", "jaccard"),
    (1, 300, 65, "none", "counts"),
    (0, 1, 7, "none", "core_acc"),
    (2, 300, 0, "equal", "core_acc"),
    (1, 590, 0, "mc", "jaccard"),
]
ORACLE_JOBS = [0, 1, 3, 5, 6, 7, 9, 10, 11]


def _mode(name):
    return {"none": H.RANDOM_NONE, "equal": H.RANDOM_EQUAL_BASE, "mc": H.RANDOM_MC}[name]


@pytest.fixture(scope="module")
def dbs():
    out = []
    for ci, (kl, s64, nr, nq) in enumerate(CLASSES):
        # references and queries from one tree, so every pair is related
        allr, alle, _ = H.synth(nr + nq, kl, s64, seed=100 + ci)
        raw, le = allr[:nr].copy(), alle[:nr].copy()
        qraw, qle = allr[nr:].copy(), alle[nr:].copy()
        can_len = max(kl) < 33
        rcl = (np.arange(nr) * 7 % 3).astype(np.uint16)
        qcl = (np.arange(nq) % 3).astype(np.uint16)
        ref = H.SketchSet(raw, kl, s64, lengths=le if can_len else None, clusters=rcl)
        qry = H.SketchSet(qraw, kl, s64, lengths=qle if can_len else None, clusters=qcl)
        table, lt = mc_table(kl, symmetric=False, seed=ci, zero_from=kl[-2])
        out.append(dict(kl=kl, s64=s64, raw=raw, le=le, qraw=qraw, qle=qle, rcl=rcl, qcl=qcl,
                        ref=ref, qry=qry, table=table, lt=lt))
    yield out
    for d in out:
        d["qry"].free()
        d["ref"].free()


def _job_sets(dbs, spec):
    ci, nr, nq, rnd, kind = spec
    d = dbs[ci]
    ref = d["ref"].prefix(nr)
    return d, ref, (ref if nq == 0 else d["qry"].prefix(nq))


def _single(d, ref, qry, rnd, kind):
    table = d["table"] if rnd == "mc" else None
    out, nf, _ = H.dense(ref, None if qry is ref else qry, random_mode=_mode(rnd), table=table,
                         out=kind)
    return out, nf


@pytest.fixture(scope="module")
def batch_run(dbs):
    views = []
    jobs = []
    for spec in JOBS:
        d, ref, qry = _job_sets(dbs, spec)
        views.append((d, ref, qry))
        jobs.append(dict(ref=ref, query=qry, random_mode=_mode(spec[3]),
                         table=d["table"] if spec[3] == "mc" else None, out=spec[4]))
    b = H.Batch(jobs)
    yield b, views
    b.free()


TILE_ROWS = {2: 64, 4: 32, 5: 16}


def _kept_cols(spec, row0):
    """Column tiles (256 wide) of the row tile starting at row0: all of them
    for a query job, those reaching past the diagonal for a self job."""
    _, nr, nq, _, _ = spec
    nct = -(-nr // 256)
    return nct if nq else nct - min(nct, (row0 + 1) // 256)


def _tile_ops(spec, v, row0):
    ci = spec[0]
    kl, s64 = CLASSES[ci][0], CLASSES[ci][1]
    return _kept_cols(spec, row0) * TILE_ROWS[v] * 256 * len(kl) * s64 * 30


def _uniform_ops(spec, v):
    """Matching ops of a job's tiles in shape v, padding rows included."""
    rows = spec[2] or spec[1]
    tr = TILE_ROWS[v]
    return sum(_tile_ops(spec, v, b * tr) for b in range(-(-rows // tr)))


def _best_split_time(spec, rate):
    """Least estimated time over every way to lay row tiles of 64, 32 and 16
    rows from row 0 (each at a multiple of its height) until the job's rows
    are covered: brute force over all such sequences."""
    rows = spec[2] or spec[1]

    @functools.lru_cache(maxsize=None)
    def best(row0):
        if row0 >= rows:
            return 0.0
        return min(_tile_ops(spec, v, row0) / rate[v] + best(row0 + tr)
                   for v, tr in TILE_ROWS.items() if row0 % tr == 0)
    return best(0)


def test_batch_autotuned_equals_single_calls(batch_run):
    b, views = batch_run
    nf, used, ms = b.run()  # first run: autotuner times every shape
    assert used in (2, 4, 5, 6) and ms > 0
    chosen, per, row_tiles = b.tuning()
    assert chosen == used and per[2] > 0 and per[4] > 0 and per[5] > 0
    # each job's split follows the tuner's model: padded work of a tile /
    # that shape's measured throughput on the whole batch's padded work,
    # least in total over every possible split (brute force here)
    rate = {v: sum(_uniform_ops(spec, v) for spec in JOBS) / per[v] for v in (2, 4, 5)}
    for x, spec in enumerate(JOBS):
        rows = spec[2] or spec[1]
        n64, n32, n16 = (int(t) for t in row_tiles[x])
        assert 64 * n64 + 32 * n32 + 16 * n16 >= rows, (x, row_tiles[x])
        # the split lays the tallest tiles first
        r, est = 0, 0.0
        for v, n in ((2, n64), (4, n32), (5, n16)):
            for _ in range(n):
                est += _tile_ops(spec, v, r) / rate[v]
                r += TILE_ROWS[v]
        want = _best_split_time(spec, rate)
        assert est == pytest.approx(want, rel=1e-9), (x, row_tiles[x], est, want)
    assert (nf == 0).all()
    for x, spec in enumerate(JOBS):
        d, ref, qry = views[x]
        want, wnf = _single(d, ref, qry, spec[3], spec[4])
        got = b.result(x)
        assert wnf == 0
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"job {x} {spec}"


@pytest.mark.parametrize("variant", [2, 4, 5, 6])
def test_batch_each_shape_and_graph_replays(batch_run, variant):
    # plans 2 / 4: every job in one shape on its own kernel; 5: every job in
    # 16-row tiles and 6: each job's rows split over the three heights, on
    # the mixed-shape kernel
    # three replays of the captured graph, with other device work between
    # them, all equal to the single calls
    b, views = batch_run
    ref0 = [None] * len(JOBS)
    for rep in range(3):
        nf, used, _ = b.run(variant)
        assert used == variant and (nf == 0).all()
        for x in range(len(JOBS)):
            got = b.result(x)
            if ref0[x] is None:
                d, ref, qry = views[x]
                ref0[x], _ = _single(d, ref, qry, JOBS[x][3], JOBS[x][4])
            assert np.array_equal(got.view(np.uint8), ref0[x].view(np.uint8)), (rep, x)
        d, ref, qry = views[0]
        H.dense(ref, qry)  # eager work between replays


def test_batch_against_oracle(batch_run):
    b, _ = batch_run
    b.run()
    for x in ORACLE_JOBS:
        ci, nr, nq, rnd, kind = JOBS[x]
        d = batch_run[1][x][0]
        kl, s64 = d["kl"], d["s64"]
        if rnd == "mc":
            random = O.Random(2, True, d["table"], d["lt"])
        elif rnd == "equal":
            random = O.Random(1, True)
        else:
            random = O.Random(0)
        q = dict(query=d["qraw"][:nq], query_len=d["qle"][:nq],
                 query_cluster=d["qcl"][:nq]) if nq else {}
        want, counts, nf, _ = O.query_db(d["raw"][:nr], kl, s64, 14, d["le"][:nr],
                                         ref_cluster=d["rcl"][:nr], random=random,
                                         jaccard=(kind == "jaccard"), want_counts=True, **q)
        got = b.result(x)
        assert nf == 0
        if kind == "counts":
            assert np.array_equal(got, counts), x
        elif kind == "jaccard":
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), x
        else:
            assert np.abs(got - want).max() <= TOL, x


def test_batch_rejects_bad_jobs(dbs):
    d0, d1 = dbs[0], dbs[1]
    with pytest.raises(H.PpskError, match="Mismatching"):
        H.Batch([dict(ref=d0["ref"], query=d1["qry"])])
    with pytest.raises(H.PpskError, match="lengths"):
        # k >= 33 class has no lengths: equal-base is refused
        H.Batch([dict(ref=dbs[3]["ref"], query=dbs[3]["qry"], random_mode=H.RANDOM_EQUAL_BASE)])
