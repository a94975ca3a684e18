"""Persistent batch kernel: many uneven query-vs-ref jobs in one launch.

Every job's output must equal the single-job ppsk_dense call on the same
sets (bit for bit: both run the same epilogue on the same counts), and a
subset is checked against the oracle (counts and Jaccards bit-exact,
core/accessory within 1e-6).  Jobs mix k-sets (including k >= 33, where only
no correction or a Monte Carlo table applies), sketch sizes, sample counts
from 1 query upwards, random modes and output kinds, and query prefixes of
resident databases.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import mc_table

pytestmark = pytest.mark.gpu

TOL = 1e-6

# klist, s64, reference DB size, query DB size
CLASSES = [
    ([15, 17, 19, 21], 16, 800, 70),
    ([13, 16, 19, 22, 25, 28], 47, 1500, 200),
    ([15, 17, 19, 21, 23, 25, 27, 29], 156, 600, 130),
    ([21, 24, 27, 30, 33, 36, 39], 30, 900, 40),
]
# class, refs, queries, random correction, output
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
]
ORACLE_JOBS = [0, 1, 3, 5, 6, 7, 9]


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
    return d, d["ref"].prefix(nr), d["qry"].prefix(nq)


def _single(d, ref, qry, rnd, kind):
    table = d["table"] if rnd == "mc" else None
    out, nf, _ = H.dense(ref, qry, random_mode=_mode(rnd), table=table, out=kind)
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


def _padded_ops(spec, v):
    """Matching ops of a job's tiles in shape v (64, 32 or 16 x 256), padding
    included: query-vs-ref jobs cover every (row tile, column tile)."""
    ci, nr, nq, _, _ = spec
    kl, s64 = CLASSES[ci][0], CLASSES[ci][1]
    tr = {2: 64, 4: 32, 5: 16}[v]
    tiles = -(-nq // tr) * -(-nr // 256)
    return tiles * tr * 256 * len(kl) * s64 * 30


def test_batch_autotuned_equals_single_calls(batch_run):
    b, views = batch_run
    nf, used, ms = b.run()  # first run: autotuner times every shape
    assert used in (2, 4, 5, 6) and ms > 0
    chosen, per, shapes = b.tuning()
    assert chosen == used and per[2] > 0 and per[4] > 0 and per[5] > 0
    # the per-job shapes follow the tuner's model: padded work of the job in a
    # shape / that shape's measured throughput on the whole batch's padded work
    ops = {v: [_padded_ops(spec, v) for spec in JOBS] for v in (2, 4, 5)}
    rate = {v: sum(ops[v]) / per[v] for v in (2, 4, 5)}
    for x in range(len(JOBS)):
        est = {v: ops[v][x] / rate[v] for v in (2, 4, 5)}
        best = min(est.values())
        assert est[int(shapes[x])] == best, (x, est, shapes[x])
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
    # 16-row tiles and 6: each job in its own shape, on the mixed-shape kernel
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
        want, counts, nf, _ = O.query_db(d["raw"][:nr], kl, s64, 14, d["le"][:nr],
                                         query=d["qraw"][:nq], query_len=d["qle"][:nq],
                                         ref_cluster=d["rcl"][:nr], query_cluster=d["qcl"][:nq],
                                         random=random, jaccard=(kind == "jaccard"),
                                         want_counts=True)
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
