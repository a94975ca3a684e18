"""Batched jobs at the extreme shapes of the uneven-jobs workload.

The uneven-jobs workload (256 jobs) draws queries from 1 to 2,048, references
from 500 to 20,000, 4 to 14 k-mer lengths within k = 13..39 and 1,000 to
20,000 bins (sketchsize64 16 to 313).  These jobs sit at those ends, in one
persistent launch:

* 14 k-mer lengths (13..39, k >= 33 included) at sketchsize64 = 313 against
  20,000 references, with no correction and with a Monte Carlo table (zero
  from its k_max on, as the reference's random_match returns for
  k >= max_k);
* 2,048 queries at 12 k-mer lengths and sketchsize64 = 250 with the
  equal-base correction;
* the smallest class (4 k-mer lengths, sketchsize64 = 16) with one query.

sketchsize64 >= 256 is where the reference's calc_intersize computes a
nonzero expected_samebits ((64 * s64) >> bbits >= 1, src/dist/dist.cpp:70-77)
and takes its early return, so the count path of that branch runs on the GPU
here.  Every job equals its single ppsk_dense call bit for bit, and a row
sample of each is checked against the oracle: counts bit-exact, Jaccards
bit-exact, core/accessory within 1e-6.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import mc_table

pytestmark = pytest.mark.gpu

TOL = 1e-6

K14 = list(range(13, 40, 2))          # 14 lengths, 13..39
K12 = list(range(13, 25))              # 12 lengths, 13..24 (equal-base defined)
K4 = [15, 19, 23, 27]

# name: klist, s64, references, queries
CLASSES = {
    "k14_s313": (K14, 313, 20000, 72),
    "k12_s250": (K12, 250, 1000, 2048),
    "k4_s16": (K4, 16, 500, 1),
}
# class, refs, queries, random correction, output, oracle row sample
JOBS = [
    ("k14_s313", 20000, 72, "none", "counts", [0, 31, 71]),
    ("k14_s313", 20000, 40, "mc", "core_acc", [0, 17, 39]),
    ("k14_s313", 20000, 9, "mc", "jaccard", [2, 8]),
    ("k12_s250", 1000, 2048, "equal", "core_acc", [0, 700, 1500, 2047]),
    ("k12_s250", 1000, 2048, "none", "counts", [5, 2046]),
    ("k4_s16", 500, 1, "none", "core_acc", [0]),
]


def _mode(name):
    return {"none": H.RANDOM_NONE, "equal": H.RANDOM_EQUAL_BASE, "mc": H.RANDOM_MC}[name]


@pytest.fixture(scope="module")
def dbs():
    out = {}
    for ci, (name, (kl, s64, nr, nq)) in enumerate(CLASSES.items()):
        allr, alle, _ = H.synth(nr + nq, kl, s64, seed=700 + ci)
        raw, le = allr[:nr], alle[:nr]
        qraw, qle = allr[nr:], alle[nr:]
        can_len = max(kl) < 33
        rcl = (np.arange(nr) * 7 % 3).astype(np.uint16)
        qcl = (np.arange(nq) % 3).astype(np.uint16)
        ref = H.SketchSet(raw, kl, s64, lengths=le if can_len else None, clusters=rcl)
        qry = H.SketchSet(qraw, kl, s64, lengths=qle if can_len else None, clusters=qcl)
        # Monte Carlo table with k_max = 31: zero for k >= 31
        table, lt = mc_table(kl, symmetric=False, seed=ci, zero_from=31)
        out[name] = dict(kl=kl, s64=s64, raw=raw, le=le, qraw=qraw, qle=qle, rcl=rcl,
                         qcl=qcl, ref=ref, qry=qry, table=table, lt=lt)
    yield out
    for d in out.values():
        d["qry"].free()
        d["ref"].free()


@pytest.fixture(scope="module")
def batch_run(dbs):
    views, jobs = [], []
    for name, nr, nq, rnd, kind, _ in JOBS:
        d = dbs[name]
        rv, qv = d["ref"].prefix(nr), d["qry"].prefix(nq)
        views.append((d, rv, qv))
        jobs.append(dict(ref=rv, query=qv, random_mode=_mode(rnd),
                         table=d["table"] if rnd == "mc" else None, out=kind))
    b = H.Batch(jobs)
    nf, used, ms = b.run()
    assert (nf == 0).all() and ms > 0
    yield b, views
    b.free()
    for _, rv, qv in views:
        rv.free()
        qv.free()


def test_extreme_jobs_equal_single_calls(batch_run):
    b, views = batch_run
    for x, (name, nr, nq, rnd, kind, _) in enumerate(JOBS):
        d, rv, qv = views[x]
        want, nf, _ = H.dense(rv, qv, random_mode=_mode(rnd),
                              table=d["table"] if rnd == "mc" else None, out=kind)
        assert nf == 0
        got = b.result(x)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (x, name, rnd, kind)


def _oracle_random(rnd, d):
    if rnd == "mc":
        return O.Random(2, True, d["table"], d["lt"])
    if rnd == "equal":
        return O.Random(1, True)
    return O.Random(0)


@pytest.mark.parametrize("x", range(len(JOBS)))
def test_extreme_jobs_row_sample_against_oracle(batch_run, x):
    b, views = batch_run
    name, nr, nq, rnd, kind, rows = JOBS[x]
    d = views[x][0]
    kl, s64 = d["kl"], d["s64"]
    rows = np.asarray(rows)
    want, counts, nf, _ = O.query_db(d["raw"][:nr], kl, s64, 14, d["le"][:nr],
                                     query=d["qraw"][rows], query_len=d["qle"][rows],
                                     ref_cluster=d["rcl"][:nr], query_cluster=d["qcl"][rows],
                                     random=_oracle_random(rnd, d), jaccard=(kind == "jaccard"),
                                     want_counts=True)
    got = b.result(x)
    got = got.reshape(nq, nr, -1)[rows].reshape(len(rows) * nr, -1)
    assert nf == 0
    if kind == "counts":
        assert np.array_equal(got, counts)
        # not a constant array
        assert counts.max() > counts.min()
    elif kind == "jaccard":
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    else:
        assert np.abs(got - want).max() <= TOL
