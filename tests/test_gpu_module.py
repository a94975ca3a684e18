"""The drop-in pp_sketchlib module on the GPU against the oracle.

Inputs: the reference's own test databases (test/small_mass*.h5: real
sketches, old format without length/base_freq attributes, no /random group,
so the equal-base correction with the 3 Mb default length applies), and
synthetic databases written by the module (with and without a Monte Carlo
/random group).  The oracle reads the same sketch words from an independent
h5dump decoding (tests/golden/small_mass*.npz).
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import max_abs_diff

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DB = os.path.join(GOLD, "small_mass")
DB1 = os.path.join(GOLD, "small_mass_plus1")
KL = [14, 17, 20, 23, 26, 29]
TOL = 1e-6


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


@pytest.fixture(scope="module")
def fx():
    a = np.load(DB + ".npz", allow_pickle=False)
    b = np.load(DB1 + ".npz", allow_pickle=False)
    return a, b


def _len(n):
    return np.full(n, 3_000_000, np.uint64)


def test_self_small_mass(pp, fx):
    a, _ = fx
    names = a["names"].tolist()
    raw = a["sketches"]
    d = pp.queryDatabase(DB, DB, names, names, KL)
    ref, _, nf, _ = O.query_db(raw, KL, 156, 14, _len(6), random=O.Random(1, True))
    assert nf == 0 and d.shape == (15, 2) and d.dtype == np.float32
    assert np.abs(d - ref).max() <= TOL
    j = pp.queryDatabase(DB, DB, names, names, KL, jaccard=True)
    rj, _, _, _ = O.query_db(raw, KL, 156, 14, _len(6), random=O.Random(1, True), jaccard=True)
    assert np.array_equal(j.view(np.uint32), rj.view(np.uint32))
    j0 = pp.queryDatabase(DB, DB, names, names, KL, random_correct=False, jaccard=True)
    rj0, _, _, _ = O.query_db(raw, KL, 156, 14, _len(6), jaccard=True)
    assert np.array_equal(j0.view(np.uint32), rj0.view(np.uint32))
    # use_gpu=True takes the same path for core/accessory
    dg = pp.queryDatabase(DB, DB, names, names, KL, use_gpu=True)
    assert np.array_equal(dg, d)


def test_query_vs_ref_small_mass(pp, fx):
    a, b = fx
    names = a["names"].tolist()
    qn = ["6925_1_55"]
    assert b["names"].tolist()[-1] == qn[0]
    d = pp.queryDatabase(DB, DB1, names, qn, KL)
    ref, _, nf, _ = O.query_db(a["sketches"], KL, 156, 14, _len(6), query=b["sketches"][6:],
                               query_len=_len(1), random=O.Random(1, True))
    assert d.shape == (6, 2)
    assert np.abs(d - ref).max() <= TOL
    # sub-list of k-mer lengths: unrequested ones are dropped on load
    d2 = pp.queryDatabase(DB, DB1, names, qn, [17, 23, 29])
    sel = [1, 3, 5]
    ref2, _, _, _ = O.query_db(a["sketches"][:, sel], [17, 23, 29], 156, 14, _len(6),
                               query=b["sketches"][6:, sel], query_len=_len(1),
                               random=O.Random(1, True))
    assert np.abs(d2 - ref2).max() <= TOL
    # a prefix of the stored lengths: the loader keeps the longer ones
    # (src/api.cpp:471-478) and only the requested ones are used
    d3 = pp.queryDatabase(DB, DB1, names, qn, [14, 20])
    ref3, _, _, _ = O.query_db(a["sketches"][:, [0, 2]], [14, 20], 156, 14, _len(6),
                               query=b["sketches"][6:, [0, 2]], query_len=_len(1),
                               random=O.Random(1, True))
    assert np.abs(d3 - ref3).max() <= TOL


def test_sparse_small_mass(pp, fx):
    a, b = fx
    names = b["names"].tolist()
    raw = b["sketches"]
    for knn, col in ((1, 0), (3, 0), (3, 1), (7, 0)):
        i, j, d = pp.querySelfSparse(DB1, names, KL, kNN=knn, dist_col=col)
        ri, rj, rd, _ = O.query_db_sparse(raw, KL, 156, 14, _len(7), knn, dist_col=col,
                                          random=O.Random(1, True))
        assert isinstance(i, list)
        assert i == ri.tolist() and j == rj.tolist()
        assert max_abs_diff(d, rd) <= TOL
    i, j, d = pp.querySelfSparse(DB1, names, KL, jaccard=True, kNN=3, dist_col=1)
    ri, rj, rd, _ = O.query_db_sparse(raw, KL, 156, 14, _len(7), 3, dist_col=1,
                                      random=O.Random(1, True), jaccard=True)
    assert i == ri.tolist() and j == rj.tolist()
    assert np.array_equal(np.array(d, np.float32).view(np.uint32), rd.view(np.uint32))
    dense, _, _, _ = O.query_db(raw, KL, 156, 14, _len(7), random=O.Random(1, True))
    for cutoff, col in ((0.01, 0), (0.02, 0), (0.2, 1)):
        i, j, d = pp.querySelfSparse(DB1, names, KL, dist_cutoff=cutoff, dist_col=col)
        ri, rj, rd = O.sparsify_threshold(O.long_to_square(dense[:, col]), cutoff)
        assert i == ri.tolist() and j == rj.tolist()
        assert max_abs_diff(d, rd) <= TOL


def test_jaccard_dist(pp, fx):
    a, _ = fx
    names = a["names"].tolist()
    for k in (14, 20, 29):
        kk = KL.index(k)
        c = O.calc_intersize(a["sketches"][1, kk], a["sketches"][4, kk], 156, 14)
        want_raw = O.jaccard(c, 156, 0.0)
        got = pp.jaccardDist(DB, names[1], names[4], k)
        assert got == want_raw
        want = O.jaccard(c, 156, O.csrs(k, True, 3_000_000, 3_000_000))
        assert pp.jaccardDist(DB, names[1], names[4], k, random_correct=True) == want


def test_synthetic_monte_carlo_db(pp, tmp_path):
    import ppsk_hip as H
    db = str(tmp_path / "mc")
    n, kl, s64 = 150, [15, 17, 19, 21, 23, 25, 27, 29], 156
    pp.writeSyntheticDatabase(db, n, kl, sketch_size=s64 * 64, seed=21)
    raw, lengths, bases = H.synth(n, kl, s64, seed=21)
    names = [f"sample{i}" for i in range(n)]
    rng = np.random.default_rng(2)
    C = 3
    keys = list(range(13, 25))
    tab = rng.uniform(0, 0.02, size=(len(keys), C, C)).astype(np.float32)
    tab = ((tab + tab.transpose(0, 2, 1)) / 2).astype(np.float32)
    cent = (0.25 + rng.uniform(-0.02, 0.02, size=(C, 4))).astype(np.float32)
    # half of the samples carry a stored cluster; the rest are added by nearest centroid
    stored = {f"sample{i}": int(rng.integers(0, C)) for i in range(0, n, 2)}
    pp.writeRandomMatches(db, list(stored.keys()), list(stored.values()), keys, tab, cent,
                          13, 25, True)
    clusters = np.array([stored.get(nm, O.closest_cluster(bases[i], True, cent))
                         for i, nm in enumerate(names)], np.uint16)
    table = np.zeros((len(kl), C, C), np.float32)
    lt = np.array([k < 25 for k in kl])
    for x, k in enumerate(kl):
        if k < 25:
            table[x] = tab[keys.index(k)]
    rnd = O.Random(2, True, table, lt)
    d = pp.queryDatabase(db, db, names, names, kl)
    ref, _, nf, _ = O.query_db(raw, kl, s64, 14, lengths, ref_cluster=clusters, random=rnd)
    assert nf == 0
    assert np.abs(d - ref).max() <= TOL
    # query mode: queries always take the nearest centroid (api.cpp:243-244)
    qn = names[:20]
    rn = names[20:]
    dq = pp.queryDatabase(db, db, rn, qn, kl)
    qcl = np.array([O.closest_cluster(bases[i], True, cent) for i in range(20)], np.uint16)
    refq, _, _, _ = O.query_db(raw[20:], kl, s64, 14, lengths[20:], query=raw[:20],
                               query_len=lengths[:20], ref_cluster=clusters[20:],
                               query_cluster=qcl, random=rnd)
    assert np.abs(dq - refq).max() <= TOL
    i, j, dd = pp.querySelfSparse(db, names, kl, kNN=5)
    ri, rj, rd, _ = O.query_db_sparse(raw, kl, s64, 14, lengths, 5, ref_cluster=clusters,
                                      random=rnd)
    assert i == ri.tolist() and j == rj.tolist()
    assert max_abs_diff(dd, rd) <= TOL
