"""constructDatabase -> queryDatabase at the shape of the reference's own CPU
workload: 64 genomes of 2 Mb, k = 15, 17, ..., 29 (8 lengths), 10,000 bins
(sketchsize64 156), bbits 14, canonical strand, all-vs-all core/accessory
over the 2,016 upper-triangle pairs.

Genomes: leaves of a seeded random binary tree over one i.i.d. ACGT ancestor
(50% GC).  Every branch applies SNPs at a rate drawn from 0.02-0.8% (a leaf
lies 0.1-5% from the root) and flips the presence of each 1 kb block of a
400-block accessory pool with probability 0-5% (0-20% accessory sequence).
Each genome is one FASTA file.

Checks, all against the oracle: the sketch words of 6 genomes at every k
(bit-exact), the length attribute of every genome, the 2,016 core/accessory
distances (within 1e-6; equal-base correction, as a database built without
random tables is queried), the per-k Jaccards (bit-exact) and the kNN lists
of querySelfSparse (same i, j, order)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

KL = list(range(15, 30, 2))
S64 = 156
N, LENGTH = 64, 2_000_000
TOL = 1e-6
CHECKED = [0, 17, 31, 32, 45, 63]


def _tree_genomes(rng, n_leaves, length, n_blocks=400, block=1000):
    anc = rng.integers(0, 4, size=length, dtype=np.uint8)
    pool = rng.integers(0, 4, size=(n_blocks, block), dtype=np.uint8)
    leaves = []
    # depth-first over the tree with an explicit stack: (core, present, leaves below)
    stack = [(anc, rng.random(n_blocks) < 0.5, n_leaves)]
    while stack:
        core, present, k = stack.pop()
        if k == 1:
            leaves.append(np.concatenate([core, pool[present].reshape(-1)]))
            continue
        left = int(rng.integers(1, k))
        for part in (k - left, left):
            c = core.copy()
            m = rng.random(length) < rng.uniform(0.0002, 0.008)
            c[m] = (c[m] + rng.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) % 4
            p = present ^ (rng.random(n_blocks) < rng.uniform(0.0, 0.05))
            stack.append((c, p, part))
    return leaves


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    import pp_sketchlib as pp
    work = tmp_path_factory.mktemp("genomes")
    g = _tree_genomes(np.random.default_rng(23), N, LENGTH)
    letters = np.frombuffer(b"ACGT", np.uint8)
    names, files = [], []
    for i, codes in enumerate(g):
        p = work / f"g{i}.fa"
        p.write_bytes(b">contig1\n" + letters[codes].tobytes() + b"\n")
        names.append(f"genome_{i}")
        files.append([str(p)])
    db = str(work / "db")
    pp.constructDatabase(db, names, files, KL, S64, calc_random=False, num_threads=16)
    raw = np.array([[pp.getSketch(db, nm)["sketch"][k] for k in KL] for nm in names], np.uint64)
    lens = np.array([pp.getSketch(db, nm)["length"] for nm in names], np.uint64)
    return pp, db, names, g, raw, lens


def test_sketch_words_and_lengths(built):
    pp, db, names, g, raw, lens = built
    assert raw.shape == (N, len(KL), S64 * 14)
    letters = np.frombuffer(b"ACGT", np.uint8)
    for i in CHECKED:
        seq = letters[g[i]].tobytes().decode()
        for x, k in enumerate(KL):
            words, _, _ = O.sketch([seq], k, S64, True)
            assert np.array_equal(raw[i, x], words), (i, k)
    assert lens.tolist() == [x.size for x in g]


def test_all_vs_all_core_accessory(built):
    pp, db, names, g, raw, lens = built
    d = pp.queryDatabase(db, db, names, names, KL)
    want, counts, nf, _ = O.query_db(raw, KL, S64, 14, lens, random=O.Random(1, True),
                                     want_counts=True)
    assert nf == 0 and d.shape == (N * (N - 1) // 2, 2)
    assert np.abs(d - want).max() <= TOL
    assert want[:, 0].std() > 0 and want[:, 1].std() > 0
    dj = pp.queryDatabase(db, db, names, names, KL, jaccard=True)
    wj, _, _, _ = O.query_db(raw, KL, S64, 14, lens, random=O.Random(1, True), jaccard=True)
    assert np.array_equal(np.asarray(dj, np.float32).view(np.uint32), wj.view(np.uint32))


def test_knn_lists(built):
    pp, db, names, g, raw, lens = built
    i, j, ds = pp.querySelfSparse(db, names, KL, kNN=5)
    ri, rj, rd, _ = O.query_db_sparse(raw, KL, S64, 14, lens, 5, random=O.Random(1, True))
    assert i == ri.tolist() and j == rj.tolist()
    assert np.abs(np.array(ds) - rd).max() <= TOL
