"""GPU sketching (ppsk_sketch) against the oracle's restatement, which is
itself checked bit-exactly against the reference's compiled sketch core
(tests/test_oracle.py).

Words bit-exact per sample and k-mer length; the inverse-minhash sums and
densification flags equal.  Inputs cover lowercase, N and U characters,
several contigs per sample (k-mers never cross them), contigs shorter than
k, lengths around the per-thread work-unit edges, sketch sizes needing
densification, k above 31 and 33 (both rotation periods), forward-only
hashing, and sketch sizes on both sides of the LDS bin-minima limit.
"""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H

pytestmark = pytest.mark.gpu


def _rseq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(list(alphabet), size=n))


def _check(samples, klist, s64, use_rc=True):
    raw, mh, dn = H.sketch(samples, klist, s64, use_rc=use_rc)
    for s, contigs in enumerate(samples):
        want_mh = 0.0
        want_dn = 0
        for x, k in enumerate(klist):
            us, m, d = O.sketch(contigs, k, s64, use_rc)
            assert np.array_equal(raw[s, x], us), (s, k)
            want_mh += m
            want_dn |= int(d != 0)
        assert mh[s] == pytest.approx(want_mh, rel=0, abs=1e-15), s
        assert dn[s] == want_dn, s


@pytest.mark.parametrize("use_rc", [True, False])
def test_sketch_mixed_inputs(use_rc):
    rng = np.random.default_rng(1)
    samples = [
        [_rseq(rng, 30000)],
        [_rseq(rng, 5000, "ACGTacgtNnU"), _rseq(rng, 12), _rseq(rng, 9000)],
        [_rseq(rng, 1024 + 14), _rseq(rng, 2048 + 15), _rseq(rng, 1023 + 16)],  # unit edges
        [_rseq(rng, 10)],                      # shorter than every k: all bins empty
        [_rseq(rng, 400)],                     # densified
    ]
    _check(samples, [15, 17, 21, 31, 33, 35], 16, use_rc)


@pytest.mark.parametrize("s64", [2, 100, 156, 313])
def test_sketch_sizes(s64):
    rng = np.random.default_rng(s64)
    samples = [[_rseq(rng, 60000)], [_rseq(rng, 20000), _rseq(rng, 20000, "ACGTN")]]
    _check(samples, [15, 29], s64)


def test_sketch_tree_genomes_distance_sanity():
    # genomes evolved from one ancestor: sketches feed the distance path, and
    # closer genomes get a smaller core distance
    rng = np.random.default_rng(3)
    base = rng.choice(list("ACGT"), size=200000)
    genomes = []
    for rate in (0.0, 0.002, 0.01, 0.03):
        g = base.copy()
        m = rng.random(g.size) < rate
        g[m] = rng.choice(list("ACGT"), size=int(m.sum()))
        genomes.append(["".join(g)])
    kl = [15, 17, 19, 21, 23, 25, 27, 29]
    raw, _, _ = H.sketch(genomes, kl, 156)
    with H.SketchSet(raw, kl, 156) as s:
        d, nf, _ = H.dense(s)
    assert nf == 0
    core = d[:3, 0]  # (0,1), (0,2), (0,3)
    assert core[0] < core[1] < core[2]


def _check_codon(samples, klist, s64, use_rc=True):
    raw, mh, dn = H.sketch(samples, klist, s64, use_rc=use_rc, codon=True)
    for s, contigs in enumerate(samples):
        want_mh, want_dn = 0.0, 0
        for x, k in enumerate(klist):
            us, m, d = O.sketch_filtered(contigs, k, s64, use_rc, codon=True)
            assert np.array_equal(raw[s, x], us), (s, k)
            want_mh += m
            want_dn |= int(d != 0)
        assert mh[s] == pytest.approx(want_mh, rel=0, abs=1e-15), s
        assert dn[s] == want_dn, s


@pytest.mark.parametrize("use_rc", [True, False])
def test_codon_phased_sketch(use_rc):
    # codon-phased seeds: a span of 3k - 2 bases, every third one hashed;
    # three interleaved rolling states per lane
    rng = np.random.default_rng(21)
    samples = [
        [_rseq(rng, 30000)],
        [_rseq(rng, 6000, "ACGTacgtNnU"), _rseq(rng, 30), _rseq(rng, 9000)],
        [_rseq(rng, 1024 + 40), _rseq(rng, 2048 + 43), _rseq(rng, 1023 + 45)],  # unit edges
        [_rseq(rng, 50)],  # shorter than every span of k >= 17
    ]
    _check_codon(samples, [5, 9, 15, 17, 21], 16, use_rc)


def test_codon_phased_large_k_and_sketch():
    rng = np.random.default_rng(22)
    samples = [[_rseq(rng, 80000)], [_rseq(rng, 40000, "ACGTN")]]
    _check_codon(samples, [29, 41, 85], 156)
