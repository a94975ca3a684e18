"""Read sketching with the count filter on the GPU (ppsk_sketch_ex,
constructDatabase with min_count > 0) against the oracle's restatement, which
is itself pinned to the reference's compiled sketch() over FASTQ files
(tests/test_reads_filter.py).

Words bit-exact per sample and k; inverse-minhash sums and densification
flags equal.  Inputs: reads sampled from 20-40 kb genomes at 5-10x coverage
with substitution errors and N calls (every N restarts the hash iterator,
which re-seeds the count-min sketch's second hash), several samples per
call, both strands and forward-only hashing, the count-min sketch and the
exact counter, min_count from 1 (the unfiltered sketch) to 255 (nothing
passes: every bin empty)."""
import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from seqfiles import fastq_reads, put

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


def _reads(seed, length, n_reads, n_rate):
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), size=length))
    text = fastq_reads(rng, genome, n_reads, error=0.01, n_rate=n_rate)
    return [line for i, line in enumerate(text.split("\n")) if i % 4 == 1]


def _check(samples, klist, s64, min_count, exact, use_rc=True):
    raw, mh, dn = H.sketch(samples, klist, s64, use_rc=use_rc, min_count=min_count, exact=exact)
    for s, contigs in enumerate(samples):
        want_mh, want_dn = 0.0, 0
        for x, k in enumerate(klist):
            us, m, d = O.sketch_filtered(contigs, k, s64, use_rc, min_count, exact)
            assert np.array_equal(raw[s, x], us), (s, k, min_count, exact)
            want_mh += m
            want_dn |= int(d != 0)
        assert mh[s] == want_mh, s
        assert dn[s] == want_dn, s


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("min_count", [1, 2, 3, 7])
def test_filtered_sketch_matches_oracle(exact, min_count):
    samples = [_reads(1, 20000, 1500, 0.0), _reads(2, 30000, 1800, 0.01),
               _reads(3, 12000, 900, 0.05)]
    _check(samples, [15, 21, 31], 16, min_count, exact)


def test_filtered_forward_strand_and_large_sketch():
    samples = [_reads(4, 40000, 3000, 0.002)]
    _check(samples, [17, 25], 156, 3, False, use_rc=False)
    _check(samples, [17], 156, 2, True, use_rc=True)


def test_nothing_passes_at_255():
    samples = [_reads(5, 20000, 400, 0.0)]
    raw, mh, dn = H.sketch(samples, [21], 8, min_count=255)
    us, m, d = O.sketch_filtered(samples[0], 21, 8, True, 255, False)
    assert d == -1 and np.array_equal(raw[0, 0], us)


def test_construct_database_from_reads(pp, tmp_path):
    rng = np.random.default_rng(6)
    genome = "".join(rng.choice(list("ACGT"), size=30000))
    fq = put(tmp_path / "reads.fq.gz", fastq_reads(rng, genome, 2500, error=0.01, n_rate=0.002),
             gz=True)
    fa = put(tmp_path / "asm.fa", f">c\n{genome}\n")
    kl = [15, 19, 23]
    for exact in (False, True):
        db = str(tmp_path / f"reads_{int(exact)}")
        pp.constructDatabase(db, ["reads", "asm"], [[fq], [fa]], kl, 16, calc_random=False,
                             min_count=3, exact=exact)
        info = pp.getSketch(db, "reads")
        _, _, _, reads, contigs = pp._readSequences([fq], kl[-1])
        assert reads and info["reads"]
        mh_sum = 0.0
        for k in kl:
            us, m, _ = O.sketch_filtered(contigs, k, 16, True, 3, exact)
            assert np.array_equal(info["sketch"][k], us), (exact, k)
            mh_sum += m
            if O.ref_sketch_lib() is not None:
                ref = O.ref_sketch_file(fq, kl[-1], k, 16, True, 3, exact)
                assert np.array_equal(info["sketch"][k], ref[0]), (exact, k)
        assert info["length"] == int(len(kl) / mh_sum)
        # the assembly is not reads: sketched without the filter
        asm = pp.getSketch(db, "asm")
        assert not asm["reads"]
        for k in kl:
            assert np.array_equal(asm["sketch"][k], O.sketch([genome], k, 16, True)[0])


@pytest.mark.parametrize("exact", [False, True])
def test_codon_phased_filtered_sketch(exact):
    # spaced seeds: the count-min sketch's second hash is the k-mer's own
    samples = [_reads(7, 20000, 1500, 0.01), _reads(8, 15000, 1200, 0.0)]
    raw, mh, dn = H.sketch(samples, [9, 13], 16, min_count=2, exact=exact, codon=True)
    for s, contigs in enumerate(samples):
        want_mh = 0.0
        for x, k in enumerate([9, 13]):
            us, m, _ = O.sketch_filtered(contigs, k, 16, True, 2, exact, codon=True)
            assert np.array_equal(raw[s, x], us), (s, k)
            want_mh += m
        assert mh[s] == want_mh
