"""The reads count filter of the sketcher (min_count > 0 on FASTQ input):
the oracle's restatement (count-min sketch or exact hash counter, counted in
sequence order before binning) against the reference's own compiled
sketch() over its reader (oracle/_ref/libref_sketch.so, built from
sketch/sketch.cpp, seqio.cpp and countmin.cpp)."""
import numpy as np
import pytest

import oracle as O
from seqfiles import fastq_reads, put


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


def _need_ref():
    if O.ref_sketch_lib() is None:
        pytest.skip("oracle/_ref/libref_sketch.so not built (needs the reference sources)")


@pytest.fixture(scope="module")
def reads_file(tmp_path_factory):
    rng = np.random.default_rng(11)
    genome = "".join(rng.choice(list("ACGT"), size=20000))
    path = tmp_path_factory.mktemp("reads") / "r.fq.gz"
    return put(path, fastq_reads(rng, genome, 1600, error=0.01, n_rate=0.001), gz=True)


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("min_count,k", [(1, 15), (2, 21), (3, 17), (5, 31)])
def test_filtered_sketch_matches_reference(pp, reads_file, exact, min_count, k):
    _need_ref()
    _, _, _, reads, contigs = pp._readSequences([reads_file], k)
    assert reads
    us, mh, dn = O.sketch_filtered(contigs, k, 32, True, min_count, exact)
    rus, rmh, rdn = O.ref_sketch_file(reads_file, k, k, 32, True, min_count, exact)
    assert np.array_equal(us, rus)
    assert mh == rmh and (dn != 0) == (rdn != 0)


def test_filter_removes_error_kmers(reads_file, pp):
    # at 8x coverage with 1% errors, requiring 3 copies drops the error k-mers:
    # the filtered sketch lies closer to the genome's own sketch
    _, _, _, _, contigs = pp._readSequences([reads_file], 21)
    rng = np.random.default_rng(11)
    genome = "".join(rng.choice(list("ACGT"), size=20000))
    g, _, _ = O.sketch([genome], 21, 32, True)
    raw, _, _ = O.sketch_filtered(contigs, 21, 32, True, 0, False)
    filt, _, _ = O.sketch_filtered(contigs, 21, 32, True, 3, False)
    assert O.calc_intersize(filt, g, 32, 14) > O.calc_intersize(raw, g, 32, 14)


def test_min_count_zero_is_the_plain_sketch(reads_file, pp):
    _, _, _, _, contigs = pp._readSequences([reads_file], 15)
    a = O.sketch(contigs, 15, 16, True)
    b = O.sketch_filtered(contigs, 15, 16, True, 0, False)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]


@pytest.mark.parametrize("exact", [False, True])
def test_codon_phased_filtered_sketch_matches_reference(pp, reads_file, exact):
    # spaced seeds recompute the second hash at every k-mer
    _need_ref()
    k = 9  # span 25
    _, _, _, _, contigs = pp._readSequences([reads_file], k)
    us, mh, _ = O.sketch_filtered(contigs, k, 32, True, 2, exact, codon=True)
    rus, rmh, _ = O.ref_sketch_file(reads_file, k, k, 32, True, 2, exact, codon=True)
    assert np.array_equal(us, rus) and mh == rmh
