"""constructDatabase end to end on the GPU: sequence files -> sketches in the
reference's HDF5 layout -> queryDatabase / querySelfSparse.

Sketch words are compared bit-exactly with the oracle's sketch of the
sequences the reader keeps, and, where oracle/_ref is built, with the
reference's own compiled reader + sketch() on the same file.  Sample
attributes (length, missing_bases, reads, base_freq) follow the reference
Reference constructor (src/reference.cpp:23-56).  Inputs: genomes evolved
from one ancestor (SNPs and accessory blocks), written as plain, gzipped,
wrapped, CRLF, lower-case and N-holding FASTA, split over two files, and as
FASTQ reads (the reads path sizes the genome from the bin-0 minhash)."""
import os

import numpy as np
import pytest

import oracle as O
from seqfiles import put as _put, samples as _samples

pytestmark = pytest.mark.gpu

KL = [15, 17, 19, 21, 25, 29]
S64 = 16
TOL = 1e-6


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


def test_construct_database_sketches_and_attributes(pp, tmp_path):
    names, files = _samples(tmp_path)
    db = str(tmp_path / "built")
    pp.constructDatabase(db, names, files, KL, S64, calc_random=False, num_threads=4)
    assert sorted(pp.getSampleNames(db)) == sorted(names)
    have_ref = O.ref_sketch_lib() is not None
    for name, fl in zip(names, files):
        info = pp.getSketch(db, name)
        bases, total, n_count, reads, contigs = pp._readSequences(fl, KL[-1])
        assert info["kmers"] == KL and info["sketchsize64"] == S64 and info["bbits"] == 14
        mh_sum = 0.0
        for k in KL:
            words, mh, _ = O.sketch(contigs, k, S64, True)
            assert np.array_equal(info["sketch"][k], words), (name, k)
            mh_sum += mh
            if have_ref and len(fl) == 1:
                ref = O.ref_sketch_file(fl[0], KL[-1], k, S64, True)
                assert np.array_equal(info["sketch"][k], ref[0]), (name, k)
        assert info["reads"] == reads == (name == "sample_4")
        assert info["missing_bases"] == n_count
        assert info["base_freq"] == bases
        want_len = int(len(KL) / mh_sum) if reads else total
        assert info["length"] == want_len, name
        if have_ref and len(fl) == 1:
            rb, rtot, rn, rreads, _ = O.ref_read_stats(fl[0], KL[-1])
            assert info["base_freq"] == rb.tolist() and info["missing_bases"] == rn
            if not rreads:
                assert info["length"] == rtot
    # the short contig of sample_1 was dropped; sample_2 holds N runs
    assert pp.getSketch(db, "sample_2")["missing_bases"] > 0


def _stored(pp, db, names):
    raw, lens = [], []
    for n in names:
        info = pp.getSketch(db, n)
        raw.append([info["sketch"][k] for k in info["kmers"]])
        lens.append(info["length"])
    return np.array(raw, np.uint64), np.array(lens, np.uint64)


def test_constructed_database_queries(pp, tmp_path):
    names, files = _samples(tmp_path, seed=8)
    db = str(tmp_path / "ref")
    pp.constructDatabase(db, names[:4], files[:4], KL, S64, calc_random=False)
    qdb = str(tmp_path / "query")
    pp.constructDatabase(qdb, names[4:], files[4:], KL, S64, calc_random=False)
    raw, lens = _stored(pp, db, names[:4])
    d = pp.queryDatabase(db, db, names[:4], names[:4], KL)
    want, _, nf, _ = O.query_db(raw, KL, S64, 14, lens, random=O.Random(1, True))
    assert nf == 0 and np.abs(d - want).max() <= TOL
    qraw, qlens = _stored(pp, qdb, names[4:])
    dq = pp.queryDatabase(db, qdb, names[:4], names[4:], KL)
    wq, _, _, _ = O.query_db(raw, KL, S64, 14, lens, query=qraw, query_len=qlens,
                             random=O.Random(1, True))
    assert np.abs(dq - wq).max() <= TOL
    i, j, ds = pp.querySelfSparse(db, names[:4], KL, kNN=2)
    ri, rj, rd, _ = O.query_db_sparse(raw, KL, S64, 14, lens, 2, random=O.Random(1, True))
    assert i == ri.tolist() and j == rj.tolist()
    assert np.abs(np.array(ds) - rd).max() <= TOL


def test_existing_database_is_reused(pp, tmp_path):
    names, files = _samples(tmp_path, seed=9)
    db = str(tmp_path / "again")
    pp.constructDatabase(db, names[:2], files[:2], KL, S64, calc_random=False)
    before = os.stat(db + ".h5").st_mtime_ns
    pp.constructDatabase(db, names[:2], files[:2], KL, S64, calc_random=False)
    assert os.stat(db + ".h5").st_mtime_ns == before


def test_construct_database_errors(pp, tmp_path):
    empty = _put(tmp_path / "empty.fa", ">x\nACG\n")
    with pytest.raises(RuntimeError, match="Errors during sketching"):
        pp.constructDatabase(str(tmp_path / "e"), ["x"], [[empty]], KL, S64)
    with pytest.raises(RuntimeError, match="Minimum k"):
        pp.constructDatabase(str(tmp_path / "k"), ["x"], [[empty]], [2, 15], S64)


def test_construct_database_codon_phased(pp, tmp_path):
    names, files = _samples(tmp_path, seed=12)
    names, files = names[:3], files[:3]
    db = str(tmp_path / "codon")
    kl = [9, 11, 13]
    pp.constructDatabase(db, names, files, kl, S64, codon_phased=True, calc_random=True)
    have_ref = O.ref_sketch_lib() is not None
    for name, fl in zip(names, files):
        info = pp.getSketch(db, name)
        _, _, _, _, contigs = pp._readSequences(fl, kl[-1])
        for k in kl:
            us, _, _ = O.sketch_filtered(contigs, k, S64, True, codon=True)
            assert np.array_equal(info["sketch"][k], us), (name, k)
            if have_ref and len(fl) == 1:
                ref = O.ref_sketch_file(fl[0], kl[-1], k, S64, True, codon=True)
                assert np.array_equal(info["sketch"][k], ref[0]), (name, k)
    assert pp._readRandom(db)["present"]
    d = pp.queryDatabase(db, db, names, names, kl)
    assert d.shape == (3, 2) and np.isfinite(d).all()
