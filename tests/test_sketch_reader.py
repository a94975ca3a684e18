"""The sequence reader behind constructDatabase against the reference's own
reader (src/sketch/seqio.cpp over kseq, compiled from its sources into
oracle/_ref): the kept sequences, composition, total and N counts and the
reads flag, on hand-written edge cases and on randomly generated FASTA/FASTQ
text (line endings, empty lines, lower case, N/U and other characters,
quality strings holding '@', '>' and '+', quality length mismatches,
records shorter than k, gzip).  The hand-written cases also carry their
expected sequences, so they pin the reader when oracle/_ref is absent."""
import gzip
import os

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


def _need_ref():
    if O.ref_sketch_lib() is None:
        pytest.skip("oracle/_ref/libref_sketch.so not built (needs the reference sources)")


def _write(path, text, gz=False):
    data = text.encode("latin-1") if isinstance(text, str) else text
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)
    return str(path)


def _compare_with_reference(pp, path, k):
    bases, total, n_count, reads, contigs = pp._readSequences([path], k)
    rb, rtot, rn, rreads, rnseq = O.ref_read_stats(path, k)
    assert contigs == O.ref_read_seqs(path, k), path
    assert len(contigs) == rnseq
    assert total == rtot and n_count == rn and reads == rreads, path
    # both divide the same integer counts by the same total
    np.testing.assert_array_equal(np.array(bases), rb)
    return contigs


HAND = [
    # (text, k, expected kept sequences, reads)
    (">a\nACGT\nacgt\n>b\nAC\n", 3, ["ACGTACGT"], False),
    (">a desc\r\nACGT\r\nNNuu\r\n", 4, ["ACGTNNUU"], False),
    ("junk\n>a\n\n\nACG\n\nTTT\n", 3, ["ACGTTT"], False),
    ("@r1\nACGTA\n+\nII@>I\n@r2\nGGG\n+r2\n+++\n", 3, ["ACGTA", "GGG"], True),
    # quality split over lines until it is as long as the sequence
    ("@r1\nACGTACGT\n+\nIIII\nIIII\n@r2\nCCCC\n+\nIIII\n", 4, ["ACGTACGT", "CCCC"], True),
    # a quality of another length ends the file
    ("@r1\nACGT\n+\nIII\n@r2\nGGGG\n+\nIIII\n", 4, [], False),
    # an empty record followed by '+': one quality line is read ('@' in it
    # does not start a record)
    ("@e\n+\n\n>a\nACGT\n", 4, ["ACGT"], False),
    ("@e\n+\n@x\n>a\nACGT\n", 4, [], False),
    # a lone '\r' line is kept (kseq drops '\r' only from longer strings)
    (">a\n\r\nACGT\n", 4, ["\rACGT"], False),
    (">a\nAC", 3, [], False),
]


@pytest.mark.parametrize("case", range(len(HAND)))
def test_hand_written_cases(pp, tmp_path, case):
    text, k, want, reads = HAND[case]
    path = _write(tmp_path / "in.fa", text)
    _, _, _, got_reads, contigs = pp._readSequences([path], k)
    assert contigs == want
    assert got_reads == reads
    if O.ref_sketch_lib() is not None:
        _compare_with_reference(pp, path, k)


def _random_text(rng):
    nl = lambda: "\r\n" if rng.random() < 0.3 else "\n"  # noqa: E731
    alphabet = list("ACGTACGTACGTacgtNnUuRY-*")
    qual = [chr(c) for c in range(33, 127)]
    out = []
    if rng.random() < 0.2:
        out.append("noise line" + nl())
    for _ in range(int(rng.integers(0, 8))):
        fastq = rng.random() < 0.4
        out.append(("@" if fastq else ">") + f"s{int(rng.integers(1000))}")
        if rng.random() < 0.3:
            out.append(" some comment")
        out.append(nl())
        seq = []
        for _ in range(int(rng.integers(0, 5))):
            if rng.random() < 0.15:
                out.append(nl())  # empty line
                continue
            line = "".join(rng.choice(alphabet, size=int(rng.integers(1, 60))))
            seq.append(line)
            out.append(line + nl())
        if fastq:
            out.append("+" + (" x" if rng.random() < 0.2 else "") + nl())
            n = sum(len(s) for s in seq)
            if rng.random() < 0.1:
                n += int(rng.integers(-2, 3))  # length mismatch
            q = "".join(rng.choice(qual, size=max(n, 0)))
            cut = int(rng.integers(0, len(q) + 1))
            for part in (q[:cut], q[cut:]):
                if part:
                    out.append(part + nl())
    text = "".join(out)
    if text and rng.random() < 0.2:
        text = text.rstrip("\n")  # no final newline
    return text


@pytest.mark.parametrize("seed", range(60))
def test_random_files_match_reference(pp, tmp_path, seed):
    _need_ref()
    rng = np.random.default_rng(seed)
    text = _random_text(rng)
    k = int(rng.choice([3, 5, 15, 31]))
    path = _write(tmp_path / ("in.fq.gz" if seed % 3 == 0 else "in.fq"), text, gz=seed % 3 == 0)
    _compare_with_reference(pp, path, k)


def test_composition_stops_after_one_million_bases(pp, tmp_path):
    _need_ref()
    rng = np.random.default_rng(5)
    a = "".join(rng.choice(list("ACGT"), size=700_000))
    b = "".join(rng.choice(list("GGGTN"), size=500_000))
    path = _write(tmp_path / "big.fa.gz", f">a\n{a}\n>b\n{b}\n", gz=True)
    contigs = _compare_with_reference(pp, path, 21)
    assert [len(c) for c in contigs] == [700_000, 500_000]
    bases, total, n_count, _, _ = pp._readSequences([path], 21)
    assert total == 1_200_000 and n_count == b.count("N")
    # only the first 999,999 bases are counted, over 1,000,000
    head = (a + b)[:999_999]
    assert bases[2] == head.count("G") / 1_000_000


def test_several_files_read_as_one_sample(pp, tmp_path):
    p1 = _write(tmp_path / "a.fa", ">x\nACGTACGTAA\n>y\nAC\n")
    p2 = _write(tmp_path / "b.fa.gz", ">z\nggggccccNN\n", gz=True)
    both = _write(tmp_path / "ab.fa", ">x\nACGTACGTAA\n>y\nAC\n>z\nggggccccNN\n")
    one = pp._readSequences([p1, p2], 5)
    cat = pp._readSequences([both], 5)
    assert one == cat
    assert one[4] == ["ACGTACGTAA", "GGGGCCCCNN"] and one[2] == 2


def test_missing_file_raises(pp, tmp_path):
    with pytest.raises(RuntimeError, match="Could not open"):
        pp._readSequences([str(tmp_path / "absent.fa")], 15)


def test_reader_feeds_the_reference_sketch(pp, tmp_path):
    # the oracle's sketch of what this reader keeps equals the reference's
    # reader + sketch() on the same file: what constructDatabase will store
    _need_ref()
    from seqfiles import samples
    _, files = samples(tmp_path, seed=2)
    kl = [15, 21, 29]
    for fl in files:
        if len(fl) != 1:
            continue
        _, _, _, _, contigs = pp._readSequences(fl, kl[-1])
        for k in kl:
            us, mh, dn = O.sketch(contigs, k, 16, True)
            rus, rmh, rdn = O.ref_sketch_file(fl[0], kl[-1], k, 16, True)
            assert np.array_equal(us, rus) and mh == rmh and (dn != 0) == (rdn != 0), (fl, k)
