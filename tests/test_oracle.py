"""Pins the CPU oracle before it is trusted as the checker.

Each check is independent of the oracle's own formulas: the reference's golden
outputs (test/ppsketch_ref*.npy, copied to tests/golden), the reference's own
compiled calc_intersize (oracle/_ref, when the reference sources were present
at build time), brute force on unpacked bin values, closed forms and textbook
routines, and the exact arrays of the reference's test-matrix.py.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = [("ppsketch_ref", [15, 19, 23, 27]), ("ppsketch_ref_phased", [15, 19, 23, 27, 31])]


def _gold(name):
    return np.load(os.path.join(GOLD, name + ".npy"), allow_pickle=False)


def _rand_sketch(rng, n, s64=156, bbits=14, share=None):
    vals = rng.integers(0, 1 << bbits, size=(n, s64 * 64), dtype=np.uint32)
    if share is not None:
        m = rng.random(s64 * 64) < share
        vals[1:, m] = vals[0, m]
    return vals, O.pack_bins(vals, s64, bbits)


# --------------------------------------------------------------------------
# calc_intersize
# --------------------------------------------------------------------------
@pytest.mark.parametrize("share,s64", [(0.0, 156), (0.3, 156), (0.9, 156), (0.5, 16), (0.5, 313)])
def test_intersize_brute_force(share, s64):
    rng = np.random.default_rng(int(share * 100) + s64)
    vals, words = _rand_sketch(rng, 2, s64=s64, share=share)
    assert np.array_equal(O.unpack_bins(words, s64, 14), vals)
    expect = int((vals[0] == vals[1]).sum())
    assert O.calc_intersize(words[0], words[1], s64, 14) == expect


def test_intersize_identical_and_complement():
    rng = np.random.default_rng(3)
    _, words = _rand_sketch(rng, 1)
    assert O.calc_intersize(words[0], words[0], 156, 14) == 156 * 64
    # flipping any one layer bit of a bin breaks that bin's match
    w2 = words[0].copy()
    w2[5] ^= np.uint64(1 << 7)
    assert O.calc_intersize(words[0], w2, 156, 14) == 156 * 64 - 1


@pytest.mark.skipif(not O.ref_calc_intersize_available(),
                    reason="reference calc_intersize not built (reference sources absent)")
def test_intersize_matches_reference_build():
    rng = np.random.default_rng(17)
    for share in (0.0, 0.2, 0.7, 1.0):
        vals, words = _rand_sketch(rng, 2, share=share)
        assert O.calc_intersize(words[0], words[1], 156, 14) == \
            O.ref_calc_intersize(words[0], words[1], 156, 14)
    # the expected-bits branch (s64 >= 256 with bbits = 14)
    vals, words = _rand_sketch(rng, 2, s64=300, share=0.4)
    assert O.calc_intersize(words[0], words[1], 300, 14) == \
        O.ref_calc_intersize(words[0], words[1], 300, 14)


# --------------------------------------------------------------------------
# Jaccard and random correction against the goldens
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name,klist", SETS)
def test_raw_jaccard_golden_bit_exact(name, klist):
    raw = _gold(name + "_raw_jaccard")
    cnt = raw.astype(np.float64) * 9984
    assert np.abs(cnt - np.round(cnt)).max() < 1e-3
    cnt = np.round(cnt).astype(np.int64)
    got = np.array([[np.float32(O.jaccard(c, 156, 0.0)) for c in row] for row in cnt])
    assert np.array_equal(got.view(np.uint32), raw.view(np.uint32))


def _groups(values, gap):
    order = np.argsort(values)
    groups, cur = [], [order[0]]
    for a, b in zip(order[:-1], order[1:]):
        if values[b] - values[a] > gap:
            groups.append(cur)
            cur = [b]
        else:
            cur.append(b)
    groups.append(cur)
    return groups


def _float_r_reproducing(counts, targets, fn, lo, hi):
    lo32, hi32 = np.float32(lo), np.float32(hi)
    cands = np.arange(lo32.view(np.int32), hi32.view(np.int32) + 1, dtype=np.int32).view(np.float32)
    return [c for c in cands
            if all(np.float32(fn(int(k), float(c))) == t for k, t in zip(counts, targets))]


def test_random_correction_golden():
    """k = 15 of ppsketch_ref: the corrected Jaccards group into the few values
    of a Monte Carlo table; within every group one float32 r reproduces all
    Jaccards bit-exactly through observed_excess.  A dropped 1/(1-r) does not."""
    raw = _gold("ppsketch_ref_raw_jaccard")
    jac = _gold("ppsketch_ref_jaccard")
    cnt = np.round(raw.astype(np.float64) * 9984).astype(np.int64)[:, 0]
    j = jac[:, 0]
    changed = np.where(j != raw[:, 0])[0]
    assert len(changed) > 300
    obs = cnt[changed] / 9984.0
    jd = j[changed].astype(np.float64)
    rhat = (obs - jd) / (1 - jd)
    groups = [g for g in _groups(rhat, 2e-6) if len(g) >= 2]
    assert 3 <= len(groups) <= 9
    right = lambda c, r: O.jaccard(c, 156, r)
    wrong = lambda c, r: max(c / 9984.0 - r, 0.0)
    for g in groups:
        rows = changed[g]
        lo, hi = rhat[g].min() - 3e-6, rhat[g].max() + 3e-6
        assert _float_r_reproducing(cnt[rows], j[rows], right, lo, hi)
        if len(g) >= 20:
            assert not _float_r_reproducing(cnt[rows], j[rows], wrong, lo, hi)


# --------------------------------------------------------------------------
# Regression
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name,klist", SETS)
def test_fit_golden(name, klist):
    jac = _gold(name + "_jaccard")
    ref = _gold(name)
    tol = 5.0 / (156 * 64)
    got = np.array([O.fit_slope(klist, row, tol)[:2] for row in jac], np.float32)
    assert np.abs(got - ref).max() <= 2e-7


def test_fit_closed_form_and_window():
    rng = np.random.default_rng(0)
    k = np.array([13, 15, 17, 19, 21, 23, 25, 27], np.float32)
    for _ in range(50):
        core = rng.uniform(0, 0.08)
        acc = rng.uniform(0, 0.4)
        noise = rng.normal(0, 0.01, size=len(k))
        jac = ((1 - acc) * (1 - core) ** k * np.exp(noise)).astype(np.float32)
        c, a, ok = O.fit_slope(k, jac, 5.0 / 9984)
        slope, icpt = np.polyfit(k.astype(np.float64), np.log(jac.astype(np.float64)), 1)
        assert ok
        assert abs(c - (1 - np.exp(slope) if slope < 0 else 0)) < 1e-6
        assert abs(a - (1 - np.exp(icpt) if icpt < 0 else 0)) < 1e-6
    tol = 5.0 / 9984
    j = np.array([1e-6, 0.5, 1e-6, 0.3, 0.2, 1e-6, 1e-7, 0.0], np.float32)
    # window is [1, 4]: leading/trailing invalid dropped, interior invalid kept
    c, a, ok = O.fit_slope(k, j, tol)
    s, b = np.polyfit(k[1:5].astype(np.float64), np.log(j[1:5].astype(np.float64)), 1)
    want_c = 1 - np.exp(s) if s < 0 else 0.0
    want_a = 1 - np.exp(b) if b < 0 else 0.0
    assert ok and abs(c - want_c) < 1e-6 and abs(a - want_a) < 1e-6
    # a window that drops the interior point would fit a different line
    s2, _ = np.polyfit(k[[1, 3, 4]].astype(np.float64), np.log(j[[1, 3, 4]].astype(np.float64)), 1)
    assert abs(s2 - s) > 0.1
    # fewer than two valid points: failure
    assert not O.fit_slope(k, np.array([0, 0, 0.4, 0, 0, 0, 0, 0], np.float32), tol)[2]
    assert not O.fit_slope(k, np.zeros(8, np.float32), tol)[2]
    # positive slope / intercept clamp to zero
    c, a, ok = O.fit_slope(k, np.linspace(0.3, 0.6, 8).astype(np.float32), tol)
    assert ok and c == 0.0
    c, a, ok = O.fit_slope(k[:2], np.array([1.0, 1.0], np.float32), tol)
    assert ok and c == 0.0 and a == 0.0


def test_identical_sketches_give_zero():
    rng = np.random.default_rng(8)
    _, w = _rand_sketch(rng, 1)
    raw = np.stack([np.stack([w[0]] * 4)] * 2)
    out, _, nf, _ = O.query_db(raw, [15, 17, 19, 21], 156, 14, np.full(2, 3_000_000, np.uint64),
                               random=O.Random(1, True))
    assert nf == 0 and out[0, 0] == 0.0 and out[0, 1] == 0.0


# --------------------------------------------------------------------------
# Random-match closed forms
# --------------------------------------------------------------------------
def test_csrs_closed_form():
    for k in (9, 13, 15, 21, 29, 31, 32):
        for rc in (0, 1):
            for l1, l2 in ((2_000_000, 2_100_000), (5_000_000, 40_000), (1, 1)):
                m = 2.0 ** ((k - 1) * 2 + rc)
                got = O.csrs(k, rc, l1, l2)
                assert got == O.csrs(k, rc, l2, l1)
                if m > 2.0 ** 53:
                    # 1 - 1/m rounds to 1 in double: the reference's direct
                    # evaluation (random_match.cpp:330-331) gives j = 0, r = 0
                    assert got == 0.0
                    continue
                j1 = -np.expm1(l1 * np.log1p(-1.0 / m))
                j2 = -np.expm1(l2 * np.log1p(-1.0 / m))
                want = j1 * j2 / (j1 + j2 - j1 * j2) if (j1 > 0 and j2 > 0) else 0.0
                # 1 - pow(1 - 1/m, l) cancels: relative error ~ eps * m / l
                rel = 4e-16 * m / min(l1, l2) + 1e-12
                assert got == pytest.approx(want, rel=rel, abs=1e-300)
    assert O.csrs(15, 1, 0, 100) == 0.0
    assert np.isnan(O.csrs_j(33, 1, 100))


def test_closest_cluster_brute_force():
    rng = np.random.default_rng(4)
    cent = rng.dirichlet(np.ones(4), size=5).astype(np.float32)
    for _ in range(200):
        b = rng.dirichlet(np.ones(4))
        for rc in (0, 1):
            bb = b.copy()
            if rc:
                bb[0] = bb[3] = 0.5 * (b[0] + b[3])
                bb[1] = bb[2] = 0.5 * (b[1] + b[2])
            d = ((cent - bb.astype(np.float32)) ** 2).sum(1)
            assert O.closest_cluster(b, rc, cent) == int(np.argmin(d))


# --------------------------------------------------------------------------
# Matrix utilities: the reference's own test arrays (test/test-matrix.py:34-67)
# --------------------------------------------------------------------------
RR = np.array([1, 2, 3, 4, 5, 6], np.float32)
QQ = np.array([8], np.float32)
QR = np.array([10, 11, 12, 13, 14, 15, 16, 17], np.float32)
SQ1 = np.array([[0, 1, 2, 3], [1, 0, 4, 5], [2, 4, 0, 6], [3, 5, 6, 0]], np.float32)
SQ2 = np.array([[0, 1, 2, 3, 10, 14], [1, 0, 4, 5, 11, 15], [2, 4, 0, 6, 12, 16],
                [3, 5, 6, 0, 13, 17], [10, 11, 12, 13, 0, 8], [14, 15, 16, 17, 8, 0]],
               np.float32)


def test_matrix_ops_reference_arrays():
    assert np.array_equal(O.long_to_square(RR), SQ1)
    assert np.array_equal(O.long_to_square(RR, QR, QQ), SQ2)
    assert np.array_equal(O.square_to_long(SQ1), RR)
    i, j, d = O.sparsify_threshold(SQ2, 5)
    dense = np.zeros_like(SQ2)
    dense[i, j] = d
    dense = dense + dense.T
    want = SQ2.copy()
    want[want >= 5] = 0
    assert np.array_equal(dense, want)
    with pytest.raises(RuntimeError):
        O.sparsify_threshold(SQ2, -1)


def test_condensed_index_round_trip():
    for n in (2, 3, 7, 100, 4097):
        ks = np.unique(np.linspace(0, n * (n - 1) // 2 - 1, 300).astype(np.int64))
        for k in ks:
            i = O.calc_row_idx(int(k), n)
            j = O.calc_col_idx(int(k), i, n)
            assert 0 <= i < j < n
            assert O.square_to_condensed(i, j, n) == k


# --------------------------------------------------------------------------
# Sparse outputs vs the reference test's numpy implementations
# (test/test-dists.py:32-61) on the golden core distances
# --------------------------------------------------------------------------
def test_knn_golden_matches_argsort():
    core = _gold("ppsketch_ref")[:, 0]
    square = O.long_to_square(core)
    n = square.shape[0]
    for knn in (1, 3, 10):
        i, j, d = O.knn_from_square(square, knn)
        ei, ej = [], []
        for r in range(n):
            order = np.argsort(square[r, :])
            got = [c for c in order if c != r][:knn]
            ei += [r] * knn
            ej += got
        assert i.tolist() == ei
        assert j.tolist() == ej
        assert np.all(np.diff(d.reshape(n, knn), axis=1) >= 0)


def test_threshold_golden_matches_scan():
    core = _gold("ppsketch_ref")[:, 0]
    square = O.long_to_square(core)
    for cutoff in (0.001, 0.01, 0.02):
        i, j, d = O.sparsify_threshold(square, cutoff)
        ei = [(r, c) for r in range(square.shape[0]) for c in range(r, square.shape[1])
              if square[r, c] < cutoff and r != c]
        assert list(zip(i.tolist(), j.tolist())) == ei


def test_knn_heap_tie_order_is_libstdcxx():
    """All-equal rows: the kept set and order come from the heap, not from a
    stable sort; pin a few shapes against a direct priority-queue model."""
    import heapq  # noqa: F401  (only for documentation of intent)
    n = 9
    square = np.zeros((n, n), np.float32)
    i, j, d = O.knn_from_square(square, 4)
    # every row keeps 4 of its 8 zero-distance neighbours and no +inf self entry
    assert np.all(d == 0)
    for r in range(n):
        row = j[r * 4:(r + 1) * 4]
        assert r not in row and len(set(row.tolist())) == 4


def test_knn_rows_matches_square_heap():
    """knn_rows over a subset of rows of a square (own column excluded) is the
    square heap restricted to those rows, ties included; with no exclusion
    and distinct distances it is a plain ascending sort."""
    core = _gold("ppsketch_ref")[:, 0]
    square = O.long_to_square(core)
    square[3, :5] = square[3, 5]  # a run of ties in one row
    n = square.shape[0]
    sel = np.array([0, 3, 7, n - 1])
    for knn in (1, 4, n):
        i, j, d = O.knn_from_square(square, knn)
        ri, rj, rd = O.knn_rows(square[sel], sel, knn)
        full = np.arange(n * knn).reshape(n, knn)[sel].ravel()
        assert np.array_equal(ri, i[full]) and np.array_equal(rj, j[full])
        assert np.array_equal(rd.view(np.uint32), d[full].view(np.uint32))
    rng = np.random.default_rng(3)
    rows = rng.permutation(40 * 25).astype(np.float32).reshape(40, 25)
    ri, rj, rd = O.knn_rows(rows, np.full(40, -1), 6)
    assert np.array_equal(ri, np.repeat(np.arange(40), 6))
    assert np.array_equal(rj.reshape(40, 6), np.argsort(rows, axis=1)[:, :6])


# --------------------------------------------------------------------------
# Sketching: restatement vs the reference's own compiled sketch core
# (oracle/_ref/libref_sketch.so, built from sketch/sketch.cpp, seqio.cpp,
# countmin.cpp with -DWEB_SKETCH)
# --------------------------------------------------------------------------
def _rseq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(list(alphabet), size=n))


def _revcomp(s):
    return s[::-1].translate(str.maketrans("ACGTacgt", "TGCAtgca"))


SKETCH_CASES = [
    # contigs, k, s64
    ((("ACGT", 6000),), 15, 16),                      # some empty bins: densified
    ((("ACGT", 3000), ("ACGT", 40), ("ACGT", 7000)), 21, 8),
    ((("ACGTacgtNnU", 9000),), 17, 16),                # lowercase, N, U
    ((("ACGT", 120),), 31, 4),                         # heavy densification
    ((("ACGT", 10),), 15, 2),                          # shorter than k: all bins empty
    ((("ACGT", 150000),), 29, 156),
]


@pytest.mark.parametrize("case", range(len(SKETCH_CASES)))
@pytest.mark.parametrize("rc", [True, False])
def test_sketch_matches_reference_core(case, rc):
    if O.ref_sketch_lib() is None:
        pytest.skip("reference sketch core not built (oracle/_ref)")
    spec, k, s64 = SKETCH_CASES[case]
    rng = np.random.default_rng(case)
    contigs = [_rseq(rng, n, alpha) for alpha, n in spec]
    us, mh, dens = O.sketch(contigs, k, s64, rc)
    rus, rmh, rdens = O.ref_sketch(contigs, k, s64, rc)
    assert np.array_equal(us, rus)
    assert mh == rmh
    assert (dens != 0) == bool(rdens)


CODON_CASES = [
    # contigs, k (span 3k - 2), s64
    ((("ACGT", 6000),), 3, 16),                       # span 7: every bin densified
    ((("ACGTacgtNnU", 9000), ("ACGT", 50)), 9, 16),  # lowercase, N, U; a short contig
    ((("ACGT", 20000),), 15, 32),                      # span 43 > 33
    ((("ACGT", 40000),), 21, 156),
]


@pytest.mark.parametrize("case", range(len(CODON_CASES)))
@pytest.mark.parametrize("rc", [True, False])
def test_codon_phased_sketch_matches_reference_core(case, rc):
    if O.ref_sketch_lib() is None:
        pytest.skip("reference sketch core not built (oracle/_ref)")
    spec, k, s64 = CODON_CASES[case]
    rng = np.random.default_rng(100 + case)
    contigs = [_rseq(rng, n, alpha) for alpha, n in spec]
    us, mh, dens = O.sketch_filtered(contigs, k, s64, rc, codon=True)
    rus, rmh, rdens = O.ref_sketch(contigs, k, s64, rc, codon=True)
    assert np.array_equal(us, rus)
    assert mh == rmh
    assert (dens != 0) == bool(rdens)


def test_codon_phased_ignores_the_wobble_positions():
    # one span-long sequence holds one k-mer: changing bases off the seed
    # (offsets not divisible by 3) keeps the sketch, changing a seed base
    # moves it; an N anywhere in the span removes the k-mer
    rng = np.random.default_rng(8)
    k, s64 = 5, 1
    span = 3 * k - 2
    s = _rseq(rng, span)
    base, _, _ = O.sketch_filtered([s], k, s64, False, codon=True)
    wob = list(s)
    for i in range(span):
        if i % 3:
            wob[i] = "ACGT"[("ACGT".index(wob[i]) + 1) % 4]
    same, _, _ = O.sketch_filtered(["".join(wob)], k, s64, False, codon=True)
    assert np.array_equal(base, same)
    seed = list(s)
    seed[3] = "ACGT"[("ACGT".index(seed[3]) + 1) % 4]
    moved, _, _ = O.sketch_filtered(["".join(seed)], k, s64, False, codon=True)
    assert not np.array_equal(base, moved)
    _, _, dens = O.sketch_filtered([s[:4] + "N" + s[5:]], k, s64, False, codon=True)
    assert dens == -1  # no k-mer: every bin empty


def test_sketch_canonical_strand_symmetry():
    # canonical k-mers: a sequence and its reverse complement sketch alike;
    # forward-only hashing does not
    rng = np.random.default_rng(9)
    s = _rseq(rng, 20000)
    a, _, _ = O.sketch([s], 21, 32, True)
    b, _, _ = O.sketch([_revcomp(s)], 21, 32, True)
    c, _, _ = O.sketch([_revcomp(s)], 21, 32, False)
    d, _, _ = O.sketch([s], 21, 32, False)
    assert np.array_equal(a, b)
    assert not np.array_equal(c, d)


def test_sketch_bins_by_brute_force():
    # bin minima and bit slicing from an independent model: a set of hash
    # values (the sketch of one k-mer equals its bin's minimum, every other
    # bin filled by probing); checked through the layer words
    rng = np.random.default_rng(4)
    kmer = _rseq(rng, 15)
    us, _, dens = O.sketch([kmer], 15, 2, True)
    assert dens == 1
    vals = O.unpack_bins(us.reshape(1, -1), 2, 14)[0]
    # a single k-mer fills every bin with the same 14-bit value
    assert len(set(vals.tolist())) == 1
