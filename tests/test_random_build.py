"""Host pieces of the Monte Carlo random-match construction against the
reference's compiled generators (random/rng.cpp, random/kmeans/
mersenneTwister2002.c, built into oracle/_ref), numpy's independent MT19937,
scikit-learn's Lloyd iterations and the structure the random genomes must
have.  The GPU half (sketching and comparing the genomes, the tables stored
by constructDatabase / addRandom) is in test_gpu_random.py."""
import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def pp():
    import pp_sketchlib
    return pp_sketchlib


def _need_ref():
    if O.ref_random_lib() is None:
        pytest.skip("oracle/_ref/libref_random.so not built (needs the reference sources)")


@pytest.mark.parametrize("seed", [1, 42, 2**63 + 5])
@pytest.mark.parametrize("jumps", [0, 1, 3])
def test_xoshiro_stream_and_jump_match_reference(pp, seed, jumps):
    _need_ref()
    got = pp._xoshiro(seed, 2000, jumps)
    assert np.array_equal(got, O.ref_xoshiro(seed, 2000, jumps))


def test_splitmix_seeding_known_answer(pp):
    # state word 0 is splitmix64(seed); splitmix64 from 0 gives 0xe220a8397b1dcdaf,
    # so with seed 0 the first output is rotl(s1 * 5, 7) * 9 where s1 = splitmix64(s0)
    def splitmix(x):
        m = (1 << 64) - 1
        z = (x + 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)
    s0 = splitmix(0)
    assert s0 == 0xE220A8397B1DCDAF
    s1 = splitmix(s0)
    m = (1 << 64) - 1
    x = (s1 * 5) & m
    want = ((((x << 7) | (x >> 57)) & m) * 9) & m
    assert int(pp._xoshiro(0, 1)[0]) == want


def test_kmeans_generator_is_mt19937(pp):
    # numpy's legacy RandomState seeds MT19937 as init_genrand does
    u = np.random.RandomState(3019).randint(0, 2**32, size=50, dtype=np.uint32)
    want = u.astype(np.float64) * (1.0 / 4294967295.0)
    if O.ref_random_lib() is not None:
        assert np.array_equal(O.ref_mt_double(3019, 50), want)
    # the seeding's first pick is row floor(n * u0)
    x = np.random.default_rng(0).random((37, 4))
    c, _, it = pp._kmeans(x, 1, 3019, 0)
    assert it == 0
    assert np.array_equal(c[0], x[int(37 * want[0])])


def _plusplus(x, k, u):
    """kmeans++ seeding with the reference's discrete_rand on given uniforms."""
    n = x.shape[0]
    picks = [int(n * u[0])]
    nearest = None
    for c in range(1, k):
        cur = ((x - x[picks[-1]]) ** 2).sum(axis=1)
        nearest = cur if nearest is None else np.minimum(cur, nearest)
        r = nearest.sum() * u[c]
        run, pick = nearest[0], 0
        while r >= run and pick < n - 1:
            pick += 1
            run += nearest[pick]
        picks.append(pick)
    return picks


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kmeans_plusplus_seeding(pp, seed):
    x = np.random.default_rng(seed).random((25, 4))
    u = np.random.RandomState(3019).randint(0, 2**32, size=8, dtype=np.uint32) / 4294967295.0
    c, _, _ = pp._kmeans(x, 4, 3019, 0)
    picks = _plusplus(x, 4, u)
    assert np.array_equal(c, x[picks])


def test_kmeans_lloyd_reaches_sklearn_fixed_point(pp):
    from sklearn.cluster import KMeans
    rng = np.random.default_rng(5)
    centres = np.array([[0.30, 0.20, 0.20, 0.30], [0.25, 0.25, 0.25, 0.25],
                        [0.18, 0.32, 0.32, 0.18]])
    x = np.concatenate([cc + rng.normal(0, 0.01, size=(20, 4)) for cc in centres])
    init, _, _ = pp._kmeans(x, 3, 3019, 0)
    c, z, it = pp._kmeans(x, 3, 3019, 300)
    assert 1 <= it < 300
    z = np.array(z)
    # a fixed point: every row joins its first nearest centre, every centre
    # is its members' mean
    d = ((x[:, None, :] - c[None, :, :]) ** 2).sum(axis=2)
    assert np.array_equal(z, d.argmin(axis=1))
    for k in range(3):
        assert np.allclose(c[k], x[z == k].mean(axis=0), rtol=0, atol=1e-15)
    km = KMeans(n_clusters=3, init=init, n_init=1, max_iter=300, tol=0,
                algorithm="lloyd").fit(x)
    assert np.array_equal(km.labels_, z)
    assert np.allclose(km.cluster_centers_, c, rtol=0, atol=1e-14)


def test_kmeans_more_centres_than_rows(pp):
    x = np.random.default_rng(1).random((2, 4))
    c, z, _ = pp._kmeans(x, 3, 3019, 300)
    # two seeded centres, the third left at zero and empty
    assert np.array_equal(np.sort(z), [0, 1])
    assert np.array_equal(c[2], np.zeros(4))


CENTROIDS = np.array([[0.31, 0.19, 0.19, 0.31], [0.2, 0.3, 0.3, 0.2], [0.25, 0.25, 0.25, 0.25]],
                     np.float32)


@pytest.mark.parametrize("use_rc", [True, False])
def test_random_genome_composition(pp, use_rc):
    L = 60001
    g = pp._randomGenomes(CENTROIDS, 2, L, use_rc, 3)
    assert len(g) == 6 and all(len(s) == L for s in g)
    comp = {"A": "T", "C": "G", "G": "C", "T": "A"}
    for i, s in enumerate(g):
        assert set(s) <= set("ACGT")
        p = CENTROIDS[i // 2].astype(np.float64)
        p /= p.sum()
        head = s[:L // 2 + 1]  # positions 0 .. L/2 are forward
        tail = s[L // 2 + 1:]
        if use_rc:
            tail = "".join(comp[b] for b in tail)
        for part in (head, tail):
            f = np.array([part.count(b) for b in "ACGT"]) / len(part)
            sd = np.sqrt(p * (1 - p) / len(part))
            assert np.all(np.abs(f - p) < 5 * sd), (i, f, p)


def test_random_genome_streams_follow_thread_split(pp):
    L = 5000
    one = pp._randomGenomes(CENTROIDS, 2, L, True, 1)
    assert one == pp._randomGenomes(CENTROIDS, 2, L, True, 1)  # deterministic
    # 6 genomes over 2 threads: thread 0 makes 0-2 from the unjumped stream,
    # as a single thread does
    two = pp._randomGenomes(CENTROIDS, 2, L, True, 2)
    assert two[:3] == one[:3] and two[3:] != one[3:]
    # genome 0 always comes from thread 0's first, unjumped draws
    for t in (3, 4, 6, 8):
        g = pp._randomGenomes(CENTROIDS, 2, L, True, t)
        assert g[0] == one[0]
        assert len(set(g)) == 6
    # 6 threads: each genome its own thread, thread t jumped t times
    six = pp._randomGenomes(CENTROIDS, 2, L, True, 6)
    assert all(six[i] != one[i] for i in range(1, 6))


def test_mc_kmer_lengths_closed_form():
    # one bin of 1,024 against two genomes of 50 kb: k = 3 .. 12
    assert O.mc_kmer_lengths(50_000, 16, True) == list(range(3, 13))
    # 2 Mb and 10,000 bins (156 words): k = 3 .. 17
    assert O.mc_kmer_lengths(2_000_000, 156, True) == list(range(3, 18))
