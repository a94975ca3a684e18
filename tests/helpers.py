"""Shared fixtures for the parity tests: seeded synthetic sketch sets and the
oracle's view of them."""
import functools

import numpy as np

import ppsk_hip as H

KLIST8 = list(range(15, 30, 2))


@functools.lru_cache(maxsize=None)
def synth(n, klist=tuple(KLIST8), s64=156, seed=11, dcore_max=0.05, dacc_max=0.30):
    raw, lengths, bf = H.synth(n, list(klist), s64, seed=seed, dcore_max=dcore_max,
                               dacc_max=dacc_max)
    raw.setflags(write=False)
    return raw, lengths, bf


def with_duplicates(raw, lengths, dup_pairs):
    """Copy sample b's sketch onto sample a for each (a, b): exact ties."""
    raw = raw.copy()
    lengths = lengths.copy()
    for a, b in dup_pairs:
        raw[a] = raw[b]
        lengths[a] = lengths[b]
    return raw, lengths


def mc_table(klist, n_clusters=3, seed=5, symmetric=True, zero_from=None):
    """A Monte Carlo random-match table [n_k][C][C] (float32) and k<max_k flags."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.0, 0.02, size=(len(klist), n_clusters, n_clusters)).astype(np.float32)
    if symmetric:
        t = ((t + t.transpose(0, 2, 1)) / 2).astype(np.float32)
    lt = np.ones(len(klist), bool)
    if zero_from is not None:
        lt = np.array([k < zero_from for k in klist])
        t[~lt] = 0
    return t, lt


def max_abs_diff(a, b):
    """max |a - b| over finite entries; infinities must coincide exactly."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape
    inf_a, inf_b = np.isinf(a), np.isinf(b)
    assert np.array_equal(inf_a, inf_b) and np.array_equal(a[inf_a], b[inf_b])
    fin = ~inf_a
    return float(np.abs(a[fin] - b[fin]).max()) if fin.any() else 0.0
