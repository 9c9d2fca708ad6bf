"""Workload generators: deterministic, right shapes and alphabets."""
import numpy as np

from psacak_amd import gen


def test_iid_dna_deterministic_and_thread_independent():
    a = gen.iid_dna(3_000_000, seed=9, threads=1)
    b = gen.iid_dna(3_000_000, seed=9, threads=8)
    assert np.array_equal(a, b)
    assert set(np.unique(a).tolist()) == {65, 67, 71, 84}
    c = np.bincount(a, minlength=256)[[65, 67, 71, 84]] / len(a)
    assert np.all(np.abs(c - 0.25) < 0.01)
    assert not np.array_equal(a, gen.iid_dna(3_000_000, seed=10))


def test_markov_dna_has_repeats_and_skew():
    n = 1 << 22
    a = gen.markov_dna(n, seed=2, threads=4)
    b = gen.markov_dna(n, seed=2, threads=1)
    assert np.array_equal(a, b)
    assert set(np.unique(a).tolist()) <= {65, 67, 71, 84}
    norep = gen.markov_dna(n, seed=2, repeat_frac=0.0)
    assert not np.array_equal(a, norep)

    def distinct_kmers(x, k=28):
        code = np.searchsorted(np.array([65, 67, 71, 84]), x).astype(np.uint64)
        v = np.zeros(len(x) - k + 1, dtype=np.uint64)
        for j in range(k):
            v |= code[j: len(x) - k + 1 + j] << np.uint64(2 * j)
        return len(np.unique(v)) / len(v)

    # copies make a visible share of the 28-mers repeat
    assert distinct_kmers(a) < distinct_kmers(norep) - 0.05


def test_zipf32_shape():
    z = gen.zipf32(1 << 20, sigma=1 << 16, s=1.1, repeat_frac=0.2, seed=3)
    assert z.dtype == np.uint32 and z.max() < (1 << 16)
    cnt = np.bincount(z)
    top = np.sort(cnt)[::-1]
    assert top[0] > 20 * top[100]


def test_lengths_and_requests():
    L = gen.lengths(10000, 1024, 4 << 20, seed=4)
    assert L.min() >= 1024 and L.max() <= 4 << 20
    assert 300e3 < L.mean() < 700e3
    g, l = gen.requests(2000, mean_ms=5.0)
    assert abs(g.mean() - 5.0) < 0.5
    assert (l >= 10 << 10).all()
