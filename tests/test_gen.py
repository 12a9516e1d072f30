"""Workload generators: determinism and the properties the configs state."""
import numpy as np

from caps_sa import gen


def test_deterministic():
    assert np.array_equal(gen.uniform_u8(1000, 4, seed=3), gen.uniform_u8(1000, 4, seed=3))
    assert not np.array_equal(gen.uniform_u8(1000, 4, seed=3), gen.uniform_u8(1000, 4, seed=4))
    assert np.array_equal(gen.markov_repeats(50000, seed=9), gen.markov_repeats(50000, seed=9))


def test_uniform_alphabet():
    t = gen.uniform_u8(100000, 4, seed=1)
    counts = np.bincount(t, minlength=4)
    assert t.max() == 3 and counts.min() > 24000


def test_markov_repeats_are_repeat_rich():
    n = 1 << 20
    t = gen.markov_repeats(n, seed=0xC1)
    assert set(np.unique(t).tolist()) <= {0, 1, 2, 3}
    # planted repeats: many 64-mers occur more than once
    k = 64
    view = np.lib.stride_tricks.sliding_window_view(t[: n - n % k], k)[::k]
    packed = [bytes(v) for v in view[:4000]]
    whole = t.tobytes()
    dup = sum(whole.count(p) > 1 for p in packed[:400])
    assert dup > 100


def test_zipf_and_lengths():
    z = gen.zipf_u16(200000, 50000, 1.1, seed=2)
    assert z.max() < 50000 and np.mean(z == 0) > 0.05
    L = gen.lognormal_lengths(4096, 65536, 1.5, 1024, 1 << 24, seed=5)
    assert L.min() >= 1024 and L.max() <= 1 << 24
    assert 40000 < np.median(L) < 100000


def test_genome_layout():
    text, ends, pos = gen.genome(1_000_000, 300, n_frac=0.058, n_mean=2000.0, seed=3)
    assert int(ends[-1]) == len(text) and np.all(np.diff(ends.astype(np.int64)) >= 2)
    assert np.all(text[ends.astype(np.int64) - 1] >= 4)
    frac_n = np.mean(text == 5)
    assert 0.04 < frac_n < 0.09
    assert np.all(text[pos.astype(np.int64)] < 4)
