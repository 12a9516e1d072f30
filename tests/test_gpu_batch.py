"""Batched builds: every text's suffix array bit-identical to the oracle's
build_in_memory of that text alone (u8 and u16 texts, uneven lengths,
empty and one-symbol texts)."""
import numpy as np
import pytest

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu


def _check_all(texts, outs):
    for i, (t, got) in enumerate(zip(texts, outs)):
        want = O.Ctx(t).build_in_memory(threads=4) if len(t) else np.zeros(0, dtype=np.uint64)
        assert np.array_equal(got.astype(np.uint64), want), (i, len(t))


def test_small_mixed_batch(cuda):
    rng = np.random.default_rng(40)
    texts = [np.frombuffer(b"banana", dtype=np.uint8), np.zeros(0, dtype=np.uint8), np.array([7], dtype=np.uint8),
             np.frombuffer(b"mississippi", dtype=np.uint8), np.zeros(5000, dtype=np.uint8)]
    for n in rng.integers(1, 3000, 60):
        texts.append(rng.integers(0, 4, n).astype(np.uint8))
    for n in rng.integers(1, 3000, 40):
        texts.append(gen.zipf_u16(int(n), 50000, 1.1, seed=int(n)))
    outs = caps_sa.build_batch(texts)
    _check_all(texts, outs)


def test_lognormal_batch(cuda):
    """A scaled-down copy of the 4,096-text workload: lognormal lengths, half
    u8 over {0..3}, half u16 Zipf(1.1) over 50,000 symbols."""
    count = 256
    lengths = gen.lognormal_lengths(count, 4096, 1.5, 64, 1 << 17, seed=7)
    texts = []
    for i, n in enumerate(lengths.tolist()):
        if i % 2 == 0:
            texts.append(gen.uniform_u8(n, 4, seed=1000 + i))
        else:
            texts.append(gen.zipf_u16(n, 50000, 1.1, seed=1000 + i))
    outs = caps_sa.build_batch(texts)
    _check_all(texts, outs)


def test_batch_device_entry(cuda):
    rng = np.random.default_rng(41)
    texts = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(100, 20000, 30)]
    ends = np.cumsum([len(t) for t in texts]).astype(np.uint64)
    cat = np.concatenate(texts)
    d_text = dev.text_to_device(cat)
    out = dev.build_batch_device(d_text, len(cat), 1, ends).cpu().numpy().view(np.uint32).astype(np.uint64)
    start = 0
    for t, e in zip(texts, ends.tolist()):
        want = O.Ctx(t).build_in_memory() + start
        assert np.array_equal(out[start:e], want)
        start = e


def test_batch_shared_stretches(cuda):
    """Texts whose tie groups share long stretches, alone and across texts:
    every text is built from copies of a few common units, so groups span
    texts and part where a text ends (a window end inside a stretch)."""
    rng = np.random.default_rng(42)
    units = [rng.integers(0, 4, 600).astype(np.uint8) for _ in range(3)]
    texts = []
    for i in range(48):
        parts = []
        for _ in range(int(rng.integers(1, 6))):
            parts.append(rng.integers(0, 4, int(rng.integers(1, 30))).astype(np.uint8))
            u = units[int(rng.integers(0, 3))]
            parts.append(u[: len(u) - int(rng.integers(0, 80))])
        texts.append(np.concatenate(parts))
    outs = caps_sa.build_batch(texts)
    _check_all(texts, outs)
