"""Graph builds (CAPS_FLAG_GRAPH): the tile sort, refinement and giant-group
phases captured into hipGraphs and launched on the build's stream.  Each
graph build is checked against the oracle, with eager builds and host
synchronizations between graph builds, over inputs whose phases launch the
same kernels (the instantiated graphs are updated in place) and inputs whose
phases differ (instantiated again)."""
import numpy as np
import pytest
import torch

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu

G = caps_sa.FLAG_GRAPH


def _batch(seed, count, median):
    lengths = gen.lognormal_lengths(count, median, 1.5, 64, 1 << 17, seed=seed)
    return [gen.uniform_u8(int(m), 4, seed=seed * 7919 + i) if i % 2 == 0 else
            gen.markov_repeats(int(m), seed=seed * 7919 + i, mean_len=200.0) for i, m in enumerate(lengths)]


def _run_batch(texts, flags):
    ends = np.cumsum([len(t) for t in texts]).astype(np.uint64)
    cat = np.concatenate(texts)
    out = dev.build_batch_device(dev.text_to_device(cat), len(cat), 1, ends, flags=flags)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32).astype(np.uint64)
    st = caps_sa.last_stats()
    return got, ends, st


def _check_batch(texts, got, ends):
    start = 0
    for t, e in zip(texts, ends.tolist()):
        want = O.Ctx(t).build_in_memory(threads=8) + start
        assert np.array_equal(got[start:e], want)
        start = e


def test_graph_batches_replayed_with_eager_work_between(cuda):
    for rep, (seed, count, median) in enumerate([(1, 120, 3000.0), (1, 120, 3000.0), (2, 120, 3000.0),
                                                 (3, 40, 20000.0), (1, 120, 3000.0)]):
        texts = _batch(seed, count, median)
        print(f"case graph batch rep={rep} seed={seed} texts={count}", flush=True)
        got, ends, st = _run_batch(texts, G)
        assert sum(st["k_graph"]) > 0
        _check_batch(texts, got, ends)
        eager, _, _ = _run_batch(texts, 0)  # eager build of the same batch between graph builds
        assert np.array_equal(got, eager)


@pytest.mark.parametrize("idx", [4, 8])
def test_graph_single_texts_with_giant_groups(cuda, idx):
    """Texts whose tie groups exceed 2,048 members (runs of one symbol, a
    long repeated unit), so the giant-group graph runs, with LCP arrays."""
    rng = np.random.default_rng(5)
    unit = rng.integers(0, 4, 300).astype(np.uint8)
    cases = [np.zeros(20000, dtype=np.uint8), np.tile(unit, 60), gen.markov_repeats(200_000, seed=8, mean_len=3000.0),
             np.tile(unit, 60)]
    for i, text in enumerate(cases):
        print(f"case graph text {i} n={len(text)} idx={idx}", flush=True)
        d_text = dev.text_to_device(text)
        want, want_lcp = O.Ctx(text).build_in_memory(threads=8, want_lcp=True)
        for flags in (G, 0, G):
            sa, lcp = dev.build_device(d_text, len(text), 1, idx_bytes=idx, want_lcp=True, flags=flags)
            torch.cuda.synchronize()
            dt = np.uint32 if idx == 4 else np.uint64
            assert np.array_equal(sa.cpu().numpy().view(dt).astype(np.uint64), want), (i, flags)
            assert np.array_equal(lcp.cpu().numpy().view(dt).astype(np.uint64), want_lcp), (i, flags)
