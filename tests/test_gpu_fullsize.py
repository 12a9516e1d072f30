"""Every workload at its own size on the GPU, checked against the oracle.

- The 268,435,456-symbol repeat-rich text (order-3 Markov chain over {0..3},
  40% planted repeats of mean length 5,000 with 1% substitutions, u32 index):
  the default build, checked by the oracle's verifier (permutation, strictly
  increasing adjacent suffixes: the unique suffix array), and the builds with
  4,096 classification buckets under binary search, the piecewise-linear
  router and the MLP router, each bit-equal to the verified one.
- The batch of 4,096 texts (lognormal lengths, median 65,536, sigma 1.5,
  clipped to [1,024, 16,777,216]; even texts u8 {0..3}, odd texts u16
  Zipf(1.1) over 50,000 symbols): 48 seeded texts equal to the oracle's
  build_in_memory of each text alone.
- The 820M genome share (175,000 segments, u64 index): the device segment
  directory (built for 256 or more ends), checked by the verifier with the
  canonical order inside comparator-Equal classes.
- Short segments in both boundary orders, with and without max_context:
  suffix and LCP arrays bit-exact against the oracle's build_in_memory.
"""
import time

import numpy as np
import pytest
import torch

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def repeat_rich():
    n = 1 << 28
    t0 = time.perf_counter()
    text = gen.markov_repeats(n, seed=0xC1)
    _log(f"repeat-rich n={n} generated in {time.perf_counter() - t0:.1f} s")
    return text


def test_repeat_rich_268M_plain_and_routers(cuda, repeat_rich):
    text = repeat_rich
    n = len(text)
    d_text = dev.text_to_device(text)
    d_out = torch.empty(n, dtype=torch.int32, device=cuda)
    _log("case repeat-rich-268M default build")
    dev.build_device(d_text, n, 1, idx_bytes=4, d_out=d_out)
    torch.cuda.synchronize()
    st = caps_sa.last_stats()
    sa = d_out.cpu().numpy().view(np.uint32).copy()
    t0 = time.perf_counter()
    bad = O.Ctx(text).verify(sa.astype(np.uint64), strict=True, canon=True, threads=16)
    _log(f"  {st['ms_total']:.1f} ms, verify {time.perf_counter() - t0:.1f} s -> {bad}")
    assert bad == -1
    for router in (caps_sa.Router.BSEARCH, caps_sa.Router.PWL, caps_sa.Router.MLP):
        _log(f"case repeat-rich-268M 4096 buckets router={router.name}")
        dev.build_device(d_text, n, 1, idx_bytes=4, d_out=d_out, router=int(router), gpu_buckets=4096)
        torch.cuda.synchronize()
        st = caps_sa.last_stats()
        _log(f"  {st['ms_total']:.1f} ms, classification buckets {st['coarse_buckets']} "
             f"(largest {st['coarse_max']}), buckets {st['buckets']}")
        assert st["coarse_buckets"] == 4096
        got = d_out.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, sa), router
        if router != caps_sa.Router.BSEARCH:
            assert st["router_count"] == n


def test_batch_4096_texts(cuda):
    count = 4096
    lengths = gen.lognormal_lengths(count, 65536.0, 1.5, 1024, 1 << 24, seed=0xC1).astype(np.int64)
    texts = []
    for i, m in enumerate(lengths.tolist()):
        sd = 0xC1 * 1_000_003 + i
        texts.append(gen.uniform_u8(m, 4, seed=sd) if i % 2 == 0 else gen.zipf_u16(m, 50000, 1.1, seed=sd))
    _log(f"case batch-4096: {int(lengths.sum())} symbols")
    sample = set(np.random.default_rng(0xC1).choice(count, size=48, replace=False).tolist())
    for w, sel in ((1, list(range(0, count, 2))), (2, list(range(1, count, 2)))):
        cat = np.concatenate([texts[i] for i in sel])
        ends = np.cumsum([len(texts[i]) for i in sel]).astype(np.uint64)
        d_text = dev.text_to_device(cat)
        out = dev.build_batch_device(d_text, len(cat), w, ends)
        torch.cuda.synchronize()
        _log(f"  sym_bytes={w}: {len(sel)} texts, {len(cat)} symbols, {caps_sa.last_stats()['ms_total']:.1f} ms")
        got = out.cpu().numpy().view(np.uint32)
        starts = np.concatenate([[0], ends[:-1]]).astype(np.int64)
        for k, i in enumerate(sel):
            if i not in sample:
                continue
            a, b = int(starts[k]), int(ends[k])
            want = O.Ctx(texts[i]).build_in_memory(threads=16)
            assert np.array_equal(got[a:b].astype(np.uint64) - a, want), (i, len(texts[i]))


def test_genome_share_820M_device_directory(cuda):
    n, n_seg = 820_000_000, 175_000
    t0 = time.perf_counter()
    text, ends, pos = gen.genome(n, n_seg, seed=0xC1, positions=True)
    _log(f"case genome-share-820M: {n_seg} segments, {len(pos)} positions ({time.perf_counter() - t0:.1f} s)")
    lp = caps_sa.SegmentedText(n, ends)
    d_text = dev.text_to_device(text)
    d_pos = torch.from_numpy(pos.view(np.int64)).to(cuda)
    sa, _ = dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=len(pos), idx_bytes=8, lp=lp)
    torch.cuda.synchronize()
    _log(f"  {caps_sa.last_stats()['ms_total']:.1f} ms")
    del d_pos
    got = sa.cpu().numpy().view(np.uint64)
    t0 = time.perf_counter()
    bad = O.Ctx(text, ends, O.ORDER_SHORTER_FIRST).verify(got, positions=pos, strict=True, canon=True, threads=16)
    _log(f"  verify {time.perf_counter() - t0:.1f} s -> {bad}")
    assert bad == -1


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("max_ctx", [caps_sa.MAX_CONTEXT, 40])
def test_many_short_segments_exact(cuda, order, max_ctx):
    """300..1,000 short segments (the device directory), shared stretches
    across segments and segment ends inside them; u32 and u64 index."""
    rng = np.random.default_rng(300 + order)
    unit = rng.integers(0, 4, 120).astype(np.uint8)
    parts = []
    for _ in range(700):
        parts.append(rng.integers(0, 4, int(rng.integers(1, 40))).astype(np.uint8))
        parts.append(unit[: int(rng.integers(20, 120))])
    text = np.concatenate(parts)
    n = len(text)
    cuts = np.unique(rng.integers(1, n, int(rng.integers(300, 1000))))
    ends = np.append(cuts, n).astype(np.uint64)
    assert len(ends) >= 256
    lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
    ctx = O.Ctx(text, ends, order)
    want, want_lcp = ctx.build_in_memory(max_ctx=max_ctx, threads=8, want_lcp=True)
    if order == 0:
        want = ctx.canonicalize(want, max_ctx)
    d_text = dev.text_to_device(text)
    for idx in (4, 8):
        _log(f"case segments={len(ends)} n={n} order={order} max_ctx={max_ctx} idx={idx}")
        sa, lcp = dev.build_device(d_text, n, 1, idx_bytes=idx, lp=lp, max_context=max_ctx, want_lcp=True)
        torch.cuda.synchronize()
        dt = np.uint32 if idx == 4 else np.uint64
        got = sa.cpu().numpy().view(dt).astype(np.uint64)
        got_lcp = lcp.cpu().numpy().view(dt).astype(np.uint64)
        assert np.array_equal(got, want)
        assert np.array_equal(got_lcp, want_lcp)


def test_device_positions_out_of_range_are_refused(cuda):
    text = gen.uniform_u8(1000, 4, seed=1)
    d_text = dev.text_to_device(text)
    for idx, dt in ((4, torch.int32), (8, torch.int64)):
        d_pos = torch.tensor([0, 5, 1000], dtype=dt, device=cuda)
        with pytest.raises(caps_sa.CapsError) as e:
            dev.build_device(d_text, 1000, 1, d_positions=d_pos, n_pos=3, idx_bytes=idx)
        assert e.value.code == caps_sa.CAPS_E_INVALID
