"""Bucket batches and kept-position bitmaps on the GPU.

A build may sort its coarse intervals a batch at a time (batch_records);
the suffix array and LCP array must not change, and must equal the oracle's
build_in_memory.  The filter entry (the crate's build_ext_mem_for_filter,
whose FilteredSource walks a bitmap of kept positions) must emit exactly what
build_ext_mem_for_positions emits for the same positions."""
import numpy as np
import pytest

import caps_sa
import oracle as O
from caps_sa import gen

pytestmark = pytest.mark.gpu


def _sa_lcp(d_text, n, w, **kw):
    from caps_sa import device as dev

    sa, lcp = dev.build_device(d_text, n, w, want_lcp=True, **kw)
    st = caps_sa.last_stats()
    return (sa.cpu().numpy().view(np.uint32 if sa.element_size() == 4 else np.uint64).astype(np.uint64),
            lcp.cpu().numpy().view(np.uint32 if lcp.element_size() == 4 else np.uint64).astype(np.uint64),
            st["batches"])


def test_batched_builds_match_one_batch(cuda):
    from caps_sa import device as dev

    for text in (gen.markov_repeats(300000, seed=61, mean_len=700.0), gen.zipf_u16(80000, 300, 1.1, seed=62),
                 gen.uniform_u8(200000, 4, seed=63)):
        w = text.dtype.itemsize
        n = len(text)
        d_text = dev.text_to_device(text)
        want, want_lcp = O.Ctx(text).build_in_memory(threads=4, want_lcp=True)
        one, one_lcp, b1 = _sa_lcp(d_text, n, w)
        assert b1 == 1
        assert np.array_equal(one, want) and np.array_equal(one_lcp, want_lcp)
        # records of all batches written by one pass 1, or pass 1 per batch
        for flags in (0, caps_sa.FLAG_SCATTER_PER_BATCH):
            for rec in (n // 3, 20000, 4096):
                print(f"case n={n} w={w} batch_records={rec} flags={flags}", flush=True)
                sa, lcp, nb = _sa_lcp(d_text, n, w, batch_records=rec, flags=flags)
                assert nb > 1
                assert np.array_equal(sa, want), (rec, flags)
                assert np.array_equal(lcp, want_lcp), (rec, flags)


def test_batched_segmented_positions(cuda):
    import torch
    from caps_sa import device as dev

    text, ends, pos = gen.genome(300000, 100, median=2500.0, n_frac=0.058, n_mean=500.0, seed=64)
    n = len(text)
    d_text = dev.text_to_device(text)
    d_pos = torch.from_numpy(pos.view(np.int64)).to("cuda:0")
    for order in (0, 1):
        lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
        ctx = O.Ctx(text, ends, order)
        want = ctx.build_in_memory(pos, threads=4)
        if order == 0:
            want = ctx.canonicalize(want, caps_sa.MAX_CONTEXT)
        for rec, flags in ((0, 0), (30000, 0), (30000, caps_sa.FLAG_SCATTER_PER_BATCH)):
            sa, _ = dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=len(pos), idx_bytes=8, lp=lp,
                                     batch_records=rec, flags=flags)
            got = sa.cpu().numpy().view(np.uint64)
            assert np.array_equal(got, want), (order, rec, flags)


def test_filter_entry_matches_positions_entry(cuda):
    text, ends, pos = gen.genome(250000, 90, median=2500.0, n_frac=0.058, n_mean=500.0, seed=65)
    n = len(text)
    lp = caps_sa.SegmentedText(n, ends)
    for rec in (0, 25000):
        opts = caps_sa.ExtMemOpts(batch_records=rec)
        a, b = [], []
        caps_sa.build_ext_mem_for_positions_with(text, pos, lp, opts, a.append)
        caps_sa.build_ext_mem_for_filter_with(text, lambda q: text[q.astype(np.int64)] < 4, lp, opts, b.append)
        assert len(a) == len(pos) and a == b, rec
    # plain text, a sparse random filter, and the predicate called per position
    t = gen.uniform_u8(3000, 3, seed=66)
    keep = np.random.default_rng(67).random(len(t)) < 0.3
    a, b = [], []
    caps_sa.build_ext_mem_for_positions(t, np.flatnonzero(keep).astype(np.uint64), None, a.append)
    caps_sa.build_ext_mem_for_filter(t, lambda q: bool(keep[int(q)]), None, b.append)
    assert a == b and len(a) == int(keep.sum())


def test_filter_edge_cases(cuda):
    t = np.frombuffer(b"mississippi", dtype=np.uint8).copy()
    for mask, want in ((np.zeros(11, bool), []), (np.eye(1, 11, 4, dtype=bool)[0], [4])):
        got = []
        caps_sa.build_ext_mem_for_filter(t, mask, None, got.append)
        assert got == want
    got = []
    caps_sa.build_ext_mem_for_filter(t, np.ones(11, bool), None, got.append)
    assert got == [10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2]


def test_device_filtered_matches_positions(cuda):
    import torch
    from caps_sa import device as dev

    text, ends, pos = gen.genome(200000, 70, median=2500.0, n_frac=0.058, n_mean=500.0, seed=68)
    n = len(text)
    lp = caps_sa.SegmentedText(n, ends)
    d_text = dev.text_to_device(text)
    bits = caps_sa.keep_bitmap(text < 4, n)
    d_keep = torch.from_numpy(bits.view(np.int64)).to("cuda:0")
    d_pos = torch.from_numpy(pos.view(np.int64)).to("cuda:0")
    want, want_lcp = dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=len(pos), idx_bytes=8, lp=lp,
                                      want_lcp=True)
    for rec, flags in ((0, 0), (20000, 0), (20000, caps_sa.FLAG_SCATTER_PER_BATCH)):
        sa, lcp, kept = dev.build_device_filtered(d_text, n, 1, d_keep, len(pos), idx_bytes=8, lp=lp,
                                                  want_lcp=True, batch_records=rec, flags=flags)
        assert kept == len(pos)
        assert torch.equal(sa[:kept], want) and torch.equal(lcp[:kept], want_lcp), (rec, flags)
    with pytest.raises(caps_sa.CapsError):
        dev.build_device_filtered(d_text, n, 1, d_keep, len(pos) - 1, idx_bytes=8, lp=lp)


def test_emit_error_stops_a_batched_stream(cuda):
    t = gen.uniform_u8(100000, 4, seed=69)
    seen = []

    def emit(p):
        seen.append(p)
        if len(seen) == 5:
            raise ValueError("stop")

    with pytest.raises(caps_sa.BuildError) as e:
        caps_sa.try_build_in_memory_sample_sort(t, caps_sa.ExtMemOpts(batch_records=10000), emit)
    assert e.value.kind == "Emit" and isinstance(e.value.inner, ValueError)
    assert len(seen) == 5


def test_filter_star_and_u16_positions_batches(cuda):
    """The filter entry under STAR segment order, and a u16 text with a
    position list sorted in bucket batches, both against the positions entry
    and the oracle."""
    import torch
    from caps_sa import device as dev

    text, ends, pos = gen.genome(120000, 60, median=2500.0, n_frac=0.058, n_mean=400.0, seed=70)
    n = len(text)
    lp = caps_sa.StarSegmented(n, ends)
    for rec in (0, 9000):
        a, b = [], []
        opts = caps_sa.ExtMemOpts(batch_records=rec)
        caps_sa.build_ext_mem_for_positions_with(text, pos, lp, opts, a.append)
        caps_sa.build_ext_mem_for_filter_with(text, text < 4, lp, opts, b.append)
        want = O.Ctx(text, ends, 1).build_in_memory(pos, threads=4)
        assert a == b and np.array_equal(np.array(a, dtype=np.uint64), want), rec
    t16 = gen.zipf_u16(60000, 2000, 1.1, seed=71)
    keep = np.random.default_rng(72).random(len(t16)) < 0.6
    p16 = np.flatnonzero(keep).astype(np.uint64)
    d_text = dev.text_to_device(t16)
    d_pos = torch.from_numpy(p16.view(np.int64)).to("cuda:0")
    want = O.Ctx(t16).build_in_memory(p16, threads=4)
    for rec in (0, 5000):
        sa, _ = dev.build_device(d_text, len(t16), 2, d_positions=d_pos, n_pos=len(p16), idx_bytes=8,
                                 batch_records=rec)
        assert np.array_equal(sa.cpu().numpy().view(np.uint64), want), rec


def test_filtered_builds_with_learned_routers(cuda):
    """The learned routers sample kept positions from the bitmap (router
    training and splitter sample); the suffix array does not change."""
    import torch
    from caps_sa import device as dev

    text, ends, pos = gen.genome(400000, 150, median=2500.0, n_frac=0.058, n_mean=500.0, seed=73)
    n = len(text)
    lp = caps_sa.SegmentedText(n, ends)
    d_text = dev.text_to_device(text)
    d_keep = torch.from_numpy(caps_sa.keep_bitmap(text < 4, n).view(np.int64)).to("cuda:0")
    base, _, kept = dev.build_device_filtered(d_text, n, 1, d_keep, len(pos), idx_bytes=8, lp=lp)
    for router in (1, 2):
        sa, _, k2 = dev.build_device_filtered(d_text, n, 1, d_keep, len(pos), idx_bytes=8, lp=lp, router=router,
                                              batch_records=60000)
        assert k2 == kept and torch.equal(sa[:k2], base[:kept]), router
