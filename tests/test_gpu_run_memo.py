"""The run memo: maximal runs of one code recorded once per build, so that
the LCP scans of groups of 65 or more tied suffixes skip the common part of
two runs of the same code.  Builds with the memo, without it
(CAPS_FLAG_NO_RUN_MEMO) and the oracle's build_in_memory agree (suffix and
LCP arrays), on texts whose tie groups run into long runs: N runs of many
lengths (equal lengths too), homopolymer runs of every nucleotide, runs that
end at a segment end or at the text end, runs shorter than the memo's
threshold, in both boundary orders and under max_context."""
import numpy as np
import pytest
import torch

import caps_sa
import oracle as O
from caps_sa import device as dev

pytestmark = pytest.mark.gpu

NO = caps_sa.FLAG_NO_RUN_MEMO


def _runs_text(seed, n_units=900):
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(n_units):
        parts.append(rng.integers(0, 4, int(rng.integers(1, 6))).astype(np.uint8))
        sym = 5 if i % 3 else int(rng.integers(0, 5))
        ln = int(rng.choice([20, 47, 48, 49, 64, 100, 500, 2000, 2000, 5000]))
        parts.append(np.full(ln, sym, dtype=np.uint8))
    parts.append(np.full(3000, 5, dtype=np.uint8))  # a run to the text end
    return np.concatenate(parts)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("max_ctx", [caps_sa.MAX_CONTEXT, 3000])
def test_run_memo_builds_agree(cuda, order, max_ctx):
    text = _runs_text(40 + order)
    n = len(text)
    rng = np.random.default_rng(9)
    cuts = np.unique(rng.integers(1, n, 120))
    ends = np.append(cuts, n).astype(np.uint64)
    lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
    ctx = O.Ctx(text, ends, order)
    pos = np.flatnonzero(text < 4).astype(np.uint64)
    want, want_lcp = ctx.build_in_memory(pos, max_ctx=max_ctx, threads=8, want_lcp=True)
    if order == 0:
        want = ctx.canonicalize(want, max_ctx)
    d_text = dev.text_to_device(text)
    d_pos = torch.from_numpy(pos.view(np.int64)).cuda()
    for flags in (0, NO):
        print(f"case runs n={n} order={order} max_ctx={max_ctx} flags={flags}", flush=True)
        sa, lcp = dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=len(pos), idx_bytes=8, lp=lp,
                                   max_context=max_ctx, want_lcp=True, flags=flags)
        torch.cuda.synchronize()
        st = caps_sa.last_stats()
        print(f"  runs={st['runs']} mid={st['mid_groups']} giant={st['giant_groups']}", flush=True)
        if flags == 0:
            assert st["runs"] > 0 and st["mid_groups"] + st["giant_groups"] > 0
        assert np.array_equal(sa.cpu().numpy().view(np.uint64), want)
        assert np.array_equal(lcp.cpu().numpy().view(np.uint64), want_lcp)


def test_run_memo_all_suffixes_plain(cuda):
    """Every suffix of a plain text made of long runs (giant groups of
    suffixes inside runs), u32 index."""
    text = _runs_text(7, 300)
    want, want_lcp = O.Ctx(text).build_in_memory(threads=8, want_lcp=True)
    d_text = dev.text_to_device(text)
    for flags in (0, NO, caps_sa.FLAG_GRAPH):
        sa, lcp = dev.build_device(d_text, len(text), 1, idx_bytes=4, want_lcp=True, flags=flags)
        torch.cuda.synchronize()
        assert np.array_equal(sa.cpu().numpy().view(np.uint32).astype(np.uint64), want), flags
        assert np.array_equal(lcp.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp), flags


def test_run_memo_small_groups(cuda):
    """Groups of 3..64 tied suffixes whose shared stretch is a long run of
    one code, of nearly equal lengths (the small-group fast-forward's run
    skip), beside the groups of suffixes inside the runs (65+ members, so the
    memo is built): memo on, memo off and the oracle agree, every suffix, in
    both boundary orders and under max_context."""
    rng = np.random.default_rng(77)
    parts = []
    for c in [3, 4, 5, 7, 8, 9, 16, 17, 32, 33, 64]:
        prefix = rng.integers(0, 4, 6).astype(np.uint8)
        ln = int(rng.choice([300, 700, 1500]))
        sym = int(rng.integers(0, 6))
        for _ in range(c):
            parts.append(rng.integers(0, 4, int(rng.integers(1, 9))).astype(np.uint8))
            parts.append(prefix)
            parts.append(np.full(ln + int(rng.integers(0, 3)), sym, dtype=np.uint8))
            parts.append(rng.integers(0, 4, 5).astype(np.uint8))
    text = np.concatenate(parts)
    n = len(text)
    d_text = dev.text_to_device(text)
    ends = np.unique(np.append(rng.integers(1, n, 40), n)).astype(np.uint64)
    for order, max_ctx, segmented in ((0, caps_sa.MAX_CONTEXT, False), (0, 1000, False), (1, caps_sa.MAX_CONTEXT, True)):
        lp = (caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)) if segmented else None
        ctx = O.Ctx(text, ends if segmented else None, order)
        want, want_lcp = ctx.build_in_memory(max_ctx=max_ctx, threads=8, want_lcp=True)
        for flags in (0, NO):
            print(f"case small-group runs n={n} order={order} max_ctx={max_ctx} flags={flags}", flush=True)
            sa, lcp = dev.build_device(d_text, n, 1, idx_bytes=4, lp=lp, max_context=max_ctx, want_lcp=True,
                                       flags=flags)
            torch.cuda.synchronize()
            st = caps_sa.last_stats()
            if flags == 0:
                assert st["runs"] > 0 and st["small_groups"] > 0
            assert np.array_equal(sa.cpu().numpy().view(np.uint32).astype(np.uint64), want)
            assert np.array_equal(lcp.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp)
