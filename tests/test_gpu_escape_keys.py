"""32-symbol escape seed keys for genome text (CAPS_FLAG_SEED_ESCAPE32):
nucleotides 0..3 in 2-bit fields, a spacer or N (code >= 4) and every field
after it the largest field value.

Pinned against the oracle's comparator and LCP: a strict key inequality must
imply the same suffix order; two suffixes that part (or end) inside their
first 32 symbols before any escape must get different keys, ordered as the
suffixes; builds with the escape keys must give the oracle's suffix and LCP
arrays, on texts whose suffix pairs share k-1, k and k+1 symbols, end at
segment ends and run into N runs, in both boundary orders, through a position
list and through a kept-position bitmap.
"""
import numpy as np
import pytest
import torch

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu

E = caps_sa.FLAG_SEED_ESCAPE32
K = 32


def _genome_like(n, seed, n_runs=40, run_len=300, segs=60):
    rng = np.random.default_rng(seed)
    text = rng.integers(0, 4, n).astype(np.uint8)
    for _ in range(n_runs):  # N runs (5)
        s = int(rng.integers(0, n - run_len))
        text[s:s + int(rng.integers(1, run_len))] = 5
    unit = rng.integers(0, 4, 200).astype(np.uint8)  # shared stretches running into escapes
    for _ in range(30):
        s = int(rng.integers(0, n - 260))
        text[s:s + 200] = unit
        text[s + 200] = 5 if rng.random() < 0.5 else 4
    cuts = np.unique(rng.integers(1, n, segs))
    ends = np.append(cuts, n).astype(np.uint64)
    text[ends[:-1] - 1] = 4  # spacers closing the segments
    return text, ends


@pytest.mark.parametrize("order", [0, 1])
def test_escape_keys_order_and_strictness(cuda, order):
    text, ends = _genome_like(60_000, 3 + order)
    n = len(text)
    lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
    ctx = O.Ctx(text, ends, order)
    keys, (bits, k) = dev.seed_keys(dev.text_to_device(text), n, 1, np.arange(n, dtype=np.uint64), lp, geometry=2,
                                    rank=order)
    assert (bits, k) == (2, K)
    rng = np.random.default_rng(7)
    a = rng.integers(0, n, 20000)
    # neighbours in suffix order too: pairs that share long prefixes
    sa = ctx.build_in_memory(threads=8)
    if order == 0:
        sa = ctx.canonicalize(sa)
    pairs = list(zip(a.tolist(), rng.integers(0, n, 20000).tolist())) + list(zip(sa[:-1].tolist(), sa[1:].tolist()))
    lim = lambda p: ctx.lim_at(p)
    esc_at = lambda p, m: next((i for i in range(m) if text[p + i] >= 4), m)
    strict = 0
    for p, q in pairs:
        c = ctx.cmp(p, q)
        if keys[p] < keys[q]:
            assert c < 0, (order, p, q)
        elif keys[p] > keys[q]:
            assert c > 0, (order, p, q)
        h = ctx.naive_lcp(p, q)  # capped by both windows
        w = min(lim(p), lim(q))
        # parted (or one window ended) inside the key, before any escape of either
        if h < K and h < w and esc_at(p, h + 1) > h and esc_at(q, h + 1) > h:
            assert keys[p] != keys[q], (order, p, q, h)
            strict += 1
    assert strict > 1000


@pytest.mark.parametrize("order", [0, 1])
def test_escape_keys_depths_and_boundaries_exact(cuda, order):
    """Pairs sharing k-1, k and k+1 symbols before an N, a spacer or a
    segment end, and pairs identical through a simultaneous segment end."""
    segs = []
    for d in (K - 1, K, K + 1):
        for tail in ([5, 0], [4], [1, 2], [3, 3, 5]):
            body = [3] * d
            segs.append(body + tail)
            segs.append(body + [0] + tail)
            segs.append(body + [5] * 3)
    segs += [[1] * (K + 2)] * 3 + [[2] * (K - 1) + [5]] * 2
    text = np.array(sum(segs, []), dtype=np.uint8)
    ends = np.cumsum([len(s) for s in segs]).astype(np.uint64)
    n = len(text)
    lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
    ctx = O.Ctx(text, ends, order)
    want, want_lcp = ctx.build_in_memory(threads=4, want_lcp=True)
    if order == 0:
        want = ctx.canonicalize(want)
    d_text = dev.text_to_device(text)
    for idx in (4, 8):
        sa, lcp = dev.build_device(d_text, n, 1, idx_bytes=idx, lp=lp, want_lcp=True, flags=E)
        torch.cuda.synchronize()
        assert caps_sa.last_stats()["key_k"] == K
        dt = np.uint32 if idx == 4 else np.uint64
        assert np.array_equal(sa.cpu().numpy().view(dt).astype(np.uint64), want)
        assert np.array_equal(lcp.cpu().numpy().view(dt).astype(np.uint64), want_lcp)


def test_escape_keys_genome_builds(cuda):
    """Genome-like texts (lognormal segments ending in a spacer, N runs):
    the ACGT positions through a position list and through a bitmap, with
    the escape keys against the default keys and the oracle."""
    n = 3_000_000
    text, ends, pos = gen.genome(n, 600, median=2500.0, n_frac=0.058, n_mean=2000.0, seed=11)
    lp = caps_sa.SegmentedText(n, ends)
    ctx = O.Ctx(text, ends, O.ORDER_SHORTER_FIRST)
    want, want_lcp = ctx.build_in_memory(pos, threads=8, want_lcp=True)
    want = ctx.canonicalize(want)
    d_text = dev.text_to_device(text)
    d_pos = torch.from_numpy(pos.view(np.int64)).cuda()
    for flags in (E, 0):
        print(f"case genome positions flags={flags}", flush=True)
        sa, lcp = dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=len(pos), idx_bytes=8, lp=lp,
                                   want_lcp=True, flags=flags)
        torch.cuda.synchronize()
        st = caps_sa.last_stats()
        print(f"  key_k={st['key_k']} pairs={st['pairs']} small={st['small_groups']} mid={st['mid_groups']} "
              f"giant={st['giant_groups']} refine={st['ms_refine'] + st['ms_giant']:.2f} ms", flush=True)
        assert np.array_equal(sa.cpu().numpy().view(np.uint64), want)
        assert np.array_equal(lcp.cpu().numpy().view(np.uint64), want_lcp)
    keep = gen.keep_below(text, 4)
    d_keep = torch.from_numpy(keep.view(np.int64)).cuda()
    sa, _, kept = dev.build_device_filtered(d_text, n, 1, d_keep, len(pos), idx_bytes=8, lp=lp, flags=E)
    torch.cuda.synchronize()
    assert kept == len(pos)
    assert np.array_equal(sa.cpu().numpy().view(np.uint64), want)
