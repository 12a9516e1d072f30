"""GPU build vs the oracle: bit-exact suffix arrays and LCP arrays.

The oracle is the C restatement of caps-sa's build_in_memory (merge sort
with the LCP merge); for plain texts it is cross-checked against brute force
and prefix doubling in test_oracle.py.  Under SegmentedText with the default
boundary order, comparator-Equal suffixes are compared in ascending position
order (the oracle's order inside those classes is canonicalized first).
"""
import itertools

import numpy as np
import pytest

import caps_sa
import oracle as O
from caps_sa import gen

pytestmark = pytest.mark.gpu

MAX = caps_sa.MAX_CONTEXT


def _lp(n, ends, order):
    if ends is None:
        return None
    return caps_sa.StarSegmented(n, ends) if order == 1 else caps_sa.SegmentedText(n, ends)


def check(text, positions=None, ends=None, order=0, max_ctx=MAX, idx=np.uint32, with_lcp=True, label=""):
    text = np.ascontiguousarray(text)
    n = len(text)
    print(f"case {label} n={n} dtype={text.dtype} pos={'all' if positions is None else len(positions)} "
          f"ends={None if ends is None else len(ends)} order={order} max_ctx={max_ctx}", flush=True)
    ctx = O.Ctx(text, ends, order)
    pos = np.arange(n, dtype=np.uint64) if positions is None else np.asarray(positions, dtype=np.uint64)
    got, got_lcp = caps_sa.build_in_memory_for_positions_with(
        text, pos.astype(idx), _lp(n, ends, order), caps_sa.Opts(max_ctx), idx, return_lcp=True)
    want = ctx.build_in_memory(pos, max_ctx=max_ctx, threads=4)
    if ends is not None and order == 0:
        want = ctx.canonicalize(want, max_ctx)
    got = got.astype(np.uint64)
    if not np.array_equal(got, want):
        i = int(np.argmax(got != want))
        raise AssertionError(f"{label}: SA differs at {i}: got {got[max(0, i - 2):i + 3]} want {want[max(0, i - 2):i + 3]}")
    if with_lcp:
        want_lcp = ctx.lcp_array(want, max_ctx)
        got_lcp = got_lcp.astype(np.uint64)
        if not np.array_equal(got_lcp, want_lcp):
            i = int(np.argmax(got_lcp != want_lcp))
            raise AssertionError(f"{label}: LCP differs at {i}: got {got_lcp[i]} want {want_lcp[i]}")


@pytest.mark.parametrize("t", [b"banana", b"mississippi", b"abracadabra", b"a", b"aa", b"ab", b"ba",
                               bytes([7]), bytes([0, 1, 2, 0, 1, 5, 0, 2, 1, 6]), b"a" * 36, b"ab" * 14,
                               b"ACGTACGTACGTACGTACGTACGTACGT" * 100 + bytes([200])],
                         ids=lambda t: f"len{len(t)}-{t[:8].hex()}")
def test_fixtures(t):
    check(np.frombuffer(t, dtype=np.uint8), label=repr(t[:20]))


def test_empty():
    out = caps_sa.build_in_memory(np.zeros(0, dtype=np.uint8))
    assert out.size == 0


@pytest.mark.parametrize("sigma", [2, 4, 6, 255])
def test_random_u8(sigma):
    for n in [2, 3, 7, 16, 17, 63, 64, 65, 200, 1000, 5000]:
        check(gen.uniform_u8(n, sigma, seed=n * 31 + sigma), label=f"u8 sigma={sigma}")


@pytest.mark.parametrize("max_ctx", [0, 1, 2, 4, 16, 40])
def test_max_context(max_ctx):
    for n in [2, 3, 7, 64, 500, 3000]:
        check(gen.uniform_u8(n, 3, seed=n + max_ctx), max_ctx=max_ctx, label=f"max_ctx={max_ctx}")


def test_u16_random():
    for sigma in [3, 16, 300, 65535]:
        for n in [5, 100, 4000]:
            check(gen.uniform_u16(n, sigma, seed=sigma + n), label=f"u16 sigma={sigma}")


def test_u16_zipf():
    check(gen.zipf_u16(20000, seed=3), label="u16 zipf")


def test_long_runs_and_periodic():
    check(np.full(2000, 7, dtype=np.uint8), label="run 2000")
    check((np.arange(2000) % 3).astype(np.uint8), label="periodic 3")
    mixed = np.concatenate([(np.arange(500) % 4), np.full(1500, 4), (np.arange(500) % 4)]).astype(np.uint8)
    check(mixed, label="mixed run")


def test_giant_groups():
    # 9000 suffixes share their first 32 symbols: exercises chunked
    # refinement and the merge-path LCP merge.
    check(np.zeros(9000, dtype=np.uint8), label="zeros 9000")
    check((np.arange(12000) % 2).astype(np.uint8), label="period 2, 12000")
    t = gen.uniform_u8(6000, 4, seed=5)
    check(np.tile(t[:700], 12), label="tandem 700x12")


def test_positions_subsets():
    rng = np.random.default_rng(0xFEED)
    for n in [33, 200, 1000, 20000]:
        t = gen.uniform_u8(n, 6, seed=n)
        pos = np.flatnonzero(rng.integers(0, 10, n) < 7).astype(np.uint64)
        rng.shuffle(pos)
        check(t, positions=pos, label="subset")
    t = np.frombuffer(b"mississippi", dtype=np.uint8)
    check(t, positions=np.arange(0, 11, 2), label="even positions")


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("max_ctx", [MAX, 9])
def test_dense_increasing_position_lists(order, max_ctx):
    """A strictly increasing list covering at least half of the text is built
    through a kept-position bitmap; the same set shuffled, and a sparser
    list, stay on the position-list path.  Every form equals the oracle."""
    rng = np.random.default_rng(0xD5 + order)
    for n, keep in [(5000, 0.95), (20000, 0.6), (20000, 0.45), (64, 0.5)]:
        lengths = rng.integers(5, 400, max(1, n // 150))
        lengths = lengths[np.cumsum(lengths) <= n]
        ends = np.concatenate([np.cumsum(lengths), [n]]).astype(np.uint64)
        ends = np.unique(ends)
        t = rng.integers(0, 3, n).astype(np.uint8)
        pos = np.flatnonzero(rng.random(n) < keep).astype(np.uint64)
        check(t, positions=pos, ends=ends, order=order, max_ctx=max_ctx, label=f"increasing {keep}")
        shuffled = pos.copy()
        rng.shuffle(shuffled)
        check(t, positions=shuffled, ends=ends, order=order, max_ctx=max_ctx, label=f"shuffled {keep}")
    t = gen.markov_repeats(1 << 16, seed=12, mean_len=800.0)
    check(t, positions=np.arange(1, 1 << 16, dtype=np.uint64), label="all but the first")
    check(t, positions=np.arange(0, 1 << 16, 2, dtype=np.uint64), label="even positions")


def test_segmented_shorter_first():
    rng = np.random.default_rng(0x5E6)
    t = np.frombuffer(b"helloworldbananamississippi", dtype=np.uint8)
    check(t, ends=np.cumsum([5, 5, 6, 11]), label="seg small")
    for _ in range(20):
        lengths = rng.integers(5, 50, rng.integers(1, 10))
        n = int(lengths.sum())
        check(rng.integers(0, 3, n).astype(np.uint8), ends=np.cumsum(lengths), label="seg random")


def test_segmented_star():
    rng = np.random.default_rng(0xCAFE)
    check(np.frombuffer(b"AAAA", dtype=np.uint8), ends=[4], order=1, label="star AAAA")
    t = np.frombuffer(b"helloworldbananamississippi", dtype=np.uint8)
    check(t, ends=np.cumsum([5, 5, 6, 11]), order=1, label="star small")
    for _ in range(20):
        lengths = rng.integers(5, 50, rng.integers(1, 10))
        n = int(lengths.sum())
        check(rng.integers(0, 3, n).astype(np.uint8), ends=np.cumsum(lengths), order=1, label="star random")


def test_segmented_genome_like():
    text, ends, pos = gen.genome(200000, 80, median=2500.0, n_frac=0.058, n_mean=500.0, seed=11)
    check(text, positions=pos, ends=ends, label="genome shorter-first", with_lcp=False)
    check(text, positions=pos, ends=ends, order=1, label="genome star", with_lcp=False)


def test_u64_index():
    check(gen.uniform_u8(5000, 4, seed=9), idx=np.uint64, label="u64")
    check(gen.markov_repeats(30000, seed=2, mean_len=300.0), idx=np.uint64, label="u64 markov")


def test_c0_uniform_1m():
    """The 1,048,576-symbol uniform {0..3} text, u32 index, vs the oracle."""
    text = gen.uniform_u8(1 << 20, 4, seed=0xC0)
    ctx = O.Ctx(text)
    got, got_lcp = caps_sa.build_in_memory_with(text, return_lcp=True)
    want = ctx.build_in_memory(threads=8)
    assert np.array_equal(got.astype(np.uint64), want)
    assert np.array_equal(got_lcp.astype(np.uint64), ctx.kasai(want))


def test_markov_repeats_4m_properties():
    """Repeat-rich text: permutation + strictly increasing adjacent suffixes
    (which pins the unique SA) + Kasai LCPs."""
    text = gen.markov_repeats(1 << 22, seed=0xC1)
    ctx = O.Ctx(text)
    got, got_lcp = caps_sa.build_in_memory_with(text, return_lcp=True)
    got = got.astype(np.uint64)
    assert ctx.verify(got, strict=True, canon=True, threads=8) == -1
    assert np.array_equal(got_lcp.astype(np.uint64), ctx.kasai(got))


@pytest.mark.parametrize("offset", [1, 3, 5])
def test_unaligned_device_text(offset):
    """A device text that starts at an odd address: the word loads of the key
    gather and of the pair compare align on absolute addresses."""
    import torch
    from caps_sa import device as dev

    for text, w in ((gen.markov_repeats(20000, seed=offset, mean_len=400.0), 1),
                    (gen.zipf_u16(9000, 300, 1.1, seed=offset), 2)):
        n = len(text)
        raw = np.zeros((n + 8) * w, dtype=np.uint8)
        raw[offset * w: offset * w + n * w] = text.view(np.uint8)
        d_raw = torch.from_numpy(raw).to("cuda:0")
        d_view = d_raw[offset * w:]
        assert d_view.data_ptr() % 8 != 0
        sa, lcp = dev.build_device(d_view, n, w, idx_bytes=4, want_lcp=True)
        got = sa.cpu().numpy().view(np.uint32).astype(np.uint64)
        ctx = O.Ctx(text)
        want, want_lcp = ctx.build_in_memory(threads=4, want_lcp=True)
        assert np.array_equal(got, want)
        assert np.array_equal(lcp.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp)


def test_persistent_refinement_matches_grid():
    """Work-queue refinement (persistent slots) gives the same SA and LCP as
    the grid launch, on texts with every small-group class."""
    import torch
    from caps_sa import device as dev

    for text in (gen.markov_repeats(300000, seed=31, mean_len=700.0), gen.zipf_u16(60000, 40, 1.1, seed=3),
                 np.frombuffer(b"ab" * 3000 + b"abc" * 2000, dtype=np.uint8).copy()):
        w = text.dtype.itemsize
        d_text = dev.text_to_device(text)
        a, la = dev.build_device(d_text, len(text), w, want_lcp=True)
        b, lb = dev.build_device(d_text, len(text), w, want_lcp=True, flags=caps_sa.FLAG_PERSISTENT_REFINE)
        torch.cuda.synchronize()
        assert torch.equal(a, b) and torch.equal(la, lb)
        want, want_lcp = O.Ctx(text).build_in_memory(threads=4, want_lcp=True)
        assert np.array_equal(b.cpu().numpy().view(np.uint32).astype(np.uint64), want)
        assert np.array_equal(lb.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp)


def test_radix_tile_sort_matches_merge():
    """Both tile sorts (merge path, 4-bit radix), with a u32 and a u64 index,
    give the oracle's SA and LCP, including tiles with all-ones keys next to
    the pads."""
    import torch
    from caps_sa import device as dev

    star_text = np.full(9000, 255, dtype=np.uint8)
    star_text[::7] = 3
    for text in (gen.markov_repeats(200000, seed=41, mean_len=500.0), star_text,
                 gen.uniform_u8(70000, 256, seed=4)):
        d_text = dev.text_to_device(text)
        want, want_lcp = O.Ctx(text).build_in_memory(threads=4, want_lcp=True)
        for (idx_bytes, view), flags in itertools.product(((4, np.uint32), (8, np.uint64)),
                                                          (0, caps_sa.FLAG_TILE_RADIX)):
            sa, lcp = dev.build_device(d_text, len(text), 1, idx_bytes=idx_bytes, want_lcp=True, flags=flags)
            assert np.array_equal(sa.cpu().numpy().view(view).astype(np.uint64), want), (idx_bytes, flags)
            assert np.array_equal(lcp.cpu().numpy().view(view).astype(np.uint64), want_lcp), (idx_bytes, flags)


def test_full_tiles_both_index_widths_both_tile_sorts():
    """1.5M suffixes in buckets of about 1,536 records: 977 classification
    intervals of one bucket each (at least 512 intervals, 20 bytes of
    splitter and cursor tables), so scatter pass 2 stages 8,192-record tiles
    for the u32 and the u64 index alike, as (key, index) records for the
    merge tile sort and as key and index arrays for the radix one.  Every
    build gives the oracle's SA and LCP."""
    import torch
    from caps_sa import device as dev

    text = gen.markov_repeats(1500000, seed=43, mean_len=500.0)
    d_text = dev.text_to_device(text)
    want, want_lcp = O.Ctx(text).build_in_memory(threads=4, want_lcp=True)
    for idx_bytes, view in ((4, np.uint32), (8, np.uint64)):
        for flags in (0, caps_sa.FLAG_TILE_RADIX):
            print(f"idx_bytes={idx_bytes} flags={flags}", flush=True)
            sa, lcp = dev.build_device(d_text, len(text), 1, idx_bytes=idx_bytes, want_lcp=True, flags=flags)
            assert np.array_equal(sa.cpu().numpy().view(view).astype(np.uint64), want), (idx_bytes, flags)
            assert np.array_equal(lcp.cpu().numpy().view(view).astype(np.uint64), want_lcp), (idx_bytes, flags)


def test_autotune_then_build_is_exact():
    """The autotuner times the candidate knobs on the input and returns one of
    them; builds with the tuned knobs (and with every candidate it timed) give
    the oracle's suffix array."""
    import torch
    from caps_sa import device as dev

    text = gen.markov_repeats(600000, seed=51, mean_len=700.0)
    d_text = dev.text_to_device(text)
    choice, report = dev.autotune(d_text, len(text), 1, reps=1)
    assert 0 < len(report["tried"]) <= caps_sa.TUNE_MAX
    assert report["best_ms"] == min(t["ms"] for t in report["tried"])
    assert any(t["tile_keys"] == choice["tile_keys"] and t["flags"] == choice["flags"] for t in report["tried"])
    want = O.Ctx(text).build_in_memory(threads=4)
    for t in report["tried"] + [choice]:
        sa, _ = dev.build_device(d_text, len(text), 1, tile_keys=t["tile_keys"], flags=t["flags"])
        assert np.array_equal(sa.cpu().numpy().view(np.uint32).astype(np.uint64), want), t


def _copies_text(counts, stretch, seed, sigma=4, dtype=np.uint8, run=None):
    """Units of `stretch` symbols, each copied counts[i] times with copies cut
    to slightly different lengths and separated by short random context, so
    tie groups of every small size share long stretches and part at
    different depths.  `run` makes every unit a run of that one symbol."""
    rng = np.random.default_rng(seed)
    parts = []
    for c in counts:
        unit = np.full(stretch, run) if run is not None else rng.integers(0, sigma, stretch)
        for _ in range(c):
            parts.append(rng.integers(0, sigma, int(rng.integers(1, 40))))
            parts.append(unit[: stretch - int(rng.integers(0, 60))])
    return np.concatenate(parts).astype(dtype)


def test_shared_stretches_small_groups():
    """Groups of 3..64 members that share hundreds of symbols (the
    refinement's fast-forward over whole code words), in every small class,
    with windows (segment ends, max_context) ending inside the stretches."""
    counts = [3, 4, 5, 8, 9, 16, 17, 32, 33, 64]
    t = _copies_text(counts, 700, seed=71)
    check(t, label="copies u8")
    check(_copies_text(counts, 900, seed=72, sigma=6, run=5), label="runs u8")
    for mx in (50, 333):
        check(t, max_ctx=mx, label=f"copies max_ctx={mx}")
    rng = np.random.default_rng(73)
    ends = np.unique(np.concatenate([np.sort(rng.integers(1, len(t), 60)), [len(t)]]))
    check(t, ends=ends, label="copies segmented", with_lcp=False)
    check(t, ends=ends, order=1, label="copies star", with_lcp=False)
    pos = np.flatnonzero(rng.integers(0, 4, len(t)) > 0).astype(np.uint64)
    check(t, positions=pos, idx=np.uint64, label="copies positions u64")
    check(_copies_text(counts, 500, seed=74, sigma=300, dtype=np.uint16), label="copies u16")


def test_shared_stretches_persistent():
    """The same texts with every small class on the work-queue kernels."""
    import torch
    from caps_sa import device as dev

    for text in (_copies_text([3, 4, 5, 8, 9, 16, 17, 32, 33, 64], 700, seed=75),
                 _copies_text([3, 7, 12, 40], 600, seed=76, sigma=300, dtype=np.uint16)):
        w = text.dtype.itemsize
        d_text = dev.text_to_device(text)
        sa, lcp = dev.build_device(d_text, len(text), w, want_lcp=True, flags=caps_sa.FLAG_PERSISTENT_REFINE)
        want, want_lcp = O.Ctx(text).build_in_memory(threads=4, want_lcp=True)
        assert np.array_equal(sa.cpu().numpy().view(np.uint32).astype(np.uint64), want)
        assert np.array_equal(lcp.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp)


def test_mid_groups_every_lds_class():
    """Tie groups of 65..2048 members in each LDS class of the mid-size
    refinement (up to 128, 129..512 and 513..2048 members): a unit copied c
    times gives groups of exactly c members at its first offsets and smaller
    ones further in, with windows (max_context, segment ends) ending inside
    the shared stretches."""
    from caps_sa import device as dev

    counts = [65, 100, 128, 129, 300, 512, 513, 1000, 2000]
    t = _copies_text(counts, 160, seed=81)
    t16 = _copies_text(counts, 160, seed=82, sigma=300, dtype=np.uint16)
    check(t, label="mid copies u8")
    check(t, max_ctx=90, label="mid copies max_ctx=90")
    rng = np.random.default_rng(83)
    ends = np.unique(np.concatenate([np.sort(rng.integers(1, len(t), 300)), [len(t)]]))
    check(t, ends=ends, label="mid copies segmented", with_lcp=False)
    check(t, ends=ends, order=1, label="mid copies star", with_lcp=False)
    check(t16, label="mid copies u16")
    # the same groups through the three launches by size (a build with
    # thousands of such groups takes them by itself)
    cases = [(t, None, 0, MAX), (t, None, 0, 90), (t16, None, 0, MAX), (t, ends, 1, MAX), (t, ends, 0, MAX)]
    for text, e, order, mx in cases:
        print(f"case by-size n={len(text)} {text.dtype} ends={e is not None} order={order} max_ctx={mx}", flush=True)
        d_text = dev.text_to_device(text)
        sa, lcp = dev.build_device(d_text, len(text), text.dtype.itemsize, lp=_lp(len(text), e, order),
                                   max_context=mx, want_lcp=True, flags=caps_sa.FLAG_MID_BY_SIZE)
        ctx = O.Ctx(text, e, order)
        want = ctx.build_in_memory(max_ctx=mx, threads=4)
        if e is not None and order == 0:
            want = ctx.canonicalize(want, mx)
        assert np.array_equal(sa.cpu().numpy().view(np.uint32).astype(np.uint64), want)
        if e is None:
            want_lcp = ctx.lcp_array(want, mx)
            assert np.array_equal(lcp.cpu().numpy().view(np.uint32).astype(np.uint64), want_lcp)


def test_tight_and_pow2_seed_fields_agree():
    """Seed keys with ceil(log2(sigma))-bit fields (narrowed from the packed
    text's power-of-two fields) and with power-of-two fields give the same
    SA and LCP, equal to the oracle's: 3-, 5-, 7-bit u8 alphabets, 10- and
    13-bit u16 alphabets, plain, segmented and STAR, with max_context."""
    import torch
    from caps_sa import device as dev

    rng = np.random.default_rng(0x71)
    cases = []
    for sigma, dt in ((6, np.uint8), (20, np.uint8), (100, np.uint8), (1000, np.uint16), (5000, np.uint16)):
        t = rng.integers(0, sigma, 30000).astype(dt)
        t[5000:9000] = t[15000:19000]  # a long repeat: deep groups
        cases.append(t)
    for t in cases:
        w = t.dtype.itemsize
        n = len(t)
        d_text = dev.text_to_device(t)
        lengths = rng.integers(1, 400, 200)
        lengths = lengths[np.cumsum(lengths) < n]
        ends = np.append(np.cumsum(lengths), n).astype(np.uint64)
        for order, lp in ((0, None), (0, caps_sa.SegmentedText(n, ends)), (1, caps_sa.StarSegmented(n, ends))):
            for mx in (MAX, 37):
                ctx = O.Ctx(t, None if lp is None else ends, order)
                want = ctx.build_in_memory(max_ctx=mx, threads=4)
                if lp is not None and order == 0:
                    want = ctx.canonicalize(want, mx)
                outs = []
                for flags in (0, caps_sa.FLAG_POW2_SEED):
                    sa, lcp = dev.build_device(d_text, n, w, lp=lp, want_lcp=True, max_context=mx, flags=flags)
                    outs.append((sa.cpu().numpy().view(np.uint32).astype(np.uint64), lcp.cpu().numpy()))
                label = (t.dtype, int(t.max()), order, lp is not None, mx)
                assert np.array_equal(outs[0][0], want), label
                assert np.array_equal(outs[1][0], want), label
                assert np.array_equal(outs[0][1], outs[1][1]), label


def test_scatter_pass1_variants_agree(cuda):
    """Scatter pass 1 staged in LDS (default for every suffix, u32 index, up
    to 1,024 classification buckets), written record by record
    (CAPS_FLAG_SCATTER_DIRECT), and with 4,096 classification buckets: the
    same suffix array as the oracle's."""
    from caps_sa import device as dev
    import torch

    text = gen.markov_repeats(3_000_017, seed=21, mean_len=400.0)
    want = O.Ctx(text).build_in_memory(threads=8)
    d_text = dev.text_to_device(text)
    for label, kw in (("staged", {}), ("direct", {"flags": caps_sa.FLAG_SCATTER_DIRECT}),
                      ("4096 buckets", {"gpu_buckets": 4096}), ("u64 index", {"idx_bytes": 8})):
        print(f"case scatter pass 1 {label}", flush=True)
        idx = kw.pop("idx_bytes", 4)
        sa, _ = dev.build_device(d_text, len(text), 1, idx_bytes=idx, **kw)
        torch.cuda.synchronize()
        got = sa.cpu().numpy().view(np.uint32 if idx == 4 else np.uint64).astype(np.uint64)
        assert np.array_equal(got, want), label
