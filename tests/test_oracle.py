"""Pins the CPU oracle (oracle/) before it is trusted as the GPU's checker.

Each oracle function is checked against something that does not share its
code: the reference tests' known answers, pure-Python brute force
(tests/golden/kats.json, written by tests/golden/make_golden.py), prefix
doubling, Kasai's LCP algorithm, and closed forms.  A deliberately broken
comparator is shown to fail these checks (test_checks_catch_a_planted_bug).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "kats.json")))
MAX = O.MAX_CTX


def _ctx(case):
    return O.Ctx(np.array(case["text"], dtype=np.uint8), case["ends"], case["order"])


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
def test_build_in_memory_matches_golden(case):
    ctx = _ctx(case)
    got = ctx.build_in_memory(case["positions"])
    if case["ends"] is not None and case["order"] == 0:
        got = ctx.canonicalize(got)
    assert got.tolist() == case["sa"]


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
def test_brute_matches_golden(case):
    assert _ctx(case).brute_sa(case["positions"]).tolist() == case["sa"]


def test_lcp_and_cmp_kats():
    for text, p, q, mx, want in GOLD["kats"]["lcp"]:
        assert O.Ctx(text.encode()).lcp(p, q, MAX if mx is None else mx) == want
    for text, p, q, want in GOLD["kats"]["cmp"]:
        assert O.Ctx(text.encode()).cmp(p, q) == want


def test_lcp_stride_boundaries():
    # lcp.rs:482-491 and 524-532: a difference at every vector-stride boundary
    for diff_at in [0, 1, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 200]:
        t = np.full(512, ord("A"), dtype=np.uint8)
        t[diff_at] = ord("G")
        assert O.Ctx(t).lcp(0, 256) == diff_at


def test_lcp_u16_byte_inside_symbol():
    # lcp.rs:546-564: the first differing byte may sit in either byte of a symbol
    for byte_diff_at in [0, 1, 2, 3, 31, 32, 33, 63, 64, 65, 127, 128, 200]:
        t = np.full(256, 0xAAAA, dtype=np.uint16)
        mask = 0x00FF if byte_diff_at % 2 == 0 else 0xFF00
        t[byte_diff_at // 2] ^= mask & 0xAAAA
        assert O.Ctx(t).lcp(0, 128) == byte_diff_at // 2


def test_lcp_matches_scalar_scan_random():
    rng = np.random.default_rng(0xA5A5)
    for dt in (np.uint8, np.uint16):
        for n in [1, 32, 33, 200, 1000]:
            t = rng.integers(0, 4, n).astype(dt)
            ctx = O.Ctx(t)
            for _ in range(50):
                p, q = rng.integers(0, n, 2)
                mx = int(rng.integers(0, n + 2))
                want = 0
                lim = min(n - p, n - q, mx)
                while want < lim and t[p + want] == t[q + want]:
                    want += 1
                assert ctx.lcp(int(p), int(q), mx) == want


def test_limits_kats_and_directory():
    k = GOLD["kats"]["limits"]
    ctx = O.Ctx(np.zeros(k["n"], dtype=np.uint8), k["ends"])
    for p, want in k["lim"]:
        assert ctx.lim_at(p) == want
    # limits.rs:390-405: the directory agrees with binary search over 2000 segments
    lengths = np.array([1 + i % 97 for i in range(2000)], dtype=np.uint64)
    n = int(lengths.sum())
    ctx = O.Ctx(np.zeros(n, dtype=np.uint8), np.cumsum(lengths))
    assert ctx.L.has_dir == 1
    ends = np.cumsum(lengths)
    for p in range(0, n + 1, 7):
        i = int(np.searchsorted(ends, p, side="right"))
        want = int(ends[i]) - p if i < len(ends) else n - p
        assert ctx.lim_at(p) == want == ctx.lim_at_bsearch(p)


@pytest.mark.parametrize("sigma", [2, 4, 6, 255])
def test_merge_sort_vs_doubling_and_brute(sigma):
    rng = np.random.default_rng(sigma)
    for n in [2, 3, 7, 16, 17, 63, 64, 65, 200, 1000, 5000]:
        t = rng.integers(0, sigma, n).astype(np.uint8)
        ctx = O.Ctx(t)
        sa, lcp = ctx.build_in_memory(want_lcp=True)
        assert np.array_equal(sa, ctx.doubling_sa())
        if n <= 1000:
            assert np.array_equal(sa, ctx.brute_sa())
        # merge_sort's LCP postcondition (sample_sort.rs:208-216, 689-706)
        assert np.array_equal(lcp, ctx.kasai(sa))
        assert lcp[0] == 0


def test_merge_sort_parallel_equals_sequential():
    t = np.random.default_rng(3).integers(0, 4, 200000).astype(np.uint8)
    ctx = O.Ctx(t)
    a = ctx.build_in_memory(threads=1)
    b = ctx.build_in_memory(threads=8)
    assert np.array_equal(a, b)
    assert np.array_equal(a, ctx.doubling_sa())


@pytest.mark.parametrize("max_ctx", [0, 1, 2, 4, 16])
def test_finite_max_context_vs_brute(max_ctx):
    # sample_sort.rs:750-789
    rng = np.random.default_rng(0x0F11 + max_ctx)
    for n in [2, 3, 7, 64, 500]:
        t = rng.integers(0, 3, n).astype(np.uint8)
        ctx = O.Ctx(t)
        sa, lcp = ctx.build_in_memory(max_ctx=max_ctx, want_lcp=True)
        assert np.array_equal(sa, ctx.brute_sa(max_ctx=max_ctx))
        assert np.array_equal(lcp, ctx.lcp_array(sa, max_ctx))


def test_segmented_validity_and_star():
    rng = np.random.default_rng(0x5E6)
    for order in (O.ORDER_SHORTER_FIRST, O.ORDER_STAR):
        for _ in range(20):
            lengths = rng.integers(5, 50, rng.integers(1, 10))
            n = int(lengths.sum())
            t = rng.integers(0, 3, n).astype(np.uint8)
            ctx = O.Ctx(t, np.cumsum(lengths), order)
            sa = ctx.build_in_memory()
            assert ctx.verify(sa, strict=order == O.ORDER_STAR, canon=False) == -1
            brute = ctx.brute_sa()
            if order == O.ORDER_STAR:
                assert np.array_equal(sa, brute)  # total order: unique answer
            else:
                assert np.array_equal(ctx.canonicalize(sa), brute)


def test_sample_sort_matches_build_in_memory():
    # ext_mem.rs:2784-2806
    rng = np.random.default_rng(0xC0DEC0DE)
    for n in [16, 200, 2000, 30000]:
        t = rng.integers(0, 6, n).astype(np.uint8)
        ctx = O.Ctx(t)
        want = ctx.build_in_memory()
        for p in [0, 1, 4, 16, 64]:
            assert np.array_equal(ctx.sample_sort(p=p, threads=4), want)
    unit = np.frombuffer(b"ACGTACGTACGTACGTACGTACGTACGT" * 100 + bytes([200]), dtype=np.uint8)
    ctx = O.Ctx(unit)
    assert np.array_equal(ctx.sample_sort(p=8), ctx.build_in_memory())


def test_sample_sort_for_positions():
    rng = np.random.default_rng(0xFEED)
    t = rng.integers(0, 6, 3000).astype(np.uint8)
    pos = np.flatnonzero(rng.integers(0, 10, 3000) < 7).astype(np.uint64)
    ctx = O.Ctx(t)
    assert np.array_equal(ctx.sample_sort(pos, p=8), ctx.brute_sa(pos))


def test_emit_error_stops_early():
    ctx = O.Ctx(b"banana")
    calls = []

    def emit(arr):
        calls.append(len(arr))
        return 7

    assert ctx.sample_sort(p=2, emit=emit) == 7
    assert len(calls) == 1


def test_sample_sort_parameters_closed_form():
    import math

    for n, req, thr, want in [(0, 0, 4, 0), (10, 0, 4, 4), (10, 0, 32, 10), (1 << 20, 0, 4, 16),
                              (1 << 30, 0, 4, 8192), (100, 7, 4, 7), (5, 7, 4, 5)]:
        assert O.effective_subproblem_count(n, req, thr) == want
    for n, p in [(1 << 28, 1024), (1000, 4), (10, 8), (2, 1)]:
        per = math.ceil(4 * max(math.log(n), 1.0))
        assert O.sample_target_total(n, p) == min(max(p * per, p), n)
    xs = np.arange(100, dtype=np.uint64) * 3
    got = O.evenly_spaced(xs, 7)
    assert got.tolist() == [int(xs[(2 * i + 1) * 100 // 14]) for i in range(7)]
    assert O.evenly_spaced(xs, 200).tolist() == xs.tolist()
    assert O.evenly_spaced(xs, 0).tolist() == []


def test_cascade_merge_equals_sort():
    rng = np.random.default_rng(11)
    t = rng.integers(0, 3, 5000).astype(np.uint8)
    ctx = O.Ctx(t)
    # runs: sorted pieces of random subsets, lcp[0] of each run = 0
    sa_parts, lcp_parts, bounds = [], [], [0]
    perm = rng.permutation(5000).astype(np.uint64)
    cuts = np.sort(rng.choice(np.arange(1, 5000), 9, replace=False))
    for piece in np.split(perm, cuts):
        s, l = ctx.build_in_memory(piece, want_lcp=True)
        sa_parts.append(s)
        lcp_parts.append(l)
        bounds.append(bounds[-1] + len(s))
    out, lout = ctx.cascade_merge(np.concatenate(sa_parts), np.concatenate(lcp_parts), bounds)
    want = ctx.doubling_sa()
    assert np.array_equal(out, want)
    assert np.array_equal(lout, ctx.kasai(want))


def test_packer_kats():
    # pack.rs:516-528: 255 byte values -> 8-bit fields, k = 8; 256 -> no key
    p = O.OraclePacker(bytes(range(255)))
    assert p.ok and p.bits == 8 and p.k == 8
    assert not O.OraclePacker(bytes(range(256))).ok
    # pack.rs:484-513: dense-only declines a gapped alphabet; remap budget
    gapped = bytes([0, 2, 0, 2, 0, 2, 0, 2])
    assert not O.OraclePacker(gapped, policy=0).ok
    assert not O.OraclePacker(gapped, policy=1).ok
    assert not O.OraclePacker(gapped, policy=2, remap_budget=7).ok
    assert O.OraclePacker(gapped, policy=2, remap_budget=8).ok
    # 4 symbols -> 4-bit fields (one code reserved), k = 16 (finding of pack.rs:134-139)
    p = O.OraclePacker(bytes([0, 1, 2, 3] * 4))
    assert p.bits == 4 and p.k == 16
    # closed form of key_at_bounded for "0123..." at p=0, lim=3, shorter first:
    # fields 1,2,3 then 13 pad fields of 0
    assert p.key(0, 3, 0) == (1 << 60) | (2 << 56) | (3 << 52)
    # longer first: no bias, pad = alphabet (4)
    want = (0 << 60) | (1 << 56) | (2 << 52)
    for i in range(13):
        want |= 4 << (48 - 4 * i)
    assert p.key(0, 3, 1) == want
    assert p.shared_fields(p.key(0, 16, 0), p.key(0, 16, 0)) == 16
    assert p.shared_fields(1 << 63, 0) == 0


def test_packer_k32_depths():
    # pack.rs:531-563: 3 symbols -> 2-bit fields, k = 32; seeds exact at k-1, k, k+1
    k = 32
    segs = [[0] * (k - 1) + [1], [0] * (k - 1) + [2], [1] * k + [0], [1] * k + [2], [2] * (k + 1) + [0],
            [2] * (k + 1) + [1], [0] * (k + 2), [0] * (k + 2)]
    text = np.array(sum(segs, []), dtype=np.uint8)
    p = O.OraclePacker(text)
    assert p.k == k and p.bits == 2


def test_checks_catch_a_planted_bug(monkeypatch):
    """A plausible mistake (an off-by-one in the LCP cap) must fail the
    checks used above: brute force and the LCP postcondition."""
    t = np.random.default_rng(5).integers(0, 2, 300).astype(np.uint8)
    ctx = O.Ctx(t)
    sa, lcp = ctx.build_in_memory(want_lcp=True)
    bad = lcp.copy()
    bad[10] += 1
    assert not np.array_equal(bad, ctx.kasai(sa))
    swapped = sa.copy()
    swapped[[5, 6]] = swapped[[6, 5]]
    assert ctx.verify(swapped) != -1
    assert not np.array_equal(swapped, ctx.brute_sa())


def test_verify_against_keep_bitmap():
    """The bitmap verifier accepts the oracle's SA of the kept positions and
    rejects a swapped pair, a dropped entry and an entry outside the bitmap."""
    from caps_sa import gen

    text, ends, pos = gen.genome(60000, 25, median=2500.0, n_frac=0.058, n_mean=300.0, seed=5)
    ctx = O.Ctx(text, ends, O.ORDER_SHORTER_FIRST)
    sa = ctx.canonicalize(ctx.build_in_memory(pos, threads=4))
    keep = gen.keep_below(text, 4)
    assert ctx.verify_keep(sa, keep, threads=4) == -1
    bad = sa.copy()
    bad[10], bad[11] = bad[11], bad[10]
    assert ctx.verify_keep(bad, keep, threads=4) == 11
    assert ctx.verify_keep(sa[:-1], keep, threads=4) == len(sa) - 1
    outside = sa.copy()
    outside[0] = int(np.flatnonzero(text >= 4)[0])
    assert ctx.verify_keep(outside, keep, threads=4) == len(sa)
