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


# ---- seed_subarray (pack.rs:241-321) and upper_bound_by_pivot
# (ext_mem.rs:1705-1734), pinned by a pure-Python comparator ---------------

def _lcg(state):
    state = (state * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
    return state, state >> 33


def _lcg_text(n, mod, seed, spacer_every=0, spacer_run=0):
    out, st = [], seed
    for i in range(n):
        if spacer_every and i % spacer_every < spacer_run:
            out.append(5)
        else:
            st, v = _lcg(st)
            out.append(v % mod)
    return np.array(out, dtype=np.uint8)


def _spacer_ends(text):
    """One past each maximal non-spacer run, closing at the text length
    (pack.rs:567-584)."""
    ends, in_run = [], False
    for i, b in enumerate(text.tolist()):
        if b == 5:
            if in_run:
                ends.append(i)
                in_run = False
        else:
            in_run = True
    if not ends or ends[-1] != len(text):
        ends.append(len(text))
    return ends


class _Brute:
    """Suffix order by slicing (sample_sort.rs:643-647), limits from the ends
    by bisection, ties by the default boundary order (limits.rs:66-75) or
    STAR's lim_b.cmp(lim_a).then(pos) (pack.rs:337-346)."""

    def __init__(self, text, ends, order):
        import bisect

        self.t = text.tolist()
        self.ends = list(ends) if ends is not None else [len(text)]
        self.order = order
        self._bis = bisect.bisect_right

    def lim(self, p):
        return self.ends[self._bis(self.ends, p)] - p

    def lcp(self, a, b):
        cap = min(self.lim(a), self.lim(b))
        i = 0
        while i < cap and self.t[a + i] == self.t[b + i]:
            i += 1
        return i

    def cmp(self, a, b):
        la, lb = self.lim(a), self.lim(b)
        h = self.lcp(a, b)
        if h < min(la, lb):
            return -1 if self.t[a + h] < self.t[b + h] else 1
        if self.order == O.ORDER_SHORTER_FIRST:
            return (la > lb) - (la < lb)
        return ((lb > la) - (lb < la)) or ((a > b) - (a < b))


def _check_seed(text, ends, order, positions, policy=1, budget=0):
    ctx = O.Ctx(text, ends, order)
    pk = O.OraclePacker(text, policy=policy, remap_budget=budget)
    assert pk.ok
    rank = 0 if order == O.ORDER_SHORTER_FIRST else 1
    sa, lcp = ctx.seed_subarray(pk, rank, positions)
    br = _Brute(text, ends, order)
    assert sorted(sa.tolist()) == sorted(int(x) for x in positions)
    assert lcp[0] == 0
    for i in range(1, len(sa)):
        a, b = int(sa[i - 1]), int(sa[i])
        c = br.cmp(a, b)
        assert c < 0 if order == O.ORDER_STAR else c <= 0, (i, a, b)
        assert int(lcp[i]) == br.lcp(a, b), (i, a, b)
    return sa, lcp, pk, br


def test_seed_subarray_plain_lcg_text_equals_brute_force():
    # pack.rs:433-439: 4096 LCG symbols over {0..3}, every position
    text = _lcg_text(4096, 4, 12345)
    sa, lcp, _, _ = _check_seed(text, None, O.ORDER_SHORTER_FIRST, np.arange(len(text)))
    ctx = O.Ctx(text)
    assert np.array_equal(sa, ctx.brute_sa())
    assert np.array_equal(lcp, ctx.kasai(sa))


def test_seed_subarray_segmented_lcg_texts_both_orders():
    # pack.rs:441-481: spacer runs every 617 / 411 symbols, ACGT positions
    for seed, every, run, order in ((999, 617, 9, O.ORDER_SHORTER_FIRST), (4242, 411, 7, O.ORDER_STAR)):
        text = _lcg_text(8192, 5, seed, every, run)
        ends = _spacer_ends(text)
        acgt = np.flatnonzero(text < 4)
        _check_seed(text, ends, order, acgt)


def test_seed_subarray_remapped_gapped_alphabet():
    # pack.rs:484-513: an exactly budgeted remap orders the gapped text
    text = np.array([0, 2, 0, 2, 0, 2, 0, 2], dtype=np.uint8)
    sa, _, _, _ = _check_seed(text, None, O.ORDER_SHORTER_FIRST, np.arange(8), policy=2, budget=8)
    assert sa.tolist() == O.Ctx(text).brute_sa().tolist()


def test_packer_k32_depths():
    """pack.rs:531-563: pairs sharing k-1, k and k+1 symbols and a pair
    identical through a simultaneous segment end; 3 symbols -> 2-bit fields,
    k = 32.  Keys of each pair share exactly min(true LCP, k) fields, key
    order is suffix order on the first k symbols, and the seed's adjacent
    LCPs are the true (limit-capped) ones under both boundary orders."""
    k = 32
    segs = [[0] * (k - 1) + [1], [0] * (k - 1) + [2], [1] * k + [0], [1] * k + [2], [2] * (k + 1) + [0],
            [2] * (k + 1) + [1], [0] * (k + 2), [0] * (k + 2)]
    text = np.array(sum(segs, []), dtype=np.uint8)
    ends = np.cumsum([len(s) for s in segs]).tolist()
    starts = [0] + ends[:-1]
    p = O.OraclePacker(text)
    assert p.k == k and p.bits == 2
    for order, rank in ((O.ORDER_SHORTER_FIRST, 0), (O.ORDER_STAR, 1)):
        br = _Brute(text, ends, order)
        keys = [p.key(s, br.lim(s), rank) for s in starts]
        for j in range(0, len(starts), 2):
            a, b = starts[j], starts[j + 1]
            true_lcp = br.lcp(a, b)
            assert true_lcp == [k - 1, k, k + 1, k + 2][j // 2]
            assert p.shared_fields(keys[j], keys[j + 1]) == min(true_lcp, k)
            if true_lcp < k:  # the keys decide the order on their own
                assert (keys[j] < keys[j + 1]) == (br.cmp(a, b) < 0)
        sa, lcp, _, _ = _check_seed(text, ends, order, np.array(starts))
        # depths k-1, k, k+1: the pairs stay adjacent with their true LCP
        for depth in (k - 1, k, k + 1):
            assert depth in lcp.tolist()


def test_upper_bound_by_pivot_semantics():
    """First index whose suffix is strictly greater than the pivot: a pivot
    inside the run lands one past itself (a lower bound would land on it),
    and a run of comparator-Equal suffixes is passed over whole."""
    text = _lcg_text(3000, 3, 77)
    ctx = O.Ctx(text)
    br = _Brute(text, None, O.ORDER_SHORTER_FIRST)
    run = ctx.brute_sa(np.arange(0, 3000, 3, dtype=np.uint64))
    for idx in (0, 1, 500, len(run) - 1):
        pivot = int(run[idx])
        assert ctx.upper_bound_by_pivot(run, pivot) == idx + 1
    for pivot in (1, 2, 1000, 2999):  # pivots outside the run
        want = sum(1 for r in run.tolist() if br.cmp(r, pivot) <= 0)
        assert ctx.upper_bound_by_pivot(run, pivot) == want
    assert ctx.upper_bound_by_pivot(run[:0], 5) == 0
    # segmented ShorterFirst: equal segments -> Equal suffixes at the same
    # offset; the bound passes every one of them
    segs = [[1, 0, 2, 1]] * 5 + [[1, 0, 2]] + [[1, 0, 2, 1, 1]]
    t2 = np.array(sum(segs, []), dtype=np.uint8)
    e2 = np.cumsum([len(s) for s in segs]).tolist()
    c2 = O.Ctx(t2, e2, O.ORDER_SHORTER_FIRST)
    b2 = _Brute(t2, e2, O.ORDER_SHORTER_FIRST)
    run2 = c2.canonicalize(c2.build_in_memory())
    starts = [0] + e2[:-1]
    for pivot in starts:
        want = sum(1 for r in run2.tolist() if b2.cmp(r, pivot) <= 0)
        lower = sum(1 for r in run2.tolist() if b2.cmp(r, pivot) < 0)
        got = c2.upper_bound_by_pivot(run2, pivot)
        assert got == want
        if pivot < 20:  # the five equal segments: Equal class of five
            assert want - lower == 5
