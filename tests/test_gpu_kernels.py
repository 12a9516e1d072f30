"""Kernel-level parity on the GPU against the oracle:
  - 64-lane wave LCP and the suffix comparator (lcp.rs:125-201);
  - packed prefix keys in the reference geometry (pack.rs:187-203) and the
    order property of the dense geometry;
  - the merge-path LCP merge (sample_sort.rs:397-522) and the cascade merge
    of runs (ext_mem.rs:2194-2389);
  - the streaming builders with their emit callback and error path
    (ext_mem.rs:711-871, 257-277)."""
import numpy as np
import pytest

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu
MAX = caps_sa.MAX_CONTEXT


def test_wave_lcp_matches_oracle(cuda):
    rng = np.random.default_rng(7)
    for dt, sigma in ((np.uint8, 4), (np.uint8, 2), (np.uint16, 5)):
        for n in (1, 33, 300, 5000):
            t = rng.integers(0, sigma, n).astype(dt)
            if n > 100:
                t[n // 2: n // 2 + 400] = t[: min(400, n - n // 2)][: len(t[n // 2: n // 2 + 400])]
            ctx = O.Ctx(t)
            pairs = rng.integers(0, n, (400, 2)).astype(np.uint64)
            for mx in (MAX, 0, 3, 70):
                got = dev.lcp_pairs(dev.text_to_device(t), n, t.dtype.itemsize, pairs.ravel(), mx)
                want = [ctx.lcp(int(p), int(q), mx) for p, q in pairs]
                assert got.tolist() == want, (dt, n, mx)


def test_wave_lcp_stride_boundaries(cuda):
    # lcp.rs:482-491 / 524-532: differences at every stride boundary
    for diff_at in [0, 1, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 200, 255, 256, 257]:
        t = np.full(1024, ord("A"), dtype=np.uint8)
        t[diff_at] = ord("G")
        got = dev.lcp_pairs(dev.text_to_device(t), len(t), 1, np.array([0, 512], dtype=np.uint64))
        assert int(got[0]) == diff_at


def test_suffix_cmp_matches_oracle(cuda):
    rng = np.random.default_rng(8)
    for order in (0, 1):
        lengths = rng.integers(3, 40, 30)
        n = int(lengths.sum())
        t = rng.integers(0, 3, n).astype(np.uint8)
        ends = np.cumsum(lengths)
        lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
        ctx = O.Ctx(t, ends, order)
        pairs = rng.integers(0, n, (500, 2)).astype(np.uint64)
        for mx in (MAX, 5):
            got = dev.suffix_cmp_pairs(dev.text_to_device(t), n, 1, pairs.ravel(), lp, mx)
            want = [ctx.cmp(int(p), int(q), mx) for p, q in pairs]
            assert got.tolist() == want


@pytest.mark.parametrize("rank", [0, 1])
def test_reference_geometry_keys(cuda, rank):
    rng = np.random.default_rng(9)
    for sigma in (2, 3, 4, 6, 15, 255):
        n = 3000
        t = rng.integers(0, sigma, n).astype(np.uint8)
        lengths = rng.integers(1, 60, 100)
        lengths = lengths[np.cumsum(lengths) <= n]
        lengths = np.append(lengths, n - lengths.sum()) if lengths.sum() < n else lengths
        ends = np.cumsum(lengths)
        lp = caps_sa.StarSegmented(n, ends) if rank else caps_sa.SegmentedText(n, ends)
        ctx = O.Ctx(t, ends, rank)
        pk = O.OraclePacker(t)
        assert pk.ok
        pos = np.arange(n, dtype=np.uint64)
        keys, (bits, k) = dev.seed_keys(dev.text_to_device(t), n, 1, pos, lp, geometry=1, rank=rank)
        assert (bits, k) == (pk.bits, pk.k)
        want = [pk.key(p, ctx.lim_at(p), rank) for p in range(n)]
        assert keys.tolist() == want


def test_reference_geometry_ineligible(cuda):
    t = np.arange(256, dtype=np.uint8)  # no spare sentinel code (pack.rs:516-528)
    with pytest.raises(caps_sa.CapsError) as e:
        dev.seed_keys(dev.text_to_device(t), 256, 1, np.arange(4, dtype=np.uint64), geometry=1)
    assert e.value.code == caps_sa.CAPS_E_UNSUPPORTED


def test_seed_eligibility_chain_matches_oracle(cuda):
    """caps_seed_params_device against the oracle's seed_params
    (pack.rs:217-234) on the pack.rs:484-528 texts, with the order of
    ext_mem.rs:1495-1513 pinned as ext_mem.rs:2534-2567 pins it: an earlier
    failing check wins over a later one."""
    D = lambda t: dev.text_to_device(np.asarray(t, dtype=np.uint8))
    gapped = [0, 2, 0, 2, 0, 2, 0, 2]
    cases = [
        (list(range(255)), 1, 0),  # 255 values: 8-bit fields, k = 8
        (list(range(256)), 1, 0),  # 256 values: no sentinel code
        (gapped, 1, 0),  # DenseAlphabetOnly declines a gapped alphabet
        (gapped, 2, 7),  # Remap over its budget
        (gapped, 2, 8),  # an exactly budgeted Remap
        ([0, 1, 2, 3] * 100, 1, 0),
        ([0, 1, 2] * 50, 2, 0),  # identity alphabet: no copy, any budget
        ([7, 9, 200] * 40, 2, 1000),
    ]
    for text, policy, budget in cases:
        got = dev.seed_params(D(text), len(text), 1, policy=policy, remap_budget=budget)
        pk = O.OraclePacker(np.asarray(text, dtype=np.uint8), policy=policy, remap_budget=budget)
        assert (got["reason"] == caps_sa.SEED_OK) == pk.ok, (text[:4], policy, budget, got)
        if pk.ok:
            assert (got["bits"], got["k"]) == (pk.bits, pk.k)
        assert got["alphabet"] == len(set(text))
    assert dev.seed_params(D(list(range(256))), 256, 1, policy=1)["reason"] == caps_sa.SEED_NO_SENTINEL
    assert dev.seed_params(D(gapped), 8, 1, policy=1)["reason"] == caps_sa.SEED_GAPPED_ALPHABET
    assert dev.seed_params(D(gapped), 8, 1, policy=2, remap_budget=7)["reason"] == caps_sa.SEED_REMAP_BUDGET
    # order: Disabled before max_context before the symbol type before the scan
    t4 = D([0, 1, 2, 3])
    r = lambda **kw: dev.seed_params(t4, 4, 1, **kw)["reason"]
    assert r(policy=0) == caps_sa.SEED_DISABLED
    assert r(policy=0, max_context=3) == caps_sa.SEED_DISABLED
    assert r(policy=1, max_context=3) == caps_sa.SEED_MAX_CONTEXT
    assert r(policy=1) == caps_sa.SEED_OK
    t16 = dev.text_to_device(np.array([0, 1, 2, 3], dtype=np.uint16))
    assert dev.seed_params(t16, 4, 2, policy=1)["reason"] == caps_sa.SEED_NOT_U8
    assert dev.seed_params(t16, 4, 2, policy=1, max_context=3)["reason"] == caps_sa.SEED_MAX_CONTEXT
    # a build under every policy is the same suffix array
    t = gen.markov_repeats(20000, seed=3, mean_len=100.0)
    want = O.Ctx(t).build_in_memory()
    for pol in caps_sa.PackedPrefixSeedPolicy:
        opts = caps_sa.ExtMemOpts(packed_prefix_seed=pol, remap_budget=1 << 20)
        got = []
        caps_sa.build_in_memory_sample_sort(t, opts, got.append)
        assert np.array_equal(np.array(got, dtype=np.uint64), want), pol


def test_dense_keys_order_property(cuda):
    """Dense (weakly order-preserving) keys: key_a < key_b must imply
    suffix_a < suffix_b under the comparator, for both boundary orders."""
    rng = np.random.default_rng(10)
    for order in (0, 1):
        for sigma in (2, 4, 5, 200):
            n = 4000
            t = rng.integers(0, sigma, n).astype(np.uint8)
            t[1000:1600] = t[2000:2600]
            lengths = rng.integers(1, 90, 200)
            lengths = lengths[np.cumsum(lengths) < n]
            ends = np.append(np.cumsum(lengths), n)
            lp = caps_sa.StarSegmented(n, ends) if order else caps_sa.SegmentedText(n, ends)
            ctx = O.Ctx(t, ends, order)
            keys, _ = dev.seed_keys(dev.text_to_device(t), n, 1, np.arange(n, dtype=np.uint64), lp, geometry=0,
                                    rank=order)
            a = rng.integers(0, n, 3000)
            b = rng.integers(0, n, 3000)
            for p, q in zip(a.tolist(), b.tolist()):
                if keys[p] < keys[q]:
                    assert ctx.cmp(p, q) < 0, (order, sigma, p, q)


def _runs(ctx, positions, cuts):
    sa_parts, lcp_parts, bounds = [], [], [0]
    for piece in np.split(positions, cuts):
        s, l = ctx.build_in_memory(piece, want_lcp=True)
        sa_parts.append(s)
        lcp_parts.append(l)
        bounds.append(bounds[-1] + len(s))
    return np.concatenate(sa_parts), np.concatenate(lcp_parts), bounds


@pytest.mark.parametrize("order", [0, 1])
def test_merge_lcp_matches_oracle(cuda, order):
    rng = np.random.default_rng(11 + order)
    for n, sigma in ((50, 2), (3000, 3), (20000, 4)):
        t = rng.integers(0, sigma, n).astype(np.uint8)
        t[n // 3: n // 3 + n // 5] = t[: n // 5]
        ends = None
        lp = None
        if order:
            lengths = rng.integers(5, 200, n)
            lengths = lengths[np.cumsum(lengths) < n]
            ends = np.append(np.cumsum(lengths), n)
            lp = caps_sa.StarSegmented(n, ends)
        ctx = O.Ctx(t, ends, order)
        perm = rng.permutation(n).astype(np.uint64)
        cut = int(rng.integers(1, n))
        x, lx = ctx.build_in_memory(perm[:cut], want_lcp=True)
        y, ly = ctx.build_in_memory(perm[cut:], want_lcp=True)
        for mx in (MAX, 7):
            if mx != MAX:
                x, lx = ctx.build_in_memory(perm[:cut], want_lcp=True, max_ctx=mx)
                y, ly = ctx.build_in_memory(perm[cut:], want_lcp=True, max_ctx=mx)
            z, lz = caps_sa.merge_lcp(t, x, lx, y, ly, lp, mx)
            wz, wlz = ctx.merge(x, lx, y, ly, mx)
            assert np.array_equal(z, wz) and np.array_equal(lz, wlz), (n, mx)


def test_cascade_merge_matches_oracle(cuda):
    rng = np.random.default_rng(13)
    for n, nruns in ((1000, 7), (30000, 40)):
        t = rng.integers(0, 4, n).astype(np.uint8)
        ctx = O.Ctx(t)
        perm = rng.permutation(n).astype(np.uint64)
        cuts = np.sort(rng.choice(np.arange(1, n), nruns - 1, replace=False))
        sa, lcp, bounds = _runs(ctx, perm, cuts)
        # an empty run between boundaries is skipped (ext_mem.rs:2215-2221)
        bounds = bounds[:3] + [bounds[2]] + bounds[3:]
        got, glcp = caps_sa.cascade_merge(t, sa, lcp, bounds)
        want, wlcp = ctx.cascade_merge(sa, lcp, bounds)
        assert np.array_equal(got, want) and np.array_equal(glcp, wlcp)


def test_sample_sort_streaming_matches_oracle(cuda):
    for text in (b"banana", b"mississippi", b"abracadabra"):
        out = []
        caps_sa.build_in_memory_sample_sort(text, caps_sa.ExtMemOpts(), out.append)
        assert out == O.Ctx(text).build_in_memory().tolist()
    t = gen.markov_repeats(200000, seed=21, mean_len=800.0)
    out = []
    caps_sa.build_in_memory_sample_sort(t, caps_sa.ExtMemOpts(subproblem_count=1024), out.append)
    assert np.array_equal(np.array(out, dtype=np.uint64), O.Ctx(t).sample_sort(p=1024, threads=8))


def test_for_positions_streaming(cuda):
    text, ends, pos = gen.genome(100000, 40, n_mean=300.0, seed=4)
    ctx = O.Ctx(text, ends, O.ORDER_STAR)
    out = []
    caps_sa.try_build_ext_mem_for_positions_with(text, pos, caps_sa.StarSegmented(len(text), ends),
                                                 caps_sa.ExtMemOpts(), out.append)
    assert np.array_equal(np.array(out, dtype=np.uint64), ctx.brute_sa(pos))
    out = []
    caps_sa.build_in_memory_sample_sort_for_positions(text, pos, caps_sa.ExtMemOpts(), out.append)
    plain = O.Ctx(text)
    assert plain.verify(np.array(out, dtype=np.uint64), positions=pos, strict=True) == -1


def test_emit_error_aborts(cuda):
    # ext_mem.rs:2895-2915: the first callback error aborts with that error
    class Stop(Exception):
        pass

    seen = []

    def emit(p):
        seen.append(p)
        if len(seen) == 2:
            raise Stop()

    with pytest.raises(caps_sa.BuildError) as e:
        caps_sa.try_build_in_memory_sample_sort(b"banana", caps_sa.ExtMemOpts(subproblem_count=2), emit)
    assert e.value.kind == "Emit" and isinstance(e.value.inner, Stop)
    assert len(seen) == 2
    def always_stop(p):
        raise Stop()

    with pytest.raises(OSError):
        caps_sa.build_in_memory_sample_sort(b"banana", caps_sa.ExtMemOpts(), always_stop)
