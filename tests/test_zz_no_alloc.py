"""The no-alloc entry point (caps_build_device_noalloc) against
caps_build_device: the same suffix and LCP arrays, called through the C ABI,
on two calls in a row with the second input smaller than the first (the
scratch array keeps the first call's contents), for u8/u16 text, u32/u64
index, segmented text and position subsets.  Sizes above the scratch limit
are refused before any device work.
"""
import numpy as np
import pytest
import torch

import caps_sa
from caps_sa import device as dev
from caps_sa import gen

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20  # CAPS_NOALLOC_MAX_SYMBOLS
NOALLOC_MIB = 128  # CAPS_NOALLOC_SCRATCH_MIB


def _pair(text, idx_bytes=4, lp=None, positions=None, want_lcp=True, router=0):
    """(noalloc sa, noalloc lcp, device sa, device lcp, scratch bytes) as numpy arrays."""
    n = len(text)
    w = text.dtype.itemsize
    d_text = dev.text_to_device(text)
    d_pos = None
    if positions is not None:
        dt = np.uint32 if idx_bytes == 4 else np.uint64
        d_pos = torch.from_numpy(np.asarray(positions, dtype=dt).view(np.int32 if idx_bytes == 4 else np.int64)).cuda()
    n_pos = n if positions is None else len(positions)
    tdt = torch.int32 if idx_bytes == 4 else torch.int64
    d_lcp = torch.empty(max(n_pos, 1), dtype=tdt, device="cuda:0") if want_lcp else None
    sa1, lcp1 = dev.build_device_noalloc(d_text, n, w, d_positions=d_pos, n_pos=n_pos, idx_bytes=idx_bytes, lp=lp,
                                         d_lcp=d_lcp, router=router)
    torch.cuda.synchronize()
    used = caps_sa.last_stats()["scratch_bytes"]
    sa2, lcp2 = dev.build_device(d_text, n, w, d_positions=d_pos, n_pos=n_pos, idx_bytes=idx_bytes, lp=lp,
                                 want_lcp=want_lcp, router=router)
    torch.cuda.synchronize()
    npdt = np.uint32 if idx_bytes == 4 else np.uint64
    f = lambda t: None if t is None else t.cpu().numpy().view(npdt)[:n_pos].copy()
    return f(sa1), f(lcp1), f(sa2), f(lcp2), used


def _same(label, a):
    print(f"case {label}: scratch {a[4] / 2**20:.1f} MiB", flush=True)
    assert np.array_equal(a[0], a[2]), label
    if a[1] is not None:
        assert np.array_equal(a[1], a[3]), label


def test_noalloc_two_calls_second_smaller(cuda):
    big = gen.markov_repeats(LIMIT, seed=71, mean_len=300.0)
    small = gen.markov_repeats(70_001, seed=72, mean_len=50.0)
    r1 = _pair(big)
    _same("u8 u32 2^20 (first call)", r1)
    r2 = _pair(small)
    _same("u8 u32 70001 (second call, smaller)", r2)
    assert 0 < r2[4] <= r1[4] <= NOALLOC_MIB * 2**20


def test_noalloc_u16_u64_segmented_positions(cuda):
    rng = np.random.default_rng(5)
    t16 = gen.zipf_u16(300_000, 5000, 1.1, seed=9)
    _same("u16 u64", _pair(t16, idx_bytes=8))
    n = 200_000
    text = gen.uniform_u8(n, 4, seed=3)
    ends = np.unique(np.concatenate([rng.integers(1, n, 400), [n]])).astype(np.uint64)
    _same("segmented 400 ends", _pair(text, lp=caps_sa.SegmentedText(n, ends)))
    pos = np.sort(rng.choice(n, 50_000, replace=False)).astype(np.uint64)
    _same("positions", _pair(text, positions=pos, idx_bytes=8, want_lcp=False))
    _same("router pwl", _pair(gen.markov_repeats(150_000, seed=4, mean_len=200.0), router=1))


def test_noalloc_largest_inputs_fit_the_scratch(cuda):
    """The heaviest shapes at the size limit: one repeated symbol (every
    bucket oversized, one giant group merged level by level) and a u16
    text over a large alphabet, u64 index with LCP."""
    r = _pair(np.zeros(LIMIT, dtype=np.uint8), idx_bytes=8)
    _same("one symbol, 2^20, u64", r)
    r = _pair(gen.zipf_u16(LIMIT, 50000, 1.1, seed=11), idx_bytes=8)
    _same("u16 zipf, 2^20, u64", r)


def test_noalloc_refuses_sizes_above_the_scratch(cuda):
    text = np.zeros(LIMIT + 1, dtype=np.uint8)
    d_text = dev.text_to_device(text)
    with pytest.raises(caps_sa.CapsError) as e:
        dev.build_device_noalloc(d_text, len(text), 1)
    assert e.value.code == caps_sa.CAPS_E_INVALID
