"""Seeded synthetic workloads (ctypes over libcaps_gen.so)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN_PATH = os.path.join(_PKG_DIR, "libcaps_gen.so")
_g = None


def _lib():
    global _g
    if _g is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: run __graft_entry__.build() first")
        g = ctypes.CDLL(GEN_PATH)
        vp, u64, u32, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double
        g.caps_gen_uniform_u8.argtypes = [vp, u64, u32, u64]
        g.caps_gen_uniform_u16.argtypes = [vp, u64, u32, u64]
        g.caps_gen_markov_repeats.argtypes = [vp, u64, u64, dbl, dbl, dbl]
        g.caps_gen_markov_repeats.restype = u64
        g.caps_gen_zipf_u16.argtypes = [vp, u64, u32, dbl, u64]
        g.caps_gen_lognormal_lengths.argtypes = [vp, u64, dbl, dbl, u64, u64, u64]
        g.caps_gen_genome.argtypes = [vp, vp, u64, u64, dbl, dbl, dbl, dbl, u64]
        g.caps_gen_genome.restype = u64
        g.caps_gen_next.argtypes = [u64, u64]
        g.caps_gen_next.restype = u64
        _g = g
    return _g


def uniform_u8(n: int, sigma: int = 4, seed: int = 1) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().caps_gen_uniform_u8(out.ctypes.data, n, sigma, seed)
    return out


def uniform_u16(n: int, sigma: int, seed: int = 1) -> np.ndarray:
    out = np.empty(n, dtype=np.uint16)
    _lib().caps_gen_uniform_u16(out.ctypes.data, n, sigma, seed)
    return out


def markov_repeats(n: int, seed: int = 1, repeat_frac: float = 0.4, mean_len: float = 5000.0,
                   sub_rate: float = 0.01) -> np.ndarray:
    """Order-3 Markov text over {0..3} with planted repeats (the 268M
    repeat-rich workload at n = 2**28)."""
    out = np.empty(n, dtype=np.uint8)
    _lib().caps_gen_markov_repeats(out.ctypes.data, n, seed, repeat_frac, mean_len, sub_rate)
    return out


def zipf_u16(n: int, alphabet: int = 50000, s: float = 1.1, seed: int = 1) -> np.ndarray:
    out = np.empty(n, dtype=np.uint16)
    _lib().caps_gen_zipf_u16(out.ctypes.data, n, alphabet, s, seed)
    return out


def lognormal_lengths(count: int, median: float, sigma: float, lo: int, hi: int, seed: int = 1) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    _lib().caps_gen_lognormal_lengths(out.ctypes.data, count, median, sigma, lo, hi, seed)
    return out


def genome(n: int, n_seg: int, median: float = 2500.0, sigma: float = 1.6, n_frac: float = 0.058,
           n_mean: float = 10000.0, seed: int = 1, positions: bool = True):
    """Segmented genome-like text: returns (text, ends, positions with text[p] < 4);
    positions is None when not asked for (a genome-scale list is 8 bytes a position)."""
    text = np.empty(n, dtype=np.uint8)
    ends = np.empty(n_seg, dtype=np.uint64)
    _lib().caps_gen_genome(text.ctypes.data, ends.ctypes.data, n, n_seg, median, sigma, n_frac, n_mean, seed)
    pos = np.flatnonzero(text < 4).astype(np.uint64) if positions else None
    return text, ends, pos


def keep_below(text: np.ndarray, limit: int, chunk: int = 1 << 26) -> np.ndarray:
    """Kept-position bitmap of {q : text[q] < limit}, built chunk by chunk
    ((n + 63) // 64 u64 words, bit q % 64 of word q // 64)."""
    n = len(text)
    words = np.zeros((n + 63) // 64, dtype=np.uint64)
    as_bytes = words.view(np.uint8)
    for s in range(0, n, chunk):  # chunk is a multiple of 64: whole words
        e = min(n, s + chunk)
        packed = np.packbits(text[s:e] < limit, bitorder="little")
        as_bytes[s // 8: s // 8 + len(packed)] = packed
    return words
