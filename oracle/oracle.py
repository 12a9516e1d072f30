"""ctypes binding of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker or the timed CPU baseline.  The product (caps-sa_amd/) never
imports it.  See oracle.h for what each function restates and how it is
pinned.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAX_CTX = (1 << 64) - 1
ORDER_SHORTER_FIRST = 0
ORDER_STAR = 1

_o = None


class Text(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_uint64), ("sym_bytes", ctypes.c_uint32)]


class Limits(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("ends", ctypes.c_void_p), ("n_ends", ctypes.c_uint64),
                ("order", ctypes.c_int), ("has_dir", ctypes.c_int), ("block_shift", ctypes.c_uint32),
                ("first_after_block_start", ctypes.c_void_p), ("n_dir", ctypes.c_uint64)]


class Packer(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8 * 256), ("ranked", ctypes.c_void_p), ("bits", ctypes.c_uint32),
                ("k", ctypes.c_uint32), ("alphabet", ctypes.c_uint32)]


EMIT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        o = ctypes.CDLL(LIB)
        P, u64, i32, vp = ctypes.POINTER, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        T, L = P(Text), P(Limits)
        sig = {
            "orc_limits_init": (i32, [L, u64, vp, u64, i32]),
            "orc_limits_free": (None, [L]),
            "orc_lim_at": (u64, [L, u64]),
            "orc_lim_at_bsearch": (u64, [L, u64]),
            "orc_boundary_order": (i32, [L, u64, u64, u64, u64]),
            "orc_lcp": (u64, [T, u64, u64, u64]),
            "orc_suffix_cmp_with": (i32, [T, L, u64, u64, u64]),
            "orc_merge": (None, [T, L, vp, u64, vp, u64, vp, vp, vp, vp, u64]),
            "orc_merge_sort": (None, [T, L, vp, vp, vp, vp, u64, u64, i32]),
            "orc_build_in_memory": (None, [T, L, vp, u64, u64, vp, i32]),
            "orc_cascade_merge": (None, [T, L, vp, vp, u64, vp, u64, u64, vp, vp]),
            "orc_effective_subproblem_count": (u64, [u64, u64, u64]),
            "orc_sample_target_total": (u64, [u64, u64]),
            "orc_evenly_spaced": (u64, [vp, u64, u64, vp]),
            "orc_upper_bound_by_pivot": (u64, [T, L, vp, u64, u64, u64]),
            "orc_build_in_memory_sample_sort": (i32, [T, L, vp, u64, u64, u64, i32, EMIT, vp]),
            "orc_seed_params": (i32, [vp, u64, i32, u64, P(Packer)]),
            "orc_packer_free": (None, [P(Packer)]),
            "orc_key_at_bounded": (u64, [P(Packer), vp, u64, u64, u64, i32]),
            "orc_shared_fields": (u64, [P(Packer), u64, u64]),
            "orc_seed_subarray": (None, [T, L, P(Packer), i32, vp, vp, u64]),
            "orc_brute_sa": (None, [T, L, vp, u64, u64]),
            "orc_doubling_sa": (i32, [T, vp]),
            "orc_kasai_lcp": (i32, [T, vp, vp]),
            "orc_naive_lcp": (u64, [T, L, u64, u64, u64]),
            "orc_verify_sa": (ctypes.c_int64, [T, L, vp, u64, vp, u64, i32, i32, i32]),
            "orc_verify_sa_keep": (ctypes.c_int64, [T, L, vp, u64, vp, u64, i32, i32, i32]),
            "orc_canonicalize": (None, [T, L, vp, u64, u64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(o, name)
            f.restype = res
            f.argtypes = args
        _o = o
    return _o


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


class Ctx:
    """Text + limit provider, kept alive for the oracle calls."""

    def __init__(self, text, ends=None, order=ORDER_SHORTER_FIRST):
        if isinstance(text, (bytes, bytearray)):
            text = np.frombuffer(bytes(text), dtype=np.uint8)
        self.text = np.ascontiguousarray(text)
        assert self.text.dtype in (np.uint8, np.uint16)
        self.n = len(self.text)
        self.T = Text(self.text.ctypes.data if self.n else None, self.n, self.text.dtype.itemsize)
        self.ends = _u64(ends) if ends is not None else np.zeros(0, dtype=np.uint64)
        self.L = Limits()
        rc = lib().orc_limits_init(ctypes.byref(self.L), self.n,
                                   self.ends.ctypes.data if len(self.ends) else None, len(self.ends), order)
        assert rc == 0

    def __del__(self):
        try:
            lib().orc_limits_free(ctypes.byref(self.L))
        except Exception:
            pass

    # -- primitives
    def lim_at(self, p):
        return lib().orc_lim_at(ctypes.byref(self.L), p)

    def lim_at_bsearch(self, p):
        return lib().orc_lim_at_bsearch(ctypes.byref(self.L), p)

    def lcp(self, p, q, max_ctx=MAX_CTX):
        return lib().orc_lcp(ctypes.byref(self.T), p, q, max_ctx)

    def cmp(self, p, q, max_ctx=MAX_CTX):
        return lib().orc_suffix_cmp_with(ctypes.byref(self.T), ctypes.byref(self.L), p, q, max_ctx)

    def naive_lcp(self, a, b, max_ctx=MAX_CTX):
        return lib().orc_naive_lcp(ctypes.byref(self.T), ctypes.byref(self.L), a, b, max_ctx)

    # -- builds
    def build_in_memory(self, positions=None, max_ctx=MAX_CTX, threads=1, want_lcp=False):
        sa = _u64(np.arange(self.n, dtype=np.uint64) if positions is None else positions).copy()
        lcp = np.zeros(len(sa), dtype=np.uint64) if want_lcp else None
        lib().orc_build_in_memory(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, len(sa), max_ctx,
                                  lcp.ctypes.data if lcp is not None else None, threads)
        return (sa, lcp) if want_lcp else sa

    def merge(self, x, lx, y, ly, max_ctx=MAX_CTX):
        x, lx, y, ly = _u64(x), _u64(lx), _u64(y), _u64(ly)
        z = np.zeros(len(x) + len(y), dtype=np.uint64)
        lz = np.zeros_like(z)
        lib().orc_merge(ctypes.byref(self.T), ctypes.byref(self.L), x.ctypes.data, len(x), y.ctypes.data, len(y),
                        lx.ctypes.data, ly.ctypes.data, z.ctypes.data, lz.ctypes.data, max_ctx)
        return z, lz

    def cascade_merge(self, sa, lcp, boundaries, max_ctx=MAX_CTX):
        sa, lcp, b = _u64(sa), _u64(lcp), _u64(boundaries)
        out = np.zeros_like(sa)
        lout = np.zeros_like(sa)
        lib().orc_cascade_merge(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, lcp.ctypes.data,
                                len(sa), b.ctypes.data, len(b), max_ctx, out.ctypes.data, lout.ctypes.data)
        return out, lout

    def sample_sort(self, positions=None, p=0, max_ctx=MAX_CTX, threads=1, emit=None):
        out = []

        def _cb(_ctx, ptr, count):
            arr = np.ctypeslib.as_array(ptr, shape=(count,)).copy()
            if emit is not None:
                return emit(arr)
            out.append(arr)
            return 0

        cb = EMIT(_cb)
        pos = _u64(positions) if positions is not None else None
        n_pos = len(pos) if pos is not None else self.n
        rc = lib().orc_build_in_memory_sample_sort(ctypes.byref(self.T), ctypes.byref(self.L),
                                                   pos.ctypes.data if pos is not None else None, n_pos, p,
                                                   max_ctx, threads, cb, None)
        if emit is not None:
            return rc
        return np.concatenate(out) if out else np.zeros(0, dtype=np.uint64)

    def seed_subarray(self, packer, rank, positions):
        """pack.rs:241-321 on a u8 text: (sorted positions, adjacent LCPs)."""
        sa = _u64(positions).copy()
        lcp = np.zeros(len(sa), dtype=np.uint64)
        lib().orc_seed_subarray(ctypes.byref(self.T), ctypes.byref(self.L), ctypes.byref(packer.P), rank,
                                sa.ctypes.data if len(sa) else None, lcp.ctypes.data if len(sa) else None, len(sa))
        return sa, lcp

    def upper_bound_by_pivot(self, recs, pivot, max_ctx=MAX_CTX):
        """ext_mem.rs:1705-1734: first index of a sorted run whose suffix is
        strictly greater than the pivot suffix."""
        recs = _u64(recs)
        return lib().orc_upper_bound_by_pivot(ctypes.byref(self.T), ctypes.byref(self.L),
                                              recs.ctypes.data if len(recs) else None, len(recs), pivot, max_ctx)

    def brute_sa(self, positions=None, max_ctx=MAX_CTX):
        sa = _u64(np.arange(self.n, dtype=np.uint64) if positions is None else positions).copy()
        lib().orc_brute_sa(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, len(sa), max_ctx)
        return sa

    def doubling_sa(self):
        sa = np.zeros(self.n, dtype=np.uint64)
        assert lib().orc_doubling_sa(ctypes.byref(self.T), sa.ctypes.data) == 0
        return sa

    def kasai(self, sa):
        sa = _u64(sa)
        lcp = np.zeros(len(sa), dtype=np.uint64)
        assert lib().orc_kasai_lcp(ctypes.byref(self.T), sa.ctypes.data, lcp.ctypes.data) == 0
        return lcp

    def verify(self, sa, positions=None, max_ctx=MAX_CTX, strict=True, canon=True, threads=8):
        sa = _u64(sa)
        pos = _u64(positions) if positions is not None else None
        return lib().orc_verify_sa(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, len(sa),
                                   pos.ctypes.data if pos is not None else None, max_ctx, int(strict), int(canon),
                                   threads)

    def verify_keep(self, sa, keep_words, max_ctx=MAX_CTX, strict=True, canon=True, threads=8):
        """verify() against a kept-position bitmap instead of a position list."""
        sa = _u64(sa)
        kw = _u64(keep_words)
        return lib().orc_verify_sa_keep(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, len(sa),
                                        kw.ctypes.data, max_ctx, int(strict), int(canon), threads)

    def canonicalize(self, sa, max_ctx=MAX_CTX):
        sa = _u64(sa).copy()
        lib().orc_canonicalize(ctypes.byref(self.T), ctypes.byref(self.L), sa.ctypes.data, len(sa), max_ctx)
        return sa

    def lcp_array(self, sa, max_ctx=MAX_CTX):
        """Adjacent LCPs capped by both limits and max_ctx (naive scan)."""
        sa = _u64(sa)
        out = np.zeros(len(sa), dtype=np.uint64)
        for i in range(1, len(sa)):
            out[i] = self.naive_lcp(int(sa[i - 1]), int(sa[i]), max_ctx)
        return out


class OraclePacker:
    """pack.rs Packer over a u8 text."""

    def __init__(self, text, policy=1, remap_budget=0):
        self.text = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8)
                                         if isinstance(text, (bytes, bytearray)) else text, dtype=np.uint8)
        self.P = Packer()
        self.ok = bool(lib().orc_seed_params(self.text.ctypes.data if len(self.text) else None, len(self.text),
                                             policy, remap_budget, ctypes.byref(self.P)))

    def __del__(self):
        try:
            lib().orc_packer_free(ctypes.byref(self.P))
        except Exception:
            pass

    @property
    def bits(self):
        return self.P.bits

    @property
    def k(self):
        return self.P.k

    def key(self, p, lim, rank=0):
        return lib().orc_key_at_bounded(ctypes.byref(self.P), self.text.ctypes.data, len(self.text), p, lim, rank)

    def shared_fields(self, a, b):
        return lib().orc_shared_fields(ctypes.byref(self.P), a, b)


def effective_subproblem_count(n, requested, threads):
    return lib().orc_effective_subproblem_count(n, requested, threads)


def sample_target_total(n, p):
    return lib().orc_sample_target_total(n, p)


def evenly_spaced(xs, count):
    xs = _u64(xs)
    out = np.zeros(max(count, 1), dtype=np.uint64)
    k = lib().orc_evenly_spaced(xs.ctypes.data if len(xs) else None, len(xs), count, out.ctypes.data)
    return out[:k]
