"""Python mirror of the caps-sa 0.7.0 public API over the MI355X C ABI.

Names, argument meaning and error behaviour follow the Rust crate
(src/lib.rs:28-45): ``build_in_memory*`` (sample_sort.rs:97-206),
``(try_)build_in_memory_sample_sort*`` and ``(try_)build_ext_mem_for_positions*``
(ext_mem.rs:370-871), ``Opts`` / ``ExtMemOpts``, ``PlainText`` /
``SegmentedText`` / ``BoundaryRank`` (limits.rs), ``PackedPrefixSeedPolicy``
(pack.rs:37-69).  Every build runs on the GPU through
``libcaps_sa_hip.so``; there is no CPU fallback, so a missing library or a
missing GPU raises.

Text is a numpy array of ``uint8`` or ``uint16`` (``bytes`` is accepted as
``uint8``).  The index type is chosen with ``index_dtype`` (``numpy.uint32`` or
``numpy.uint64``), the counterpart of the crate's ``I: Index`` generic.
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum
import os
import warnings
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libcaps_sa_hip.so")

CAPS_OK = 0
CAPS_E_INVALID = 1
CAPS_E_DEVICE = 2
CAPS_E_EMIT = 3
CAPS_E_UNSUPPORTED = 4
CAPS_E_NOMEM = 5

MAX_CONTEXT = (1 << 64) - 1  # usize::MAX


class _Limits(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("ends", ctypes.POINTER(ctypes.c_uint64)),
        ("n_ends", ctypes.c_uint64),
    ]


class _Opts(ctypes.Structure):
    _fields_ = [
        ("max_context", ctypes.c_uint64),
        ("subproblem_count", ctypes.c_uint64),
        ("router", ctypes.c_uint32),
        ("key_geometry", ctypes.c_uint32),
        ("gpu_buckets", ctypes.c_uint32),
        ("tile_keys", ctypes.c_uint32),
        ("want_lcp", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("batch_records", ctypes.c_uint64),
        ("packed_prefix_seed", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("remap_budget", ctypes.c_uint64),
    ]


class SeedParams(ctypes.Structure):
    """caps_seed_params_t: the reference's packed-seed decision."""

    _fields_ = [("reason", ctypes.c_uint32), ("bits", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("alphabet", ctypes.c_uint32), ("remap", ctypes.c_uint32)]


SEED_OK, SEED_DISABLED, SEED_MAX_CONTEXT, SEED_NOT_U8, SEED_GAPPED_ALPHABET, SEED_REMAP_BUDGET, SEED_NO_SENTINEL = range(7)


class TuneReport(ctypes.Structure):
    _fields_ = [("n_tried", ctypes.c_uint32), ("best_ms", ctypes.c_double), ("ms", ctypes.c_double * 16),
                ("tile_keys", ctypes.c_uint32 * 16), ("flags", ctypes.c_uint32 * 16)]

    def as_dict(self) -> dict:
        k = self.n_tried
        return {"best_ms": self.best_ms,
                "tried": [{"tile_keys": self.tile_keys[i], "flags": self.flags[i], "ms": self.ms[i]}
                          for i in range(k)]}


class Stats(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_double)
        for name in (
            "ms_total", "ms_alphabet", "ms_sample", "ms_classify", "ms_scatter", "ms_tile_sort", "ms_merge",
            "ms_detect", "ms_refine", "ms_giant", "k_classify", "k_scatter", "k_tile_sort", "k_detect", "k_pairs",
            "k_groups_small", "k_groups_mid", "k_boundary_lcp",
        )
    ] + [
        (name, ctypes.c_uint64)
        for name in (
            "buckets", "tiles", "oversized", "pairs", "small_groups", "mid_groups", "giant_groups",
            "router_err_sum", "router_err_max", "router_fallbacks", "router_count",
        )
    ] + [("key_bits", ctypes.c_uint32), ("key_k", ctypes.c_uint32), ("bucket_max", ctypes.c_uint64),
         ("k_scatter1", ctypes.c_double), ("k_scatter2", ctypes.c_double), ("batches", ctypes.c_uint64),
         ("coarse_buckets", ctypes.c_uint64), ("coarse_max", ctypes.c_uint64), ("scratch_bytes", ctypes.c_uint64),
         ("k_tile", ctypes.c_double * 3), ("tile_records", ctypes.c_uint64 * 3), ("cmp_symbols", ctypes.c_uint64 * 3),
         ("k_graph", ctypes.c_double * 3), ("runs", ctypes.c_uint64), ("router_bound_lo", ctypes.c_int64),
         ("router_bound_hi", ctypes.c_int64)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, ctypes.Array) else v
        return out


EMIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64)

_lib = None

# Every entry point of include/caps_sa.h with its ctypes signature.
_SIGNATURES = {
    "caps_opts_default": (None, [ctypes.POINTER(_Opts)]),
    "caps_last_error": (ctypes.c_char_p, []),
    "caps_last_emit_code": (ctypes.c_int, []),
    "caps_device_count": (ctypes.c_int, []),
    "caps_last_stats": (ctypes.c_int, [ctypes.POINTER(Stats)]),
    "caps_build_in_memory": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.POINTER(_Limits), ctypes.POINTER(_Opts), ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "caps_build_in_memory_for_positions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                           ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                           ctypes.POINTER(_Limits), ctypes.POINTER(_Opts),
                                                           ctypes.c_void_p, ctypes.c_void_p]),
    "caps_build_in_memory_sample_sort": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                         ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Limits),
                                                         ctypes.POINTER(_Opts), EMIT_FN, ctypes.c_void_p]),
    "caps_build_ext_mem_for_positions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                         ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Limits),
                                                         ctypes.POINTER(_Opts), EMIT_FN, ctypes.c_void_p]),
    "caps_build_ext_mem_for_filter": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                      ctypes.c_void_p, ctypes.POINTER(_Limits),
                                                      ctypes.POINTER(_Opts), EMIT_FN, ctypes.c_void_p]),
    "caps_build_device_filtered": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(_Limits),
                                                   ctypes.POINTER(_Opts), ctypes.c_void_p, ctypes.c_uint64,
                                                   ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p,
                                                   ctypes.c_void_p]),
    "caps_build_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.POINTER(_Opts)]),
    "caps_build_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(_Opts),
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "caps_autotune_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                             ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                             ctypes.POINTER(_Opts), ctypes.c_void_p, ctypes.c_void_p]),
    "caps_autotune_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(_Opts),
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "caps_build_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(_Limits),
                                          ctypes.POINTER(_Opts), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "caps_seed_params_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                ctypes.POINTER(_Limits), ctypes.POINTER(_Opts),
                                                ctypes.POINTER(SeedParams), ctypes.c_void_p]),
    "caps_router_pwl_fit_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p]),
    "caps_router_spline_fit_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_void_p]),
    "caps_build_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                                  ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(_Limits),
                                                  ctypes.POINTER(_Opts), ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "caps_lcp_pairs_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
    "caps_suffix_cmp_pairs_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                     ctypes.POINTER(_Limits), ctypes.c_void_p, ctypes.c_uint64,
                                                     ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
    "caps_seed_keys_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                              ctypes.POINTER(_Limits), ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p]),
    "caps_merge_lcp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(_Limits),
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "caps_router_mlp_grad_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
    "caps_router_mlp_forward_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                       ctypes.c_void_p, ctypes.c_void_p]),
    "caps_cascade_merge": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.POINTER(_Limits), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_void_p, ctypes.c_void_p]),
}


class CapsError(RuntimeError):
    """A nonzero return code of the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"caps error {code}: {message}")
        self.code = code


def lib() -> ctypes.CDLL:
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(rc: int) -> None:
    if rc != CAPS_OK:
        raise CapsError(rc, lib().caps_last_error().decode())


def last_stats() -> dict:
    st = Stats()
    _check(lib().caps_last_stats(ctypes.byref(st)))
    return st.as_dict()


# ---------------------------------------------------------------------------
# Limit providers (limits.rs)
# ---------------------------------------------------------------------------
class BoundaryRank(enum.Enum):
    """limits.rs:113-123."""

    SHORTER_FIRST = 0
    LONGER_FIRST = 1


class LimitProvider:
    """Base of the providers the GPU build has a device descriptor for.

    A user-defined ``boundary_order`` cannot run on the GPU; only the default
    order (shorter first) and the STAR convention are available."""

    kind = 0
    order = 0
    ends: Optional[np.ndarray] = None

    def boundary_rank(self) -> Optional[BoundaryRank]:
        return BoundaryRank.LONGER_FIRST if self.order == 1 else BoundaryRank.SHORTER_FIRST

    def _c(self) -> _Limits:
        lim = _Limits()
        lim.kind = self.kind
        lim.order = self.order
        if self.ends is not None and len(self.ends):
            self._keep = np.ascontiguousarray(self.ends, dtype=np.uint64)
            lim.ends = self._keep.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
            lim.n_ends = len(self._keep)
        return lim


class PlainText(LimitProvider):
    """limits.rs:130-156: lim_at(p) = n - p."""

    def __init__(self, n: int):
        self.n = n

    def lim_at(self, p: int) -> int:
        return self.n - p


class SegmentedText(LimitProvider):
    """limits.rs:182-336 with the default (shorter first) boundary order."""

    kind = 1

    def __init__(self, text_len: int, ends: Sequence[int]):
        ends = np.asarray(ends, dtype=np.uint64)
        if len(ends) and np.any(ends[1:] <= ends[:-1]):
            raise ValueError("SegmentedText::from_ends: ends must be strictly increasing")
        if len(ends) == 0 and text_len != 0:
            raise ValueError(f"SegmentedText::from_ends: empty ends but text_len ({text_len}) != 0")
        if len(ends) and int(ends[-1]) != text_len:
            raise ValueError(f"SegmentedText::from_ends: last end ({int(ends[-1])}) != text_len ({text_len})")
        self.n = text_len
        self.ends = ends

    @classmethod
    def from_lengths(cls, text_len: int, lengths: Sequence[int]) -> "SegmentedText":
        ends = np.cumsum(np.asarray(lengths, dtype=np.uint64), dtype=np.uint64)
        total = int(ends[-1]) if len(ends) else 0
        if total != text_len:
            raise ValueError(
                f"SegmentedText::from_lengths: per-segment lengths sum to {total} but text_len is {text_len}")
        return cls(text_len, ends)

    @classmethod
    def from_ends(cls, text_len: int, ends: Sequence[int]) -> "SegmentedText":
        return cls(text_len, ends)

    def n_segments(self) -> int:
        return len(self.ends)

    def lim_at(self, p: int) -> int:
        i = int(np.searchsorted(self.ends, p, side="right"))
        return int(self.ends[i]) - p if i < len(self.ends) else self.n - p


class StarSegmented(SegmentedText):
    """STAR's provider: lim_b.cmp(&lim_a).then(p_a.cmp(&p_b)),
    BoundaryRank::LongerFirst (sample_sort.rs:998-1015, pack.rs:329-351)."""

    order = 1


# ---------------------------------------------------------------------------
# Options
# ---------------------------------------------------------------------------
class PackedPrefixSeedPolicy(enum.Enum):
    """pack.rs:37-69.  On the GPU every build seeds with packed keys: the
    reference's geometry where the reference would seed
    (seed_params_device; the remap budget is ExtMemOpts.remap_budget),
    dense fields otherwise.  Output never changes."""

    DISABLED = 0
    DENSE_ALPHABET_ONLY = 1
    REMAP = 2


class Router(enum.IntEnum):
    BSEARCH = 0
    PWL = 1
    MLP = 2
    SPLINE = 3


@dataclass
class Opts:
    """sample_sort.rs:71-85."""

    max_context: int = MAX_CONTEXT


@dataclass(frozen=True)
class GeometricMemoizationConfig:
    """lcp_memo.rs:64-145: tuning of the crate's geometric LCP memo, each
    value non-zero (NonZeroUsize there: ValueError here).  Memoization never
    changes the output; the GPU build accepts the policy and does not use it
    (its long-LCP scans read packed code words and fast-forward over shared
    stretches instead)."""

    probe_symbols: int = 256
    min_lcp_symbols: int = 1024
    activate_after_entries: int = 64
    max_entries_per_partition: int = 4096

    def __post_init__(self):
        for name in ("probe_symbols", "min_lcp_symbols", "activate_after_entries", "max_entries_per_partition"):
            v = getattr(self, name)
            if not isinstance(v, int) or v <= 0:
                raise ValueError(f"{name} must be a positive integer")

    def with_probe_symbols(self, v: int) -> "GeometricMemoizationConfig":
        return dataclasses.replace(self, probe_symbols=v)

    def with_min_lcp_symbols(self, v: int) -> "GeometricMemoizationConfig":
        return dataclasses.replace(self, min_lcp_symbols=v)

    def with_activate_after_entries(self, v: int) -> "GeometricMemoizationConfig":
        return dataclasses.replace(self, activate_after_entries=v)

    def with_max_entries_per_partition(self, v: int) -> "GeometricMemoizationConfig":
        return dataclasses.replace(self, max_entries_per_partition=v)


class LcpMemoizationPolicy:
    """lcp_memo.rs:16-40: Disabled is None, Geometric is a configuration."""

    DISABLED = None

    @staticmethod
    def geometric(config: Optional[GeometricMemoizationConfig] = None) -> GeometricMemoizationConfig:
        return config if config is not None else GeometricMemoizationConfig()


def _env_usize(name: str) -> Optional[int]:
    """read_env_usize (ext_mem.rs:241-243): usize::from_str of the value, an
    optional '+' then ASCII digits, else None."""
    v = os.environ.get(name)
    if v is None:
        return None
    body = v[1:] if v.startswith("+") else v
    if not body or not all("0" <= ch <= "9" for ch in body):
        return None
    x = int(body)
    return x if x < (1 << 64) else None


def _env_bool(name: str) -> bool:
    """read_env_bool (ext_mem.rs:249-253)."""
    return os.environ.get(name) in ("1", "true", "yes", "on")


@dataclass
class ExtMemOpts:
    """ext_mem.rs:58-239 (the fields with a GPU meaning) plus GPU knobs.
    lcp_memoization (None: LcpMemoizationPolicy::Disabled, else the geometric
    configuration) and collect_lcp_memoization_stats are accepted for
    callers of the crate and leave the GPU build unchanged."""

    max_context: int = MAX_CONTEXT
    subproblem_count: int = 0
    packed_prefix_seed: PackedPrefixSeedPolicy = PackedPrefixSeedPolicy.DISABLED
    remap_budget: int = 0  # PackedPrefixSeedPolicy::Remap { max_extra_bytes }
    router: Router = Router.BSEARCH
    gpu_buckets: int = 0
    tile_keys: int = 0
    batch_records: int = 0  # records per bucket batch (0: one batch while it fits in HBM)
    lcp_memoization: Optional[GeometricMemoizationConfig] = None
    collect_lcp_memoization_stats: bool = False

    @classmethod
    def from_env(cls) -> "ExtMemOpts":
        """ext_mem.rs:154-196, for the variables with a GPU meaning and the
        memoization policy (its tuning variables count only when
        CAPS_SA_GEOMETRIC_MEMO enables it; zero or invalid values keep the
        defaults)."""
        o = cls()
        v = _env_usize("CAPS_SA_SUBPROBLEMS")
        if v is not None:
            o.subproblem_count = v
        v = _env_usize("CAPS_SA_MAX_CONTEXT")
        if v is not None:
            o.max_context = v
        v = _env_usize("CAPS_SA_PACKED_PREFIX_REMAP_BYTES")
        if v is not None:
            o.packed_prefix_seed = PackedPrefixSeedPolicy.REMAP
            o.remap_budget = v
        elif _env_bool("CAPS_SA_PACKED_PREFIX_SEED"):
            o.packed_prefix_seed = PackedPrefixSeedPolicy.DENSE_ALPHABET_ONLY
        if _env_bool("CAPS_SA_GEOMETRIC_MEMO"):
            c = GeometricMemoizationConfig()
            v = _env_usize("CAPS_SA_MEMO_PROBE")
            if v:
                c = c.with_probe_symbols(v)
            v = _env_usize("CAPS_SA_MEMO_MIN_LCP")
            if v:
                c = c.with_min_lcp_symbols(v)
            v = _env_usize("CAPS_SA_MEMO_ACTIVATE_ENTRIES")
            if v:
                c = c.with_activate_after_entries(v)
            v = _env_usize("CAPS_SA_MEMO_CAPACITY")
            if v:
                c = c.with_max_entries_per_partition(v)
            o.lcp_memoization = c
        o.collect_lcp_memoization_stats = _env_bool("CAPS_SA_MEMO_STATS")
        return o


FLAG_ROUTER_UNTRAINED = 1
FLAG_PERSISTENT_REFINE = 2
FLAG_TILE_RADIX = 4
FLAG_MERGE_W16 = 8
FLAG_MERGE_W64 = 16
FLAG_POW2_SEED = 32
FLAG_SCATTER_PER_BATCH = 64
FLAG_COUNT_WORK = 128
FLAG_SCATTER_DIRECT = 256
FLAG_GRAPH = 512
FLAG_SEED_ESCAPE32 = 1024
FLAG_NO_RUN_MEMO = 2048
FLAG_MID_BY_SIZE = 4096
TUNE_MAX = 16


def _c_opts(max_context=MAX_CONTEXT, subproblem_count=0, router=0, key_geometry=0, gpu_buckets=0, tile_keys=0,
            want_lcp=0, flags=0, batch_records=0, packed_prefix_seed=0, remap_budget=0) -> _Opts:
    o = _Opts()
    lib().caps_opts_default(ctypes.byref(o))
    o.max_context = min(int(max_context), MAX_CONTEXT)
    o.subproblem_count = int(subproblem_count)
    o.router = int(router)
    o.key_geometry = int(key_geometry)
    o.gpu_buckets = int(gpu_buckets)
    o.tile_keys = int(tile_keys)
    o.want_lcp = int(want_lcp)
    o.flags = int(flags)
    o.batch_records = int(batch_records)
    o.packed_prefix_seed = int(packed_prefix_seed)
    o.remap_budget = int(remap_budget)
    return o


def _ext_c_opts(opts: Optional[ExtMemOpts]) -> _Opts:
    opts = opts or ExtMemOpts()
    geom = 0 if opts.packed_prefix_seed == PackedPrefixSeedPolicy.DISABLED else 1
    return _c_opts(opts.max_context, opts.subproblem_count, int(opts.router), geom, opts.gpu_buckets,
                   opts.tile_keys, batch_records=opts.batch_records,
                   packed_prefix_seed=opts.packed_prefix_seed.value, remap_budget=opts.remap_budget)


def _as_text(text) -> np.ndarray:
    if isinstance(text, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(text), dtype=np.uint8)
    arr = np.ascontiguousarray(text)
    if arr.dtype not in (np.uint8, np.uint16):
        raise CapsError(CAPS_E_UNSUPPORTED, f"symbol type {arr.dtype} is outside the GPU build (u8, u16)")
    return arr


def _limits(lp: Optional[LimitProvider], n: int) -> _Limits:
    return (lp or PlainText(n))._c()


# ---------------------------------------------------------------------------
# In-memory builders (sample_sort.rs:97-206)
# ---------------------------------------------------------------------------
def build_in_memory_for_positions_with(text, positions, lp: Optional[LimitProvider] = None,
                                       opts: Optional[Opts] = None, index_dtype=np.uint32,
                                       return_lcp: bool = False, router: int = 0, gpu_buckets: int = 0,
                                       key_geometry: int = 0):
    """sample_sort.rs:168-206.  Returns the sorted positions (and the LCP
    array, merge_sort's postcondition, when ``return_lcp``)."""
    t = _as_text(text)
    n = len(t)
    idt = np.dtype(index_dtype)
    pos = np.ascontiguousarray(positions, dtype=idt)
    out = np.empty(len(pos), dtype=idt)
    lcp = np.empty(len(pos), dtype=idt) if return_lcp else None
    lim = _limits(lp, n)
    o = _c_opts((opts or Opts()).max_context, router=router, gpu_buckets=gpu_buckets, key_geometry=key_geometry,
                want_lcp=int(return_lcp))
    rc = lib().caps_build_in_memory_for_positions(
        t.ctypes.data, n, t.dtype.itemsize, pos.ctypes.data, len(pos), idt.itemsize, ctypes.byref(lim),
        ctypes.byref(o), out.ctypes.data, lcp.ctypes.data if lcp is not None else None)
    _check(rc)
    return (out, lcp) if return_lcp else out


def build_in_memory_with(text, lp: Optional[LimitProvider] = None, opts: Optional[Opts] = None,
                         index_dtype=np.uint32, return_lcp: bool = False, router: int = 0, gpu_buckets: int = 0,
                         key_geometry: int = 0):
    """sample_sort.rs:118-127."""
    t = _as_text(text)
    n = len(t)
    idt = np.dtype(index_dtype)
    out = np.empty(n, dtype=idt)
    lcp = np.empty(n, dtype=idt) if return_lcp else None
    lim = _limits(lp, n)
    o = _c_opts((opts or Opts()).max_context, router=router, gpu_buckets=gpu_buckets, key_geometry=key_geometry,
                want_lcp=int(return_lcp))
    rc = lib().caps_build_in_memory(t.ctypes.data, n, t.dtype.itemsize, idt.itemsize, ctypes.byref(lim),
                                    ctypes.byref(o), out.ctypes.data,
                                    lcp.ctypes.data if lcp is not None else None)
    _check(rc)
    return (out, lcp) if return_lcp else out


def build_in_memory(text, index_dtype=np.uint32):
    """sample_sort.rs:97-103."""
    return build_in_memory_with(text, None, None, index_dtype)


def build_in_memory_with_opts(text, opts: Opts, index_dtype=np.uint32):
    """sample_sort.rs:106-112."""
    return build_in_memory_with(text, None, opts, index_dtype)


def build_in_memory_for_positions(text, positions, index_dtype=np.uint32):
    """sample_sort.rs:143-149."""
    return build_in_memory_for_positions_with(text, positions, None, None, index_dtype)


def build_in_memory_for_positions_with_opts(text, positions, opts: Opts, index_dtype=np.uint32):
    """sample_sort.rs:152-165."""
    return build_in_memory_for_positions_with(text, positions, None, opts, index_dtype)


class _TextItem(ctypes.Structure):
    _fields_ = [("text", ctypes.c_void_p), ("n", ctypes.c_uint64), ("out_sa", ctypes.c_void_p)]


def build_batch(texts, index_dtype=np.uint32):
    """build_in_memory for every text of the list, as one GPU build per symbol
    width.  Returns the suffix arrays in input order."""
    arrs = [_as_text(t) for t in texts]
    idt = np.dtype(index_dtype)
    outs = [np.empty(len(a), dtype=idt) for a in arrs]
    for width in (1, 2):
        sel = [i for i, a in enumerate(arrs) if a.dtype.itemsize == width]
        if not sel:
            continue
        items = (_TextItem * len(sel))()
        for j, i in enumerate(sel):
            items[j].text = arrs[i].ctypes.data if len(arrs[i]) else None
            items[j].n = len(arrs[i])
            items[j].out_sa = outs[i].ctypes.data if len(arrs[i]) else None
        o = _c_opts()
        _check(lib().caps_build_batch(items, len(sel), width, idt.itemsize, ctypes.byref(o)))
    return outs


# ---------------------------------------------------------------------------
# Streaming builders (ext_mem.rs)
# ---------------------------------------------------------------------------
class BuildError(Exception):
    """ext_mem.rs:257-277: Io (device failure) or Emit (callback error)."""

    def __init__(self, kind: str, inner):
        super().__init__(f"{kind}: {inner}")
        self.kind = kind
        self.inner = inner


def _emit_callback(emit, err_box):
    def _cb(_ctx, ptr, count):
        try:
            arr = np.ctypeslib.as_array(ptr, shape=(count,))
            for v in arr.tolist():
                emit(v)
            return 0
        except Exception as exc:  # the callback's error aborts the build
            err_box.append(exc)
            return 1

    return EMIT_FN(_cb)


def _stream_result(rc, err_box):
    if rc == CAPS_E_EMIT:
        raise BuildError("Emit", err_box[0] if err_box else lib().caps_last_emit_code())
    if rc != CAPS_OK:
        raise BuildError("Io", CapsError(rc, lib().caps_last_error().decode()))


def keep_bitmap(keep, n: int) -> np.ndarray:
    """The kept-position bitmap the crate's FilteredSource builds by walking
    the predicate once (ext_mem.rs:940-990): (n + 63) // 64 u64 words,
    position q kept iff bit q % 64 of word q // 64 is set.  `keep` is a
    boolean array of length n, or a predicate over positions: it is first
    called once on the whole position array (a vectorised predicate), and
    per position if that does not give n booleans."""
    if callable(keep):
        mask = None
        try:
            with warnings.catch_warnings():  # a per-position predicate may reject the array
                warnings.simplefilter("ignore")
                m = np.asarray(keep(np.arange(n, dtype=np.uint64)))
            if m.shape == (n,):
                mask = m.astype(bool)
        except Exception:
            mask = None
        if mask is None:
            mask = np.fromiter((bool(keep(q)) for q in range(n)), dtype=bool, count=n)
    else:
        mask = np.asarray(keep, dtype=bool)
        if mask.shape != (n,):
            raise ValueError("keep mask must have one entry per position")
    words = (n + 63) // 64
    packed = np.packbits(mask, bitorder="little")
    buf = np.zeros(words * 8, dtype=np.uint8)
    buf[: len(packed)] = packed
    return buf.view(np.uint64)


def try_build_ext_mem_for_filter_with(text, keep, lp, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:522-547: the kept positions come from `keep` (see
    keep_bitmap), walked once into a bitmap; no position list is built."""
    t = _as_text(text)
    n = len(t)
    bits = keep_bitmap(keep, n)
    lim = _limits(lp, n)
    o = _ext_c_opts(opts)
    err_box = []
    cb = _emit_callback(emit, err_box)
    rc = lib().caps_build_ext_mem_for_filter(t.ctypes.data, n, t.dtype.itemsize, bits.ctypes.data,
                                             ctypes.byref(lim), ctypes.byref(o), cb, None)
    _stream_result(rc, err_box)


def try_build_ext_mem_for_filter(text, keep, opts, emit):
    """ext_mem.rs:487-499."""
    try_build_ext_mem_for_filter_with(text, keep, None, opts, emit)


def build_ext_mem_for_filter(text, keep, opts, emit):
    """ext_mem.rs:472-484: errors surface as OSError (io::Error)."""
    try:
        try_build_ext_mem_for_filter_with(text, keep, None, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def build_ext_mem_for_filter_with(text, keep, lp, opts, emit):
    """ext_mem.rs:503-519."""
    try:
        try_build_ext_mem_for_filter_with(text, keep, lp, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def _stream(entry, text, positions, lp, opts, emit):
    t = _as_text(text)
    n = len(t)
    lim = _limits(lp, n)
    o = _ext_c_opts(opts)
    err_box = []

    def _cb(_ctx, ptr, count):
        try:
            arr = np.ctypeslib.as_array(ptr, shape=(count,))
            for v in arr.tolist():
                emit(v)
            return 0
        except Exception as exc:  # the callback's error aborts the build
            err_box.append(exc)
            return 1

    cb = EMIT_FN(_cb)
    if positions is None:
        pos_ptr, n_pos = None, n
    else:
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        pos_ptr, n_pos = pos.ctypes.data, len(pos)
    rc = entry(t.ctypes.data, n, t.dtype.itemsize, pos_ptr, n_pos, ctypes.byref(lim), ctypes.byref(o), cb, None)
    if rc == CAPS_E_EMIT:
        raise BuildError("Emit", err_box[0] if err_box else lib().caps_last_emit_code())
    if rc != CAPS_OK:
        raise BuildError("Io", CapsError(rc, lib().caps_last_error().decode()))


def try_build_in_memory_sample_sort_with(text, lp, opts: Optional[ExtMemOpts], emit: Callable[[int], None]):
    """ext_mem.rs:755-783."""
    _stream(lib().caps_build_in_memory_sample_sort, text, None, lp, opts, emit)


def try_build_in_memory_sample_sort(text, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:726-736."""
    try_build_in_memory_sample_sort_with(text, None, opts, emit)


def build_in_memory_sample_sort(text, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:717-723: errors surface as OSError (io::Error)."""
    try:
        try_build_in_memory_sample_sort(text, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def build_in_memory_sample_sort_with(text, lp, opts, emit):
    """ext_mem.rs:740-752."""
    try:
        try_build_in_memory_sample_sort_with(text, lp, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def try_build_ext_mem_with(text, lp, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:324-356: every position of the text.  On the GPU this is the
    same build as the in-memory sample sort over all positions (the crate's
    two paths differ only in where buckets are kept)."""
    _stream(lib().caps_build_in_memory_sample_sort, text, None, lp, opts, emit)


def try_build_ext_mem(text, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:298-307."""
    try_build_ext_mem_with(text, None, opts, emit)


def build_ext_mem(text, opts: Optional[ExtMemOpts], emit):
    """ext_mem.rs:289-295: errors surface as OSError (io::Error)."""
    try:
        try_build_ext_mem(text, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def build_ext_mem_with(text, lp, opts, emit):
    """ext_mem.rs:315-321."""
    try:
        try_build_ext_mem_with(text, lp, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def try_build_in_memory_sample_sort_for_positions_with(text, positions, lp, opts, emit):
    """ext_mem.rs:842-871."""
    _stream(lib().caps_build_in_memory_sample_sort, text, positions, lp, opts, emit)


def build_in_memory_sample_sort_for_positions(text, positions, opts, emit):
    """ext_mem.rs:787-800."""
    try:
        try_build_in_memory_sample_sort_for_positions_with(text, positions, None, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def try_build_ext_mem_for_positions_with(text, positions, lp, opts, emit):
    """ext_mem.rs:417-449 (records stay in HBM; no disk spill)."""
    _stream(lib().caps_build_ext_mem_for_positions, text, positions, lp, opts, emit)


def build_ext_mem_for_positions(text, positions, opts, emit):
    """ext_mem.rs:370-381."""
    try:
        try_build_ext_mem_for_positions_with(text, positions, None, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


def build_ext_mem_for_positions_with(text, positions, lp, opts, emit):
    """ext_mem.rs:399-414."""
    try:
        try_build_ext_mem_for_positions_with(text, positions, lp, opts, emit)
    except BuildError as e:
        raise OSError(str(e)) from e


# ---------------------------------------------------------------------------
# Kernel-level entry points (for parity tests)
# ---------------------------------------------------------------------------
def merge_lcp(text, x, lcp_x, y, lcp_y, lp=None, max_context=MAX_CONTEXT):
    """sample_sort.rs:505-522 on the GPU (merge-path LCP merge)."""
    t = _as_text(text)
    x, lx, y, ly = (np.ascontiguousarray(a, dtype=np.uint64) for a in (x, lcp_x, y, lcp_y))
    z = np.empty(len(x) + len(y), dtype=np.uint64)
    lz = np.empty_like(z)
    lim = _limits(lp, len(t))
    _check(lib().caps_merge_lcp(t.ctypes.data, len(t), t.dtype.itemsize, ctypes.byref(lim), x.ctypes.data,
                                lx.ctypes.data, len(x), y.ctypes.data, ly.ctypes.data, len(y),
                                min(max_context, MAX_CONTEXT), z.ctypes.data, lz.ctypes.data))
    return z, lz


def cascade_merge(text, sa, lcp, boundaries, lp=None, max_context=MAX_CONTEXT):
    """ext_mem.rs:2144-2243 on the GPU."""
    t = _as_text(text)
    sa, lcp, b = (np.ascontiguousarray(a, dtype=np.uint64) for a in (sa, lcp, boundaries))
    out = np.empty_like(sa)
    lout = np.empty_like(sa)
    lim = _limits(lp, len(t))
    _check(lib().caps_cascade_merge(t.ctypes.data, len(t), t.dtype.itemsize, ctypes.byref(lim), sa.ctypes.data,
                                    lcp.ctypes.data, len(sa), b.ctypes.data, len(b),
                                    min(max_context, MAX_CONTEXT), out.ctypes.data, lout.ctypes.data))
    return out, lout


__all__ = [name for name in dir() if not name.startswith("_")]
