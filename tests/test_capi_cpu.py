"""CPU checks of the C ABI library and the host-side mirror: the library
loads, exports every entry point include/caps_sa.h declares, and rejects
invalid arguments before touching a device.  No compute call runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import caps_sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "caps_sa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(caps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = caps_sa.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), f"{name} declared in caps_sa.h but not exported"


def test_python_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(caps_sa._SIGNATURES)


def test_struct_layouts():
    assert ctypes.sizeof(caps_sa._Limits) == 24
    assert ctypes.sizeof(caps_sa._Opts) == 64


def test_opts_default():
    o = caps_sa._c_opts()
    assert o.max_context == caps_sa.MAX_CONTEXT and o.subproblem_count == 0


def _call_invalid(sym_bytes=1, idx_bytes=4, n=4, text=b"abcd", lim=None):
    t = np.frombuffer(text, dtype=np.uint8) if text is not None else None
    out = np.zeros(max(n, 1), dtype=np.uint64)
    lim = lim or caps_sa.PlainText(n)._c()
    o = caps_sa._c_opts()
    return caps_sa.lib().caps_build_in_memory(t.ctypes.data if t is not None else None, n, sym_bytes, idx_bytes,
                                             ctypes.byref(lim), ctypes.byref(o), out.ctypes.data, None)


def test_argument_validation_before_device():
    assert _call_invalid(sym_bytes=4) == caps_sa.CAPS_E_UNSUPPORTED
    assert _call_invalid(idx_bytes=2) == caps_sa.CAPS_E_UNSUPPORTED
    assert _call_invalid(text=None) == caps_sa.CAPS_E_INVALID
    assert "NULL" in caps_sa.lib().caps_last_error().decode()


def test_positions_are_checked_before_device():
    """Host positions of n or more, or NULL positions with n_pos != n, are
    refused with CAPS_E_INVALID before any device work (so also without a
    GPU); the reference panics on such an index."""
    t = np.frombuffer(b"abcdabcd", dtype=np.uint8)
    out = np.zeros(8, dtype=np.uint64)
    lim = caps_sa.PlainText(8)._c()
    o = caps_sa._c_opts()
    L = caps_sa.lib()
    for idx_bytes, dt in ((4, np.uint32), (8, np.uint64)):
        pos = np.array([0, 3, 8], dtype=dt)
        rc = L.caps_build_in_memory_for_positions(t.ctypes.data, 8, 1, pos.ctypes.data, 3, idx_bytes,
                                                  ctypes.byref(lim), ctypes.byref(o), out.ctypes.data, None)
        assert rc == caps_sa.CAPS_E_INVALID
        assert "out of range" in L.caps_last_error().decode()
    rc = L.caps_build_in_memory_for_positions(t.ctypes.data, 8, 1, None, 5, 4, ctypes.byref(lim), ctypes.byref(o),
                                              out.ctypes.data, None)
    assert rc == caps_sa.CAPS_E_INVALID


def test_signed_symbols_are_unsupported():
    with pytest.raises(caps_sa.CapsError) as e:
        caps_sa.build_in_memory(np.array([0, 1, -1], dtype=np.int8))
    assert e.value.code == caps_sa.CAPS_E_UNSUPPORTED


def test_segmented_text_constructor_checks():
    # limits.rs:247-281 panics -> ValueError here
    s = caps_sa.SegmentedText.from_lengths(15, [3, 5, 7])
    assert s.ends.tolist() == [3, 8, 15] and s.n_segments() == 3
    assert [s.lim_at(p) for p in (0, 1, 2, 3, 5, 7, 8, 14, 15)] == [3, 2, 1, 5, 3, 1, 7, 1, 0]
    with pytest.raises(ValueError, match="sum to"):
        caps_sa.SegmentedText.from_lengths(20, [3, 5, 7])
    with pytest.raises(ValueError):
        caps_sa.SegmentedText.from_ends(15, [3, 3, 15])
    with pytest.raises(ValueError):
        caps_sa.SegmentedText.from_ends(16, [3, 8, 15])
    assert caps_sa.SegmentedText.from_lengths(0, []).n_segments() == 0


def test_boundary_rank():
    assert caps_sa.PlainText(3).boundary_rank() == caps_sa.BoundaryRank.SHORTER_FIRST
    assert caps_sa.StarSegmented(4, [4]).boundary_rank() == caps_sa.BoundaryRank.LONGER_FIRST


def test_ext_mem_opts_from_env(monkeypatch):
    monkeypatch.setenv("CAPS_SA_SUBPROBLEMS", "17")
    monkeypatch.setenv("CAPS_SA_PACKED_PREFIX_SEED", "on")
    o = caps_sa.ExtMemOpts.from_env()
    assert o.subproblem_count == 17
    assert o.packed_prefix_seed == caps_sa.PackedPrefixSeedPolicy.DENSE_ALPHABET_ONLY
    monkeypatch.setenv("CAPS_SA_PACKED_PREFIX_REMAP_BYTES", "100")
    assert caps_sa.ExtMemOpts.from_env().packed_prefix_seed == caps_sa.PackedPrefixSeedPolicy.REMAP


def test_cpp_api_mirror_header_compiles():
    """The C++ mirror of the crate's API is self-contained C++17 over the C ABI."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = "#include \"caps_sa.hpp\"\nint main() { return 0; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I",
                        os.path.join(root, "caps-sa_amd", "host"), "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_struct_sizes_match_the_bindings():
    """caps_opts_t / caps_stats_t as the C compiler lays them out equal the
    ctypes mirrors."""
    import subprocess
    import tempfile

    src = ("#include <stdio.h>\n#include \"caps_sa.h\"\n"
           "int main(void) { printf(\"%zu %zu %zu\\n\", sizeof(caps_opts_t), sizeof(caps_stats_t), "
           "sizeof(caps_limits_t)); return 0; }\n")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "sizes")
        r = subprocess.run(["gcc", "-x", "c", "-I", os.path.join(ROOT, "include"), "-o", exe, "-"], input=src,
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [ctypes.sizeof(caps_sa._Opts), ctypes.sizeof(caps_sa.Stats),
                                     ctypes.sizeof(caps_sa._Limits)]


def test_keep_bitmap_layout():
    """Position q kept iff bit q % 64 of word q // 64 (the crate's
    FilteredSource bitmap, ext_mem.rs:940-990), from a mask, a vectorised
    predicate or a per-position predicate."""
    rng = np.random.default_rng(3)
    for n in (0, 1, 63, 64, 65, 1000):
        mask = rng.random(n) < 0.4
        for keep in (mask, lambda q: mask[q.astype(np.int64)], lambda q: bool(mask[int(q)])):
            bits = caps_sa.keep_bitmap(keep, n)
            assert bits.dtype == np.uint64 and len(bits) == (n + 63) // 64
            back = [(int(bits[q >> 6]) >> (q & 63)) & 1 for q in range(n)]
            assert back == mask.astype(int).tolist()
            assert all(int(bits[-1]) >> b == 0 for b in range(n % 64, 64) if n % 64) if n else True


def test_memoization_policy_from_env(monkeypatch):
    """ext_mem.rs:154-196 and lcp_memo.rs:11-14, 64-145: the geometric memo
    is off by default; CAPS_SA_GEOMETRIC_MEMO turns it on with the defaults
    (256, 1024, 64, 4096); its tuning variables count only then, and zero or
    malformed values keep the defaults."""
    for var in ("CAPS_SA_GEOMETRIC_MEMO", "CAPS_SA_MEMO_PROBE", "CAPS_SA_MEMO_MIN_LCP",
                "CAPS_SA_MEMO_ACTIVATE_ENTRIES", "CAPS_SA_MEMO_CAPACITY", "CAPS_SA_MEMO_STATS"):
        monkeypatch.delenv(var, raising=False)
    o = caps_sa.ExtMemOpts.from_env()
    assert o.lcp_memoization is None and not o.collect_lcp_memoization_stats
    monkeypatch.setenv("CAPS_SA_MEMO_PROBE", "9")  # ignored while the memo is off
    assert caps_sa.ExtMemOpts.from_env().lcp_memoization is None
    monkeypatch.setenv("CAPS_SA_GEOMETRIC_MEMO", "yes")
    monkeypatch.setenv("CAPS_SA_MEMO_MIN_LCP", "0")
    monkeypatch.setenv("CAPS_SA_MEMO_ACTIVATE_ENTRIES", "12x")
    monkeypatch.setenv("CAPS_SA_MEMO_CAPACITY", "+77")
    monkeypatch.setenv("CAPS_SA_MEMO_STATS", "on")
    o = caps_sa.ExtMemOpts.from_env()
    assert o.lcp_memoization == caps_sa.GeometricMemoizationConfig(9, 1024, 64, 77)
    assert o.collect_lcp_memoization_stats
    monkeypatch.setenv("CAPS_SA_GEOMETRIC_MEMO", "TRUE")  # only 1/true/yes/on
    assert caps_sa.ExtMemOpts.from_env().lcp_memoization is None
    d = caps_sa.LcpMemoizationPolicy.geometric()
    assert (d.probe_symbols, d.min_lcp_symbols, d.activate_after_entries, d.max_entries_per_partition) == \
        (256, 1024, 64, 4096)
    assert d.with_probe_symbols(5).probe_symbols == 5 and d.probe_symbols == 256
    with pytest.raises(ValueError):
        d.with_max_entries_per_partition(0)


def test_cpp_mirror_memoization_from_env():
    """The C++ mirror parses the same variables the same way."""
    import subprocess
    import tempfile

    src = r'''
#include <cstdio>
#include "caps_sa.hpp"
int main() {
    auto o = caps_sa::ExtMemOpts::from_env();
    if (!o.lcp_memoization) { std::printf("off %d\n", (int)o.collect_lcp_memoization_stats); return 0; }
    const auto &c = *o.lcp_memoization;
    std::printf("%llu %llu %llu %llu %d\n", (unsigned long long)c.probe_symbols, (unsigned long long)c.min_lcp_symbols,
                (unsigned long long)c.activate_after_entries, (unsigned long long)c.max_entries_per_partition,
                (int)o.collect_lcp_memoization_stats);
    try { c.with_probe_symbols(0); } catch (const std::invalid_argument &) { std::printf("zero rejected\n"); }
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "memo")
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-x", "c++", "-I", os.path.join(ROOT, "caps-sa_amd", "host"),
                            "-o", exe, "-"], input=src, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        base = {k: v for k, v in os.environ.items() if not k.startswith("CAPS_SA_")}
        run = lambda env: subprocess.run([exe], env={**base, **env}, capture_output=True, text=True).stdout.split("\n")
        assert run({"CAPS_SA_MEMO_PROBE": "9"})[0] == "off 0"
        out = run({"CAPS_SA_GEOMETRIC_MEMO": "1", "CAPS_SA_MEMO_PROBE": "9", "CAPS_SA_MEMO_MIN_LCP": "0",
                   "CAPS_SA_MEMO_ACTIVATE_ENTRIES": "12x", "CAPS_SA_MEMO_CAPACITY": "+77", "CAPS_SA_MEMO_STATS": "on"})
        assert out[0] == "9 1024 64 77 1" and out[1] == "zero rejected"
