"""CPU checks of the on-disk formats and of the command-line tool's argument
handling (no build runs here).

Formats: suffix-array files are packed little-endian u32[] / u64[]
(examples/caps_sa.rs:117-190); SaLcp records are (pos, lcp) little-endian,
8 or 16 bytes (ext_bucket.rs:65-106), a run's first lcp stored as 0
(ext_bucket.rs:203-218).  The byte images below are written out by hand from
that layout, and the C++ header (host/formats.hpp) and the Python module
(caps_sa.formats) are checked against them and against each other.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from caps_sa import formats as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "caps-sa_amd", "host")
CLI = os.path.join(HOST, "caps_sa")


def test_salcp_byte_layout():
    # pos first, then lcp, each little-endian
    r32 = F.salcp_run([0x01020304, 0x11223344], [0xDEAD, 0x0A0B0C0D], 4)
    assert r32.tobytes() == bytes([4, 3, 2, 1, 0, 0, 0, 0, 0x44, 0x33, 0x22, 0x11, 0x0D, 0x0C, 0x0B, 0x0A])
    r64 = F.salcp_run([0x0102030405060708], [7], 8)
    assert r64.tobytes() == bytes([8, 7, 6, 5, 4, 3, 2, 1] + [0] * 8)
    assert F.SALCP_U32.itemsize == 8 and F.SALCP_U64.itemsize == 16


def test_sa_file_layout(tmp_path):
    sa = np.array([3, 0x01020304, 0], dtype=np.uint64)
    F.write_sa(tmp_path / "a4", sa, 4)
    assert open(tmp_path / "a4", "rb").read() == bytes([3, 0, 0, 0, 4, 3, 2, 1, 0, 0, 0, 0])
    F.write_sa(tmp_path / "a8", sa, 8)
    assert len(open(tmp_path / "a8", "rb").read()) == 24
    assert F.read_sa(tmp_path / "a8", 8).tolist() == sa.tolist()
    with pytest.raises(ValueError):
        F.read_sa(tmp_path / "a4", 8)  # 12 bytes is not a whole number of u64
    with pytest.raises(ValueError):
        F.write_sa(tmp_path / "x", sa, 2)


def test_sa_width_choice():
    # examples/caps_sa.rs:117: u32 only for the in-memory path on n <= 2^32
    assert F.sa_width(10, False) == 4 and F.sa_width(1 << 32, False) == 4
    assert F.sa_width((1 << 32) + 1, False) == 8
    assert F.sa_width(10, True) == 8


def test_salcp_soa_round_trip(tmp_path):
    """The crate's soa round trip (ext_bucket.rs:1180-1197): two runs, the
    first lcp of each stored as 0, boundaries [0, 3, 5000]."""
    positions = np.arange(5000, dtype=np.uint64)
    lcps = positions + 10
    for w in (4, 8):
        path = tmp_path / f"soa{w}"
        bounds = F.write_salcp(path, [(positions[:3], lcps[:3]), (positions[3:], lcps[3:])], w)
        assert bounds == [0, 3, 5000]
        got_pos, got_lcp = F.read_salcp(path, w)
        assert got_pos.tolist() == positions.tolist()
        assert got_lcp[0] == 0 and got_lcp[3] == 0
        got_lcp[0], got_lcp[3] = lcps[0], lcps[3]
        assert got_lcp.tolist() == lcps.tolist()


def test_salcp_round_trip_below_buffer(tmp_path):
    """ext_bucket.rs:1218-1237: ten records (i, 2i) added one by one (no
    lcp reset) read back unchanged."""
    recs = np.zeros(10, dtype=F.SALCP_U64)
    recs["pos"] = np.arange(10)
    recs["lcp"] = np.arange(10) * 2
    recs.tofile(tmp_path / "r")
    pos, lcp = F.read_salcp(tmp_path / "r", 8)
    assert pos.tolist() == list(range(10)) and lcp.tolist() == [2 * i for i in range(10)]


CPP = r"""
#include <cstdio>
#include <vector>
#include "formats.hpp"
using namespace caps_sa::formats;
int main(int argc, char **argv) {
    const std::string d = argv[1];
    std::vector<uint32_t> a = {3u, 0x01020304u, 0u};
    write_sa<uint32_t>(d + "/cpp_sa4", a.data(), a.size());
    std::vector<uint64_t> b = {0x0102030405060708ull, 5ull};
    write_sa<uint64_t>(d + "/cpp_sa8", b.data(), b.size());
    {
        LeWriter w(d + "/cpp_salcp4");
        std::vector<uint32_t> p = {9, 8, 7}, l = {55, 1, 2};
        write_salcp_run<uint32_t>(w, p.data(), l.data(), 3);
        write_salcp_run<uint32_t>(w, p.data(), l.data(), 2);
        w.close();
    }
    std::vector<uint64_t> pos, lcp;
    read_salcp<uint64_t>(d + "/py_salcp8", pos, lcp);
    for (size_t i = 0; i < pos.size(); ++i) std::printf("%llu %llu\n", (unsigned long long)pos[i], (unsigned long long)lcp[i]);
    std::vector<uint32_t> back = read_sa<uint32_t>(d + "/cpp_sa4");
    std::printf("sa4 %zu %u\n", back.size(), back[1]);
    try { read_sa<uint64_t>(d + "/cpp_sa4"); } catch (const std::exception &) { std::printf("bad width rejected\n"); }
    return 0;
}
"""


def test_cpp_and_python_formats_agree(tmp_path):
    src = tmp_path / "f.cpp"
    src.write_text(CPP)
    exe = tmp_path / "f"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-I", HOST, "-o", str(exe), str(src)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    F.write_salcp(tmp_path / "py_salcp8", [([1 << 40, 2], [5, 6]), ([3], [9])], 8)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert out[:3] == [f"{1 << 40} 0", "2 6", "3 0"]
    assert out[3] == f"sa4 3 {0x01020304}" and out[4] == "bad width rejected"
    assert F.read_sa(tmp_path / "cpp_sa4", 4).tolist() == [3, 0x01020304, 0]
    assert F.read_sa(tmp_path / "cpp_sa8", 8).tolist() == [0x0102030405060708, 5]
    pos, lcp = F.read_salcp(tmp_path / "cpp_salcp4", 4)
    assert pos.tolist() == [9, 8, 7, 9, 8] and lcp.tolist() == [0, 1, 2, 0, 1]


def _cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(CLI), reason="caps_sa tool not built")
def test_cli_arguments():
    """examples/caps_sa.rs:32-92: usage on --help (exit 0), exactly two
    positional args (exit 2), integer values (the crate panics: exit 101),
    unknown flags count as positional."""
    r = _cli("--help")
    assert r.returncode == 0 and r.stderr.startswith("usage: caps_sa <input> <output>")
    assert _cli("-h").returncode == 0
    r = _cli("in")
    assert r.returncode == 2 and "expected 2 positional args (input, output), got 1" in r.stderr
    r = _cli("in", "out", "--bogus")
    assert r.returncode == 2 and "got 3" in r.stderr
    for flag in ("--subproblem-count", "--threads"):
        r = _cli("in", "out", flag, "x7")
        assert r.returncode == 101 and f"{flag} expects a positive integer" in r.stderr
        assert _cli("in", "out", flag).returncode == 101
    with tempfile.TemporaryDirectory() as d:
        r = _cli(os.path.join(d, "missing"), os.path.join(d, "out"))
        assert r.returncode == 1 and r.stderr.startswith("Error: cannot open")
