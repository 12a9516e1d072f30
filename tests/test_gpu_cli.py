"""The command-line tool (host/caps_sa) end to end on the GPU: its output file
equals the oracle's suffix array in the width the crate's tool writes
(examples/caps_sa.rs:117-190: u32 for the in-memory path, u64 streamed for
--ext-mem / --in-mem-ss), and its stderr report has the crate's three lines."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from caps_sa import formats as F
from caps_sa import gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "caps-sa_amd", "host", "caps_sa")

READ = re.compile(r"^read: n=(\d+) bytes in \d+\.\d{3}s$")
BUILD = re.compile(r"^build: mode=(in-mem\(u32\)|in-mem\(u64\)|ext-mem\(u64,stream\)|in-mem-ss\(u64,stream\)) "
                   r"n=(\d+) entries in \d+\.\d{3}s$")
WRITE = re.compile(r"^write: \d+\.\d{3}s$")


def _texts():
    rng = np.random.default_rng(11)
    return [
        ("markov-repeats-200k", gen.markov_repeats(200_000, seed=5)),
        ("bytes-256-100k", rng.integers(0, 256, 100_000, dtype=np.uint8)),
        ("runs-50k", np.repeat(rng.integers(0, 3, 500, dtype=np.uint8), 100)),
        ("one-symbol", np.array([7], dtype=np.uint8)),
        ("empty", np.zeros(0, dtype=np.uint8)),
    ]


@pytest.mark.parametrize("mode", [[], ["--ext-mem"], ["--in-mem-ss", "--subproblem-count", "16", "--threads", "4"]],
                         ids=["in-mem", "ext-mem", "in-mem-ss"])
def test_cli_output_matches_oracle(tmp_path, mode):
    assert os.path.exists(CLI), "caps_sa tool not built"
    for label, text in _texts():
        n = len(text)
        print(f"case cli {label} n={n} mode={mode}", flush=True)
        src, dst = tmp_path / f"{label}.bin", tmp_path / f"{label}.sa"
        text.tofile(src)
        r = subprocess.run([CLI, str(src), str(dst), *mode], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = r.stderr.strip().split("\n")
        assert len(lines) == 3, r.stderr
        m_read, m_build = READ.match(lines[0]), BUILD.match(lines[1])
        assert m_read and int(m_read.group(1)) == n
        assert m_build and int(m_build.group(2)) == n
        assert WRITE.match(lines[2])
        sample_sort = bool(mode)
        width = F.sa_width(n, sample_sort)
        want_mode = ("ext-mem(u64,stream)" if "--ext-mem" in mode else "in-mem-ss(u64,stream)") if sample_sort \
            else "in-mem(u32)"
        assert m_build.group(1) == want_mode
        assert os.path.getsize(dst) == n * width
        got = F.read_sa(dst, width).astype(np.uint64)
        want = O.Ctx(text).build_in_memory(threads=4) if n else np.zeros(0, dtype=np.uint64)
        assert np.array_equal(got, want), label


def test_build_ext_mem_every_position():
    """build_ext_mem / build_ext_mem_with (ext_mem.rs:289-356) emit every
    position in suffix order."""
    import caps_sa

    text = gen.markov_repeats(60_000, seed=9)
    want = O.Ctx(text).build_in_memory(threads=4)
    got = []
    caps_sa.build_ext_mem(text, caps_sa.ExtMemOpts(), got.append)
    assert np.array_equal(np.array(got, dtype=np.uint64), want)
    got = []
    caps_sa.build_ext_mem_with(text, caps_sa.PlainText(len(text)), caps_sa.ExtMemOpts(subproblem_count=5), got.append)
    assert np.array_equal(np.array(got, dtype=np.uint64), want)
    # the geometric memo policy is accepted and leaves the output unchanged
    got = []
    caps_sa.build_ext_mem(text, caps_sa.ExtMemOpts(lcp_memoization=caps_sa.LcpMemoizationPolicy.geometric(),
                                                   collect_lcp_memoization_stats=True), got.append)
    assert np.array_equal(np.array(got, dtype=np.uint64), want)
