"""The reference crate's fixture tests through the C++ mirror of its API
(caps-sa_amd/host/caps_sa.hpp), built by __graft_entry__.build() as
caps-sa_amd/host/test_host and run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "caps-sa_amd", "host", "test_host")


def test_cpp_api_mirror_fixtures():
    assert os.path.exists(EXE), "caps-sa_amd/host/test_host is not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
