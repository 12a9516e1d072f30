"""Builds the in-tree native libraries of the package.

  libcaps_sa_hip.so  hipcc, gfx950: kernels + C ABI (csrc/engine.hip)
  libcaps_gen.so     gcc: seeded workload generators (gen/gen.c)
  host/test_host     g++: test program of the C++ API mirror
  host/caps_sa       g++: command-line builder (host/caps_sa_cli.cpp)
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_hip() -> str:
    out = os.path.join(HERE, "libcaps_sa_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join("csrc", "engine.hip")])
    os.replace(tmp, out)
    return out


def build_gen() -> str:
    out = os.path.join(HERE, "libcaps_gen.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-o", out, os.path.join("gen", "gen.c"), "-lm"])
    return out


def build_host_test() -> str:
    """The C++ API mirror's test program (host/test_host.cpp), linked against
    the HIP library built above."""
    out = os.path.join(HERE, "host", "test_host")
    _run(["g++", "-O2", "-std=c++17", "-o", out, os.path.join("host", "test_host.cpp"), "-L" + HERE, "-lcaps_sa_hip",
          "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_cli() -> str:
    """The command-line builder (host/caps_sa_cli.cpp), linked against the
    HIP library built above."""
    out = os.path.join(HERE, "host", "caps_sa")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out, os.path.join("host", "caps_sa_cli.cpp"), "-L" + HERE,
          "-lcaps_sa_hip", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all() -> None:
    build_gen()
    build_hip()
    build_host_test()
    build_cli()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libcaps_sa_hip.so"))
