"""Build recipe of libpis_hip.so (gfx950), in-tree so the built object travels with the
source tree.  Run as `python -m pis_amd.build`."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["csrc/kernels.hip", "csrc/engine.hip", "csrc/server.hip", "csrc/noalloc.hip", "csrc/dump.cpp"]
OUT = os.path.join(HERE, "libpis_hip.so")

# -ffp-contract=off: the reference (Rust) never fuses a*b+c; kernels ask for fma() where
# they want it.  Packed FP32 instructions are disabled for every HIP translation unit.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False, out: str = OUT, defines=()) -> str:
    """Compile libpis_hip.so; `defines` are extra -D tuning knobs (e.g. PIS_FORCE_BATCH=4)."""
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libpis_hip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


def build_cpp_test() -> str:
    """The C++ host-mirror parity test (tests/cpp/test_manager.cpp), host code only: g++
    against libpis_hip.so and the oracle, with $ORIGIN-relative run paths."""
    src = os.path.join(ROOT, "tests", "cpp", "test_manager.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "test_manager")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", out, src,
           "-L", HERE, "-lpis_hip", "-L", os.path.join(ROOT, "oracle"), "-l:liboracle.so",
           "-Wl,-rpath,$ORIGIN/../../../pis_amd:$ORIGIN/../../../oracle"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for tests/cpp/test_manager")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
