"""The C++ host mirror (pis_amd/host/pis.hpp) on the GPU: runs tests/cpp/test_manager,
built by __graft_entry__.build(), which compares LJGpuManager with the CPU oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "test_manager")


@pytest.mark.gpu
def test_cpp_host_mirror_parity():
    assert os.path.exists(EXE), "build it with __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "OK" in r.stdout
