"""The C++ SimpleAssembler mirror (nuto_amd/host) against the oracle, through its own C++ test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_host_shim")


def test_cpp_simple_assembler_mirror():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
