"""The C++ host interface (parallelcmacionize_amd/host/cmi.hpp) used the way
the reference's programs use DensitySubGrid: tests/cpp/test_facade.cpp (built
by __graft_entry__.build()) runs the reference's 54 direction known answers,
per-packet vs per-buffer interact, the MemorySpace spill and a Stromgren run
through libcmi.so."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_facade_program():
    exe = os.path.join(HERE, "cpp", "test_facade")
    assert os.path.exists(exe), "tests/cpp/test_facade not built"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=180)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "All tests successful." in r.stdout
