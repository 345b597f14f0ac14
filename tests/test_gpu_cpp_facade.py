"""Runs the C++ facade test (tests/cpp/test_facade.cpp, built by
__graft_entry__.build()): the reference's BVH and SignedDistance known
answers through axom::spin::BVH<3, MI355X_EXEC> and
axom::quest::SignedDistance<3, MI355X_EXEC>."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_facade")


def test_cpp_facade_known_answers():
    assert os.path.exists(BIN), "tests/cpp/test_facade is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
