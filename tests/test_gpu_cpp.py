"""Runs the C++ facade test (tests/cpp/test_facade_render.cpp, built by
__graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "test_facade_render")


def test_cpp_facade_render_matches_oracle():
    assert os.path.exists(EXE), "C++ test not built: run __graft_entry__.build()"
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "passed" in r.stdout
