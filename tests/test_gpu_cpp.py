"""Runs the C++ facade tests (tests/cpp/*.cpp, built by __graft_entry__.build())
on the GPU: the MultiRenderer render test and the dr::MultiRendererDR test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build")


def _run(name):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), "C++ test not built: run __graft_entry__.build()"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "passed" in r.stdout and "FAILED" not in r.stdout


def test_cpp_facade_render_matches_oracle():
    _run("test_facade_render")


def test_cpp_facade_dr():
    _run("test_facade_dr")
