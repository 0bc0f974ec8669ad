"""The header-only C++ front end (include/lnt/SparseMIA.hpp): it compiles and
links here; on the GPU its test program checks sort, merge, contraction and
Hadamard results against dense brute force."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_front_end.cpp")
LIBDIR = os.path.join(ROOT, "libnt_amd")


def _build(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", SRC, "-I", os.path.join(ROOT, "include"),
           "-L", LIBDIR, "-llnt", "-Wl,-rpath," + LIBDIR, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_front_end_compiles_and_links(tmp_path):
    _build(str(tmp_path / "fe"))


@pytest.mark.gpu
def test_front_end_on_gpu(tmp_path):
    exe = _build(str(tmp_path / "fe"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "all checks passed" in r.stdout
