"""CPU checks of the C-ABI library: it loads without a GPU, exports every
function that include/*.h declares, and fails cleanly where it needs one."""
import ctypes
import os
import subprocess

import pytest

import libnt_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_is_built_and_loads():
    assert os.path.exists(libnt_amd.LIB_PATH), "run python -m libnt_amd.build"
    libnt_amd.lib()


def test_every_declared_function_is_exported():
    names = libnt_amd.exported_symbols_declared_in_headers()
    assert "lnt_change_sort_order" in names and "lnt_merge_into" in names
    out = subprocess.run(["nm", "-D", "--defined-only", libnt_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_binding_covers_the_header():
    names = set(libnt_amd.exported_symbols_declared_in_headers())
    assert names <= set(libnt_amd.SIGNATURES), names - set(libnt_amd.SIGNATURES)


def test_abi_version_and_null_context_errors():
    L = libnt_amd.lib()
    assert L.lnt_abi_version() == 2
    assert L.lnt_context_synchronize(None) == 1
    assert b"null context" in L.lnt_last_error()
    assert L.lnt_context_destroy(None) == 0
    assert L.lnt_free(None, None, 0) == 0
    p = ctypes.c_void_p()
    assert L.lnt_alloc(None, 8, 1, ctypes.byref(p)) == 1 and not p.value


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "lnt.h"\n#include "lnt_testgen.h"\n#include "lnt_mex.h"\nint main(void){return lnt_abi_version()==LNT_ABI_VERSION?0:1;}\n')
    for cc in ("gcc", "g++"):
        exe = tmp_path / ("t_" + cc)
        cmd = [cc, "-x", "c" if cc == "gcc" else "c++", str(src), "-I", os.path.join(ROOT, "include"),
               "-L", os.path.dirname(libnt_amd.LIB_PATH), "-llnt", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
