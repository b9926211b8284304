"""The C++ drop-in header include/gfa_action.hxx compiles against include/gfa.h, links libgfa.so and
builds a plan through the C-ABI (host only)."""
import os
import shutil
import subprocess

import pytest

from angstromcube_amd import gfa as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_cxx_drop_in(tmp_path):
    exe = tmp_path / "drop_in"
    libdir = os.path.dirname(G.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "drop_in.cpp"), "-L", libdir, "-lgfa",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "nnzb 512" in out.stdout and "rejected" in out.stdout
    assert "keys 64021 32021 64022 32022" in out.stdout and "noco2 refused" in out.stdout
