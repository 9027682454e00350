"""The library's host C++ (lattice, exact Delaunay insertion, Voronoi dual, grad_coeff, ILU(0)
ordering, CSV / VTU output) built with AddressSanitizer and UndefinedBehaviorSanitizer and run
through csrc/host_check.cpp (make sanitize): no memory error, no undefined behaviour, no leak."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "sanitize"], cwd=os.path.join(ROOT, "vfvchem-vfvm3d_amd"), capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host check ok" in r.stdout
