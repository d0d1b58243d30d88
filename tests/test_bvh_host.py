"""Host BVH builder invariants (no GPU): compiled with g++ from the product's
bvh.cpp and run on random, degenerate and flat triangle sets."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe():
    d = tempfile.mkdtemp()
    out = os.path.join(d, "bvh_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", out,
                    os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"),
                    os.path.join(ROOT, "xitils_amd", "csrc", "bvh.cpp")], check=True)
    return out


@pytest.mark.parametrize("n,mode", [(1, 0), (3, 0), (5, 0), (1000, 0), (200000, 0), (5000, 1), (20000, 2)])
def test_bvh_invariants(exe, n, mode):
    r = subprocess.run([exe, str(n), str(mode)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    nodes, depth = map(int, r.stdout.split()[1:3])
    # the traversal stack bound used by the kernels: (width - 1) * depth + 1 entries
    assert depth >= 1 and nodes >= 1
