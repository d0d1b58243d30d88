"""Host BVH builder invariants (no GPU): compiled with g++ from the product's
bvh.cpp and run on random, degenerate and flat triangle sets."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# the default layout (quantised 4-wide nodes), the same with the top four
# levels numbered breadth-first for LDS staging, and the float 4- and 8-wide
# ones
LAYOUTS = {"quant4": [], "quant4_top4": ["-DXT_BVH_TOP_LEVELS=4"], "float4": ["-DXT_BVH_QUANT=0"],
           "float8": ["-DXT_BVH_QUANT=0", "-DXT_BVH_WIDTH=8"]}


@pytest.fixture(scope="module", params=list(LAYOUTS))
def exe(request):
    d = tempfile.mkdtemp()
    out = os.path.join(d, "bvh_check_" + request.param)
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", *LAYOUTS[request.param], "-o", out,
                    os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"),
                    os.path.join(ROOT, "xitils_amd", "csrc", "bvh.cpp")], check=True)
    return out


@pytest.mark.parametrize("n,mode", [(0, 0), (1, 0), (3, 0), (5, 0), (1000, 0), (200000, 0), (5000, 1), (20000, 2),
                                    (20000, 3), (5000, 4), (5000, 5)])
def test_bvh_invariants(exe, n, mode):
    r = subprocess.run([exe, str(n), str(mode)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    nodes, depth = map(int, r.stdout.split()[1:3])
    # the traversal stack bound used by the kernels: (width - 1) * depth + 1 entries
    assert depth >= 1 and nodes >= 1
