"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer (g++; no GPU code is built): the product's BVH builder
(xitils_amd/csrc/bvh.cpp, threaded for large inputs) through its structural
checker, and the CPU oracle plus the host walk through a driver that renders,
traces and fits on small scenes."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BVH = os.path.join(ROOT, "xitils_amd", "csrc", "bvh.cpp")
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", TSAN_OPTIONS="halt_on_error=1")


def _build(name, srcs, flags):
    out = os.path.join(tempfile.mkdtemp(), name)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", *flags, "-o", out, *srcs, "-pthread"],
                   check=True)
    return out


def _run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, env=SAN_ENV, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout[-2000:] + r.stderr[-4000:]


ASAN_UBSAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


@pytest.mark.slow
def test_bvh_builder_asan_ubsan():
    exe = _build("bvh_check_asan", [os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"), BVH], ASAN_UBSAN)
    for n, mode in [(0, 0), (1, 0), (5, 0), (3000, 0), (120000, 0), (4000, 1), (8000, 2), (8000, 3), (3000, 4),
                    (3000, 5)]:
        _run(exe, n, mode)


@pytest.mark.slow
def test_bvh_builder_tsan():
    # inputs above 65,536 triangles build subtrees on other threads
    exe = _build("bvh_check_tsan", [os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"), BVH], ["-fsanitize=thread"])
    _run(exe, 300000, 0)


@pytest.mark.slow
def test_oracle_and_host_walk_asan_ubsan():
    srcs = [os.path.join(ROOT, "tests", "cpp", "san_driver.cpp"), os.path.join(ROOT, "oracle", "xo_capi.cpp"),
            os.path.join(ROOT, "tests", "cpp", "trav_count.cpp"), BVH]
    exe = _build("san_driver", srcs, ASAN_UBSAN + ["-fopenmp", "-ffp-contract=off"])
    _run(exe)
