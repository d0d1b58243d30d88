"""The path's transcendentals -- the oracle's (xo_core.hpp) and the HIP
library's (xt_libm.h, built for the host with g++) -- against float64 numpy:
sin and cos within 2 ulp on [-256, 256] (dense around the angles the path
uses, [-2 pi, 2 pi]), x^5 equal to the correctly rounded value, atan2
within 2.5 ulp and acos within 3 ulp (including the zeros' signs and the
ends of acos's domain); and the two implementations equal bit for bit, which is
what makes GPU and oracle sample the same directions."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KINDS = {"sin": 0, "cos": 1, "pow5": 2, "atan2": 3, "acos": 4}


@pytest.fixture(scope="module")
def product_libm():
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "libm_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "libm_check.cpp")], check=True)

    def run(kind, x):
        src, dst = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.ascontiguousarray(x, np.float32).tofile(src)
        subprocess.run([exe, str(KINDS[kind]), src, dst], check=True)
        return np.fromfile(dst, np.float32)
    return run


def ulp_err(got, want64):
    want32 = want64.astype(np.float32)
    ulp = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want64) / ulp


def _args():
    rng = np.random.default_rng(5)
    return np.concatenate([np.linspace(-2 * np.pi, 2 * np.pi, 400001), rng.uniform(-256, 256, 200000),
                           (rng.random(100000) * np.float32(2 * np.pi)),
                           np.array([0.0, -0.0, 1e-30, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi])]).astype(np.float32)


def test_sin_cos_within_2_ulp_and_identical(product_libm):
    x = _args()
    for kind, ref in (("sin", np.sin), ("cos", np.cos)):
        o = po.libm(kind, x)
        p = product_libm(kind, x)
        assert np.array_equal(o.view(np.uint32), p.view(np.uint32)), kind
        want = ref(x.astype(np.float64))
        # relative accuracy away from the function's zeros, absolute near them
        # (where the reduced argument's last rounding dominates an ulp of the result)
        err = ulp_err(o, want)
        big = np.abs(want) > 1e-3
        assert err[big].max() <= 2.0, (kind, err[big].max())
        assert np.abs(o.astype(np.float64) - want).max() < 2e-7
    s, c = po.libm("sin", x).astype(np.float64), po.libm("cos", x).astype(np.float64)
    assert np.abs(s * s + c * c - 1).max() < 5e-7


def test_pow5_correctly_rounded_and_identical(product_libm):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.random(300000), np.linspace(0, 1, 10001), [0.0, 1.0, 2.0, 0.5]]).astype(np.float32)
    o, p = po.libm("pow5", x), product_libm("pow5", x)
    assert np.array_equal(o, p)
    from fractions import Fraction
    sample = rng.choice(x.size, 2000, replace=False)
    for i in sample:
        exact = Fraction(float(x[i])) ** 5
        assert o[i] == np.float32(float(exact))   # float64 of the exact value, then to float32
    assert np.all(ulp_err(o, x.astype(np.float64) ** 5) <= 0.5000001)


def test_atan2_acos_pinned_and_identical(product_libm):
    rng = np.random.default_rng(3)
    th = np.linspace(-np.pi, np.pi, 400001)
    yx = np.concatenate([rng.normal(size=(200000, 2)),
                         np.stack([np.sin(th), np.cos(th)], 1),                 # unit directions, every octant
                         rng.normal(size=(50000, 2)) * 10.0 ** rng.integers(-6, 6, size=(50000, 1)),
                         [[0.0, 1.0], [-0.0, 1.0], [0.0, -1.0], [-0.0, -1.0], [1.0, 0.0], [-1.0, 0.0],
                          [0.0, 0.0], [-0.0, 0.0], [0.0, -0.0], [-0.0, -0.0], [1.0, 1.0], [-1.0, -1.0]]]).astype(np.float32)
    o, p = po.libm("atan2", yx), product_libm("atan2", yx)
    assert np.array_equal(o.view(np.uint32), p.view(np.uint32))
    want = np.arctan2(yx[:, 0].astype(np.float64), yx[:, 1].astype(np.float64))
    assert ulp_err(o, want).max() <= 2.5
    assert np.array_equal(np.signbit(o), np.signbit(want))                      # IEEE zero and quadrant signs
    c = np.concatenate([rng.uniform(-1, 1, 200000), np.linspace(-1, 1, 400001),
                        1.0 - np.float32(2.0 ** -24) * np.arange(200), [-1.0, 0.0, -0.0, 1.0]]).astype(np.float32)
    o, p = po.libm("acos", c), product_libm("acos", c)
    assert np.array_equal(o.view(np.uint32), p.view(np.uint32))
    assert ulp_err(o, np.arccos(c.astype(np.float64))).max() <= 3.0
    assert o[-1] == 0.0 and o[-4] == np.float32(np.pi)
