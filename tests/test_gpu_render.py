"""GPU render parity against the oracle at the same seed and spp.

Tolerance: relMSE <= 1e-4 (mean over pixels and channels of
(x - r)^2 / (r^2 + 0.01)), and at least 99.9% of the pixels bit-identical.
Both sides draw the same counter-based random numbers, share the operation
order and evaluate sin, cos, x^5, atan2 and acos the same deterministic way
(xt_libm.h / the oracle's restatement), so the films are equal bit for bit;
the 0.1% allowance covers an atan2 or acos whose double results fall on
either side of a float rounding boundary."""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import rel_mse
from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    print("[gpu] opening device 0", flush=True)
    return api.Device(0)


def _compare(dev, desc, w, h, spp, **kw):
    print(f"[gpu] render {desc.name} {w}x{h} spp={spp} {kw}", flush=True)
    g = api.Scene(dev, desc)
    fg = g.render(w, h, spp, **kw)
    o = po.OracleScene(desc)
    fo = o.render(w, h, spp, **kw)
    assert np.isfinite(fg).all()
    err = rel_mse(fg / spp, fo / spp)
    exact = np.mean(np.all(fg == fo, axis=2))
    print(f"[gpu]   relMSE={err:.3e} bit-exact pixels={exact:.4f}", flush=True)
    assert err <= 1e-4 and exact >= 0.999
    return fg, fo, exact, g


def test_cornell_standard(dev):
    _, _, exact, g = _compare(dev, scenes.cornell(64, 64), 64, 64, 8, seed=1, max_depth=8)
    st = g.stats()
    assert st["camera_rays"] == 64 * 64 * 8 and st["shadow_rays"] > 0


def test_cornell_unbounded_depth_roulette(dev):
    _compare(dev, scenes.cornell(48, 48), 48, 48, 4, seed=5, max_depth=0)


def test_cornell_naive(dev):
    _compare(dev, scenes.cornell(48, 48), 48, 48, 8, seed=2, max_depth=8, integrator=1)


def test_cornell_uniform_barycentrics(dev):
    _compare(dev, scenes.cornell(48, 48), 48, 48, 4, seed=3, max_depth=8, flags=scenes.FLAG_UNIFORM_BARY)


def test_material_zoo(dev):
    _compare(dev, scenes.material_zoo(64, 64), 64, 64, 8, seed=7, max_depth=8)


def test_heightfield_env_nee(dev):
    desc = scenes.heightfield(cells=64, width=96, height=54)
    _compare(dev, desc, 96, 54, 4, seed=11, max_depth=16, flags=scenes.FLAG_ENV_NEE)


def test_heightfield_sky_only_reference_mode(dev):
    desc = scenes.heightfield(cells=64, width=96, height=54)
    _compare(dev, desc, 96, 54, 4, seed=12, max_depth=16)


def test_progressive_accumulation_and_f64_film(dev):
    desc = scenes.cornell(32, 32)
    g = api.Scene(dev, desc)
    a = g.render(32, 32, 6, seed=4, max_depth=8)
    b = np.zeros((32, 32, 3), np.float32)
    for k in range(3):
        g.render(32, 32, 2, film=b, seed=4, max_depth=8, spp_offset=2 * k)
    # the same samples in a different batching give the same film up to
    # float summation order
    assert np.allclose(a, b, rtol=1e-5, atol=1e-5)
    c = g.render(32, 32, 6, f64=True, seed=4, max_depth=8)
    assert c.dtype == np.float64 and np.allclose(c, a, rtol=1e-5, atol=1e-5)


def test_furnace_exact_on_gpu(dev):
    a = 0.6
    desc = scenes.furnace_object(a)
    g = api.Scene(dev, desc)
    for integ in (0, 1):
        v = g.render(32, 32, 1, seed=3, integrator=integ)
        is_a = np.all(np.abs(v - a) <= 2e-6 * a, axis=2)
        is_1 = np.all(v == 1.0, axis=2)
        assert np.all(is_a | is_1) and is_a.mean() > 0.1


def test_render_longer_than_one_batch(dev):
    # 1024^2 x 80 spp = 84 M paths: more than one batch of the default path
    # capacity (64 Mi), so the render runs two batches; the same samples in
    # five 16-spp calls (one batch each) give the same film up to the float
    # summation order of the per-pixel sums
    desc = scenes.cornell(1024, 1024)
    g = api.Scene(dev, desc)
    print("[gpu] 1024^2 x 80 spp in one call", flush=True)
    a = g.render(1024, 1024, 80, seed=6, max_depth=4)
    b = np.zeros_like(a)
    for k in range(5):
        g.render(1024, 1024, 16, film=b, seed=6, max_depth=4, spp_offset=16 * k)
    assert np.isfinite(a).all()
    assert np.allclose(a, b, rtol=1e-5, atol=1e-4)


def test_many_batches_large_image(dev):
    # more paths than one batch holds: the batching over samples/pixels must
    # not change the result
    desc = scenes.cornell(2048, 2048)
    g = api.Scene(dev, desc)
    f = g.render(2048, 2048, 2, seed=1, max_depth=4)
    assert np.isfinite(f).all() and f.min() >= 0
    o = po.OracleScene(desc)
    sub = o.render(2048, 2048, 2, seed=1, max_depth=4)
    assert rel_mse(f / 2, sub / 2) <= 1e-4
    assert np.mean(np.all(f == sub, axis=2)) >= 0.999


@pytest.mark.gpu
def test_large_render_with_chunked_traversal_matches_smaller_batches():
    """A render whose path queue holds at least 8 Mi paths traces its rays in
    chunks handed out from a cursor (extension and shadow kernels); the same
    samples rendered in two calls below that size trace in even per-wave
    shares.  Every path's radiance is the same either way, so the FP64 films
    agree to the rounding of summing a pixel's samples in two parts."""
    desc = scenes.heightfield(cells=256)
    sc = api.Scene(api.Device(0), desc)
    W = H = 1024
    kw = dict(seed=5, max_depth=8, flags=api.XT_FLAG_ENV_NEE)
    print("[gpu] chunked render: 1024^2 x 8 spp vs 2 x 4 spp", flush=True)
    big = sc.render(W, H, 8, f64=True, **kw)                  # 8,388,608 paths in one batch
    parts = np.zeros((H, W, 3), np.float64)
    sc.render(W, H, 4, film=parts, spp_offset=0, **kw)
    sc.render(W, H, 4, film=parts, spp_offset=4, **kw)
    assert np.isfinite(big).all() and big.mean() > 0
    np.testing.assert_allclose(big, parts, rtol=1e-9, atol=1e-9)
