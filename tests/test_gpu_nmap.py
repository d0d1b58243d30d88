"""GPU parity of normal-mapped shading against the oracle at the same seed
and spp: the microfacet-based normal mapping BSDF (Sandbox/
MicrofacetBasedNormalMapping/Main.cpp:86-340) over each base, and the plain
normal map (Object.h:59-68) on the displaced plane under a lat-long sky.

Tolerance as the other render parity tests: relMSE <= 1e-4.  Both sides draw
the walk's uniforms from the same Philox streams, so pixels match bit for bit
except where a last-ulp difference of a transcendental (ocml vs glibc) flips
a walk step."""
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


def _plane(**kw):
    args = dict(grid=96, width=64, height=48, normal_map_size=256, sky_size=(256, 128))
    args.update(kw)
    return scenes.displaced_plane(**args)


def _compare(dev, desc, w, h, spp, **kw):
    print(f"[gpu] render {desc.name} {w}x{h} spp={spp} {kw}", flush=True)
    g = api.Scene(dev, desc)
    fg = g.render(w, h, spp, **kw)
    fo = po.OracleScene(desc).render(w, h, spp, **kw)
    assert np.isfinite(fg).all() and np.isfinite(fo).all()
    err = rel_mse(fg / spp, fo / spp)
    exact = np.mean(np.all(fg == fo, axis=2))
    print(f"[gpu]   relMSE={err:.3e} bit-exact pixels={exact:.4f} mean={fg.mean() / spp:.4f}", flush=True)
    assert err <= 1e-4 and exact >= 0.999   # same random numbers, operation order and transcendentals
    return fg, fo


@pytest.mark.parametrize("base", ["metal", "diffuse", "ggx_lambert"])
def test_microfacet_normal_mapping_matches_oracle(dev, base):
    _compare(dev, _plane(base=base, variant=3), 64, 48, 8, seed=5, max_depth=12, flags=scenes.FLAG_ENV_NEE)


def test_plain_normal_map_matches_oracle(dev):
    _compare(dev, _plane(base="diffuse", microfacet=False, variant=9), 64, 48, 8, seed=6, max_depth=12,
             flags=scenes.FLAG_ENV_NEE)


def test_microfacet_normal_mapping_sky_only_and_naive(dev):
    desc = _plane(variant=12)
    _compare(dev, desc, 64, 48, 8, seed=7, max_depth=12)
    _compare(dev, desc, 64, 48, 8, seed=8, max_depth=12, integrator=scenes.INTEGRATOR_NAIVE)


def test_guided_render_on_normal_mapped_plane(dev):
    """Guided sampling over the stochastic BSDF stays finite and converges to
    the plain estimate (one-sample MIS with the walk's getPDF)."""
    desc = _plane(grid=64, width=32, height=24, variant=5)
    g = api.Scene(dev, desc)
    kw = dict(max_depth=12, flags=scenes.FLAG_ENV_NEE)
    plain = g.render(32, 24, 512, seed=1, **kw) / 512
    g.guide_init()
    film = np.zeros((24, 32, 3), np.float32)
    for k in range(8):
        g.render(32, 24, 1, film=film, seed=2, spp_offset=k,
                 **dict(kw, flags=kw["flags"] | api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN))
    guided = g.render(32, 24, 512, seed=3, **dict(kw, flags=kw["flags"] | api.XT_FLAG_GUIDE)) / 512
    assert np.isfinite(guided).all()
    print(f"[gpu] guided mean {guided.mean():.5f} plain mean {plain.mean():.5f}", flush=True)
    assert guided.mean() == pytest.approx(plain.mean(), rel=0.03)
