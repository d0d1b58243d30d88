"""The instrumented host walk of the product's BVH (tests/cpp/trav_count.cpp)
finds the brute-force closest and any hits, so its node and triangle counts
are those of a correct traversal; the GPU test (test_gpu_trace_counts.py)
then requires the kernels' fetch counters to equal them."""
import numpy as np
import pytest

import pyoracle as po
import travcount
from helpers import camera_rays, scene_rays, hit_prims
from xitils_amd import scenes


@pytest.fixture(scope="module", autouse=True)
def _built():
    travcount.build()
    po.build()


@pytest.mark.parametrize("which", ["cornell", "heightfield"])
def test_host_walk_finds_the_brute_force_hits(which):
    desc = scenes.cornell() if which == "cornell" else scenes.heightfield(cells=48, materials=False)
    rng = np.random.default_rng(3)
    rays = np.concatenate([scene_rays(rng, desc, 4000, 0.02, 0.98), camera_rays(desc, 40, 30)])
    hb = travcount.HostBVH(desc)
    hits, cnt, pr = hb.trace(rays, per_ray=True)
    ref = po.OracleScene(desc).trace_closest(rays, mode=1)
    assert np.array_equal(hit_prims(hits), hit_prims(ref))
    assert np.array_equal(hits[:, 0], ref[:, 0])
    assert pr[:, 0].sum() == cnt["nodes"] and pr[:, 1].sum() == cnt["tris"]
    assert np.all(pr[:, 0] >= 1)                       # the root is always visited
    assert cnt["tris"] >= (hit_prims(hits) != 0xFFFFFFFF).sum()
    occ, cnt_any, _ = hb.trace(rays[:2000], any_hit=True)
    assert np.array_equal(occ, po.OracleScene(desc).trace_any(rays[:2000], mode=1))
    # an any-hit walk stops at the first blocker: never more work than closest
    _, cnt_cl, _ = hb.trace(rays[:2000])
    assert cnt_any["nodes"] <= cnt_cl["nodes"]
