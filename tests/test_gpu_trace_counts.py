"""The traversal kernels' fetch counters (the counting variant behind the
roofline's algorithmic bytes) against the instrumented host walk of the same
BVH (tests/cpp/trav_count.cpp) on the same rays: node visits and triangle
tests must be equal, as must the hits, on samples of the ray-query
benchmark's three batches over the 2,097,152-triangle height field."""
import numpy as np
import pytest

import travcount
from helpers import hit_prims
from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field():
    print("[gpu] counts: building the 1024^2 height field", flush=True)
    desc = scenes.heightfield(cells=1024)
    g = api.Scene(api.Device(0), desc)
    g.set_counting(True)
    hb = travcount.HostBVH(desc)
    info = g.info()
    assert (hb.nodes, hb.depth) == (info["bvh_nodes"], info["bvh_depth"])
    return desc, g, hb


@pytest.mark.parametrize("kind", ["primary", "diffuse", "shadow"])
def test_fetch_counts_equal_the_host_walk(field, kind):
    desc, g, hb = field
    rays = np.ascontiguousarray(scenes.ray_batch(desc, kind)[::32])
    print(f"[gpu] counts: {kind}, {rays.shape[0]} rays", flush=True)
    if kind == "shadow":
        occ = g.trace_any(rays)
        st = g.stats()
        gn, gt = st["shadow_node_fetches"], st["shadow_tri_fetches"]
        hocc, cnt, _ = hb.trace(rays, any_hit=True)
        assert np.array_equal(occ, hocc)
    else:
        hits = g.trace_closest(rays)
        st = g.stats()
        gn, gt = st["ext_node_fetches"], st["ext_tri_fetches"]
        hh, cnt, _ = hb.trace(rays)
        assert np.array_equal(hit_prims(hits), hit_prims(hh))
        assert np.array_equal(hits[:, 0], hh[:, 0])
    n = rays.shape[0]
    print(f"[gpu]   gpu nodes {gn} tris {gt}; host nodes {cnt['nodes']} tris {cnt['tris']}; "
          f"per ray {cnt['nodes'] / n:.3f} nodes {cnt['tris'] / n:.3f} tris", flush=True)
    assert (gn, gt) == (cnt["nodes"], cnt["tris"])
