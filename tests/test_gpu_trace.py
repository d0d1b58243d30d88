"""GPU parity of the ray queries (xt_trace_closest / xt_trace_any) against the
oracle: closest-hit prim ids bit-exact, t / barycentrics bit-exact (same
arithmetic, -ffp-contract=off on both sides), occlusion bits exact."""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import rays_from, random_dirs, unit, camera_rays, scene_rays, hit_prims
from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    print("[gpu] opening device 0", flush=True)
    return api.Device(0)


def _check_closest(dev, desc, rays, exact_t=True):
    print(f"[gpu] trace_closest {desc.name} n={len(rays)}", flush=True)
    g = api.Scene(dev, desc)
    o = po.OracleScene(desc)
    hg = g.trace_closest(rays)
    hb = o.trace_closest(rays, mode=1)          # brute force
    hr = o.trace_closest(rays, mode=0)          # reference-shaped BVH
    pg, pb, pr = hit_prims(hg), hit_prims(hb), hit_prims(hr)
    assert np.array_equal(pg, pb)
    assert np.array_equal(pg, pr)
    hit = pg != 0xFFFFFFFF
    if exact_t:
        assert np.array_equal(hg[hit, 0], hb[hit, 0])
        assert np.array_equal(hg[hit, 2:4], hb[hit, 2:4])
    else:
        assert np.all(np.abs(hg[hit, 0] - hb[hit, 0]) <= 1e-5 * np.abs(hb[hit, 0]))
    assert np.all(np.isinf(hg[~hit, 0]))
    return g, o, hg


def test_cornell_closest_bit_exact(dev):
    desc = scenes.cornell()
    rng = np.random.default_rng(7)
    rays = np.concatenate([camera_rays(desc, 64, 64), scene_rays(rng, desc, 20000, 0.01, 0.99)])
    _check_closest(dev, desc, rays)


def test_heightfield_closest_bit_exact(dev):
    desc = scenes.heightfield(cells=96, materials=False)
    rng = np.random.default_rng(8)
    cam = camera_rays(desc, 96, 54)
    # diffuse-bounce-like rays from points just above the surface
    pts = rng.uniform(0.05, 0.95, size=(20000, 3)).astype(np.float32)
    pts[:, 1] = rng.uniform(-0.05, 0.12, size=20000)
    d = random_dirs(rng, 20000)
    _check_closest(dev, desc, np.concatenate([cam, rays_from(pts, d)]))


def test_zoo_closest_with_normals_and_tmax(dev):
    desc = scenes.material_zoo(32, 32)
    rng = np.random.default_rng(9)
    rays = scene_rays(rng, desc, 20000, 0.02, 0.98)
    rays[::3, 3] = rng.uniform(0.01, 0.5, size=rays[::3].shape[0]).astype(np.float32)   # finite tmax
    _check_closest(dev, desc, rays)


def test_any_hit_exact(dev):
    for desc in (scenes.cornell(), scenes.heightfield(cells=64, materials=False)):
        print(f"[gpu] trace_any {desc.name}", flush=True)
        rng = np.random.default_rng(10)
        rays = scene_rays(rng, desc, 20000, 0.02, 0.98)
        rays[:, 3] = rng.uniform(0.0, 1.0, size=20000).astype(np.float32)
        rays[::4, 3] = np.inf
        g = api.Scene(dev, desc)
        o = po.OracleScene(desc)
        assert np.array_equal(g.trace_any(rays), o.trace_any(rays, mode=1))


def test_shared_vertex_ties_follow_min_prim_rule(dev):
    # axis-aligned grid; rays aimed exactly at shared vertices and edges, where
    # several triangles give the same t: the smaller global prim id must win
    desc = scenes.heightfield(cells=8, amplitude=0.0, materials=False)
    xs = np.linspace(0, 1, 9, dtype=np.float32)
    targets = np.array([[x, 0.0, z] for x in xs[1:-1] for z in xs[1:-1]], np.float32)
    targets = np.concatenate([targets, targets + np.array([0.0625, 0, 0], np.float32)])
    o = np.array([0.5, 1.0, 0.5], np.float32)
    d = unit(targets - o)
    rays = rays_from(np.broadcast_to(o, d.shape), d)
    _check_closest(dev, desc, rays)


def test_empty_and_degenerate_inputs(dev):
    desc = scenes.cornell()
    g = api.Scene(dev, desc)
    assert g.trace_closest(np.zeros((0, 8), np.float32)).shape == (0, 4)
    # zero direction, NaN origin: no hit, no fault
    bad = rays_from([[0.5, 0.5, 0.5], [np.nan, 0.5, 0.5]], [[0, 0, 0], [0, 1, 0]])
    h = g.trace_closest(bad)
    assert np.all(hit_prims(h) == 0xFFFFFFFF)
    # a scene with an empty object list
    empty = scenes.SceneDesc(materials=[scenes.diffuse(0.5)], camera=desc.camera)
    ge = api.Scene(dev, empty)
    assert np.all(hit_prims(ge.trace_closest(camera_rays(desc, 4, 4))) == 0xFFFFFFFF)


def test_unsupported_material_is_rejected(dev):
    desc = scenes.cornell()
    desc.materials[0] = scenes.Material(99, (0.5,))
    with pytest.raises(api.XtError) as e:
        api.Scene(dev, desc)
    assert e.value.status == 2


def test_non_finite_vertices_are_rejected(dev):
    """Positions the BVH's quantisation grid cannot hold (not finite, or
    beyond 1e30) fail scene creation with XT_ERR_INVALID_ARGUMENT."""
    for bad in (np.nan, np.inf, 3e30):
        desc = scenes.cornell()
        obj = desc.objects[0]
        pos = obj.positions.copy()
        pos[obj.indices[0, 0], 1] = bad
        obj.positions = pos
        print(f"[gpu] scene with a vertex coordinate {bad}", flush=True)
        with pytest.raises(api.XtError) as e:
            api.Scene(dev, desc)
        assert e.value.status == 1 and "finite" in str(e.value)


def test_large_heightfield_properties(dev):
    # full 1024^2-cell heightfield (2.1 M triangles): closest hits agree with
    # brute force on a sample, and any-hit with tmax = t_hit * (1 + 2e-6) is
    # always occluded (size-independent invariant)
    desc = scenes.heightfield(cells=1024, materials=False)
    print("[gpu] building 2.1M-triangle scene", flush=True)
    g = api.Scene(dev, desc)
    info = g.info()
    assert info["num_prims"] == 2 * 1024 * 1024
    rays = camera_rays(desc, 480, 270)
    h = g.trace_closest(rays)
    hit = hit_prims(h) != 0xFFFFFFFF
    assert hit.mean() > 0.5
    r2 = rays[hit].copy()
    r2[:, 3] = h[hit, 0] * np.float32(1 + 2e-6)
    assert g.trace_any(r2).all()
    r2[:, 3] = h[hit, 0] * np.float32(0.999)
    assert not g.trace_any(r2).any()
    # brute force on 64 rays (2.1 M triangles each)
    o = po.OracleScene(desc)
    sel = np.nonzero(hit)[0][::2000][:64]
    hb = o.trace_closest(rays[sel], mode=1)
    assert np.array_equal(hit_prims(h[sel]), hit_prims(hb))
    assert np.array_equal(h[sel, 0], hb[:, 0])


def test_ray_batches_at_full_size_match_oracle(dev):
    """The microbenchmark's three batches on the full 1024^2-cell height field
    (2,097,152 triangles): GPU hit ids and occlusion bits equal the oracle's
    reference-shaped BVH on every 103rd ray of each (20,133 rays); the query reports
    its ray count, kernel time and fetch counts."""
    desc = scenes.heightfield(cells=1024, materials=False)
    g = api.Scene(dev, desc)
    o = po.OracleScene(desc)
    for kind in scenes.RAY_BATCH_KINDS:
        full = scenes.ray_batch(desc, kind)
        rays = np.ascontiguousarray(full[:: len(full) // 20000])   # spread over the frame / surface
        print(f"[gpu] ray batch {kind} n={len(rays)}", flush=True)
        g.set_counting(True)
        if kind == "shadow":
            got = g.trace_any(rays)
            st = g.stats()
            assert np.array_equal(got, o.trace_any(rays, mode=0))
            assert st["shadow_rays"] == len(rays) and st["ms_shadow"] > 0 and st["shadow_node_fetches"] >= len(rays)
        else:
            got = g.trace_closest(rays)
            st = g.stats()
            assert np.array_equal(hit_prims(got), hit_prims(o.trace_closest(rays, mode=0)))
            assert st["extension_rays"] == len(rays) and st["ms_extend"] > 0 and st["ext_node_fetches"] >= len(rays)
        g.set_counting(False)


def test_large_queries_take_rays_in_chunks(dev):
    """A query of at least 8 Mi rays hands its rays out in chunks from a cursor
    (a wave whose rays end early takes more), smaller ones in an even share
    per wave: each ray's hit and occlusion bit are the same either way, bit
    for bit, and the same large query twice in a row (the cursor resets
    itself at the end of a launch) answers the same."""
    desc = scenes.heightfield(cells=256, materials=False)
    g = api.Scene(dev, desc)
    rays = camera_rays(desc, 3072, 2816)          # 8,650,752 rays, distinct per pixel
    assert len(rays) >= 8 << 20
    print(f"[gpu] chunked query: {len(rays)} rays", flush=True)
    half = len(rays) // 2                          # each half below the chunked size
    bits = lambda h: h.view(np.uint32)             # misses carry prim bits 0xFFFFFFFF (a NaN as float)
    h_big = g.trace_closest(rays)
    h_parts = np.concatenate([g.trace_closest(rays[:half]), g.trace_closest(rays[half:])])
    assert np.array_equal(bits(h_big), bits(h_parts))
    assert np.array_equal(bits(g.trace_closest(rays)), bits(h_big))
    hit = hit_prims(h_big) != 0xFFFFFFFF
    assert 0.3 < hit.mean() < 1.0
    r2 = rays.copy()
    r2[:, 3] = np.where(hit, h_big[:, 0] * np.float32(0.5), np.float32(np.inf))
    o_big = g.trace_any(r2)
    o_parts = np.concatenate([g.trace_any(r2[:half]), g.trace_any(r2[half:])])
    assert np.array_equal(o_big, o_parts)
    assert np.array_equal(g.trace_any(r2), o_big)
