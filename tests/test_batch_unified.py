"""Batches through the wavefront: xt_render_batch renders its small unguided
items (by default; every unguided item under XITILS_AMD_BATCH=unified) in
one wavefront -- the items' paths numbered one after another and walked in
windows of the path capacity, each window one replay of a captured graph,
each path traced and shaded in its own item's scene -- and the rest per item.
Every film must equal xt_render's bit for bit -- across windows, scenes,
depths, sample offsets, film types and integrators, in every form -- and
items sharing a film must accumulate in item order."""
import os

import numpy as np
import pytest
import torch

from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    print("[gpu] unified batch: opening device 0", flush=True)
    return api.Device(0)


@pytest.fixture
def batch_mode(request):
    mode = request.param
    if mode:
        os.environ["XITILS_AMD_BATCH"] = mode
    yield mode
    os.environ.pop("XITILS_AMD_BATCH", None)


@pytest.mark.parametrize("batch_mode", ["unified", None], indirect=True, ids=["unified", "default"])
def test_unified_batch_equals_single_renders_across_windows(dev, batch_mode):
    # 2 x 1024^2 x 6 spp + the rest: more paths than one window (16 Mi)
    cor = api.Scene(dev, scenes.cornell(64, 64))
    zoo = api.Scene(dev, scenes.material_zoo(48, 48))
    sky = api.Scene(dev, scenes.sky_scene(scenes.ENV_MAP_LATLONG, 48, 36))
    hf = api.Scene(dev, scenes.heightfield(cells=64, width=96, height=54))
    nm = api.Scene(dev, scenes.displaced_plane(grid=96, width=64, height=48, normal_map_size=256,
                                               sky_size=(256, 128), variant=3))
    room_desc, *_ = scenes.batch_scene(5)
    room = api.Scene(dev, room_desc)
    jobs = [(cor, 1024, 1024, 6, dict(seed=1, max_depth=5)),
            (zoo, 48, 48, 8, dict(seed=2, max_depth=8)),
            (sky, 48, 36, 8, dict(seed=3, max_depth=6, flags=scenes.FLAG_ENV_NEE)),
            (hf, 96, 54, 4, dict(seed=4, max_depth=16, flags=scenes.FLAG_ENV_NEE, spp_offset=7)),
            (cor, 40, 40, 6, dict(seed=5, max_depth=5, integrator=1)),
            (room, 1024, 1024, 6, dict(seed=6, max_depth=29)),
            (nm, 64, 48, 4, dict(seed=7, max_depth=12, flags=scenes.FLAG_ENV_NEE))]
    ref = []
    for sc, w, h, spp, kw in jobs:
        print(f"[gpu] xt_render {sc.desc.name} {w}x{h} spp={spp}", flush=True)
        ref.append(sc.render(w, h, spp, **kw))
    f64 = np.zeros((48, 48, 3), np.float64)
    ref64 = zoo.render(48, 48, 8, f64=True, seed=2, max_depth=8)
    dfilm = torch.zeros((36, 48, 3), dtype=torch.float32, device="cuda:0")
    print(f"[gpu] xt_render_batch of all ({batch_mode or 'default'})", flush=True)
    films, ms = api.render_batch([(sc, w, h, spp, None, kw) for sc, w, h, spp, kw in jobs] +
                                 [(zoo, 48, 48, 8, f64, dict(seed=2, max_depth=8)),
                                  (sky, 48, 36, 8, (dfilm.data_ptr(), False), dict(seed=3, max_depth=6,
                                                                                   flags=scenes.FLAG_ENV_NEE))])
    torch.cuda.synchronize()
    for (sc, w, h, spp, kw), a, b, t in zip(jobs, ref, films, ms):
        print(f"[gpu]   {sc.desc.name} {w}x{h}: {t:.2f} ms, identical={np.array_equal(a, b)}", flush=True)
        assert np.array_equal(a, b)
        assert t > 0
    assert np.array_equal(f64, ref64)
    assert np.array_equal(dfilm.cpu().numpy(), ref[2])
    if batch_mode == "unified":
        # latency is the end of the window holding an item's last path: monotone in item order
        assert all(ms[k] <= ms[k + 1] + 1e-3 for k in range(len(jobs) - 1))


def test_unified_items_sharing_a_film_accumulate_in_order(dev):
    a = api.Scene(dev, scenes.cornell(48, 48))
    b = api.Scene(dev, scenes.material_zoo(48, 48))
    ref = np.zeros((48, 48, 3), np.float32)
    a.render(48, 48, 4, film=ref, seed=8, max_depth=6)
    b.render(48, 48, 4, film=ref, seed=9, max_depth=6)
    a.render(48, 48, 4, film=ref, seed=8, max_depth=6, spp_offset=4)
    film = np.zeros((48, 48, 3), np.float32)
    api.render_batch([(a, 48, 48, 4, film, dict(seed=8, max_depth=6)), (b, 48, 48, 4, film, dict(seed=9, max_depth=6)),
                      (a, 48, 48, 4, film, dict(seed=8, max_depth=6, spp_offset=4))])
    assert np.array_equal(ref, film)


def test_shared_film_with_a_guided_item_renders_in_item_order(dev):
    # a guided item cannot join the wavefront: every item through xt_render in turn
    a = api.Scene(dev, scenes.cornell(32, 32))
    b = api.Scene(dev, scenes.material_zoo(32, 32))
    b.guide_init(seed=3)
    for k in range(2):
        b.render(32, 32, 4, seed=40 + k, max_depth=6, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
    kb = dict(seed=41, max_depth=6, flags=api.XT_FLAG_GUIDE)
    ref = np.zeros((32, 32, 3), np.float32)
    a.render(32, 32, 4, film=ref, seed=8, max_depth=6)
    b.render(32, 32, 4, film=ref, **kb)
    a.render(32, 32, 2, film=ref, seed=8, max_depth=6, spp_offset=4)
    film = np.zeros((32, 32, 3), np.float32)
    api.render_batch([(a, 32, 32, 4, film, dict(seed=8, max_depth=6)), (b, 32, 32, 4, film, kb),
                      (a, 32, 32, 2, film, dict(seed=8, max_depth=6, spp_offset=4))])
    assert np.array_equal(ref, film)


def test_large_guided_items_on_two_scenes_render_concurrently(dev):
    # guided items of more than one graph replay on different scenes: the
    # lagged loop, from two host threads at once; films as xt_render's
    a = api.Scene(dev, scenes.cornell(32, 32))
    b = api.Scene(dev, scenes.material_zoo(32, 32))
    for sc, seed in ((a, 3), (b, 4)):
        sc.guide_init(seed=seed)
        sc.render(64, 64, 4, seed=seed, max_depth=6, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
    kw = dict(max_depth=6, flags=api.XT_FLAG_GUIDE)
    jobs = [(a, 1024, 1024, 5, dict(seed=21, **kw)), (b, 1024, 1024, 5, dict(seed=22, **kw)),
            (a, 512, 512, 20, dict(seed=23, **kw))]
    ref = [sc.render(w, h, spp, **k) for sc, w, h, spp, k in jobs]
    print("[gpu] guided batch of 3 large items on 2 scenes", flush=True)
    films, ms = api.render_batch([(sc, w, h, spp, None, k) for sc, w, h, spp, k in jobs])
    for r, f, t in zip(ref, films, ms):
        assert np.array_equal(r, f)
        assert t > 0


def test_per_item_graphs_stay_available(dev):
    sc = api.Scene(dev, scenes.cornell(48, 48))
    kw = dict(seed=12, max_depth=6)
    ref = sc.render(48, 48, 8, **kw)
    os.environ["XITILS_AMD_BATCH"] = "items"
    try:
        films, _ = api.render_batch([(sc, 48, 48, 8, None, kw)])
    finally:
        del os.environ["XITILS_AMD_BATCH"]
    assert np.array_equal(ref, films[0])
