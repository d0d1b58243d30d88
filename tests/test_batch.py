"""Batched, graph-replayed renders (xt_render_batch) and the autotuner.

The counter-based RNG keys every random number by (pixel, sample, vertex),
and every path writes only its own radiance slot, so a render does not
depend on queue order, launch shapes or batching: a graph-replayed render
must give xt_render's film bit for bit, for every scene kind, for guided
sampling, for progressive offsets, across several batches, and after the
autotuner has changed the launch shapes."""
import numpy as np
import pytest

from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    print("[gpu] batch: opening device 0", flush=True)
    return api.Device(0)


def _jobs(dev):
    out = []
    d = scenes.cornell(64, 64)
    out.append((api.Scene(dev, d), 64, 64, 8, dict(seed=1, max_depth=8)))
    d = scenes.material_zoo(48, 48)
    out.append((api.Scene(dev, d), 48, 48, 8, dict(seed=2, max_depth=8)))
    d = scenes.sky_scene(scenes.ENV_MAP_LATLONG, 48, 36)
    out.append((api.Scene(dev, d), 48, 36, 8, dict(seed=3, max_depth=6, flags=scenes.FLAG_ENV_NEE)))
    d = scenes.heightfield(cells=64, width=96, height=54)
    out.append((api.Scene(dev, d), 96, 54, 4, dict(seed=4, max_depth=16, flags=scenes.FLAG_ENV_NEE)))
    d = scenes.cornell(40, 40)
    out.append((api.Scene(dev, d), 40, 40, 6, dict(seed=5, max_depth=5, integrator=1)))
    return out


def test_batch_matches_single_renders(dev):
    jobs = _jobs(dev)
    ref = []
    for sc, w, h, spp, kw in jobs:
        print(f"[gpu] xt_render {sc.desc.name} {w}x{h} spp={spp}", flush=True)
        ref.append(sc.render(w, h, spp, **kw))
    print("[gpu] xt_render_batch of all", flush=True)
    films, ms = api.render_batch([(sc, w, h, spp, None, kw) for sc, w, h, spp, kw in jobs])
    for (sc, *_), a, b, t in zip(jobs, ref, films, ms):
        print(f"[gpu]   {sc.desc.name}: {t:.2f} ms, identical={np.array_equal(a, b)}", flush=True)
        assert np.array_equal(a, b)
        assert t > 0


def test_batch_progressive_repeat_and_graph_reuse(dev):
    sc = api.Scene(dev, scenes.cornell(48, 48))
    kw = dict(seed=9, max_depth=8)
    ref = np.zeros((48, 48, 3), np.float32)
    sc.render(48, 48, 4, film=ref, spp_offset=0, **kw)
    sc.render(48, 48, 4, film=ref, spp_offset=4, **kw)
    sc.render(48, 48, 4, film=ref, spp_offset=8, **kw)
    film = np.zeros((48, 48, 3), np.float32)
    for off in (0, 4, 8):      # the same graph replayed with a moving sample offset
        api.render_batch([(sc, 48, 48, 4, film, dict(spp_offset=off, **kw))])
    assert np.array_equal(ref, film)


def test_batch_spans_several_batches(dev):
    # more paths than one graph replay: the item renders through the bounce loop
    # 1024 x 512 at 40 spp: 21 M paths, more than one batch of the path buffers
    sc = api.Scene(dev, scenes.cornell(1024, 512))
    kw = dict(seed=11, max_depth=4)
    print("[gpu] multi-batch render", flush=True)
    ref = sc.render(1024, 512, 40, **kw)
    films, _ = api.render_batch([(sc, 1024, 512, 40, None, kw)])
    assert np.array_equal(ref, films[0])


def test_batch_mixes_graph_and_loop_items_on_one_scene(dev):
    # a small item (one graph replay) and a large one (the bounce loop, which
    # grows the scene's path buffers after the graph was enqueued) on the same
    # scene, plus a small item on another scene: every film as xt_render's
    a = api.Scene(dev, scenes.cornell(64, 64))
    b = api.Scene(dev, scenes.material_zoo(32, 32))
    jobs = [(a, 64, 64, 8, dict(seed=31, max_depth=6)), (a, 1024, 1024, 6, dict(seed=32, max_depth=5)),
            (b, 32, 32, 16, dict(seed=33, max_depth=6))]
    print("[gpu] mixed graph / loop batch", flush=True)
    films, ms = api.render_batch([(sc, w, h, spp, None, kw) for sc, w, h, spp, kw in jobs])
    for (sc, w, h, spp, kw), f, t in zip(jobs, films, ms):
        assert np.array_equal(sc.render(w, h, spp, **kw), f), (w, h, spp)
        assert t > 0


def test_guided_graph_render_matches(dev):
    sc = api.Scene(dev, scenes.material_zoo(48, 48))
    sc.guide_init(seed=3)
    for k in range(3):
        sc.render(48, 48, 4, seed=20 + k, max_depth=8, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
    kw = dict(seed=21, max_depth=8, flags=api.XT_FLAG_GUIDE)
    ref = sc.render(48, 48, 8, **kw)
    films, _ = api.render_batch([(sc, 48, 48, 8, None, kw)])
    assert np.array_equal(ref, films[0])
    # a new network (new allocation) must not be read through the graph cached for the old one
    print("[gpu] guided batch after guide_init again", flush=True)
    sc.guide_init(seed=4)
    sc.render(48, 48, 4, seed=30, max_depth=8, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
    ref2 = sc.render(48, 48, 8, **kw)
    films2, _ = api.render_batch([(sc, 48, 48, 8, None, kw)])
    assert np.array_equal(ref2, films2[0])
    assert not np.array_equal(ref, ref2)


def test_autotune_changes_shapes_not_results(dev):
    sc = api.Scene(dev, scenes.cornell(128, 128))
    kw = dict(seed=13, max_depth=8)
    ref = sc.render(128, 128, 8, **kw)
    tb, sb = sc.autotune(128, 128, 8, **kw)
    print(f"[gpu] autotune: {tb} traversal blocks/CU, {sb} shade blocks/CU", flush=True)
    # candidates 2, 4, 8 traversal blocks per CU, the last clamped to what stays resident
    assert (tb in (2, 4) or 5 <= tb <= 8) and sb in (2, 4, 8)
    assert np.array_equal(ref, sc.render(128, 128, 8, **kw))
    films, _ = api.render_batch([(sc, 128, 128, 8, None, kw)])
    assert np.array_equal(ref, films[0])


def test_batch_rejects_uncapped_and_training_renders(dev):
    sc = api.Scene(dev, scenes.cornell(16, 16))
    with pytest.raises(api.XtError):
        api.render_batch([(sc, 16, 16, 1, None, dict(max_depth=0))])
    sc.guide_init()
    with pytest.raises(api.XtError):
        api.render_batch([(sc, 16, 16, 1, None, dict(max_depth=4, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN))])
