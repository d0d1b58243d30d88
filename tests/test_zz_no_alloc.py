"""xt_render_static (the render entry point that allocates nothing: its
bounce loop runs in the library's static device scratch) against xt_render,
through the C ABI: two calls in a row, the second with a smaller input, and
a render that needs several batches of the scratch's capacity.  The counter-
based RNG makes every path's radiance independent of batching and each
pixel's samples are summed in sample order, so the films are equal bit for
bit."""
import numpy as np
import pytest
import torch

from xitils_amd import api, scenes

pytestmark = pytest.mark.gpu


def _static(sc, w, h, spp, **kw):
    film = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda:0")
    sc.render_static(film.data_ptr(), w, h, spp, **kw)
    torch.cuda.synchronize()
    return film.cpu().numpy()


def test_static_render_equals_xt_render():
    print("[gpu] static render: opening device 0", flush=True)
    dev = api.Device(0)
    big = api.Scene(dev, scenes.cornell(512, 512))
    small = api.Scene(dev, scenes.material_zoo(40, 30))
    # 512 x 512 x 4 = 1 Mi paths: several batches of the scratch's ~220 k
    for sc, w, h, spp, kw in [(big, 512, 512, 4, dict(seed=3, max_depth=8)),
                              (small, 40, 30, 3, dict(seed=4, max_depth=6, spp_offset=5))]:
        print(f"[gpu] static render {sc.desc.name} {w}x{h} spp={spp}", flush=True)
        a = _static(sc, w, h, spp, **kw)
        b = sc.render(w, h, spp, **kw)
        print(f"[gpu]   identical={np.array_equal(a, b)} mean={a.mean():.5f}", flush=True)
        assert np.isfinite(a).all() and a.max() > 0
        assert np.array_equal(a, b)


def test_static_render_rejections():
    dev = api.Device(0)
    sc = api.Scene(dev, scenes.cornell(16, 16))
    film = torch.zeros((16, 16, 3), dtype=torch.float32, device="cuda:0")
    with pytest.raises(api.XtError, match="bounce cap"):
        sc.render_static(film.data_ptr(), 16, 16, 1, max_depth=0)
    with pytest.raises(api.XtError, match="NULL"):
        sc.render_static(0, 16, 16, 1, max_depth=4)
    sc.guide_init(seed=1)
    with pytest.raises(api.XtError, match="unguided"):
        sc.render_static(film.data_ptr(), 16, 16, 1, max_depth=4, flags=api.XT_FLAG_GUIDE)
    assert float(film.abs().sum()) == 0.0
