"""Image skies: SkySphereFromImage's angular-map lookup (SkySphere.h:14-34)
with Texture::rgb's bilinear filter (Texture.h:83-103), and the lat-long map
with its 2D importance distribution (new).

Oracle pins (CPU): bilinear filtering is exact on linear ramps, so the
looked-up value recovers the texel coordinate (u W - 0.5, v H - 0.5) that the
mapping formulas give in float64; the angular map's radius is angle / pi and
its exact centre is black (sq == 0); the lat-long pdf integrates to one and
its samples follow it (chi^2 over texel blocks); a Lambert plane under a
constant lat-long sky of 1 renders to its albedo with the map sampled as a
light.  GPU: renders match the oracle (relMSE <= 1e-4, as the other render
parity tests), and map sampling agrees with BSDF-only rendering in the
mean."""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import rel_mse
from xitils_amd import api, scenes


def ramp(W, H):
    m = np.zeros((H, W, 3), np.float32)
    m[..., 0] = np.arange(W)[None, :]
    m[..., 1] = np.arange(H)[:, None]
    m[..., 2] = 1.0
    return m


def sky_only(kind, m):
    s = scenes.SceneDesc(name="sky_only")
    s.add_material(scenes.diffuse(0.5))
    s.env = scenes.Env(True, (0.0, 0.0, 0.0), map_kind=kind, map=m)
    s.camera = scenes.Camera(scenes.look_at((0, 0, 0), (0, 0, 1)), math.radians(60), 1.0)
    return s


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_angular_lookup_recovers_coordinates():
    W = H = 64
    o = po.OracleScene(sky_only(scenes.ENV_MAP_ANGULAR, ramp(W, H)))
    rng = np.random.default_rng(1)
    wi = unit(rng.normal(size=(4000, 3)))
    d = -wi
    sq = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    r = np.arccos(d[:, 2]) / np.pi / sq
    u = (d[:, 0] * r + 1) / 2
    v = (d[:, 1] * r + 1) / 2
    L = o.env_radiance(wi.astype(np.float32)).astype(np.float64)
    inner = (u * W > 1) & (u * W < W - 1) & (v * H > 1) & (v * H < H - 1)
    assert inner.sum() > 3000
    np.testing.assert_allclose(L[inner, 0], u[inner] * W - 0.5, atol=2e-3)
    np.testing.assert_allclose(L[inner, 1], v[inner] * H - 0.5, atol=2e-3)
    # radius in the map = angle from the probe axis / pi
    ang = np.arccos(np.clip(d[:, 2], -1, 1))
    np.testing.assert_allclose(np.hypot(2 * u - 1, 2 * v - 1), ang / np.pi, atol=1e-5)
    # the exact centre direction is black (sq == 0, SkySphere.h:23-25)
    assert np.all(o.env_radiance(np.array([[0, 0, -1.0]], np.float32)) == 0)


def test_latlong_lookup_recovers_coordinates():
    W, H = 128, 64
    o = po.OracleScene(sky_only(scenes.ENV_MAP_LATLONG, ramp(W, H)))
    rng = np.random.default_rng(2)
    d = unit(rng.normal(size=(4000, 3)))
    u = (np.arctan2(d[:, 2], d[:, 0]) + np.pi) / (2 * np.pi)
    v = np.arccos(d[:, 1]) / np.pi
    L = o.env_radiance(d.astype(np.float32)).astype(np.float64)
    inner = (u * W > 1) & (u * W < W - 1) & (v * H > 1) & (v * H < H - 1)
    np.testing.assert_allclose(L[inner, 0], u[inner] * W - 0.5, atol=3e-3)
    np.testing.assert_allclose(L[inner, 1], v[inner] * H - 0.5, atol=3e-3)
    # u wraps: just past the seam the lookup blends the last and first columns
    seam = (u * W < 0.5) & inner.__invert__() & (v * H > 1) & (v * H < H - 1)
    assert np.all(L[seam, 0] >= 0) and np.all(L[seam, 0] <= W - 1)


def test_latlong_pdf_normalised_and_sampled():
    W, H = 64, 32
    m = scenes.procedural_sky_latlong(W, H, sun_half_angle_deg=6.0, sun_radiance=50.0)
    o = po.OracleScene(sky_only(scenes.ENV_MAP_LATLONG, m))
    # integral over the sphere by a fine (cos theta, phi) midpoint grid
    n = 1200
    c = -1 + (np.arange(n) + 0.5) * 2 / n
    ph = (np.arange(2 * n) + 0.5) * np.pi / n
    C, P = np.meshgrid(c, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    d = np.stack([S * np.cos(P), C, S * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    integral = o.env_map_pdf(d).astype(np.float64).sum() * (2 / n) * (np.pi / n)
    assert abs(integral - 1) < 5e-3, integral
    # samples: their pdf matches the evaluated pdf, and texel frequencies match
    rng = np.random.default_rng(3)
    N = 200000
    ds, pdf = o.env_map_sample(rng.random((N, 2)).astype(np.float32))
    np.testing.assert_allclose(np.linalg.norm(ds, axis=1), 1, atol=1e-5)
    pe = o.env_map_pdf(ds)
    ok = pe > 0
    assert ok.mean() > 0.999
    # equal except where rounding puts a sample on the neighbouring texel
    agree = np.abs(pdf[ok] / pe[ok] - 1) < 2e-3
    assert agree.mean() > 0.999
    u = (np.arctan2(ds[:, 2], ds[:, 0]) + np.pi) / (2 * np.pi)
    v = np.arccos(np.clip(ds[:, 1], -1, 1)) / np.pi
    bx = np.minimum((u * W).astype(int), W - 1) // 8
    by = np.minimum((v * H).astype(int), H - 1) // 8
    counts = np.bincount(by * (W // 8) + bx, minlength=(W // 8) * (H // 8))
    # expected block probabilities: integral of the pdf over each block's texels
    cu = (np.arange(W) + 0.5) / W * 2 * np.pi - np.pi
    cv = (np.arange(H) + 0.5) / H * np.pi
    Pp, Tt = np.meshgrid(cu, cv)
    dc = np.stack([np.sin(Tt) * np.cos(Pp), np.cos(Tt), np.sin(Tt) * np.sin(Pp)], -1).reshape(-1, 3)
    ptex = o.env_map_pdf(dc.astype(np.float32)).reshape(H, W) * np.sin(Tt) * (2 * np.pi / W) * (np.pi / H)
    pblk = ptex.reshape(H // 8, 8, W // 8, 8).sum(axis=(1, 3)).ravel()
    assert abs(pblk.sum() - 1) < 1e-3
    e = pblk * N
    keep = e > 20
    chi2 = ((counts[keep] - e[keep]) ** 2 / e[keep]).sum()
    assert chi2 < 2.0 * keep.sum() + 30, (chi2, keep.sum())


def plane_under_sky(albedo=0.6):
    s = scenes.SceneDesc(name="plane_sky")
    g = s.add_material(scenes.diffuse(albedo))
    qv, qi = scenes.quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    s.objects.append(scenes.ObjectDesc(qv, qi, material=g))
    s.env = scenes.Env(True, (0.0, 0.0, 0.0), map_kind=scenes.ENV_MAP_LATLONG,
                       map=np.ones((32, 64, 3), np.float32))
    s.camera = scenes.Camera(scenes.look_at((0, 2, 0), (0.01, 0, 0.0), up=(0, 0, 1)), math.radians(40), 1.0)
    return s


def test_plane_under_constant_map_renders_albedo():
    a = 0.6
    o = po.OracleScene(plane_under_sky(a))
    for flags in (0, scenes.FLAG_ENV_NEE):
        f = o.render(16, 16, 256, seed=4, max_depth=1, flags=flags) / 256
        assert abs(f.mean() - a) < 0.01 * a, (flags, f.mean())


# ---------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def dev():
    print("[gpu] env maps: opening device 0", flush=True)
    return api.Device(0)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [scenes.ENV_MAP_ANGULAR, scenes.ENV_MAP_LATLONG])
@pytest.mark.parametrize("flags", [0, scenes.FLAG_ENV_NEE])
def test_sky_renders_match_oracle(dev, kind, flags):
    desc = scenes.sky_scene(kind, 64, 48)
    print(f"[gpu] sky render kind={kind} flags={flags}", flush=True)
    g = api.Scene(dev, desc)
    fg = g.render(64, 48, 8, seed=3, max_depth=6, flags=flags)
    fo = po.OracleScene(desc).render(64, 48, 8, seed=3, max_depth=6, flags=flags)
    err = rel_mse(fg / 8, fo / 8)
    exact = np.mean(np.all(fg == fo, axis=2))
    print(f"[gpu]   relMSE {err:.2e}, bit-exact pixels {exact:.3f}", flush=True)
    assert np.isfinite(fg).all() and err <= 1e-4 and exact >= 0.999


@pytest.mark.gpu
def test_map_sampling_agrees_with_bsdf_sampling(dev):
    desc = scenes.sky_scene(scenes.ENV_MAP_LATLONG, 48, 36)
    g = api.Scene(dev, desc)
    spp = 8192
    a = g.render(48, 36, spp, seed=5, max_depth=6, flags=0) / spp
    b = g.render(48, 36, spp, seed=6, max_depth=6, flags=scenes.FLAG_ENV_NEE) / spp
    print(f"[gpu] means bsdf-only {a.mean():.5f} map-sampled {b.mean():.5f}; relMSE {rel_mse(b, a):.2e}", flush=True)
    assert abs(a.mean() - b.mean()) < 0.01 * a.mean()
    assert rel_mse(b, a) < 0.02


@pytest.mark.gpu
def test_plane_under_constant_map_on_gpu(dev):
    a = 0.6
    g = api.Scene(dev, plane_under_sky(a))
    for flags in (0, scenes.FLAG_ENV_NEE):
        f = g.render(32, 32, 1024, seed=4, max_depth=1, flags=flags) / 1024
        assert abs(f.mean() - a) < 2e-3 * a, (flags, f.mean())
