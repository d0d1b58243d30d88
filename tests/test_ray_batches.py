"""Inputs of the ray-query microbenchmark (scenes.ray_batch): the direction
distributions against their closed forms and the origins against the
surface, on a flat height field where the geometric normal is +y."""
import math

import numpy as np
import pytest

from xitils_amd import scenes


@pytest.fixture(scope="module")
def flat():
    return scenes.heightfield(cells=32, width=64, height=36, amplitude=0.0, materials=False)


def _ks_uniform(x, lo, hi):
    """Kolmogorov-Smirnov distance of x from U[lo, hi]."""
    x = np.sort((np.asarray(x, np.float64) - lo) / (hi - lo))
    n = x.size
    return max(np.max(np.arange(1, n + 1) / n - x), np.max(x - np.arange(n) / n))


def test_layout_and_unit_directions(flat):
    for kind in scenes.RAY_BATCH_KINDS:
        r = scenes.ray_batch(flat, kind, 64, 36)
        assert r.shape == (64 * 36, 8) and r.dtype == np.float32
        assert np.all(np.isinf(r[:, 3])) and np.all(r[:, 7] == 0)
        assert np.allclose(np.linalg.norm(r[:, 4:7], axis=1), 1.0, atol=1e-6)
    with pytest.raises(ValueError):
        scenes.ray_batch(flat, "bogus")


def test_diffuse_is_cosine_weighted(flat):
    """For cosine-weighted directions about n, cos^2(theta) is U[0, 1] and
    E[cos theta] = 2/3; origins lie on the plane y = 0, offset by 1e-5 d."""
    r = scenes.ray_batch(flat, "diffuse", 256, 256)
    cy = r[:, 5].astype(np.float64)
    assert np.all(cy >= 0)
    assert _ks_uniform(cy ** 2, 0.0, 1.0) < 0.01
    assert abs(cy.mean() - 2.0 / 3.0) < 0.005
    assert np.allclose(r[:, 1], 1e-5 * r[:, 5], atol=1e-9)
    assert np.all((r[:, 0] >= 0) & (r[:, 0] <= 1) & (r[:, 2] >= 0) & (r[:, 2] <= 1))


def test_shadow_rays_fill_the_sun_cone(flat):
    """Uniform over the cone's solid angle: cos(theta) to the sun axis is
    U[cos(theta_max), 1]."""
    r = scenes.ray_batch(flat, "shadow", 256, 256)
    sun = np.asarray(flat.env.sun_dir, np.float64)
    sun /= np.linalg.norm(sun)
    c = r[:, 4:7].astype(np.float64) @ sun
    cmax = math.cos(math.radians(flat.env.sun_half_angle_deg))
    assert np.all(c >= cmax - 1e-6)
    # float32 directions quantise cos near 1 to ~6e-8 steps, coarse against the
    # cone's 1 - cos(theta_max) = 9.5e-6: compare the angle instead (theta^2 ~ U)
    th2 = np.arccos(np.clip(c, -1, 1)) ** 2
    assert _ks_uniform(th2, 0.0, math.radians(flat.env.sun_half_angle_deg) ** 2) < 0.02


def test_primary_rays_jitter_around_the_pixel_corner(flat):
    """The film position of each primary ray, recovered from its camera-space
    direction, is x/W + (u - 1/2)/W and 1 - y/H + (v - 1/2)/H
    (RenderTarget.h:68-80: the jitter is centred on the pixel corner), with u
    and v uniform."""
    W, H = 64, 36
    r = scenes.ray_batch(flat, "primary", W, H)
    cam = flat.camera
    M = cam.cam_to_world.astype(np.float64)
    assert np.allclose(r[:, 0:3], M[:3, 3])
    dc = r[:, 4:7].astype(np.float64) @ M[:3, :3]          # world -> camera (orthonormal rotation)
    tx, ty = (float(v) for v in cam.tans())
    fx = dc[:, 0] / dc[:, 2] / (2 * tx) + 0.5
    fy = dc[:, 1] / dc[:, 2] / (2 * ty) + 0.5
    ys, xs = np.divmod(np.arange(W * H), W)
    u = (fx - xs / W) * W + 0.5
    v = (fy - (1.0 - ys / H)) * H + 0.5
    assert np.all((u > -1e-4) & (u < 1 + 1e-4)) and np.all((v > -1e-4) & (v < 1 + 1e-4))
    assert _ks_uniform(u, 0, 1) < 0.05 and _ks_uniform(v, 0, 1) < 0.05
