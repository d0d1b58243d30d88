"""Shared test helpers: seeded ray batches and image metrics."""
from __future__ import annotations

import numpy as np


def rays_from(o, d, tmax=None):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = max(o.shape[0], d.shape[0])
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = np.inf if tmax is None else tmax
    r[:, 4:7] = d
    return r


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def random_dirs(rng, n):
    return unit(rng.normal(size=(n, 3)))


def camera_rays(desc, width, height, jitter_seed=None):
    """Pinhole camera rays through pixel centres (Camera.h:45-55), float32."""
    cam = desc.camera
    tx, ty = cam.tans()
    ys, xs = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    fx = (xs.astype(np.float32) + np.float32(0.5)) / np.float32(width)
    fy = np.float32(1) - (ys.astype(np.float32) + np.float32(0.5)) / np.float32(height)
    d = np.stack([tx * (fx - np.float32(0.5)) * 2, ty * (fy - np.float32(0.5)) * 2, np.ones_like(fx)], -1)
    d = unit(d.reshape(-1, 3))
    M = cam.cam_to_world.astype(np.float32)
    dw = (d @ M[:3, :3].T).astype(np.float32)
    o = np.broadcast_to(M[:3, 3], dw.shape)
    return rays_from(o, dw)


def scene_rays(rng, desc, n, lo, hi, tmax=None):
    """Random origins in the box [lo, hi] with random directions."""
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    return rays_from(o, random_dirs(rng, n), tmax)


def rel_mse(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def hit_prims(h):
    return np.ascontiguousarray(h[:, 1]).view(np.uint32)
