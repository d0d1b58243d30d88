"""Image-quality measurement: relative MSE against a converged reference.

relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 0.01), with r
rendered by the same library at many samples per pixel into an FP64 film
(16,384 spp by default).
"""
from __future__ import annotations

import numpy as np


def rel_mse(x, ref, eps=0.01):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


def reference_image(scene, width, height, spp=16384, chunk=256, seed=0x5EED, **kw):
    """Mean radiance per pixel from `spp` samples accumulated in float64 on the
    GPU (chunks of `chunk` spp, distinct sample indices, distinct seed from
    the images it judges)."""
    film = np.zeros((height, width, 3), np.float64)
    done = 0
    while done < spp:
        n = min(chunk, spp - done)
        scene.render(width, height, n, film=film, spp_offset=done, seed=seed, **kw)
        done += n
    return film / spp
