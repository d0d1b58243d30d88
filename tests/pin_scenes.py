"""Scenes with closed-form answers, shared by the oracle pins (CPU) and the
GPU parity tests: instanced objects with explicit inverse transforms
(Object.h:16-111, Transform.h:13-30), several emitters (Scene.h:40-55), a
stack of pass-through surfaces for the roulette weighting (PathTracer.h:149-155)
and Lambert's polygon irradiance formula."""
from __future__ import annotations

import math

import numpy as np

from xitils_amd import scenes


def trs(t=(0.0, 0.0, 0.0), axis=(0.0, 1.0, 0.0), angle_deg=0.0, scale=1.0):
    """objectToWorld = translate(t) * rotate(axis, angle) * scale(s) and its
    explicit inverse (as Transform keeps matInv, Transform.h:13-30), both
    computed in float64 and rounded to float32."""
    ax = np.asarray(axis, np.float64)
    ax = ax / np.linalg.norm(ax)
    th = math.radians(angle_deg)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * (K @ K)   # Rodrigues
    M = np.eye(4)
    M[:3, :3] = R * scale
    M[:3, 3] = t
    Mi = np.eye(4)
    Mi[:3, :3] = R.T / scale
    Mi[:3, 3] = -(R.T / scale) @ np.asarray(t, np.float64)
    return M.astype(np.float32), Mi.astype(np.float32)


def world_points(obj: scenes.ObjectDesc, pts=None) -> np.ndarray:
    p = obj.positions if pts is None else np.asarray(pts)
    p = np.asarray(p, np.float64)
    if obj.xf is None:
        return p
    M = obj.xf.astype(np.float64)
    return p @ M[:3, :3].T + M[:3, 3]


def polygon_irradiance(p, n, poly) -> float:
    """Irradiance at point p with unit normal n from a unit-radiance planar
    polygon (vertices in order, either winding), fully above p's tangent
    plane: Lambert's formula E = 1/2 |sum_i beta_i (g_i . n)| with beta_i the
    angle subtended by edge i and g_i the unit normal of the plane through p
    and that edge."""
    v = np.asarray(poly, np.float64) - np.asarray(p, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    n = np.asarray(n, np.float64)
    E = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        g = np.cross(a, b)
        E += math.acos(float(np.clip(a @ b, -1, 1))) * float(g @ n) / np.linalg.norm(g)
    return abs(E) / 2.0


def strip_mesh(w, d, cuts):
    """A w x d rectangle in the plane y = 0, facing -y, triangulated as a fan of
    quads split at the x positions `cuts` (fractions of w), so its triangles
    have unequal areas.  Returns (positions, indices, outline)."""
    xs = [0.0] + sorted(cuts) + [1.0]
    P, I = [], []
    for k in range(len(xs) - 1):
        x0, x1 = xs[k] * w - w / 2, xs[k + 1] * w - w / 2
        b = len(P)
        P += [(x0, 0, -d / 2), (x0, 0, d / 2), (x1, 0, d / 2), (x1, 0, -d / 2)]
        # (p1 - p0) x (p2 - p0) points to -y
        I += [(b, b + 2, b + 1), (b, b + 3, b + 2)]
    outline = np.array([(-w / 2, 0, -d / 2), (-w / 2, 0, d / 2), (w / 2, 0, d / 2), (w / 2, 0, -d / 2)], np.float64)
    return np.array(P, np.float32), np.array(I, np.int32), outline


def two_lights(albedo=0.5, LA=(3.0, 3.0, 3.0), LB=(2.0, 1.5, 1.0), scale=2.0):
    """A Lambert floor lit by two emitters: A, a 0.4 x 0.3 quad of 2 triangles
    with the identity transform, and B, a 0.2 x 0.15 strip of 6 triangles of
    unequal areas instanced under a uniform scale, a rotation about y and a
    translation.  A camera with a tiny field of view looks straight down at
    the floor point q.  Returns (desc, q, expected radiance at q with one
    bounce) where expected = a / pi * (LA * E_A + LB * E_B)."""
    s = scenes.SceneDesc(name="two_lights")
    fl = s.add_material(scenes.diffuse(albedo))
    ma = s.add_material(scenes.emission(*LA))
    mb = s.add_material(scenes.emission(*LB))
    v, i = scenes.quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    s.objects.append(scenes.ObjectDesc(v, i, material=fl))
    ya = 1.0
    va, ia = scenes.quad((-0.5, ya, 0.1), (-0.1, ya, 0.1), (-0.1, ya, 0.4), (-0.5, ya, 0.4))   # faces -y
    s.objects.append(scenes.ObjectDesc(va, ia, material=ma))
    vb, ib, outline = strip_mesh(0.2, 0.15, [0.1, 0.45])
    M, Mi = trs((0.35, 0.8, -0.2), (0, 1, 0), 30.0, scale)
    ob = scenes.ObjectDesc(vb, ib, material=mb, xf=M, xf_inv=Mi)
    s.objects.append(ob)
    q = np.array([0.05, 0.0, 0.05])
    s.camera = scenes.Camera(scenes.look_at(q + (0, 0.5, 0), q + (0, 0, 1e-4), up=(0, 0, 1)), math.radians(0.01), 1.0)
    Ea = polygon_irradiance(q, (0, 1, 0), va.astype(np.float64))
    Eb = polygon_irradiance(q, (0, 1, 0), world_points(ob, outline))
    want = albedo / math.pi * (np.asarray(LA) * Ea + np.asarray(LB) * Eb)
    return s, q, want


def passthrough_stack(n_planes=12, width=32, height=32):
    """n parallel 100 x 100 planes at y = 1 .. n of SpecularRefraction with
    index 1 (Material.h:191-218: the ray passes straight through with weight
    1 and pdf -1), under a unit sky, seen by a camera at the origin looking up
    +y.  Every path crosses all n planes and escapes: it is a path of n + 1
    segments whose only weight changes are the roulette's (PathTracer.h:
    149-155)."""
    s = scenes.SceneDesc(name="passthrough_stack")
    m = s.add_material(scenes.Material(scenes.MAT_SPEC_REFR, (1.0,)))
    for k in range(1, n_planes + 1):
        v, i = scenes.quad((-50, k, -50), (50, k, -50), (50, k, 50), (-50, k, 50))
        s.objects.append(scenes.ObjectDesc(v, i, material=m))
    s.env = scenes.Env(True, (1.0, 1.0, 1.0))
    s.camera = scenes.Camera(scenes.look_at((0, 0, 0), (0, 1, 1e-4), up=(0, 0, 1)), math.radians(20.0),
                             height / width)
    return s


def instanced_room(width=64, height=64):
    """A Cornell-style room whose contents are instances: a unit box and a
    smooth sphere (vertex normals) placed by rotations about tilted axes,
    uniform scales and translations, and two emitters, the ceiling quad
    (identity) and a 6-triangle strip instanced under a uniform scale of 1.5
    and a rotation, hanging under the ceiling."""
    s = scenes.cornell(width, height)
    s.name = "instanced_room"
    light2 = s.add_material(scenes.emission(4.0, 6.0, 9.0))
    mirror = s.add_material(scenes.metal(0.3, (0.9, 0.8, 0.6)))
    # replace the two pre-rotated boxes by instances of object-space meshes
    s.objects = [o for k, o in enumerate(s.objects) if k not in (1, 2)]
    bv, bi = scenes.box_mesh((1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 0.0)
    M, Mi = trs((0.3, 0.24, 0.35), (0.2, 1.0, 0.1), -25.0, 0.3)
    s.objects.append(scenes.ObjectDesc(bv, bi, material=0, xf=M, xf_inv=Mi))
    M, Mi = trs((0.68, 0.31, 0.62), (0.0, 1.0, 0.0), 18.0, 0.6)
    tv, ti = scenes.box_mesh((0.5, 1.0, 0.5), (0.0, 0.0, 0.0), 0.0)
    s.objects.append(scenes.ObjectDesc(tv, ti, material=0, xf=M, xf_inv=Mi))
    sp, si, sn = scenes.uv_sphere((0.0, 0.0, 0.0), 1.0, nu=32, nv=16)
    M, Mi = trs((0.62, 0.72, 0.3), (1.0, 0.3, 0.2), 40.0, 0.12)
    s.objects.append(scenes.ObjectDesc(sp, si, material=mirror, normals=sn, xf=M, xf_inv=Mi))
    vb, ib, _ = strip_mesh(0.12, 0.08, [0.2, 0.65])
    M, Mi = trs((0.3, 0.85, 0.6), (0.0, 1.0, 0.0), 35.0, 1.5)
    s.objects.append(scenes.ObjectDesc(vb, ib, material=light2, xf=M, xf_inv=Mi))
    return s
