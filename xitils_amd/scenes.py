"""Synthetic scenes for the path-tracing workloads (pure numpy, seeded).

Every scene is returned as a :class:`SceneDesc`: a list of objects in the
reference's sense (one mesh, one transform, one material per object, as
``Object(shape, material, objectToWorld)`` in Object.h:16-40, with an optional
per-triangle material override), a material table, an environment and a
pinhole camera.  The same description feeds the HIP library and the CPU
checker, so both see bit-identical inputs.

Material parameter layout (7 floats, shared with include/xitils_amd.h):
  DIFFUSE       albedo r, g, b                          (Material.h:50-86)
  SPEC_REFL     -                                       (Material.h:177-189)
  SPEC_REFR     index                                   (Material.h:191-218)
  SPEC_FRESNEL  index                                   (Material.h:220-273)
  METAL         alpha (= roughness^2), f0 r, g, b       (Material.h:275-384)
  EMISSION      power r, g, b                           (Material.h:386-411)
  GGX_LAMBERT   albedo r, g, b, alpha, f0, p_spec       (new; lobe mixture as
                                                          GlossyWithHighLight,
                                                          Material.h:134-175)
  MICROFACET_NMAP  base type, base p[0..5]              (MicrofacetNormalMapping,
                                                          Sandbox/MicrofacetBasedNormal
                                                          Mapping/Main.cpp:86-340)
A material's normal map (Material::normalmap, applied as Object.h:59-68) is an
index into SceneDesc.textures.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

(MAT_DIFFUSE, MAT_SPEC_REFL, MAT_SPEC_REFR, MAT_SPEC_FRESNEL, MAT_METAL, MAT_EMISSION, MAT_GGX_LAMBERT,
 MAT_MICROFACET_NMAP) = range(8)

FLAG_UNIFORM_BARY = 1
FLAG_ENV_NEE = 2
FLAG_SORT_MATERIALS = 16

INTEGRATOR_STANDARD = 0
INTEGRATOR_NAIVE = 1


@dataclass
class Material:
    type: int
    params: tuple = (0.0,) * 7
    normal_map: int = -1     # index into SceneDesc.textures (Material::normalmap), -1 = none

    def p7(self) -> np.ndarray:
        p = np.zeros(7, np.float32)
        p[: len(self.params)] = self.params
        return p


def diffuse(r, g=None, b=None):
    g = r if g is None else g
    b = r if b is None else b
    return Material(MAT_DIFFUSE, (r, g, b))


def emission(r, g, b):
    return Material(MAT_EMISSION, (r, g, b))


def metal(roughness, f0):
    # Metal(roughness, f0): alpha = roughness * roughness (Material.h:282-286)
    a = np.float32(roughness) * np.float32(roughness)
    return Material(MAT_METAL, (float(a), f0[0], f0[1], f0[2]))


def ggx_lambert(albedo, alpha, f0, p_spec=0.5):
    return Material(MAT_GGX_LAMBERT, (albedo[0], albedo[1], albedo[2], alpha, f0, p_spec))


def microfacet_nmap(base: Material, normal_map: int = -1):
    """MicrofacetNormalMapping(base) (Main.cpp:86-92); base is DIFFUSE, METAL
    or GGX_LAMBERT."""
    if base.type not in (MAT_DIFFUSE, MAT_METAL, MAT_GGX_LAMBERT):
        raise ValueError("microfacet normal mapping needs a Diffuse, Metal or GGX+Lambert base")
    return Material(MAT_MICROFACET_NMAP, (float(base.type),) + tuple(base.p7()[:6]), normal_map)


@dataclass
class ObjectDesc:
    positions: np.ndarray                 # (nv, 3) float32, object space
    indices: np.ndarray                   # (nt, 3) int32
    material: int = 0
    normals: Optional[np.ndarray] = None  # (nv, 3) float32
    xf: Optional[np.ndarray] = None       # (4, 4) float32 Transform::mat
    xf_inv: Optional[np.ndarray] = None   # (4, 4) float32 Transform::matInv
    tri_materials: Optional[np.ndarray] = None  # (nt,) uint32
    texcoords: Optional[np.ndarray] = None      # (nv, 2) float32
    tangents: Optional[np.ndarray] = None       # (nv, 3) float32
    bitangents: Optional[np.ndarray] = None     # (nv, 3) float32

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])


ENV_MAP_NONE, ENV_MAP_ANGULAR, ENV_MAP_LATLONG = 0, 1, 2


@dataclass
class Env:
    has_sky: bool = False
    sky: tuple = (0.0, 0.0, 0.0)
    sun_dir: tuple = (0.0, 1.0, 0.0)
    sun_half_angle_deg: float = 0.0
    sun_radiance: tuple = (0.0, 0.0, 0.0)
    # image sky: ENV_MAP_ANGULAR (SkySphereFromImage) or ENV_MAP_LATLONG, with
    # map an (H, W, 3) float32 array, row 0 first
    map_kind: int = ENV_MAP_NONE
    map: np.ndarray = None

    def sun_terms(self):
        """(omc, sin2) of the disk half angle, in float64 then rounded."""
        th = math.radians(self.sun_half_angle_deg)
        omc = 2.0 * math.sin(th / 2.0) ** 2  # 1 - cos, without cancellation
        sin2 = math.sin(th) ** 2
        return np.float32(omc), np.float32(sin2)


@dataclass
class Camera:
    cam_to_world: np.ndarray   # (4, 4) float32
    fov: float                 # radians
    aspect: float              # height / width, as PinholeCamera (Camera.h:32-36)

    def tans(self):
        # Camera.h:49-50: tanf(fov/2) and tanf(fov/2 * aspect) (quirk Q9)
        f = np.float32(self.fov)
        a = np.float32(self.aspect)
        tx = np.tan(np.float32(f / np.float32(2)), dtype=np.float32)
        ty = np.tan(np.float32(np.float32(f / np.float32(2)) * a), dtype=np.float32)
        return np.float32(tx), np.float32(ty)

    def packed(self) -> np.ndarray:
        tx, ty = self.tans()
        return np.concatenate([self.cam_to_world.astype(np.float32).reshape(16), np.array([tx, ty], np.float32)])


@dataclass
class SceneDesc:
    objects: List[ObjectDesc] = field(default_factory=list)
    materials: List[Material] = field(default_factory=list)
    env: Env = field(default_factory=Env)
    camera: Optional[Camera] = None
    name: str = ""
    textures: List[np.ndarray] = field(default_factory=list)   # (H, W, 3) float32, row 0 first

    @property
    def num_triangles(self) -> int:
        return sum(o.num_triangles for o in self.objects)

    def add_material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1


# --------------------------------------------------------------------------- helpers

def translate(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    mi = np.eye(4, dtype=np.float32)
    mi[:3, 3] = (-np.float32(x), -np.float32(y), -np.float32(z))
    return m, mi


def look_at(pos, target, up=(0.0, 1.0, 0.0)):
    """Camera-to-world with columns (right, up, forward), translation pos,
    as Transform.h lookAt (camera space looks down +z)."""
    pos = np.asarray(pos, np.float64)
    d = np.asarray(target, np.float64) - pos
    d /= np.linalg.norm(d)
    right = np.cross(np.asarray(up, np.float64), d)
    right /= np.linalg.norm(right)
    nup = np.cross(d, right)
    m = np.eye(4, dtype=np.float64)
    m[:3, 0] = right
    m[:3, 1] = nup
    m[:3, 2] = d
    m[:3, 3] = pos
    return m.astype(np.float32)


def box_mesh(size, center, rot_y_deg):
    """Axis box of extent size=(sx, sy, sz) resting on its center, rotated
    about +y, emitted directly in world space (12 triangles, outward faces)."""
    sx, sy, sz = (np.float64(v) / 2 for v in size)
    c = np.array([[-sx, -sy, -sz], [sx, -sy, -sz], [sx, sy, -sz], [-sx, sy, -sz],
                  [-sx, -sy, sz], [sx, -sy, sz], [sx, sy, sz], [-sx, sy, sz]], np.float64)
    t = math.radians(rot_y_deg)
    rot = np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])
    v = c @ rot.T + np.asarray(center, np.float64)
    # faces (outward normals via cross(p1-p0, p2-p0))
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2)]
    idx = []
    for a, b, cc, d in quads:
        idx += [(a, b, cc), (a, cc, d)]
    return v.astype(np.float32), np.array(idx, np.int32)


def quad(p0, p1, p2, p3):
    v = np.array([p0, p1, p2, p3], np.float32)
    return v, np.array([(0, 1, 2), (0, 2, 3)], np.int32)


# --------------------------------------------------------------------------- Cornell box

def cornell(width=256, height=256):
    """Unit-cube room, 5 walls as 10 triangles (left red, right green, others
    0.73 grey), a 0.3 short box rotated -18 deg and a 0.3x0.6x0.3 tall box
    rotated +18 deg, and a 0.25 x 0.2 ceiling quad light of radiance
    (17, 12, 4) facing down.  Pinhole camera at (0.5, 0.5, -1.35) looking +z,
    fov 40 deg."""
    s = SceneDesc(name="cornell")
    white = s.add_material(diffuse(0.73))
    red = s.add_material(diffuse(0.63, 0.065, 0.05))
    green = s.add_material(diffuse(0.14, 0.45, 0.09))
    light = s.add_material(emission(17.0, 12.0, 4.0))
    # normals point into the room
    walls = [
        (quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), white),   # floor y=0, normal +y
        (quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), white),   # ceiling y=1, normal -y
        (quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), white),   # back z=1, normal -z
        (quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), red),     # left x=0, normal +x
        (quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), green),   # right x=1, normal -x
    ]
    wv = np.concatenate([w[0][0] for w in walls])
    wi = np.concatenate([w[0][1] + 4 * k for k, w in enumerate(walls)])
    wm = np.concatenate([[w[1], w[1]] for w in walls]).astype(np.uint32)
    s.objects.append(ObjectDesc(wv, wi, material=white, tri_materials=wm))
    bv, bi = box_mesh((0.3, 0.3, 0.3), (0.33, 0.15, 0.35), -18.0)
    s.objects.append(ObjectDesc(bv, bi, material=white))
    tv, ti = box_mesh((0.3, 0.6, 0.3), (0.66, 0.3, 0.66), 18.0)
    s.objects.append(ObjectDesc(tv, ti, material=white))
    lx0, lx1, lz0, lz1, ly = 0.375, 0.625, 0.4, 0.6, 0.999
    lv, li = quad((lx0, ly, lz0), (lx1, ly, lz0), (lx1, ly, lz1), (lx0, ly, lz1))  # normal -y
    s.objects.append(ObjectDesc(lv, li, material=light))
    m, _ = translate(0.5, 0.5, -1.35)
    s.camera = Camera(m, math.radians(40.0), height / width)
    return s


# --------------------------------------------------------------------------- heightfield

def _hash2(ix, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iz.astype(np.uint64) * np.uint64(0xD8163841)
         ^ np.uint64(seed) * np.uint64(0xCB1AB31F)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0x5BD1E995)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(x, z, seed):
    ix, iz = np.floor(x), np.floor(z)
    fx, fz = x - ix, z - iz
    ux = fx * fx * (3 - 2 * fx)
    uz = fz * fz * (3 - 2 * fz)
    ix = ix.astype(np.int64)
    iz = iz.astype(np.int64)
    a = _hash2(ix, iz, seed)
    b = _hash2(ix + 1, iz, seed)
    c = _hash2(ix, iz + 1, seed)
    d = _hash2(ix + 1, iz + 1, seed)
    return (a * (1 - ux) + b * ux) * (1 - uz) + (c * (1 - ux) + d * ux) * uz


def fbm(x, z, seed, octaves=6, base=4.0, lacunarity=2.0, gain=0.5):
    out = np.zeros_like(x)
    amp, freq = 0.5, base
    for o in range(octaves):
        out += amp * value_noise(x * freq, z * freq, seed + 7919 * o)
        amp *= gain
        freq *= lacunarity
    return out


def heightfield(cells=1024, width=1920, height=1080, seed=2, amplitude=0.1, materials=True):
    """cells x cells grid over [0,1]^2 (2 * cells^2 triangles), y =
    amplitude * fBm_6(x, z); per-triangle GGX-over-Lambert with alpha ~
    U[0.05, 0.5], F0 ~ U[0.02, 0.08], albedo ~ U[0.2, 0.8]; constant sky 1.0 plus
    a 0.5 deg sun of radiance 5e4 at elevation 30 deg, azimuth 45 deg."""
    s = SceneDesc(name=f"heightfield{cells}")
    n = cells
    g = np.linspace(0.0, 1.0, n + 1, dtype=np.float64)
    X, Z = np.meshgrid(g, g, indexing="ij")          # X[i, j] = i/n, Z[i, j] = j/n
    Y = amplitude * fbm(X, Z, seed)
    pos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    v00 = (i * (n + 1) + j).reshape(-1)
    v01 = (i * (n + 1) + j + 1).reshape(-1)
    v10 = ((i + 1) * (n + 1) + j).reshape(-1)
    v11 = ((i + 1) * (n + 1) + j + 1).reshape(-1)
    t0 = np.stack([v00, v01, v11], axis=1)
    t1 = np.stack([v00, v11, v10], axis=1)
    idx = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.int32)
    nt = idx.shape[0]
    if materials:
        rng = np.random.default_rng(seed)
        alb = rng.uniform(0.2, 0.8, size=(nt, 3)).astype(np.float32)
        alpha = rng.uniform(0.05, 0.5, size=nt).astype(np.float32)
        f0 = rng.uniform(0.02, 0.08, size=nt).astype(np.float32)
        s.materials = _ggx_lambert_table(alb, alpha, f0)
        tri_m = np.arange(nt, dtype=np.uint32)
    else:
        s.add_material(diffuse(0.5))
        tri_m = None
    s.objects.append(ObjectDesc(pos, idx, material=0, tri_materials=tri_m))
    el, az = math.radians(30.0), math.radians(45.0)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    s.env = Env(True, (1.0, 1.0, 1.0), sun, 0.25, (5e4, 5e4, 5e4))
    cam = look_at((0.5, 0.42, -0.28), (0.5, 0.0, 0.55))
    s.camera = Camera(cam, math.radians(50.0), height / width)
    return s


class MaterialTable(list):
    """Material list backed by a packed (n, 8) array, so a per-triangle table
    of millions of materials is not built as Python objects."""

    def __init__(self, packed: np.ndarray):
        super().__init__()
        self.packed = packed

    def __len__(self):
        return self.packed.shape[0]

    def __getitem__(self, i):
        row = self.packed[i]
        return Material(int(row[0].view(np.uint32)) if row.dtype == np.float32 else int(row[0]), tuple(row[1:8]))


def _ggx_lambert_table(alb, alpha, f0, p_spec=0.5):
    nt = alb.shape[0]
    packed = np.zeros((nt, 8), np.float32)
    packed[:, 0] = np.array([MAT_GGX_LAMBERT], np.uint32).view(np.float32)[0]
    packed[:, 1:4] = alb
    packed[:, 4] = alpha
    packed[:, 5] = f0
    packed[:, 6] = p_spec
    return MaterialTable(packed)


def material_normal_maps(mats):
    """(n,) uint32 of 1 + texture index per material (0 = none), or None."""
    if isinstance(mats, MaterialTable):
        return None
    nm = np.array([getattr(m, "normal_map", -1) + 1 for m in mats], np.uint32)
    return nm if nm.any() else None


def pack_materials(mats) -> np.ndarray:
    """(n, 8) float32 records {type bits, p[7]} as xt_material."""
    if isinstance(mats, MaterialTable):
        return np.ascontiguousarray(mats.packed)
    out = np.zeros((len(mats), 8), np.float32)
    for k, m in enumerate(mats):
        out[k, 0] = np.array([m.type], np.uint32).view(np.float32)[0]
        out[k, 1:] = m.p7()
    return out


# --------------------------------------------------------------------------- ray batches

RAY_BATCH_KINDS = ("primary", "diffuse", "shadow")


def _frame(n):
    """Two unit vectors completing each row of n (unit, (m, 3) float64) to an
    orthonormal basis."""
    a = np.where(np.abs(n[:, :1]) < 0.5, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return t, np.cross(n, t)


def _surface_points(desc, rng, n):
    """n points uniform over each triangle of the first object, triangles
    chosen uniformly by index; returns positions and geometric normals
    oriented to +y (height-field-like scenes)."""
    obj = desc.objects[0]
    pos = obj.positions.astype(np.float64)
    tri = obj.indices[rng.integers(0, obj.indices.shape[0], n)]
    p0, p1, p2 = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    su = np.sqrt(rng.random(n))[:, None]
    v = rng.random(n)[:, None]
    p = (1.0 - su) * p0 + su * (1.0 - v) * p1 + su * v * p2
    ng = np.cross(p1 - p0, p2 - p0)
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    ng *= np.where(ng[:, 1:2] < 0.0, -1.0, 1.0)
    return p, ng


def ray_batch(desc, kind, width=1920, height=1080, seed=7, eps=1e-5):
    """width*height rays of one kind as (n, 8) float32 {o, tmax, d, 0}, the
    inputs of the ray-query microbenchmark on the height field:
      primary -- the scene camera's rays (PinholeCamera::generateRay,
                 Camera.h:45-55), one per pixel, jittered inside the pixel as
                 RenderTargetTile::GenerateFilmPosition (RenderTarget.h:68-80);
      diffuse -- cosine-weighted bounce directions (Vector.h:817-824) about the
                 geometric normal at uniformly chosen surface points;
      shadow  -- directions uniform in the sun's cone from surface points that
                 face the sun, tmax = inf (any-hit queries).
    Bounce and shadow origins are offset by eps along d (PathTracer.h:254)."""
    if kind not in RAY_BATCH_KINDS:
        raise ValueError(f"ray batch kind must be one of {RAY_BATCH_KINDS}")
    n = width * height
    rng = np.random.default_rng([seed, RAY_BATCH_KINDS.index(kind)])
    if kind == "primary":
        cam = desc.camera
        tx, ty = (float(v) for v in cam.tans())
        ys, xs = np.divmod(np.arange(n), width)
        fx = xs / width + (rng.random(n) - 0.5) / width
        fy = 1.0 - ys / height + (rng.random(n) - 0.5) / height
        dc = np.stack([tx * (fx - 0.5) * 2, ty * (fy - 0.5) * 2, np.ones(n)], axis=1)
        dc /= np.linalg.norm(dc, axis=1, keepdims=True)
        M = cam.cam_to_world.astype(np.float64)
        d = dc @ M[:3, :3].T
        o = np.broadcast_to(M[:3, 3], d.shape)
    elif kind == "diffuse":
        p, ng = _surface_points(desc, rng, n)
        u, v = rng.random(n), rng.random(n)
        r, phi = np.sqrt(u), 2.0 * np.pi * v
        t, b = _frame(ng)
        d = np.sqrt(1.0 - u)[:, None] * ng + (r * np.cos(phi))[:, None] * t + (r * np.sin(phi))[:, None] * b
        o = p + eps * d
    else:
        env = desc.env
        sun = np.asarray(env.sun_dir, np.float64)
        sun /= np.linalg.norm(sun)
        p, ng = _surface_points(desc, rng, 2 * n)
        keep = np.nonzero(ng @ sun > 0.0)[0]
        if keep.size < n:
            raise ValueError("too few surface points face the sun for a shadow-ray batch")
        p = p[keep[:n]]
        cmax = math.cos(math.radians(env.sun_half_angle_deg))
        ct = 1.0 - rng.random(n) * (1.0 - cmax)
        st, phi = np.sqrt(np.maximum(0.0, 1.0 - ct * ct)), 2.0 * np.pi * rng.random(n)
        t, b = _frame(np.broadcast_to(sun, (n, 3)).copy())
        d = ct[:, None] * sun + (st * np.cos(phi))[:, None] * t + (st * np.sin(phi))[:, None] * b
        o = p + eps * d
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = np.inf
    rays[:, 4:7] = d
    return rays


# --------------------------------------------------------------------------- small test scenes

def uv_sphere(center, radius, nu=24, nv=12, smooth=True):
    th = np.linspace(0, math.pi, nv + 1)
    ph = np.linspace(0, 2 * math.pi, nu + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pos = np.asarray(center) + radius * nrm
    idx = []
    for a in range(nv):
        for b in range(nu):
            v0 = a * (nu + 1) + b
            v1 = v0 + 1
            v2 = v0 + (nu + 1)
            v3 = v2 + 1
            if a != 0:
                idx.append((v0, v1, v2))
            if a != nv - 1:
                idx.append((v1, v3, v2))
    return pos.astype(np.float32), np.array(idx, np.int32), (nrm.astype(np.float32) if smooth else None)


def material_zoo(width=64, height=64):
    """Cornell room whose contents cover every material kind: a smooth-normal
    sphere of GGX-over-Lambert, a metal box, a mirror quad, a glass sphere and
    a Fresnel sphere; sky off (closed room)."""
    s = cornell(width, height)
    s.name = "material_zoo"
    s.objects = s.objects[:1] + s.objects[3:]   # walls + light
    ggx = s.add_material(ggx_lambert((0.6, 0.5, 0.3), 0.2, 0.04))
    met = s.add_material(metal(0.4, (0.9, 0.6, 0.3)))
    mir = s.add_material(Material(MAT_SPEC_REFL))
    glass = s.add_material(Material(MAT_SPEC_REFR, (1.5,)))
    fres = s.add_material(Material(MAT_SPEC_FRESNEL, (1.5,)))
    p, i, nrm = uv_sphere((0.3, 0.18, 0.4), 0.17)
    s.objects.append(ObjectDesc(p, i, material=ggx, normals=nrm))
    bv, bi = box_mesh((0.25, 0.45, 0.25), (0.7, 0.225, 0.7), 25.0)
    s.objects.append(ObjectDesc(bv, bi, material=met))
    qv, qi = quad((0.05, 0.3, 0.95), (0.05, 0.8, 0.95), (0.45, 0.8, 0.95), (0.45, 0.3, 0.95))
    s.objects.append(ObjectDesc(qv, qi, material=mir))
    p, i, nrm = uv_sphere((0.72, 0.62, 0.35), 0.1, 16, 8)
    s.objects.append(ObjectDesc(p, i, material=glass, normals=nrm))
    p, i, nrm = uv_sphere((0.25, 0.55, 0.3), 0.08, 16, 8)
    s.objects.append(ObjectDesc(p, i, material=fres, normals=nrm))
    return s


def furnace_object(albedo=0.6, width=32, height=32):
    """A convex mesh (box) of Lambert albedo a under a uniform sky of 1 with
    nothing else: every camera ray that hits it returns exactly a (one bounce
    escapes with weight a), every miss returns 1."""
    s = SceneDesc(name="furnace")
    m = s.add_material(diffuse(albedo))
    bv, bi = box_mesh((0.6, 0.6, 0.6), (0.0, 0.0, 0.0), 30.0)
    s.objects.append(ObjectDesc(bv, bi, material=m))
    s.env = Env(True, (1.0, 1.0, 1.0))
    cam = look_at((0.9, 0.7, -1.6), (0.0, 0.0, 0.0))
    s.camera = Camera(cam, math.radians(45.0), height / width)
    return s


# --------------------------------------------------------------------------- image skies

def latlong_directions(width, height):
    """Unit direction at each lat-long texel centre (y up, row 0 at the zenith,
    u = (atan2(d.z, d.x) + pi) / 2pi)."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = 2 * np.pi * u - np.pi
    theta = np.pi * v
    P, T = np.meshgrid(phi, theta)
    return np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)


def procedural_sky_latlong(width=2048, height=1024, sun_dir=(0.4, 0.6, -0.7), sun_half_angle_deg=0.5,
                           sun_radiance=2e4, seed=5):
    """FP32 lat-long HDR: a zenith-to-horizon gradient, a dim ground and a sun
    disk of the given radiance baked into the texels (the importance sampler
    finds it; there is no analytic sun)."""
    d = latlong_directions(width, height)
    up = np.clip(d[..., 1], -1, 1)
    t = np.clip(up, 0, 1)[..., None]
    zen = np.array([0.25, 0.45, 1.0])
    hor = np.array([1.0, 0.95, 0.85])
    sky = (1 - t) * hor + t * zen
    ground = np.array([0.12, 0.1, 0.08]) * np.ones_like(sky)
    img = np.where((up >= 0)[..., None], sky, ground)
    sd = np.asarray(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    c = d @ sd
    img = img + (c >= math.cos(math.radians(sun_half_angle_deg)))[..., None] * sun_radiance
    rng = np.random.default_rng(seed)
    img = img * (1 + 0.05 * rng.random(img.shape))      # fine texture so texels differ
    return img.astype(np.float32)


def procedural_angular_probe(size=256, seed=6):
    """Angular-map light probe (the layout SkySphereFromImage reads, SkySphere.h:
    20-33): texel (u, v) in [0, 1]^2 -> (a, b) = (2u - 1, 2v - 1), r = |(a, b)|,
    theta = pi r about -z; outside the unit disk stays black."""
    u = (np.arange(size) + 0.5) / size
    A, B = np.meshgrid(2 * u - 1, 2 * u - 1)
    r = np.sqrt(A * A + B * B)
    th = np.pi * np.minimum(r, 1)
    s = np.where(r > 0, np.sin(th) / np.maximum(r, 1e-12), 0.0)
    d = np.stack([A * s, B * s, np.cos(th)], -1)
    base = 0.5 + 0.5 * d[..., 1:2] * np.array([0.3, 0.6, 1.0]) + 0.3 * (d[..., 0:1] > 0.5)
    spot = np.exp(40 * (d @ np.array([0.0, 0.6, -0.8]) - 1))[..., None] * np.array([30.0, 25.0, 20.0])
    img = base + spot
    img = img * (1 + 0.05 * np.random.default_rng(seed).random(img.shape))
    img[r > 1] = 0.0
    return img.astype(np.float32)


def sky_scene(kind=ENV_MAP_LATLONG, width=64, height=48, map_size=None, albedo=0.6):
    """A Lambert ground quad with a GGX sphere, a metal box and a diffuse
    sphere under an image sky (angular probe or lat-long HDR); the camera
    sees both objects and sky."""
    s = SceneDesc(name=f"sky_{'angular' if kind == ENV_MAP_ANGULAR else 'latlong'}")
    g = s.add_material(diffuse(albedo))
    ggx = s.add_material(ggx_lambert((0.7, 0.4, 0.3), 0.25, 0.04))
    met = s.add_material(metal(0.3, (0.9, 0.8, 0.6)))
    dif = s.add_material(diffuse(0.3, 0.6, 0.4))
    qv, qi = quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4))
    s.objects.append(ObjectDesc(qv, qi, material=g))
    p, i, n = uv_sphere((0.0, 0.6, 0.0), 0.6)
    s.objects.append(ObjectDesc(p, i, material=ggx, normals=n))
    bv, bi = box_mesh((0.7, 0.7, 0.7), (1.3, 0.35, 0.5), 20.0)
    s.objects.append(ObjectDesc(bv, bi, material=met))
    p, i, n = uv_sphere((-1.2, 0.4, 0.6), 0.4)
    s.objects.append(ObjectDesc(p, i, material=dif, normals=n))
    if kind == ENV_MAP_ANGULAR:
        m = procedural_angular_probe(map_size or 128)
    else:
        ms = map_size or 256
        m = procedural_sky_latlong(ms, ms // 2, sun_half_angle_deg=3.0, sun_radiance=300.0)
    s.env = Env(True, (0.0, 0.0, 0.0), map_kind=kind, map=m)
    cam = look_at((0.5, 1.6, -4.2), (0.0, 0.5, 0.0))
    s.camera = Camera(cam, math.radians(60.0), height / width)
    return s


# --------------------------------------------------------------------------- batch of small scenes

def batch_scene(index, master_seed=3):
    """Scene `index` of the 64-scene batch workload: a unit room (Lambert
    walls, one ceiling area light) holding tessellated spheres and boxes.
    Drawn from the master seed: triangle count log-uniform on [1e4, 2e6],
    image side uniform on [128, 1024] (square), spp log-uniform on [4, 256],
    depth uniform on [4, 32].  Returns (desc, width, height, spp, depth)."""
    rng = np.random.default_rng([master_seed, index])
    tris = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6)))))
    side = int(rng.integers(128, 1025))
    spp = int(round(math.exp(rng.uniform(math.log(4), math.log(256)))))
    depth = int(rng.integers(4, 33))
    s = SceneDesc(name=f"batch{index:02d}")
    wall = s.add_material(diffuse(0.7))
    red = s.add_material(diffuse(0.6, 0.1, 0.08))
    light = s.add_material(emission(12.0, 11.0, 9.0))
    mats = [s.add_material(diffuse(*rng.uniform(0.2, 0.8, 3))) for _ in range(3)]
    mats += [s.add_material(metal(float(rng.uniform(0.2, 0.6)), tuple(rng.uniform(0.5, 0.95, 3)))) for _ in range(2)]
    # room: floor, ceiling, back, left (red), right; open front
    for k, (a, b, c, d) in enumerate([((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)),
                                      ((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),
                                      ((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)),
                                      ((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)),
                                      ((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))]):
        v, i = quad(a, b, c, d)
        s.objects.append(ObjectDesc(v, i, material=red if k == 3 else wall))
    lv, li = quad((0.4, 0.999, 0.4), (0.6, 0.999, 0.4), (0.6, 0.999, 0.6), (0.4, 0.999, 0.6))
    s.objects.append(ObjectDesc(lv, li, material=light))
    nobj = int(rng.integers(3, 13))
    per = max(12, (tris - 12) // nobj)
    for k in range(nobj):
        m = mats[int(rng.integers(0, len(mats)))]
        c = rng.uniform(0.15, 0.85, 3)
        c[1] = rng.uniform(0.1, 0.6)
        if k % 3 == 2:
            bv, bi = box_mesh(tuple(rng.uniform(0.08, 0.25, 3)), tuple(c), float(rng.uniform(0, 90)))
            s.objects.append(ObjectDesc(bv, bi, material=m))
        else:
            nv = max(4, int(math.sqrt(per / 4)))
            p, i, n = uv_sphere(tuple(c), float(rng.uniform(0.05, 0.15)), nu=2 * nv, nv=nv)
            s.objects.append(ObjectDesc(p, i, material=m, normals=n))
    cam = look_at((0.5, 0.5, -1.1), (0.5, 0.45, 0.5))
    s.camera = Camera(cam, math.radians(45.0), 1.0)
    return s, side, side, spp, depth


# --------------------------------------------------------------------------- occluded room

def _box_faces(lo, hi, inward=False):
    """12 triangles of an axis-aligned box; normals outward (or inward)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2)]
    idx = []
    for a, b, cc, d in quads:
        t = [(a, b, cc), (a, cc, d)]
        if inward:
            t = [(a, cc, b), (a, d, cc)]
        idx += t
    return c.astype(np.float32), np.array(idx, np.int32)


def occluded_room(width=1280, height=720, n_tris=500_000, seed=4, light_radiance=500.0):
    """Closed 8 x 4 x 3 room (x, z horizontal, y up) lit only through a 0.05-wide
    vertical slit in the x = 8 wall, behind which an exterior area light of
    radiance 500 faces the slit from a closed light box; ~n_tris triangles of
    seeded boxes and tessellated spheres inside; 50% Lambert, 50% GGX (Metal,
    alpha ~ U[0.1, 0.4]); no sky."""
    s = SceneDesc(name="occluded_room")
    rng = np.random.default_rng(seed)
    wall = s.add_material(diffuse(0.7))
    light = s.add_material(emission(light_radiance, light_radiance, light_radiance))
    X, Y, Z = 8.0, 4.0, 3.0
    # room shell, inward normals, except the x = X wall which has the slit
    slit_z0, slit_z1 = 1.475, 1.525
    verts, idx = [], []

    def add_quad(p0, p1, p2, p3):
        b = sum(v.shape[0] for v in verts)
        verts.append(np.array([p0, p1, p2, p3], np.float32))
        idx.append(np.array([(0, 1, 2), (0, 2, 3)], np.int32) + b)
    add_quad((0, 0, 0), (0, 0, Z), (X, 0, Z), (X, 0, 0))          # floor, +y
    add_quad((0, Y, 0), (X, Y, 0), (X, Y, Z), (0, Y, Z))          # ceiling, -y
    add_quad((0, 0, 0), (0, Y, 0), (0, Y, Z), (0, 0, Z))          # x = 0, +x
    add_quad((0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0))          # z = 0, +z
    add_quad((0, 0, Z), (0, Y, Z), (X, Y, Z), (X, 0, Z))          # z = Z, -z
    # x = X wall (normal -x) with a full-height slit z in [slit_z0, slit_z1]
    add_quad((X, 0, 0), (X, 0, slit_z0), (X, Y, slit_z0), (X, Y, 0))
    add_quad((X, 0, slit_z1), (X, 0, Z), (X, Y, Z), (X, Y, slit_z1))
    s.objects.append(ObjectDesc(np.concatenate(verts), np.concatenate(idx), material=wall))
    # exterior light box behind the slit: the emitter at x = X + 1 faces -x
    lb_v, lb_i = _box_faces((X + 0.001, 0.0, slit_z0 - 0.5), (X + 1.0, Y, slit_z1 + 0.5), inward=True)
    keep = [k for k in range(12) if k not in (4, 5, 6, 7)]   # open toward the slit; x = X + 1 is the emitter
    s.objects.append(ObjectDesc(lb_v, lb_i[keep], material=wall))
    lv, li = quad((X + 1.0, 0.0, slit_z0 - 0.5), (X + 1.0, 0.0, slit_z1 + 0.5),
                  (X + 1.0, Y, slit_z1 + 0.5), (X + 1.0, Y, slit_z0 - 0.5))   # normal -x
    s.objects.append(ObjectDesc(lv, li, material=light))
    # clutter: boxes and spheres, half Lambert, half GGX
    lam = [s.add_material(diffuse(*rng.uniform(0.2, 0.8, 3))) for _ in range(8)]
    ggx = [s.add_material(metal(float(np.sqrt(rng.uniform(0.1, 0.4))), tuple(rng.uniform(0.5, 0.95, 3))))
           for _ in range(8)]
    tris = 0
    k = 0
    while tris < n_tris:
        mat = (lam if k % 2 == 0 else ggx)[int(rng.integers(0, 8))]
        c = rng.uniform((0.5, 0.0, 0.3), (X - 1.5, 1.5, Z - 0.3))
        if k % 3 == 2:
            r = float(rng.uniform(0.08, 0.35))
            nu = int(rng.integers(48, 160))
            p, i, nrm = uv_sphere((c[0], r + c[1] * 0.5, c[2]), r, nu, nu // 2)
            s.objects.append(ObjectDesc(p, i, material=mat, normals=nrm))
            tris += i.shape[0]
        else:
            size = rng.uniform(0.1, 0.6, 3)
            bv, bi = box_mesh(tuple(size), (c[0], size[1] / 2, c[2]), float(rng.uniform(0, 90)))
            s.objects.append(ObjectDesc(bv, bi, material=mat))
            tris += 12
        k += 1
    cam = look_at((0.6, 2.2, 0.4), (6.0, 0.8, 2.2))
    s.camera = Camera(cam, math.radians(70.0), height / width)
    return s


# --------------------------------------------------------------------------- displaced plane (normal mapping)

def _sines(n, f_lo, f_hi, rng):
    """n plane waves: frequency (cycles per unit uv), direction angle, phase."""
    f = rng.uniform(f_lo, f_hi, n)
    ang = rng.uniform(0.0, 2 * np.pi, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    return f, ang, ph


def _sine_height_grad(u, v, f, ang, ph):
    """h = sum_k sin(2 pi f_k (cos a_k u + sin a_k v) + p_k) / f_k and its
    gradient (float64)."""
    h = np.zeros(np.broadcast(u, v).shape)
    hu = np.zeros_like(h)
    hv = np.zeros_like(h)
    for fk, ak, pk in zip(f, ang, ph):
        arg = 2 * np.pi * fk * (np.cos(ak) * u + np.sin(ak) * v) + pk
        h += np.sin(arg) / fk
        c = np.cos(arg) * 2 * np.pi
        hu += c * np.cos(ak)
        hv += c * np.sin(ak)
    return h, hu, hv


def sine_normal_map(size=1024, n_sines=32, f_lo=4.0, f_hi=64.0, seed=5, bump=0.03):
    """Tangent-space normal map of the height field sum_k sin(...)/f_k over
    uv in [0, 1]^2 (32 sines, frequency U[4, 64], amplitude 1/f): texel
    (x, y) at uv = ((x + 0.5) / size, (y + 0.5) / size) stores
    (n + 1) / 2 with n = normalize(-bump dh/du, -bump dh/dv, 1) in
    (tangent, bitangent, normal) order, the layout Object.h:59-68 decodes."""
    rng = np.random.default_rng(seed)
    f, ang, ph = _sines(n_sines, f_lo, f_hi, rng)
    t = (np.arange(size) + 0.5) / size
    U, V = np.meshgrid(t, t)            # row y <-> v
    _, hu, hv = _sine_height_grad(U, V, f, ang, ph)
    n = np.stack([-bump * hu, -bump * hv, np.ones_like(hu)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return ((n + 1) / 2).astype(np.float32)


def displaced_plane_mesh(grid=4096, extent=4.0, height=0.08, n_sines=8, f_lo=1.0, f_hi=4.0, seed=5):
    """A grid x grid vertex plane over [-extent/2, extent/2]^2 in xz,
    displaced along y by `height` times a low-frequency sine sum (the
    geometric displacement), with analytic vertex normals, texCoords (u along
    x, v along z) and the tangent frame dP/du, dP/dv (TriangleMesh.h:58-105
    builds the same attributes for a displaced mesh).  Returns (positions,
    indices, normals, texcoords, tangents, bitangents)."""
    rng = np.random.default_rng(seed + 1000)
    f, ang, ph = _sines(n_sines, f_lo, f_hi, rng)
    t = np.linspace(0.0, 1.0, grid)
    U, V = np.meshgrid(t, t)            # row j <-> v
    h, hu, hv = _sine_height_grad(U, V, f, ang, ph)
    P = np.stack([(U - 0.5) * extent, height * h, (V - 0.5) * extent], -1)
    Pu = np.stack([np.full_like(U, extent), height * hu, np.zeros_like(U)], -1)
    Pv = np.stack([np.zeros_like(U), height * hv, np.full_like(U, extent)], -1)
    N = np.cross(Pv, Pu)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    T = Pu / np.linalg.norm(Pu, axis=-1, keepdims=True)
    B = Pv / np.linalg.norm(Pv, axis=-1, keepdims=True)
    del h, hu, hv, Pu, Pv
    j, i = np.meshgrid(np.arange(grid - 1, dtype=np.int64), np.arange(grid - 1, dtype=np.int64), indexing="ij")
    v00 = (j * grid + i).reshape(-1)
    v10, v01, v11 = v00 + 1, v00 + grid, v00 + grid + 1
    # wound so the face normal points +y, like the vertex normals
    idx = np.empty((2 * v00.size, 3), np.int32)
    idx[0::2] = np.stack([v00, v01, v10], -1)
    idx[1::2] = np.stack([v10, v01, v11], -1)
    f32 = lambda a: np.ascontiguousarray(a.reshape(-1, a.shape[-1]), np.float32)
    return (f32(P), idx, f32(N), f32(np.stack([U, V], -1)), f32(T), f32(B))


def displaced_plane(grid=4096, width=3840, height=2160, alpha=0.2, f0=(0.95, 0.9, 0.85), base="metal",
                    microfacet=True, normal_map_size=1024, sky_size=(2048, 1024), variant=0, seed=5):
    """The normal-mapping workload: a displaced plane (grid^2 vertices) whose
    material is MicrofacetNormalMapping over GGX (Metal, alpha) with the
    32-sine normal map, under a procedural lat-long HDR sky with a sun of
    radiance 2e4.  microfacet=False keeps the base material with the same
    normal map (the sandbox's naive mode, Main.cpp:411-413).  `variant`
    picks one of 16 camera placements around the plane."""
    s = SceneDesc(name=f"displaced_plane_{grid}")
    pos, idx, nrm, uv, tan, bit = displaced_plane_mesh(grid, seed=seed)
    s.textures.append(sine_normal_map(normal_map_size, seed=seed))
    if base == "metal":
        b = Material(MAT_METAL, (float(alpha), f0[0], f0[1], f0[2]))
    elif base == "diffuse":
        b = diffuse(*f0)
    else:
        b = ggx_lambert(f0, float(alpha), 0.04)
    m = microfacet_nmap(b, normal_map=0) if microfacet else Material(b.type, b.params, normal_map=0)
    mat = s.add_material(m)
    s.objects.append(ObjectDesc(pos, idx, material=mat, normals=nrm, texcoords=uv, tangents=tan, bitangents=bit))
    sw, sh = sky_size
    sky = procedural_sky_latlong(sw, sh, sun_dir=(0.3, 0.5, 0.8), sun_half_angle_deg=0.5, sun_radiance=2e4, seed=seed)
    s.env = Env(True, (0.0, 0.0, 0.0), map_kind=ENV_MAP_LATLONG, map=sky)
    s.camera = displaced_plane_camera(variant, width, height)
    return s


def displaced_plane_camera(variant, width, height):
    """Camera `variant` (0..15) of the displaced-plane workload: eight
    azimuths at two elevations, looking at the plane's centre."""
    az = 2 * np.pi * (variant % 8) / 8
    el = 0.55 if variant < 8 else 0.9
    r = 2.6
    eye = (r * np.sin(az) * np.cos(el), r * np.sin(el), -r * np.cos(az) * np.cos(el))
    return Camera(look_at(eye, (0.0, 0.0, 0.0)), math.radians(45.0), height / width)
