"""Pins for the CPU oracle (oracle/xo_core.hpp).

The reference ships no tests, golden vectors or fixtures (it has no test
directory; its only artefacts are 8-bit screenshots), so every function of the
oracle is pinned here against what the mathematics fixes: published
known-answer vectors, closed forms, invariants, brute force, and statistical
tests chosen so a plausible slip (a dropped term, a wrong sign, a swapped
operand, a transposed index) fails at least one of them.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import rays_from, random_dirs, unit, camera_rays, scene_rays, hit_prims
from xitils_amd import scenes

pytestmark = []


@pytest.fixture(scope="module", autouse=True)
def _built():
    po.build()


# ------------------------------------------------------------------ RNG (R24)

# Known-answer vectors of Philox4x32-10 published with Random123
# (kat_vectors: counter, key -> output).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(x) for x in po.philox(ctr, key)) == want


def test_uniform_range_and_chi2():
    u = np.array([po.uniform(1, 0, p, s, 1, d) for p in range(64) for s in range(8) for d in range(16)])
    assert u.min() >= 0.0 and u.max() < 1.0
    hist, _ = np.histogram(u, bins=16, range=(0, 1))
    exp = len(u) / 16
    chi2 = float(((hist - exp) ** 2 / exp).sum())
    assert chi2 < 40.0   # 15 dof, p ~ 5e-4
    # dimensions are distinct streams
    a = np.array([po.uniform(1, 0, 7, 3, 2, d) for d in range(16)])
    assert len(set(a.tolist())) == 16


# ------------------------------------------------------------------ triangle test (R5)

def test_triangle_known_answer():
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    ok, out = po.tri_test(p, rays_from([0.25, 0.25, -2.0], [0, 0, 1])[0])
    assert ok
    t, b0, b1 = out
    assert t == 2.0
    # hit (0.25, 0.25) = b0*p0 + b1*p1 + (1-b0-b1)*p2  -> b1 = 0.25, b2 = 0.25, b0 = 0.5
    assert abs(b0 - 0.5) < 1e-7 and abs(b1 - 0.25) < 1e-7
    # outside the triangle
    assert not po.tri_test(p, rays_from([0.75, 0.75, -2.0], [0, 0, 1])[0])[0]
    # behind the origin: t <= 0 rejected
    assert not po.tri_test(p, rays_from([0.25, 0.25, 1.0], [0, 0, 1])[0])[0]
    # tMax bound is inclusive (TriangleIndexed.h:123 accepts t == tMax)
    assert po.tri_test(p, rays_from([0.25, 0.25, -2.0], [0, 0, 1], tmax=2.0)[0])[0]
    assert not po.tri_test(p, rays_from([0.25, 0.25, -2.0], [0, 0, 1], tmax=1.999)[0])[0]
    # back face is hit too (no culling)
    ok, out = po.tri_test(p, rays_from([0.25, 0.25, 2.0], [0, 0, -1])[0])
    assert ok and out[0] == 2.0


def test_triangle_random_against_double_precision():
    rng = np.random.default_rng(5)
    n_ok = 0
    for _ in range(2000):
        p = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, size=3).astype(np.float32)
        # aim at a point of the triangle's plane, inside or outside it
        w = rng.uniform(-0.3, 1.0, size=3)
        w /= w.sum()
        d = unit(w @ p.astype(np.float64) - o)
        ok, out = po.tri_test(p, rays_from(o, d)[0])
        # Moller-Trumbore in float64 as the independent answer
        P = p.astype(np.float64)
        e1, e2 = P[1] - P[0], P[2] - P[0]
        h = np.cross(d.astype(np.float64), e2)
        a = e1 @ h
        nrm = np.cross(e1, e2)
        if abs(a) < 1e-9 or abs(nrm @ d) < 0.05 * np.linalg.norm(nrm):
            continue   # grazing: float32 conditioning, not the formula, dominates
        s = o - P[0]
        u = (s @ h) / a
        q = np.cross(s, e1)
        v = (d.astype(np.float64) @ q) / a
        t = (e2 @ q) / a
        inside = u > 1e-5 and v > 1e-5 and u + v < 1 - 1e-5 and t > 1e-5
        outside = u < -1e-5 or v < -1e-5 or u + v > 1 + 1e-5 or t < -1e-5
        if inside:
            assert ok
            assert abs(out[0] - t) <= 2e-5 * max(1.0, abs(t))
            # b0 weights p0: barycentric of p0 is 1 - u - v
            assert abs(out[1] - (1 - u - v)) < 1e-4 and abs(out[2] - u) < 1e-4
            n_ok += 1
        elif outside:
            assert not ok
    assert n_ok > 500


def test_watertight_shared_edge():
    # two triangles sharing the diagonal of the unit square; rays through the
    # diagonal never miss both
    sq = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    A, B = sq[[0, 1, 2]], sq[[0, 2, 3]]
    rng = np.random.default_rng(3)
    for s in np.linspace(0.01, 0.99, 199).astype(np.float32):
        o = np.array([s, s, -1.0], np.float32) + rng.normal(scale=1e-3, size=3).astype(np.float32) * [1, 1, 0]
        target = np.array([s, s, 0.0], np.float32)
        d = unit(target - o)
        r = rays_from(o, d)[0]
        assert po.tri_test(A, r)[0] or po.tri_test(B, r)[0]


# ------------------------------------------------------------------ BVH (R2-R4, R6-R8)

def _mixed_scene():
    s = scenes.material_zoo(32, 32)
    # one transformed instance (rotation + translation + uniform scale)
    v, i = scenes.box_mesh((1, 1, 1), (0, 0, 0), 0)
    th = math.radians(30)
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] = 0.2 * np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    M[:3, 3] = (0.5, 0.7, 0.2)
    Mi = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    s.objects.append(scenes.ObjectDesc(v, i, material=0, xf=M, xf_inv=Mi))
    return s


@pytest.mark.parametrize("which", ["cornell", "zoo", "heightfield"])
def test_reference_bvh_equals_brute_force(which):
    rng = np.random.default_rng(11)
    if which == "cornell":
        desc = scenes.cornell()
    elif which == "zoo":
        desc = _mixed_scene()
    else:
        desc = scenes.heightfield(cells=24, materials=False)
    o = po.OracleScene(desc)
    rays = np.concatenate([scene_rays(rng, desc, 3000, 0.02, 0.98), camera_rays(desc, 48, 48)])
    a = o.trace_closest(rays, mode=0)
    b = o.trace_closest(rays, mode=1)
    pa, pb = hit_prims(a), hit_prims(b)
    assert (pa != 0xFFFFFFFF).mean() > 0.3
    if which == "zoo":
        # the transformed instance is tested in object space by the reference
        # (Object.h:107) and in world space by brute force: ids agree, t to 1e-5
        same = pa == pb
        assert same.mean() > 0.999
        both = (pa != 0xFFFFFFFF) & same
        assert np.all(np.abs(a[both, 0] - b[both, 0]) <= 1e-5 * np.maximum(1, b[both, 0]))
    else:
        assert np.array_equal(pa, pb)
        assert np.array_equal(a[:, 0], b[:, 0])
    occ_a = o.trace_any(rays[:1000], mode=0)
    occ_b = o.trace_any(rays[:1000], mode=1)
    assert np.array_equal(occ_a, occ_b)


def test_any_hit_matches_closest_within_tmax():
    desc = scenes.cornell()
    o = po.OracleScene(desc)
    rng = np.random.default_rng(2)
    rays = scene_rays(rng, desc, 2000, 0.05, 0.95)
    h = o.trace_closest(rays, mode=1)
    t = h[:, 0]
    finite = np.isfinite(t)
    r2 = rays.copy()
    r2[finite, 3] = t[finite] * np.float32(0.5)     # strictly before the closest hit
    assert not o.trace_any(r2[finite], mode=1).any()
    r2[finite, 3] = t[finite] * np.float32(1 + 2e-6)   # just past it reaches it
    assert o.trace_any(r2[finite], mode=1).all()
    # at exactly t the scaled comparison (tScaled vs tMax * det,
    # TriangleIndexed.h:120-125) accepts almost always; rounding of t = tScaled / det
    # can put t one ulp above tMax * det / det
    r2[finite, 3] = t[finite]
    assert o.trace_any(r2[finite], mode=1).mean() > 0.9


def test_reference_bvh_structure():
    desc = scenes.heightfield(cells=16, materials=False)
    o = po.OracleScene(desc)
    nodes, depth = o.ref_bvh_info(0)
    n = desc.num_triangles
    assert nodes == 2 * n - 1          # binary tree, one primitive per leaf (AccelerationStructure.h:191-195)
    assert depth >= math.ceil(math.log2(n))


# ------------------------------------------------------------------ BSDFs (R12-R18)

N_UP = np.array([0, 0, 1], np.float32)


def _isect(wo, n=N_UP):
    return np.concatenate([[0, 0, 0], n, wo]).astype(np.float32)


def test_lambert_closed_forms():
    rng = np.random.default_rng(1)
    alb = (0.6, 0.3, 0.9)
    p7 = [*alb, 0, 0, 0, 0]
    wo = unit([0.3, 0.1, 0.8])
    cos_list = []
    for k in range(4000):
        u = rng.random(3).astype(np.float32)
        out = po.bsdf_sample(scenes.MAT_DIFFUSE, p7, _isect(wo), u)
        wi, pdf, val = out[:3], out[3], out[4:]
        c = float(wi @ N_UP)
        assert np.allclose(val, alb)                         # weight = albedo (Material.h:71-75)
        assert abs(pdf - c / math.pi) < 1e-6
        ev = po.bsdf_eval(scenes.MAT_DIFFUSE, p7, _isect(wo), wi)
        assert np.allclose(ev[:3], np.array(alb) / math.pi * c, rtol=1e-5)
        assert abs(ev[3] - c / math.pi) < 1e-6
        cos_list.append(c)
    # cos(theta) of a cosine-weighted hemisphere: cos^2 ~ U[0, 1]
    c2 = np.array(cos_list) ** 2
    hist, _ = np.histogram(c2, bins=10, range=(0, 1))
    chi2 = float(((hist - 400) ** 2 / 400).sum())
    assert chi2 < 30
    # below the horizon the BSDF is zero (clampPositive, Material.h:61)
    assert np.all(po.bsdf_eval(scenes.MAT_DIFFUSE, p7, _isect(wo), unit([0.2, 0, -1]))[:3] == 0)


def _sphere_grid(n_theta=256, n_phi=512):
    th = (np.arange(n_theta) + 0.5) * math.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * math.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (math.pi / n_theta) * (2 * math.pi / n_phi)).reshape(-1)
    return d.astype(np.float32), dw


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.6])
def test_ggx_D_normalised(alpha):
    # integral of D(h) (n.h) over the hemisphere = 1 (projected-area normalisation)
    d, dw = _sphere_grid(512, 8)
    up = d[:, 2] > 0
    D = np.array([po.ggx_D(alpha, h, N_UP) for h in d[up]])
    integral = float(np.sum(D * d[up, 2] * dw[up]))
    assert abs(integral - 1.0) < 2e-2
    # lower hemisphere is zero (Heaviside, Material.h:301)
    assert po.ggx_D(alpha, unit([0.1, 0.2, -0.9]), N_UP) == 0.0


def test_ggx_G1_closed_form():
    for alpha in (0.05, 0.3, 0.8):
        for th in (0.1, 0.6, 1.2):
            x = np.array([math.sin(th), 0, math.cos(th)], np.float32)
            lam = (-1 + math.sqrt(1 + alpha ** 2 * math.tan(th) ** 2)) / 2
            assert abs(po.ggx_G1(alpha, x, N_UP) - 1 / (1 + lam)) < 1e-5
        assert abs(po.ggx_G1(alpha, N_UP, N_UP) - 1.0) < 1e-7


@pytest.mark.parametrize("alpha,wo", [(0.2, [0.3, 0.1, 0.9]), (0.5, [0.8, 0.0, 0.3])])
def test_metal_vndf_weight_and_pdf(alpha, wo):
    # With F0 = 1 Schlick is 1, and VNDF sampling gives f cos / pdf = G1(wi)
    # (the G2 / G1(wo) identity of separable Smith masking).
    wo = unit(wo)
    p7 = [alpha, 1, 1, 1, 0, 0, 0]
    rng = np.random.default_rng(4)
    for _ in range(500):
        out = po.bsdf_sample(scenes.MAT_METAL, p7, _isect(wo), rng.random(3).astype(np.float32))
        wi, pdf, val = out[:3], out[3], out[4:]
        if not pdf > 0:
            continue
        g1 = po.ggx_G1(alpha, wi, N_UP)
        assert abs(val[0] - g1) < 2e-4 * max(1, g1)
        ev = po.bsdf_eval(scenes.MAT_METAL, p7, _isect(wo), wi)
        assert abs(ev[3] - pdf) <= 1e-5 * pdf
    # VNDF pdf integrates to 1 over the sphere of wi
    d, dw = _sphere_grid(384, 384)
    pdfs = np.array([po.bsdf_eval(scenes.MAT_METAL, p7, _isect(wo), x)[3] for x in d])
    pdfs = np.where(np.isfinite(pdfs), pdfs, 0)
    assert abs(float(np.sum(pdfs * dw)) - 1.0) < 3e-2


def test_schlick_endpoints():
    # F(wo.h = 1) = f0 ; F(wo.h = 0) = 1   (Material.h:288-292)
    wo = N_UP
    p7 = [0.3, 0.25, 0.5, 0.75, 0, 0, 0]
    # wi = wo -> h = n -> wo.h = 1
    ev = po.bsdf_eval(scenes.MAT_METAL, p7, _isect(wo), wo)
    D = po.ggx_D(0.3, N_UP, N_UP)
    G = po.ggx_G1(0.3, wo, N_UP) ** 2
    assert np.allclose(ev[:3], np.array([0.25, 0.5, 0.75]) * D * G / 4, rtol=1e-5)


def test_ggx_lambert_mixture():
    alpha, f0, ps = 0.25, 0.04, 0.5
    alb = (0.5, 0.4, 0.3)
    p7 = [*alb, alpha, f0, ps, 0]
    wo = unit([0.4, 0.2, 0.7])
    d, dw = _sphere_grid(256, 256)
    ev = np.array([po.bsdf_eval(scenes.MAT_GGX_LAMBERT, p7, _isect(wo), x) for x in d])
    # pdf normalised over the upper hemisphere (mixture of two normalised
    # pdfs, with the GGX lobe's below-horizon mass cut)
    total = float(np.sum(np.where(np.isfinite(ev[:, 3]), ev[:, 3], 0) * dw))
    assert 0.9 < total <= 1.0 + 2e-2
    # energy: directional albedo below 1 (white furnace bound)
    alb_dir = np.sum(ev[:, :3] * dw[:, None], axis=0)
    assert np.all(alb_dir < 1.0)
    # diffuse term alone below the GGX lobe: at grazing-ish wi far from the
    # mirror direction the value is (1 - f0) a / pi cos
    wi = unit([-0.7, -0.3, 0.2])
    e = po.bsdf_eval(scenes.MAT_GGX_LAMBERT, p7, _isect(wo), wi)
    c = float(wi @ N_UP)
    assert np.all(e[:3] >= np.array(alb) * (1 - f0) / math.pi * c * (1 - 1e-6))
    # sampled weight = eval / pdf
    rng = np.random.default_rng(8)
    for _ in range(300):
        out = po.bsdf_sample(scenes.MAT_GGX_LAMBERT, p7, _isect(wo), rng.random(3).astype(np.float32))
        wi, pdf, val = out[:3], out[3], out[4:]
        if pdf <= 0 or not np.any(val):
            continue
        e = po.bsdf_eval(scenes.MAT_GGX_LAMBERT, p7, _isect(wo), wi)
        assert np.allclose(val, e[:3] / e[3], rtol=1e-5)


def test_specular_reflection_and_refraction():
    wo = unit([0.3, -0.2, 0.8])
    out = po.bsdf_sample(scenes.MAT_SPEC_REFL, [0] * 7, _isect(wo), [0.5, 0.5, 0.5])
    wi = out[:3]
    assert out[3] == -1.0 and np.allclose(out[4:], 1)
    assert np.allclose(wi, [-0.3 / np.linalg.norm([0.3, -0.2, 0.8]) * 1, 0.2 / np.linalg.norm([0.3, -0.2, 0.8]),
                            0.8 / np.linalg.norm([0.3, -0.2, 0.8])], atol=1e-6)
    # refraction entering index 1.5: Snell sin_t = sin_i / 1.5, transmitted below the surface
    out = po.bsdf_sample(scenes.MAT_SPEC_REFR, [1.5, 0, 0, 0, 0, 0, 0], _isect(wo), [0.5, 0.5, 0.5])
    wt = out[:3]
    sin_i = math.sqrt(1 - float(wo @ N_UP) ** 2)
    sin_t = math.sqrt(max(0.0, 1 - float(wt @ N_UP) ** 2 / float(wt @ wt)))
    assert wt[2] < 0 and abs(sin_t - sin_i / 1.5) < 1e-5
    # Fresnel at normal incidence from air to n = 1.5 reflects ((n-1)/(n+1))^2 = 0.04
    n_inc = N_UP
    refl = sum(1 for u in np.linspace(0, 0.999, 1000)
               if po.bsdf_sample(scenes.MAT_SPEC_FRESNEL, [1.5, 0, 0, 0, 0, 0, 0], _isect(n_inc),
                                 [u, 0.5, 0.5])[2] > 0)
    assert abs(refl / 1000 - 0.04) < 2e-3
    # Material.h:265-271 evaluates g^2 = (eta_o / eta_i)^2 - 1 + c^2 with eta_o
    # on the wo side: outside the glass (eta_o = 1, eta_i = 1.5) grazing
    # incidence with c^2 < 1 - 1/1.5^2 gives g^2 <= 0 and the lobe always
    # reflects (the reference inverts the usual ratio; kept as written).
    wo_graze = unit([0.9, 0.0, 0.2])
    for u in (0.0, 0.5, 0.9999):
        r = po.bsdf_sample(scenes.MAT_SPEC_FRESNEL, [1.5, 0, 0, 0, 0, 0, 0], _isect(wo_graze), [u, 0.5, 0.5])
        assert r[2] > 0
    # inside (geometric normal away from wo) the same angle transmits sometimes
    isect = np.concatenate([[0, 0, 0], -N_UP, wo_graze]).astype(np.float32)
    dirs = [po.bsdf_sample(scenes.MAT_SPEC_FRESNEL, [1.5, 0, 0, 0, 0, 0, 0], isect, [u, 0.5, 0.5])[2]
            for u in np.linspace(0, 0.999, 200)]
    assert any(z > 0 for z in dirs) and any(z <= 0 for z in dirs)


# ------------------------------------------------------------------ lights (R19-R22)

def test_light_pdf_is_uniform_area_for_identity_objects():
    desc = scenes.cornell()
    o = po.OracleScene(desc)
    lv = desc.objects[3].positions
    area = 0.25 * 0.2
    for k in range(50):
        out = o.sample_light(1, k, 0, 1)
        p, n, pdf = out[:3], out[3:6], out[9]
        assert abs(pdf - 1 / area) < 1e-3 * (1 / area)
        assert abs(p[1] - 0.999) < 1e-6 and 0.375 <= p[0] <= 0.625 and 0.4 <= p[2] <= 0.6
        assert np.allclose(n, [0, -1, 0], atol=1e-6)


def test_faithful_vs_uniform_barycentrics():
    # reference barycentrics t_i/sum are not uniform on the triangle: the
    # density of b0 piles up around 1/3; the uniform mode is uniform in area
    desc = scenes.SceneDesc()
    m = desc.add_material(scenes.emission(1, 1, 1))
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    desc.objects.append(scenes.ObjectDesc(tri, np.array([[0, 1, 2]], np.int32), material=m))
    desc.camera = scenes.cornell().camera
    o = po.OracleScene(desc)
    def draw(ub):
        return np.array([o.sample_light(3, k, 0, 1, uniform_bary=ub)[:2] for k in range(6000)])
    pu = draw(True)
    pf = draw(False)
    # uniform: P(x < 0.5) over the triangle = 1 - 0.25 = 0.75
    assert abs((pu[:, 0] < 0.5).mean() - 0.75) < 0.02
    # faithful: x = t1/(t0+t1+t2) with i.i.d. uniforms; its mean is 1/3 and
    # P(x < 0.5) ~ 0.80 (Monte Carlo of the closed form below)
    rng = np.random.default_rng(0)
    t = rng.random((400000, 3))
    x = t[:, 1] / t.sum(1)
    assert abs((pf[:, 0] < 0.5).mean() - (x < 0.5).mean()) < 0.02
    assert abs(pf[:, 0].mean() - 1 / 3) < 0.01


# ------------------------------------------------------------------ estimator (R10, R11)

def test_furnace_convex_object_under_uniform_sky():
    a = 0.6
    desc = scenes.furnace_object(a)
    o = po.OracleScene(desc)
    for integ in (0, 1):
        hits = 0
        for k in range(3):
            v = o.render(32, 32, 1, seed=3, integrator=integ, spp_offset=k)
            # one bounce, then every BSDF ray escapes: each sample is exactly a
            # (hit) or 1 (miss)
            is_a = np.all(np.abs(v - a) <= 2e-6 * a, axis=2)
            is_1 = np.all(v == 1.0, axis=2)
            assert np.all(is_a | is_1)
            hits += is_a.sum()
        assert hits > 0.1 * 3 * 32 * 32


def test_direct_view_of_emitter():
    desc = scenes.cornell(64, 64)
    o = po.OracleScene(desc)
    cam = scenes.Camera(scenes.look_at((0.5, 0.5, 0.5), (0.5, 1.0, 0.5), up=(0, 0, 1)), math.radians(5), 1.0)
    f = o.render(8, 8, 1, seed=1, max_depth=0, camera=cam)
    # looking straight up at the light: Le + reflected light from the room
    assert np.all(f[..., 0] >= 17.0) and np.all(f[..., 1] >= 12.0)


def _irradiance_rect(p, x0, x1, z0, z1, y):
    """Irradiance at point p (normal +y) from a unit-radiance rectangle on the
    plane y, by Lambert's polygon formula E = 1/2 sum beta_i cos(gamma_i)."""
    verts = np.array([[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float64) - p
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    E = 0.0
    for i in range(4):
        a, b = verts[i], verts[(i + 1) % 4]
        g = np.cross(a, b)
        beta = math.acos(np.clip(a @ b, -1, 1))
        E += beta * (g / np.linalg.norm(g))[1]
    return abs(E) / 2.0


def test_nee_mis_matches_polygon_irradiance():
    # floor point under the ceiling light, only the light in the scene (no
    # walls): one bounce of a Lambert floor gives L = a/pi * Le * E
    desc = scenes.SceneDesc()
    a = 0.5
    fl = desc.add_material(scenes.diffuse(a))
    li = desc.add_material(scenes.emission(1, 1, 1))
    v, i = scenes.quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    desc.objects.append(scenes.ObjectDesc(v, i, material=fl))
    v, i = scenes.quad((0.375, 0.999, 0.4), (0.625, 0.999, 0.4), (0.625, 0.999, 0.6), (0.375, 0.999, 0.6))
    desc.objects.append(scenes.ObjectDesc(v, i, material=li))
    cam = scenes.Camera(scenes.look_at((0.3, 0.5, 0.3), (0.3, 0.0, 0.3001)), math.radians(0.01), 1.0)
    desc.camera = cam
    o = po.OracleScene(desc)
    f = o.render(1, 1, 20000, seed=9, max_depth=1, flags=scenes.FLAG_UNIFORM_BARY)
    L = f[0, 0, 0] / 20000
    E = _irradiance_rect(np.array([0.3, 0.0, 0.3]), 0.375, 0.625, 0.4, 0.6, 0.999)
    assert abs(L - a / math.pi * E) < 0.02 * a / math.pi * E
    # BSDF-only sampling converges to the same value (noisier)
    fn = o.render(1, 1, 200000, seed=9, max_depth=1, integrator=1)
    Ln = fn[0, 0, 0] / 200000
    assert abs(Ln - a / math.pi * E) < 0.1 * a / math.pi * E


def test_sun_irradiance_closed_form():
    # Lambert plane under sky 1 + sun disk: L = a (sky) + a/pi * Ls * Omega * cos(theta_sun)
    desc = scenes.SceneDesc()
    a = 0.5
    m = desc.add_material(scenes.diffuse(a))
    v, i = scenes.quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    desc.objects.append(scenes.ObjectDesc(v, i, material=m))
    el = math.radians(30)
    sun = (math.cos(el), math.sin(el), 0.0)
    desc.env = scenes.Env(True, (1, 1, 1), sun, 0.25, (5e4, 5e4, 5e4))
    desc.camera = scenes.Camera(scenes.look_at((0.0, 1.0, 0.0), (0.0, 0.0, 0.0001)), math.radians(0.01), 1.0)
    o = po.OracleScene(desc)
    f = o.render(1, 1, 40000, seed=5, max_depth=1, flags=scenes.FLAG_ENV_NEE)
    L = f[0, 0, 0] / 40000
    omega = 2 * math.pi * (1 - math.cos(math.radians(0.25)))
    want = a + a / math.pi * 5e4 * omega * math.sin(el)
    assert abs(L - want) < 0.01 * want


def test_standard_and_naive_agree_in_the_mean():
    desc = scenes.cornell(16, 16)
    o = po.OracleScene(desc)
    fs = o.render(16, 16, 512, seed=4, max_depth=6)
    fn = o.render(16, 16, 4096, seed=4, max_depth=6, integrator=1)
    ms = fs.mean(axis=(0, 1)) / 512
    mn = fn.mean(axis=(0, 1)) / 4096
    assert np.allclose(ms, mn, rtol=0.05)


def test_max_depth_zero_bounce_is_direct_emission_only():
    desc = scenes.cornell(16, 16)
    o = po.OracleScene(desc)
    f = o.render(16, 16, 2, seed=1, max_depth=-1)   # <= 0: unlimited (roulette)
    assert np.isfinite(f).all() and f.min() >= 0
    g = o.render(16, 16, 2, seed=1, max_depth=1)
    h = o.render(16, 16, 2, seed=1, max_depth=2)
    assert np.all(h >= g - 1e-5)     # more bounces never remove energy at a fixed seed


# ------------------------------------------------------------------ camera (R25, R26)

def test_pinhole_camera_closed_form():
    fov, aspect = math.radians(40), 0.5
    cam = scenes.Camera(scenes.translate(1, 2, 3)[0], fov, aspect)
    c18 = cam.packed()
    r = po.camera_ray(c18, 0.5, 0.5)
    assert np.allclose(r[:3], [1, 2, 3]) and np.allclose(r[4:7], [0, 0, 1])
    r = po.camera_ray(c18, 1.0, 1.0)
    d = r[4:7]
    # x = tan(fov/2), y = tan(fov/2 * aspect) (quirk: not tan(fov/2) * aspect)
    assert abs(d[0] / d[2] - math.tan(fov / 2)) < 1e-6
    assert abs(d[1] / d[2] - math.tan(fov / 2 * aspect)) < 1e-6
    assert abs(np.linalg.norm(d) - 1) < 1e-6


# ------------------------------------------------------------------ film position (R26)

def check_film_position(L=None):
    """RenderTargetTile::GenerateFilmPosition (RenderTarget.h:68-80): the
    jitter span of pixel x is one pixel wide and centred on its corner x / W
    (quirk Q8), neighbouring spans abut so the spans tile the film, and y is
    flipped (row 0 at the top, fy near 1)."""
    W, H = 7, 5
    ulast = float(np.float32(1 - 2.0 ** -24))       # the largest uniform the RNG returns
    xs, ys = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    xs, ys = xs.ravel(), ys.ravel()
    def fp(ux, uy):
        rows = np.stack([xs, ys, np.full(xs.shape, ux), np.full(xs.shape, uy)], 1)
        return po.film_position(rows, W, H, L=L).astype(np.float64)
    mid, lo, hi = fp(0.5, 0.5), fp(0.0, 0.0), fp(ulast, ulast)
    # centre of the span: the pixel corner, y flipped
    assert np.allclose(mid[:, 0], xs / W, atol=1e-7)
    assert np.allclose(mid[:, 1], 1 - ys / H, atol=1e-7)
    # span: one pixel, centred on the corner
    assert np.allclose(lo[:, 0], xs / W - 0.5 / W, atol=1e-6)
    assert np.allclose(hi[:, 0] - lo[:, 0], 1.0 / W, atol=1e-6)
    assert np.allclose(lo[:, 1], 1 - ys / H - 0.5 / H, atol=1e-6)
    assert np.allclose(hi[:, 1] - lo[:, 1], 1.0 / H, atol=1e-6)
    # the spans of a row tile [-1/2W, 1 - 1/2W): pixel x + 1 starts where x ends
    row = ys == 0
    order = np.argsort(xs[row])
    assert np.allclose(lo[row][order][1:, 0], hi[row][order][:-1, 0], atol=1e-6)
    # row 0 is the top of the image: camera rays of pixel (0, 0) point up-left
    cam = scenes.cornell().camera.packed()
    r = po.camera_ray(cam, float(mid[0, 0]), float(mid[0, 1]))
    assert r[4] < 0 and r[5] > 0


def test_film_position():
    check_film_position()


# ------------------------------------------------------------------ roulette weighting (R10, R11)

def check_roulette_passthrough(L=None, integrator=0):
    """Every path through a stack of n pass-through planes (weight 1 each)
    meets one roulette test per vertex beyond pathLength 5 (PathTracer.h:
    149-155): m = n - 4 tests for n >= 5, in both estimators.  A surviving
    path carries exactly 1 / 0.9^m of sky radiance 1, a killed one 0, so each
    sample is 0 or 0.9^-m, the surviving fraction is 0.9^m and the mean is 1
    (the roulette is unbiased).  With a bounce cap below n nothing reaches
    the sky."""
    from pin_scenes import passthrough_stack
    n = 12
    m = n - 4
    desc = passthrough_stack(n, 64, 64)
    o = po.OracleScene(desc, L=L)
    vals = np.concatenate([o.render(64, 64, 1, seed=21, integrator=integrator, spp_offset=k)[..., 0].ravel()
                           for k in range(16)])
    v_m = 0.9 ** -m
    alive = vals != 0
    assert np.all(np.abs(vals[alive] - v_m) <= 1e-5 * v_m)
    p = 0.9 ** m
    sigma = math.sqrt(p * (1 - p) / vals.size)
    assert abs(alive.mean() - p) < 5 * sigma
    assert abs(vals.mean() - 1.0) < 5 * sigma * v_m
    # bounce cap: n planes need max_depth >= n
    assert o.render(16, 16, 4, seed=2, integrator=integrator, max_depth=n - 1).max() == 0.0
    assert o.render(16, 16, 4, seed=2, integrator=integrator, max_depth=n).max() > 0.0


@pytest.mark.parametrize("integrator", [0, 1])
def test_roulette_passthrough_stack(integrator):
    check_roulette_passthrough(integrator=integrator)


# ------------------------------------------------------------------ several lights, instanced emitter (R19-R22)

def check_two_lights(L=None):
    """Direct light at a floor point from two emitters with unequal triangle
    counts (2 and 6, of unequal areas) and areas, one an instance under a
    uniform scale 2, a rotation and a translation: the NEE + BSDF MIS
    estimate (uniform area barycentrics, so unbiased) converges to
    a / pi * sum_k Le_k E_k with E_k Lambert's polygon irradiance.  This pins
    the light-object choice (1 / lights.size(), Scene.h:40-55), the triangle
    choice (1 / ntri, TriangleMesh.h:216-235) and the instance's area scaling
    (Object.h:91-111)."""
    from pin_scenes import two_lights
    desc, q, want = two_lights()
    o = po.OracleScene(desc, L=L)
    spp = 40000
    f = o.render(1, 1, spp, seed=9, max_depth=1, flags=scenes.FLAG_UNIFORM_BARY)
    got = f[0, 0] / spp
    assert np.allclose(got, want, rtol=0.02), (got, want)
    # the sampled light pdf of the instanced strip is 1 / (world area * ntri * nlights)
    areas = []
    for k in range(200):
        out = o.sample_light(5, k, 0, 1, uniform_bary=True)
        p, pdf = out[:3], out[9]
        if abs(p[1] - 1.0) > 1e-4:      # a point of B
            areas.append(1.0 / (pdf * 6 * 2))
    assert len(areas) > 50
    tri_world = []
    from pin_scenes import world_points
    ob = desc.objects[2]
    for t in ob.indices:
        w = world_points(ob, ob.positions[t])
        tri_world.append(np.linalg.norm(np.cross(w[1] - w[0], w[2] - w[0])) / 2)
    for a in areas:
        assert min(abs(a - t) for t in tri_world) <= 1e-4 * a


def test_two_lights_instanced_emitter():
    check_two_lights()
