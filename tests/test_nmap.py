"""Pins for the oracle's microfacet-based normal mapping and normal-map
application (Sandbox/MicrofacetBasedNormalMapping/Main.cpp:86-340,
Object.h:59-68), and the CPU side of the displaced-plane workload.

What fixes each part without the reference running:
* the two-facet shadowing G1 and the facet choice lambda_p have closed forms
  (Schuessler et al. 2017): a_p(w) + a_t(w) >= w.wg with equality when both
  facets face w, so G1 = 1 there and (w.wg) cos(theta_p) / (w.wp) when the
  tangent facet is back-facing;
* with an unperturbed normal (wg == wp) the material is its base, exactly;
* the sampling walk and the evaluation walk estimate the same BSDF: the mean
  sample weight equals the integral of bsdfCos over the sphere and so does
  the first moment (a wrong G1 argument, a missing factor or a dropped walk
  step on one side breaks it);
* the model conserves energy: over a white Lambertian base the integral of
  bsdfCos is 1 for every wo;
* the normal map turns a constant texel c into shading.n =
  normalize((2c-1).b n + (2c-1).r t + (2c-1).g b), with n face-forwarded and
  t, b not, and the lookup is Texture::rgb's bilinear filter.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import rays_from, unit
from xitils_amd import scenes

NMAP = scenes.MAT_MICROFACET_NMAP


@pytest.fixture(scope="module", autouse=True)
def _built():
    po.build()


def _frame(rng, theta):
    """(wg, wp, wt) with wp tilted by theta from wg toward u, wt = -u."""
    wg = unit(rng.normal(size=3)).astype(np.float64)
    a = unit(rng.normal(size=3)).astype(np.float64)
    u = a - wg * (a @ wg)
    u /= np.linalg.norm(u)
    wp = math.cos(theta) * wg + math.sin(theta) * u
    return wg, wp, -u


def test_two_facet_shadowing_closed_forms():
    rng = np.random.default_rng(11)
    checked = {"both": 0, "p_only": 0, "t_only": 0}
    for trial in range(400):
        th = rng.uniform(0.05, 1.2)
        wg, wp, wt = _frame(rng, th)
        w = unit(rng.normal(size=3)).astype(np.float64)
        if w @ wg <= 0.05:
            w = w - 2 * (w @ wg) * wg + 0.1 * wg
            w /= np.linalg.norm(w)
        f = lambda v: np.asarray(v, np.float32)
        g_p, lam = po.nmap_terms(f(w), f(wp), f(wp), f(wt), f(wg))
        g_t, _ = po.nmap_terms(f(w), f(wt), f(wp), f(wt), f(wg))
        ap = max(w @ wp, 0) / math.cos(th)
        at = max(w @ wt, 0) * math.tan(th)
        assert lam == pytest.approx(ap / (ap + at), rel=2e-5, abs=2e-6)
        if w @ wp > 0 and w @ wt > 0:
            checked["both"] += 1
            assert g_p == pytest.approx(1.0, abs=2e-5) and g_t == pytest.approx(1.0, abs=2e-5)
            assert ap + at == pytest.approx(w @ wg, rel=1e-9)          # projected area conserved
        elif w @ wp > 0:
            checked["p_only"] += 1
            assert g_t == 0.0
            assert g_p == pytest.approx(min(1.0, (w @ wg) * math.cos(th) / (w @ wp)), rel=2e-5, abs=2e-6)
        elif w @ wt > 0:
            checked["t_only"] += 1
            assert g_p == 0.0
            assert g_t == pytest.approx(min(1.0, (w @ wg) / at), rel=2e-5, abs=2e-6)
    assert min(checked.values()) > 10
    # seen along the geometric normal only the perturbed facet is visible
    wg, wp, wt = _frame(rng, 0.5)
    f = lambda v: np.asarray(v, np.float32)
    assert po.nmap_terms(f(wg), f(wp), f(wp), f(wt), f(wg))[1] == pytest.approx(1.0, abs=1e-6)


def _isect(n, wo, sn):
    return np.concatenate([[0, 0, 0], n, wo, sn]).astype(np.float32)


BASES = [
    (scenes.MAT_METAL, (0.3, 0.9, 0.7, 0.5)),
    (scenes.MAT_DIFFUSE, (0.8, 0.6, 0.4)),
    (scenes.MAT_GGX_LAMBERT, (0.5, 0.6, 0.7, 0.25, 0.04, 0.5)),
]


def _nmap_p7(btype, bp):
    p = np.zeros(7, np.float32)
    p[0] = btype
    p[1:1 + len(bp)] = bp
    return p


def _base_p7(bp):
    p = np.zeros(7, np.float32)
    p[:len(bp)] = bp
    return p


@pytest.mark.parametrize("btype,bp", BASES)
def test_unperturbed_normal_is_the_base_material(btype, bp):
    rng = np.random.default_rng(3)
    for k in range(50):
        n = unit(rng.normal(size=3))
        wo = unit(rng.normal(size=3))
        sn = n if n @ wo > 0 else -n          # face-forwarded, as an unmapped hit has it
        wi = unit(rng.normal(size=3))
        rng6 = (7, 0, k, 2, 3, 64)
        e = po.bsdf_eval_walk(NMAP, _nmap_p7(btype, bp), _isect(n, wo, sn), wi, rng6)
        b = po.bsdf_eval(btype, _base_p7(bp), np.concatenate([[0, 0, 0], n, wo]).astype(np.float32), wi)
        np.testing.assert_array_equal(e, b)
        s = po.bsdf_sample_walk(NMAP, _nmap_p7(btype, bp), _isect(n, wo, sn), rng6)
        u3 = [po.uniform(7, 0, k, 2, 3, 68 + j) for j in range(3)]
        sb = po.bsdf_sample(btype, _base_p7(bp), np.concatenate([[0, 0, 0], n, wo]).astype(np.float32), u3)
        np.testing.assert_array_equal(s, sb)


def _perturbed_isect(theta, wo_angle, azim=0.7):
    n = np.array([0.0, 0.0, 1.0])
    sn = np.array([math.sin(theta), 0.0, math.cos(theta)])
    wo = np.array([math.sin(wo_angle) * math.cos(azim), math.sin(wo_angle) * math.sin(azim), math.cos(wo_angle)])
    return _isect(n, wo, sn)


@pytest.mark.parametrize("btype,bp", BASES)
@pytest.mark.parametrize("theta,wo_angle", [(0.35, 0.4), (0.8, 1.1), (1.1, 0.2)])
def test_sampling_walk_and_evaluation_walk_agree(btype, bp, theta, wo_angle):
    isect = _perturbed_isect(theta, wo_angle)
    p7 = _nmap_p7(btype, bp)
    n = 1 << 21
    ms, fs, vs, delta = po.bsdf_moments(NMAP, p7, isect, n, 0, seed=5)
    me, fe, ve, _ = po.bsdf_moments(NMAP, p7, isect, n, 1, seed=6)
    assert delta == 0.0
    se = math.sqrt(vs / n + ve / n)
    assert abs(ms.mean() - me.mean()) < 5 * se + 1e-4, (ms, me, se)
    np.testing.assert_allclose(ms, me, atol=6 * se + 2e-4, rtol=0.02)
    np.testing.assert_allclose(fs, fe, atol=8 * se + 2e-4)
    # and no energy is created
    assert ms.max() <= 1.0 + 1e-6


@pytest.mark.parametrize("theta,wo_angle", [(0.3, 0.2), (0.7, 0.9), (1.0, 1.3)])
def test_white_lambert_base_conserves_energy(theta, wo_angle):
    isect = _perturbed_isect(theta, wo_angle)
    p7 = _nmap_p7(scenes.MAT_DIFFUSE, (1.0, 1.0, 1.0))
    n = 1 << 21
    ms, _, _, _ = po.bsdf_moments(NMAP, p7, isect, n, 0, seed=9)
    # every walk throughput is 1 unless the 100-step cap drops it
    assert 0.999 <= ms.mean() <= 1.0
    me, _, ve, _ = po.bsdf_moments(NMAP, p7, isect, n, 1, seed=10)
    assert me.mean() == pytest.approx(1.0, abs=5 * math.sqrt(ve / n) + 1e-3)


def test_pdf_is_the_documented_approximation():
    """getPDF integrates to about 1 over the sphere and is positive where
    the walk's exit directions land."""
    isect = _perturbed_isect(0.6, 0.7)
    p7 = _nmap_p7(scenes.MAT_METAL, (0.3, 0.9, 0.7, 0.5))
    rng = np.random.default_rng(1)
    d = unit(rng.normal(size=(20000, 3)))
    pdf = np.array([po.bsdf_eval_walk(NMAP, p7, isect, w, (1, 0, 0, 0, 1, 1024))[3] for w in d], np.float64)
    integral = pdf.mean() * 4 * math.pi
    assert 0.5 < integral < 1.5          # an approximation (Main.cpp:245), not a density
    for k in range(200):
        s = po.bsdf_sample_walk(NMAP, p7, isect, (2, 0, k, 0, 1, 64))
        if np.any(s[4:7] != 0):
            assert s[3] > 0


def _quad_scene(texel, sn_flip=False, tangents=True, texture=None, uvs=None):
    s = scenes.SceneDesc(name="nmap_quad")
    tex = np.array(texel, np.float32).reshape(1, 1, 3) if texture is None else texture
    s.textures.append(tex)
    m = s.add_material(scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5), normal_map=0))
    pos = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32) if uvs is None else uvs
    tan = np.tile(np.array([1, 0, 0], np.float32), (4, 1)) if tangents else None
    bit = np.tile(np.array([0, 0, 1], np.float32), (4, 1)) if tangents else None
    s.objects.append(scenes.ObjectDesc(pos, idx, material=m, normals=nrm, texcoords=uv, tangents=tan, bitangents=bit))
    return s


def test_normal_map_application_constant_texel():
    c = np.array([0.8, 0.35, 0.9], np.float32)
    sc = po.OracleScene(_quad_scene(c))
    rays = np.concatenate([rays_from([[0.2, 1, 0.1]], [[0, -1, 0]]), rays_from([[0.2, -1, 0.1]], [[0, 1, 0]])])
    sn = sc.shading_normals(rays)
    m = c.astype(np.float64) * 2 - 1
    t, b = np.array([1.0, 0, 0]), np.array([0, 0, 1.0])
    for k, nff in enumerate([np.array([0, 1.0, 0]), np.array([0, -1.0, 0])]):
        want = m[2] * nff + m[0] * t + m[1] * b
        want /= np.linalg.norm(want)
        np.testing.assert_allclose(sn[k], want, atol=2e-6)
    # without a tangent frame only the blue channel scales the normal
    sc = po.OracleScene(_quad_scene(c, tangents=False))
    np.testing.assert_allclose(sc.shading_normals(rays[:1])[0], [0, 1, 0], atol=1e-7)


def test_normal_map_bilinear_lookup():
    """A 2x1 texture read at u between the texel centres blends the two
    (Texture.h:83-103), clamped outside."""
    tex = np.zeros((1, 2, 3), np.float32)
    tex[0, 0] = (0.5, 0.5, 1.0)        # unperturbed
    tex[0, 1] = (1.0, 0.5, 0.5)        # tilted toward +t
    sc = po.OracleScene(_quad_scene(None, texture=tex))
    # hit points along x map to u = (x + 1) / 2
    xs = np.array([-0.9, -0.2, 0.0, 0.3, 0.9], np.float32)
    rays = rays_from(np.stack([xs, np.ones_like(xs), np.full_like(xs, 0.1)], -1), [[0, -1, 0]])
    sn = sc.shading_normals(rays)
    for x, got in zip(xs, sn):
        u = (x + 1) / 2
        w1 = min(max(u * 2 - 0.5, 0.0), 1.0)
        c = (1 - w1) * tex[0, 0] + w1 * tex[0, 1]
        m = c * 2 - 1
        want = m[2] * np.array([0, 1, 0]) + m[0] * np.array([1, 0, 0]) + m[1] * np.array([0, 0, 1])
        np.testing.assert_allclose(got, want / np.linalg.norm(want), atol=3e-6)


def test_displaced_plane_geometry_and_map():
    pos, idx, nrm, uv, tan, bit = scenes.displaced_plane_mesh(grid=33)
    assert pos.shape == (33 * 33, 3) and idx.shape == (2 * 32 * 32, 3)
    fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
    assert (fn[:, 1] > 0).all() and (nrm[:, 1] > 0).all()
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)
    # the tangent frame is orthogonal to the normal (dP/du, dP/dv of a height field)
    assert np.abs(np.sum(nrm * tan, 1)).max() < 1e-6 and np.abs(np.sum(nrm * bit, 1)).max() < 1e-6
    m = scenes.sine_normal_map(64)
    n = m * 2 - 1
    np.testing.assert_allclose(np.linalg.norm(n, axis=-1), 1, atol=1e-6)
    assert (n[..., 2] > 0).all()
