// Microfacet-based normal mapping (Schuessler, Hanika, Jung, Dachsbacher
// 2017), as the MicrofacetNormalMapping material of
// Sandbox/MicrofacetBasedNormalMapping/Main.cpp:86-340 in its
// SameMaterialExplicit mode (Main.cpp:22): the surface is modelled as two
// facets, the normal-mapped wp and the tangent facet wt, both carrying the
// base material, and light bounces between them in a random walk.
//
// The material's p[0] holds the base type (DIFFUSE, METAL or GGX_LAMBERT),
// p[1..6] the base parameters.  bsdfCos and evalAndSample are stochastic: the
// walk draws its uniforms from a Philox stream of its own per vertex and call
// site (WalkKey + base dim).  Dim `base` picks the first facet; walk step k
// reads Philox block base/4 + 1 + k (lobe select, u1, u2, escape test).
// The walk is kept out of line: it runs only on these materials, and inlined
// into the shading kernel its registers would lower the occupancy of every
// shading pass.
#pragma once

#include "xt_device.h"

namespace xt {

// uniform streams per vertex: BSDF sampling, light-sample evaluation, and the
// evaluation of a guide-sampled direction
enum : int { D_WALK_SAMPLE = 64, D_WALK_EVAL = 1024, D_WALK_EVAL2 = 2048 };

struct WalkKey {
    uint2 key;
    uint32_t pixel, sample, slot;
};
XT_HD uint4 walk_block(const WalkKey& k, int block) {
    return philox(make_uint4(k.pixel, k.sample, k.slot, (uint32_t)block), k.key);
}

// BASE: the base type as a compile-time constant (the walk is instantiated
// per base, so each instance carries only its base's code and registers).
template <uint32_t BASE>
XT_HD xt_material nmap_base(const xt_material& m) {
    xt_material b;
    b.type = BASE;
    for (int k = 0; k < 6; ++k) b.p[k] = m.p[k + 1];
    b.p[6] = 0.0f;
    return b;
}

// Facet frame (Main.cpp:96-117): 0 = wg == wp (the base alone), 1 = wt has a
// NaN, 2 = the walk.  wg = geometric side of wo, wp = shading normal on wg's
// side, wt = the tangent facet, facing away from wp.
XT_HD int nmap_frame(const Surf& s, f3* wg, f3* wp, f3* wt) {
    *wg = s.n;
    *wp = s.sn;
    if (dot(s.wo, *wg) < 0.0f) *wg = *wg * -1.0f;
    if (dot(*wg, *wp) < 0.0f) *wp = *wp * -1.0f;
    if (wg->x == wp->x && wg->y == wp->y && wg->z == wp->z) return 0;   // Vector.h:344-349
    *wt = normalize(cross(normalize(cross(*wg, *wp)), *wg));
    if (__builtin_isnan(wt->x) || __builtin_isnan(wt->y) || __builtin_isnan(wt->z)) return 1;
    if (dot(*wp, *wt) > 0) *wt = *wt * -1.0f;
    return 2;
}
// projected areas and shadowing of the two-facet surface (Main.cpp:323-337);
// min is Utils.h:37 (x < y ? x : y), so a NaN ratio passes through
XT_HD float nmap_ap(f3 w, f3 wp, f3 wg) { return clamp_pos(dot(w, wp)) / fabsf(dot(wp, wg)); }
XT_HD float nmap_at(f3 w, f3 wp, f3 wt, f3 wg) {
    const float d = dot(wp, wg);
    return clamp_pos(dot(w, wt)) * safe_sqrt(1 - d * d) / fabsf(d);   // pow(d, 2): the exact square, rounded once
}
XT_HD float nmap_lambda_p(f3 w, f3 wp, f3 wt, f3 wg) {
    const float ap = nmap_ap(w, wp, wg);
    return ap / (ap + nmap_at(w, wp, wt, wg));
}
XT_HD float nmap_G1(f3 w, f3 wm, f3 wp, f3 wt, f3 wg) {
    if (!(dot(w, wm) > 0.0f)) return 0.0f;
    const float g = clamp_pos(dot(w, wg)) / (nmap_ap(w, wp, wg) + nmap_at(w, wp, wt, wg));
    return 1.0f < g ? 1.0f : g;
}
// the isect the base sees on facet wm (Main.cpp:125-130)
XT_HD Surf nmap_facet(const Surf& s, f3 wm, f3 wr) {
    Surf f = s;
    f.n = wm;
    f.sn = wm;
    f.wo = -wr;
    return f;
}
// The base on a facet is one-sided in both directions, as GlossyClamped, the
// base the reference pairs with this material (Main.cpp:31-84): zero unless
// wo and wi are both above the facet.
XT_HD bool nmap_sides(const Surf& f, f3 wi) { return dot(f.wo, f.sn) > 0.0f && dot(wi, f.sn) > 0.0f; }
XT_HD f3 nmap_facet_eval(const xt_material& base, const Surf& f, f3 wi, float* pdf) {
    if (!nmap_sides(f, wi)) {
        *pdf = 0.0f;
        return splat(0.0f);
    }
    return bsdf_eval(base, f, wi, pdf);
}
XT_HD f3 nmap_facet_sample(const xt_material& base, const Surf& f, f3 wm, uint4 u, f3* wr, float* pdf) {
    const f3 v = bsdf_sample(base, f, make_frame(wm), Rng::to_unit(u.x), Rng::to_unit(u.y), Rng::to_unit(u.z), wr, pdf);
    if (!nmap_sides(f, *wr)) {
        *pdf = 0.0f;
        return splat(0.0f);
    }
    return v;
}

// getPDF (Main.cpp:243-318): the approximation that treats the facets'
// inter-reflection as diffuse.  The tangent-facet terms carry probability_wp
// as the reference writes them (Main.cpp:295-296).
XT_HD float nmap_pdf_frame(const xt_material& base, const Surf& s, f3 wi, f3 wg, f3 wp, f3 wt) {
    const float probability_wp = nmap_lambda_p(s.wo, wp, wt, wg);
    float pdf = 0.0f, diff_pdf = 0.0f, pb;
    const float g_wi_wp = nmap_G1(wi, wp, wp, wt, wg);
    if (probability_wp > 0.0f) {
        nmap_facet_eval(base, nmap_facet(s, wp, -s.wo), wi, &pb);
        pdf += probability_wp * g_wi_wp * pb;
        diff_pdf += probability_wp * (1.0f - g_wi_wp);
    }
    if (probability_wp < 1.0f) {
        const float g_wi_wt = nmap_G1(wi, wt, wp, wt, wg);
        nmap_facet_eval(base, nmap_facet(s, wt, -s.wo), wi, &pb);
        pdf += probability_wp * g_wi_wt * pb;
        diff_pdf += probability_wp * g_wi_wt;
    }
    return (float)((double)pdf + (double)(diff_pdf * clamp_pos(dot(wi, wg))) / 3.14159265358979323846);
}

// Results travel by value: out-of-line functions writing through pointers
// would put the caller's variables in scratch memory.
struct NmapEval {
    f3 f;
    float pdf;
};
struct NmapSample {
    f3 v, wi;
    float pdf;
};

// bsdfCos (Main.cpp:94-170) and getPDF of one direction.
template <uint32_t BASE>
static __host__ __device__ __noinline__ NmapEval nmap_eval_t(xt_material m, Surf s, f3 wi, WalkKey wk, int base_dim) {
    const xt_material base = nmap_base<BASE>(m);
    f3 wg, wp, wt;
    const int mode = nmap_frame(s, &wg, &wp, &wt);
    NmapEval r;
    if (mode == 0) {
        r.f = bsdf_eval(base, s, wi, &r.pdf);
        return r;
    }
    if (mode == 1) {
        r.f = splat(0.0f);
        r.pdf = 1.0f;
        return r;
    }
    r.pdf = nmap_pdf_frame(base, s, wi, wg, wp, wt);
    f3 result = splat(0.0f), weight = splat(1.0f);
    const int b0 = base_dim >> 2;
    bool d = Rng::to_unit(walk_block(wk, b0).x) < nmap_lambda_p(s.wo, wp, wt, wg);
    f3 wr = -s.wo;
    for (int i = 0;; ++i) {
        const f3 wm = d ? wp : wt;
        const Surf f = nmap_facet(s, wm, wr);
        const uint4 u = walk_block(wk, b0 + 1 + i);
        float pb;
        result += weight * nmap_G1(wi, wm, wp, wt, wg) * nmap_facet_eval(base, f, wi, &pb);
        weight = weight * nmap_facet_sample(base, f, wm, u, &wr, &pb);
        if (Rng::to_unit(u.w) < nmap_G1(wr, wm, wp, wt, wg)) break;
        d = !d;
        if (is_zero(weight)) break;
        if (i + 1 > 100) break;
    }
    r.f = result;
    return r;
}

// evalAndSample (Main.cpp:172-241): the walk's throughput, exit direction
// and getPDF of it (-1 when every step sampled a delta lobe).  A NaN tangent
// facet leaves the reference's pdf unset; here it is 0.
template <uint32_t BASE>
static __host__ __device__ __noinline__ NmapSample nmap_sample_t(xt_material m, Surf s, WalkKey wk, int base_dim) {
    const xt_material base = nmap_base<BASE>(m);
    f3 wg, wp, wt;
    const int mode = nmap_frame(s, &wg, &wp, &wt);
    const int b0 = base_dim >> 2;
    NmapSample r;
    if (mode == 0) {
        const uint4 u = walk_block(wk, b0 + 1);
        r.v = bsdf_sample(base, s, make_frame(s.sn), Rng::to_unit(u.x), Rng::to_unit(u.y), Rng::to_unit(u.z), &r.wi,
                          &r.pdf);
        return r;
    }
    if (mode == 1) {
        r.v = splat(0.0f);
        r.wi = s.sn;
        r.pdf = 0.0f;
        return r;
    }
    f3 weight = splat(1.0f);
    bool d = Rng::to_unit(walk_block(wk, b0).x) < nmap_lambda_p(s.wo, wp, wt, wg);
    f3 wr = -s.wo;
    bool specular = true;
    for (int i = 0;; ++i) {
        const f3 wm = d ? wp : wt;
        const Surf f = nmap_facet(s, wm, wr);
        const uint4 u = walk_block(wk, b0 + 1 + i);
        float tmppdf;
        weight = weight * nmap_facet_sample(base, f, wm, u, &wr, &tmppdf);
        if (tmppdf >= 0.0f) specular = false;
        if (Rng::to_unit(u.w) < nmap_G1(wr, wm, wp, wt, wg)) break;
        d = !d;
        if (is_zero(weight)) break;
        if (i + 1 > 100) {
            weight = weight * splat(0.0f);
            break;
        }
    }
    r.v = weight;
    r.wi = wr;
    r.pdf = specular ? -1.0f : nmap_pdf_frame(base, s, wr, wg, wp, wt);
    return r;
}

static __host__ __device__ __noinline__ NmapEval nmap_eval(xt_material m, Surf s, f3 wi, WalkKey wk, int base_dim) {
    const uint32_t b = (uint32_t)m.p[0];
    if (b == XT_MAT_METAL) return nmap_eval_t<XT_MAT_METAL>(m, s, wi, wk, base_dim);
    if (b == XT_MAT_DIFFUSE) return nmap_eval_t<XT_MAT_DIFFUSE>(m, s, wi, wk, base_dim);
    return nmap_eval_t<XT_MAT_GGX_LAMBERT>(m, s, wi, wk, base_dim);
}
static __host__ __device__ __noinline__ NmapSample nmap_sample(xt_material m, Surf s, WalkKey wk, int base_dim) {
    const uint32_t b = (uint32_t)m.p[0];
    if (b == XT_MAT_METAL) return nmap_sample_t<XT_MAT_METAL>(m, s, wk, base_dim);
    if (b == XT_MAT_DIFFUSE) return nmap_sample_t<XT_MAT_DIFFUSE>(m, s, wk, base_dim);
    return nmap_sample_t<XT_MAT_GGX_LAMBERT>(m, s, wk, base_dim);
}

// Material dispatch at the shading call sites: the stochastic material reads
// the walk stream `base_dim` of this vertex, the others are deterministic.
// NMAP = false compiles the walk out (scenes without the material).
template <bool NMAP>
__device__ __forceinline__ f3 mat_eval(const xt_material& m, const Surf& s, f3 wi, float* pdf, const WalkKey& wk,
                                       int base_dim) {
    if (NMAP && m.type == XT_MAT_MICROFACET_NMAP) {
        const NmapEval r = nmap_eval(m, s, wi, wk, base_dim);
        *pdf = r.pdf;
        return r.f;
    }
    return bsdf_eval(m, s, wi, pdf);
}
template <bool NMAP>
__device__ __forceinline__ f3 mat_sample(const xt_material& m, const Surf& s, const Frame& fr, uint4 blk0,
                                         const WalkKey& wk, f3* wi, float* pdf) {
    if (NMAP && m.type == XT_MAT_MICROFACET_NMAP) {
        const NmapSample r = nmap_sample(m, s, wk, D_WALK_SAMPLE);
        *wi = r.wi;
        *pdf = r.pdf;
        return r.v;
    }
    return bsdf_sample(m, s, fr, Rng::to_unit(blk0.y), Rng::to_unit(blk0.z), Rng::to_unit(blk0.w), wi, pdf);
}

}  // namespace xt
