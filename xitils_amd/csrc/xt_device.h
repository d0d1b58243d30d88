// Device-resident scene layout, counter-based RNG, BSDFs and light sampling
// shared by the gfx950 kernels (and host code that prepares the tables).
#pragma once

#include "bvh_layout.h"
#include "xt_math.h"
#include "../../include/xitils_amd.h"

namespace xt {

// ------------------------------------------------------------------ HBM layout
//
// BVH node (W = BVH_WIDTH children, bvh_layout.h): 8 rows of W floats, 32 W
// bytes (256 bytes, two L2 lines, for W = 8):
//   rows lo.x[W] hi.x[W] lo.y[W] hi.y[W] lo.z[W] hi.z[W],
//   child[W] as int: >= 0 internal node index; < 0 leaf ~(first << 3 | (count - 1)),
//   reserved
// Empty child slots have lo = +inf, hi = -inf and never pass the slab test.
//
// Triangles, in leaf order ("slots"), 3 float4 = 48 bytes each:
//   [0] p0.xyz, global prim id   [1] p1.xyz, material index   [2] p2.xyz, flags
// Optional vertex normals per slot, 3 float4 (n0, n1, n2).
// Optional normal-map attributes per slot (slots whose material has a normal
// map), 6 float4: (tangent k.xyz, uv) for k = 0..2 then (bitangent k.xyz, uv):
//   [0] t0, u0  [1] t1, v0  [2] t2, u1  [3] b0, v1  [4] b1, u2  [5] b2, v2
// Flags bits 8..31 hold 1 + the normal map's texture index (0 = none).
constexpr int NODE_ROW_F4 = BVH_WIDTH / 4;     // float4 per row of a float node
constexpr int NODE_F4 = BVH_NODE_WORDS / 4;    // float4 per node (bvh_layout.h)
constexpr int TRI_F4 = 3;
constexpr int NMAP_F4 = 6;
constexpr uint32_t TRI_FLAG_NORMALS = 1u;
constexpr uint32_t TRI_FLAG_EMISSIVE = 2u;
constexpr uint32_t TRI_FLAG_TANGENTS = 4u;   // the object supplied tangents and bitangents
constexpr uint32_t TRI_NMAP_SHIFT = 8u;
constexpr uint32_t INVALID = 0xFFFFFFFFu;

XT_HD int leaf_code(uint32_t first, uint32_t count) { return ~(int)((first << 3) | (count - 1)); }

struct DevEnv {
    int has_sky;
    f3 sky;
    f3 sun_dir;
    float sun_omc;   // 1 - cos(half angle)
    float sun_sin2;  // sin^2(half angle)
    f3 sun_L;
    Frame sun_frame; // make_frame(sun_dir), built once on the host
    // image sky: XT_ENV_MAP_ANGULAR / XT_ENV_MAP_LATLONG (0: constant sky)
    int map_kind;
    int map_w, map_h;
    const float4* map;       // texels (rgb, pad), row 0 first
    const float* map_pdf;    // lat-long: per texel solid-angle pdf times sin(theta)
    const float* map_marg;   // lat-long: row CDF (map_h + 1)
    const float* map_cond;   // lat-long: per-row column CDFs (map_h x (map_w + 1))
};

// an RGB texture (rgb, pad per texel), row 0 first
struct DevTexture {
    int w, h;
    const float4* texels;
};

struct DevScene {
    const float4* nodes;
    const float4* tris;
    const float4* normals;      // nullptr when no object has vertex normals
    const float4* nmap_attr;    // NMAP_F4 per slot; nullptr without normal maps
    const DevTexture* textures; // normal maps
    const xt_material* mats;
    const float* light_pdf;     // per slot; area pdf of emissive slots (Scene.h:49-55)
    const uint2* light_objs;    // {first, count} into light_slots, one per emissive object
    const uint32_t* light_slots;
    uint32_t num_lights;
    uint32_t num_prims;
    uint32_t top_nodes;         // nodes [0, top_nodes) are the tree's top levels, staged in LDS by the traversal
    DevEnv env;
};

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter = (pixel, sample, vertex slot, block), key =
// (seed, stream); block b yields dims 4b..4b+3.
XT_HD uint4 philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// Dimension map (per vertex slot; slot 0 = camera, slot v = v-th surface vertex)
enum : int {
    D_RR = 0, D_BSDF_SEL = 1, D_BSDF_U1 = 2, D_BSDF_U2 = 3,
    D_LIGHT_KIND = 4, D_LIGHT_OBJ = 5, D_LIGHT_TRI = 6, D_ENV_SEL = 7,
    D_BARY0 = 8, D_BARY1 = 9, D_BARY2 = 10,
    D_ENV_U1 = 5, D_ENV_U2 = 6,   // an env sample uses the area-light object/triangle dims: one block
    D_FILM_X = 0, D_FILM_Y = 1
};

struct Rng {
    uint2 key;
    uint32_t pixel, sample;
    XT_HD uint4 block(uint32_t slot, int b) const { return philox(make_uint4(pixel, sample, slot, (uint32_t)b), key); }
    XT_HD static uint32_t pick(uint4 v, int w) { return w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w)); }
    XT_HD static float to_unit(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
};

// ------------------------------------------------------------------ surface
struct Surf {
    f3 p, n, wo, sn;   // position, normal, outgoing dir, shading normal (world)
    uint32_t mat;
    uint32_t slot;
};

// ------------------------------------------------------------------ BSDFs
// Metal / GGX pieces (Material.h:288-378).  pow(float, int) promotes to
// double in the reference (Material.h:297, 301, 312): those steps are double.
XT_HD float ggx_D(float alpha, f3 x, f3 n) {
    float alpha2 = alpha * alpha;
    float dn = dot(n, x);
    double dd = (double)dn;
    float c2 = (float)(dd * dd);
    float c4 = c2 * c2;
    float tan2 = 1 / c2 - 1;
    double s = (double)(alpha2 + tan2);
    float h = dn > 0 ? 1.0f : 0.0f;
    return (float)((double)(alpha2 * h) / (3.14159265358979323846 * (double)c4 * (s * s)));
}
XT_HD float ggx_G1(float alpha, f3 x, f3 n) {
    float alpha2 = alpha * alpha;
    double dd = (double)dot(n, x);
    float c2 = (float)(dd * dd);
    float tan2 = 1 / c2 - 1;
    float lambda = (-1 + safe_sqrt(1 + alpha2 * tan2)) / 2;
    return 1 / (1 + lambda);
}
// Disk point shared by the sampling routines: r = sqrt(u1), phi = 2 pi u2
// (Material.h:335-338, Vector.h:821-823 compute the same values).
struct DiskPt {
    float r, c, s;
};
XT_HD DiskPt disk_pt(float u1, float u2) {
    const float PI = 3.14159265358979323846f;
    DiskPt d;
    d.r = sqrtf(u1);
    float phi = u2 * (2.0f * PI);
    xt_sincosf(phi, &d.s, &d.c);
    return d;
}

// Visible-normal sampling, Heitz 2018 (Material.h:323-345); fr = make_frame(n).
XT_HD f3 ggx_sample_vndf(float alpha, f3 eye, const Frame& fr, DiskPt dp) {
    f3 Ve = to_local(fr, eye);
    f3 Vh = normalize(mk(Ve.x, alpha * Ve.y, alpha * Ve.z));
    float lensq = Vh.y * Vh.y + Vh.z * Vh.z;
    f3 T1 = lensq > 0 ? div(mk(0, -Vh.z, Vh.y), safe_sqrt(lensq)) : mk(0, 0, 1);
    f3 T2 = cross(Vh, T1);
    float t1 = dp.r * dp.c;
    float t2 = dp.r * dp.s;
    float s = 0.5f * (1.0f + Vh.x);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    f3 Nh = t1 * T1 + t2 * T2 + safe_sqrt(1.0f - t1 * t1 - t2 * t2) * Vh;
    f3 Ne = normalize(mk(clamp_pos(Nh.x), alpha * Nh.y, alpha * Nh.z));
    return to_world(fr, Ne.x, Ne.y, Ne.z);
}
XT_HD f3 schlick(f3 f0, float c) { return f0 + (splat(1.0f) - f0) * xt_pow5f(1 - c); }

// cosine-weighted hemisphere in frame fr (Vector.h:817-824)
XT_HD f3 sample_cos_hemisphere(const Frame& fr, float u, DiskPt dp) {
    return to_world(fr, safe_sqrt(1.0f - u), dp.r * dp.c, dp.r * dp.s);
}

XT_HD float cos_pdf(float c) { return (float)((double)c / 3.14159265358979323846); }

// Material::bsdfCos (Material.h:17) and Material::getPDF (Material.h:31) of
// one direction, sharing the half vector, D and G1(wo) between the two.
XT_HD f3 bsdf_eval(const xt_material& m, const Surf& s, f3 wi, float* pdf) {
    const float PI = 3.14159265358979323846f;
    switch (m.type) {
    case XT_MAT_DIFFUSE: {  // Material.h:59-65, 78-81
        float c = clamp_pos(dot(s.sn, wi));
        *pdf = cos_pdf(c);
        return div(mk(m.p[0], m.p[1], m.p[2]), PI) * c;
    }
    case XT_MAT_METAL: {  // Material.h:347-356, 372-378
        float a = m.p[0];
        f3 h = normalize(wi + s.wo);
        float D = ggx_D(a, h, s.sn);
        float g1o = ggx_G1(a, s.wo, s.sn);
        float dv = g1o * clamp_pos(dot(s.wo, h)) * D / dot(s.wo, s.sn);
        *pdf = dv / (4 * dot(s.wo, h));
        f3 Fr = schlick(mk(m.p[1], m.p[2], m.p[3]), dot(s.wo, h));
        float G = ggx_G1(a, wi, s.sn) * g1o;
        return div(Fr * D * G, clamp_pos(4 * dot(s.wo, s.sn)));
    }
    case XT_MAT_GGX_LAMBERT: {
        float cosI = dot(s.sn, wi);
        if (!(cosI > 0.0f)) {
            *pdf = 0.0f;
            return splat(0.0f);
        }
        float a = m.p[3];
        f3 h = normalize(wi + s.wo);
        float D = ggx_D(a, h, s.sn);
        float g1o = ggx_G1(a, s.wo, s.sn);
        float dv = g1o * clamp_pos(dot(s.wo, h)) * D / dot(s.wo, s.sn);
        float ps = m.p[5];
        *pdf = ps * (dv / (4 * dot(s.wo, h))) + (1.0f - ps) * cos_pdf(cosI);
        f3 Fr = schlick(splat(m.p[4]), dot(s.wo, h));
        float G = ggx_G1(a, wi, s.sn) * g1o;
        f3 spec = div(Fr * D * G, clamp_pos(4 * dot(s.wo, s.sn)));
        f3 diff = mk(m.p[0], m.p[1], m.p[2]) * ((1.0f - m.p[4]) / PI) * cosI;
        return spec + diff;
    }
    default:
        *pdf = 0.0f;
        return splat(0.0f);
    }
}

XT_HD float fresnel_dielectric(f3 wo, f3 n, float eta_o, float eta_i) {  // Material.h:265-271
    const float c = fabsf(dot(wo, n));
    float r = eta_o / eta_i;
    float g_sq = r * r - 1.0f + c * c;
    if (g_sq <= 0.0f) return 1.0f;
    float g = sqrtf(g_sq);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * (a * a) * (1.0f + b * b);
}

// Material::evalAndSample (Material.h:25): f*cos/pdf, wi, pdf (-1 = delta).
// fr = make_frame(s.sn), shared with the light sampler.
XT_HD f3 bsdf_sample(const xt_material& m, const Surf& s, const Frame& fr, float usel, float u1, float u2, f3* wi,
                     float* pdf) {
    const DiskPt dp = disk_pt(u1, u2);
    switch (m.type) {
    case XT_MAT_DIFFUSE: {  // Material.h:67-76
        *wi = sample_cos_hemisphere(fr, u1, dp);
        *pdf = cos_pdf(dot(*wi, s.sn));
        return mk(m.p[0], m.p[1], m.p[2]);
    }
    case XT_MAT_SPEC_REFL: {  // Material.h:180-187
        *wi = -normalize(s.wo - 2 * dot(s.sn, s.wo) * s.sn);
        *pdf = -1.0f;
        return splat(1.0f);
    }
    case XT_MAT_SPEC_REFR:
    case XT_MAT_SPEC_FRESNEL: {  // Material.h:196-250
        float index = m.p[0];
        float eta_i, eta_o;
        if (dot(s.n, s.wo) > 0.0f) {
            eta_o = 1.0f;
            eta_i = index;
        } else {
            eta_o = index;
            eta_i = 1.0f;
        }
        if (m.type == XT_MAT_SPEC_FRESNEL && fresnel_dielectric(s.wo, s.sn, eta_o, eta_i) >= usel) {
            *wi = -normalize(s.wo - 2 * dot(s.sn, s.wo) * s.sn);
        } else {
            float ri = eta_o / eta_i;
            float c = dot(-s.wo, s.sn);
            *wi = ri * (-s.wo - c * s.sn) - safe_sqrt(1 - ri * ri * (1 - c * c)) * s.sn;
        }
        *pdf = -1.0f;
        return splat(1.0f);
    }
    case XT_MAT_METAL: {  // Material.h:358-370
        float a = m.p[0];
        f3 h = ggx_sample_vndf(a, s.wo, fr, dp);
        *wi = 2 * dot(s.wo, h) * h - s.wo;
        f3 fc = bsdf_eval(m, s, *wi, pdf);
        f3 r = div(fc, *pdf);
        if (!(*pdf > 0.0f) || !finite3(r)) return splat(0.0f);
        return r;
    }
    case XT_MAT_GGX_LAMBERT: {
        float a = m.p[3];
        if (usel < m.p[5]) {
            f3 h = ggx_sample_vndf(a, s.wo, fr, dp);
            *wi = 2 * dot(s.wo, h) * h - s.wo;
        } else {
            *wi = sample_cos_hemisphere(fr, u1, dp);
        }
        f3 fc = bsdf_eval(m, s, *wi, pdf);
        if (!(*pdf > 0.0f)) return splat(0.0f);
        f3 r = div(fc, *pdf);
        if (!finite3(r)) return splat(0.0f);
        return r;
    }
    default:  // Emission (Material.h:400-402)
        *pdf = 0.0f;
        return splat(0.0f);
    }
}

XT_HD bool mat_is_delta(uint32_t t) { return t == XT_MAT_SPEC_REFL || t == XT_MAT_SPEC_REFR || t == XT_MAT_SPEC_FRESNEL; }

// Emission::getEmission, one-sided on n (Material.h:408-410).
XT_HD f3 emitted(const xt_material& m, f3 wo, f3 n) { return dot(wo, n) > 0.0f ? mk(m.p[0], m.p[1], m.p[2]) : splat(0.0f); }

// ------------------------------------------------------------------ environment
XT_HD bool env_has_sun(const DevEnv& e) { return !is_zero(e.sun_L) && e.sun_omc > 0.0f; }
XT_HD bool env_in_sun(const DevEnv& e, f3 du) {
    f3 c = cross(du, e.sun_dir);
    return dot(du, e.sun_dir) > 0.0f && dot(c, c) <= e.sun_sin2;
}
// A texture as the lookups see it (plain values, so the out-of-line lookups
// take no address of a kernel argument); kind XT_ENV_MAP_LATLONG wraps in x,
// any other kind clamps.
struct TexRef {
    int kind, w, h;
    const float4* texels;
};
// Texture::rgb(x, y) (Texture.h:76-80) with warpClamp (Texture.h:161-171);
// the lat-long map wraps in x.
XT_HD f3 tex_texel(const TexRef& m, int x, int y) {
    if (m.kind == XT_ENV_MAP_LATLONG) {
        x %= m.w;
        if (x < 0) x += m.w;
    } else {
        x = min(max(x, 0), m.w - 1);
    }
    y = min(max(y, 0), m.h - 1);
    const float4 t = m.texels[(size_t)y * m.w + x];
    return mk(t.x, t.y, t.z);
}
// Texture::rgb(uv), bilinear filter (Texture.h:83-103)
XT_HD f3 tex_bilinear(const TexRef& m, float u, float v) {
    const int x0 = (int)floorf(u * m.w - 0.5f);
    const int y0 = (int)floorf(v * m.h - 0.5f);
    const float wx1 = (u * m.w - 0.5f) - x0;
    const float wy1 = (v * m.h - 0.5f) - y0;
    const float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1;
    return wx0 * wy0 * tex_texel(m, x0, y0) + wx0 * wy1 * tex_texel(m, x0, y0 + 1) + wx1 * wy0 * tex_texel(m, x0 + 1, y0) +
           wx1 * wy1 * tex_texel(m, x0 + 1, y0 + 1);
}
// lat-long coordinates of a unit direction (y up, row 0 at the zenith).  Out
// of line: the double-precision atan2 and acos would otherwise add their
// registers to every kernel that can evaluate a sky pdf.  Returned by value
// (in registers; output pointers would go through the call's stack).
struct LatLongUV {
    float u, v;
};
static __host__ __device__ __noinline__ LatLongUV env_latlong_uv(f3 d) {
    const float PI = 3.14159265358979323846f;
    return {(xt_atan2f(d.z, d.x) + PI) / (2.0f * PI), xt_acosf(fminf(fmaxf(d.y, -1.0f), 1.0f)) / PI};
}
// SkySphereFromImage::getRadiance (SkySphere.h:20-33) or the lat-long map.
// Kept out of line: inlined, the lookup's registers lower the occupancy of
// every kernel that evaluates the sky.
static __host__ __device__ __noinline__ f3 env_sky_map(int kind, int w, int h, const float4* texels, f3 wi) {
    const TexRef m{kind, w, h, texels};
    if (kind == XT_ENV_MAP_ANGULAR) {
        const f3 d = normalize(-wi);
        const float sq = sqrtf(d.x * d.x + d.y * d.y);
        if (sq == 0.0f) return splat(0.0f);
        const float r = (float)((1.0 / 3.14159265358979323846) * (double)xt_acosf(d.z) / (double)sq);
        return tex_bilinear(m, (d.x * r + 1.0f) / 2.0f, (d.y * r + 1.0f) / 2.0f);
    }
    const LatLongUV uv = env_latlong_uv(normalize(wi));
    return tex_bilinear(m, uv.u, uv.v);
}
// the image sky, or SkySphereUniform::getRadiance (SkySphere.h:40-42)
XT_HD f3 env_sky(const DevEnv& e, f3 wi) {
    return e.map_kind != XT_ENV_MAP_NONE ? env_sky_map(e.map_kind, e.map_w, e.map_h, e.map, wi) : e.sky;
}
// the sky plus the sun disk
XT_HD f3 env_radiance(const DevEnv& e, f3 d) {
    f3 L = env_sky(e, d);
    if (env_has_sun(e) && env_in_sun(e, normalize(d))) L += e.sun_L;
    return L;
}
// largest i in [0, n) with cdf[i] <= u (cdf[0] = 0, cdf[n] = 1)
XT_HD int cdf_find(const float* cdf, int n, float u) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// lat-long importance sample: row by the marginal CDF, column by the row's
// CDF, the remainders as offsets inside the texel.
XT_HD f3 env_map_sample(const DevEnv& e, float u1, float u2) {
    const float PI = 3.14159265358979323846f;
    const int W = e.map_w, H = e.map_h;
    const int y = cdf_find(e.map_marg, H, u1);
    const float dy = e.map_marg[y + 1] - e.map_marg[y];
    const float fy = dy > 0.0f ? (u1 - e.map_marg[y]) / dy : 0.5f;
    const float* c = e.map_cond + (size_t)y * (W + 1);
    const int x = cdf_find(c, W, u2);
    const float dx = c[x + 1] - c[x];
    const float fx = dx > 0.0f ? (u2 - c[x]) / dx : 0.5f;
    const float u = (x + fx) / (float)W, v = (y + fy) / (float)H;
    const float phi = 2.0f * PI * u - PI, theta = PI * v;
    float st, ct, sp, cp;
    xt_sincosf(theta, &st, &ct);
    xt_sincosf(phi, &sp, &cp);
    return mk(st * cp, ct, st * sp);
}
XT_HD float env_map_pdf(const DevEnv& e, f3 du) {
    const LatLongUV uv = env_latlong_uv(du);
    const int x = min((int)(uv.u * e.map_w), e.map_w - 1), y = min((int)(uv.v * e.map_h), e.map_h - 1);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - du.y * du.y));
    return st > 0.0f ? e.map_pdf[(size_t)y * e.map_w + x] / st : 0.0f;
}

// Environment light strategy: pick env vs area lights, then sun cone vs
// cosine sky.
struct EnvNee {
    bool on;
    float p_select, p_sun;
};
XT_HD EnvNee env_nee_setup(const DevEnv& e, bool enabled, bool have_area) {
    EnvNee n;
    bool sky = !is_zero(e.sky) || e.map_kind != 0;
    n.on = enabled && e.has_sky && (env_has_sun(e) || sky);
    n.p_select = n.on ? (have_area ? 0.5f : 1.0f) : 0.0f;
    n.p_sun = env_has_sun(e) ? (sky ? 0.5f : 1.0f) : 0.0f;
    return n;
}
XT_HD float env_nee_pdf(const DevEnv& e, const EnvNee& s, f3 du, f3 n) {
    const float PI = 3.14159265358979323846f;
    float p = 0.0f;
    if (s.p_sun > 0.0f && env_in_sun(e, du)) p += s.p_sun * (1.0f / (2.0f * PI * e.sun_omc));
    if (s.p_sun < 1.0f) {
        if (e.map_kind == XT_ENV_MAP_LATLONG) p += (1.0f - s.p_sun) * env_map_pdf(e, du);
        else p += (1.0f - s.p_sun) * cos_pdf(clamp_pos(dot(du, n)));
    }
    return p;
}
// Sun cone (uniform in solid angle) or cosine sky about the shading normal
// (frame fr); both branches share phi and the frame arithmetic, so the
// choice is a select, not a divergent branch.
XT_HD f3 env_nee_sample(const DevEnv& e, const EnvNee& s, const Frame& fr, float usel, float u1, float u2) {
    const bool sun = usel < s.p_sun;
    if (!sun && e.map_kind == XT_ENV_MAP_LATLONG) return env_map_sample(e, u1, u2);
    const DiskPt dp = disk_pt(u1, u2);
    float uo = u1 * e.sun_omc;
    float a = sun ? 1.0f - uo : safe_sqrt(1.0f - u1);
    float b = sun ? safe_sqrt(uo * (2.0f - uo)) : dp.r;
    Frame f;
    f.e1 = sun ? e.sun_frame.e1 : fr.e1;
    f.e2 = sun ? e.sun_frame.e2 : fr.e2;
    f.e3 = sun ? e.sun_frame.e3 : fr.e3;
    return to_world(f, a, b * dp.c, b * dp.s);
}

__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// Object.h:59-68: shading.n = normalize(n.b * shading.n + n.r * tangent +
// n.g * bitangent) with n = rgb(texCoord) * 2 - 1, from the slot's
// interpolated texCoord and normalised tangent frame (TriangleIndexed.h:
// 139-168); without tangents the frame is zero.  Out of line: only
// normal-mapped hits run it.
static __device__ __noinline__ f3 normal_mapped(const float4* attr, bool tangents, int w, int h, const float4* texels,
                                                float b0, float b1, f3 sn) {
    const float4 a0 = attr[0], a1 = attr[1], a2 = attr[2], a3 = attr[3], a4 = attr[4], a5 = attr[5];
    const f3 uv = lerp3(mk(a0.w, a1.w, 0.0f), mk(a2.w, a3.w, 0.0f), mk(a4.w, a5.w, 0.0f), b0, b1);
    f3 tg = splat(0.0f), bt = splat(0.0f);
    if (tangents) {
        tg = normalize(lerp3(xyz(a0), xyz(a1), xyz(a2), b0, b1));
        bt = normalize(lerp3(xyz(a3), xyz(a4), xyz(a5), b0, b1));
    }
    const TexRef m{XT_ENV_MAP_NONE, w, h, texels};
    const f3 c = tex_bilinear(m, uv.x, uv.y) * 2.0f - splat(1.0f);
    return normalize(c.z * sn + c.x * tg + c.y * bt);
}

// SurfaceIntersection at a hit (TriangleIndexed.h:133-172, then the identity
// pass of Object.h:106-144 where Transform::asNormal normalises n, wo and
// shading.n once more).
// NMAP: the scene has normal maps (the lookup is compiled in only then).
template <bool NMAP = true>
__device__ __forceinline__ Surf make_surf(const DevScene& sc, uint32_t slot, f3 d, float b0, float b1) {
    const float4* tr = sc.tris + (size_t)slot * TRI_F4;
    float4 v0 = tr[0], v1 = tr[1], v2 = tr[2];
    f3 p0 = xyz(v0), p1 = xyz(v1), p2 = xyz(v2);
    Surf s;
    s.slot = slot;
    s.mat = __float_as_uint(v1.w);
    uint32_t flags = __float_as_uint(v2.w);
    s.p = lerp3(p0, p1, p2, b0, b1);
    f3 wo = normalize(-d);
    f3 n;
    if (flags & TRI_FLAG_NORMALS) {
        const float4* nr = sc.normals + (size_t)slot * TRI_F4;
        n = normalize(lerp3(xyz(nr[0]), xyz(nr[1]), xyz(nr[2]), b0, b1));
    } else {
        n = normalize(cross(p1 - p0, p2 - p0));
    }
    f3 sn = face_forward(n, wo);
    const uint32_t nm = flags >> TRI_NMAP_SHIFT;
    if (NMAP && nm != 0) {
        const DevTexture t = sc.textures[nm - 1];
        sn = normal_mapped(sc.nmap_attr + (size_t)slot * NMAP_F4, (flags & TRI_FLAG_TANGENTS) != 0, t.w, t.h, t.texels,
                           b0, b1, sn);
    }
    s.n = normalize(n);
    s.wo = normalize(wo);
    s.sn = normalize(sn);
    return s;
}

__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

}  // namespace xt
