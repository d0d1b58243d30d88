// ORACLE — test infrastructure only.
//
// A CPU restatement, in plain C++17 (g++, -ffp-contract=off, no fast-math), of
// the reference's path-tracing hot path: BVH closest/any-hit traversal, the
// PBRT-style watertight triangle test, the Material BSDFs, light sampling and
// the StandardPathTracer / NaivePathTracer estimators.  It is the checker the
// tests compare the HIP product against; nothing in the product links it.
//
// Every function cites the reference file:line it follows.  Where the
// reference depends on a library absent from this image (libsimdpp lane order,
// MSVC <random>), the choice made here is stated at the function.
//
// Deviations, all shared with the product by specification (not by code):
//  * Sampler (Sampler.h:77-93, one MT19937 per 16x16 tile) is replaced by a
//    counter-based Philox4x32-10 stream keyed by (seed, stream) and counted by
//    (pixel, sample, vertex slot, 4-word block).  Dimension map: RngDim.
//  * Tie rule: a triangle passes the reference test (TriangleIndexed.h:75-131)
//    against the ray's ORIGINAL tMax; the closest hit is the lexicographic
//    minimum of (t, global prim id) over passing triangles, so the result does
//    not depend on BVH shape or traversal order.  The reference keeps "last
//    visited wins" (it accepts t == tMax, TriangleIndexed.h:123).
//  * max_depth (bounce cap) is added; <= 0 keeps the reference's unbounded,
//    roulette-only behaviour (PathTracer.h:149-155).
//  * Optional environment NEE (sun cone + cosine sky) MIS-combined with BSDF
//    escapes, off by default (the reference adds sky only on BSDF escapes,
//    PathTracer.h:192-197).
//  * Optional uniform barycentric light sampling (the reference's
//    TriangleIndexed.h:184-192 normalised-uniforms scheme is the default).
//  * GGX-over-Lambert material (new, needed by the heightfield workload) uses
//    the one-sample lobe mixture of GlossyWithHighLight (Material.h:134-175).
//
// Planted slips: building with -DXO_SLIP=k (k > 0) puts one deliberate
// mistake into the restatement, so the pin tests can show that they fail
// under it (tests/test_oracle_slips.py).  The oracle itself is built with
// XO_SLIP = 0, which changes nothing.
//   1  film position: y not flipped               (RenderTarget.h:77)
//   2  film position: jitter not centred on the pixel corner (RenderTarget.h:78-79)
//   3  roulette: survivor weight not divided by the survival probability (PathTracer.h:153)
//   4  light pdf: no division by the number of light objects (Scene.h:46, 53)
//   5  light pdf: no division by the mesh's triangle count (TriangleMesh.h:225, 234)
//   6  light pdf: no division by the object's area scaling (Object.h:95, 110)
#pragma once

#ifndef XO_SLIP
#define XO_SLIP 0
#endif

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace xo {

static const float Pi = (float)M_PI;                              // Utils.h:32
static const float Infinity = std::numeric_limits<float>::infinity();  // Utils.h:35

// ---------------------------------------------------------------- math (L0)
// Vector3f restated without libsimdpp: component-wise ops in float, the pad
// lane is always 0 (Vector.h:216-394).
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() {}
    explicit V3(float v) : x(v), y(v), z(v) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
    V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    V3 operator*(const V3& v) const { return V3(x * v.x, y * v.y, z * v.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator-() const { return V3(-x, -y, -z); }
    // Vector.h:325-335: v / s is v * (1/s)
    V3 operator/(float s) const { float inv = 1.0f / s; return *this * inv; }
    V3& operator+=(const V3& v) { x += v.x; y += v.y; z += v.z; return *this; }
    V3& operator*=(const V3& v) { x *= v.x; y *= v.y; z *= v.z; return *this; }
    V3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
    V3& operator/=(float s) { float inv = 1.0f / s; return *this *= inv; }
    bool isZero() const { return x == 0.0f && y == 0.0f && z == 0.0f; }  // Vector.h:379-384
    // Vector.h:357-360; the 4-lane reduce_add is taken as (x + y) + z (the
    // SSE3 horizontal-add order with a zero pad lane).  libsimdpp is absent,
    // so this lane order is a stated choice.
    float lengthSq() const { return (x * x + y * y) + z * z; }
    float length() const { return sqrtf(lengthSq()); }
    V3& normalize() { *this /= length(); return *this; }                // Vector.h:365
    int maxDimension() const { return (x > y && x > z) ? 0 : ((y > z) ? 1 : 2); }  // Vector.h:376
};
inline V3 operator*(float s, const V3& v) { return v * s; }               // Vector.h:609-611
inline float dot(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }  // Vector.h:628-632
inline V3 normalize(const V3& v) { return v / v.length(); }               // Vector.h:613
inline V3 vabs(const V3& v) { return V3(fabsf(v.x), fabsf(v.y), fabsf(v.z)); }
// Vector.h:637-654: computed in float64 then rounded to float.
inline V3 cross(const V3& a, const V3& b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return V3((float)(ay * bz - az * by), (float)(az * bx - ax * bz), (float)(ax * by - ay * bx));
}
inline V3 permute(const V3& v, int a, int b, int c) { return V3(v[a], v[b], v[c]); }
inline V3 vmin(const V3& a, const V3& b) { return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
inline V3 vmax(const V3& a, const V3& b) { return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
// Utils.h:81-83 / Vector.h:745-757: t1*a + t2*b + (1 - t1 - t2)*c
inline V3 lerp3(const V3& a, const V3& b, const V3& c, float t1, float t2) {
    return t1 * a + t2 * b + (1.0f - t1 - t2) * c;
}
inline V3 faceForward(const V3& n, const V3& v) { return (dot(n, v) < 0.0f) ? -n : n; }  // Vector.h:762
inline float clampPositive(float x) { return x > 0.0f ? x : 0.0f; }       // Utils.h:44
inline V3 clampPositive(const V3& v) { return V3(clampPositive(v.x), clampPositive(v.y), clampPositive(v.z)); }
inline float safeSqrt(float x) { return sqrtf(clampPositive(x)); }        // Utils.h:89
inline float heaviside(float x) { return x > 0 ? 1.0f : 0.0f; }          // Utils.h:45

// Transcendentals.  The reference calls the platform libm (std::sin, std::cos,
// std::pow, std::atan2, std::acos), whose last bits vary between libraries;
// here they are evaluated deterministically in IEEE arithmetic, the way the
// HIP library evaluates them, so both sides sample the same directions.
// sin/cos (|x| <= 256): Cody-Waite reduction x = k pi/2 + r, pi/2 split into
// three floats (the first two with 16 significant bits, so k * part is exact
// for |k| < 2^8), Taylor series of sin to r^9 and cos to r^10 on |r| <= pi/4
// (truncation < 2.3e-9 relative), quadrant k mod 4.  Pinned against float64
// sin/cos within 2 ulp (tests/test_libm.py).
inline void sinCos(float x, float* s, float* c) {
    const float k = rintf(x * 0x1.45f306p-1f);                        // x * 2/pi
    float r = x - k * 0x1.921ep+0f;
    r = r - k * 0x1.b544p-16f;
    r = r - k * 0x1.0b4612p-34f;
    const float r2 = r * r;
    const float ps = ((0x1.71de3ap-19f * r2 + -0x1.a01a02p-13f) * r2 + 0x1.111112p-7f) * r2 + -0x1.555556p-3f;
    const float sp = r + (r * r2) * ps;
    const float pc = (((-0x1.27e4fcp-22f * r2 + 0x1.a01a02p-16f) * r2 + -0x1.6c16c2p-10f) * r2 + 0x1.555556p-5f) * r2 +
                     -0.5f;
    const float cp = 1.0f + r2 * pc;
    const float q = k - 4.0f * floorf(k * 0.25f);
    if (q == 1.0f) { *s = cp; *c = -sp; }
    else if (q == 2.0f) { *s = -sp; *c = -cp; }
    else if (q == 3.0f) { *s = -cp; *c = sp; }
    else { *s = sp; *c = cp; }
}
// std::pow(float, int) promotes to double: x^5 by double products, one float rounding.
inline float pow5(float x) { const double d = x; return (float)(d * d * d * d * d); }
// atan2 and acos in float (the device's method, xt_libm.h): atan on [0, 1]
// as a P(a^2) with a degree-8 polynomial (near-minimax fit, float
// coefficients, fma Horner), then the octant; acos x = atan2(sqrt((1 - x)
// (1 + x)), x).  Pinned against float64 (2.4 / 2.8 ulp).
inline float atanPoly(float t) {
    float p = 0x1.758192p-9f;
    p = std::fma(p, t, -0x1.073e90p-6f);
    p = std::fma(p, t, 0x1.5db3fap-5f);
    p = std::fma(p, t, -0x1.335c98p-4f);
    p = std::fma(p, t, 0x1.b3d812p-4f);
    p = std::fma(p, t, -0x1.22e3a2p-3f);
    p = std::fma(p, t, 0x1.9972d8p-3f);
    p = std::fma(p, t, -0x1.5554a6p-2f);
    return std::fma(p, t, 1.0f);
}
inline float atan2F(float y, float x) {
    const float Pi = 0x1.921fb6p+1f, HalfPi = 0x1.921fb6p+0f;
    const float ax = std::fabs(x), ay = std::fabs(y);
    const float mx = std::fmax(ax, ay), mn = std::fmin(ax, ay);
    const float a = mx > 0.0f ? mn / mx : 0.0f;
    float r = a * atanPoly(a * a);
    if (ay > ax) r = HalfPi - r;
    if (std::signbit(x)) r = Pi - r;
    return std::copysign(r, y);
}
inline float acosF(float x) { return atan2F(std::sqrt((1.0f - x) * (1.0f + x)), x); }

// Vector.h:781-795
struct Basis {
    V3 e1, e2, e3;
    explicit Basis(const V3& v) {
        e1 = normalize(v);
        e2 = cross(v, fabsf(v.x) < 0.5f ? V3(1, 0, 0) : (fabsf(v.y) < 0.5f ? V3(0, 1, 0) : V3(0, 0, 1))).normalize();
        e3 = cross(e1, e2).normalize();
    }
    V3 fromGlobal(const V3& v) const { return V3(dot(v, e1), dot(v, e2), dot(v, e3)); }
    V3 toGlobal(float a, float b, float c) const { return e1 * a + e2 * b + e3 * c; }
    V3 toGlobal(const V3& v) const { return toGlobal(v.x, v.y, v.z); }
};

// Transform.h:11-160 (points, vectors, normals) with an explicit inverse.
struct Xf {
    float m[4][4];
    float mi[4][4];
    static Xf identity() {
        Xf t;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) t.m[i][j] = t.mi[i][j] = (i == j) ? 1.0f : 0.0f;
        return t;
    }
    static V3 point(const float (&a)[4][4], const V3& v) {               // Transform.h:62-73
        float xp = a[0][0] * v.x + a[0][1] * v.y + a[0][2] * v.z + a[0][3];
        float yp = a[1][0] * v.x + a[1][1] * v.y + a[1][2] * v.z + a[1][3];
        float zp = a[2][0] * v.x + a[2][1] * v.y + a[2][2] * v.z + a[2][3];
        float wp = a[3][0] * v.x + a[3][1] * v.y + a[3][2] * v.z + a[3][3];
        if (wp == 1) return V3(xp, yp, zp);
        return V3(xp, yp, zp) / wp;
    }
    static V3 vec(const float (&a)[4][4], const V3& v) {                 // Transform.h:105-111
        return V3(a[0][0] * v.x + a[0][1] * v.y + a[0][2] * v.z,
                  a[1][0] * v.x + a[1][1] * v.y + a[1][2] * v.z,
                  a[2][0] * v.x + a[2][1] * v.y + a[2][2] * v.z);
    }
    V3 operator()(const V3& p) const { return point(m, p); }
    V3 inversePoint(const V3& p) const { return point(mi, p); }
    V3 asVector(const V3& v) const { return vec(m, v); }
    V3 asVectorInverse(const V3& v) const { return vec(mi, v); }
    V3 asNormal(const V3& v) const { return vec(m, v).normalize(); }      // Transform.h:121-127 (forward matrix, Q12)
};

struct Ray {                                                               // Ray.h:8-32
    V3 o, d;
    float tMax = Infinity;
};

// ---------------------------------------------------------------- RNG (R24)
// Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
// 1, 2, 3"): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key bumps 0x9E3779B9 /
// 0xBB67AE85, 10 rounds.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Dimension map per (pixel, sample, vertex slot): 16 dims in 4 Philox blocks.
enum RngDim : int {
    DIM_RR = 0, DIM_BSDF_SEL = 1, DIM_BSDF_U1 = 2, DIM_BSDF_U2 = 3,
    DIM_LIGHT_KIND = 4, DIM_LIGHT_OBJ = 5, DIM_LIGHT_TRI = 6, DIM_ENV_SEL = 7,
    DIM_BARY0 = 8, DIM_BARY1 = 9, DIM_BARY2 = 10,
    DIM_ENV_U1 = 5, DIM_ENV_U2 = 6,   // an NEE sample is either an area point or an env direction
    DIM_FILM_X = 0, DIM_FILM_Y = 1,  // vertex slot 0 (camera)
    // uniform streams of the stochastic (random-walk) BSDF: sampling, the
    // light-sample evaluation, and a second evaluation of the same vertex
    DIM_WALK_SAMPLE = 64, DIM_WALK_EVAL = 1024, DIM_WALK_EVAL2 = 2048
};

struct PathRng {
    uint32_t key[2];
    uint32_t pixel, sample;
    uint32_t bits(uint32_t slot, int dim) const {
        uint32_t ctr[4] = {pixel, sample, slot, (uint32_t)(dim >> 2)};
        uint32_t out[4];
        philox4x32_10(ctr, key, out);
        return out[dim & 3];
    }
    // [0, 1 - 2^-24]: never returns 1 (the reference's randf can, Q6).
    float uni(uint32_t slot, int dim) const { return (float)(bits(slot, dim) >> 8) * 0x1p-24f; }
};

// ---------------------------------------------------------------- materials (L3)
enum MatType : uint32_t {
    MAT_DIFFUSE = 0, MAT_SPEC_REFL = 1, MAT_SPEC_REFR = 2, MAT_SPEC_FRESNEL = 3,
    MAT_METAL = 4, MAT_EMISSION = 5, MAT_GGX_LAMBERT = 6,
    MAT_MICROFACET_NMAP = 7   // MicrofacetNormalMapping over a Metal base (p[0] alpha, p[1..3] f0)
};
struct Material {
    uint32_t type = MAT_DIFFUSE;
    float p[7] = {0, 0, 0, 0, 0, 0, 0};
    bool emissive() const { return type == MAT_EMISSION; }
    V3 rgb(int off = 0) const { return V3(p[off], p[off + 1], p[off + 2]); }
};

// Interaction.h:16-32 (the parts the estimator reads).
struct Isect {
    V3 p, wo, n, sn;      // sn = shading.n
    int obj = -1;         // object index
    int tri = -1;         // triangle index within the object
    uint32_t prim = 0xFFFFFFFFu;  // global prim id
    const Material* mat = nullptr;
    float t = Infinity;
    float b0 = 0, b1 = 0;  // barycentrics of p0, p1 from the triangle test
};

// Texture::rgb with the default bilinear filter and clamped addressing
// (Texture.h:83-103, 161-171); RGB floats, row 0 first.
struct Texture2D {
    int w = 0, h = 0;
    std::vector<float> rgb;
    V3 texel(int x, int y) const {
        x = std::min(std::max(x, 0), w - 1);
        y = std::min(std::max(y, 0), h - 1);
        const float* t = &rgb[3 * ((size_t)y * w + x)];
        return V3(t[0], t[1], t[2]);
    }
    V3 bilinear(float u, float v) const {
        int x0 = (int)floorf(u * w - 0.5f);
        int y0 = (int)floorf(v * h - 0.5f);
        float wx1 = (u * w - 0.5f) - x0;
        float wy1 = (v * h - 0.5f) - y0;
        float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1;
        return wx0 * wy0 * texel(x0, y0) + wx0 * wy1 * texel(x0, y0 + 1) + wx1 * wy0 * texel(x0 + 1, y0) +
               wx1 * wy1 * texel(x0 + 1, y0 + 1);
    }
};

// GGX pieces, Material.h:288-378 (Metal).  Arithmetic types follow the
// reference: pow(float, int) promotes to double (Material.h:297,301,312).
struct GGX {
    float alpha;
    float D(const V3& x, const V3& n) const {                               // Material.h:294-302
        float alpha2 = alpha * alpha;
        double dn = (double)dot(n, x);
        float cosThetaH2 = (float)(dn * dn);
        float cosThetaH4 = cosThetaH2 * cosThetaH2;
        float tanThetaH2 = 1 / cosThetaH2 - 1;
        double s = (double)(alpha2 + tanThetaH2);
        return (float)((double)(alpha2 * heaviside(dot(n, x))) / (M_PI * (double)cosThetaH4 * (s * s)));
    }
    float G1(const V3& x, const V3& n) const {                              // Material.h:309-316
        float alpha2 = alpha * alpha;
        double dn = (double)dot(n, x);
        float cosThetaX2 = (float)(dn * dn);
        float tanThetaX2 = 1 / cosThetaX2 - 1;
        float lambda = (-1 + safeSqrt(1 + alpha2 * tanThetaX2)) / 2;
        return 1 / (1 + lambda);
    }
    float DV(const V3& x, const V3& n, const V3& eye) const {               // Material.h:304-307
        return G1(eye, n) * clampPositive(dot(eye, x)) * D(x, n) / dot(eye, n);
    }
    V3 sampleVNDF(const V3& eye, const V3& n, float U1, float U2) const {   // Material.h:323-345 (Heitz 2018)
        Basis basis(n);
        V3 Ve = basis.fromGlobal(eye);
        V3 Vh = V3(Ve.x, alpha * Ve.y, alpha * Ve.z).normalize();
        float lensq = Vh.y * Vh.y + Vh.z * Vh.z;
        V3 T1 = lensq > 0 ? V3(0, -Vh.z, Vh.y) / safeSqrt(lensq) : V3(0, 0, 1);
        V3 T2 = cross(Vh, T1);
        float r = sqrtf(U1);
        float phi = 2.0f * Pi * U2;
        float sphi, cphi;
        sinCos(phi, &sphi, &cphi);
        float t1 = r * cphi;
        float t2 = r * sphi;
        float s = 0.5f * (1.0f + Vh.x);
        t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
        V3 Nh = t1 * T1 + t2 * T2 + safeSqrt(1.0f - t1 * t1 - t2 * t2) * Vh;
        V3 Ne = V3(clampPositive(Nh.x), alpha * Nh.y, alpha * Nh.z).normalize();
        return basis.toGlobal(Ne);
    }
    float pdf(const V3& wo, const V3& n, const V3& wi) const {              // Material.h:372-378
        V3 h = (wi + wo).normalize();
        return DV(h, n, wo) / (4 * dot(wo, h));
    }
};
inline V3 schlick(const V3& f0, float c) {                                // Material.h:288-292
    return f0 + (V3(1.0f) - f0) * pow5(1 - c);
}
inline bool finite3(const V3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// Uniform stream of one call of a stochastic BSDF (the random walk of
// MicrofacetNormalMapping, which draws from the Sampler inside bsdfCos and
// evalAndSample): dim `base` picks the first facet, dims base + 4 + 4k ..
// base + 7 + 4k are walk step k (lobe select, u1, u2, escape test), so each
// step is one Philox block.
struct Walk {
    const PathRng* rng = nullptr;
    uint32_t slot = 0;
    int base = 0;
    float first() const { return rng->uni(slot, base); }
    float step(int k, int j) const { return rng->uni(slot, base + 4 + 4 * k + j); }
};
inline V3 nmapBsdfCos(const Material& m, const Isect& is, const V3& wi, const Walk* walk);
inline float nmapPdf(const Material& m, const Isect& is, const V3& wi);
inline V3 nmapEvalAndSample(const Material& m, const Isect& is, const Walk* walk, V3* wi, float* pdf);

// Material::bsdfCos (Material.h:17) per subclass.  `walk` is read only by
// the stochastic material.
inline V3 bsdfCos(const Material& m, const Isect& is, const V3& wi, const Walk* walk = nullptr) {
    switch (m.type) {
    case MAT_DIFFUSE:                                                       // Material.h:59-65
        return m.rgb() / (float)M_PI * clampPositive(dot(is.sn, wi));
    case MAT_METAL: {                                                       // Material.h:347-356
        GGX g{m.p[0]};
        V3 f0 = m.rgb(1);
        V3 h = (wi + is.wo).normalize();
        V3 Fr = schlick(f0, dot(is.wo, h));
        float D = g.D(h, is.sn);
        float G = g.G1(wi, is.sn) * g.G1(is.wo, is.sn);
        return Fr * D * G / clampPositive(4 * dot(is.wo, is.sn));
    }
    case MAT_GGX_LAMBERT: {
        // New material: GGX specular (Fresnel Schlick, scalar F0) plus a
        // Lambert base weighted by (1 - F0); rejects wi below the shading
        // horizon.
        float cosI = dot(is.sn, wi);
        if (!(cosI > 0.0f)) return V3(0.0f);
        GGX g{m.p[3]};
        V3 h = (wi + is.wo).normalize();
        V3 Fr = schlick(V3(m.p[4]), dot(is.wo, h));
        float D = g.D(h, is.sn);
        float G = g.G1(wi, is.sn) * g.G1(is.wo, is.sn);
        V3 spec = Fr * D * G / clampPositive(4 * dot(is.wo, is.sn));
        V3 diff = m.rgb() * ((1.0f - m.p[4]) / (float)M_PI) * cosI;
        return spec + diff;
    }
    case MAT_MICROFACET_NMAP:
        return nmapBsdfCos(m, is, wi, walk);
    default:                                                                // delta / emission: Material.h:396-398
        return V3(0.0f);
    }
}

// Material::getPDF (Material.h:31).  Delta materials are never asked (the
// estimator guards with pdf < 0).
inline float bsdfPdf(const Material& m, const Isect& is, const V3& wi) {
    switch (m.type) {
    case MAT_DIFFUSE:                                                       // Material.h:78-81 (double division)
        return (float)((double)clampPositive(dot(wi, is.sn)) / M_PI);
    case MAT_METAL: {
        GGX g{m.p[0]};
        return g.pdf(is.wo, is.sn, wi);
    }
    case MAT_GGX_LAMBERT: {
        float cosI = dot(is.sn, wi);
        if (!(cosI > 0.0f)) return 0.0f;
        GGX g{m.p[3]};
        float ps = m.p[5];
        float pd = (float)((double)cosI / M_PI);
        return ps * g.pdf(is.wo, is.sn, wi) + (1.0f - ps) * pd;
    }
    case MAT_MICROFACET_NMAP:
        return nmapPdf(m, is, wi);
    default:
        return 0.0f;
    }
}

// Vector.h:817-824
inline V3 sampleCosHemisphere(const V3& v, float u, float uphi) {
    Basis basis(v);
    float r = sqrtf(u);
    float phi = uphi * (2.0f * Pi);
    float sphi, cphi;
    sinCos(phi, &sphi, &cphi);
    return basis.toGlobal(safeSqrt(1.0f - u), r * cphi, r * sphi);
}

// Material.h:265-271
inline float fresnelDielectric(const V3& wo, const V3& n, float eta_o, float eta_i) {
    const float c = fabsf(dot(wo, n));
    float r = eta_o / eta_i;
    float g_sq = r * r - 1.0f + c * c;
    if (g_sq <= 0.0f) return 1.0f;
    float g = sqrtf(g_sq);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * (a * a) * (1.0f + b * b);
}
inline V3 refractDir(const V3& wo, const V3& n, float rindex) {           // Material.h:210-212
    float c = dot(-wo, n);
    return rindex * (-wo - c * n) - safeSqrt(1 - rindex * rindex * (1 - c * c)) * n;
}

// Material::evalAndSample (Material.h:25): returns f*cos/pdf, sets wi and pdf
// (pdf = -1 for delta lobes).  u = (sel, u1, u2) from dims 1..3; the
// stochastic material draws from `walk` instead.
inline V3 evalAndSample(const Material& m, const Isect& is, float usel, float u1, float u2, V3* wi, float* pdf,
                        const Walk* walk = nullptr) {
    switch (m.type) {
    case MAT_DIFFUSE: {                                                     // Material.h:67-76
        *wi = sampleCosHemisphere(is.sn, u1, u2);
        *pdf = (float)((double)dot(*wi, is.sn) / M_PI);
        return m.rgb();
    }
    case MAT_SPEC_REFL: {                                                   // Material.h:180-187
        const V3& wo = is.wo; const V3& n = is.sn;
        *wi = -(wo - 2 * dot(n, wo) * n).normalize();
        *pdf = -1.0f;
        return V3(1.0f);
    }
    case MAT_SPEC_REFR:                                                     // Material.h:196-216
    case MAT_SPEC_FRESNEL: {                                                // Material.h:225-250
        const V3& wo = is.wo; const V3& n = is.sn;
        float index = m.p[0];
        float eta_i, eta_o;
        if (dot(is.n, wo) > 0.0f) { eta_o = 1.0f; eta_i = index; } else { eta_o = index; eta_i = 1.0f; }
        if (m.type == MAT_SPEC_FRESNEL && fresnelDielectric(wo, n, eta_o, eta_i) >= usel) {
            *wi = -(wo - 2 * dot(n, wo) * n).normalize();
        } else {
            *wi = refractDir(wo, n, eta_o / eta_i);
        }
        *pdf = -1.0f;
        return V3(1.0f);
    }
    case MAT_METAL: {                                                       // Material.h:358-370
        GGX g{m.p[0]};
        V3 h = g.sampleVNDF(is.wo, is.sn, u1, u2);
        *wi = 2 * dot(is.wo, h) * h - is.wo;
        *pdf = g.pdf(is.wo, is.sn, *wi);
        V3 r = bsdfCos(m, is, *wi) / *pdf;
        if (!(*pdf > 0.0f) || !finite3(r)) return V3(0.0f);  // Q10 guard: drop non-finite samples
        return r;
    }
    case MAT_GGX_LAMBERT: {
        // One-sample lobe mixture after GlossyWithHighLight (Material.h:148-162):
        // lobe chosen with probability p_spec, value and pdf are the full sums.
        GGX g{m.p[3]};
        float ps = m.p[5];
        if (usel < ps) {
            V3 h = g.sampleVNDF(is.wo, is.sn, u1, u2);
            *wi = 2 * dot(is.wo, h) * h - is.wo;
        } else {
            *wi = sampleCosHemisphere(is.sn, u1, u2);
        }
        *pdf = bsdfPdf(m, is, *wi);
        if (!(*pdf > 0.0f)) return V3(0.0f);
        V3 r = bsdfCos(m, is, *wi) / *pdf;
        if (!finite3(r)) return V3(0.0f);
        return r;
    }
    case MAT_MICROFACET_NMAP:
        return nmapEvalAndSample(m, is, walk, wi, pdf);
    default:                                                                // Emission: Material.h:400-402
        *pdf = 0.0f;
        return V3(0.0f);
    }
}

// ---------------------------------------------------------------- microfacet-based normal mapping
// Schuessler et al. 2017, the MicrofacetNormalMapping material of
// Sandbox/MicrofacetBasedNormalMapping/Main.cpp:86-340 in its
// SameMaterialExplicit mode (Main.cpp:22): the normal-mapped facet wp and the
// tangent facet wt both carry the base material.  Parameters: p[0] = base
// material type (Diffuse, Metal or GGX+Lambert), p[1..6] = the base's p[0..5].
inline Material nmapBase(const Material& m) {
    Material b;
    b.type = (uint32_t)m.p[0];
    for (int k = 0; k < 6; ++k) b.p[k] = m.p[k + 1];
    return b;
}
// Facet frame (Main.cpp:96-117, 178-194, 244-260): 0 = wg == wp (the base
// material alone), 1 = wt has a NaN, 2 = the walk.
inline int nmapFrame(const Isect& is, V3* wg, V3* wp, V3* wt) {
    *wg = is.n;
    *wp = is.sn;
    if (dot(is.wo, *wg) < 0.0f) *wg *= -1.0f;
    if (dot(*wg, *wp) < 0.0f) *wp *= -1.0f;
    // Vector3f::operator== compares the lanes (Vector.h:344-349)
    if (wg->x == wp->x && wg->y == wp->y && wg->z == wp->z) return 0;
    *wt = cross(cross(*wg, *wp).normalize(), *wg).normalize();
    if (std::isnan(wt->x) || std::isnan(wt->y) || std::isnan(wt->z)) return 1;
    if (dot(*wp, *wt) > 0) *wt *= -1.0f;
    return 2;
}
// Main.cpp:323-337; min is Utils.h:37 (x < y ? x : y, so a NaN y passes).
inline float nmapAp(const V3& w, const V3& wp, const V3& wg) { return clampPositive(dot(w, wp)) / fabsf(dot(wp, wg)); }
inline float nmapAt(const V3& w, const V3& wp, const V3& wt, const V3& wg) {
    float d = dot(wp, wg);
    return clampPositive(dot(w, wt)) * safeSqrt(1 - d * d) / fabsf(d);   // pow(d, 2): the exact square, rounded once
}
inline float nmapLambdaP(const V3& w, const V3& wp, const V3& wt, const V3& wg) {
    return nmapAp(w, wp, wg) / (nmapAp(w, wp, wg) + nmapAt(w, wp, wt, wg));
}
inline float nmapG1(const V3& w, const V3& wm, const V3& wp, const V3& wt, const V3& wg) {
    if (!(dot(w, wm) > 0.0f)) return 0.0f;
    float g = clampPositive(dot(w, wg)) / (nmapAp(w, wp, wg) + nmapAt(w, wp, wt, wg));
    return 1.0f < g ? 1.0f : g;
}
inline bool isZero3(const V3& v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
// the isect the base material sees on facet wm (Main.cpp:125-130)
inline Isect nmapFacet(const Isect& is, const V3& wm, const V3& wr) {
    Isect f = is;
    f.n = wm;
    f.sn = wm;
    f.wo = -wr;
    return f;
}
// The base on a facet is one-sided in both directions, as GlossyClamped, the
// base the reference pairs with this material (Main.cpp:31-84): zero unless
// wo and wi are both above the facet.  (Unclamped, GGX's bsdfCos divides by
// a non-positive wo.n once the walk reaches a facet from behind.)
inline bool nmapSides(const Isect& f, const V3& wi) { return dot(f.wo, f.sn) > 0.0f && dot(wi, f.sn) > 0.0f; }
inline V3 nmapFacetCos(const Material& base, const Isect& f, const V3& wi) {
    return nmapSides(f, wi) ? bsdfCos(base, f, wi) : V3(0.0f);
}
inline float nmapFacetPdf(const Material& base, const Isect& f, const V3& wi) {
    return nmapSides(f, wi) ? bsdfPdf(base, f, wi) : 0.0f;
}
inline V3 nmapFacetSample(const Material& base, const Isect& f, const Walk* walk, int k, V3* wr, float* pdf) {
    V3 v = evalAndSample(base, f, walk->step(k, 0), walk->step(k, 1), walk->step(k, 2), wr, pdf);
    if (!nmapSides(f, *wr)) {
        *pdf = 0.0f;
        return V3(0.0f);
    }
    return v;
}
// bsdfCos as a random walk on the two facets (Main.cpp:94-170).
inline V3 nmapBsdfCos(const Material& m, const Isect& is, const V3& wi, const Walk* walk) {
    const Material base = nmapBase(m);
    V3 wg, wp, wt;
    int mode = nmapFrame(is, &wg, &wp, &wt);
    if (mode == 0) return bsdfCos(base, is, wi);
    if (mode == 1) return V3(0.0f);
    V3 result(0.0f), weight(1.0f);
    bool d = walk->first() < nmapLambdaP(is.wo, wp, wt, wg);
    V3 wr = -is.wo;
    int i = 0;
    float pdf;
    while (true) {
        const V3& wm = d ? wp : wt;
        Isect f = nmapFacet(is, wm, wr);
        result += weight * nmapG1(wi, wm, wp, wt, wg) * nmapFacetCos(base, f, wi);
        weight *= nmapFacetSample(base, f, walk, i, &wr, &pdf);
        float g = nmapG1(wr, wm, wp, wt, wg);
        if (walk->step(i, 3) < g) break;
        d = !d;
        if (isZero3(weight)) break;
        ++i;
        if (i > 100) break;
    }
    return result;
}
// evalAndSample as the same walk (Main.cpp:172-241); pdf = -1 only when every
// step sampled a delta lobe.  A NaN tangent facet leaves the reference's pdf
// unset; here it is 0.
inline V3 nmapEvalAndSample(const Material& m, const Isect& is, const Walk* walk, V3* wi, float* pdf) {
    const Material base = nmapBase(m);
    V3 wg, wp, wt;
    int mode = nmapFrame(is, &wg, &wp, &wt);
    if (mode == 0) return evalAndSample(base, is, walk->step(0, 0), walk->step(0, 1), walk->step(0, 2), wi, pdf);
    if (mode == 1) {
        *pdf = 0.0f;
        return V3(0.0f);
    }
    V3 weight(1.0f);
    bool d = walk->first() < nmapLambdaP(is.wo, wp, wt, wg);
    V3 wr = -is.wo;
    int i = 0;
    float tmppdf;
    bool specular = true;
    while (true) {
        const V3& wm = d ? wp : wt;
        Isect f = nmapFacet(is, wm, wr);
        weight *= nmapFacetSample(base, f, walk, i, &wr, &tmppdf);
        if (tmppdf >= 0.0f) specular = false;
        float g = nmapG1(wr, wm, wp, wt, wg);
        if (walk->step(i, 3) < g) break;
        d = !d;
        if (isZero3(weight)) break;
        ++i;
        if (i > 100) {
            weight *= V3(0.0f);
            break;
        }
    }
    *wi = wr;
    *pdf = specular ? -1.0f : nmapPdf(m, is, *wi);
    return weight;
}
// getPDF (Main.cpp:243-318): the approximation that treats the facets'
// inter-reflection as diffuse.  The tangent-facet terms are weighted by
// probability_wp as the reference writes them (Main.cpp:295-296), not by
// 1 - probability_wp.
inline float nmapPdf(const Material& m, const Isect& is, const V3& wi) {
    const Material base = nmapBase(m);
    V3 wg, wp, wt;
    int mode = nmapFrame(is, &wg, &wp, &wt);
    if (mode == 0) return bsdfPdf(base, is, wi);
    if (mode == 1) return 1.0f;
    float probability_wp = nmapLambdaP(is.wo, wp, wt, wg);
    float pdf = 0.0f, diff_pdf = 0.0f;
    float g_wi_wp = nmapG1(wi, wp, wp, wt, wg);
    if (probability_wp > 0.0f) {
        Isect f = is;
        f.n = wp;
        f.sn = wp;
        pdf += probability_wp * g_wi_wp * nmapFacetPdf(base, f, wi);
        diff_pdf += probability_wp * (1.0f - g_wi_wp);
    }
    if (probability_wp < 1.0f) {
        Isect f = is;
        f.n = wt;
        f.sn = wt;
        float g_wi_wt = nmapG1(wi, wt, wp, wt, wg);
        pdf += probability_wp * g_wi_wt * nmapFacetPdf(base, f, wi);
        diff_pdf += probability_wp * g_wi_wt;
    }
    return (float)((double)pdf + (double)(diff_pdf * clampPositive(dot(wi, wg))) / M_PI);
}

// Emission::getEmission (Material.h:408-410), one-sided on n.
inline V3 emission(const Material& m, const V3& wo, const V3& n) {
    return dot(wo, n) > 0.0f ? m.rgb() : V3(0.0f);
}

// ---------------------------------------------------------------- geometry (L2)
struct Mesh {                                                              // TriangleMesh.h:237-251
    std::vector<V3> pos, nrm;
    std::vector<float> uv;          // 2 per vertex (texCoords), optional
    std::vector<V3> tan, bitan;     // per vertex, optional
    std::vector<int> idx;
    int ntri() const { return (int)idx.size() / 3; }
    const V3& P(int t, int k) const { return pos[idx[t * 3 + k]]; }
    const V3& N(int t, int k) const { return nrm[idx[t * 3 + k]]; }
    float triArea(int t) const {                                           // TriangleIndexed.h:56-63
        return cross(P(t, 1) - P(t, 0), P(t, 2) - P(t, 0)).length() / 2.0f;
    }
};

// TriangleIndexed::intersect arithmetic (TriangleIndexed.h:80-131).  Passing
// means the reference would accept the triangle against tMax; outputs t and
// the barycentrics b0, b1 of p0, p1.
inline bool triTest(const V3& p0, const V3& p1, const V3& p2, const Ray& ray, float tMax,
                    float* tOut, float* b0Out, float* b1Out) {
    V3 p0t = p0 - ray.o, p1t = p1 - ray.o, p2t = p2 - ray.o;
    int kz = vabs(ray.d).maxDimension();
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    V3 d = permute(ray.d, kx, ky, kz);
    p0t = permute(p0t, kx, ky, kz); p1t = permute(p1t, kx, ky, kz); p2t = permute(p2t, kx, ky, kz);
    float sx = -d.x / d.z, sy = -d.y / d.z, sz = 1.0f / d.z;
    p0t.x += sx * p0t.z; p0t.y += sy * p0t.z;
    p1t.x += sx * p1t.z; p1t.y += sy * p1t.z;
    p2t.x += sx * p2t.z; p2t.y += sy * p2t.z;
    float e0 = p1t.x * p2t.y - p1t.y * p2t.x;
    float e1 = p2t.x * p0t.y - p2t.y * p0t.x;
    float e2 = p0t.x * p1t.y - p0t.y * p1t.x;
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;
    p0t.z *= sz; p1t.z *= sz; p2t.z *= sz;
    float tScaled = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
    if (det < 0 && (tScaled >= 0 || tScaled < tMax * det)) return false;
    if (det > 0 && (tScaled <= 0 || tScaled > tMax * det)) return false;
    float invDet = 1 / det;
    *b0Out = e0 * invDet;
    *b1Out = e1 * invDet;
    *tOut = tScaled * invDet;
    return true;
}

// Bounds3::intersect slab test (Bounds.h:176-191) over [0, tMax].
struct AABB {
    V3 mn = V3(std::numeric_limits<float>::max()), mx = V3(std::numeric_limits<float>::lowest());  // Bounds.h:98-103
    void merge(const AABB& b) { mn = vmin(mn, b.mn); mx = vmax(mx, b.mx); }
    void merge(const V3& p) { mn = vmin(mn, p); mx = vmax(mx, p); }
    V3 size() const { return mx - mn; }
    V3 center() const { return (mx + mn) / 2; }
    float refSurfaceArea() const { V3 s = size(); return 2 * s.x * s.y * s.z; }  // Bounds.h:144-147 (a volume, Q1)
    // The far ends carry a relative slack of 2^-20 (PBRT's 1 + 2*gamma(3)
    // argument) so rounding never culls a box holding a hit at exactly the
    // current closest t: with the (t, prim id) tie rule, equal-t hits in
    // another box must still be visited.  The reference has no slack.
    bool intersect(const Ray& ray, float tMax) const {
        const float slack = 1.0f + 0x1p-20f;
        float t1 = 0, t2 = tMax * slack;
        for (int i = 0; i < 3; ++i) {
            float invD = 1.0f / ray.d[i];
            float tNear = (mn[i] - ray.o[i]) * invD;
            float tFar = (mx[i] - ray.o[i]) * invD;
            if (tNear > tFar) std::swap(tNear, tFar);
            tFar *= slack;
            t1 = tNear > t1 ? tNear : t1;
            t2 = tFar < t2 ? tFar : t2;
            if (t1 > t2) return false;
        }
        return true;
    }
};

// The reference's binary BVH (AccelerationStructure.h:65-291): sort by
// centroid on the max-extent axis, <= 8 prims -> median split, else 8
// equal-count buckets with cost SA(L) nL + SA(R) nR where SA is the reference's
// 2xyz "surface area" (Q1); one primitive per leaf; near child first by
// sign(d[axis]) (AccelerationStructure.h:119-156).
struct RefBVH {
    struct Node { AABB box; int axis = 0; int prim = -1; int child[2] = {-1, -1}; };
    std::vector<Node> nodes;
    struct GB { int prim; AABB b; V3 c; };

    void build(const std::vector<AABB>& primBounds) {
        nodes.clear();
        std::vector<GB> g(primBounds.size());
        for (size_t i = 0; i < primBounds.size(); ++i) g[i] = GB{(int)i, primBounds[i], primBounds[i].center()};
        if (g.empty()) return;
        nodes.reserve(2 * g.size());
        nodes.emplace_back();
        AABB root;
        for (auto& x : g) root.merge(x.b);
        nodes[0].box = root;
        sub(g, 0, (int)g.size(), 0, 0, root);
    }
    void sub(std::vector<GB>& g, int b, int e, int node, int depth, AABB aabb) {
        if (e - b == 1) { nodes[node].prim = g[b].prim; nodes[node].box = g[b].b; return; }
        if (depth != 0) nodes[node].box = aabb;
        int axis = nodes[node].box.size().maxDimension();
        nodes[node].axis = axis;
        // std::sort is unstable in the reference (Q2); stable_sort plus the
        // prim id makes this build deterministic.
        std::stable_sort(g.begin() + b, g.begin() + e, [axis](const GB& x, const GB& y) { return x.c[axis] < y.c[axis]; });
        int n = e - b, split = 0;
        AABB s1, s2;
        if (n <= 8) {
            split = n / 2;
            for (int i = 0; i < split; ++i) s1.merge(g[b + i].b);
            for (int i = split; i < n; ++i) s2.merge(g[b + i].b);
        } else {
            const int nb = std::min(8, n);
            std::vector<AABB> bb(nb), l(nb - 1), r(nb - 1);
            std::vector<int> bs(nb), ls(nb - 1), rs(nb - 1);
            int beg = 0;
            for (int i = 0; i < nb; ++i) {
                int end = n * (i + 1) / nb;
                for (int k = beg; k < end; ++k) bb[i].merge(g[b + k].b);
                bs[i] = end - beg; beg = end;
            }
            for (int i = 0; i < nb - 1; ++i) {
                l[i] = bb[i]; ls[i] = bs[i];
                if (i > 0) { l[i].merge(l[i - 1]); ls[i] += ls[i - 1]; }
                r[i] = bb[nb - 1 - i]; rs[i] = bs[nb - 1 - i];
                if (i > 0) { r[i].merge(r[i - 1]); rs[i] += rs[i - 1]; }
            }
            float best = Infinity; int bi = 0;
            split = ls[0];
            for (int i = 0; i < nb - 1; ++i) {
                float sah = l[i].refSurfaceArea() * ls[i] + r[nb - 2 - i].refSurfaceArea() * rs[nb - 2 - i];
                if (sah < best) { split = ls[i]; bi = i; best = sah; }
            }
            s1 = l[bi]; s2 = r[nb - 2 - bi];
        }
        int c0 = (int)nodes.size(); nodes.emplace_back();
        int c1 = (int)nodes.size(); nodes.emplace_back();
        nodes[node].child[0] = c0; nodes[node].child[1] = c1;
        sub(g, b, b + split, c0, depth + 1, s1);
        sub(g, b + split, e, c1, depth + 1, s2);
    }
    // Visits leaves in the reference's order; f(prim, tCull) returns true to
    // stop (any-hit) and may lower tCull (closest hit).
    template <class F> bool traverse(const Ray& ray, float& tCull, F&& f, int node = 0) const {
        if (nodes.empty()) return false;
        const Node& nd = nodes[node];
        if (!nd.box.intersect(ray, tCull)) return false;
        if (nd.prim >= 0) return f(nd.prim, tCull);
        int first = ray.d[nd.axis] > 0 ? 0 : 1;
        if (traverse(ray, tCull, f, nd.child[first])) return true;
        return traverse(ray, tCull, f, nd.child[1 - first]);
    }
};

// ---------------------------------------------------------------- scene (L4)
struct Object {                                                            // Object.h:16-111
    Mesh mesh;
    Xf xf = Xf::identity();
    uint32_t material = 0;
    std::vector<uint32_t> triMaterial;   // optional per-triangle override (heightfield workload)
    uint32_t primBase = 0;               // global prim id of triangle 0
    RefBVH blas;
    float area = 0;                      // object-space mesh area (TriangleMesh.h:253-258)
    float areaScaling = 1;               // TriangleMesh.h:193-200
    uint32_t matOf(int t) const { return triMaterial.empty() ? material : triMaterial[t]; }
};

struct Env {
    V3 sky = V3(0.0f);
    bool hasSky = false;      // Scene::skySphere != nullptr
    // Sun disk: unit direction, 1 - cos and sin^2 of the half angle (both
    // rounded from float64 on the host, so the 0.25 deg disk keeps full
    // precision), radiance.
    V3 sunDir = V3(0, 1, 0);
    float sunOmc = 0.0f;
    float sunSin2 = 0.0f;
    V3 sunL = V3(0.0f);
    // Image sky (replaces the constant colour): 1 = angular map as
    // SkySphereFromImage (SkySphere.h:14-34), 2 = latitude-longitude map
    // (new: u = (atan2(d.z, d.x) + pi) / 2pi, v = acos(d.y) / pi, y up, row 0
    // at the zenith), with a 2D piecewise-constant importance distribution.
    int mapKind = 0;
    int mapW = 0, mapH = 0;
    std::vector<float> map;                      // rgb, row-major, row 0 first
    std::vector<float> margCdf, condCdf, pdfPix; // lat-long distribution (buildMapDistribution)

    bool hasSun() const { return !sunL.isZero() && sunOmc > 0.0f; }
    // Inside the disk: d . s > 0 and |d x s|^2 <= sin^2(half angle).
    bool inSun(const V3& dUnit) const {
        return dot(dUnit, sunDir) > 0.0f && cross(dUnit, sunDir).lengthSq() <= sunSin2;
    }
    // Texture::rgb(x, y) (Texture.h:76-80) with warpClamp (Texture.h:161-171);
    // the lat-long map wraps in x.
    V3 texel(int x, int y) const {
        if (mapKind == 2) {
            x %= mapW;
            if (x < 0) x += mapW;
        } else {
            x = std::min(std::max(x, 0), mapW - 1);
        }
        y = std::min(std::max(y, 0), mapH - 1);
        const float* t = &map[3 * ((size_t)y * mapW + x)];
        return V3(t[0], t[1], t[2]);
    }
    // Texture::rgb(uv), bilinear filter (Texture.h:83-103)
    V3 bilinear(float u, float v) const {
        int x0 = (int)floorf(u * mapW - 0.5f);
        int y0 = (int)floorf(v * mapH - 0.5f);
        int x1 = x0 + 1, y1 = y0 + 1;
        float wx1 = (u * mapW - 0.5f) - x0;
        float wy1 = (v * mapH - 0.5f) - y0;
        float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1;
        return wx0 * wy0 * texel(x0, y0) + wx0 * wy1 * texel(x0, y1) + wx1 * wy0 * texel(x1, y0) +
               wx1 * wy1 * texel(x1, y1);
    }
    // lat-long coordinates of a unit direction
    static void latlongUV(const V3& d, float* u, float* v) {
        *u = (atan2F(d.z, d.x) + Pi) / (2.0f * Pi);
        *v = acosF(std::min(std::max(d.y, -1.0f), 1.0f)) / Pi;
    }
    V3 skyRadiance(const V3& wi) const {
        if (mapKind == 1) {                         // SkySphere.h:20-33
            V3 d = normalize(-wi);
            float sq = sqrtf(d.x * d.x + d.y * d.y);
            if (sq == 0.0f) return V3(0.0f);
            float r = (float)((1.0 / M_PI) * (double)acosF(d.z) / (double)sq);
            float u = (d.x * r + 1.0f) / 2.0f;
            float v = (d.y * r + 1.0f) / 2.0f;
            return bilinear(u, v);
        }
        if (mapKind == 2) {
            float u, v;
            latlongUV(normalize(wi), &u, &v);
            return bilinear(u, v);
        }
        return sky;                                 // SkySphere.h:40-42
    }
    // SkySphere::getRadiance plus the sun disk.
    V3 radiance(const V3& d) const {
        V3 L = skyRadiance(d);
        if (hasSun() && inSun(normalize(d))) L += sunL;
        return L;
    }

    // Lat-long importance distribution: texel weight = the largest luminance
    // of its 3x3 neighbourhood (the bilinear lookup inside a pixel reads
    // those texels) times sin(theta) at the row centre; rows by a marginal
    // CDF, columns by per-row CDFs, sums in double.  pdfPix folds in the
    // Jacobian W H / (2 pi^2); the solid-angle pdf is pdfPix / sin(theta).
    void buildMapDistribution() {
        margCdf.clear(); condCdf.clear(); pdfPix.clear();
        if (mapKind != 2 || mapW <= 0 || mapH <= 0) return;
        const int W = mapW, H = mapH;
        auto lum = [&](int x, int y) {
            V3 t = texel(x, y);
            return 0.2126 * t.x + 0.7152 * t.y + 0.0722 * t.z;
        };
        std::vector<double> w((size_t)W * H);
        for (int y = 0; y < H; ++y) {
            double st = std::sin(M_PI * (y + 0.5) / H);
            for (int x = 0; x < W; ++x) {
                double m = 0.0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) m = std::max(m, lum(x + dx, y + dy));
                w[(size_t)y * W + x] = m * st;
            }
        }
        std::vector<double> row(H, 0.0);
        double total = 0.0;
        condCdf.assign((size_t)H * (W + 1), 0.0f);
        for (int y = 0; y < H; ++y) {
            double acc = 0.0;
            for (int x = 0; x < W; ++x) acc += w[(size_t)y * W + x];
            row[y] = acc;
            total += acc;
            double run = 0.0;
            for (int x = 0; x < W; ++x) {
                condCdf[(size_t)y * (W + 1) + x] = acc > 0.0 ? (float)(run / acc) : (float)x / (float)W;
                run += w[(size_t)y * W + x];
            }
            condCdf[(size_t)y * (W + 1) + W] = 1.0f;
        }
        margCdf.assign(H + 1, 0.0f);
        double run = 0.0;
        for (int y = 0; y < H; ++y) {
            margCdf[y] = total > 0.0 ? (float)(run / total) : (float)y / (float)H;
            run += row[y];
        }
        margCdf[H] = 1.0f;
        pdfPix.assign((size_t)W * H, 0.0f);
        for (size_t k = 0; k < w.size(); ++k)
            pdfPix[k] = total > 0.0 ? (float)(w[k] / total * W * H / (2.0 * M_PI * M_PI)) : 0.0f;
    }
    // largest i in [0, n) with cdf[i] <= u (cdf[0] = 0, cdf[n] = 1)
    static int cdfFind(const float* cdf, int n, float u) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (cdf[mid] <= u) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
    V3 sampleMap(float u1, float u2, float* pdf) const {
        const int W = mapW, H = mapH;
        int y = cdfFind(margCdf.data(), H, u1);
        float dy = margCdf[y + 1] - margCdf[y];
        float fy = dy > 0.0f ? (u1 - margCdf[y]) / dy : 0.5f;
        const float* c = &condCdf[(size_t)y * (W + 1)];
        int x = cdfFind(c, W, u2);
        float dx = c[x + 1] - c[x];
        float fx = dx > 0.0f ? (u2 - c[x]) / dx : 0.5f;
        float u = (x + fx) / (float)W, v = (y + fy) / (float)H;
        float phi = 2.0f * Pi * u - Pi, theta = Pi * v;
        float st, ct, sp, cp;
        sinCos(theta, &st, &ct);
        sinCos(phi, &sp, &cp);
        *pdf = st > 0.0f ? pdfPix[(size_t)y * W + x] / st : 0.0f;
        return V3(st * cp, ct, st * sp);
    }
    float mapPdf(const V3& dUnit) const {
        float u, v;
        latlongUV(dUnit, &u, &v);
        int x = std::min((int)(u * mapW), mapW - 1), y = std::min((int)(v * mapH), mapH - 1);
        float st = sqrtf(std::max(0.0f, 1.0f - dUnit.y * dUnit.y));
        return st > 0.0f ? pdfPix[(size_t)y * mapW + x] / st : 0.0f;
    }
};

struct SampledLight { V3 p, n, sn; int obj = -1, tri = -1; float pdf = 0; };

struct Scene {
    std::vector<Object> objects;
    std::vector<Material> materials;
    std::vector<int> matNormalMap;        // per material: index into textures, -1 = none (Material::normalmap)
    std::vector<Texture2D> textures;
    std::vector<int> lights;     // emissive objects in insertion order (Scene.h:17-22)
    Env env;
    RefBVH tlas;
    uint32_t nprim = 0;
    std::vector<int> primObj;    // global prim -> object
    std::vector<std::array<V3, 3>> worldTri;  // world-space vertices per global prim (brute force)

    void build() {
        nprim = 0;
        lights.clear();
        primObj.clear();
        worldTri.clear();
        std::vector<AABB> objBounds;
        for (size_t oi = 0; oi < objects.size(); ++oi) {
            Object& o = objects[oi];
            o.primBase = nprim;
            int nt = o.mesh.ntri();
            std::vector<AABB> tb(nt);
            o.area = 0;
            float scaled = 0;
            for (int t = 0; t < nt; ++t) {
                tb[t].merge(o.mesh.P(t, 0)); tb[t].merge(o.mesh.P(t, 1)); tb[t].merge(o.mesh.P(t, 2));
                float a = o.mesh.triArea(t);
                o.area += a;
                V3 w0 = o.xf(o.mesh.P(t, 0)), w1 = o.xf(o.mesh.P(t, 1)), w2 = o.xf(o.mesh.P(t, 2));
                float aw = cross(w1 - w0, w2 - w0).length() / 2.0f;
                scaled += a * (aw / a);                                     // TriangleIndexed.h:65-73
                primObj.push_back((int)oi);
                worldTri.push_back({w0, w1, w2});
            }
            o.areaScaling = scaled / o.area;
            o.blas.build(tb);
            nprim += nt;
            if (materials[o.material].emissive()) lights.push_back((int)oi);
            // Object::bound (Object.h:98-100): world box of the object box;
            // the reference builds it from a corner plus basis vectors
            // (Transform.h:160-185), here the 8 corners are transformed.  It
            // only bounds TLAS culling; hits follow the tie rule.
            AABB ob;
            for (int t = 0; t < nt; ++t) { ob.merge(tb[t]); }
            AABB wb;
            for (int c = 0; c < 8; ++c) {
                V3 q((c & 1) ? ob.mx.x : ob.mn.x, (c & 2) ? ob.mx.y : ob.mn.y, (c & 4) ? ob.mx.z : ob.mn.z);
                wb.merge(o.xf(q));
            }
            objBounds.push_back(wb);
        }
        tlas.build(objBounds);
    }

    // Builds the SurfaceIntersection of a hit (TriangleIndexed.h:133-172 then
    // Object.h:106-144 with no normal map).
    void fillIsect(int oi, int t, const Ray& objRay, float b0, float b1, Isect* is) const {
        const Object& o = objects[oi];
        const Mesh& m = o.mesh;
        V3 p = lerp3(m.P(t, 0), m.P(t, 1), m.P(t, 2), b0, b1);
        V3 wo = normalize(-objRay.d);
        V3 n;
        if (!m.nrm.empty()) n = lerp3(m.N(t, 0), m.N(t, 1), m.N(t, 2), b0, b1).normalize();
        else n = cross(m.P(t, 1) - m.P(t, 0), m.P(t, 2) - m.P(t, 0)).normalize();
        V3 sn = faceForward(n, wo);
        const uint32_t mi = o.matOf(t);
        const int nm = mi < matNormalMap.size() ? matNormalMap[mi] : -1;
        if (nm >= 0) {
            // TriangleIndexed.h:139-168 (texCoord, tangent, bitangent), then the
            // normal map in object space, Object.h:59-68
            V3 uv(0.0f), tg(0.0f), bt(0.0f);
            if (!m.uv.empty()) {
                auto UV = [&](int k) { int v = m.idx[t * 3 + k]; return V3(m.uv[2 * v], m.uv[2 * v + 1], 0.0f); };
                uv = lerp3(UV(0), UV(1), UV(2), b0, b1);
            }
            if (!m.tan.empty()) {
                auto T = [&](int k) { return m.tan[m.idx[t * 3 + k]]; };
                auto B = [&](int k) { return m.bitan[m.idx[t * 3 + k]]; };
                tg = lerp3(T(0), T(1), T(2), b0, b1).normalize();
                bt = lerp3(B(0), B(1), B(2), b0, b1).normalize();
            }
            V3 c = textures[nm].bilinear(uv.x, uv.y) * 2.0f - V3(1.0f);
            sn = (c.z * sn + c.x * tg + c.y * bt).normalize();
        }
        is->p = o.xf(p);
        is->n = o.xf.asNormal(n);
        is->wo = o.xf.asNormal(wo);
        is->sn = o.xf.asNormal(sn);
        is->obj = oi; is->tri = t; is->prim = o.primBase + (uint32_t)t;
        is->mat = &materials[o.matOf(t)];
    }

    // Scene::intersect (Scene.h:30) -> TLAS -> Object::intersect (object-space
    // ray, Object.h:106-107) -> BLAS -> triangle test, with the (t, prim) tie
    // rule.  Returns the hit and its t.
    bool intersect(const Ray& ray, Isect* is) const {
        float tCull = ray.tMax;
        float bestT = Infinity; uint32_t bestPrim = 0xFFFFFFFFu;
        int bestObj = -1, bestTri = -1; float bb0 = 0, bb1 = 0;
        tlas.traverse(ray, tCull, [&](int oi, float& tc) -> bool {
            const Object& o = objects[oi];
            Ray r;
            r.o = o.xf.inversePoint(ray.o);                                 // Transform.h:155-157
            r.d = o.xf.asVectorInverse(ray.d);
            r.tMax = ray.tMax;
            float tcl = tc;
            o.blas.traverse(r, tcl, [&](int t, float& tc2) -> bool {
                float th, b0, b1;
                if (triTest(o.mesh.P(t, 0), o.mesh.P(t, 1), o.mesh.P(t, 2), r, ray.tMax, &th, &b0, &b1)) {
                    uint32_t gp = o.primBase + (uint32_t)t;
                    if (th < bestT || (th == bestT && gp < bestPrim)) {
                        bestT = th; bestPrim = gp; bestObj = oi; bestTri = t; bb0 = b0; bb1 = b1;
                        tc2 = th;
                    }
                }
                return false;
            });
            if (bestT < tc) tc = bestT;
            return false;
        });
        if (bestObj < 0) return false;
        Ray r;
        const Object& o = objects[bestObj];
        r.o = o.xf.inversePoint(ray.o);
        r.d = o.xf.asVectorInverse(ray.d);
        fillIsect(bestObj, bestTri, r, bb0, bb1, is);
        is->t = bestT;
        is->b0 = bb0;
        is->b1 = bb1;
        return true;
    }

    // Scene::intersectAny (Scene.h:34-36) -> Geometry::intersectAny per
    // triangle (Geometry.h:19-23): any triangle passing against tMax.
    bool intersectAny(const Ray& ray) const {
        float tCull = ray.tMax;
        return tlas.traverse(ray, tCull, [&](int oi, float&) -> bool {
            const Object& o = objects[oi];
            Ray r;
            r.o = o.xf.inversePoint(ray.o);
            r.d = o.xf.asVectorInverse(ray.d);
            r.tMax = ray.tMax;
            float tcl = ray.tMax;
            return o.blas.traverse(r, tcl, [&](int t, float&) -> bool {
                float th, b0, b1;
                return triTest(o.mesh.P(t, 0), o.mesh.P(t, 1), o.mesh.P(t, 2), r, ray.tMax, &th, &b0, &b1);
            });
        });
    }

    // Brute force over world-space triangles (test oracle for the BVH).
    bool intersectBrute(const Ray& ray, float* tOut, uint32_t* primOut, float* b0Out, float* b1Out) const {
        float bestT = Infinity; uint32_t bestPrim = 0xFFFFFFFFu; float bb0 = 0, bb1 = 0;
        for (uint32_t p = 0; p < nprim; ++p) {
            float th, b0, b1;
            if (triTest(worldTri[p][0], worldTri[p][1], worldTri[p][2], ray, ray.tMax, &th, &b0, &b1)) {
                if (th < bestT || (th == bestT && p < bestPrim)) { bestT = th; bestPrim = p; bb0 = b0; bb1 = b1; }
            }
        }
        if (bestPrim == 0xFFFFFFFFu) return false;
        *tOut = bestT; *primOut = bestPrim; *b0Out = bb0; *b1Out = bb1;
        return true;
    }
    bool intersectAnyBrute(const Ray& ray) const {
        for (uint32_t p = 0; p < nprim; ++p) {
            float th, b0, b1;
            if (triTest(worldTri[p][0], worldTri[p][1], worldTri[p][2], ray, ray.tMax, &th, &b0, &b1)) return true;
        }
        return false;
    }

    // Light pdf of a point on triangle t of emissive object oi
    // (Scene.h:49-55 -> Object.h:169-172 -> TriangleMesh.h:232-235 ->
    // TriangleIndexed.h:211-213).
    float lightPdfArea(int oi, int t) const {
        const Object& o = objects[oi];
        if (!materials[o.material].emissive()) return 0.0f;   // not in the light list
        float pdf = 1.0f / o.mesh.triArea(t);
        if (XO_SLIP != 5) pdf = pdf / (float)o.mesh.ntri();
        if (XO_SLIP != 6) pdf = pdf / o.areaScaling;
        return XO_SLIP == 4 ? pdf : pdf / (float)lights.size();
    }

    // Scene::sampleSurface (Scene.h:40-47) -> Object::sampleSurface
    // (Object.h:153-167) -> TriangleMesh::sampleSurface (TriangleMesh.h:216-230)
    // -> TriangleIndexed::sampleSurface (TriangleIndexed.h:179-209).
    SampledLight sampleLight(const PathRng& rng, uint32_t slot, bool uniformBary) const {
        SampledLight s;
        int oi = lights[rng.bits(slot, DIM_LIGHT_OBJ) % (uint32_t)lights.size()];
        const Object& o = objects[oi];
        int t = (int)(rng.bits(slot, DIM_LIGHT_TRI) % (uint32_t)o.mesh.ntri());
        float t0, t1;
        if (uniformBary) {
            // Uniform area sampling (corrected Q5): sqrt parameterisation.
            float su = sqrtf(rng.uni(slot, DIM_BARY0));
            t0 = 1.0f - su;
            t1 = rng.uni(slot, DIM_BARY1) * su;
        } else {
            float a = rng.uni(slot, DIM_BARY0), b = rng.uni(slot, DIM_BARY1), c = rng.uni(slot, DIM_BARY2);
            float sum = a + b + c;
            if (sum == 0.0f) { a = b = c = 1.0f; sum = 3.0f; }  // probability 2^-72
            float inv = 1.0f / sum;
            t0 = a * inv; t1 = b * inv;
        }
        const Mesh& m = o.mesh;
        V3 p = lerp3(m.P(t, 0), m.P(t, 1), m.P(t, 2), t0, t1);
        V3 n = cross(m.P(t, 1) - m.P(t, 0), m.P(t, 2) - m.P(t, 0)).normalize();
        V3 sn = m.nrm.empty() ? n : lerp3(m.N(t, 0), m.N(t, 1), m.N(t, 2), t0, t1).normalize();
        s.p = o.xf(p);
        s.n = o.xf.asNormal(n);
        s.sn = o.xf.asNormal(sn);
        s.obj = oi; s.tri = t;
        s.pdf = lightPdfArea(oi, t);
        return s;
    }
};

// ---------------------------------------------------------------- camera / film
struct Camera {                                                            // Camera.h:17-55
    Xf camToWorld = Xf::identity();
    float tanX = 1, tanY = 1;   // tanf(fov/2), tanf(fov/2 * aspect) (Q9)
    Ray generate(float fx, float fy) const {
        Ray r;
        r.o = V3(0, 0, 0);
        r.d = V3(tanX * (fx - 0.5f) * 2, tanY * (fy - 0.5f) * 2, 1).normalize();
        Ray w;
        w.o = camToWorld(r.o);                                              // Transform.h:147-149
        w.d = camToWorld.asVector(r.d);
        w.tMax = r.tMax;
        return w;
    }
};

struct RenderParams {
    int width = 0, height = 0, spp = 1, sppOffset = 0, maxDepth = 0;
    uint32_t seed = 1, stream = 0;
    int integrator = 0;          // 0 = StandardPathTracer, 1 = NaivePathTracer
    uint32_t flags = 0;
    int rrStart = 5;             // PathTracer.h:255
    float rrProb = 0.9f;         // PathTracer.h:256
    float eps = 0.00001f;        // PathTracer.h:254
    float shadowMargin = 0.0001f;// PathTracer.h:257
};
enum : uint32_t { FLAG_UNIFORM_BARY = 1u, FLAG_ENV_NEE = 2u };

// RenderTargetTile::GenerateFilmPosition (RenderTarget.h:68-80).
inline void filmPosition(int x, int y, int W, int H, float ux, float uy, float* fx, float* fy) {
    float px = (float)x, py = (float)y;
    px /= (float)W;
    py /= (float)H;
    if (XO_SLIP != 1) py = 1.0f - py;
    const float half = XO_SLIP == 2 ? 0.0f : 0.5f;
    px += (ux - half) / (float)W;
    py += (uy - half) / (float)H;
    *fx = px; *fy = py;
}

struct PathStats { uint64_t extRays = 0, shadowRays = 0; };

// Environment NEE strategy (extension): probability of picking the
// environment over area lights, and the mixture pdf of a direction.
struct EnvNee {
    const Env& env;
    bool on;
    float pSelect;     // probability of sampling the environment
    float pSun;        // within env: probability of the sun cone
    EnvNee(const Env& e, bool enabled, bool haveArea) : env(e) {
        bool sky = !e.sky.isZero() || e.mapKind != 0;
        on = enabled && e.hasSky && (e.hasSun() || sky);
        pSelect = on ? (haveArea ? 0.5f : 1.0f) : 0.0f;
        pSun = e.hasSun() ? (sky ? 0.5f : 1.0f) : 0.0f;
    }
    float sunPdf() const { return 1.0f / (2.0f * Pi * env.sunOmc); }
    // solid-angle pdf of the env mixture (before pSelect): sun cone, then the
    // lat-long map distribution or a cosine lobe about n
    float pdf(const V3& dUnit, const V3& n) const {
        float p = 0.0f;
        if (pSun > 0.0f && env.inSun(dUnit)) p += pSun * sunPdf();
        if (pSun < 1.0f) {
            if (env.mapKind == 2) p += (1.0f - pSun) * env.mapPdf(dUnit);
            else p += (1.0f - pSun) * (float)((double)clampPositive(dot(dUnit, n)) / M_PI);
        }
        return p;
    }
    V3 sample(const V3& n, float usel, float u1, float u2) const {
        if (usel < pSun) {
            // uniform cone: cos = 1 - u*omc, sin^2 = u*omc*(2 - u*omc)
            float uo = u1 * env.sunOmc;
            float cosT = 1.0f - uo;
            float sinT = safeSqrt(uo * (2.0f - uo));
            float phi = 2.0f * Pi * u2;
            Basis b(env.sunDir);
            float sphi, cphi;
            sinCos(phi, &sphi, &cphi);
            return b.toGlobal(cosT, sinT * cphi, sinT * sphi);
        }
        if (env.mapKind == 2) {
            float pdfMap;
            return env.sampleMap(u1, u2, &pdfMap);
        }
        return sampleCosHemisphere(n, u1, u2);
    }
};

// StandardPathTracer::eval (PathTracer.h:117-251) from a camera ray.
inline V3 evalStandard(const Scene& sc, const PathRng& rng, const Ray& camRay, const RenderParams& rp, PathStats* st) {
    V3 radiance(0.0f), weight(1.0f);
    Ray cur = camRay;
    Isect isect, next;
    int pathLength = 1;
    const bool canSample = !sc.lights.empty();
    EnvNee envn(sc.env, (rp.flags & FLAG_ENV_NEE) != 0, canSample);
    const float pArea = canSample ? 1.0f - envn.pSelect : 0.0f;
    st->extRays++;
    if (!sc.intersect(cur, &isect)) {
        if (sc.env.hasSky) radiance += weight * sc.env.radiance(cur.d);
        return radiance;
    }
    ++pathLength;
    if (isect.mat->emissive()) radiance += weight * emission(*isect.mat, -cur.d, isect.n);
    uint32_t slot = 1;
    while (true) {
        if (rp.maxDepth > 0 && (int)slot > rp.maxDepth) break;   // added bounce cap
        if (pathLength > rp.rrStart) {
            if (rng.uni(slot, DIM_RR) >= rp.rrProb) break;
            if (XO_SLIP != 3) weight /= rp.rrProb;
        }
        bool nextSampled = true;
        float pdfB;
        V3 wi;
        const Walk walkS{&rng, slot, DIM_WALK_SAMPLE}, walkE{&rng, slot, DIM_WALK_EVAL};
        V3 meval = evalAndSample(*isect.mat, isect, rng.uni(slot, DIM_BSDF_SEL), rng.uni(slot, DIM_BSDF_U1),
                                 rng.uni(slot, DIM_BSDF_U2), &wi, &pdfB, &walkS);
        if (!meval.isZero()) {
            cur.d = wi;
            cur.o = isect.p + rp.eps * cur.d;
            cur.tMax = Infinity;
            st->extRays++;
            if (sc.intersect(cur, &next)) {
                if (next.mat->emissive()) {
                    float mis;
                    if (pdfB >= 0.0f && canSample) {
                        float pl = sc.lightPdfArea(next.obj, next.tri) * pArea;
                        float cosL = fabsf(dot(cur.d, next.sn));
                        float distSq = next.t * next.t;
                        pl *= distSq / cosL;
                        mis = (pdfB * pdfB) / (pdfB * pdfB + pl * pl);
                    } else {
                        mis = 1.0f;
                    }
                    if (mis > 0.0f) radiance += weight * mis * meval * emission(*next.mat, -cur.d, next.n);
                }
            } else {
                if (sc.env.hasSky) {
                    V3 Le = sc.env.radiance(cur.d);
                    if (envn.on && pdfB >= 0.0f) {
                        float pe = envn.pSelect * envn.pdf(normalize(cur.d), isect.sn);
                        // power heuristic; a direction the sky sampler cannot produce
                        // (pe = 0) or an unbounded BSDF pdf keeps the full weight
                        float mis = (pe > 0.0f && pdfB < Infinity) ? (pdfB * pdfB) / (pdfB * pdfB + pe * pe) : 1.0f;
                        radiance += weight * mis * meval * Le;
                    } else {
                        radiance += weight * meval * Le;
                    }
                }
                nextSampled = false;
            }
        }
        // NEE (PathTracer.h:202-233), with the environment as an optional
        // second light kind.
        bool sampleEnv = false;
        if (envn.on) sampleEnv = !canSample || rng.uni(slot, DIM_LIGHT_KIND) < envn.pSelect;
        if (sampleEnv) {
            if (pdfB >= 0.0f) {
                V3 d = envn.sample(isect.sn, rng.uni(slot, DIM_ENV_SEL), rng.uni(slot, DIM_ENV_U1), rng.uni(slot, DIM_ENV_U2));
                float pe = envn.pSelect * envn.pdf(d, isect.sn);
                V3 fc = bsdfCos(*isect.mat, isect, d, &walkE);
                float pb = bsdfPdf(*isect.mat, isect, d);
                float mis = (pe * pe) / (pb * pb + pe * pe);
                if (pe > 0.0f && !fc.isZero() && mis > 0.0f) {   // a NaN pdf drops the sample
                    Ray sr;
                    sr.d = d;
                    sr.o = isect.p + rp.eps * sr.d;
                    sr.tMax = Infinity;
                    st->shadowRays++;
                    if (!sc.intersectAny(sr)) radiance += weight * mis * fc * sc.env.radiance(d) / pe;
                }
            }
        } else if (canSample) {
            SampledLight ls = sc.sampleLight(rng, slot, (rp.flags & FLAG_UNIFORM_BARY) != 0);
            float pdfL = ls.pdf * pArea;
            float dist = (ls.p - isect.p).length();
            Ray sr;
            sr.d = (ls.p - isect.p) / dist;
            sr.o = isect.p + rp.eps * sr.d;
            sr.tMax = dist - rp.shadowMargin;
            if (dot(sr.d, ls.n) < 0) {
                st->shadowRays++;
                if (!sc.intersectAny(sr)) {
                    float mis, distSq = dist * dist;
                    if (pdfB >= 0.0f) {
                        float pb = bsdfPdf(*isect.mat, isect, sr.d);
                        float cosL = fabsf(dot(-sr.d, ls.sn));
                        pb *= cosL / distSq;
                        mis = (pdfL * pdfL) / (pb * pb + pdfL * pdfL);
                    } else {
                        mis = 0.0f;
                    }
                    if (mis > 0.0f) {
                        float G = fabsf(dot(-sr.d, ls.sn)) / distSq;
                        const Material& lm = sc.materials[sc.objects[ls.obj].matOf(ls.tri)];
                        radiance += weight * mis * bsdfCos(*isect.mat, isect, sr.d, &walkE) * emission(lm, -sr.d, ls.n) * G / pdfL;
                    }
                }
            }
        }
        if (!nextSampled) break;
        if (meval.isZero()) break;          // weight *= 0 below (PathTracer.h:240-245)
        if (next.mat->emissive()) break;    // Emission scatters nothing (Material.h:400-402)
        isect = next;
        weight *= meval;
        ++pathLength;
        ++slot;
        if (weight.isZero()) break;
    }
    return radiance;
}

// NaivePathTracer::eval (PathTracer.h:50-104).
inline V3 evalNaive(const Scene& sc, const PathRng& rng, const Ray& camRay, const RenderParams& rp, PathStats* st) {
    V3 radiance(0.0f), weight(1.0f);
    Ray cur = camRay;
    Isect is;
    int pathLength = 1;
    while (true) {
        if (rp.maxDepth > 0 && pathLength > rp.maxDepth + 1) break;
        uint32_t slot = (uint32_t)pathLength;
        if (pathLength > rp.rrStart) {
            if (rng.uni(slot, DIM_RR) >= rp.rrProb) break;
            if (XO_SLIP != 3) weight /= rp.rrProb;
        }
        st->extRays++;
        if (sc.intersect(cur, &is)) {
            if (is.mat->emissive()) radiance += weight * emission(*is.mat, -cur.d, is.n);
            float pdf;
            V3 wi;
            const Walk walkS{&rng, slot, DIM_WALK_SAMPLE};
            weight *= evalAndSample(*is.mat, is, rng.uni(slot, DIM_BSDF_SEL), rng.uni(slot, DIM_BSDF_U1),
                                    rng.uni(slot, DIM_BSDF_U2), &wi, &pdf, &walkS);
            if (weight.isZero()) break;
            cur.d = wi;
            cur.o = is.p + rp.eps * cur.d;
            cur.tMax = Infinity;
        } else {
            if (sc.env.hasSky) radiance += clampPositive(weight * sc.env.radiance(cur.d));
            break;
        }
        ++pathLength;
    }
    return radiance;
}

}  // namespace xo
