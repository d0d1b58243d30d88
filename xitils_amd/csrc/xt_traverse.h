// BVH4 traversal and the watertight triangle test on gfx950.
//
// One lane per ray, wave64.  The traversal stack lives in LDS
// (stack[level][lane], conflict-free) for the first LDS_DEPTH entries and
// spills to a per-lane global column beyond that; the host sizes the column
// from the built tree (3 * depth + 1 entries bound any BVH4 traversal), so no
// push can leave its buffer.
#pragma once

#include "xt_device.h"

namespace xt {

// Slab-test slack: the box interval's far end (and the cull distance) is
// widened by 8 ulp-of-1 so float rounding in the slab test can never cull a
// box that holds the exact closest hit (PBRT's 1 + 2*gamma(3) argument).
constexpr float BOX_MARGIN = 1.0f + 0x1p-20f;

struct RayPre {
    f3 o, d;
    float ix, iy, iz;      // 1 / d (Bounds.h:179)
    bool sx_neg, sy_neg, sz_neg;
    int px, py, pz;        // float4 index of the near plane per axis: x 0|1, y 2|3, z 4|5
    int kx, ky, kz;        // permutation (TriangleIndexed.h:88-91)
    float shx, shy, shz;   // shear (TriangleIndexed.h:96-98)
};

__device__ __forceinline__ RayPre ray_setup(f3 o, f3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    r.sx_neg = signbit(d.x);
    r.sy_neg = signbit(d.y);
    r.sz_neg = signbit(d.z);
    r.px = r.sx_neg ? 1 : 0;
    r.py = r.sy_neg ? 3 : 2;
    r.pz = r.sz_neg ? 5 : 4;
    int kz = max_dim(mk(fabsf(d.x), fabsf(d.y), fabsf(d.z)));
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dx = comp(d, kx), dy = comp(d, ky), dz = comp(d, kz);
    r.shx = -dx / dz;
    r.shy = -dy / dz;
    r.shz = 1.0f / dz;
    return r;
}

// TriangleIndexed::intersect (TriangleIndexed.h:80-131), same operation order.
// Passing means the reference would accept the triangle against tmax.
__device__ __forceinline__ bool tri_test(f3 p0, f3 p1, f3 p2, const RayPre& r, float tmax, float* t, float* b0,
                                         float* b1) {
    f3 a = p0 - r.o, b = p1 - r.o, c = p2 - r.o;
    float ax = comp(a, r.kx), ay = comp(a, r.ky), az = comp(a, r.kz);
    float bx = comp(b, r.kx), by = comp(b, r.ky), bz = comp(b, r.kz);
    float cx = comp(c, r.kx), cy = comp(c, r.ky), cz = comp(c, r.kz);
    ax += r.shx * az;
    ay += r.shy * az;
    bx += r.shx * bz;
    by += r.shy * bz;
    cx += r.shx * cz;
    cy += r.shy * cz;
    float e0 = bx * cy - by * cx;
    float e1 = cx * ay - cy * ax;
    float e2 = ax * by - ay * bx;
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;
    az *= r.shz;
    bz *= r.shz;
    cz *= r.shz;
    float ts = e0 * az + e1 * bz + e2 * cz;
    if (det < 0 && (ts >= 0 || ts < tmax * det)) return false;
    if (det > 0 && (ts <= 0 || ts > tmax * det)) return false;
    float inv = 1 / det;
    *b0 = e0 * inv;
    *b1 = e1 * inv;
    *t = ts * inv;
    return true;
}

struct TraceResult {
    float t;
    uint32_t slot;
    uint32_t prim;
    float b0, b1;
};


// Four child slab tests; returns entry distances (inf = miss).  The near and
// far planes of each axis are picked by address (the ray's direction signs),
// so no per-child selects are needed.
__device__ __forceinline__ void node_slabs(const float4* __restrict__ nd, const RayPre& r, float cull, float tn[4]) {
    float4 nx = nd[r.px], fx = nd[r.px ^ 1];
    float4 ny = nd[r.py], fy = nd[r.py ^ 1];
    float4 nz = nd[r.pz], fz = nd[r.pz ^ 1];
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float t0x = (nxa[k] - r.o.x) * r.ix, t1x = (fxa[k] - r.o.x) * r.ix;
        float t0y = (nya[k] - r.o.y) * r.iy, t1y = (fya[k] - r.o.y) * r.iy;
        float t0z = (nza[k] - r.o.z) * r.iz, t1z = (fza[k] - r.o.z) * r.iz;
        float tmin = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
        float tmax = fminf(fminf(t1x, t1y), fminf(t1z, cull)) * BOX_MARGIN;
        tn[k] = tmin <= tmax ? tmin : INFINITY;
    }
}

// Fetch counters of the counting build (algorithmic traffic accounting).
struct VisitCount {
    uint32_t nodes = 0, tris = 0;
};

// ---------------------------------------------------------------- persistent traversal
//
// A wave owns the contiguous ray range [begin, end) and keeps its 64 lanes
// busy: whenever at least REFILL_IDLE lanes have finished their ray, the idle
// lanes take the next rays of the range (a wave-uniform cursor, no atomics).
// Each loop iteration advances every busy lane by one node visit or one leaf.
constexpr int REFILL_IDLE = 16;

struct Trav {
    int node, sp;
    float best_t, b0, b1, cull, tmax;
    uint32_t best_prim, best_slot;
};

__device__ __forceinline__ void trav_init(Trav& t, float tmax) {
    t.node = 0;
    t.sp = 0;
    t.best_t = INFINITY;
    t.b0 = t.b1 = 0.0f;
    t.cull = tmax;
    t.tmax = tmax;
    t.best_prim = INVALID;
    t.best_slot = INVALID;
}

// One traversal step; returns true when the ray is finished (for ANY, *occ
// tells whether a blocker was found).
template <bool ANY, int LDS_DEPTH, bool COUNT>
__device__ __forceinline__ bool trav_step(const DevScene& sc, const RayPre& r, Trav& t, uint32_t* lds, int lds_stride,
                                          uint32_t* ovf, size_t ovf_stride, VisitCount& vc, bool* occ) {
    if (t.node >= 0) {
        const float4* nd = sc.nodes + (size_t)t.node * NODE_F4;
        if (COUNT) vc.nodes++;
        float tn[4];
        node_slabs(nd, r, t.cull, tn);
        float4 chf = nd[6];
        int ch[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z), __float_as_int(chf.w)};
#define XT_CSWAP(i, j)                                   \
    if (tn[j] < tn[i]) {                                 \
        float tt = tn[i]; tn[i] = tn[j]; tn[j] = tt;     \
        int cc = ch[i]; ch[i] = ch[j]; ch[j] = cc;       \
    }
        XT_CSWAP(0, 1) XT_CSWAP(2, 3) XT_CSWAP(0, 2) XT_CSWAP(1, 3) XT_CSWAP(1, 2)
#undef XT_CSWAP
        int nhit = (tn[0] < INFINITY) + (tn[1] < INFINITY) + (tn[2] < INFINITY) + (tn[3] < INFINITY);
        if (nhit > 0) {
            for (int k = nhit - 1; k >= 1; --k) {
                if (t.sp < LDS_DEPTH) lds[t.sp * lds_stride] = (uint32_t)ch[k];
                else ovf[(size_t)(t.sp - LDS_DEPTH) * ovf_stride] = (uint32_t)ch[k];
                ++t.sp;
            }
            t.node = ch[0];
            return false;
        }
    } else {
        // one triangle per step: a divergent wave then pays for one node
        // visit plus one triangle test per iteration, not for a whole leaf
        uint32_t code = (uint32_t)(~t.node);
        uint32_t first = code >> 3, cnt = (code & 7u) + 1u;
        if (COUNT) vc.tris++;
        const float4* tr = sc.tris + (size_t)first * TRI_F4;
        float4 v0 = tr[0], v1 = tr[1], v2 = tr[2];
        float th, b0, b1;
        if (tri_test(xyz(v0), xyz(v1), xyz(v2), r, t.tmax, &th, &b0, &b1)) {
            if (ANY) {
                *occ = true;
                return true;
            }
            uint32_t prim = __float_as_uint(v0.w);
            if (th < t.best_t || (th == t.best_t && prim < t.best_prim)) {
                t.best_t = th;
                t.best_prim = prim;
                t.best_slot = first;
                t.b0 = b0;
                t.b1 = b1;
                t.cull = fminf(t.tmax, th);
            }
        }
        if (cnt > 1) {  // rest of the leaf: first + 1, count - 1
            t.node = ~(int)(((first + 1u) << 3) | (cnt - 2u));
            return false;
        }
    }
    if (t.sp == 0) return true;
    --t.sp;
    t.node = (int)(t.sp < LDS_DEPTH ? lds[t.sp * lds_stride] : ovf[(size_t)(t.sp - LDS_DEPTH) * ovf_stride]);
    return false;
}

}  // namespace xt
