// BVH traversal (BVH_WIDTH-wide nodes) and the watertight triangle test on
// gfx950.
//
// One lane per ray, wave64.  The traversal stack lives in LDS
// (stack[level][lane], conflict-free) for the first LDS_DEPTH entries and
// spills to a per-lane global column beyond that; the host sizes the column
// from the built tree ((W - 1) * depth + 1 entries bound any traversal of a
// W-wide tree), so no push can leave its buffer.
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


// One float of row `row` (0..7) for child k of a node.
__device__ __forceinline__ float node_row(const float4* nd, int row, int k) {
    const float4 v = nd[row * NODE_ROW_F4 + (k >> 2)];
    const int c = k & 3;
    return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

// The node's child slab tests; returns entry distances (inf = miss).  The
// near and far planes of each axis are picked by address (the ray's
// direction signs), so no per-child selects are needed.
__device__ __forceinline__ void node_slabs(const float4* __restrict__ nd, const RayPre& r, float cull,
                                           float tn[BVH_WIDTH]) {
    float4 rows[6][NODE_ROW_F4];
    const int pr[6] = {r.px, r.px ^ 1, r.py, r.py ^ 1, r.pz, r.pz ^ 1};
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
        for (int h = 0; h < NODE_ROW_F4; ++h) rows[q][h] = nd[pr[q] * NODE_ROW_F4 + h];
#pragma unroll
    for (int k = 0; k < BVH_WIDTH; ++k) {
        auto at = [&](int q) {
            const float4 v = rows[q][k >> 2];
            const int c = k & 3;
            return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
        };
        float t0x = (at(0) - r.o.x) * r.ix, t1x = (at(1) - r.o.x) * r.ix;
        float t0y = (at(2) - r.o.y) * r.iy, t1y = (at(3) - r.o.y) * r.iy;
        float t0z = (at(4) - r.o.z) * r.iz, t1z = (at(5) - r.o.z) * r.iz;
        float tmin = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
        float tmax = fminf(fminf(t1x, t1y), fminf(t1z, cull)) * BOX_MARGIN;
        tn[k] = tmin <= tmax ? tmin : INFINITY;
    }
}

// Quantised node (bvh_layout.h): entry distances and child codes of its
// four children from four 16-byte loads.  Each plane is origin + q 2^e; the
// node-wide offset origin - ray origin is computed once per axis, so a plane
// costs a byte convert, an fma and a multiply.  That offset carries one
// rounding (at most 2^-24 of its magnitude) that node_slabs' per-plane
// difference does not, so the near and far offsets are pushed outward by
// 2^-22 of it, signed by the ray direction: every decoded box contains the
// box node_slabs would test, and the remaining roundings are those
// BOX_MARGIN covers there.  A zero direction component gives +-inf (or NaN,
// which fminf / fmaxf ignore, for an offset of exactly zero) as in
// node_slabs.  Planes are picked per axis by the ray's direction sign.
// Empty slots are inverted boxes (bytes 255 / 0) that carry a real child
// code (quantize_node), so they are not masked here.
__device__ __forceinline__ void node_slabs_q(const float4 w0, const float4 w1, const float4 w2, const float4 w3,
                                             const RayPre& r, float cull, float tn[4], int ch[4]) {
    const uint32_t meta = __float_as_uint(w0.w);
    const int ex = (int)(int8_t)(meta & 255u), ey = (int)(int8_t)((meta >> 8) & 255u),
              ez = (int)(int8_t)((meta >> 16) & 255u);
    const uint32_t lx = __float_as_uint(w2.x), hx = __float_as_uint(w2.y);
    const uint32_t ly = __float_as_uint(w2.z), hy = __float_as_uint(w2.w);
    const uint32_t lz = __float_as_uint(w3.x), hz = __float_as_uint(w3.y);
    const uint32_t nx = r.sx_neg ? hx : lx, fx = r.sx_neg ? lx : hx;
    const uint32_t ny = r.sy_neg ? hy : ly, fy = r.sy_neg ? ly : hy;
    const uint32_t nz = r.sz_neg ? hz : lz, fz = r.sz_neg ? lz : hz;
    ch[0] = __float_as_int(w1.x);
    ch[1] = __float_as_int(w1.y);
    ch[2] = __float_as_int(w1.z);
    ch[3] = __float_as_int(w1.w);
    const float stx = ldexpf(1.0f, ex), sty = ldexpf(1.0f, ey), stz = ldexpf(1.0f, ez);
    const float ox = w0.x - r.o.x, oy = w0.y - r.o.y, oz = w0.z - r.o.z;
    const float sx = copysignf(fabsf(ox) * 0x1p-22f, r.ix), sy = copysignf(fabsf(oy) * 0x1p-22f, r.iy),
                sz = copysignf(fabsf(oz) * 0x1p-22f, r.iz);
    const float nox = ox - sx, fox = ox + sx, noy = oy - sy, foy = oy + sy, noz = oz - sz, foz = oz + sz;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        auto q = [&](uint32_t v) { return (float)((v >> (8 * k)) & 255u); };
        const float t0x = fmaf(q(nx), stx, nox) * r.ix, t1x = fmaf(q(fx), stx, fox) * r.ix;
        const float t0y = fmaf(q(ny), sty, noy) * r.iy, t1y = fmaf(q(fy), sty, foy) * r.iy;
        const float t0z = fmaf(q(nz), stz, noz) * r.iz, t1z = fmaf(q(fz), stz, foz) * r.iz;
        const float tmin = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
        const float tmax = fminf(fminf(t1x, t1y), fminf(t1z, cull)) * BOX_MARGIN;
        tn[k] = tmin <= tmax ? tmin : INFINITY;
    }
}

// LDS copy of the tree's top levels (nodes [0, top_nodes), numbered
// breadth-first by the builder): every ray starts there, so the first few
// node visits of each ray are LDS reads instead of global gathers.
constexpr int TOP_F4 = BVH_QUANT && BVH_TOP_MAX_NODES > 0 ? BVH_TOP_MAX_NODES * NODE_F4 : 1;

// Block-wide copy into LDS; every thread of the block must call it.  A build
// without staged levels (XT_BVH_TOP_LEVELS = 0, the default) copies nothing
// and needs no barrier.
__device__ __forceinline__ void stage_top_nodes(float4* top, const DevScene& sc) {
    if constexpr (BVH_QUANT && BVH_TOP_MAX_NODES > 0) {
        const uint32_t n = min(sc.top_nodes, (uint32_t)BVH_TOP_MAX_NODES) * NODE_F4;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) top[i] = sc.nodes[i];
        __syncthreads();
    }
}

// What the traversal reads of a scene: nodes, leaf-ordered triangles and the
// number of LDS-staged top nodes.  DevScene has the same members, so the
// traversal takes either (SV = DevScene for one scene, uniform across the
// grid; SV = TravView per lane when a queue mixes scenes).
struct TravView {
    const float4* nodes;
    const float4* tris;
    uint32_t top_nodes;
};

// Fetch counters of the counting build (algorithmic traffic accounting).
struct VisitCount {
    uint32_t nodes = 0, tris = 0;
};

// ---------------------------------------------------------------- persistent traversal
//
// A wave keeps its 64 lanes busy: whenever at least REFILL_IDLE lanes have
// finished their ray, the idle lanes take the next rays of the wave's chunk
// (RayChunks below).
// Each loop iteration advances every busy lane by one node visit or one leaf.
#ifndef XT_REFILL_IDLE
#define XT_REFILL_IDLE 16
#endif
constexpr int REFILL_IDLE = XT_REFILL_IDLE;   // 8 and 24: 1080p frame 1.2% and 0.3% slower; 32 no better

// Which rays a wave traces.  DYN = false: an even contiguous share per wave.
// DYN = true (large queues): the queue is cut into chunks of a quarter of an
// even share per wave (64..TRAV_CHUNK rays); wave w first takes chunk w,
// then, each time its chunk runs out while lanes wait for rays, the next
// chunk from a cursor in device memory (one atomic per chunk), so waves whose
// rays end early take work from the rest instead of idling to the end of the
// launch.  The last block to finish resets the cursor and its own block
// counter for the next launch that uses them (dyn[0], dyn[1]; zero before the
// first).  The cursor costs some instructions and registers, which small
// queues do not win back, so the host picks DYN from the queue size
// (TRAV_DYN_MIN).  A ray's result does not depend on which wave traces it.
// The chunk cap of 256 rays (4 per lane) measured best on the 1080p frame:
// 1% faster than 512, 2% faster than 128 and 1024.
#ifndef XT_TRAV_CHUNK
#define XT_TRAV_CHUNK 256
#endif
constexpr uint32_t TRAV_CHUNK = XT_TRAV_CHUNK;
struct RayChunks {
    uint32_t next, end;   // this wave's untaken rays [next, end)
    uint32_t base;        // first ray handed out by the cursor
    uint32_t chunk;       // rays per chunk
    uint32_t n;
    bool exhausted;
};
template <bool DYN>
__device__ __forceinline__ RayChunks chunks_first(uint32_t wave, uint32_t nwaves, uint32_t n) {
    RayChunks c;
    if constexpr (!DYN) {
        const uint32_t per = (n + nwaves - 1) / nwaves;
        c.chunk = per;
        c.next = min(wave * per, n);
        c.end = min(c.next + per, n);
        c.base = n;
        c.n = n;
        c.exhausted = true;
        return c;
    }
    const uint64_t quarter = ((uint64_t)n + 4ull * nwaves - 1) / (4ull * nwaves);
    c.chunk = (uint32_t)min<uint64_t>(max<uint64_t>(quarter, 64), TRAV_CHUNK);
    const uint64_t b = (uint64_t)wave * c.chunk;
    c.next = (uint32_t)min<uint64_t>(b, n);
    c.end = (uint32_t)min<uint64_t>(b + c.chunk, n);
    c.base = (uint32_t)min<uint64_t>((uint64_t)nwaves * c.chunk, n);
    c.n = n;
    c.exhausted = c.base >= n;   // no chunk is left for the cursor
    return c;
}
// Wave-uniform: the whole wave calls it, with every lane active.
__device__ __forceinline__ void chunks_refill(RayChunks& c, uint32_t* dyn) {
    uint32_t got = 0;
    if (__lane_id() == 0) got = atomicAdd(&dyn[0], c.chunk);
    const uint64_t g = (uint64_t)__builtin_amdgcn_readfirstlane(got) + c.base;
    if (g >= c.n) {
        c.exhausted = true;
    } else {
        c.next = (uint32_t)g;
        c.end = (uint32_t)min<uint64_t>(g + c.chunk, c.n);
    }
}
// Every thread of the block calls it once its wave has left the loop.
__device__ __forceinline__ void chunks_release(uint32_t* dyn) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&dyn[1], 1u) == gridDim.x - 1) {
            atomicExch(&dyn[0], 0u);
            atomicExch(&dyn[1], 0u);
        }
    }
}


struct Trav {
    int node, sp;
    float best_t, b0, b1, cull, tmax;
    uint32_t best_slot;   // the hit's primitive id is read back from the slot (prim_of_slot)
};

__device__ __forceinline__ void trav_init(Trav& t, float tmax) {
    t.node = 0;
    t.sp = 0;
    t.best_t = INFINITY;
    t.b0 = t.b1 = 0.0f;
    t.cull = tmax;
    t.tmax = tmax;
    t.best_slot = INVALID;
}

// Visits internal node t.node: true when t.node is now the next child to
// visit (the nearest hit one; the other hit children are pushed), false when
// no child was hit.
template <bool ANY, int LDS_DEPTH, bool COUNT, class SV>
__device__ __forceinline__ bool visit_node(const SV& sc, const RayPre& r, Trav& t, const float4* top,
                                           uint32_t* lds, int lds_stride, uint32_t* ovf, size_t ovf_stride,
                                           VisitCount& vc) {
    const float4* nd = sc.nodes + (size_t)t.node * NODE_F4;
    if (COUNT) vc.nodes++;
    float tn[BVH_WIDTH];
    int ch[BVH_WIDTH];
    if constexpr (BVH_QUANT) {
        float4 w0, w1, w2, w3;
        if (BVH_TOP_MAX_NODES > 0 && (uint32_t)t.node < sc.top_nodes) {
            const float4* p = top + t.node * NODE_F4;
            w0 = p[0];
            w1 = p[1];
            w2 = p[2];
            w3 = p[3];
        } else {
            w0 = nd[0];
            w1 = nd[1];
            w2 = nd[2];
            w3 = nd[3];
        }
        node_slabs_q(w0, w1, w2, w3, r, t.cull, tn, ch);
    } else {
        node_slabs(nd, r, t.cull, tn);
#pragma unroll
        for (int k = 0; k < BVH_WIDTH; ++k) ch[k] = __float_as_int(node_row(nd, 6, k));
    }
    if constexpr (ANY) {
        // any hit: the answer does not depend on the visiting order, so
        // children go in node order (the first hit one next) and the
        // sorting network is skipped (5% faster on sun shadow rays)
        int nxt = -1;
        if (BVH_WIDTH == 4 && t.sp + BVH_WIDTH <= LDS_DEPTH) {
            // room in LDS: every child is written, the pushed ones (hit,
            // not next) compacted from the stack top, the rest into the
            // free slot above them -- no per-push branches
            int pos = t.sp;
#pragma unroll
            for (int k = 0; k < BVH_WIDTH; ++k) {
                const bool hit = tn[k] < INFINITY;
                const bool push = hit && nxt != -1;
                if (hit && nxt == -1) nxt = k;
                lds[(push ? pos : t.sp + BVH_WIDTH - 1) * lds_stride] = (uint32_t)ch[k];
                pos += push ? 1 : 0;
            }
            t.sp = pos;
        } else {
#pragma unroll
            for (int k = 0; k < BVH_WIDTH; ++k) {
                if (tn[k] < INFINITY) {
                    if (nxt == -1) {
                        nxt = k;
                    } else {
                        if (t.sp < LDS_DEPTH) lds[t.sp * lds_stride] = (uint32_t)ch[k];
                        else ovf[(size_t)(t.sp - LDS_DEPTH) * ovf_stride] = (uint32_t)ch[k];
                        ++t.sp;
                    }
                }
            }
        }
        if (nxt >= 0) {
            int c = ch[0];
#pragma unroll
            for (int k = 1; k < BVH_WIDTH; ++k) c = nxt == k ? ch[k] : c;
            t.node = c;
            return true;
        }
    } else {
#define XT_CSWAP(i, j)                                   \
    if (tn[j] < tn[i]) {                                 \
        float tt = tn[i]; tn[i] = tn[j]; tn[j] = tt;     \
        int cc = ch[i]; ch[i] = ch[j]; ch[j] = cc;       \
    }
        static_assert(BVH_WIDTH == 4 || BVH_WIDTH == 8, "nodes are 4 or 8 wide");
        if constexpr (BVH_WIDTH == 4) {
            XT_CSWAP(0, 1) XT_CSWAP(2, 3) XT_CSWAP(0, 2) XT_CSWAP(1, 3) XT_CSWAP(1, 2)
        } else {   // Batcher's odd-even merge sort for 8 keys (19 comparators)
            XT_CSWAP(0, 1) XT_CSWAP(2, 3) XT_CSWAP(4, 5) XT_CSWAP(6, 7)
            XT_CSWAP(0, 2) XT_CSWAP(1, 3) XT_CSWAP(4, 6) XT_CSWAP(5, 7)
            XT_CSWAP(1, 2) XT_CSWAP(5, 6)
            XT_CSWAP(0, 4) XT_CSWAP(1, 5) XT_CSWAP(2, 6) XT_CSWAP(3, 7)
            XT_CSWAP(2, 4) XT_CSWAP(3, 5)
            XT_CSWAP(1, 2) XT_CSWAP(3, 4) XT_CSWAP(5, 6)
        }
#undef XT_CSWAP
        int nhit = 0;
#pragma unroll
        for (int k = 0; k < BVH_WIDTH; ++k) nhit += tn[k] < INFINITY;
        if (nhit > 0) {
            if (BVH_WIDTH == 4 && t.sp + BVH_WIDTH - 1 <= LDS_DEPTH) {
                // room in LDS: the far-to-near order ch[nhit-1] .. ch[1] goes
                // to fixed slots sp, sp+1, sp+2 by selects; slots at or above
                // the new top hold leftovers that are never read
                const int v0 = nhit == 4 ? ch[3 % BVH_WIDTH] : (nhit == 3 ? ch[2] : ch[1]);
                const int v1 = nhit == 4 ? ch[2] : ch[1];
                uint32_t* s = lds + t.sp * lds_stride;
                s[0] = (uint32_t)v0;
                s[lds_stride] = (uint32_t)v1;
                s[2 * lds_stride] = (uint32_t)ch[1];
                t.sp += nhit - 1;
                t.node = ch[0];
                return true;
            }
            // far children first; static indices keep ch[] in registers
#pragma unroll
            for (int k = BVH_WIDTH - 1; k >= 1; --k) {
                if (k < nhit) {
                    if (t.sp < LDS_DEPTH) lds[t.sp * lds_stride] = (uint32_t)ch[k];
                    else ovf[(size_t)(t.sp - LDS_DEPTH) * ovf_stride] = (uint32_t)ch[k];
                    ++t.sp;
                }
            }
            t.node = ch[0];
            return true;
        }
    }
    return false;
}

// Primitive id of the triangle in sorted slot `slot` (the w of its first vertex).
template <class SV>
__device__ __forceinline__ uint32_t prim_of_slot(const SV& sc, uint32_t slot) {
    return __float_as_uint(sc.tris[(size_t)slot * TRI_F4].w);
}

// Tests the first triangle of leaf code `leaf` (< 0); returns the code of the
// rest of the leaf, or 0 when the leaf is done (or, for ANY, when *occ was set).
template <bool ANY, bool COUNT, class SV>
__device__ __forceinline__ int visit_tri(const SV& sc, const RayPre& r, Trav& t, int leaf, VisitCount& vc,
                                         bool* occ) {
    uint32_t code = (uint32_t)(~leaf);
    uint32_t first = code >> 3, cnt = (code & 7u) + 1u;
    if (COUNT) vc.tris++;
    const float4* tr = sc.tris + (size_t)first * TRI_F4;
    float4 v0 = tr[0], v1 = tr[1], v2 = tr[2];
    float th, b0, b1;
    if (tri_test(xyz(v0), xyz(v1), xyz(v2), r, t.tmax, &th, &b0, &b1)) {
        if (ANY) {
            *occ = true;
            return 0;
        }
        // equal distances go to the lower primitive id; the current best's id is
        // read back from its slot only then (exact ties are rare), which keeps
        // it out of the registers the traversal loop holds
        if (th < t.best_t ||
            (th == t.best_t && (t.best_slot == INVALID || __float_as_uint(v0.w) < prim_of_slot(sc, t.best_slot)))) {
            t.best_t = th;
            t.best_slot = first;
            t.b0 = b0;
            t.b1 = b1;
            t.cull = fminf(t.tmax, th);
        }
    }
    // rest of the leaf: first + 1, count - 1
    return cnt > 1 ? ~(int)(((first + 1u) << 3) | (cnt - 2u)) : 0;
}

// Pops the next entry into t.node; false when the stack is empty.
template <int LDS_DEPTH>
__device__ __forceinline__ bool trav_pop(Trav& t, const uint32_t* lds, int lds_stride, const uint32_t* ovf,
                                         size_t ovf_stride) {
    if (t.sp == 0) return false;
    --t.sp;
    t.node = (int)(t.sp < LDS_DEPTH ? lds[t.sp * lds_stride] : ovf[(size_t)(t.sp - LDS_DEPTH) * ovf_stride]);
    return true;
}

// One traversal step of a lane (`active`: the lane holds a ray; the call
// sites make it for every lane); returns true when the ray is finished (for ANY,
// *occ tells whether a blocker was found).  A step visits the lane's node,
// and if that leaves the lane at a leaf (the nearest hit child is one, or
// the popped entry is), tests the leaf's next triangle in the same step, so
// the lanes that reach leaves share the triangle code with those already in
// leaves.  The visits and their order per ray do not change (identical hits
// and fetch counts), only how they are grouped into wave iterations.
#ifndef XT_TRAV_FUSED
#define XT_TRAV_FUSED 1
#endif
template <bool ANY, int LDS_DEPTH, bool COUNT, class SV>
__device__ __forceinline__ bool trav_step(bool active, const SV& sc, const RayPre& r, Trav& t, const float4* top,
                                          uint32_t* lds, int lds_stride, uint32_t* ovf, size_t ovf_stride,
                                          VisitCount& vc, bool* occ) {
#if XT_TRAV_FUSED
    bool done = false;
    if (active && t.node >= 0) {
        if (!visit_node<ANY, LDS_DEPTH, COUNT>(sc, r, t, top, lds, lds_stride, ovf, ovf_stride, vc) &&
            !trav_pop<LDS_DEPTH>(t, lds, lds_stride, ovf, ovf_stride))
            done = true;
    }
    if (active && !done && t.node < 0) {
        const int rest = visit_tri<ANY, COUNT>(sc, r, t, t.node, vc, occ);
        if (ANY && *occ) return true;
        if (rest) t.node = rest;
        else done = !trav_pop<LDS_DEPTH>(t, lds, lds_stride, ovf, ovf_stride);
    }
    return done;
#else
    if (!active) return false;
    if (t.node >= 0) {
        if (visit_node<ANY, LDS_DEPTH, COUNT>(sc, r, t, top, lds, lds_stride, ovf, ovf_stride, vc)) return false;
    } else {
        // one triangle per step: a divergent wave then pays for one node
        // visit plus one triangle test per iteration, not for a whole leaf
        const int rest = visit_tri<ANY, COUNT>(sc, r, t, t.node, vc, occ);
        if (ANY && *occ) return true;
        if (rest) {
            t.node = rest;
            return false;
        }
    }
    return !trav_pop<LDS_DEPTH>(t, lds, lds_stride, ovf, ovf_stride);
#endif
}

}  // namespace xt
