// Host-side structural checks of the BVH builder (xitils_amd/csrc/bvh.cpp):
// every triangle lands in exactly one leaf, every child box (as the traversal
// decodes it, bvh_child) contains the exact bounds of its subtree's triangles
// with the builder's guaranteed outward margin, leaf sizes respect the 3-bit
// count, and the reported depth bounds the traversal stack.  Prints "OK
// <nodes> <depth>" or the first violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../xitils_amd/csrc/bvh.h"

using namespace xt;

struct Ctx {
    const BVHResult& b;
    const std::vector<BuildTri>& t;
    std::vector<int> seen;
    std::vector<int> level_of;   // per node, 0 = not reached
    int max_level = 0;
};

// Walks the subtree of `code`; writes the exact bounds of its triangles.
static int walk(Ctx& c, int code, int level, double lo[3], double hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    if (code < 0) {
        unsigned u = ~(unsigned)code;
        unsigned first = u >> 3, cnt = (u & 7u) + 1;
        for (unsigned k = 0; k < cnt; ++k) {
            unsigned prim = c.b.slot_prim[first + k];
            c.seen[prim]++;
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::fmin(lo[a], (double)c.t[prim].p[v][a]);
                    hi[a] = std::fmax(hi[a], (double)c.t[prim].p[v][a]);
                }
        }
        return 0;
    }
    if (level > c.max_level) c.max_level = level;
    if (c.level_of[code]) {
        printf("FAIL node %d reached twice\n", code);
        return 1;
    }
    c.level_of[code] = level;
    for (int k = 0; k < BVH_WIDTH; ++k) {
        double clo[3], chi[3], m[3], slo[3], shi[3];
        int ch;
        if (!bvh_child(c.b, (uint32_t)code, k, clo, chi, &ch, m)) continue;
        if (walk(c, ch, level + 1, slo, shi)) return 1;
        for (int a = 0; a < 3; ++a) {
            if (!(clo[a] <= slo[a] - m[a] && chi[a] >= shi[a] + m[a])) {
                printf("FAIL child box node %d slot %d axis %d: [%.9g, %.9g] margin %.3g vs [%.9g, %.9g]\n", code, k, a,
                       clo[a], chi[a], m[a], slo[a], shi[a]);
                return 1;
            }
            lo[a] = std::fmin(lo[a], slo[a]);
            hi[a] = std::fmax(hi[a], shi[a]);
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    int n = argc > 1 ? atoi(argv[1]) : 10000;
    int mode = argc > 2 ? atoi(argv[2]) : 0;
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0, 1);
    std::vector<BuildTri> t(n);
    for (int i = 0; i < n; ++i) {
        float c[3] = {U(rng), U(rng), U(rng)};
        if (mode == 1) c[0] = c[1] = c[2] = 0.5f;          // all centroids equal
        if (mode == 2) c[1] = 0.0f;                        // flat sheet
        // 3: far from the origin (the grid origin's rounding is coarse against the step)
        // 4 / 5: huge / tiny coordinates
        const float off = mode == 3 ? 1000.0f : 0.0f, scale = mode == 4 ? 1e20f : (mode == 5 ? 1e-20f : 1.0f);
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k)
                t[i].p[v][k] = off + scale * (c[k] + (mode == 1 ? 0.0f : 0.01f * (U(rng) - 0.5f)));
    }
    BVHBuildOptions opt;
    BVHResult b = build_bvh(t, opt);
    Ctx c{b, t, std::vector<int>(n, 0), std::vector<int>(b.num_nodes, 0)};
    if ((int)b.slot_prim.size() != n) { printf("FAIL slot count\n"); return 1; }
    double rlo[3], rhi[3];
    if (walk(c, 0, 1, rlo, rhi)) return 1;
    for (int i = 0; i < n; ++i)
        if (c.seen[i] != 1) { printf("FAIL prim %d seen %d\n", i, c.seen[i]); return 1; }
    if (c.max_level != (int)b.depth) { printf("FAIL depth %d vs %u\n", c.max_level, b.depth); return 1; }
    if (b.max_leaf > 8) { printf("FAIL leaf size\n"); return 1; }
    // the top levels come first, breadth-first (the kernels stage [0, top_nodes) in LDS)
    const int top_levels = (BVH_QUANT && BVH_TOP_LEVELS > 0) ? BVH_TOP_LEVELS : 0;
    uint32_t in_top = 0;
    for (uint32_t v = 0; v < b.num_nodes; ++v) {
        if (!c.level_of[v]) { printf("FAIL node %u unreachable\n", v); return 1; }
        if (c.level_of[v] <= top_levels) {
            ++in_top;
            if (v >= b.top_nodes) { printf("FAIL top-level node %u at index >= %u\n", v, b.top_nodes); return 1; }
        }
        if (v > 0 && v < b.top_nodes && c.level_of[v] < c.level_of[v - 1]) { printf("FAIL top levels not breadth-first\n"); return 1; }
    }
    if (in_top != b.top_nodes || (int)b.top_nodes > (BVH_QUANT ? BVH_TOP_MAX_NODES : 0)) {
        printf("FAIL top_nodes %u vs %u\n", b.top_nodes, in_top);
        return 1;
    }
    // quantised empty slots: inverted boxes (bytes 255 / 0 on every axis) whose
    // code is child 0's (the traversal does not mask them), -1 in an all-empty node
    if constexpr (BVH_QUANT) {
        for (uint32_t v = 0; v < b.num_nodes; ++v) {
            uint32_t w[16];
            std::memcpy(w, &b.nodes[(size_t)v * BVH_NODE_WORDS], sizeof w);
            const int nvalid = (int)(w[3] >> 24);
            for (int k = nvalid; k < 4; ++k) {
                bool inverted = true;
                for (int a = 0; a < 3; ++a)
                    inverted = inverted && ((w[8 + 2 * a] >> (8 * k)) & 255u) == 255u &&
                               ((w[9 + 2 * a] >> (8 * k)) & 255u) == 0u;
                if (w[4 + k] != (nvalid > 0 ? w[4] : ~0u) || !inverted) {
                    printf("FAIL empty slot %d of node %u\n", k, v);
                    return 1;
                }
            }
        }
    }
    printf("OK %u %u\n", b.num_nodes, b.depth);
    return 0;
}
