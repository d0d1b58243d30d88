// Host-side structural checks of the BVH builder (xitils_amd/csrc/bvh.cpp):
// every triangle lands in exactly one leaf, every child box contains its
// subtree, leaf sizes respect the 3-bit count, and the reported depth bounds
// the traversal stack.  Prints "OK <nodes> <depth>" or the first violation.
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
    int max_level = 0;
};

static bool inside(const float* lo, const float* hi, const float* p) {
    for (int k = 0; k < 3; ++k)
        if (p[k] < lo[k] || p[k] > hi[k]) return false;
    return true;
}

static int walk(Ctx& c, int code, const float* lo, const float* hi, int level) {
    if (code < 0) {
        unsigned u = ~(unsigned)code;
        unsigned first = u >> 3, cnt = (u & 7u) + 1;
        for (unsigned k = 0; k < cnt; ++k) {
            unsigned prim = c.b.slot_prim[first + k];
            c.seen[prim]++;
            for (int v = 0; v < 3; ++v)
                if (!inside(lo, hi, c.t[prim].p[v])) { printf("FAIL leaf box\n"); return 1; }
        }
        return 0;
    }
    if (level > c.max_level) c.max_level = level;
    const int W = BVH_WIDTH;
    const float* n = &c.b.nodes[(size_t)code * BVH_NODE_FLOATS];
    for (int k = 0; k < W; ++k) {
        float clo[3] = {n[0 * W + k], n[2 * W + k], n[4 * W + k]};
        float chi[3] = {n[1 * W + k], n[3 * W + k], n[5 * W + k]};
        int ch;
        memcpy(&ch, &n[6 * W + k], 4);
        if (clo[0] > chi[0]) continue;  // empty slot
        if (lo && (!inside(lo, hi, clo) || !inside(lo, hi, chi))) { printf("FAIL child box\n"); return 1; }
        if (walk(c, ch, clo, chi, level + 1)) return 1;
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
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) t[i].p[v][k] = c[k] + (mode == 1 ? 0.0f : 0.01f * (U(rng) - 0.5f));
    }
    BVHBuildOptions opt;
    BVHResult b = build_bvh(t, opt);
    Ctx c{b, t, std::vector<int>(n, 0)};
    if ((int)b.slot_prim.size() != n) { printf("FAIL slot count\n"); return 1; }
    if (walk(c, 0, nullptr, nullptr, 1)) return 1;
    for (int i = 0; i < n; ++i)
        if (c.seen[i] != 1) { printf("FAIL prim %d seen %d\n", i, c.seen[i]); return 1; }
    if (c.max_level != (int)b.depth) { printf("FAIL depth %d vs %u\n", c.max_level, b.depth); return 1; }
    if (b.max_leaf > 8) { printf("FAIL leaf size\n"); return 1; }
    printf("OK %u %u\n", b.num_nodes, b.depth);
    return 0;
}
