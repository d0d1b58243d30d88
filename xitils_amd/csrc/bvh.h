// Host BVH builder: binned SAH binary tree collapsed to an XT_BVH_WIDTH-wide
// BVH (bvh_layout.h) with leaf-ordered triangles (layout in xt_device.h).
//
// Replaces the reference's build (AccelerationStructure.h:163-291: sort +
// 8 equal-count buckets, "SAH" on 2xyz, one primitive per leaf).  Hit results
// do not depend on the tree: traversal returns the lexicographic minimum of
// (t, prim id) over triangles passing the reference test.
#pragma once

#include <cstdint>
#include <vector>

#include "bvh_layout.h"

namespace xt {

struct BuildTri {
    float p[3][3];   // world-space vertices
};

struct BVHResult {
    std::vector<float> nodes;          // BVH_NODE_WORDS 32-bit words per node (bvh_layout.h)
    std::vector<uint32_t> slot_prim;   // leaf slot -> prim index (input order)
    uint32_t num_nodes = 0;
    uint32_t top_nodes = 0;            // nodes of the first BVH_TOP_LEVELS levels: indices [0, top_nodes), breadth-first
    uint32_t depth = 0;                // levels of nodes (root = 1)
    uint32_t max_leaf = 0;
};

#ifndef XT_BVH_MAX_LEAF
#define XT_BVH_MAX_LEAF 4
#endif
#ifndef XT_BVH_NODE_COST
#define XT_BVH_NODE_COST 1.0f
#endif

struct BVHBuildOptions {
    int bins = 32;
    int max_leaf = XT_BVH_MAX_LEAF;         // <= 8 (3-bit count in the leaf code)
    float cost_node = XT_BVH_NODE_COST;     // SAH traversal cost relative to one triangle test
    int threads = 0;           // 0 = hardware concurrency
};

BVHResult build_bvh(const std::vector<BuildTri>& tris, const BVHBuildOptions& opt);

// Child k of a node as the traversal sees it: the box (decoded exactly in
// double for quantised nodes), the child code and the outward margin the
// builder guarantees beyond the child's exact bounds per axis (0 for both
// layouts: boxes are only rounded outward); false for an empty slot.
bool bvh_child(const BVHResult& b, uint32_t node, int k, double lo[3], double hi[3], int* code,
               double margin[3] = nullptr);

}  // namespace xt
