// Width of the BVH nodes shared by the host builder and the traversal
// kernels: a node holds XT_BVH_WIDTH children as 8 rows of XT_BVH_WIDTH
// floats (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, child codes, reserved), i.e.
// 32 * XT_BVH_WIDTH bytes -- 128 bytes, one L2 line, for 4 children.
// Width 8 also builds and traverses (two L2 lines per node); on the 2M-triangle
// heightfield it fetches 30% fewer nodes per ray but moves more bytes and
// renders 14% slower, so 4 is the default.
#pragma once

#ifndef XT_BVH_WIDTH
#define XT_BVH_WIDTH 4
#endif
static_assert(XT_BVH_WIDTH == 4 || XT_BVH_WIDTH == 8, "BVH nodes are 4 or 8 wide");

namespace xt {
constexpr int BVH_WIDTH = XT_BVH_WIDTH;
}
