// BVH node layout shared by the host builder and the traversal kernels.
//
// XT_BVH_QUANT = 1 (default, 4-wide only): 64-byte nodes, 16 32-bit words
//   w0..2   origin (float) of the node's quantisation grid
//   w3      bytes 0..2: signed exponents e_x, e_y, e_z (grid step 2^e per
//           axis); byte 3: number of valid children (valid slots first)
//   w4..7   child codes (>= 0 node index, < 0 leaf ~(first << 3 | count - 1))
//   w8..13  planes lo_x, hi_x, lo_y, hi_y, lo_z, hi_z: byte k = child k's
//           plane in grid steps above the origin
//   w14..15 zero
//   Child k's box is [origin + lo_k 2^e, origin + hi_k 2^e] per axis,
//   rounded outward to the grid on the host.  The origin is a multiple of
//   the step and the step at least 2^-22 of the node's magnitude, so every
//   grid plane is a float: the traversal decodes it exactly (one fma) and
//   runs the float layout's slab test on it.  Four 16-byte loads per node
//   visit.
// XT_BVH_QUANT = 0: float nodes of XT_BVH_WIDTH (4 or 8) children as 8 rows
//   of XT_BVH_WIDTH floats (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, child codes,
//   reserved) -- 128 bytes for 4 children, 256 for 8.
//
// Measured on the 1080p height field: 8-wide float nodes fetch 30% fewer
// nodes per ray than 4-wide ones but render 14% slower (more bytes, more
// registers); the traversal's throughput follows the number of load
// instructions per node visit more than its arithmetic.
#pragma once

#ifndef XT_BVH_WIDTH
#define XT_BVH_WIDTH 4
#endif
#ifndef XT_BVH_QUANT
#define XT_BVH_QUANT 1
#endif
// Levels of the tree (root = 1) numbered breadth-first at the front of the
// node array and staged in LDS by the traversal kernels (quantised layout).
// Off by default: on the 1080p height field staging 3, 4 or 5 levels made
// the frame 2%, 2% and 16% slower -- the top nodes already hit in L1, and a
// wave whose lanes straddle the boundary issues both the LDS and the global
// loads (5 levels, 21.8 KB, also costs occupancy).
#ifndef XT_BVH_TOP_LEVELS
#define XT_BVH_TOP_LEVELS 0
#endif
static_assert(XT_BVH_WIDTH == 4 || XT_BVH_WIDTH == 8, "BVH nodes are 4 or 8 wide");
static_assert(!XT_BVH_QUANT || XT_BVH_WIDTH == 4, "quantised nodes are 4 wide");

namespace xt {
constexpr int BVH_WIDTH = XT_BVH_WIDTH;
constexpr bool BVH_QUANT = XT_BVH_QUANT != 0;
constexpr int BVH_FLOAT_NODE_WORDS = 8 * BVH_WIDTH;                    // float layout
constexpr int BVH_NODE_WORDS = BVH_QUANT ? 16 : BVH_FLOAT_NODE_WORDS;  // layout in HBM
constexpr int BVH_TOP_LEVELS = XT_BVH_TOP_LEVELS;
// nodes in the top levels of a 4-wide tree: 1 + 4 + ... + 4^(L-1)
constexpr int BVH_TOP_MAX_NODES = BVH_TOP_LEVELS <= 0 ? 0 : ((1 << (2 * BVH_TOP_LEVELS)) - 1) / 3;
}
