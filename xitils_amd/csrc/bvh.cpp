// Host binned-SAH BVH build, collapsed to a BVH_WIDTH-wide tree (see bvh.h).
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace xt {

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Box box;
    int left = -1, right = -1;  // left < 0: leaf
    uint32_t first = 0, count = 0;
};

struct Builder {
    const BVHBuildOptions& opt;
    std::vector<Box> pbox;
    std::vector<float> cen;  // 3 per prim
    std::vector<uint32_t> idx;
    std::vector<BNode> nodes;
    std::atomic<uint32_t> next{0};
    std::atomic<int> spare_threads{0};

    explicit Builder(const BVHBuildOptions& o) : opt(o) {}

    uint32_t alloc() { return next.fetch_add(1); }

    // Builds the subtree over idx[b, e) into node ni.
    void build(uint32_t ni, uint32_t b, uint32_t e) {
        BNode& nd = nodes[ni];
        Box box, cb;
        for (uint32_t i = b; i < e; ++i) {
            box.grow(pbox[idx[i]]);
            cb.grow(&cen[3 * idx[i]]);
        }
        nd.box = box;
        const uint32_t n = e - b;
        auto make_leaf = [&]() {
            nd.left = nd.right = -1;
            nd.first = b;
            nd.count = n;
        };
        if (n <= 1) {
            make_leaf();
            return;
        }
        const int B = opt.bins;
        float best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        const float parent_area = box.area();
        for (int axis = 0; axis < 3; ++axis) {
            float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0f)) continue;
            float scale = (float)B / ext;
            std::vector<Box> bb(B);
            std::vector<uint32_t> bc(B, 0);
            for (uint32_t i = b; i < e; ++i) {
                uint32_t p = idx[i];
                int k = std::min(B - 1, (int)((cen[3 * p + axis] - cb.lo[axis]) * scale));
                bb[k].grow(pbox[p]);
                bc[k]++;
            }
            std::vector<float> ra(B);
            std::vector<uint32_t> rc(B);
            Box acc;
            uint32_t cnt = 0;
            for (int k = B - 1; k > 0; --k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                ra[k] = acc.area();
                rc[k] = cnt;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (int k = 0; k < B - 1; ++k) {
                lacc.grow(bb[k]);
                lcnt += bc[k];
                if (lcnt == 0 || rc[k + 1] == 0) continue;
                float c = opt.cost_node + (lacc.area() * lcnt + ra[k + 1] * rc[k + 1]) / std::max(parent_area, 1e-30f);
                if (c < best_cost) {
                    best_cost = c;
                    best_axis = axis;
                    best_split = k;
                }
            }
        }
        if ((int)n <= opt.max_leaf && (best_axis < 0 || (float)n <= best_cost)) {
            make_leaf();
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = b + n / 2;  // identical centroids: split by index
        } else {
            const int axis = best_axis;
            const float lo = cb.lo[axis], scale = (float)B / (cb.hi[axis] - cb.lo[axis]);
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](uint32_t p) {
                int k = std::min(B - 1, (int)((cen[3 * p + axis] - lo) * scale));
                return k <= best_split;
            });
            mid = (uint32_t)(it - idx.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
        uint32_t l = alloc(), r = alloc();
        nd.left = (int)l;
        nd.right = (int)r;
        // large subtrees go to another thread while spare threads remain
        if (n > 65536 && spare_threads.fetch_sub(1) > 0) {
            auto f = std::async(std::launch::async, [this, l, b, mid]() { build(l, b, mid); });
            build(r, mid, e);
            f.get();
            spare_threads.fetch_add(1);
        } else {
            if (n > 65536) spare_threads.fetch_add(1);
            build(l, b, mid);
            build(r, mid, e);
        }
    }
};

struct Collapser {
    const std::vector<BNode>& bn;
    const std::vector<uint32_t>& idx;
    BVHResult& out;
    uint32_t depth = 0;

    static void set_child(float* node, int k, const Box& b, int code) {
        constexpr int W = BVH_WIDTH;
        node[0 * W + k] = b.lo[0];
        node[1 * W + k] = b.hi[0];
        node[2 * W + k] = b.lo[1];
        node[3 * W + k] = b.hi[1];
        node[4 * W + k] = b.lo[2];
        node[5 * W + k] = b.hi[2];
        std::memcpy(&node[6 * W + k], &code, 4);
    }

    int leaf(const BNode& n) {
        uint32_t first = (uint32_t)out.slot_prim.size();
        for (uint32_t i = 0; i < n.count; ++i) out.slot_prim.push_back(idx[n.first + i]);
        out.max_leaf = std::max(out.max_leaf, n.count);
        return ~(int)((first << 3) | (n.count - 1));
    }

    // Emits a node whose children are the binary subtrees `kids`.
    uint32_t emit(std::vector<int> kids, uint32_t level) {
        depth = std::max(depth, level);
        uint32_t me = out.num_nodes++;
        out.nodes.resize((size_t)out.num_nodes * BVH_FLOAT_NODE_WORDS);
        {
            float* node = &out.nodes[(size_t)me * BVH_FLOAT_NODE_WORDS];
            for (int k = 0; k < BVH_WIDTH; ++k) {
                Box empty;
                set_child(node, k, empty, 0);
            }
            std::memset(&node[7 * BVH_WIDTH], 0, 4 * BVH_WIDTH);
        }
        for (size_t k = 0; k < kids.size(); ++k) {
            const BNode& c = bn[kids[k]];
            int code;
            if (c.left < 0) {
                code = leaf(c);
            } else {
                std::vector<int> g{c.left, c.right};
                expand(g);
                code = (int)emit(g, level + 1);
            }
            set_child(&out.nodes[(size_t)me * BVH_FLOAT_NODE_WORDS], (int)k, c.box, code);
        }
        return me;
    }

    // Replaces the largest-area internal child by its two children until
    // the node is full.
    void expand(std::vector<int>& g) {
        while (g.size() < (size_t)BVH_WIDTH) {
            int best = -1;
            float ba = -1.0f;
            for (size_t k = 0; k < g.size(); ++k) {
                const BNode& c = bn[g[k]];
                if (c.left >= 0 && c.box.area() > ba) {
                    ba = c.box.area();
                    best = (int)k;
                }
            }
            if (best < 0) break;
            int l = bn[g[best]].left, r = bn[g[best]].right;
            g[best] = l;
            g.push_back(r);
        }
    }
};

// Float node (8 rows of 4 floats) -> quantised node (bvh_layout.h).  Per
// axis the grid step is the smallest power of two 2^e for which every valid
// child fits in bytes: lo = floor((lo_c - origin) / 2^e) >= 0, hi =
// ceil((hi_c - origin) / 2^e) <= 255, with the origin the node's lower bound
// rounded down to a multiple of the step.  The step is at least 2^-22 of the
// node's magnitude, so origin / 2^e + 255 < 2^24 and every grid plane is a
// float.
void quantize_node(const float* f, uint32_t* q) {
    constexpr int W = 4;
    int nvalid = 0;
    for (int k = 0; k < W; ++k)
        if (f[k] <= f[W + k]) nvalid = k + 1;   // valid slots come first (Collapser::emit)
    std::memset(q, 0, 16 * sizeof(uint32_t));
    uint32_t meta = (uint32_t)nvalid << 24;
    for (int a = 0; a < 3; ++a) {
        const float* lo_r = f + (2 * a) * W;
        const float* hi_r = f + (2 * a + 1) * W;
        double lo = 0.0, hi = 0.0;
        if (nvalid > 0) {
            lo = lo_r[0];
            hi = hi_r[0];
            for (int k = 1; k < nvalid; ++k) {
                lo = std::min(lo, (double)lo_r[k]);
                hi = std::max(hi, (double)hi_r[k]);
            }
        }
        const double mag = std::max(std::fabs(lo), std::fabs(hi));
        // step >= 2^-22 of the magnitude: origin / step and origin / step + 255 stay below 2^24, so every
        // grid plane origin + q step is a float (exact layout)
        int e = std::max(-126, (mag > 0.0 ? std::ilogb(mag) : -126) - 22);
        if (hi - lo > 0.0) e = std::max(e, (int)std::ceil(std::log2((hi - lo) / 253.0)));
        float org = 0.0f;
        uint32_t lo_b = 0, hi_b = 0;
        for (; e <= 127; ++e) {
            const double step = std::ldexp(1.0, e);
            org = (float)(std::floor(lo / step) * step);   // a multiple of the step, exact in float
            bool fits = true;
            lo_b = hi_b = 0;
            for (int k = 0; k < W; ++k) {
                uint32_t ql = 255, qh = 0;   // empty slot: lo > hi
                if (k < nvalid) {
                    const double l = std::floor(((double)lo_r[k] - (double)org) / step);
                    const double h = std::ceil(((double)hi_r[k] - (double)org) / step);
                    if (!(l >= 0.0 && h <= 255.0)) {
                        fits = false;
                        break;
                    }
                    ql = (uint32_t)l;
                    qh = (uint32_t)h;
                }
                lo_b |= ql << (8 * k);
                hi_b |= qh << (8 * k);
            }
            if (fits) break;
        }
        std::memcpy(&q[a], &org, 4);
        meta |= (uint32_t)(uint8_t)(int8_t)e << (8 * a);
        q[8 + 2 * a] = lo_b;
        q[9 + 2 * a] = hi_b;
    }
    q[3] = meta;
    for (int k = 0; k < W; ++k) std::memcpy(&q[4 + k], &f[6 * W + k], 4);
    // Empty slots (inverted boxes) take child 0's code, so the traversal needs
    // no per-slot validity test: only a ray whose slab distances are all NaN,
    // or one whose widened plane offset exceeds the box's 255 steps, can
    // accept one, and then visits a real subtree a second time (closest-hit
    // and any-hit answers do not change).  The all-empty root of a scene
    // without triangles takes leaf code -1 (slot 0, a zero-area record)
    // instead of looping back to itself.
    uint32_t fill = ~0u;
    if (nvalid > 0) fill = q[4];
    for (int k = nvalid; k < W; ++k) q[4 + k] = fill;
}

// Renumbers the float-layout nodes so the first `levels` levels come first,
// breadth-first (the traversal kernels copy nodes [0, top_nodes) into LDS);
// every other node keeps its depth-first position relative to the rest.
void order_top_levels(BVHResult& out, int levels) {
    constexpr int W = BVH_WIDTH;
    const uint32_t n = out.num_nodes;
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<uint8_t> placed(n, 0);
    std::vector<uint32_t> frontier{0};
    for (int lvl = 1; lvl <= levels && !frontier.empty(); ++lvl) {
        std::vector<uint32_t> next;
        for (uint32_t v : frontier) {
            order.push_back(v);
            placed[v] = 1;
            const float* f = &out.nodes[(size_t)v * BVH_FLOAT_NODE_WORDS];
            for (int k = 0; k < W; ++k) {
                int code;
                std::memcpy(&code, &f[6 * W + k], 4);
                if (f[k] <= f[W + k] && code >= 0) next.push_back((uint32_t)code);
            }
        }
        frontier.swap(next);
    }
    out.top_nodes = (uint32_t)order.size();
    for (uint32_t v = 0; v < n; ++v)
        if (!placed[v]) order.push_back(v);
    std::vector<uint32_t> index(n);
    for (uint32_t i = 0; i < n; ++i) index[order[i]] = i;
    std::vector<float> re(out.nodes.size());
    for (uint32_t i = 0; i < n; ++i) {
        float* d = &re[(size_t)i * BVH_FLOAT_NODE_WORDS];
        std::memcpy(d, &out.nodes[(size_t)order[i] * BVH_FLOAT_NODE_WORDS], BVH_FLOAT_NODE_WORDS * sizeof(float));
        for (int k = 0; k < W; ++k) {
            int code;
            std::memcpy(&code, &d[6 * W + k], 4);
            if (d[k] <= d[W + k] && code >= 0) {
                code = (int)index[(uint32_t)code];
                std::memcpy(&d[6 * W + k], &code, 4);
            }
        }
    }
    out.nodes.swap(re);
}

}  // namespace

bool bvh_child(const BVHResult& b, uint32_t node, int k, double lo[3], double hi[3], int* code, double margin[3]) {
    const float* f = &b.nodes[(size_t)node * BVH_NODE_WORDS];
    if constexpr (BVH_QUANT) {
        uint32_t w[16];
        std::memcpy(w, f, sizeof w);
        if (k >= (int)(w[3] >> 24)) return false;
        for (int a = 0; a < 3; ++a) {
            float org;
            std::memcpy(&org, &w[a], 4);
            const int e = (int)(int8_t)(uint8_t)(w[3] >> (8 * a));
            lo[a] = (double)org + std::ldexp((double)((w[8 + 2 * a] >> (8 * k)) & 255u), e);
            hi[a] = (double)org + std::ldexp((double)((w[9 + 2 * a] >> (8 * k)) & 255u), e);
            if (margin) margin[a] = 0.0;
        }
        std::memcpy(code, &w[4 + k], 4);
        return true;
    } else {
        constexpr int W = BVH_WIDTH;
        if (!(f[k] <= f[W + k])) return false;
        for (int a = 0; a < 3; ++a) {
            lo[a] = f[(2 * a) * W + k];
            hi[a] = f[(2 * a + 1) * W + k];
            if (margin) margin[a] = 0.0;
        }
        std::memcpy(code, &f[6 * W + k], 4);
        return true;
    }
}

BVHResult build_bvh(const std::vector<BuildTri>& tris, const BVHBuildOptions& opt) {
    BVHResult out;
    const uint32_t n = (uint32_t)tris.size();
    Builder bld(opt);
    int hw = opt.threads > 0 ? opt.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    bld.spare_threads = hw - 1;
    bld.pbox.resize(n);
    bld.cen.resize(3 * (size_t)n);
    bld.idx.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        Box b;
        for (int v = 0; v < 3; ++v) b.grow(tris[i].p[v]);
        bld.pbox[i] = b;
        for (int k = 0; k < 3; ++k) bld.cen[3 * i + k] = 0.5f * (b.lo[k] + b.hi[k]);
        bld.idx[i] = i;
    }
    if (n == 0) {
        out.nodes.assign(BVH_FLOAT_NODE_WORDS, 0.0f);
        float* node = out.nodes.data();
        for (int k = 0; k < BVH_WIDTH; ++k) Collapser::set_child(node, k, Box(), 0);
        out.num_nodes = 1;
        out.depth = 1;
        if constexpr (BVH_QUANT && BVH_TOP_LEVELS > 0) out.top_nodes = 1;
        if constexpr (BVH_QUANT) {
            std::vector<float> qn(BVH_NODE_WORDS);
            quantize_node(out.nodes.data(), reinterpret_cast<uint32_t*>(qn.data()));
            out.nodes.swap(qn);
        }
        return out;
    }
    bld.nodes.resize(2 * (size_t)n);
    uint32_t root = bld.alloc();
    bld.build(root, 0, n);
    bld.nodes.resize(bld.next.load());
    Collapser col{bld.nodes, bld.idx, out};
    out.slot_prim.reserve(n);
    const BNode& r = bld.nodes[root];
    std::vector<int> g;
    if (r.left < 0) {
        g.push_back((int)root);
    } else {
        g = {r.left, r.right};
        col.expand(g);
    }
    col.emit(g, 1);
    out.depth = col.depth;
    if constexpr (BVH_QUANT && BVH_TOP_LEVELS > 0) order_top_levels(out, BVH_TOP_LEVELS);
    if constexpr (BVH_QUANT) {
        std::vector<float> qn((size_t)out.num_nodes * BVH_NODE_WORDS);
        for (uint32_t i = 0; i < out.num_nodes; ++i)
            quantize_node(&out.nodes[(size_t)i * BVH_FLOAT_NODE_WORDS], reinterpret_cast<uint32_t*>(&qn[(size_t)i * BVH_NODE_WORDS]));
        out.nodes.swap(qn);
    }
    return out;
}

}  // namespace xt
