// Kernel-side types and launch declarations shared by kernels.hip and the
// host runtime (runtime.hip).
#pragma once

#include "xt_device.h"
#include "xt_guide.h"

namespace xt {

#ifndef XT_TRACE_BLOCK
#define XT_TRACE_BLOCK 256
#endif
#ifndef XT_TRACE_LDS_DEPTH
#define XT_TRACE_LDS_DEPTH 16
#endif
#ifndef XT_SHADE_BLOCK
#define XT_SHADE_BLOCK 512
#endif
constexpr int SHADE_BLOCK = XT_SHADE_BLOCK;          // threads per block of the shade kernel
constexpr int TRACE_BLOCK = XT_TRACE_BLOCK;          // lanes per block of the traversal kernels
constexpr int TRACE_LDS_DEPTH = XT_TRACE_LDS_DEPTH;  // LDS stack entries per lane
// queue length (an upper bound the host knows) from which the traversal
// kernels hand out rays in chunks from a cursor (DYN, xt_traverse.h)
#ifndef XT_TRAV_DYN_MIN
#define XT_TRAV_DYN_MIN (8u << 20)
#endif
constexpr uint32_t TRAV_DYN_MIN = XT_TRAV_DYN_MIN;

// Where the current batch lies in the render: pixel range, paths, first
// sample.  Host-driven renders pass it in WaveParams; graph replays read it
// from device memory (WaveParams::cursor), advanced by k_advance_cursor.
struct BatchCursor {
    uint32_t first_pixel, num_pixels;
    uint32_t num_paths;
    uint32_t sample0;      // global index of the batch's first sample
    uint32_t s_rel;        // the same, counted from the render's first sample
};

// Per-batch constants of the wavefront pipeline.
struct WaveParams {
    uint32_t width, height;
    uint32_t first_pixel, num_pixels;   // pixel range of the batch
    uint32_t num_paths;                 // num_pixels * samples in the batch
    uint32_t sample0;                   // global sample index of the batch's first sample
    uint32_t seed, stream;
    uint32_t integrator, flags;
    int32_t max_depth, rr_start;
    float rr_prob, eps, shadow_margin;
    float tan_x, tan_y;
    float cam[16];
    const BatchCursor* cursor;          // non-null: the batch fields above are read from here
};

__device__ __forceinline__ BatchCursor batch_of(const WaveParams& wp) {
    if (wp.cursor) return *wp.cursor;
    return BatchCursor{wp.first_pixel, wp.num_pixels, wp.num_paths, wp.sample0, 0u};
}

// Batch schedule of one render: S samples of pix_chunk pixels per batch.
struct BatchPlan {
    uint32_t npix, pix_chunk, S, spp, spp_offset;
};

// Extension-ray queue, SoA of float4 (64 B per entry):
//   o = (origin, pdf of the BSDF sample that made the ray; -1 = delta)
//   d = (direction, path id)
//   w = (path weight, vertex index v | flags << 24)   v = vertex the ray will hit (1 = camera hit)
//   e = (material_eval of the sample, env-light pdf of the direction)
//   rec = training-record index of the vertex that sampled the ray (INVALID: none);
//         only touched while training records are kept
struct PathQueue {
    float4* o;
    float4* d;
    float4* w;
    float4* e;
    uint32_t* rec;
};
constexpr uint32_t QF_PENDING = 1u;   // the previous vertex's NEE result waits in pending[id]

// Shadow-ray queue (48 B per entry):
//   o = (origin, tmax), d = (direction, path id), c = (contribution, mode: 1 = store to pending)
struct ShadowQueue {
    float4* o;
    float4* d;
    float4* c;
};

// Per-vertex training record written while rendering (96 B, six float4):
//   a = (p.xyz, n.x)  b = (n.yz, wo.xy)  c = (wo.z, dir.xyz)
//   d = (pdf, W.xyz)  e = (E.xyz, path id | 1 << 31 if the vertex was guided)
//   f = (pre.xyz, 0)
// p, n, wo: the query; dir and pdf: the sampled direction and its strategy
// pdf; W: path throughput before the vertex; E: the vertex's sample weight
// f cos / pdf; pre: the path radiance collected up to and including the
// vertex's own light sample (written when the next vertex is shaded).
// After the path ends, C = (L_path - pre) / W is the vertex's estimate of
// f cos L_i / pdf along dir, and L_i = C / E.
struct TrainRec {
    float4 a, b, c, d, e, f;
};
constexpr uint32_t TREC_GUIDED = 1u << 31;

// Decoded mixture per path-queue entry, SoA per lobe j:
// lobes_mk[j * stride + i] = (mu, kappa) as four halves, lobes2[j * stride +
// i] = (alpha, alpha * vmf_norm(kappa)) in fp32 with kappa the stored half
// value; 128 bytes per entry.  Readers renormalise mu, so sampling and pdf
// see the same distribution.
constexpr int G_LOBE_FIELDS = 4 * G_LOBES;
__host__ __device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    const _Float16 ha = (_Float16)a, hb = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}
__host__ __device__ __forceinline__ float half_lo(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xFFFFu)); }
__host__ __device__ __forceinline__ float half_hi(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16)); }
// (normalised mu, kappa) of one stored lobe
__host__ __device__ __forceinline__ float4 unpack_lobe(uint2 v) {
    const f3 mu = normalize(mk(half_lo(v.x), half_hi(v.x), half_lo(v.y)));
    return make_float4(mu.x, mu.y, mu.z, half_hi(v.y));
}

struct GuideRender {
    const uint2* lobes_mk; // null: guiding off
    const float2* lobes2;
    uint32_t stride;
    float prob;            // probability of sampling the guide (one-sample MIS with the BSDF)
    TrainRec* recs;        // null: no training records
    uint32_t rec_base;     // record index of this bounce's first queue entry
    uint32_t rec_cap;
    uint32_t max_vertex;   // guide surface vertices 1..max_vertex only (0: all)
};

// ---- unified batches: many renders (items) share one wavefront
//
// The items' paths are numbered one after another (item k owns
// [path_base, path_base + paths), sample-major inside the item as a single
// render numbers them); the batch walks that path space in windows of up to
// `window` paths, each window one pass of the count-driven bounce loop.  A
// queue entry's path id is window-relative; item_of[id] names its item, whose
// record gives the render's parameters, scene and film.

struct MultiItem {
    WaveParams wp;          // the render's parameters (the batch fields are unused)
    const float4* nodes;    // its scene's BVH nodes and leaf-ordered triangles
    const float4* tris;
    uint32_t scene;         // index into MultiBatch::scenes
    uint32_t npix;          // width * height
    uint32_t spp_offset;
    uint32_t film_f64;
    void* film;             // device film (+= radiance sums, as xt_render)
    uint64_t path_base;     // first path of the item in the batch
    uint64_t paths;         // npix * spp
};

// The current window and the items it overlaps (written by k_window_plan).
struct MultiCursor {
    uint64_t g0;            // first path of the window
    uint32_t n;             // paths in the window (0: past the end)
    uint32_t first_item;    // first item overlapping the window
    uint32_t num_overlap;   // items overlapping the window
    uint32_t pad;
};

struct MultiBatch {
    const MultiItem* items;
    const DevScene* scenes;
    uint32_t num_items;
    uint32_t window;
    uint64_t total_paths;
    MultiCursor* cursor;
    uint32_t* item_of;      // per path of the window
    uint32_t* units;        // window plan: accumulation units before each overlapping item (num_overlap + 1)
};

template <bool ANY, int LDS_DEPTH, bool COUNT, bool MULTI = false, bool DYN = false>
__global__ void k_trace(DevScene sc, const float4* ray_o, const float4* ray_d, int ray_stride,
                        const uint32_t* count_ptr, uint32_t count_fixed, float4* hits, uint8_t* occluded,
                        int hit_prim_ids, int tmax_from_ray, uint32_t* ovf, uint32_t* dyn,
                        unsigned long long* visit_totals,
                        uint32_t* clear_a, uint32_t* clear_b, const uint32_t* item_of = nullptr,
                        const MultiItem* items = nullptr);
template <bool DYN>
__global__ void k_shadow_multi(ShadowQueue sq, const uint32_t* count, float4* radiance, float4* pending, uint32_t* ovf,
                               uint32_t* dyn, const uint32_t* item_of, const MultiItem* items);
__global__ void k_window_init(MultiBatch mb);
__global__ void k_window_plan(MultiBatch mb);
__global__ void k_window_advance(MultiBatch mb);
__global__ void k_count_from_window(uint32_t* count, const MultiCursor* c);
__global__ void k_generate_multi(MultiBatch mb, PathQueue q, float4* radiance);
template <bool NMAP>
__global__ void k_shade_multi(MultiBatch mb, PathQueue qin, const uint32_t* count_in, const float4* hits,
                              PathQueue qout, uint32_t* count_out, ShadowQueue sq, uint32_t* count_shadow,
                              float4* radiance, float4* pending);
__global__ void k_accumulate_multi(MultiBatch mb, const float4* radiance);

__global__ void k_generate(WaveParams wp, PathQueue q, float4* radiance);
template <bool GUIDE, bool NMAP, bool SORT>
__global__ void k_shade(WaveParams wp, const DevScene* scene, PathQueue qin, const uint32_t* count_in, const float4* hits,
                        PathQueue qout, uint32_t* count_out, ShadowQueue sq, uint32_t* count_shadow,
                        float4* radiance, float4* pending, GuideRender gr);
template <bool COUNT, bool DYN = false>
__global__ void k_shadow(DevScene sc, ShadowQueue sq, const uint32_t* count, float4* radiance, float4* pending,
                         uint32_t* ovf, uint32_t* dyn, unsigned long long* visit_totals, const uint32_t* counts3,
                         uint32_t* counts_out);
__global__ void k_accumulate(WaveParams wp, const float4* radiance, void* film, int film_f64);
__global__ void k_zero_u32(uint32_t* p, uint32_t n);
constexpr int VALU_PROBE_CHAINS = 16;   // FMA chains per lane of k_valu_probe
__global__ void k_valu_probe(uint32_t iters, float* out);
__global__ void k_cursor_init(BatchCursor* c, BatchPlan plan);
__global__ void k_advance_cursor(BatchCursor* c, BatchPlan plan);
__global__ void k_count_from_cursor(uint32_t* count, const BatchCursor* c);

}  // namespace xt
