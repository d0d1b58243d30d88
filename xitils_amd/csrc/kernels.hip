// gfx950 kernels: ray queries and the wavefront path tracer.
//
// Wavefront pipeline per batch of paths (one path = one pixel sample):
//   generate -> [extend -> shade -> shadow]* -> accumulate
// extend  : closest hit for every queued ray         (Scene::intersect, Scene.h:30)
// shade   : one StandardPathTracer / NaivePathTracer loop step
//           (PathTracer.h:147-246 / 65-101): emission at the new vertex,
//           roulette, BSDF sample -> next queue, light sample -> shadow queue
// shadow  : any hit for every shadow ray             (Scene::intersectAny, Scene.h:34)
// Queues are compacted with one atomic per queue per 512-thread block (wave
// ballots + an LDS prefix); per-wave atomics on the two queue counters were
// the shade kernel's bottleneck.
// Radiance is kept per path id and summed in the reference's order (BSDF-hit
// emission of vertex v+1 before the light sample of vertex v), so the image
// does not depend on queue order.
#include "kernels.h"
#include "xt_traverse.h"
#include "xt_nmap.h"

namespace xt {

// Block-aggregated appends to two queues at once (extension and shadow): one
// atomic per queue per block instead of per wave.  Every thread of the block
// must call it (the callers' loops are block-uniform).  Returns this thread's
// slots (valid when its predicate is set).
template <int BLOCK>
__device__ __forceinline__ uint2 block_append2(uint32_t* counter_a, bool pa, uint32_t* counter_b, bool pb,
                                               uint32_t* sh /* 2 * (BLOCK / 64) + 2 words of LDS */) {
    constexpr int NW = BLOCK / 64;
    const uint64_t ma = __ballot(pa), mb = __ballot(pb);
    const uint32_t wave = threadIdx.x / 64, lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    if (lane == 0) {
        sh[wave] = (uint32_t)__popcll(ma);
        sh[NW + wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (int w = 0; w < NW; ++w) {
            uint32_t ca = sh[w], cb = sh[NW + w];
            sh[w] = ta;
            sh[NW + w] = tb;
            ta += ca;
            tb += cb;
        }
        sh[2 * NW] = ta ? atomicAdd(counter_a, ta) : 0u;
        sh[2 * NW + 1] = tb ? atomicAdd(counter_b, tb) : 0u;
    }
    __syncthreads();
    uint2 r = make_uint2(sh[2 * NW] + sh[wave] + (uint32_t)__popcll(ma & below),
                         sh[2 * NW + 1] + sh[NW + wave] + (uint32_t)__popcll(mb & below));
    __syncthreads();   // sh is reused by the next iteration
    return r;
}

// VALU issue probe: VALU_PROBE_CHAINS independent fp32 FMA chains per lane
// (one VALU instruction per step, enough independent work to cover the FMA
// latency at 8 waves per SIMD) and no memory traffic; `out` is written only
// if the sums take a value they never take, which keeps the chains live.
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, float* out) {
    float a[VALU_PROBE_CHAINS];
#pragma unroll
    for (int k = 0; k < VALU_PROBE_CHAINS; ++k) a[k] = 1.0f + 1e-6f * (float)(threadIdx.x + 64 * k);
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < VALU_PROBE_CHAINS; ++k) a[k] = __builtin_fmaf(a[k], 0.9999999f, 1e-7f);
    }
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < VALU_PROBE_CHAINS; ++k) sum += a[k];
    if (sum == -1.0f) out[blockIdx.x] = sum;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- ray queries

// Adds a wave's visit counts to the 64-bit totals (counting build only).
__device__ __forceinline__ void flush_counts(unsigned long long* totals, const VisitCount& vc) {
    unsigned long long a = vc.nodes, b = vc.tris;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(&totals[0], a);
        atomicAdd(&totals[1], b);
    }
}

// Ray queries over a ray array (stride in float4 between rays): closest hit
// (ANY = false) -> hits[i] = (t, slot or prim id, b0, b1); any hit ->
// occluded[i].  Persistent waves with in-place lane refill (xt_traverse.h).
// MULTI: the rays of a unified batch's queue, each traced in its own item's
// scene (item_of[path id] -> the item's nodes and triangles, per lane).
// Bounded to 8 waves per SIMD, which every form but the chunked multi-scene
// and counting ones reaches without spilling; the chunked multi-scene form
// takes 7, the chunked counting form (run once per scene) is unbounded.
template <bool ANY, int LDS_DEPTH, bool COUNT, bool MULTI, bool DYN>
__global__ __launch_bounds__(TRACE_BLOCK, DYN && COUNT ? 1 : DYN && MULTI ? 7 : 8) void k_trace(DevScene sc, const float4* __restrict__ ray_o,
                                                        const float4* __restrict__ ray_d, int ray_stride,
                                                        const uint32_t* __restrict__ count_ptr, uint32_t count_fixed,
                                                        float4* __restrict__ hits, uint8_t* __restrict__ occluded,
                                                        int hit_prim_ids, int tmax_from_ray, uint32_t* __restrict__ ovf,
                                                        uint32_t* __restrict__ dyn,
                                                        unsigned long long* __restrict__ visit_totals,
                                                        uint32_t* __restrict__ clear_a, uint32_t* __restrict__ clear_b,
                                                        const uint32_t* __restrict__ item_of,
                                                        const MultiItem* __restrict__ items) {
    __shared__ uint32_t lds[LDS_DEPTH * TRACE_BLOCK];
    // the next round's queue counters (their last readers ran before this
    // kernel in stream order), so the bounce loop needs no clearing launches
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (clear_a) *clear_a = 0u;
        if (clear_b) *clear_b = 0u;
    }
    __shared__ float4 top[TOP_F4];
    stage_top_nodes(top, sc);
    const uint32_t n = count_ptr ? *count_ptr : count_fixed;
    const size_t nthreads = (size_t)gridDim.x * blockDim.x;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* my_lds = lds + threadIdx.x;
    uint32_t* my_ovf = ovf ? ovf + gtid : nullptr;
    const uint32_t nwaves = (uint32_t)(nthreads / 64), wave = (uint32_t)(gtid / 64);
    RayChunks ch = chunks_first<DYN>(wave, nwaves, n);
    const uint64_t below = (1ull << __lane_id()) - 1ull;
    VisitCount vc;
    bool active = false;
    uint32_t rid = 0;
    RayPre r;
    Trav t;
    TravView view{sc.nodes, sc.tris, MULTI ? 0u : sc.top_nodes};
    while (true) {
        const uint64_t busy = __ballot(active);
        const uint32_t nidle = 64u - (uint32_t)__popcll(busy);
        const bool refill = nidle >= (uint32_t)REFILL_IDLE || busy == 0;
        if (DYN && refill && ch.next >= ch.end && !ch.exhausted) chunks_refill(ch, dyn);
        if (refill && ch.next < ch.end) {
            const uint32_t my = ch.next + (uint32_t)__popcll(~busy & below);
            if (!active && my < ch.end) {
                rid = my;
                float4 o4 = ray_o[(size_t)rid * ray_stride];
                float4 d4 = ray_d[(size_t)rid * ray_stride];
                if constexpr (MULTI) {
                    const MultiItem& it = items[item_of[__float_as_uint(d4.w)]];
                    view.nodes = it.nodes;
                    view.tris = it.tris;
                }
                r = ray_setup(xyz(o4), xyz(d4));
                trav_init(t, tmax_from_ray ? o4.w : INFINITY);   // extension rays: tMax = Infinity (PathTracer.h:167)
                active = true;
            }
            ch.next = min(ch.next + nidle, ch.end);
        }
        if (__ballot(active) == 0) break;
        bool occ = false;
        bool done;
        if constexpr (MULTI)
            done = trav_step<ANY, LDS_DEPTH, COUNT>(active, view, r, t, top, my_lds, TRACE_BLOCK, my_ovf, nthreads, vc, &occ);
        else
            done = trav_step<ANY, LDS_DEPTH, COUNT>(active, sc, r, t, top, my_lds, TRACE_BLOCK, my_ovf, nthreads, vc, &occ);
        if (active) {
            if (done) {
                if (ANY) {
                    occluded[rid] = occ ? 1 : 0;
                } else {
                    hits[rid] = t.best_slot != INVALID
                                    ? make_float4(t.best_t, __uint_as_float(hit_prim_ids ? prim_of_slot(view, t.best_slot)
                                                                                            : t.best_slot),
                                                  t.b0, t.b1)
                                    : make_float4(INFINITY, __uint_as_float(INVALID), 0.0f, 0.0f);
                }
                active = false;
            }
        }
    }
    if constexpr (DYN) chunks_release(dyn);
    if (COUNT) flush_counts(visit_totals, vc);
}

#define XT_INST_TRACE(ANY, COUNT, MULTI, DYN)                                                                      \
    template __global__ void k_trace<ANY, TRACE_LDS_DEPTH, COUNT, MULTI, DYN>(                                     \
        DevScene, const float4*, const float4*, int, const uint32_t*, uint32_t, float4*, uint8_t*, int, int,       \
        uint32_t*, uint32_t*, unsigned long long*, uint32_t*, uint32_t*, const uint32_t*, const MultiItem*);
XT_INST_TRACE(false, false, false, false)
XT_INST_TRACE(true, false, false, false)
XT_INST_TRACE(false, true, false, false)
XT_INST_TRACE(true, true, false, false)
XT_INST_TRACE(false, false, true, false)
XT_INST_TRACE(false, false, true, true)
XT_INST_TRACE(false, false, false, true)
XT_INST_TRACE(true, false, false, true)
XT_INST_TRACE(false, true, false, true)
XT_INST_TRACE(true, true, false, true)
#undef XT_INST_TRACE

// ---------------------------------------------------------------- wavefront PT

// Camera ray of a path (PinholeCamera::generateRay, Camera.h:45-55) at the
// jittered film position of RenderTargetTile::GenerateFilmPosition
// (RenderTarget.h:68-80).
__global__ __launch_bounds__(256) void k_generate(WaveParams wp, PathQueue q, float4* __restrict__ radiance) {
    const BatchCursor bc = batch_of(wp);
    const uint32_t P = bc.num_paths;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < P; id += gridDim.x * blockDim.x) {
        uint32_t s_local = id / bc.num_pixels;
        uint32_t pixel = bc.first_pixel + (id - s_local * bc.num_pixels);
        uint32_t x = pixel % wp.width, y = pixel / wp.width;
        Rng rng{make_uint2(wp.seed, wp.stream), pixel, bc.sample0 + s_local};
        uint4 b0 = rng.block(0, 0);
        float ux = Rng::to_unit(b0.x), uy = Rng::to_unit(b0.y);
        float fx = (float)x, fy = (float)y;
        fx /= (float)wp.width;
        fy /= (float)wp.height;
        fy = 1.0f - fy;
        fx += (ux - 0.5f) / (float)wp.width;
        fy += (uy - 0.5f) / (float)wp.height;
        f3 dc = normalize(mk(wp.tan_x * (fx - 0.5f) * 2, wp.tan_y * (fy - 0.5f) * 2, 1));
        f3 o = xf_point(wp.cam, mk(0, 0, 0));
        f3 d = xf_vector(wp.cam, dc);
        // e (the sampling record) and the radiance slot of a camera ray are
        // implied by its vertex index 1: the shade step reads neither
        q.o[id] = make_float4(o.x, o.y, o.z, 0.0f);
        q.d[id] = make_float4(d.x, d.y, d.z, __uint_as_float(id));
        q.w[id] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(1u));
        if (q.rec) q.rec[id] = INVALID;
    }
}

// Material-sorted shading (XT_FLAG_SORT_MATERIALS): the block's queue
// entries [base, base + SHADE_BLOCK) are partitioned by shading key (miss,
// then 1 + material type of the hit) before they are shaded, stably, so each
// wave of the block mostly shades one kind of BSDF instead of every kind the
// queue mixes.  Returns this thread's entry offset within the block; every
// thread of the block must call it (the callers' loops are block-uniform).
constexpr int SHADE_KEYS = 1 + XT_MAT_COUNT_ + 1;   // miss, material types, entries past the queue end
template <int BLOCK>
__device__ __forceinline__ uint32_t block_sort_by_material(const DevScene& sc, const float4* __restrict__ hits,
                                                           uint32_t base, uint32_t n,
                                                           uint32_t* sh /* (BLOCK / 64 + 1) * SHADE_KEYS + BLOCK */) {
    constexpr int NW = BLOCK / 64;
    const uint32_t i = base + threadIdx.x;
    uint32_t key = SHADE_KEYS - 1;
    if (i < n) {
        const uint32_t slot = __float_as_uint(hits[i].y);
        key = 0;
        if (slot != INVALID) {
            const uint32_t m = __float_as_uint(sc.tris[(size_t)slot * TRI_F4 + 1].w);
            key = 1 + min(sc.mats[m].type, (uint32_t)XT_MAT_COUNT_);
        }
    }
    const uint32_t wave = threadIdx.x / 64, lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < SHADE_KEYS; ++k) {
        const uint64_t m = __ballot(key == (uint32_t)k);
        if (key == (uint32_t)k) rank = (uint32_t)__popcll(m & below);
        if (lane == 0) sh[wave * SHADE_KEYS + k] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // key-major, then wave order: every wave's share of a key after the
        // earlier waves' share of it, and after every entry of a smaller key
        uint32_t total = 0;
        for (int k = 0; k < SHADE_KEYS; ++k)
            for (int w = 0; w < NW; ++w) {
                const uint32_t c = sh[w * SHADE_KEYS + k];
                sh[w * SHADE_KEYS + k] = total;
                total += c;
            }
    }
    __syncthreads();
    uint32_t* order = sh + (NW + 1) * SHADE_KEYS;
    order[sh[wave * SHADE_KEYS + key] + rank] = threadIdx.x;
    __syncthreads();
    const uint32_t mine = order[threadIdx.x];
    __syncthreads();   // sh is reused by the next iteration
    return mine;
}

// The result of shading one queue entry: the extension ray and the light
// sample it pushes, if any.
struct ShadeOut {
    bool push_ext = false, push_sh = false, guided = false;
    f3 ext_o{0.0f, 0.0f, 0.0f}, ext_d{0.0f, 0.0f, 0.0f}, ext_w{0.0f, 0.0f, 0.0f}, ext_e{0.0f, 0.0f, 0.0f};
    float ext_pdf = 0.0f, ext_pe = 0.0f;
    uint32_t ext_v = 0;
    Surf s;
    f3 sh_o{0.0f, 0.0f, 0.0f}, sh_d{0.0f, 0.0f, 0.0f}, sh_c{0.0f, 0.0f, 0.0f};
    float sh_tmax = 0.0f;
    uint32_t id = 0;
};

// Path id -> (pixel, sample) of a single render's batch (as k_generate).
struct CursorMap {
    const WaveParams& wp;
    __device__ __forceinline__ uint2 operator()(uint32_t id) const {
        const BatchCursor bc = batch_of(wp);
        const uint32_t s_local = id / bc.num_pixels;
        return make_uint2(bc.first_pixel + (id - s_local * bc.num_pixels), bc.sample0 + s_local);
    }
};

// One StandardPathTracer / NaivePathTracer loop step (PathTracer.h:147-246 /
// 65-101) for queue entry i: emission at the new vertex, roulette, BSDF
// sample -> extension ray, light sample -> shadow ray.
// Per-render constants of the shade step.
struct ShadeConst {
    bool standard, can_sample, uniform_bary;
    EnvNee envn;
    float p_area;
    uint2 key;
};
__device__ __forceinline__ ShadeConst shade_const(const WaveParams& wp, const DevScene& sc) {
    ShadeConst c;
    c.standard = wp.integrator == XT_INTEGRATOR_STANDARD;
    c.can_sample = sc.num_lights > 0;
    c.envn = env_nee_setup(sc.env, c.standard && (wp.flags & XT_FLAG_ENV_NEE) != 0, c.can_sample);
    c.p_area = c.can_sample ? 1.0f - c.envn.p_select : 0.0f;
    c.uniform_bary = (wp.flags & XT_FLAG_UNIFORM_BARY) != 0;
    c.key = make_uint2(wp.seed, wp.stream);
    return c;
}

template <bool GUIDE, bool NMAP, class PathMap>
__device__ __forceinline__ void shade_entry(const WaveParams& wp, const DevScene& sc, const ShadeConst& sc0,
                                            const PathQueue& qin, const float4* __restrict__ hits,
                                            float4* __restrict__ radiance, float4* __restrict__ pending,
                                            const GuideRender& gr, uint32_t i, const PathMap& pmap, ShadeOut& out) {
    const bool standard = sc0.standard;
    const bool can_sample = sc0.can_sample;
    const EnvNee envn = sc0.envn;
    const float p_area = sc0.p_area;
    const bool uniform_bary = sc0.uniform_bary;
    const uint2 key = sc0.key;
    bool& push_ext = out.push_ext;
    bool& push_sh = out.push_sh;
    bool& guided = out.guided;
    f3& ext_o = out.ext_o;
    f3& ext_d = out.ext_d;
    f3& ext_w = out.ext_w;
    f3& ext_e = out.ext_e;
    float& ext_pdf = out.ext_pdf;
    float& ext_pe = out.ext_pe;
    uint32_t& ext_v = out.ext_v;
    Surf& s = out.s;
    f3& sh_o = out.sh_o;
    f3& sh_d = out.sh_d;
    f3& sh_c = out.sh_c;
    float& sh_tmax = out.sh_tmax;
    uint32_t& id = out.id;
        const float4 o4 = qin.o[i], d4 = qin.d[i], w4 = qin.w[i], h = hits[i];
        id = __float_as_uint(d4.w);
        const uint32_t vf = __float_as_uint(w4.w);
        const uint32_t v = vf & 0xFFFFFFu, qflags = vf >> 24;
        // a camera ray (vertex 1) has no sampling record and no radiance yet:
        // the generate step writes neither
        const float4 e4 = v > 1 ? qin.e[i] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        const float pdf_prev = o4.w, pe_prev = e4.w;
        const f3 d = xyz(d4), meval = xyz(e4);
        f3 weight = xyz(w4);
        f3 rad = v > 1 ? xyz(radiance[id]) : splat(0.0f);
        const f3 pend = (qflags & QF_PENDING) ? xyz(pending[id]) : splat(0.0f);
        if (GUIDE && gr.recs) {
            // the vertex that sampled this ray: radiance up to and including its light sample
            const uint32_t rin = qin.rec[i];
            if (rin != INVALID) gr.recs[rin].f = f4(rad + pend, 0.0f);
        }
        const uint32_t slot = __float_as_uint(h.y);
        const bool hit = slot != INVALID;
        xt_material mat;
        bool emissive = false;
        if (hit) {
            s = make_surf<NMAP>(sc, slot, d, h.z, h.w);
            mat = sc.mats[s.mat];
            emissive = mat.type == XT_MAT_EMISSION;
        }
        // ---- radiance arriving along the ray that reached this vertex
        if (v == 1 && standard) {
            if (!hit) {
                if (sc.env.has_sky) rad += weight * env_radiance(sc.env, d);          // PathTracer.h:133-136
            } else if (emissive) {
                rad += weight * emitted(mat, -d, s.n);                                // PathTracer.h:143-145
            }
        } else if (standard) {
            if (!hit) {
                if (sc.env.has_sky) {                                                  // PathTracer.h:192-195
                    f3 Le = env_radiance(sc.env, d);
                    if (envn.on && pdf_prev >= 0.0f) {
                        // power heuristic; a direction the sky sampler cannot produce
                        // (pe = 0) or an unbounded BSDF pdf keeps the full weight
                        const float mis = (pe_prev > 0.0f && pdf_prev < INFINITY)
                                              ? (pdf_prev * pdf_prev) / (pdf_prev * pdf_prev + pe_prev * pe_prev)
                                              : 1.0f;
                        rad += weight * mis * meval * Le;
                    } else {
                        rad += weight * meval * Le;
                    }
                }
            } else if (emissive) {                                                     // PathTracer.h:171-189
                float mis;
                if (pdf_prev >= 0.0f && can_sample) {
                    float pl = sc.light_pdf[slot] * p_area;
                    float cosL = fabsf(dot(d, s.sn));
                    float distSq = h.x * h.x;
                    pl *= distSq / cosL;
                    mis = (pdf_prev * pdf_prev) / (pdf_prev * pdf_prev + pl * pl);
                } else {
                    mis = 1.0f;
                }
                if (mis > 0.0f) rad += weight * mis * meval * emitted(mat, -d, s.n);
            }
        } else {                                                                       // PathTracer.h:75-98
            if (!hit) {
                if (sc.env.has_sky) rad += clamp_pos(weight * env_radiance(sc.env, d));
            } else if (emissive) {
                rad += weight * emitted(mat, -d, s.n);
            }
        }
        if (qflags & QF_PENDING) rad += pend;  // light sample of the previous vertex
        radiance[id] = f4(rad, 0.0f);

        bool live = hit && !emissive;
        Rng rng{key, 0u, 0u};
        if (live) {
            const uint2 ps = pmap(id);   // path id -> (pixel, sample)
            rng.pixel = ps.x;
            rng.sample = ps.y;
        }
        if (live && standard) {
            if (v > 1) {                                       // PathTracer.h:239-245
                weight = weight * meval;
                if (is_zero(weight)) live = false;
            }
            if (live && wp.max_depth > 0 && (int)v > wp.max_depth) live = false;
            uint4 blk0 = make_uint4(0, 0, 0, 0);
            if (live) blk0 = rng.block(v, 0);
            if (live && (int)(v + 1) > wp.rr_start) {          // PathTracer.h:149-155
                if (Rng::to_unit(blk0.x) >= wp.rr_prob) live = false;
                else weight = div(weight, wp.rr_prob);
            }
            if (live) {
                const Frame fs = make_frame(s.sn);   // BasisVectors(shading.n), shared by both samplers
                f3 wi = splat(0.0f);
                float pdfB = 0.0f;
                f3 me;
                // Guided vertex: one-sample MIS of the learned mixture (probability
                // gr.prob) and the BSDF sampler; the strategy pdf is the mixture
                // pdf wherever the BSDF pdf is positive, and zero elsewhere.
                guided = GUIDE && gr.lobes_mk != nullptr && !mat_is_delta(mat.type) &&
                         (gr.max_vertex == 0 || v <= gr.max_vertex);
                const WalkKey wk{key, rng.pixel, rng.sample, v};
                auto l4 = [&](int j) { return unpack_lobe(gr.lobes_mk[(size_t)j * gr.stride + i]); };
                auto l2 = [&](int j) { return gr.lobes2[(size_t)j * gr.stride + i]; };
                // (a NaN BSDF pdf -- GGX D at a grazing half vector -- stays NaN, so
                // the light-sample MIS below drops that sample as the plain path does)
                auto strategy_pdf = [&](f3 dir, float pb) {
                    if (pb != pb) return pb;
                    return pb > 0.0f ? gr.prob * guide_pdf_lobes(l4, l2, dir) + (1.0f - gr.prob) * pb : 0.0f;
                };
                if (guided) {
                    const uint4 blk3 = rng.block(v, 3);
                    // The random walk's sampling density is not its getPDF (an
                    // approximation, Main.cpp:245), so a walk-sampled direction is
                    // weighted with its throughput, the unbiased estimate of f / p:
                    // me = throughput * p^ / (c q + (1 - c) p^), the one-sample MIS
                    // whose weights (computed with p^) still sum to one.
                    const bool walk = NMAP && mat.type == XT_MAT_MICROFACET_NMAP;
                    float pb;
                    f3 fc;
                    if (Rng::to_unit(blk3.x) < gr.prob) {
                        wi = guide_sample_lobes(l4, l2, Rng::to_unit(blk3.y), Rng::to_unit(blk3.z),
                                                Rng::to_unit(blk3.w));
                        fc = mat_eval<NMAP>(mat, s, wi, &pb, wk, D_WALK_EVAL2);
                    } else if (walk) {
                        fc = mat_sample<NMAP>(mat, s, fs, blk0, wk, &wi, &pb) * pb;
                    } else {
                        mat_sample<NMAP>(mat, s, fs, blk0, wk, &wi, &pdfB);
                        fc = mat_eval<NMAP>(mat, s, wi, &pb, wk, D_WALK_EVAL2);
                    }
                    pdfB = strategy_pdf(wi, pb);
                    if (!(pdfB > 0.0f)) pdfB = 0.0f;
                    me = pdfB > 0.0f ? div(fc, pdfB) : splat(0.0f);
                    if (!finite3(me)) me = splat(0.0f);
                } else {
                    me = mat_sample<NMAP>(mat, s, fs, blk0, wk, &wi, &pdfB);
                }
                if (!is_zero(me)) {                            // PathTracer.h:165-167
                    push_ext = true;
                    ext_d = wi;
                    ext_o = s.p + wp.eps * wi;
                    ext_w = weight;
                    ext_e = me;
                    ext_pdf = pdfB;
                    ext_v = v + 1;
                    ext_pe = (envn.on && pdfB >= 0.0f) ? envn.p_select * env_nee_pdf(sc.env, envn, normalize(wi), s.sn)
                                                       : 0.0f;
                }
                // ---- light sample (PathTracer.h:202-233)
                uint4 blk1 = rng.block(v, 1);
                bool sample_env = false;
                if (envn.on) sample_env = !can_sample || Rng::to_unit(blk1.x) < envn.p_select;
                if (sample_env) {
                    if (pdfB >= 0.0f) {
                        f3 ld = env_nee_sample(sc.env, envn, fs, Rng::to_unit(blk1.w), Rng::to_unit(blk1.y),
                                               Rng::to_unit(blk1.z));
                        float pe = envn.p_select * env_nee_pdf(sc.env, envn, ld, s.sn);
                        float pb;
                        f3 fc = mat_eval<NMAP>(mat, s, ld, &pb, wk, D_WALK_EVAL);
                        if (guided) pb = strategy_pdf(ld, pb);
                        const float mis = (pe * pe) / (pb * pb + pe * pe);
                        if (pe > 0.0f && !is_zero(fc) && mis > 0.0f) {   // NaN mis: dropped
                            f3 c = div(weight * mis * fc * env_radiance(sc.env, ld), pe);
                            push_sh = true;
                            sh_d = ld;
                            sh_o = s.p + wp.eps * ld;
                            sh_tmax = INFINITY;
                            sh_c = c;
                        }
                    }
                } else if (can_sample) {
                    uint2 lo = sc.light_objs[blk1.y % sc.num_lights];
                    uint32_t lslot = sc.light_slots[lo.x + blk1.z % lo.y];
                    uint4 blk2 = rng.block(v, 2);
                    float t0, t1;
                    if (uniform_bary) {
                        float su = sqrtf(Rng::to_unit(blk2.x));
                        t0 = 1.0f - su;
                        t1 = Rng::to_unit(blk2.y) * su;
                    } else {                                   // TriangleIndexed.h:184-192
                        float a = Rng::to_unit(blk2.x), b = Rng::to_unit(blk2.y), c = Rng::to_unit(blk2.z);
                        float sum = a + b + c;
                        if (sum == 0.0f) {
                            a = b = c = 1.0f;
                            sum = 3.0f;
                        }
                        float inv = 1.0f / sum;
                        t0 = a * inv;
                        t1 = b * inv;
                    }
                    const float4* lt = sc.tris + (size_t)lslot * TRI_F4;
                    float4 l0 = lt[0], l1 = lt[1], l2 = lt[2];
                    f3 q0 = xyz(l0), q1 = xyz(l1), q2 = xyz(l2);
                    f3 lp = lerp3(q0, q1, q2, t0, t1);
                    f3 ln = normalize(cross(q1 - q0, q2 - q0));
                    f3 lsn = ln;
                    if (__float_as_uint(l2.w) & TRI_FLAG_NORMALS) {
                        const float4* nr = sc.normals + (size_t)lslot * TRI_F4;
                        lsn = normalize(lerp3(xyz(nr[0]), xyz(nr[1]), xyz(nr[2]), t0, t1));
                    }
                    ln = normalize(ln);
                    lsn = normalize(lsn);
                    float pdfL = sc.light_pdf[lslot] * p_area;
                    float dist = length(lp - s.p);
                    f3 sd = div(lp - s.p, dist);
                    if (dot(sd, ln) < 0) {
                        float distSq = dist * dist;
                        float mis;
                        float pb;
                        f3 fc = mat_eval<NMAP>(mat, s, sd, &pb, wk, D_WALK_EVAL);
                        if (guided) pb = strategy_pdf(sd, pb);
                        if (pdfB >= 0.0f) {
                            float cosL = fabsf(dot(-sd, lsn));
                            pb *= cosL / distSq;
                            mis = (pdfL * pdfL) / (pb * pb + pdfL * pdfL);
                        } else {
                            mis = 0.0f;
                        }
                        if (mis > 0.0f) {
                            float G = fabsf(dot(-sd, lsn)) / distSq;
                            const xt_material& lm = sc.mats[__float_as_uint(l1.w)];
                            f3 c = div(weight * mis * fc * emitted(lm, -sd, ln) * G, pdfL);
                            push_sh = true;
                            sh_d = sd;
                            sh_o = s.p + wp.eps * sd;
                            sh_tmax = dist - wp.shadow_margin;
                            sh_c = c;
                        }
                    }
                }
            }
        } else if (live) {  // NaivePathTracer (PathTracer.h:86-100), then the roulette of the next loop head
            uint4 blk0 = rng.block(v, 0);
            f3 wi = splat(0.0f);
            float pdfB = 0.0f;
            const WalkKey wk{key, rng.pixel, rng.sample, v};
            weight = weight * mat_sample<NMAP>(mat, s, make_frame(s.sn), blk0, wk, &wi, &pdfB);
            if (is_zero(weight)) live = false;
            if (live && wp.max_depth > 0 && (int)v > wp.max_depth) live = false;
            if (live && (int)(v + 1) > wp.rr_start) {
                uint4 nb = rng.block(v + 1, 0);
                if (Rng::to_unit(nb.x) >= wp.rr_prob) live = false;
                else weight = div(weight, wp.rr_prob);
            }
            if (live) {
                push_ext = true;
                ext_d = wi;
                ext_o = s.p + wp.eps * wi;
                ext_w = weight;
                ext_e = splat(1.0f);
                ext_pdf = pdfB;
                ext_v = v + 1;
            }
        }
}

// Compacted appends of a block's shaded entries (every thread of the block
// calls it): the extension ray with its training record, the shadow ray.
template <bool GUIDE>
__device__ __forceinline__ void shade_emit(ShadeOut& o, const PathQueue& qout, uint32_t* __restrict__ count_out,
                                           const ShadowQueue& sq, uint32_t* __restrict__ count_shadow,
                                           const GuideRender& gr, uint32_t* append_lds) {
    // guided light samples: a non-finite contribution is dropped, never summed
    if (GUIDE && o.push_sh && !finite3(o.sh_c)) o.push_sh = false;
    const uint2 slots = block_append2<SHADE_BLOCK>(count_out, o.push_ext, count_shadow, o.push_sh, append_lds);
    const uint32_t ei = slots.x, si = slots.y;
    if (o.push_ext) {
        uint32_t fl = o.push_sh ? QF_PENDING : 0u;
        qout.o[ei] = f4(o.ext_o, o.ext_pdf);
        qout.d[ei] = f4(o.ext_d, __uint_as_float(o.id));
        qout.w[ei] = f4(o.ext_w, __uint_as_float(o.ext_v | (fl << 24)));
        qout.e[ei] = f4(o.ext_e, o.ext_pe);
        if (GUIDE && gr.recs) {
            const uint32_t ri = gr.rec_base + ei;
            const bool keep = ri < gr.rec_cap && ri >= gr.rec_base;
            qout.rec[ei] = keep ? ri : INVALID;
            if (keep) {
                TrainRec& rr = gr.recs[ri];
                const Surf& s = o.s;
                rr.a = make_float4(s.p.x, s.p.y, s.p.z, s.sn.x);
                rr.b = make_float4(s.sn.y, s.sn.z, s.wo.x, s.wo.y);
                rr.c = make_float4(s.wo.z, o.ext_d.x, o.ext_d.y, o.ext_d.z);
                rr.d = make_float4(o.ext_pdf, o.ext_w.x, o.ext_w.y, o.ext_w.z);
                rr.e = f4(o.ext_e, __uint_as_float(o.id | (o.guided ? TREC_GUIDED : 0u)));
            }
        }
    }
    if (o.push_sh) {
        sq.o[si] = f4(o.sh_o, o.sh_tmax);
        sq.d[si] = f4(o.sh_d, __uint_as_float(o.id));
        sq.c[si] = f4(o.sh_c, o.push_ext ? 1.0f : 0.0f);
    }
}

// The forms without normal mapping (but the guided sorted one, at 144 VGPRs)
// are bounded to 4 waves per SIMD, which they reach without spilling.
template <bool GUIDE, bool NMAP, bool SORT>
__global__ __launch_bounds__(SHADE_BLOCK, NMAP || (GUIDE && SORT) ? 1 : 4) void k_shade(WaveParams wp, const DevScene* __restrict__ scene, PathQueue qin,
                                               const uint32_t* __restrict__ count_in, const float4* __restrict__ hits,
                                               PathQueue qout, uint32_t* __restrict__ count_out, ShadowQueue sq,
                                               uint32_t* __restrict__ count_shadow, float4* __restrict__ radiance,
                                               float4* __restrict__ pending, GuideRender gr) {
    // the scene record is read through a pointer: by value its fields
    // occupy scalar registers for the whole kernel and spill (111 SGPR
    // spills, 556 lane reads, 112 VGPRs; through the pointer 32, 80, 103)
    const DevScene& sc = *scene;
    const uint32_t n = *count_in;
    const uint32_t stride = gridDim.x * blockDim.x;
    __shared__ uint32_t append_lds[2 * (SHADE_BLOCK / 64) + 2];
    __shared__ uint32_t sort_lds[SORT ? (SHADE_BLOCK / 64 + 1) * SHADE_KEYS + SHADE_BLOCK : 1];
    const CursorMap pmap{wp};
    const ShadeConst c0 = shade_const(wp, sc);
    // whole-block iterations so the appends below always see every thread
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + (SORT ? block_sort_by_material<SHADE_BLOCK>(sc, hits, base, n, sort_lds) : threadIdx.x);
        ShadeOut o;
        if (i < n) shade_entry<GUIDE, NMAP>(wp, sc, c0, qin, hits, radiance, pending, gr, i, pmap, o);
        shade_emit<GUIDE>(o, qout, count_out, sq, count_shadow, gr, append_lds);
    }
}

#define XT_INST_SHADE(G, N, S)                                                                                   \
    template __global__ void k_shade<G, N, S>(WaveParams, const DevScene*, PathQueue, const uint32_t*, const float4*,        \
                                              PathQueue, uint32_t*, ShadowQueue, uint32_t*, float4*, float4*,         \
                                              GuideRender);
XT_INST_SHADE(false, false, false)
XT_INST_SHADE(true, false, false)
XT_INST_SHADE(false, true, false)
XT_INST_SHADE(true, true, false)
XT_INST_SHADE(false, false, true)
XT_INST_SHADE(true, false, true)
XT_INST_SHADE(false, true, true)
XT_INST_SHADE(true, true, true)
#undef XT_INST_SHADE

// Any-hit for queued shadow rays; unoccluded contributions go straight to the
// path radiance, or to pending[] when the path continues (see k_shade).
// Bounded to 8 waves per SIMD: every form fits 64 VGPRs without spilling (the
// chunked ones take 71 unbounded and ran 6% slower at 7 waves).
template <bool COUNT, bool DYN>
__global__ __launch_bounds__(TRACE_BLOCK, 8) void k_shadow(DevScene sc, ShadowQueue sq, const uint32_t* __restrict__ count,
                                                        float4* __restrict__ radiance, float4* __restrict__ pending,
                                                        uint32_t* __restrict__ ovf, uint32_t* __restrict__ dyn,
                                                        unsigned long long* __restrict__ visit_totals,
                                                        const uint32_t* __restrict__ counts3,
                                                        uint32_t* __restrict__ counts_out) {
    __shared__ uint32_t lds[TRACE_LDS_DEPTH * TRACE_BLOCK];
    // the round's final queue counts (shade, which appends them, ran before
    // this kernel) to the host's pinned ring, in place of a copy launch
    if (counts_out && blockIdx.x == 0 && threadIdx.x < 3) counts_out[threadIdx.x] = counts3[threadIdx.x];
    __shared__ float4 top[TOP_F4];
    stage_top_nodes(top, sc);
    const uint32_t n = *count;
    const size_t nthreads = (size_t)gridDim.x * blockDim.x;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* my_lds = lds + threadIdx.x;
    uint32_t* my_ovf = ovf ? ovf + gtid : nullptr;
    const uint32_t nwaves = (uint32_t)(nthreads / 64), wave = (uint32_t)(gtid / 64);
    RayChunks ch = chunks_first<DYN>(wave, nwaves, n);
    const uint64_t below = (1ull << __lane_id()) - 1ull;
    VisitCount vc;
    bool active = false;
    uint32_t rid = 0;
    RayPre r;
    Trav t;
    while (true) {
        const uint64_t busy = __ballot(active);
        const uint32_t nidle = 64u - (uint32_t)__popcll(busy);
        const bool refill = nidle >= (uint32_t)REFILL_IDLE || busy == 0;
        if (DYN && refill && ch.next >= ch.end && !ch.exhausted) chunks_refill(ch, dyn);
        if (refill && ch.next < ch.end) {
            const uint32_t my = ch.next + (uint32_t)__popcll(~busy & below);
            if (!active && my < ch.end) {
                rid = my;
                float4 o4 = sq.o[rid], d4 = sq.d[rid];
                r = ray_setup(xyz(o4), xyz(d4));
                trav_init(t, o4.w);
                active = true;
            }
            ch.next = min(ch.next + nidle, ch.end);
        }
        if (__ballot(active) == 0) break;
        bool occ = false;
        const bool done = trav_step<true, TRACE_LDS_DEPTH, COUNT>(active, sc, r, t, top, my_lds, TRACE_BLOCK, my_ovf,
                                                                   nthreads, vc, &occ);
        if (active) {
            if (done) {
                // unoccluded contributions go straight to the path radiance, or to
                // pending[] when the path continues (see k_shade)
                float4 c4 = sq.c[rid];
                uint32_t id = __float_as_uint(sq.d[rid].w);
                f3 c = occ ? splat(0.0f) : xyz(c4);
                if (c4.w != 0.0f) {
                    pending[id] = f4(c, 0.0f);
                } else if (!occ) {
                    radiance[id] = f4(xyz(radiance[id]) + c, 0.0f);
                }
                active = false;
            }
        }
    }
    if constexpr (DYN) chunks_release(dyn);
    if (COUNT) flush_counts(visit_totals, vc);
}
#define XT_INST_SHADOW(COUNT, DYN)                                                                               \
    template __global__ void k_shadow<COUNT, DYN>(DevScene, ShadowQueue, const uint32_t*, float4*, float4*, uint32_t*, \
                                                  uint32_t*, unsigned long long*, const uint32_t*, uint32_t*);
XT_INST_SHADOW(false, false)
XT_INST_SHADOW(true, false)
XT_INST_SHADOW(false, true)
XT_INST_SHADOW(true, true)
#undef XT_INST_SHADOW

// film[pixel] += radiance of each of the batch's samples, in sample order.
__global__ __launch_bounds__(256) void k_accumulate(WaveParams wp, const float4* __restrict__ radiance, void* film,
                                                    int film_f64) {
    const BatchCursor bc = batch_of(wp);
    const uint32_t np = bc.num_pixels, S = bc.num_paths / bc.num_pixels;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        size_t o = 3 * (size_t)(bc.first_pixel + p);
        if (film_f64) {
            double* f = (double*)film;
            double r = f[o], g = f[o + 1], b = f[o + 2];
            for (uint32_t s = 0; s < S; ++s) {
                float4 L = radiance[(size_t)s * np + p];
                r += (double)L.x;
                g += (double)L.y;
                b += (double)L.z;
            }
            f[o] = r;
            f[o + 1] = g;
            f[o + 2] = b;
        } else {
            float* f = (float*)film;
            float r = f[o], g = f[o + 1], b = f[o + 2];
            for (uint32_t s = 0; s < S; ++s) {
                float4 L = radiance[(size_t)s * np + p];
                r += L.x;
                g += L.y;
                b += L.z;
            }
            f[o] = r;
            f[o + 1] = g;
            f[o + 2] = b;
        }
    }
}

// ---------------------------------------------------------------- device-side batch schedule

__device__ __forceinline__ BatchCursor cursor_at(const BatchPlan& pl, uint32_t px0, uint32_t s0, uint32_t base) {
    BatchCursor c;
    c.first_pixel = px0;
    c.num_pixels = min(pl.pix_chunk, pl.npix - px0);
    c.s_rel = s0;
    c.sample0 = base + s0;
    c.num_paths = c.num_pixels * min(pl.S, pl.spp - s0);
    return c;
}

// First batch of a render whose first sample is plan.spp_offset.
__global__ void k_cursor_init(BatchCursor* c, BatchPlan plan) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = cursor_at(plan, 0, 0, plan.spp_offset);
}

// Next batch: the next pixel chunk, or the first chunk of the next S samples.
// Past the last batch the cursor holds zero paths.
__global__ void k_advance_cursor(BatchCursor* c, BatchPlan plan) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    BatchCursor cur = *c;
    const uint32_t base = cur.sample0 - cur.s_rel;   // the render's first sample (plan.spp_offset is not used here)
    uint32_t px0 = cur.first_pixel + plan.pix_chunk;
    uint32_t s0 = cur.s_rel;
    if (px0 >= plan.npix) {
        px0 = 0;
        s0 += plan.S;
    }
    if (s0 >= plan.spp) {
        cur.num_paths = 0;
        *c = cur;
        return;
    }
    *c = cursor_at(plan, px0, s0, base);
}

__global__ void k_count_from_cursor(uint32_t* count, const BatchCursor* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = c->num_paths;
}

}  // namespace xt

// ---------------------------------------------------------------- unified batches

namespace xt {

// Window [0, window) of the batch.
__global__ void k_window_init(MultiBatch mb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MultiCursor c{};
    c.g0 = 0;
    c.n = (uint32_t)min((uint64_t)mb.window, mb.total_paths);
    *mb.cursor = c;
}

// The next window (n = 0 past the end).
__global__ void k_window_advance(MultiBatch mb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MultiCursor c = *mb.cursor;
    c.g0 += mb.window;
    c.n = c.g0 < mb.total_paths ? (uint32_t)min((uint64_t)mb.window, mb.total_paths - c.g0) : 0u;
    *mb.cursor = c;
}

// The items the window overlaps and their accumulation units: item
// first_item + j accumulates min(npix, its paths in the window) pixels,
// numbered from units[j].
__global__ void k_window_plan(MultiBatch mb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MultiCursor c = *mb.cursor;
    const uint64_t lo = c.g0, hi = c.g0 + c.n;
    uint32_t k = 0;
    while (k < mb.num_items && mb.items[k].path_base + mb.items[k].paths <= lo) ++k;
    c.first_item = k;
    uint32_t j = 0, u = 0;
    for (; c.n > 0 && k < mb.num_items && mb.items[k].path_base < hi; ++k, ++j) {
        const MultiItem& it = mb.items[k];
        const uint64_t a = max(it.path_base, lo), b = min(it.path_base + it.paths, hi);
        mb.units[j] = u;
        u += (uint32_t)min((uint64_t)it.npix, b - a);
    }
    mb.units[j] = u;
    c.num_overlap = j;
    *mb.cursor = c;
}

__global__ void k_count_from_window(uint32_t* count, const MultiCursor* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = c->n;
}

// The window's overlapping item holding window path id (largest j with
// path_base <= g0 + id).
__device__ __forceinline__ uint32_t window_item(const MultiBatch& mb, const MultiCursor& c, uint32_t id) {
    const uint64_t g = c.g0 + id;
    uint32_t lo = c.first_item, hi = c.first_item + c.num_overlap - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (mb.items[mid].path_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Camera rays of the window's paths (k_generate per item).
__global__ __launch_bounds__(256) void k_generate_multi(MultiBatch mb, PathQueue q, float4* __restrict__ radiance) {
    const MultiCursor c = *mb.cursor;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += gridDim.x * blockDim.x) {
        const uint32_t k = window_item(mb, c, id);
        const MultiItem& it = mb.items[k];
        const uint64_t local = c.g0 + id - it.path_base;
        const uint32_t s_local = (uint32_t)(local / it.npix);
        const uint32_t pixel = (uint32_t)(local - (uint64_t)s_local * it.npix);
        const WaveParams& wp = it.wp;
        const uint32_t x = pixel % wp.width, y = pixel / wp.width;
        Rng rng{make_uint2(wp.seed, wp.stream), pixel, it.spp_offset + s_local};
        uint4 b0 = rng.block(0, 0);
        float ux = Rng::to_unit(b0.x), uy = Rng::to_unit(b0.y);
        float fx = (float)x, fy = (float)y;
        fx /= (float)wp.width;
        fy /= (float)wp.height;
        fy = 1.0f - fy;
        fx += (ux - 0.5f) / (float)wp.width;
        fy += (uy - 0.5f) / (float)wp.height;
        f3 dc = normalize(mk(wp.tan_x * (fx - 0.5f) * 2, wp.tan_y * (fy - 0.5f) * 2, 1));
        f3 o = xf_point(wp.cam, mk(0, 0, 0));
        f3 d = xf_vector(wp.cam, dc);
        q.o[id] = make_float4(o.x, o.y, o.z, 0.0f);   // e and radiance: implied by vertex 1, as above
        q.d[id] = make_float4(d.x, d.y, d.z, __uint_as_float(id));
        q.w[id] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(1u));
        mb.item_of[id] = k;
    }
}

// Path id -> (pixel, sample) of a unified batch's path.
struct ItemMap {
    const MultiItem& it;
    uint64_t g0;
    __device__ __forceinline__ uint2 operator()(uint32_t id) const {
        const uint64_t local = g0 + id - it.path_base;
        const uint32_t s_local = (uint32_t)(local / it.npix);
        return make_uint2((uint32_t)(local - (uint64_t)s_local * it.npix), it.spp_offset + s_local);
    }
};

// The shade step of a unified batch: each entry with its item's parameters
// and scene (unguided).
template <bool NMAP>
__global__ __launch_bounds__(SHADE_BLOCK, NMAP ? 1 : 4) void k_shade_multi(MultiBatch mb, PathQueue qin,
                                                             const uint32_t* __restrict__ count_in,
                                                             const float4* __restrict__ hits, PathQueue qout,
                                                             uint32_t* __restrict__ count_out, ShadowQueue sq,
                                                             uint32_t* __restrict__ count_shadow,
                                                             float4* __restrict__ radiance, float4* __restrict__ pending) {
    const uint32_t n = *count_in;
    const uint64_t g0 = mb.cursor->g0;
    const uint32_t stride = gridDim.x * blockDim.x;
    __shared__ uint32_t append_lds[2 * (SHADE_BLOCK / 64) + 2];
    const GuideRender gr{};
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        ShadeOut o;
        if (i < n) {
            const MultiItem& it = mb.items[mb.item_of[__float_as_uint(qin.d[i].w)]];
            const DevScene& sc = mb.scenes[it.scene];
            shade_entry<false, NMAP>(it.wp, sc, shade_const(it.wp, sc), qin, hits, radiance, pending, gr, i,
                                     ItemMap{it, g0}, o);
        }
        shade_emit<false>(o, qout, count_out, sq, count_shadow, gr, append_lds);
    }
}
template __global__ void k_shade_multi<false>(MultiBatch, PathQueue, const uint32_t*, const float4*, PathQueue,
                                              uint32_t*, ShadowQueue, uint32_t*, float4*, float4*);
template __global__ void k_shade_multi<true>(MultiBatch, PathQueue, const uint32_t*, const float4*, PathQueue,
                                             uint32_t*, ShadowQueue, uint32_t*, float4*, float4*);

// Any-hit for a unified batch's shadow queue, each ray in its item's scene.
// 8 waves per SIMD for the split form, 7 for the chunked one (65 VGPRs; 8
// would spill).
template <bool DYN>
__global__ __launch_bounds__(TRACE_BLOCK, DYN ? 7 : 8) void k_shadow_multi(ShadowQueue sq, const uint32_t* __restrict__ count,
                                                              float4* __restrict__ radiance, float4* __restrict__ pending,
                                                              uint32_t* __restrict__ ovf, uint32_t* __restrict__ dyn,
                                                              const uint32_t* __restrict__ item_of,
                                                              const MultiItem* __restrict__ items) {
    __shared__ uint32_t lds[TRACE_LDS_DEPTH * TRACE_BLOCK];
    const uint32_t n = *count;
    const size_t nthreads = (size_t)gridDim.x * blockDim.x;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* my_lds = lds + threadIdx.x;
    uint32_t* my_ovf = ovf ? ovf + gtid : nullptr;
    const uint32_t nwaves = (uint32_t)(nthreads / 64), wave = (uint32_t)(gtid / 64);
    RayChunks ch = chunks_first<DYN>(wave, nwaves, n);
    const uint64_t below = (1ull << __lane_id()) - 1ull;
    VisitCount vc;
    bool active = false;
    uint32_t rid = 0, id = 0;
    RayPre r;
    Trav t;
    TravView view{nullptr, nullptr, 0u};
    while (true) {
        const uint64_t busy = __ballot(active);
        const uint32_t nidle = 64u - (uint32_t)__popcll(busy);
        const bool refill = nidle >= (uint32_t)REFILL_IDLE || busy == 0;
        if (DYN && refill && ch.next >= ch.end && !ch.exhausted) chunks_refill(ch, dyn);
        if (refill && ch.next < ch.end) {
            const uint32_t my = ch.next + (uint32_t)__popcll(~busy & below);
            if (!active && my < ch.end) {
                rid = my;
                float4 o4 = sq.o[rid], d4 = sq.d[rid];
                id = __float_as_uint(d4.w);
                const MultiItem& it = items[item_of[id]];
                view.nodes = it.nodes;
                view.tris = it.tris;
                r = ray_setup(xyz(o4), xyz(d4));
                trav_init(t, o4.w);
                active = true;
            }
            ch.next = min(ch.next + nidle, ch.end);
        }
        if (__ballot(active) == 0) break;
        bool occ = false;
        const bool done = trav_step<true, TRACE_LDS_DEPTH, false>(active, view, r, t, nullptr, my_lds, TRACE_BLOCK,
                                                                   my_ovf, nthreads, vc, &occ);
        if (active) {
            if (done) {
                const float4 c4 = sq.c[rid];
                const f3 c = occ ? splat(0.0f) : xyz(c4);
                if (c4.w != 0.0f) {
                    pending[id] = f4(c, 0.0f);
                } else if (!occ) {
                    radiance[id] = f4(xyz(radiance[id]) + c, 0.0f);
                }
                active = false;
            }
        }
    }
    if constexpr (DYN) chunks_release(dyn);
}
template __global__ void k_shadow_multi<false>(ShadowQueue, const uint32_t*, float4*, float4*, uint32_t*, uint32_t*,
                                               const uint32_t*, const MultiItem*);
template __global__ void k_shadow_multi<true>(ShadowQueue, const uint32_t*, float4*, float4*, uint32_t*, uint32_t*,
                                              const uint32_t*, const MultiItem*);

// film[pixel] += the window's samples of that pixel, in sample order, for
// every item the window overlaps (one thread per item pixel).
__global__ __launch_bounds__(256) void k_accumulate_multi(MultiBatch mb, const float4* __restrict__ radiance) {
    const MultiCursor c = *mb.cursor;
    if (c.n == 0) return;
    const uint32_t total = mb.units[c.num_overlap];
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = c.num_overlap - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (mb.units[mid] <= u) lo = mid;
            else hi = mid - 1;
        }
        const MultiItem& it = mb.items[c.first_item + lo];
        const uint64_t a = max(it.path_base, c.g0), b = min(it.path_base + it.paths, c.g0 + c.n);
        const uint64_t first = a - it.path_base + (u - mb.units[lo]);   // local index of this pixel's first sample
        const uint64_t last = b - it.path_base;
        const uint32_t pixel = (uint32_t)(first % it.npix);
        const size_t o = 3 * (size_t)pixel;
        if (it.film_f64) {
            double* f = (double*)it.film;
            double rr = f[o], gg = f[o + 1], bb = f[o + 2];
            for (uint64_t l = first; l < last; l += it.npix) {
                const float4 L = radiance[it.path_base + l - c.g0];
                rr += (double)L.x;
                gg += (double)L.y;
                bb += (double)L.z;
            }
            f[o] = rr;
            f[o + 1] = gg;
            f[o + 2] = bb;
        } else {
            float* f = (float*)it.film;
            float rr = f[o], gg = f[o + 1], bb = f[o + 2];
            for (uint64_t l = first; l < last; l += it.npix) {
                const float4 L = radiance[it.path_base + l - c.g0];
                rr += L.x;
                gg += L.y;
                bb += L.z;
            }
            f[o] = rr;
            f[o + 1] = gg;
            f[o + 2] = bb;
        }
    }
}

}  // namespace xt
