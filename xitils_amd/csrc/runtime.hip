// Host runtime behind include/xitils_amd.h: device handles, scene upload
// (instance flattening, light tables, BVH4 build), the ray-query entry points
// and the wavefront render loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bvh.h"
#include "kernels.h"

using namespace xt;

namespace {

thread_local std::string g_last_error;

xt_status fail(xt_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

#define XT_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(XT_ERR_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(e_));       \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

}  // namespace

struct xt_device {
    int ordinal = 0;
    int num_cu = 0;
    std::string arch;
};

struct RenderBuffers {
    uint32_t capacity = 0;
    PathQueue q[2] = {};
    ShadowQueue sq = {};
    float4* hits = nullptr;
    float4* radiance = nullptr;
    float4* pending = nullptr;
    uint32_t* counts = nullptr;      // [0], [1]: path queues, [2]: shadow queue
    uint32_t* host_counts = nullptr; // pinned
    unsigned long long* visits = nullptr;  // node, tri fetches: extension [0..1], shadow [2..3]
    void release() {
        for (int k = 0; k < 2; ++k) {
            hipFree(q[k].o); hipFree(q[k].d); hipFree(q[k].w); hipFree(q[k].e);
        }
        hipFree(sq.o); hipFree(sq.d); hipFree(sq.c);
        hipFree(hits); hipFree(radiance); hipFree(pending); hipFree(counts); hipFree(visits);
        if (host_counts) hipHostFree(host_counts);
        *this = RenderBuffers();
    }
};

struct xt_scene {
    xt_device* dev = nullptr;
    hipStream_t stream = nullptr;
    DevScene ds{};
    // device allocations
    float4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_normals = nullptr;
    xt_material* d_mats = nullptr;
    float* d_light_pdf = nullptr;
    uint2* d_light_objs = nullptr;
    uint32_t* d_light_slots = nullptr;
    uint32_t* d_ovf = nullptr;        // traversal stack spill columns
    uint32_t ovf_depth = 0;           // entries per lane beyond the LDS part
    uint32_t trace_grid_max = 0;      // blocks of the traversal kernels
    uint32_t bvh_nodes = 0, bvh_depth = 0;
    uint64_t bytes_nodes = 0, bytes_tris = 0;
    double build_ms = 0;
    RenderBuffers rb;
    xt_render_stats stats{};
    int counting = 0;                 // render with the fetch-counting traversal
    // query scratch
    float4* d_qrays = nullptr;
    size_t qrays_cap = 0;
    float4* d_qhits = nullptr;
    uint8_t* d_qocc = nullptr;
    size_t qout_cap = 0;
};

extern "C" {

const char* xt_version(void) { return "xitils_amd 0.1.0 (gfx950, BVH4 wavefront)"; }
const char* xt_last_error(void) { return g_last_error.c_str(); }

xt_status xt_device_open(int ordinal, xt_device** out) {
    if (!out) return fail(XT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= ordinal || ordinal < 0)
        return fail(XT_ERR_NO_DEVICE, "no HIP device at ordinal " + std::to_string(ordinal));
    hipDeviceProp_t prop;
    XT_HIP(hipGetDeviceProperties(&prop, ordinal));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) return fail(XT_ERR_NO_DEVICE, "device is " + arch + ", this build targets gfx950");
    XT_HIP(hipSetDevice(ordinal));
    xt_device* d = new (std::nothrow) xt_device();
    if (!d) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
    d->ordinal = ordinal;
    d->num_cu = prop.multiProcessorCount;
    d->arch = arch;
    *out = d;
    return XT_OK;
}

xt_status xt_device_close(xt_device* dev) {
    delete dev;
    return XT_OK;
}

xt_status xt_scene_destroy(xt_scene* s) {
    if (!s) return XT_OK;
    hipSetDevice(s->dev->ordinal);
    if (s->stream) hipStreamSynchronize(s->stream);
    hipFree(s->d_nodes); hipFree(s->d_tris); hipFree(s->d_normals); hipFree(s->d_mats);
    hipFree(s->d_light_pdf); hipFree(s->d_light_objs); hipFree(s->d_light_slots); hipFree(s->d_ovf);
    hipFree(s->d_qrays); hipFree(s->d_qhits); hipFree(s->d_qocc);
    s->rb.release();
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return XT_OK;
}

xt_status xt_scene_create(xt_device* dev, const xt_scene_desc* desc, xt_scene** out) {
    if (!dev || !desc || !out) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->num_objects > 0 && !desc->objects) return fail(XT_ERR_INVALID_ARGUMENT, "objects is NULL");
    if (desc->num_materials == 0 || !desc->materials) return fail(XT_ERR_INVALID_ARGUMENT, "no materials");
    for (uint32_t m = 0; m < desc->num_materials; ++m)
        if (desc->materials[m].type >= XT_MAT_COUNT_)
            return fail(XT_ERR_UNSUPPORTED_MATERIAL, "material " + std::to_string(m) + " has unknown type " +
                                                         std::to_string(desc->materials[m].type));
    auto t_start = std::chrono::steady_clock::now();
    // ---- flatten objects to world space (Object.h:106-144 applied once per vertex)
    uint64_t total = 0;
    for (uint32_t o = 0; o < desc->num_objects; ++o) total += desc->objects[o].num_triangles;
    if (total >= (1ull << 28)) return fail(XT_ERR_TOO_LARGE, "more than 2^28 triangles");
    const uint32_t nprim = (uint32_t)total;
    std::vector<BuildTri> bt(nprim);
    std::vector<uint32_t> prim_mat(nprim);
    std::vector<float> prim_nrm;   // 9 per prim when any object has normals
    std::vector<uint8_t> prim_has_n(nprim, 0);
    std::vector<float> prim_light_pdf(nprim, 0.0f);
    std::vector<uint32_t> prim_obj(nprim);
    bool any_normals = false;
    for (uint32_t o = 0; o < desc->num_objects; ++o) any_normals |= desc->objects[o].normals != nullptr;
    if (any_normals) prim_nrm.assign(9 * (size_t)nprim, 0.0f);
    // emissive objects form the light list in insertion order (Scene.h:17-22)
    std::vector<uint32_t> light_objects;
    for (uint32_t o = 0; o < desc->num_objects; ++o)
        if (desc->materials[desc->objects[o].material < desc->num_materials ? desc->objects[o].material : 0].type ==
            XT_MAT_EMISSION)
            light_objects.push_back(o);
    uint32_t base = 0;
    for (uint32_t o = 0; o < desc->num_objects; ++o) {
        const xt_object_desc& od = desc->objects[o];
        if (od.num_triangles && (!od.positions || !od.indices))
            return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " lacks positions or indices");
        if (od.material >= desc->num_materials)
            return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " material out of range");
        if (od.object_to_world && !od.world_to_object)
            return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " has a transform without its inverse");
        const float* M = od.object_to_world;
        const bool emissive_obj = desc->materials[od.material].type == XT_MAT_EMISSION;
        float area = 0.0f, scaled = 0.0f;
        for (uint32_t t = 0; t < od.num_triangles; ++t) {
            f3 p[3], w[3];
            for (int k = 0; k < 3; ++k) {
                int32_t vi = od.indices[3 * t + k];
                if (vi < 0 || (uint32_t)vi >= od.num_vertices)
                    return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " index out of range");
                p[k] = mk(od.positions[3 * vi], od.positions[3 * vi + 1], od.positions[3 * vi + 2]);
                w[k] = M ? xf_point(M, p[k]) : p[k];
                bt[base + t].p[k][0] = w[k].x;
                bt[base + t].p[k][1] = w[k].y;
                bt[base + t].p[k][2] = w[k].z;
                if (od.normals) {
                    f3 nn = mk(od.normals[3 * vi], od.normals[3 * vi + 1], od.normals[3 * vi + 2]);
                    if (M) nn = normalize(xf_vector(M, nn));
                    prim_nrm[9 * (size_t)(base + t) + 3 * k + 0] = nn.x;
                    prim_nrm[9 * (size_t)(base + t) + 3 * k + 1] = nn.y;
                    prim_nrm[9 * (size_t)(base + t) + 3 * k + 2] = nn.z;
                }
            }
            uint32_t m = od.tri_materials ? od.tri_materials[t] : od.material;
            if (m >= desc->num_materials)
                return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " triangle material out of range");
            prim_mat[base + t] = m;
            prim_has_n[base + t] = od.normals ? 1 : 0;
            prim_obj[base + t] = o;
            if (emissive_obj) {
                // TriangleMesh::surfaceAreaScaling (TriangleMesh.h:193-200) pieces
                float a = tri_area(p[0], p[1], p[2]);
                area += a;
                float aw = tri_area(w[0], w[1], w[2]);
                scaled += a * (aw / a);
            }
        }
        if (emissive_obj && od.num_triangles) {
            // Scene.h:49-55 -> Object.h:169-172 -> TriangleMesh.h:232-235 -> TriangleIndexed.h:211-213
            float scaling = scaled / area;
            for (uint32_t t = 0; t < od.num_triangles; ++t) {
                f3 p[3];
                for (int k = 0; k < 3; ++k) {
                    int32_t vi = od.indices[3 * t + k];
                    p[k] = mk(od.positions[3 * vi], od.positions[3 * vi + 1], od.positions[3 * vi + 2]);
                }
                float pdf = 1.0f / tri_area(p[0], p[1], p[2]);
                pdf = pdf / (float)od.num_triangles;
                pdf = pdf / scaling;
                pdf = pdf / (float)light_objects.size();
                prim_light_pdf[base + t] = pdf;
            }
        }
        base += od.num_triangles;
    }
    // ---- BVH
    BVHBuildOptions bopt;
    BVH4Result bvh = build_bvh4(bt, bopt);
    std::vector<uint32_t> prim_slot(nprim);
    for (uint32_t s = 0; s < nprim; ++s) prim_slot[bvh.slot_prim[s]] = s;
    // ---- leaf-ordered triangle records
    std::vector<float> tris(12 * (size_t)std::max(nprim, 1u), 0.0f), nrms;
    std::vector<float> slot_pdf(std::max(nprim, 1u), 0.0f);
    if (any_normals) nrms.assign(12 * (size_t)nprim, 0.0f);
    for (uint32_t s = 0; s < nprim; ++s) {
        uint32_t p = bvh.slot_prim[s];
        float* r = &tris[12 * (size_t)s];
        for (int k = 0; k < 3; ++k) {
            r[4 * k + 0] = bt[p].p[k][0];
            r[4 * k + 1] = bt[p].p[k][1];
            r[4 * k + 2] = bt[p].p[k][2];
        }
        uint32_t flags = (prim_has_n[p] ? TRI_FLAG_NORMALS : 0u) |
                         (desc->materials[prim_mat[p]].type == XT_MAT_EMISSION ? TRI_FLAG_EMISSIVE : 0u);
        std::memcpy(&r[3], &p, 4);
        std::memcpy(&r[7], &prim_mat[p], 4);
        std::memcpy(&r[11], &flags, 4);
        if (any_normals)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) nrms[12 * (size_t)s + 4 * k + c] = prim_nrm[9 * (size_t)p + 3 * k + c];
        slot_pdf[s] = prim_light_pdf[p];
    }
    std::vector<uint2> lobjs;
    std::vector<uint32_t> lslots;
    {
        uint32_t b2 = 0;
        std::vector<uint32_t> obj_base(desc->num_objects);
        for (uint32_t o = 0; o < desc->num_objects; ++o) {
            obj_base[o] = b2;
            b2 += desc->objects[o].num_triangles;
        }
        for (uint32_t o : light_objects) {
            const xt_object_desc& od = desc->objects[o];
            if (od.num_triangles == 0) continue;
            lobjs.push_back(make_uint2((uint32_t)lslots.size(), od.num_triangles));
            for (uint32_t t = 0; t < od.num_triangles; ++t) lslots.push_back(prim_slot[obj_base[o] + t]);
        }
    }
    // ---- upload
    xt_scene* s = new (std::nothrow) xt_scene();
    if (!s) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->dev = dev;
    auto bail = [&](xt_status st, const std::string& m) {
        xt_scene_destroy(s);
        return fail(st, m);
    };
#define XT_HIP_S(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return bail(e_ == hipErrorOutOfMemory ? XT_ERR_OUT_OF_MEMORY : XT_ERR_DEVICE, \
                                          std::string(#call) + " failed: " + hipGetErrorString(e_)); \
    } while (0)
    XT_HIP_S(hipSetDevice(dev->ordinal));
    XT_HIP_S(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    XT_HIP_S(dalloc(&s->d_nodes, (size_t)bvh.num_nodes * NODE_F4));
    XT_HIP_S(hipMemcpy(s->d_nodes, bvh.nodes.data(), bvh.nodes.size() * 4, hipMemcpyHostToDevice));
    XT_HIP_S(dalloc(&s->d_tris, tris.size() / 4));
    XT_HIP_S(hipMemcpy(s->d_tris, tris.data(), tris.size() * 4, hipMemcpyHostToDevice));
    if (any_normals) {
        XT_HIP_S(dalloc(&s->d_normals, nrms.size() / 4));
        XT_HIP_S(hipMemcpy(s->d_normals, nrms.data(), nrms.size() * 4, hipMemcpyHostToDevice));
    }
    XT_HIP_S(dalloc(&s->d_mats, desc->num_materials));
    XT_HIP_S(hipMemcpy(s->d_mats, desc->materials, desc->num_materials * sizeof(xt_material), hipMemcpyHostToDevice));
    XT_HIP_S(dalloc(&s->d_light_pdf, slot_pdf.size()));
    XT_HIP_S(hipMemcpy(s->d_light_pdf, slot_pdf.data(), slot_pdf.size() * 4, hipMemcpyHostToDevice));
    if (!lobjs.empty()) {
        XT_HIP_S(dalloc(&s->d_light_objs, lobjs.size()));
        XT_HIP_S(hipMemcpy(s->d_light_objs, lobjs.data(), lobjs.size() * sizeof(uint2), hipMemcpyHostToDevice));
        XT_HIP_S(dalloc(&s->d_light_slots, lslots.size()));
        XT_HIP_S(hipMemcpy(s->d_light_slots, lslots.data(), lslots.size() * 4, hipMemcpyHostToDevice));
    }
    // traversal stack: at most 3 pushes per level plus one
    const uint32_t stack_bound = 3 * bvh.depth + 1;
    s->ovf_depth = stack_bound > (uint32_t)TRACE_LDS_DEPTH ? stack_bound - TRACE_LDS_DEPTH : 0;
    s->trace_grid_max = (uint32_t)dev->num_cu * 8;
    if (s->ovf_depth) XT_HIP_S(dalloc(&s->d_ovf, (size_t)s->ovf_depth * s->trace_grid_max * TRACE_BLOCK));
#undef XT_HIP_S
    DevScene& ds = s->ds;
    ds.nodes = s->d_nodes;
    ds.tris = s->d_tris;
    ds.normals = s->d_normals;
    ds.mats = s->d_mats;
    ds.light_pdf = s->d_light_pdf;
    ds.light_objs = s->d_light_objs;
    ds.light_slots = s->d_light_slots;
    ds.num_lights = (uint32_t)lobjs.size();
    ds.num_prims = nprim;
    const xt_env& e = desc->env;
    ds.env.has_sky = e.has_sky;
    ds.env.sky = mk(e.sky[0], e.sky[1], e.sky[2]);
    ds.env.sun_dir = mk(e.sun_dir[0], e.sun_dir[1], e.sun_dir[2]);
    ds.env.sun_omc = e.sun_one_minus_cos;
    ds.env.sun_sin2 = e.sun_sin2;
    ds.env.sun_L = mk(e.sun_radiance[0], e.sun_radiance[1], e.sun_radiance[2]);
    ds.env.sun_frame = make_frame(ds.env.sun_dir);
    s->bvh_nodes = bvh.num_nodes;
    s->bvh_depth = bvh.depth;
    s->bytes_nodes = (uint64_t)bvh.num_nodes * 128;
    s->bytes_tris = (uint64_t)nprim * 48;
    s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    *out = s;
    return XT_OK;
}

xt_status xt_scene_info(const xt_scene* s, uint64_t out[6]) {
    if (!s || !out) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    out[0] = s->ds.num_prims;
    out[1] = s->bvh_nodes;
    out[2] = s->bvh_depth;
    out[3] = s->bytes_nodes;
    out[4] = s->bytes_tris;
    std::memcpy(&out[5], &s->build_ms, 8);
    return XT_OK;
}

static uint32_t trace_blocks(const xt_scene* s, size_t n) {
    size_t b = (n + TRACE_BLOCK - 1) / TRACE_BLOCK;
    return (uint32_t)std::max<size_t>(1, std::min<size_t>(b, s->trace_grid_max));
}

static xt_status query(xt_scene* s, const xt_ray* rays, uint32_t n, void* out, int on_device, bool any) {
    if (!s || (n && (!rays || !out))) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0) return XT_OK;
    XT_HIP(hipSetDevice(s->dev->ordinal));
    const float4* d_rays = (const float4*)rays;
    void* d_out = out;
    if (!on_device) {
        if (s->qrays_cap < n) {
            hipFree(s->d_qrays);
            XT_HIP(dalloc(&s->d_qrays, 2 * (size_t)n));
            s->qrays_cap = n;
        }
        if (s->qout_cap < n) {
            hipFree(s->d_qhits);
            hipFree(s->d_qocc);
            XT_HIP(dalloc(&s->d_qhits, (size_t)n));
            XT_HIP(dalloc(&s->d_qocc, (size_t)n));
            s->qout_cap = n;
        }
        XT_HIP(hipMemcpyAsync(s->d_qrays, rays, (size_t)n * sizeof(xt_ray), hipMemcpyHostToDevice, s->stream));
        d_rays = s->d_qrays;
        d_out = any ? (void*)s->d_qocc : (void*)s->d_qhits;
    }
    uint32_t blocks = trace_blocks(s, n);
    if (any)
        hipLaunchKernelGGL((k_trace<true, TRACE_LDS_DEPTH, false>), dim3(blocks), dim3(TRACE_BLOCK), 0, s->stream, s->ds,
                           d_rays, d_rays + 1, 2, nullptr, n, nullptr, (uint8_t*)d_out, 1, 1, s->d_ovf, nullptr);
    else
        hipLaunchKernelGGL((k_trace<false, TRACE_LDS_DEPTH, false>), dim3(blocks), dim3(TRACE_BLOCK), 0, s->stream,
                           s->ds, d_rays, d_rays + 1, 2, nullptr, n, (float4*)d_out, nullptr, 1, 1, s->d_ovf, nullptr);
    XT_HIP(hipGetLastError());
    if (!on_device)
        XT_HIP(hipMemcpyAsync(out, d_out, (size_t)n * (any ? 1 : sizeof(xt_hit)), hipMemcpyDeviceToHost, s->stream));
    XT_HIP(hipStreamSynchronize(s->stream));
    return XT_OK;
}

xt_status xt_trace_closest(xt_scene* s, const xt_ray* rays, uint32_t n, xt_hit* hits, int on_device) {
    return query(s, rays, n, hits, on_device, false);
}

xt_status xt_trace_any(xt_scene* s, const xt_ray* rays, uint32_t n, uint8_t* occluded, int on_device) {
    return query(s, rays, n, occluded, on_device, true);
}

static xt_status ensure_buffers(xt_scene* s, uint32_t capacity) {
    RenderBuffers& rb = s->rb;
    if (rb.capacity >= capacity) return XT_OK;
    rb.release();
    size_t c = capacity;
    for (int k = 0; k < 2; ++k) {
        XT_HIP(dalloc(&rb.q[k].o, c));
        XT_HIP(dalloc(&rb.q[k].d, c));
        XT_HIP(dalloc(&rb.q[k].w, c));
        XT_HIP(dalloc(&rb.q[k].e, c));
    }
    XT_HIP(dalloc(&rb.sq.o, c));
    XT_HIP(dalloc(&rb.sq.d, c));
    XT_HIP(dalloc(&rb.sq.c, c));
    XT_HIP(dalloc(&rb.hits, c));
    XT_HIP(dalloc(&rb.radiance, c));
    XT_HIP(dalloc(&rb.pending, c));
    XT_HIP(dalloc(&rb.counts, 4));
    XT_HIP(dalloc(&rb.visits, 4));
    XT_HIP(hipHostMalloc((void**)&rb.host_counts, 4 * sizeof(uint32_t), hipHostMallocDefault));
    rb.capacity = capacity;
    return XT_OK;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

struct StageTimer {
    hipStream_t st;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
    hipEvent_t mark() {
        hipEvent_t e;
        hipEventCreate(&e);
        hipEventRecord(e, st);
        return e;
    }
    void add(int stage, hipEvent_t a, hipEvent_t b) { ev[stage].push_back({a, b}); }
    double total(int stage) {
        double ms = 0;
        for (auto& p : ev[stage]) {
            float m = 0;
            hipEventElapsedTime(&m, p.first, p.second);
            ms += m;
        }
        return ms;
    }
    ~StageTimer() {
        for (auto& v : ev)
            for (auto& p : v) {
                hipEventDestroy(p.first);
                hipEventDestroy(p.second);
            }
    }
};

// Paths in flight per batch (224 B each, 3.6 GB at the default 16 Mi);
// XITILS_AMD_PATHS overrides.
static uint32_t path_capacity() {
    static uint32_t cap = [] {
        const char* e = getenv("XITILS_AMD_PATHS");
        long v = e ? atol(e) : 0;
        return v >= 65536 && v <= (1l << 26) ? (uint32_t)v : (1u << 24);
    }();
    return cap;
}

// Records the first launch failure of a render (bad configuration, missing
// code object) so it is reported instead of leaving the film untouched.
#define XT_LAUNCHED(res)                                                                            \
    do {                                                                                            \
        hipError_t e_ = hipGetLastError();                                                          \
        if (e_ != hipSuccess && (res) == XT_OK)                                                     \
            (res) = fail(XT_ERR_DEVICE, std::string("kernel launch failed at line ") +               \
                                            std::to_string(__LINE__) + ": " + hipGetErrorString(e_)); \
    } while (0)

xt_status xt_render(xt_scene* s, const xt_camera* cam, const xt_render_params* p, void* film, int film_f64,
                    int film_on_device) {
    if (!s || !cam || !p || !film) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->width == 0 || p->height == 0) return fail(XT_ERR_INVALID_ARGUMENT, "empty image");
    if (p->integrator > XT_INTEGRATOR_NAIVE) return fail(XT_ERR_INVALID_ARGUMENT, "unknown integrator");
    if ((uint64_t)p->width * p->height >= (1ull << 31)) return fail(XT_ERR_TOO_LARGE, "image too large");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return fail(XT_ERR_INVALID_ARGUMENT, "rr_prob must be in (0, 1]");
    XT_HIP(hipSetDevice(s->dev->ordinal));
    const uint64_t npix_all = (uint64_t)p->width * p->height;
    const size_t film_bytes = npix_all * 3 * (film_f64 ? 8 : 4);
    s->stats = xt_render_stats{};
    if (p->spp == 0) return XT_OK;
    const uint32_t cap = (uint32_t)std::min<uint64_t>(path_capacity(), npix_all * p->spp);
    xt_status st = ensure_buffers(s, cap);
    if (st != XT_OK) return st;
    void* d_film = film;
    if (!film_on_device) {
        XT_HIP(hipMalloc(&d_film, film_bytes));
        XT_HIP(hipMemcpyAsync(d_film, film, film_bytes, hipMemcpyHostToDevice, s->stream));
    }
    RenderBuffers& rb = s->rb;
    if (s->counting) XT_HIP(hipMemsetAsync(rb.visits, 0, 4 * sizeof(unsigned long long), s->stream));
    WaveParams wp{};
    wp.width = p->width;
    wp.height = p->height;
    wp.seed = p->seed;
    wp.stream = p->stream;
    wp.integrator = p->integrator;
    wp.flags = p->flags;
    wp.max_depth = p->max_depth;
    wp.rr_start = p->rr_start;
    wp.rr_prob = p->rr_prob;
    wp.eps = p->ray_eps;
    wp.shadow_margin = p->shadow_margin;
    wp.tan_x = cam->tan_x;
    wp.tan_y = cam->tan_y;
    std::memcpy(wp.cam, cam->cam_to_world, sizeof(wp.cam));
    StageTimer tm{s->stream};
    hipEvent_t t_begin = tm.mark();
    const uint32_t gblocks_max = (uint32_t)s->dev->num_cu * 8;
    auto blocks256 = [&](uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, gblocks_max)); };
    const uint32_t shade_max = (uint32_t)s->dev->num_cu * (2048 / SHADE_BLOCK);
    auto shade_blocks = [&](uint64_t n) {
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + SHADE_BLOCK - 1) / SHADE_BLOCK, shade_max));
    };
    // batches: S samples of a pixel range, S * pixels <= capacity
    const uint32_t S = npix_all <= cap ? (uint32_t)std::min<uint64_t>(cap / npix_all, p->spp) : 1u;
    const uint32_t pix_chunk = (uint32_t)std::min<uint64_t>(npix_all, cap);
    xt_status result = XT_OK;
    for (uint32_t s0 = 0; s0 < p->spp && result == XT_OK; s0 += S) {
        const uint32_t Sb = std::min(S, p->spp - s0);
        for (uint64_t px0 = 0; px0 < npix_all; px0 += pix_chunk) {
            const uint32_t np = (uint32_t)std::min<uint64_t>(pix_chunk, npix_all - px0);
            wp.first_pixel = (uint32_t)px0;
            wp.num_pixels = np;
            wp.num_paths = np * Sb;
            wp.sample0 = p->spp_offset + s0;
            hipEvent_t a = tm.mark();
            hipLaunchKernelGGL(k_generate, dim3(blocks256(wp.num_paths)), dim3(256), 0, s->stream, wp, rb.q[0], rb.radiance);
            XT_LAUNCHED(result);
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s->stream, rb.counts + 0, wp.num_paths);
            XT_LAUNCHED(result);
            tm.add(3, a, tm.mark());
            s->stats.camera_rays += wp.num_paths;
            uint32_t live = wp.num_paths;
            int cur = 0;
            while (live > 0) {
                s->stats.extension_rays += live;
                s->stats.extend_launches++;
                a = tm.mark();
                if (s->counting)
                    hipLaunchKernelGGL((k_trace<false, TRACE_LDS_DEPTH, true>), dim3(trace_blocks(s, live)),
                                       dim3(TRACE_BLOCK), 0, s->stream, s->ds, rb.q[cur].o, rb.q[cur].d, 1,
                                       rb.counts + cur, 0u, rb.hits, nullptr, 0, 0, s->d_ovf, rb.visits);
                else
                    hipLaunchKernelGGL((k_trace<false, TRACE_LDS_DEPTH, false>), dim3(trace_blocks(s, live)),
                                       dim3(TRACE_BLOCK), 0, s->stream, s->ds, rb.q[cur].o, rb.q[cur].d, 1,
                                       rb.counts + cur, 0u, rb.hits, nullptr, 0, 0, s->d_ovf, nullptr);
                XT_LAUNCHED(result);
                hipEvent_t b = tm.mark();
                tm.add(0, a, b);
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s->stream, rb.counts + (1 - cur), 1u);
                XT_LAUNCHED(result);
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s->stream, rb.counts + 2, 1u);
                XT_LAUNCHED(result);
                a = tm.mark();
                hipLaunchKernelGGL(k_shade, dim3(shade_blocks(live)), dim3(SHADE_BLOCK), 0, s->stream, wp, s->ds, rb.q[cur],
                                   rb.counts + cur, rb.hits, rb.q[1 - cur], rb.counts + (1 - cur), rb.sq, rb.counts + 2,
                                   rb.radiance, rb.pending);
                XT_LAUNCHED(result);
                b = tm.mark();
                tm.add(1, a, b);
                hipMemcpyAsync(rb.host_counts, rb.counts, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
                if (hipStreamSynchronize(s->stream) != hipSuccess) {
                    result = fail(XT_ERR_DEVICE, std::string("render failed: ") + hipGetErrorString(hipGetLastError()));
                    break;
                }
                const uint32_t next = rb.host_counts[1 - cur], nsh = rb.host_counts[2];
                if (nsh > 0) {
                    s->stats.shadow_rays += nsh;
                    a = tm.mark();
                    if (s->counting)
                        hipLaunchKernelGGL(k_shadow<true>, dim3(trace_blocks(s, nsh)), dim3(TRACE_BLOCK), 0, s->stream,
                                           s->ds, rb.sq, rb.counts + 2, rb.radiance, rb.pending, s->d_ovf, rb.visits + 2);
                    else
                        hipLaunchKernelGGL(k_shadow<false>, dim3(trace_blocks(s, nsh)), dim3(TRACE_BLOCK), 0, s->stream,
                                           s->ds, rb.sq, rb.counts + 2, rb.radiance, rb.pending, s->d_ovf, nullptr);
                    XT_LAUNCHED(result);
                    tm.add(2, a, tm.mark());
                }
                live = next;
                cur = 1 - cur;
            }
            if (result != XT_OK) break;
            a = tm.mark();
            hipLaunchKernelGGL(k_accumulate, dim3(blocks256(np)), dim3(256), 0, s->stream, wp, rb.radiance, d_film, film_f64);
            XT_LAUNCHED(result);
            tm.add(3, a, tm.mark());
        }
    }
    hipEvent_t t_end = tm.mark();
    hipError_t e = hipStreamSynchronize(s->stream);
    if (result == XT_OK && e != hipSuccess) result = fail(XT_ERR_DEVICE, std::string("render failed: ") + hipGetErrorString(e));
    if (result == XT_OK && hipGetLastError() != hipSuccess) result = fail(XT_ERR_DEVICE, "kernel launch failed");
    if (result == XT_OK) {
        float ms = 0;
        hipEventElapsedTime(&ms, t_begin, t_end);
        s->stats.ms_total = ms;
        s->stats.ms_extend = tm.total(0);
        s->stats.ms_shade = tm.total(1);
        s->stats.ms_shadow = tm.total(2);
        s->stats.ms_other = tm.total(3);
        if (s->counting) {
            unsigned long long v[4] = {0, 0, 0, 0};
            if (hipMemcpy(v, rb.visits, sizeof v, hipMemcpyDeviceToHost) == hipSuccess) {
                s->stats.ext_node_fetches = v[0];
                s->stats.ext_tri_fetches = v[1];
                s->stats.shadow_node_fetches = v[2];
                s->stats.shadow_tri_fetches = v[3];
            }
        }
    }
    hipEventDestroy(t_begin);
    hipEventDestroy(t_end);
    if (!film_on_device) {
        if (result == XT_OK) {
            hipError_t ce = hipMemcpy(film, d_film, film_bytes, hipMemcpyDeviceToHost);
            if (ce != hipSuccess) result = fail(XT_ERR_DEVICE, "film copy failed");
        }
        hipFree(d_film);
    }
    return result;
}

xt_status xt_scene_set_counting(xt_scene* s, int on) {
    if (!s) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    s->counting = on ? 1 : 0;
    return XT_OK;
}

xt_status xt_render_get_stats(const xt_scene* s, xt_render_stats* out) {
    if (!s || !out) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = s->stats;
    return XT_OK;
}

}  // extern "C"
