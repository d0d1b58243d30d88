// Host runtime behind include/xitils_amd.h: device handles, scene upload
// (instance flattening, light tables, BVH build), the ray-query entry points
// and the wavefront render loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bvh.h"
#include "guide.h"
#include "kernels.h"
#include "runtime.h"

using namespace xt;

namespace {
thread_local std::string g_last_error;
}  // namespace

xt_status xt_fail(xt_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

static xt_status fail(xt_status s, const std::string& msg) { return xt_fail(s, msg); }

constexpr int RB_RING = 4;   // bounce rounds whose queue counts can be in flight to the host

struct RenderBuffers {
    uint32_t capacity = 0;
    PathQueue q[2] = {};
    ShadowQueue sq = {};
    float4* hits = nullptr;
    float4* radiance = nullptr;
    float4* pending = nullptr;
    uint32_t* rec[2] = {};           // training-record index per path-queue entry (allocated on demand)
    float* lobes = nullptr;          // decoded guide mixtures per path-queue entry (allocated on demand)
    uint32_t* counts = nullptr;      // [0], [1]: path queues, [2]: shadow queue
    uint32_t* host_counts = nullptr; // pinned, RB_RING slots of 4 counts (bounce rounds in flight)
    uint32_t* host_counts_dev = nullptr;  // the same memory's device address (written by k_shadow)
    hipEvent_t round_ev[RB_RING] = {};
    unsigned long long* visits = nullptr;  // node, tri fetches: extension [0..1], shadow [2..3]
    uint32_t* dyn = nullptr;               // chunk cursors of the traversal kernels: extension [0..1], shadow [2..3]
    void release() {
        for (int k = 0; k < 2; ++k) {
            hipFree(q[k].o); hipFree(q[k].d); hipFree(q[k].w); hipFree(q[k].e);
        }
        hipFree(sq.o); hipFree(sq.d); hipFree(sq.c);
        hipFree(hits); hipFree(radiance); hipFree(pending); hipFree(counts); hipFree(visits); hipFree(dyn);
        hipFree(rec[0]); hipFree(rec[1]); hipFree(lobes);
        if (host_counts) hipHostFree(host_counts);
        for (hipEvent_t& e : round_ev)
            if (e) hipEventDestroy(e);
        *this = RenderBuffers();
    }
};

// Guiding network and training state of a scene.
struct GuideState {
    bool ready = false;
    xt_guide_config cfg{};
    GuideBounds gb{};
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    float* P = nullptr;              // fp32 master parameters
    float* m = nullptr;              // Adam moments
    float* v = nullptr;
    float* partial = nullptr;        // per-block gradient rows of k_guide_grad
    uint32_t partial_rows = 0;
    void* net_mem = nullptr;         // GuideChain + GuideBack + the two published GuideChains
    GuideNet net{};                  // the copy training reads and Adam writes
    uint64_t step = 0;               // Adam steps taken
    TrainRec* recs = nullptr;
    GuideSample* samples[2] = {};    // finalised samples of a training render (bursts alternate)
    uint32_t rec_cap = 0;
    uint32_t* d_count = nullptr;     // finalised samples
    float* d_wsum = nullptr;         // their weight sum
    float* host_io = nullptr;        // pinned: count + weight sum read-back
    // A training render's Adam steps (a burst) run on train_stream while the
    // next render runs: burst b reads samples[b % 2] and ends by publishing
    // the network into pub[b % 2].  A training render reads pub[bursts % 2]
    // (the burst before last: the last one may still run), any other render
    // pub[(bursts + 1) % 2] (the last burst, waited for).
    hipStream_t train_stream = nullptr;
    GuideChain* pub[2] = {};
    hipEvent_t pub_ev[2] = {}, t0[2] = {}, t1[2] = {};
    bool pub_pending[2] = {}, unreported[2] = {};
    uint64_t bursts = 0;
    double train_ms_carry = 0.0;     // burst time folded into the next render's stats
    void release() {
        if (train_stream) hipStreamSynchronize(train_stream);
        hipFree(P); hipFree(m); hipFree(v); hipFree(partial); hipFree(net_mem);
        hipFree(recs); hipFree(samples[0]); hipFree(samples[1]); hipFree(d_count); hipFree(d_wsum);
        if (host_io) hipHostFree(host_io);
        for (hipEvent_t e : {pub_ev[0], pub_ev[1], t0[0], t0[1], t1[0], t1[1]})
            if (e) hipEventDestroy(e);
        if (train_stream) hipStreamDestroy(train_stream);
        *this = GuideState();
    }
};

// One render's bounce loop captured as a HIP graph: generate, max_depth + 1
// rounds of extend / shade / shadow, accumulate, advance the batch cursor.
// Every kernel reads its queue count from device memory, so one replay
// renders one batch with no host round trip; replays walk the batches.
struct GraphKey {
    xt_camera cam;
    xt_render_params p;
    void* film;
    int32_t f64;
    uint32_t trace_grid, shade_grid, capacity;
};
struct RenderGraph {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
    uint32_t nbatches = 0;
    BatchPlan plan{};
};

struct xt_scene {
    xt_device* dev = nullptr;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev_pool;   // StageTimer events, reused across renders
    DevScene ds{};
    DevScene* d_ds = nullptr;         // device copy of ds (the shade kernel reads it through a pointer)
    // device allocations
    float4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_normals = nullptr;
    float4* d_nmap_attr = nullptr;    // normal-map attributes per slot
    bool nmap = false;                // normal maps or microfacet normal mapping: the NMAP shade kernels
    DevTexture* d_textures = nullptr;
    std::vector<float4*> d_texels;    // one allocation per texture
    xt_material* d_mats = nullptr;
    float* d_light_pdf = nullptr;
    uint2* d_light_objs = nullptr;
    uint32_t* d_light_slots = nullptr;
    uint32_t* d_ovf = nullptr;        // traversal stack spill columns
    float4* d_map = nullptr;          // image sky texels
    float* d_map_pdf = nullptr;       // lat-long importance distribution
    float* d_map_marg = nullptr;
    float* d_map_cond = nullptr;
    uint32_t ovf_depth = 0;           // entries per lane beyond the LDS part
    uint32_t trace_grid_max = 0;      // blocks of the traversal kernels (sizes the spill columns)
    uint32_t trace_grid = 0;          // blocks launched (<= trace_grid_max; xt_scene_autotune)
    uint32_t shade_grid = 0;          // blocks of the shade kernel (xt_scene_autotune)
    BatchCursor* d_cursor = nullptr;  // device-side batch schedule of graph replays
    std::vector<struct RenderGraph> graphs;   // instantiated bounce loops (xt_render_batch)
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
    unsigned long long* d_qvisits = nullptr;   // node, tri fetches of the last counted ray query
    uint32_t* d_qdyn = nullptr;                // chunk cursor of the ray-query traversal
    float bounds_lo[3] = {0, 0, 0}, bounds_hi[3] = {0, 0, 0};
    GuideState guide;
};

using ShadeKernel = void (*)(WaveParams, const DevScene*, PathQueue, const uint32_t*, const float4*, PathQueue, uint32_t*,
                            ShadowQueue, uint32_t*, float4*, float4*, GuideRender);
// The shade kernel for a render: guided or not, with the normal-mapping
// materials or not, and with the block-local material partition
// (XT_FLAG_SORT_MATERIALS) or not -- separate instantiations, so the plain
// kernel carries none of the others' registers.
static ShadeKernel shade_kernel(bool guide, bool nmap, uint32_t flags) {
    if (flags & XT_FLAG_SORT_MATERIALS) {
        if (nmap) return guide ? k_shade<true, true, true> : k_shade<false, true, true>;
        return guide ? k_shade<true, false, true> : k_shade<false, false, true>;
    }
    if (nmap) return guide ? k_shade<true, true, false> : k_shade<false, true, false>;
    return guide ? k_shade<true, false, false> : k_shade<false, false, false>;
}

extern "C" {

#define XT_STR2(x) #x
#define XT_STR(x) XT_STR2(x)
const char* xt_version(void) { return "xitils_amd 0.1.0 (gfx950, BVH" XT_STR(XT_BVH_WIDTH) " wavefront)"; }
const char* xt_last_error(void) { return g_last_error.c_str(); }

// Static scratch of xt_render_static: every buffer of its bounce loop is
// carved from this array, so that entry point allocates nothing.  Each GPU
// the library runs on holds its own copy from code-object load until the
// process ends.  In host code `scratch` names a host-side stand-in, not
// device memory: the device address comes from hipGetSymbolAddress, once per
// device at xt_device_open, and only that address is used.
static __device__ __align__(256) unsigned char scratch[XT_STATIC_SCRATCH_BYTES];

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
    void* scratch_ptr = nullptr;   // this device's copy of the static scratch (the device is current)
    XT_HIP(hipGetSymbolAddress(&scratch_ptr, HIP_SYMBOL(scratch)));
    if (!scratch_ptr) return fail(XT_ERR_DEVICE, "hipGetSymbolAddress returned no address for the static scratch");
    xt_device* d = new (std::nothrow) xt_device();
    if (!d) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
    d->scratch = (unsigned char*)scratch_ptr;
    d->ordinal = ordinal;
    d->num_cu = prop.multiProcessorCount;
    d->arch = arch;
    *out = d;
    return XT_OK;
}

static void multi_release(xt_device* dev);

xt_status xt_device_close(xt_device* dev) {
    if (dev) {
        multi_release(dev);
        hipSetDevice(dev->ordinal);
        hipDeviceSynchronize();
        delete dev->pool;
    }
    delete dev;
    return XT_OK;
}

xt_status xt_device_valu_rate(xt_device* dev, double* rate) {
    if (!dev || !rate) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    XT_HIP(hipSetDevice(dev->ordinal));
    const uint32_t blocks = (uint32_t)dev->num_cu * 8, iters = 4096;   // 256 threads: 8 waves per SIMD
    float* out = nullptr;
    XT_HIP(dalloc(&out, blocks));
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    float best = INFINITY;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {   // the first launch warms the clocks
        hipEventRecord(a, nullptr);
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, nullptr, iters, out);
        e = hipGetLastError();
        hipEventRecord(b, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess && rep > 0 && hipEventElapsedTime(&ms, a, b) == hipSuccess && ms > 0.0f)
            best = std::min(best, ms);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipFree(out);
    if (e != hipSuccess) return fail(XT_ERR_DEVICE, std::string("VALU probe failed: ") + hipGetErrorString(e));
    if (!std::isfinite(best)) return fail(XT_ERR_DEVICE, "VALU probe could not be timed");
    const double waves = (double)blocks * 256.0 / 64.0;
    *rate = waves * (double)iters * VALU_PROBE_CHAINS / (best * 1e-3);
    return XT_OK;
}

xt_status xt_scene_destroy(xt_scene* s) {
    if (!s) return XT_OK;
    hipSetDevice(s->dev->ordinal);
    if (s->stream) hipStreamSynchronize(s->stream);
    hipFree(s->d_nodes); hipFree(s->d_tris); hipFree(s->d_normals); hipFree(s->d_mats);
    hipFree(s->d_nmap_attr); hipFree(s->d_textures);
    for (float4* t : s->d_texels) hipFree(t);
    hipFree(s->d_light_pdf); hipFree(s->d_light_objs); hipFree(s->d_light_slots); hipFree(s->d_ovf);
    hipFree(s->d_qrays); hipFree(s->d_qhits); hipFree(s->d_qocc); hipFree(s->d_qvisits); hipFree(s->d_qdyn);
    hipFree(s->d_map); hipFree(s->d_map_pdf); hipFree(s->d_map_marg); hipFree(s->d_map_cond);
    for (RenderGraph& g : s->graphs) hipGraphExecDestroy(g.exec);
    hipFree(s->d_cursor);
    hipFree(s->d_ds);
    s->rb.release();
    s->guide.release();
    for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return XT_OK;
}

// Lat-long importance distribution of an image sky: texel weight = the
// largest luminance of its 3x3 neighbourhood (the bilinear lookup inside a
// texel reads those texels) times sin(theta) at the row centre; a marginal
// row CDF and per-row column CDFs, summed in double; pdf = weight / total *
// W H / (2 pi^2), divided by sin(theta) at lookup.
struct MapDistribution {
    std::vector<float> pdf, marg, cond;
};
static MapDistribution build_map_distribution(const float* rgb, int W, int H) {
    auto lum = [&](int x, int y) {
        x %= W;
        if (x < 0) x += W;
        y = std::min(std::max(y, 0), H - 1);
        const float* t = rgb + 3 * ((size_t)y * W + x);
        return 0.2126 * t[0] + 0.7152 * t[1] + 0.0722 * t[2];
    };
    std::vector<double> w((size_t)W * H);
    for (int y = 0; y < H; ++y) {
        const double st = std::sin(M_PI * (y + 0.5) / H);
        for (int x = 0; x < W; ++x) {
            double m = 0.0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) m = std::max(m, lum(x + dx, y + dy));
            w[(size_t)y * W + x] = m * st;
        }
    }
    MapDistribution d;
    std::vector<double> row(H, 0.0);
    double total = 0.0;
    d.cond.assign((size_t)H * (W + 1), 0.0f);
    for (int y = 0; y < H; ++y) {
        double acc = 0.0;
        for (int x = 0; x < W; ++x) acc += w[(size_t)y * W + x];
        row[y] = acc;
        total += acc;
        double run = 0.0;
        for (int x = 0; x < W; ++x) {
            d.cond[(size_t)y * (W + 1) + x] = acc > 0.0 ? (float)(run / acc) : (float)x / (float)W;
            run += w[(size_t)y * W + x];
        }
        d.cond[(size_t)y * (W + 1) + W] = 1.0f;
    }
    d.marg.assign(H + 1, 0.0f);
    double run = 0.0;
    for (int y = 0; y < H; ++y) {
        d.marg[y] = total > 0.0 ? (float)(run / total) : (float)y / (float)H;
        run += row[y];
    }
    d.marg[H] = 1.0f;
    d.pdf.assign((size_t)W * H, 0.0f);
    for (size_t k = 0; k < w.size(); ++k)
        d.pdf[k] = total > 0.0 ? (float)(w[k] / total * W * H / (2.0 * M_PI * M_PI)) : 0.0f;
    return d;
}

xt_status xt_scene_create(xt_device* dev, const xt_scene_desc* desc, xt_scene** out) {
    if (!dev || !desc || !out) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->num_objects > 0 && !desc->objects) return fail(XT_ERR_INVALID_ARGUMENT, "objects is NULL");
    if (desc->num_materials == 0 || !desc->materials) return fail(XT_ERR_INVALID_ARGUMENT, "no materials");
    const xt_env& env_in = desc->env;
    if (env_in.map_kind > XT_ENV_MAP_LATLONG) return fail(XT_ERR_INVALID_ARGUMENT, "unknown env map kind");
    if (env_in.map_kind != XT_ENV_MAP_NONE) {
        if (!env_in.map_rgb || env_in.map_width == 0 || env_in.map_height == 0)
            return fail(XT_ERR_INVALID_ARGUMENT, "env map without texels");
        if ((uint64_t)env_in.map_width * env_in.map_height > (1ull << 28))
            return fail(XT_ERR_TOO_LARGE, "env map larger than 2^28 texels");
    }
    for (uint32_t m = 0; m < desc->num_materials; ++m) {
        const xt_material& mm = desc->materials[m];
        if (mm.type >= XT_MAT_COUNT_)
            return fail(XT_ERR_UNSUPPORTED_MATERIAL, "material " + std::to_string(m) + " has unknown type " +
                                                         std::to_string(mm.type));
        if (mm.type == XT_MAT_MICROFACET_NMAP) {
            const float bt = mm.p[0];
            if (!(bt == (float)XT_MAT_DIFFUSE || bt == (float)XT_MAT_METAL || bt == (float)XT_MAT_GGX_LAMBERT))
                return fail(XT_ERR_UNSUPPORTED_MATERIAL, "material " + std::to_string(m) +
                                                             ": microfacet normal mapping needs a Diffuse, Metal or "
                                                             "GGX+Lambert base");
        }
    }
    if (desc->num_textures > 0 && !desc->textures) return fail(XT_ERR_INVALID_ARGUMENT, "textures is NULL");
    for (uint32_t k = 0; k < desc->num_textures; ++k) {
        const xt_texture& t = desc->textures[k];
        if (!t.rgb || t.width == 0 || t.height == 0)
            return fail(XT_ERR_INVALID_ARGUMENT, "texture " + std::to_string(k) + " has no texels");
        if ((uint64_t)t.width * t.height > (1ull << 28))
            return fail(XT_ERR_TOO_LARGE, "texture " + std::to_string(k) + " larger than 2^28 texels");
    }
    const uint32_t* mat_nmap = desc->material_normal_maps;
    if (mat_nmap)
        for (uint32_t m = 0; m < desc->num_materials; ++m)
            if (mat_nmap[m] > desc->num_textures)
                return fail(XT_ERR_INVALID_ARGUMENT, "material " + std::to_string(m) + " normal map out of range");
    auto nmap_of = [&](uint32_t m) -> uint32_t { return mat_nmap ? mat_nmap[m] : 0u; };
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
        if ((od.tangents != nullptr) != (od.bitangents != nullptr))
            return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + " has tangents without bitangents");
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
                // the BVH's quantisation grid needs finite world positions of sane magnitude
                if (!(std::fabs(w[k].x) <= 1e30f && std::fabs(w[k].y) <= 1e30f && std::fabs(w[k].z) <= 1e30f))
                    return fail(XT_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) +
                                                             " has a vertex that is not finite or exceeds 1e30 in world space");
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
    BVHResult bvh = build_bvh(bt, bopt);
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
        if (const uint32_t nm = nmap_of(prim_mat[p])) {
            flags |= nm << TRI_NMAP_SHIFT;
            if (desc->objects[prim_obj[p]].tangents) flags |= TRI_FLAG_TANGENTS;
        }
        std::memcpy(&r[3], &p, 4);
        std::memcpy(&r[7], &prim_mat[p], 4);
        std::memcpy(&r[11], &flags, 4);
        if (any_normals)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) nrms[12 * (size_t)s + 4 * k + c] = prim_nrm[9 * (size_t)p + 3 * k + c];
        slot_pdf[s] = prim_light_pdf[p];
    }
    std::vector<uint32_t> obj_base(desc->num_objects);
    {
        uint32_t b2 = 0;
        for (uint32_t o = 0; o < desc->num_objects; ++o) {
            obj_base[o] = b2;
            b2 += desc->objects[o].num_triangles;
        }
    }
    // ---- normal-map attributes per slot (tangent frame in world space, as the normals)
    bool any_nmap = false;
    for (uint32_t p = 0; p < nprim && !any_nmap; ++p) any_nmap = nmap_of(prim_mat[p]) != 0;
    std::vector<float4> nmap_attr;
    if (any_nmap) {
        nmap_attr.assign((size_t)nprim * NMAP_F4, make_float4(0, 0, 0, 0));
        for (uint32_t sl = 0; sl < nprim; ++sl) {
            const uint32_t p = bvh.slot_prim[sl];
            if (!nmap_of(prim_mat[p])) continue;
            const uint32_t o = prim_obj[p];
            const xt_object_desc& od = desc->objects[o];
            const uint32_t t = p - obj_base[o];
            const float* M = od.object_to_world;
            float4* a = &nmap_attr[(size_t)sl * NMAP_F4];
            float uvs[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 3; ++k) {
                const int32_t vi = od.indices[3 * t + k];
                if (od.texcoords) {
                    uvs[2 * k] = od.texcoords[2 * vi];
                    uvs[2 * k + 1] = od.texcoords[2 * vi + 1];
                }
                if (od.tangents) {
                    f3 tg = mk(od.tangents[3 * vi], od.tangents[3 * vi + 1], od.tangents[3 * vi + 2]);
                    f3 bn = mk(od.bitangents[3 * vi], od.bitangents[3 * vi + 1], od.bitangents[3 * vi + 2]);
                    if (M) {
                        tg = normalize(xf_vector(M, tg));
                        bn = normalize(xf_vector(M, bn));
                    }
                    a[k] = make_float4(tg.x, tg.y, tg.z, 0.0f);
                    a[3 + k] = make_float4(bn.x, bn.y, bn.z, 0.0f);
                }
            }
            for (int k = 0; k < 6; ++k) a[k].w = uvs[k];
        }
    }
    std::vector<uint2> lobjs;
    std::vector<uint32_t> lslots;
    {
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
    if (any_nmap) {
        XT_HIP_S(dalloc(&s->d_nmap_attr, nmap_attr.size()));
        XT_HIP_S(hipMemcpy(s->d_nmap_attr, nmap_attr.data(), nmap_attr.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (desc->num_textures > 0) {
        std::vector<DevTexture> dt(desc->num_textures);
        for (uint32_t k = 0; k < desc->num_textures; ++k) {
            const xt_texture& t = desc->textures[k];
            std::vector<float4> tex((size_t)t.width * t.height);
            for (size_t q = 0; q < tex.size(); ++q)
                tex[q] = make_float4(t.rgb[3 * q], t.rgb[3 * q + 1], t.rgb[3 * q + 2], 0.0f);
            float4* d = nullptr;
            XT_HIP_S(dalloc(&d, tex.size()));
            s->d_texels.push_back(d);
            XT_HIP_S(hipMemcpy(d, tex.data(), tex.size() * sizeof(float4), hipMemcpyHostToDevice));
            dt[k] = DevTexture{(int)t.width, (int)t.height, d};
        }
        XT_HIP_S(dalloc(&s->d_textures, dt.size()));
        XT_HIP_S(hipMemcpy(s->d_textures, dt.data(), dt.size() * sizeof(DevTexture), hipMemcpyHostToDevice));
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
    const uint32_t stack_bound = (BVH_WIDTH - 1) * bvh.depth + 1;
    s->ovf_depth = stack_bound > (uint32_t)TRACE_LDS_DEPTH ? stack_bound - TRACE_LDS_DEPTH : 0;
    // Persistent traversal grid: the waves split the queue into fixed ranges,
    // so launch no more blocks per CU than can be resident at once (the
    // kernels' registers and LDS, capped at 8).
    int resident = 8;
    {
        int a = 0, b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_trace<false, TRACE_LDS_DEPTH, false>, TRACE_BLOCK, 0) ==
                hipSuccess && a > 0)
            resident = std::min(resident, a);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow<false>, TRACE_BLOCK, 0) == hipSuccess && b > 0)
            resident = std::min(resident, b);
    }
    s->trace_grid_max = (uint32_t)dev->num_cu * (uint32_t)resident;
    s->trace_grid = s->trace_grid_max;
    s->shade_grid = (uint32_t)dev->num_cu * (2048 / SHADE_BLOCK);
    if (s->ovf_depth) XT_HIP_S(dalloc(&s->d_ovf, (size_t)s->ovf_depth * s->trace_grid_max * TRACE_BLOCK));
#undef XT_HIP_S
    DevScene& ds = s->ds;
    ds.nodes = s->d_nodes;
    ds.tris = s->d_tris;
    ds.normals = s->d_normals;
    ds.nmap_attr = s->d_nmap_attr;
    s->nmap = any_nmap;
    for (uint32_t m = 0; m < desc->num_materials; ++m) s->nmap |= desc->materials[m].type == XT_MAT_MICROFACET_NMAP;
    ds.textures = s->d_textures;
    ds.mats = s->d_mats;
    ds.light_pdf = s->d_light_pdf;
    ds.light_objs = s->d_light_objs;
    ds.light_slots = s->d_light_slots;
    ds.num_lights = (uint32_t)lobjs.size();
    ds.num_prims = nprim;
    ds.top_nodes = std::min<uint32_t>(bvh.top_nodes, (uint32_t)BVH_TOP_MAX_NODES);
    const xt_env& e = desc->env;
    ds.env.has_sky = e.has_sky;
    ds.env.sky = mk(e.sky[0], e.sky[1], e.sky[2]);
    ds.env.sun_dir = mk(e.sun_dir[0], e.sun_dir[1], e.sun_dir[2]);
    ds.env.sun_omc = e.sun_one_minus_cos;
    ds.env.sun_sin2 = e.sun_sin2;
    ds.env.sun_L = mk(e.sun_radiance[0], e.sun_radiance[1], e.sun_radiance[2]);
    ds.env.sun_frame = make_frame(ds.env.sun_dir);
    ds.env.map_kind = (int)e.map_kind;
    if (e.map_kind != XT_ENV_MAP_NONE) {
        const int W = (int)e.map_width, H = (int)e.map_height;
        std::vector<float4> tex((size_t)W * H);
        for (size_t k = 0; k < tex.size(); ++k)
            tex[k] = make_float4(e.map_rgb[3 * k], e.map_rgb[3 * k + 1], e.map_rgb[3 * k + 2], 0.0f);
        auto up = [&](auto** dst, const auto& v) -> hipError_t {
            hipError_t r = dalloc(dst, v.size());
            if (r == hipSuccess) r = hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
            return r;
        };
        hipError_t r = up(&s->d_map, tex);
        if (r == hipSuccess && e.map_kind == XT_ENV_MAP_LATLONG) {
            MapDistribution md = build_map_distribution(e.map_rgb, W, H);
            r = up(&s->d_map_pdf, md.pdf);
            if (r == hipSuccess) r = up(&s->d_map_marg, md.marg);
            if (r == hipSuccess) r = up(&s->d_map_cond, md.cond);
        }
        if (r != hipSuccess) {
            xt_scene_destroy(s);
            return fail(r == hipErrorOutOfMemory ? XT_ERR_OUT_OF_MEMORY : XT_ERR_DEVICE,
                        std::string("env map upload failed: ") + hipGetErrorString(r));
        }
        ds.env.map_w = W;
        ds.env.map_h = H;
        ds.env.map = s->d_map;
        ds.env.map_pdf = s->d_map_pdf;
        ds.env.map_marg = s->d_map_marg;
        ds.env.map_cond = s->d_map_cond;
    }
    s->bvh_nodes = bvh.num_nodes;
    s->bvh_depth = bvh.depth;
    s->bytes_nodes = (uint64_t)bvh.num_nodes * NODE_F4 * 16;
    s->bytes_tris = (uint64_t)nprim * 48;
    for (int a = 0; a < 3; ++a) {
        s->bounds_lo[a] = INFINITY;
        s->bounds_hi[a] = -INFINITY;
    }
    for (const BuildTri& t : bt)
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) {
                s->bounds_lo[a] = std::min(s->bounds_lo[a], t.p[k][a]);
                s->bounds_hi[a] = std::max(s->bounds_hi[a], t.p[k][a]);
            }
    {
        hipError_t r = dalloc(&s->d_ds, 1);
        if (r == hipSuccess) r = hipMemcpy(s->d_ds, &ds, sizeof(DevScene), hipMemcpyHostToDevice);
        if (r != hipSuccess) {
            xt_scene_destroy(s);
            return fail(r == hipErrorOutOfMemory ? XT_ERR_OUT_OF_MEMORY : XT_ERR_DEVICE,
                        std::string("scene record upload failed: ") + hipGetErrorString(r));
        }
    }
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
    return (uint32_t)std::max<size_t>(1, std::min<size_t>(b, s->trace_grid));
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
    const bool count = s->counting != 0;
    if (!s->d_qdyn) {
        XT_HIP(dalloc(&s->d_qdyn, 2));
        XT_HIP(hipMemset(s->d_qdyn, 0, 2 * sizeof(uint32_t)));
    }
    if (count) {
        if (!s->d_qvisits) XT_HIP(dalloc(&s->d_qvisits, 2));
        XT_HIP(hipMemsetAsync(s->d_qvisits, 0, 2 * sizeof(unsigned long long), s->stream));
    }
    hipEvent_t ev[2];
    XT_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        hipEventDestroy(ev[0]);
        return fail(XT_ERR_DEVICE, "hipEventCreate failed");
    }
    uint32_t blocks = trace_blocks(s, n);
    float4* hits = any ? nullptr : (float4*)d_out;
    uint8_t* occ = any ? (uint8_t*)d_out : nullptr;
    hipEventRecord(ev[0], s->stream);
    const bool dyn = n >= TRAV_DYN_MIN;
#define XT_QUERY(ANY, COUNT)                                                                                      \
    hipLaunchKernelGGL((dyn ? k_trace<ANY, TRACE_LDS_DEPTH, COUNT, false, true>                                    \
                            : k_trace<ANY, TRACE_LDS_DEPTH, COUNT, false, false>),                                 \
                       dim3(blocks), dim3(TRACE_BLOCK), 0, s->stream, s->ds,                                      \
                       d_rays, d_rays + 1, 2, nullptr, n, hits, occ, 1, 1, s->d_ovf, s->d_qdyn, s->d_qvisits, nullptr, \
                       nullptr, nullptr, nullptr)
    if (any && count) {
        XT_QUERY(true, true);
    } else if (any) {
        XT_QUERY(true, false);
    } else if (count) {
        XT_QUERY(false, true);
    } else {
        XT_QUERY(false, false);
    }
#undef XT_QUERY
    hipEventRecord(ev[1], s->stream);
    hipError_t le = hipGetLastError();
    if (le == hipSuccess && !on_device)
        le = hipMemcpyAsync(out, d_out, (size_t)n * (any ? 1 : sizeof(xt_hit)), hipMemcpyDeviceToHost, s->stream);
    hipError_t se = hipStreamSynchronize(s->stream);
    float ms = 0.0f;
    if (le == hipSuccess && se == hipSuccess) hipEventElapsedTime(&ms, ev[0], ev[1]);
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (le != hipSuccess) return fail(XT_ERR_DEVICE, std::string("ray query failed: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(XT_ERR_DEVICE, std::string("ray query failed: ") + hipGetErrorString(se));
    s->stats = xt_render_stats{};
    s->stats.ms_total = ms;
    unsigned long long v[2] = {0, 0};
    if (count) XT_HIP(hipMemcpy(v, s->d_qvisits, sizeof v, hipMemcpyDeviceToHost));
    if (any) {
        s->stats.shadow_rays = n;
        s->stats.ms_shadow = ms;
        s->stats.shadow_node_fetches = v[0];
        s->stats.shadow_tri_fetches = v[1];
    } else {
        s->stats.extension_rays = n;
        s->stats.ms_extend = ms;
        s->stats.ext_node_fetches = v[0];
        s->stats.ext_tri_fetches = v[1];
    }
    return XT_OK;
}

xt_status xt_trace_closest(xt_scene* s, const xt_ray* rays, uint32_t n, xt_hit* hits, int on_device) {
    return query(s, rays, n, hits, on_device, false);
}

xt_status xt_trace_any(xt_scene* s, const xt_ray* rays, uint32_t n, uint8_t* occluded, int on_device) {
    return query(s, rays, n, occluded, on_device, true);
}

static void drop_graphs(xt_scene* s) {
    if (!s->graphs.empty()) hipStreamSynchronize(s->stream);   // none may still be replaying
    for (RenderGraph& g : s->graphs) hipGraphExecDestroy(g.exec);
    s->graphs.clear();
}

// (Re)allocates rb for `capacity` paths; its previous contents must be idle.
static xt_status alloc_buffers(RenderBuffers& rb, uint32_t capacity) {
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
    XT_HIP(dalloc(&rb.dyn, 4));
    XT_HIP(hipMemset(rb.dyn, 0, 4 * sizeof(uint32_t)));
    XT_HIP(hipHostMalloc((void**)&rb.host_counts, RB_RING * 4 * sizeof(uint32_t),
                         hipHostMallocMapped | hipHostMallocCoherent));
    XT_HIP(hipHostGetDevicePointer((void**)&rb.host_counts_dev, rb.host_counts, 0));
    for (hipEvent_t& e : rb.round_ev) XT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    rb.capacity = capacity;
    return XT_OK;
}

// The scene's own buffers (its captured graphs use them).
static xt_status ensure_buffers(xt_scene* s, uint32_t capacity) {
    if (s->rb.capacity >= capacity) return XT_OK;
    drop_graphs(s);   // graphs hold the old buffer addresses
    hipStreamSynchronize(s->stream);
    return alloc_buffers(s->rb, capacity);
}

// Path buffers of xt_render's bounce loop, shared by the device's scenes: a
// render takes a free set (grown to the render's capacity if smaller) and
// gives it back when it ends, so memory follows the renders in flight (a
// batch's concurrent renders), not the number of scenes.  xt_render is
// synchronous, so a returned set is idle.
struct RenderPool {
    std::mutex mu;
    std::vector<RenderBuffers*> sets;
    std::vector<char> busy;
    ~RenderPool() {
        for (RenderBuffers* b : sets) {
            b->release();
            delete b;
        }
    }
};
struct PooledBuffers {
    RenderPool* pool = nullptr;
    size_t idx = 0;
    RenderBuffers* rb = nullptr;
    PooledBuffers() = default;
    PooledBuffers(const PooledBuffers&) = delete;
    PooledBuffers& operator=(const PooledBuffers&) = delete;
    ~PooledBuffers() {
        if (rb) {
            std::lock_guard<std::mutex> lock(pool->mu);
            pool->busy[idx] = 0;
        }
    }
};
static xt_status acquire_buffers(xt_device* dev, uint32_t capacity, PooledBuffers& out) {
    RenderPool* pool;
    {
        std::lock_guard<std::mutex> lock(dev->pool_mu);
        if (!dev->pool) dev->pool = new (std::nothrow) RenderPool();
        pool = dev->pool;
    }
    if (!pool) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
    {
        std::lock_guard<std::mutex> lock(pool->mu);
        // the smallest free set that holds the render, else the largest free set, else a new one
        size_t pick = pool->sets.size();
        for (size_t i = 0; i < pool->sets.size(); ++i) {
            if (pool->busy[i]) continue;
            if (pick == pool->sets.size()) {
                pick = i;
                continue;
            }
            const uint32_t ci = pool->sets[i]->capacity, cp = pool->sets[pick]->capacity;
            const bool fits_i = ci >= capacity, fits_p = cp >= capacity;
            if ((fits_i && (!fits_p || ci < cp)) || (!fits_i && !fits_p && ci > cp)) pick = i;
        }
        if (pick == pool->sets.size()) {
            RenderBuffers* b = new (std::nothrow) RenderBuffers();
            if (!b) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
            pool->sets.push_back(b);
            pool->busy.push_back(0);
        }
        pool->busy[pick] = 1;
        out.pool = pool;
        out.idx = pick;
        out.rb = pool->sets[pick];
    }
    if (out.rb->capacity < capacity) return alloc_buffers(*out.rb, capacity);
    return XT_OK;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// Per-stage kernel time of one render from event pairs on its stream.  The
// events come from the scene's pool (created once, reused by every render
// after the previous one synchronised), so a bounce round costs the host
// event records, not creations.
struct StageTimer {
    hipStream_t st;
    std::vector<hipEvent_t>* pool;
    size_t used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[6];   // extend, shade, shadow, other, guide infer, guide train
    hipEvent_t mark() {
        hipEvent_t e = nullptr;
        if (used < pool->size()) {
            e = (*pool)[used];
        } else {
            hipEventCreate(&e);
            pool->push_back(e);
        }
        ++used;
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
};

// Paths in flight per batch of xt_render (224 B each, 30 GB at the default
// 128 Mi; the guiding lobes add 128 B per path to a guided render's set);
// XITILS_AMD_PATHS overrides (65536 .. 256 Mi).  Larger batches spend less
// of a render in its last, sparse bounce rounds: on the 1080p heightfield at
// 64 spp, 1349 Mpaths/s at 16 Mi paths (8 spp per batch), 1468 at 32 Mi,
// 1543-1559 at 64 Mi, 1605-1608 at 128 Mi (the whole frame in one batch).
static uint32_t path_capacity() {
    static uint32_t cap = [] {
        const char* e = getenv("XITILS_AMD_PATHS");
        long v = e ? atol(e) : 0;
        return v >= 65536 && v <= (1l << 28) ? (uint32_t)v : (1u << 27);
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

// ---------------------------------------------------------------- guiding (host side)

static uint64_t splitmix64(uint64_t& x) {   // Steele, Lea, Flood: SplitMix64
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void guide_init_params(uint32_t seed, std::vector<float>& P) {
    P.assign(G_NPARAMS, 0.0f);
    uint64_t st = 0x5EEDull * 0x100000001ull + seed;
    auto uni = [&](float bound) { return bound * (2.0f * (float)((splitmix64(st) >> 40) * 0x1.0p-24) - 1.0f); };
    for (int i = 0; i < G_HID * G_IN; ++i) P[G_OFF_W0 + i] = uni(std::sqrt(6.0f / G_IN));
    for (int i = 0; i < G_HID * G_HID; ++i) P[G_OFF_W1 + i] = uni(std::sqrt(6.0f / G_HID));
    for (int i = 0; i < G_HID * G_HID; ++i) P[G_OFF_W2 + i] = uni(std::sqrt(6.0f / G_HID));
    for (int i = 0; i < G_OUT * G_HID; ++i) P[G_OFF_W3 + i] = uni(0.1f * std::sqrt(6.0f / G_HID));
}

// Both published copies <- the training copy, on s->stream (the train
// stream must be idle).
static xt_status guide_publish_all(xt_scene* s) {
    GuideState& g = s->guide;
    for (int k = 0; k < 2; ++k) {
        XT_HIP(hipMemcpyAsync(g.pub[k], g.net.chain, sizeof(GuideChain), hipMemcpyDeviceToDevice, s->stream));
        g.pub_pending[k] = false;
    }
    return XT_OK;
}

// The network copies from the fp32 parameters, both published copies
// included (the train stream must be idle).
static xt_status guide_prepare(xt_scene* s) {
    GuideState& g = s->guide;
    hipLaunchKernelGGL(k_guide_prepare, dim3((G_HID * GC_LD0 + 255) / 256), dim3(256), 0, s->stream, g.P, g.net);
    if (hipGetLastError() != hipSuccess) return fail(XT_ERR_DEVICE, "k_guide_prepare launch failed");
    return guide_publish_all(s);
}

// Folds the times of finished training bursts into the stats (wait: block
// until they finish; slot >= 0: only that slot's burst).
static void guide_collect_bursts(xt_scene* s, bool wait, int slot = -1) {
    GuideState& g = s->guide;
    for (int k = 0; k < 2; ++k) {
        if (!g.unreported[k] || (slot >= 0 && k != slot)) continue;
        if ((wait ? hipEventSynchronize(g.t1[k]) : hipEventQuery(g.t1[k])) != hipSuccess) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, g.t0[k], g.t1[k]) == hipSuccess) g.train_ms_carry += ms;
        g.unreported[k] = false;
    }
}

// Every guide API call first lets the training bursts in flight finish.
static xt_status guide_need(xt_scene* s) {
    if (!s) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!s->guide.ready) return fail(XT_ERR_INVALID_ARGUMENT, "guiding not initialised (xt_guide_init)");
    XT_HIP(hipSetDevice(s->dev->ordinal));
    XT_HIP(hipStreamSynchronize(s->guide.train_stream));
    guide_collect_bursts(s, true);
    return XT_OK;
}

// s->stream waits (on the device) for the last training burst: captured
// guided renders read the training copy, final once that burst has ended.
static xt_status guide_wait_last_burst(xt_scene* s) {
    GuideState& g = s->guide;
    const int k = (int)((g.bursts + 1) % 2);
    if (g.ready && g.pub_pending[k]) XT_HIP(hipStreamWaitEvent(s->stream, g.pub_ev[k], 0));
    return XT_OK;
}

// Gradient of the minibatch d_samples[0..n) into g.partial rows [0, *rows):
// one fused kernel, a block per 128 samples (at most partial_rows blocks,
// each taking every rows-th sample group).
static xt_status guide_gradient_rows(xt_scene* s, hipStream_t stream, const GuideSample* d_samples, uint32_t n,
                                     float weight_scale, float weight_clip, uint32_t* rows) {
    GuideState& g = s->guide;
    const uint32_t groups = (n + GG_BLOCK_SAMPLES - 1) / GG_BLOCK_SAMPLES;
    const uint32_t blocks = std::max(1u, std::min(groups, g.partial_rows));
    hipLaunchKernelGGL(k_guide_grad, dim3(blocks), dim3(GG_BLOCK), 0, stream, g.net, g.gb, d_samples, n, weight_scale,
                       weight_clip, g.partial);
    XT_HIP(hipGetLastError());
    *rows = blocks;
    return XT_OK;
}

// Adam steps over d_samples[0..n) in minibatches, weights scaled by weight_scale.
static xt_status guide_train(xt_scene* s, hipStream_t stream, const GuideSample* d_samples, uint32_t n,
                             float weight_scale, uint32_t max_steps, uint64_t* steps_taken) {
    GuideState& g = s->guide;
    const uint32_t M = std::max(64u, g.cfg.minibatch);
    const uint32_t per_epoch = (n + M - 1) / M;
    const uint32_t steps = std::min<uint32_t>(per_epoch * std::max(1u, g.cfg.epochs), std::max(1u, max_steps));
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t off = (k % per_epoch) * M, cnt = std::min(M, n - off);
        uint32_t grid = 0;
        xt_status gst = guide_gradient_rows(s, stream, d_samples + off, cnt, weight_scale, g.cfg.weight_clip, &grid);
        if (gst != XT_OK) return gst;
        s->stats.guide_grad_samples += cnt;
        g.step++;
        const float b1t = (float)std::pow((double)g.cfg.beta1, (double)g.step);
        const float b2t = (float)std::pow((double)g.cfg.beta2, (double)g.step);
        // Adam also re-packs the bf16 copies the next step and the renders read
        hipLaunchKernelGGL(k_guide_adam, dim3((G_NPARAMS + 31) / 32), dim3(GA_BLOCK), 0, stream, g.P, g.m, g.v, g.partial,
                           grid, g.cfg.learning_rate, g.cfg.beta1, g.cfg.beta2, g.cfg.adam_eps, b1t, b2t,
                           1.0f / (float)cnt, g.net);
        if (hipGetLastError() != hipSuccess) return fail(XT_ERR_DEVICE, "guide training launch failed");
    }
    if (steps_taken) *steps_taken += steps;
    return XT_OK;
}

static xt_status guide_ensure_training(xt_scene* s) {
    GuideState& g = s->guide;
    if (g.recs) return XT_OK;
    g.rec_cap = std::max(1024u, g.cfg.max_records);
    XT_HIP(dalloc(&g.recs, g.rec_cap));
    XT_HIP(dalloc(&g.samples[0], g.rec_cap));
    XT_HIP(dalloc(&g.samples[1], g.rec_cap));
    return XT_OK;
}

// Host samples -> device GuideSample array (grows a scratch buffer).
static xt_status guide_upload_samples(xt_scene* s, const float* queries, const float* dirs, const float* weights,
                                      uint32_t n, GuideSample** out) {
    GuideState& g = s->guide;
    if (g.rec_cap < n || !g.samples[0]) {
        hipFree(g.recs);
        hipFree(g.samples[0]);
        hipFree(g.samples[1]);
        g.recs = nullptr;
        g.samples[0] = g.samples[1] = nullptr;
        g.rec_cap = std::max(n, g.cfg.max_records);
        XT_HIP(dalloc(&g.recs, g.rec_cap));
        XT_HIP(dalloc(&g.samples[0], g.rec_cap));
        XT_HIP(dalloc(&g.samples[1], g.rec_cap));
    }
    std::vector<GuideSample> hs(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 9; ++k) hs[i].q[k] = queries[9 * (size_t)i + k];
        for (int k = 0; k < 3; ++k) hs[i].dir[k] = dirs[3 * (size_t)i + k];
        hs[i].weight = weights[i];
    }
    XT_HIP(hipMemcpyAsync(g.samples[0], hs.data(), n * sizeof(GuideSample), hipMemcpyHostToDevice, s->stream));
    XT_HIP(hipStreamSynchronize(s->stream));
    *out = g.samples[0];
    return XT_OK;
}

xt_status xt_render(xt_scene* s, const xt_camera* cam, const xt_render_params* p, void* film, int film_f64,
                    int film_on_device) {
    if (!s || !cam || !p || !film) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->width == 0 || p->height == 0) return fail(XT_ERR_INVALID_ARGUMENT, "empty image");
    if (p->integrator > XT_INTEGRATOR_NAIVE) return fail(XT_ERR_INVALID_ARGUMENT, "unknown integrator");
    if ((uint64_t)p->width * p->height >= (1ull << 31)) return fail(XT_ERR_TOO_LARGE, "image too large");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return fail(XT_ERR_INVALID_ARGUMENT, "rr_prob must be in (0, 1]");
    const bool guide = (p->flags & XT_FLAG_GUIDE) != 0;
    const bool train = guide && (p->flags & XT_FLAG_GUIDE_TRAIN) != 0;
    if (guide && p->integrator != XT_INTEGRATOR_STANDARD)
        return fail(XT_ERR_INVALID_ARGUMENT, "guiding needs the standard integrator");
    if (guide && !s->guide.ready) return fail(XT_ERR_INVALID_ARGUMENT, "guiding not initialised (xt_guide_init)");
    XT_HIP(hipSetDevice(s->dev->ordinal));
    const uint64_t npix_all = (uint64_t)p->width * p->height;
    const size_t film_bytes = npix_all * 3 * (film_f64 ? 8 : 4);
    s->stats = xt_render_stats{};
    if (s->guide.ready) guide_collect_bursts(s, false);
    if (p->spp == 0) return XT_OK;
    const uint32_t cap = (uint32_t)std::min<uint64_t>(path_capacity(), npix_all * p->spp);
    PooledBuffers pooled;
    xt_status st = acquire_buffers(s->dev, cap, pooled);
    if (st != XT_OK) return st;
    RenderBuffers& rbg = *pooled.rb;
    if (guide && !rbg.lobes) XT_HIP(dalloc(&rbg.lobes, (size_t)G_LOBE_FIELDS * rbg.capacity));
    if (train) {
        if (!rbg.rec[0]) {
            XT_HIP(dalloc(&rbg.rec[0], rbg.capacity));
            XT_HIP(dalloc(&rbg.rec[1], rbg.capacity));
        }
        st = guide_ensure_training(s);
        if (st != XT_OK) return st;
    }
    for (int k = 0; k < 2; ++k) rbg.q[k].rec = train ? rbg.rec[k] : nullptr;
    GuideState& g = s->guide;
    void* d_film = film;
    if (!film_on_device) {
        XT_HIP(hipMalloc(&d_film, film_bytes));
        XT_HIP(hipMemcpyAsync(d_film, film, film_bytes, hipMemcpyHostToDevice, s->stream));
    }
    RenderBuffers& rb = rbg;
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
    StageTimer tm{s->stream, &s->ev_pool};
    hipEvent_t t_begin = tm.mark();
    const uint32_t gblocks_max = (uint32_t)s->dev->num_cu * 8;
    auto blocks256 = [&](uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, gblocks_max)); };
    const uint32_t shade_max = s->shade_grid;
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
            // The network this batch samples with: a training batch the one
            // published by the burst before last (its own samples and the
            // burst it starts take that slot next), any other the last one.
            GuideNet net_r{};
            int tslot = 0;
            if (guide) {
                tslot = (int)(g.bursts % 2);
                const int rd = train ? tslot : 1 - tslot;
                if (g.pub_pending[rd] && hipStreamWaitEvent(s->stream, g.pub_ev[rd], 0) != hipSuccess) {
                    result = fail(XT_ERR_DEVICE, "hipStreamWaitEvent failed");
                    break;
                }
                net_r = GuideNet{g.pub[rd], g.net.back};
            }
            hipEvent_t a = tm.mark();
            hipLaunchKernelGGL(k_generate, dim3(blocks256(wp.num_paths)), dim3(256), 0, s->stream, wp, rb.q[0], rb.radiance);
            XT_LAUNCHED(result);
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s->stream, rb.counts + 0, wp.num_paths);
            XT_LAUNCHED(result);
            tm.add(3, a, tm.mark());
            s->stats.camera_rays += wp.num_paths;
            // Bounce rounds.  Every kernel reads its queue length from device
            // memory; the host reads each round's counts from a pinned ring
            // `lag` rounds late, so the GPU does not wait for the host between
            // bounces.  Grids are sized from the last count read, an upper
            // bound of every later queue (a path adds at most one extension
            // and one shadow ray per round).  Rounds launched after the queue
            // ran empty exit at once.  Training renders read each round's
            // counts before the next (the record offsets must be exact).
            const int lag = train ? 0 : 2;
            uint32_t bound = wp.num_paths;         // >= every queue length from here on
            uint32_t queue_len[RB_RING];           // queue length of each unverified round
            int round = 0, verified = 0, cur = 0;
            uint32_t rec_base = 0;
            bool launch = true;
            queue_len[0] = wp.num_paths;
            // a path adds one extension ray per round, so max_depth + 1 rounds hold every live queue
            const int max_rounds = p->max_depth > 0 ? p->max_depth + 2 + lag : (1 << 16);
            // Once a read count is zero every later round is empty: the host
            // moves on (accumulate, next batch) without waiting for them.
            while (result == XT_OK && launch) {
                if (round >= max_rounds) {
                    result = fail(XT_ERR_DEVICE, "bounce loop did not drain its queue");
                    break;
                }
                {
                    const int slot = round % RB_RING;
                    a = tm.mark();
                    // large queues: rays handed out in chunks from a cursor (the
                    // round's shadow queue is no longer than its path queue)
                    const bool dyn = bound >= TRAV_DYN_MIN;
                    auto trace = s->counting ? (dyn ? k_trace<false, TRACE_LDS_DEPTH, true, false, true>
                                                    : k_trace<false, TRACE_LDS_DEPTH, true, false, false>)
                                             : (dyn ? k_trace<false, TRACE_LDS_DEPTH, false, false, true>
                                                    : k_trace<false, TRACE_LDS_DEPTH, false, false, false>);
                    hipLaunchKernelGGL(trace, dim3(trace_blocks(s, bound)), dim3(TRACE_BLOCK), 0, s->stream, s->ds,
                                       rb.q[cur].o, rb.q[cur].d, 1, rb.counts + cur, 0u, rb.hits, nullptr, 0, 0,
                                       s->d_ovf, rb.dyn, s->counting ? rb.visits : nullptr, rb.counts + (1 - cur),
                                       rb.counts + 2, nullptr, nullptr);   // also clears the next counters
                    XT_LAUNCHED(result);
                    tm.add(0, a, tm.mark());
                    GuideRender gr{};
                    if (guide) {
                        a = tm.mark();
                        const uint32_t ib = std::max(1u, std::min((bound + GI_BLOCK - 1) / GI_BLOCK,
                                                                  (uint32_t)s->dev->num_cu * 4));
                        uint2* l4 = (uint2*)rb.lobes;
                        float2* l2 = (float2*)(rb.lobes + (size_t)2 * G_LOBES * rb.capacity);
                        hipLaunchKernelGGL(k_guide_infer, dim3(ib), dim3(GI_BLOCK), 0, s->stream, net_r, g.gb, s->ds,
                                           rb.q[cur], rb.counts + cur, rb.hits, l4, l2, rb.capacity,
                                           g.cfg.max_guided_vertex);
                        XT_LAUNCHED(result);
                        tm.add(4, a, tm.mark());
                        gr.lobes_mk = l4;
                        gr.lobes2 = l2;
                        gr.stride = rb.capacity;
                        gr.prob = g.cfg.guide_prob;
                        gr.max_vertex = g.cfg.max_guided_vertex;
                        if (train) {
                            gr.recs = g.recs;
                            gr.rec_base = rec_base;
                            gr.rec_cap = g.rec_cap;
                        }
                    }
                    a = tm.mark();
                    hipLaunchKernelGGL(shade_kernel(guide, s->nmap, p->flags), dim3(shade_blocks(bound)), dim3(SHADE_BLOCK), 0,
                                       s->stream, wp, s->d_ds, rb.q[cur], rb.counts + cur, rb.hits, rb.q[1 - cur],
                                       rb.counts + (1 - cur), rb.sq, rb.counts + 2, rb.radiance, rb.pending, gr);
                    XT_LAUNCHED(result);
                    tm.add(1, a, tm.mark());
                    a = tm.mark();
                    auto shadow = s->counting ? (dyn ? k_shadow<true, true> : k_shadow<true, false>)
                                              : (dyn ? k_shadow<false, true> : k_shadow<false, false>);
                    hipLaunchKernelGGL(shadow, dim3(trace_blocks(s, bound)), dim3(TRACE_BLOCK), 0, s->stream, s->ds,
                                       rb.sq, rb.counts + 2, rb.radiance, rb.pending, s->d_ovf, rb.dyn + 2,
                                       s->counting ? rb.visits + 2 : nullptr, rb.counts,
                                       rb.host_counts_dev + 4 * slot);   // counts -> pinned ring
                    XT_LAUNCHED(result);
                    tm.add(2, a, tm.mark());
                    if (hipEventRecord(rb.round_ev[slot], s->stream) != hipSuccess) {
                        result = fail(XT_ERR_DEVICE, "hipEventRecord failed");
                        break;
                    }
                    ++round;
                    cur = 1 - cur;
                }
                // read back every round at least `lag` behind the last launch
                while (launch && verified <= round - 1 - lag) {
                    const int vs = verified % RB_RING;
                    if (hipEventSynchronize(rb.round_ev[vs]) != hipSuccess) {
                        result = fail(XT_ERR_DEVICE, std::string("render failed: ") + hipGetErrorString(hipGetLastError()));
                        break;
                    }
                    const uint32_t* hc = rb.host_counts + 4 * vs;
                    const uint32_t len = queue_len[vs];
                    const uint32_t next = hc[(verified & 1) ? 0 : 1], nsh = hc[2];
                    if (len > 0) {
                        s->stats.extension_rays += len;
                        s->stats.extend_launches++;
                        if (guide) s->stats.guide_queries += len;
                    }
                    s->stats.shadow_rays += nsh;
                    if (train) rec_base = (uint32_t)std::min<uint64_t>((uint64_t)rec_base + next, g.rec_cap);
                    queue_len[(verified + 1) % RB_RING] = next;
                    bound = next;
                    if (next == 0) launch = false;
                    ++verified;
                }
            }
            if (result != XT_OK) break;
            a = tm.mark();
            hipLaunchKernelGGL(k_accumulate, dim3(blocks256(np)), dim3(256), 0, s->stream, wp, rb.radiance, d_film, film_f64);
            XT_LAUNCHED(result);
            tm.add(3, a, tm.mark());
            if (train && rec_base > 0 && result == XT_OK) {
                // finished paths -> training samples -> Adam steps on the
                // train stream, which then publish the network into
                // pub[tslot]; the next render runs meanwhile
                a = tm.mark();
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s->stream, g.d_count, 1u);
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s->stream, (uint32_t*)g.d_wsum, 1u);
                hipLaunchKernelGGL(k_guide_finalize, dim3((rec_base + GF_BLOCK - 1) / GF_BLOCK), dim3(GF_BLOCK), 0, s->stream,
                                   g.recs, rec_base,
                                   rb.radiance, g.cfg.train_target, g.samples[tslot], g.d_count, g.d_wsum);
                XT_LAUNCHED(result);
                hipMemcpyAsync(g.host_io, g.d_count, 4, hipMemcpyDeviceToHost, s->stream);
                hipMemcpyAsync(g.host_io + 1, g.d_wsum, 4, hipMemcpyDeviceToHost, s->stream);
                if (hipStreamSynchronize(s->stream) != hipSuccess) {
                    result = fail(XT_ERR_DEVICE, "guide finalisation failed");
                    break;
                }
                uint32_t ns;
                std::memcpy(&ns, g.host_io, 4);
                const float wsum = g.host_io[1];
                tm.add(3, a, tm.mark());
                if (ns > 0 && wsum > 0.0f && std::isfinite(wsum)) {
                    // the finalisation has ended (synchronised above); so has
                    // the burst that last used this slot (waited for before
                    // this batch)
                    guide_collect_bursts(s, true, tslot);
                    hipStream_t ts_ = g.train_stream;
                    xt_status ts = hipEventRecord(g.t0[tslot], ts_) == hipSuccess ? XT_OK
                                   : fail(XT_ERR_DEVICE, "hipEventRecord failed");
                    if (ts == XT_OK)
                        ts = guide_train(s, ts_, g.samples[tslot], ns, (float)ns / wsum, g.cfg.max_steps_per_batch,
                                         &s->stats.guide_steps);
                    if (ts == XT_OK && (hipMemcpyAsync(g.pub[tslot], g.net.chain, sizeof(GuideChain),
                                                       hipMemcpyDeviceToDevice, ts_) != hipSuccess ||
                                        hipEventRecord(g.pub_ev[tslot], ts_) != hipSuccess ||
                                        hipEventRecord(g.t1[tslot], ts_) != hipSuccess))
                        ts = fail(XT_ERR_DEVICE, "guide publication failed");
                    if (ts != XT_OK) {
                        if (result == XT_OK) result = ts;
                        break;
                    }
                    g.pub_pending[tslot] = true;
                    g.unreported[tslot] = true;
                    g.bursts++;
                    s->stats.guide_samples += ns;
                }
            }
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
        s->stats.ms_guide_infer = tm.total(4);
        if (guide) {   // training bursts that have ended
            guide_collect_bursts(s, false);
            s->stats.ms_guide_train = s->guide.train_ms_carry;
            s->guide.train_ms_carry = 0.0;
        }
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
    if (!film_on_device) {
        if (result == XT_OK) {
            hipError_t ce = hipMemcpy(film, d_film, film_bytes, hipMemcpyDeviceToHost);
            if (ce != hipSuccess) result = fail(XT_ERR_DEVICE, "film copy failed");
        }
        hipFree(d_film);
    }
    return result;
}

// ---------------------------------------------------------------- graph-replayed renders

static xt_status check_render_args(xt_scene* s, const xt_camera* cam, const xt_render_params* p) {
    if (!s || !cam || !p) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->width == 0 || p->height == 0) return fail(XT_ERR_INVALID_ARGUMENT, "empty image");
    if (p->integrator > XT_INTEGRATOR_NAIVE) return fail(XT_ERR_INVALID_ARGUMENT, "unknown integrator");
    if ((uint64_t)p->width * p->height >= (1ull << 31)) return fail(XT_ERR_TOO_LARGE, "image too large");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return fail(XT_ERR_INVALID_ARGUMENT, "rr_prob must be in (0, 1]");
    if (p->flags & XT_FLAG_GUIDE_TRAIN) return fail(XT_ERR_INVALID_ARGUMENT, "graph renders do not train (use xt_render)");
    if ((p->flags & XT_FLAG_GUIDE) && (p->integrator != XT_INTEGRATOR_STANDARD || !s->guide.ready))
        return fail(XT_ERR_INVALID_ARGUMENT, "guiding needs the standard integrator and xt_guide_init");
    if (p->max_depth <= 0 || p->max_depth > 1024)
        return fail(XT_ERR_INVALID_ARGUMENT, "graph renders need a bounce cap (max_depth in 1..1024)");
    return XT_OK;
}

// Paths in flight per batch of a graph render: at most 4 Mi (0.9 GB of path
// state), so many scenes of a batch can hold their buffers at once.
static uint32_t graph_capacity(uint64_t npix, uint32_t spp) {
    return (uint32_t)std::min<uint64_t>(std::min<uint32_t>(path_capacity(), 1u << 22), npix * spp);
}

// Path state of one count-driven batch (the graph renders' buffers, or the
// static scratch of xt_render_static).
struct BounceBufs {
    PathQueue q[2];
    ShadowQueue sq;
    float4* hits;
    float4* radiance;
    float4* pending;
    uint32_t* counts;        // [0], [1]: path queues, [2]: shadow queue
    uint32_t* ovf;           // traversal spill columns for tblocks blocks
    uint32_t capacity;       // paths
};

// WaveParams of a count-driven render: the batch fields come from `cursor`.
static WaveParams cursor_wave_params(const xt_camera* cam, const xt_render_params* p, const BatchCursor* cursor) {
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
    wp.cursor = cursor;
    return wp;
}

// Batch schedule of a count-driven render with `cap` paths per batch.
static BatchPlan cursor_plan(const xt_render_params* p, uint32_t cap, uint32_t* nbatches) {
    const uint64_t npix = (uint64_t)p->width * p->height;
    BatchPlan plan;
    plan.npix = (uint32_t)npix;
    plan.S = npix <= cap ? (uint32_t)std::min<uint64_t>(cap / npix, p->spp) : 1u;
    plan.pix_chunk = (uint32_t)std::min<uint64_t>(npix, cap);
    plan.spp = p->spp;
    plan.spp_offset = p->spp_offset;
    *nbatches = ((p->spp + plan.S - 1) / plan.S) * (uint32_t)((npix + plan.pix_chunk - 1) / plan.pix_chunk);
    return plan;
}

// Guide inputs of a count-driven batch (lobes_mk == nullptr: unguided).
struct BounceGuide {
    uint2* lobes_mk = nullptr;
    float2* lobes2 = nullptr;
    uint32_t stride = 0;
};

// Enqueues the bounce loop of the cursor's batch on s->stream: generate,
// max_depth + 1 rounds of extend / (guide) / shade / shadow, accumulate,
// advance the cursor.  Every kernel reads its queue length from device memory.
static void enqueue_cursor_batch(xt_scene* s, const xt_render_params* p, const WaveParams& wp, const BounceBufs& b,
                                 const BatchPlan& plan, BatchCursor* cursor, uint32_t tblocks, const BounceGuide& gd,
                                 void* d_film, int f64) {
    const uint32_t maxpaths = plan.pix_chunk * plan.S;
    const uint32_t gblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((maxpaths + 255) / 256, (uint64_t)s->dev->num_cu * 8));
    const uint32_t ablocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((plan.pix_chunk + 255) / 256, (uint64_t)s->dev->num_cu * 8));
    const uint32_t sblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((maxpaths + SHADE_BLOCK - 1) / SHADE_BLOCK, s->shade_grid));
    const bool guide = gd.lobes_mk != nullptr;
    GuideState& g = s->guide;
    const uint32_t ib = std::max(1u, std::min((maxpaths + GI_BLOCK - 1) / GI_BLOCK, (uint32_t)s->dev->num_cu * 4));
    hipLaunchKernelGGL(k_generate, dim3(gblocks), dim3(256), 0, s->stream, wp, b.q[0], b.radiance);
    hipLaunchKernelGGL(k_count_from_cursor, dim3(1), dim3(1), 0, s->stream, b.counts + 0, cursor);
    int cur = 0;
    for (int it = 0; it <= p->max_depth; ++it) {
        hipLaunchKernelGGL((k_trace<false, TRACE_LDS_DEPTH, false>), dim3(tblocks), dim3(TRACE_BLOCK), 0, s->stream, s->ds,
                           b.q[cur].o, b.q[cur].d, 1, b.counts + cur, 0u, b.hits, nullptr, 0, 0, b.ovf, nullptr, nullptr,
                           b.counts + (1 - cur), b.counts + 2);
        GuideRender gr{};
        if (guide) {
            hipLaunchKernelGGL(k_guide_infer, dim3(ib), dim3(GI_BLOCK), 0, s->stream, g.net, g.gb, s->ds, b.q[cur],
                               b.counts + cur, b.hits, gd.lobes_mk, gd.lobes2, gd.stride, g.cfg.max_guided_vertex);
            gr.lobes_mk = gd.lobes_mk;
            gr.lobes2 = gd.lobes2;
            gr.stride = gd.stride;
            gr.prob = g.cfg.guide_prob;
            gr.max_vertex = g.cfg.max_guided_vertex;
        }
        hipLaunchKernelGGL(shade_kernel(guide, s->nmap, p->flags), dim3(sblocks), dim3(SHADE_BLOCK), 0, s->stream, wp, s->d_ds,
                           b.q[cur], b.counts + cur, b.hits, b.q[1 - cur], b.counts + (1 - cur), b.sq,
                           b.counts + 2, b.radiance, b.pending, gr);
        hipLaunchKernelGGL(k_shadow<false>, dim3(tblocks), dim3(TRACE_BLOCK), 0, s->stream, s->ds, b.sq, b.counts + 2,
                           b.radiance, b.pending, b.ovf, nullptr, nullptr, b.counts, nullptr);
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_accumulate, dim3(ablocks), dim3(256), 0, s->stream, wp, b.radiance, d_film, f64);
    hipLaunchKernelGGL(k_advance_cursor, dim3(1), dim3(1), 0, s->stream, cursor, plan);
}

// The bounce loop of one batch, captured from s->stream into a graph.
static xt_status capture_graph(xt_scene* s, const xt_camera* cam, const xt_render_params* p, void* d_film, int f64,
                               RenderGraph* out) {
    const uint64_t npix = (uint64_t)p->width * p->height;
    const uint32_t cap = graph_capacity(npix, p->spp);
    xt_status st = ensure_buffers(s, cap);
    if (st != XT_OK) return st;
    RenderBuffers& rb = s->rb;
    const bool guide = (p->flags & XT_FLAG_GUIDE) != 0;
    if (guide && !rb.lobes) {
        drop_graphs(s);
        XT_HIP(dalloc(&rb.lobes, (size_t)G_LOBE_FIELDS * rb.capacity));
    }
    if (!s->d_cursor) XT_HIP(dalloc(&s->d_cursor, 1));
    for (int k = 0; k < 2; ++k) rb.q[k].rec = nullptr;
    uint32_t nbatches = 0;
    const BatchPlan plan = cursor_plan(p, cap, &nbatches);
    const WaveParams wp = cursor_wave_params(cam, p, s->d_cursor);
    BounceBufs b{{rb.q[0], rb.q[1]}, rb.sq, rb.hits, rb.radiance, rb.pending, rb.counts, s->d_ovf, rb.capacity};
    BounceGuide gd;
    if (guide) {
        gd.lobes_mk = (uint2*)rb.lobes;
        gd.lobes2 = (float2*)(rb.lobes + (size_t)2 * G_LOBES * rb.capacity);
        gd.stride = rb.capacity;
    }
    const uint32_t tblocks = trace_blocks(s, plan.pix_chunk * plan.S);
    XT_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    enqueue_cursor_batch(s, p, wp, b, plan, s->d_cursor, tblocks, gd, d_film, f64);
    hipError_t launch_err = hipGetLastError();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (launch_err != hipSuccess || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(XT_ERR_DEVICE, std::string("graph capture failed: ") +
                                       hipGetErrorString(launch_err != hipSuccess ? launch_err : e));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(XT_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(e));
    out->exec = exec;
    out->plan = plan;
    out->nbatches = nbatches;
    return XT_OK;
}

static xt_status find_graph(xt_scene* s, const xt_camera* cam, const xt_render_params* p, void* d_film, int f64,
                            RenderGraph** out) {
    GraphKey key;
    std::memset(&key, 0, sizeof key);
    key.cam = *cam;
    key.p = *p;
    key.p.spp_offset = 0;   // only the cursor's start depends on it
    key.film = d_film;
    key.f64 = f64;
    key.trace_grid = s->trace_grid;
    key.shade_grid = s->shade_grid;
    const uint64_t npix = (uint64_t)p->width * p->height;
    key.capacity = graph_capacity(npix, p->spp);
    for (RenderGraph& g : s->graphs)
        if (std::memcmp(&g.key, &key, sizeof key) == 0 && s->rb.capacity >= key.capacity) {
            *out = &g;
            return XT_OK;
        }
    if (s->graphs.size() >= 64) drop_graphs(s);
    RenderGraph g;
    g.key = key;
    xt_status st = capture_graph(s, cam, p, d_film, f64, &g);
    if (st != XT_OK) return st;
    s->graphs.push_back(g);
    *out = &s->graphs.back();
    return XT_OK;
}

// Enqueues one render on s->stream: reset the cursor, replay the graph once
// per batch.
static xt_status enqueue_graph_render(xt_scene* s, const RenderGraph& g, uint32_t spp_offset) {
    if (g.key.p.flags & XT_FLAG_GUIDE) {
        xt_status st = guide_wait_last_burst(s);
        if (st != XT_OK) return st;
    }
    BatchPlan plan = g.plan;
    plan.spp_offset = spp_offset;
    hipLaunchKernelGGL(k_cursor_init, dim3(1), dim3(1), 0, s->stream, s->d_cursor, plan);
    XT_HIP(hipGetLastError());
    for (uint32_t b = 0; b < g.nbatches; ++b) XT_HIP(hipGraphLaunch(g.exec, s->stream));
    return XT_OK;
}

// ---------------------------------------------------------------- unified batches

// Buffers of the unified wavefront, one set per device: the path state of a
// window (as a render's RenderBuffers, plus each path's item), the traversal
// spill columns, and the batch's item and scene tables.
struct MultiState {
    hipStream_t stream = nullptr;
    uint32_t capacity = 0;
    PathQueue q[2] = {};
    ShadowQueue sq = {};
    float4* hits = nullptr;
    float4* radiance = nullptr;
    float4* pending = nullptr;
    uint32_t* counts = nullptr;
    uint32_t* dyn = nullptr;           // traversal chunk cursors: extension [0..1], shadow [2..3]
    uint32_t* item_of = nullptr;
    uint32_t* ovf = nullptr;
    size_t ovf_entries = 0;
    MultiCursor* cursor = nullptr;
    MultiItem* items = nullptr;
    uint32_t items_cap = 0;
    DevScene* scenes = nullptr;
    uint32_t scenes_cap = 0;
    uint32_t* units = nullptr;
    std::vector<hipEvent_t> events;   // one per window, plus the start
    void release_paths() {
        for (int k = 0; k < 2; ++k) {
            hipFree(q[k].o); hipFree(q[k].d); hipFree(q[k].w); hipFree(q[k].e);
        }
        hipFree(sq.o); hipFree(sq.d); hipFree(sq.c);
        hipFree(hits); hipFree(radiance); hipFree(pending); hipFree(item_of);
        for (int k = 0; k < 2; ++k) q[k] = PathQueue{};
        sq = ShadowQueue{};
        hits = radiance = pending = nullptr;
        item_of = nullptr;
        capacity = 0;
    }
};

static void multi_release(xt_device* dev) {
    MultiState* m = dev->multi;
    if (!m) return;
    hipSetDevice(dev->ordinal);
    if (m->stream) hipStreamSynchronize(m->stream);
    m->release_paths();
    hipFree(m->counts); hipFree(m->ovf); hipFree(m->cursor); hipFree(m->items); hipFree(m->scenes); hipFree(m->units);
    hipFree(m->dyn);
    for (hipEvent_t e : m->events) hipEventDestroy(e);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    dev->multi = nullptr;
}

// Which items share a wavefront.  Unguided items of at most one graph
// replay (small renders, whose bounce loops are mostly launch and tail
// overhead) by default; XITILS_AMD_BATCH=unified puts every unguided item in
// it, =items none (measurement switches: the bench compares the forms).
enum class BatchMode { Hybrid, Unified, Items };
static BatchMode batch_mode() {
    const char* e = getenv("XITILS_AMD_BATCH");
    if (e && std::strcmp(e, "items") == 0) return BatchMode::Items;
    if (e && std::strcmp(e, "unified") == 0) return BatchMode::Unified;
    return BatchMode::Hybrid;
}

// One launched unified run: its graph, events and tables, finished after
// the batch's final synchronisation.
struct UnifiedRun {
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;     // start, then the end of each window
    std::vector<uint32_t> item_index;   // batch item of each table entry
    std::vector<uint64_t> last_window;  // window holding each table entry's last path
    std::vector<MultiItem> mi;          // host tables, kept until the copies are done
    std::vector<DevScene> ds;
    void release() {
        if (exec) hipGraphExecDestroy(exec);
        for (hipEvent_t e : events) hipEventDestroy(e);
        exec = nullptr;
        events.clear();
    }
};

// One wavefront for every item: the items' paths numbered one after another
// and walked in windows of up to the path capacity; each window is one
// replay of a captured graph (plan, generate, max_depth + 1 rounds of extend
// / shade / shadow over whichever items' paths are in it, accumulate,
// advance).  Each film equals xt_render's bit for bit: every path draws its
// own random numbers and each pixel's samples are summed in sample order.
static xt_status render_unified(xt_batch_item* items, const std::vector<uint32_t>& which, const std::vector<void*>& d_films,
                                UnifiedRun& run) {
    xt_device* dev = items[0].scene->dev;
    XT_HIP(hipSetDevice(dev->ordinal));
    if (!dev->multi) {
        dev->multi = new (std::nothrow) MultiState();
        if (!dev->multi) return fail(XT_ERR_OUT_OF_MEMORY, "host allocation failed");
        XT_HIP(hipStreamCreateWithFlags(&dev->multi->stream, hipStreamNonBlocking));
        XT_HIP(dalloc(&dev->multi->counts, 4));
        XT_HIP(dalloc(&dev->multi->cursor, 1));
        XT_HIP(dalloc(&dev->multi->dyn, 4));
        XT_HIP(hipMemset(dev->multi->dyn, 0, 4 * sizeof(uint32_t)));
    }
    MultiState& m = *dev->multi;
    // item and scene tables (items without samples are left out)
    std::vector<MultiItem>& mi = run.mi;
    std::vector<uint32_t>& item_index = run.item_index;
    std::vector<xt_scene*> sc_list;
    std::vector<DevScene>& ds = run.ds;
    uint64_t total = 0;
    int max_depth = 1;
    bool nmap = false;
    uint32_t ovf_depth = 0;
    for (uint32_t i : which) {
        const xt_batch_item& it = items[i];
        if (it.params.spp == 0) continue;
        xt_scene* s = it.scene;
        uint32_t si = 0;
        while (si < sc_list.size() && sc_list[si] != s) ++si;
        if (si == sc_list.size()) {
            sc_list.push_back(s);
            ds.push_back(s->ds);
            nmap |= s->nmap;
            ovf_depth = std::max(ovf_depth, s->ovf_depth);
        }
        MultiItem e{};
        e.wp = cursor_wave_params(&it.camera, &it.params, nullptr);
        e.nodes = s->ds.nodes;
        e.tris = s->ds.tris;
        e.scene = si;
        e.npix = it.params.width * it.params.height;
        e.spp_offset = it.params.spp_offset;
        e.film_f64 = it.film_is_f64 ? 1u : 0u;
        e.film = d_films[i];
        e.path_base = total;
        e.paths = (uint64_t)e.npix * it.params.spp;
        total += e.paths;
        max_depth = std::max(max_depth, it.params.max_depth);
        mi.push_back(e);
        item_index.push_back(i);
    }
    if (mi.empty()) return XT_OK;
    XT_HIP(hipStreamSynchronize(m.stream));   // an earlier batch's replays may still read the tables
    const uint32_t window = (uint32_t)std::min<uint64_t>(path_capacity(), total);
    const uint64_t nwin = (total + window - 1) / window;
    if (nwin >= (1ull << 32)) return fail(XT_ERR_TOO_LARGE, "batch of more than 2^32 windows");
    // buffers
    if (m.capacity < window) {
        hipStreamSynchronize(m.stream);
        m.release_paths();
        const size_t c = window;
        for (int k = 0; k < 2; ++k) {
            XT_HIP(dalloc(&m.q[k].o, c)); XT_HIP(dalloc(&m.q[k].d, c));
            XT_HIP(dalloc(&m.q[k].w, c)); XT_HIP(dalloc(&m.q[k].e, c));
        }
        XT_HIP(dalloc(&m.sq.o, c)); XT_HIP(dalloc(&m.sq.d, c)); XT_HIP(dalloc(&m.sq.c, c));
        XT_HIP(dalloc(&m.hits, c)); XT_HIP(dalloc(&m.radiance, c)); XT_HIP(dalloc(&m.pending, c));
        XT_HIP(dalloc(&m.item_of, c));
        m.capacity = window;
    }
    xt_scene* s0 = sc_list[0];
    const uint32_t tblocks = std::max(1u, std::min(s0->trace_grid, (window + TRACE_BLOCK - 1) / TRACE_BLOCK));
    const size_t ovf_need = (size_t)ovf_depth * tblocks * TRACE_BLOCK;
    if (m.ovf_entries < ovf_need) {
        hipStreamSynchronize(m.stream);
        hipFree(m.ovf);
        m.ovf = nullptr;
        XT_HIP(dalloc(&m.ovf, ovf_need));
        m.ovf_entries = ovf_need;
    }
    if (m.items_cap < mi.size()) {
        hipStreamSynchronize(m.stream);
        hipFree(m.items);
        hipFree(m.units);
        m.items = nullptr;
        m.units = nullptr;
        XT_HIP(dalloc(&m.items, mi.size()));
        XT_HIP(dalloc(&m.units, mi.size() + 1));
        m.items_cap = (uint32_t)mi.size();
    }
    if (m.scenes_cap < ds.size()) {
        hipStreamSynchronize(m.stream);
        hipFree(m.scenes);
        m.scenes = nullptr;
        XT_HIP(dalloc(&m.scenes, ds.size()));
        m.scenes_cap = (uint32_t)ds.size();
    }
    // the tables are uploaded on the batch's stream (ordered after an earlier
    // run's replays, which read the same buffers)
    XT_HIP(hipMemcpyAsync(m.items, mi.data(), mi.size() * sizeof(MultiItem), hipMemcpyHostToDevice, m.stream));
    XT_HIP(hipMemcpyAsync(m.scenes, ds.data(), ds.size() * sizeof(DevScene), hipMemcpyHostToDevice, m.stream));
    MultiBatch mb{};
    mb.items = m.items;
    mb.scenes = m.scenes;
    mb.num_items = (uint32_t)mi.size();
    mb.window = window;
    mb.total_paths = total;
    mb.cursor = m.cursor;
    mb.item_of = m.item_of;
    mb.units = m.units;
    const uint32_t ncu = (uint32_t)dev->num_cu;
    const uint32_t gblocks = std::max(1u, std::min((window + 255u) / 256u, ncu * 8));
    const uint32_t sblocks = std::max(1u, std::min((window + SHADE_BLOCK - 1) / SHADE_BLOCK, s0->shade_grid));
    // capture one window's bounce loop (windows of many paths hand the
    // traversal kernels' rays out in chunks from a cursor)
    const bool dyn = window >= TRAV_DYN_MIN;
    XT_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_window_plan, dim3(1), dim3(1), 0, m.stream, mb);
    hipLaunchKernelGGL(k_count_from_window, dim3(1), dim3(1), 0, m.stream, m.counts, m.cursor);
    hipLaunchKernelGGL(k_generate_multi, dim3(gblocks), dim3(256), 0, m.stream, mb, m.q[0], m.radiance);
    int cur = 0;
    for (int it = 0; it <= max_depth; ++it) {
        hipLaunchKernelGGL((dyn ? k_trace<false, TRACE_LDS_DEPTH, false, true, true>
                                : k_trace<false, TRACE_LDS_DEPTH, false, true, false>),
                           dim3(tblocks), dim3(TRACE_BLOCK), 0, m.stream, s0->ds, m.q[cur].o, m.q[cur].d, 1,
                           m.counts + cur, 0u, m.hits, nullptr, 0, 0, m.ovf, m.dyn, nullptr, m.counts + (1 - cur),
                           m.counts + 2, m.item_of, m.items);
        if (nmap)
            hipLaunchKernelGGL(k_shade_multi<true>, dim3(sblocks), dim3(SHADE_BLOCK), 0, m.stream, mb, m.q[cur],
                               m.counts + cur, m.hits, m.q[1 - cur], m.counts + (1 - cur), m.sq, m.counts + 2,
                               m.radiance, m.pending);
        else
            hipLaunchKernelGGL(k_shade_multi<false>, dim3(sblocks), dim3(SHADE_BLOCK), 0, m.stream, mb, m.q[cur],
                               m.counts + cur, m.hits, m.q[1 - cur], m.counts + (1 - cur), m.sq, m.counts + 2,
                               m.radiance, m.pending);
        hipLaunchKernelGGL((dyn ? k_shadow_multi<true> : k_shadow_multi<false>), dim3(tblocks), dim3(TRACE_BLOCK), 0,
                           m.stream, m.sq, m.counts + 2, m.radiance, m.pending, m.ovf, m.dyn + 2, m.item_of, m.items);
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_accumulate_multi, dim3(gblocks), dim3(256), 0, m.stream, mb, m.radiance);
    hipLaunchKernelGGL(k_window_advance, dim3(1), dim3(1), 0, m.stream, mb);
    hipError_t launch_err = hipGetLastError();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(m.stream, &graph);
    if (launch_err != hipSuccess || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(XT_ERR_DEVICE, std::string("unified batch capture failed: ") +
                                       hipGetErrorString(launch_err != hipSuccess ? launch_err : e));
    }
    e = hipGraphInstantiate(&run.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(XT_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(e));
    run.events.assign(nwin + 1, nullptr);
    for (hipEvent_t& ev : run.events) XT_HIP(hipEventCreate(&ev));
    hipEventRecord(run.events[0], m.stream);
    hipLaunchKernelGGL(k_window_init, dim3(1), dim3(1), 0, m.stream, mb);
    for (uint64_t w = 0; w < nwin; ++w) {
        e = hipGraphLaunch(run.exec, m.stream);
        if (e != hipSuccess) return fail(XT_ERR_DEVICE, std::string("graph launch failed: ") + hipGetErrorString(e));
        hipEventRecord(run.events[w + 1], m.stream);
    }
    for (size_t k = 0; k < mi.size(); ++k) run.last_window.push_back((mi[k].path_base + mi[k].paths - 1) / window);
    return XT_OK;
}

// After the batch's synchronisation: each unified item's latency, batch
// start to the end of the window holding its last path.
static void unified_finish(xt_batch_item* items, UnifiedRun& run) {
    for (size_t k = 0; k < run.item_index.size(); ++k) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, run.events[0], run.events[run.last_window[k] + 1]);
        items[run.item_index[k]].ms = ms;
    }
}

// Host threads that render a batch's lagged-loop items concurrently
// (XITILS_AMD_BATCH_THREADS, default 3; 1 renders them one after another).
static uint32_t batch_loop_threads() {
    const char* e = getenv("XITILS_AMD_BATCH_THREADS");
    const int v = e ? std::atoi(e) : 3;
    return (uint32_t)std::max(1, std::min(v, 16));
}

// Items with more paths than one graph replay holds render through
// xt_render's lagged bounce loop: per item on the 64-room batch it is 15-60%
// faster than replaying the fixed max_depth + 1 rounds per 4 Mi paths, while
// graphs win below one replay (no host in the loop, concurrent streams).
static bool batch_item_uses_loop(const xt_batch_item& it) {
    const uint64_t paths = (uint64_t)it.params.width * it.params.height * it.params.spp;
    return paths > graph_capacity((uint64_t)it.params.width * it.params.height, it.params.spp);
}

xt_status xt_render_batch(xt_batch_item* items, uint32_t n) {
    if (n && !items) return fail(XT_ERR_INVALID_ARGUMENT, "items is NULL");
    for (uint32_t i = 0; i < n; ++i) {
        xt_status st = check_render_args(items[i].scene, &items[i].camera, &items[i].params);
        if (st != XT_OK) return fail(st, "item " + std::to_string(i) + ": " + xt_last_error());
        if (!items[i].film) return fail(XT_ERR_INVALID_ARGUMENT, "item " + std::to_string(i) + ": film is NULL");
    }
    struct Pending {
        void* d_film = nullptr;
        bool owned = false;        // a device copy of a host film, copied back at the end
        size_t bytes = 0;
        hipEvent_t a = nullptr, b = nullptr;
    };
    std::vector<Pending> pend(n);
    std::vector<UnifiedRun> runs;
    xt_status result = XT_OK;
    auto cleanup = [&]() {
        for (Pending& q : pend) {
            if (q.owned) hipFree(q.d_film);
            if (q.a) hipEventDestroy(q.a);
            if (q.b) hipEventDestroy(q.b);
        }
        for (UnifiedRun& r : runs) r.release();
    };
    // films on the device: one device copy per distinct host film
    std::vector<void*> d_films(n, nullptr);
    for (uint32_t i = 0; i < n && result == XT_OK; ++i) {
        xt_batch_item& it = items[i];
        it.ms = 0.0f;
        if (hipSetDevice(it.scene->dev->ordinal) != hipSuccess) {
            result = fail(XT_ERR_DEVICE, "hipSetDevice failed");
            break;
        }
        Pending& q = pend[i];
        q.bytes = (size_t)it.params.width * it.params.height * 3 * (it.film_is_f64 ? 8 : 4);
        q.d_film = it.film;
        if (!it.film_on_device) {
            uint32_t j = 0;
            while (j < i && !(!items[j].film_on_device && items[j].film == it.film)) ++j;
            if (j < i) {
                q.d_film = d_films[j];
                d_films[i] = d_films[j];
                continue;
            }
            if (hipMalloc(&q.d_film, q.bytes) != hipSuccess) {
                result = fail(XT_ERR_OUT_OF_MEMORY, "film allocation failed");
                break;
            }
            q.owned = true;
            if (hipMemcpy(q.d_film, it.film, q.bytes, hipMemcpyHostToDevice) != hipSuccess)
                result = fail(XT_ERR_DEVICE, "film copy failed");
        }
        d_films[i] = q.d_film;
    }
    // which items share the wavefront; items that share a film with another
    // item all go through it (it keeps their order), when they can
    const BatchMode mode = batch_mode();
    std::vector<uint8_t> uni(n, 0);
    bool shared_film = false, any_guided = false, one_device = true;
    for (uint32_t i = 0; i < n; ++i) {
        any_guided |= (items[i].params.flags & XT_FLAG_GUIDE) != 0;
        one_device &= items[i].scene->dev == items[0].scene->dev;
        for (uint32_t j = 0; j < i; ++j) shared_film |= d_films[j] == d_films[i];
    }
    // a film shared by items that cannot all share one wavefront: every item
    // through xt_render, in item order
    const bool in_order = shared_film && (any_guided || !one_device);
    for (uint32_t i = 0; i < n && in_order && result == XT_OK; ++i) {
        xt_batch_item& it = items[i];
        if (it.params.spp == 0) continue;
        result = xt_render(it.scene, &it.camera, &it.params, d_films[i], it.film_is_f64, 1);
        if (result == XT_OK) it.ms = (float)it.scene->stats.ms_total;
    }
    // larger unguided items that share their scene with another such item
    // would render one after another (a scene renders on its one stream):
    // they share the wavefront's bounce rounds instead
    std::vector<uint32_t> same_scene(n, 0);
    for (uint32_t i = 0; i < n && !in_order; ++i) {
        const xt_batch_item& it = items[i];
        const bool can = one_device && !(it.params.flags & XT_FLAG_GUIDE) && it.params.spp > 0;
        if (!can || mode == BatchMode::Items) continue;
        const bool small = (uint64_t)it.params.width * it.params.height * it.params.spp <=
                           graph_capacity((uint64_t)it.params.width * it.params.height, it.params.spp);
        uni[i] = mode == BatchMode::Unified || small || (shared_film && !any_guided);
        if (!uni[i])
            for (uint32_t j = 0; j < n; ++j) same_scene[j] += items[j].scene == it.scene;
    }
    for (uint32_t i = 0; i < n && !in_order && mode == BatchMode::Hybrid; ++i) {
        const xt_batch_item& it = items[i];
        const bool can = one_device && !(it.params.flags & XT_FLAG_GUIDE) && it.params.spp > 0;
        if (can && !uni[i] && same_scene[i] >= 2) uni[i] = 1;
    }
    // per-item graphs first (their scenes' streams), so the unified runs and
    // the lagged-loop items below overlap them
    for (uint32_t i = 0; i < n && result == XT_OK && !in_order; ++i) {
        xt_batch_item& it = items[i];
        if (uni[i] || it.params.spp == 0 || batch_item_uses_loop(it)) continue;
        xt_scene* s = it.scene;
        hipSetDevice(s->dev->ordinal);
        RenderGraph* g = nullptr;
        result = find_graph(s, &it.camera, &it.params, d_films[i], it.film_is_f64, &g);
        if (result != XT_OK) break;
        Pending& q = pend[i];
        hipEventCreate(&q.a);
        hipEventCreate(&q.b);
        hipEventRecord(q.a, s->stream);
        result = enqueue_graph_render(s, *g, it.params.spp_offset);
        hipEventRecord(q.b, s->stream);
    }
    // the wavefront: in runs whose films are distinct (an item whose film an
    // earlier item of the run accumulates into starts the next run, so a
    // shared film receives the items' samples in item order)
    {
        std::vector<uint32_t> cur;
        auto flush = [&]() {
            if (cur.empty() || result != XT_OK) return;
            runs.emplace_back();
            result = render_unified(items, cur, d_films, runs.back());
            cur.clear();
        };
        for (uint32_t i = 0; i < n && result == XT_OK; ++i) {
            if (!uni[i]) continue;
            bool clash = false;
            for (uint32_t k : cur) clash |= d_films[k] == d_films[i];
            if (clash) flush();
            cur.push_back(i);
        }
        flush();
    }
    // items larger than one graph replay: the lagged bounce loop of xt_render,
    // while the graphs and the wavefront run.  Items on different scenes
    // render concurrently from a few host threads (a scene's items in item
    // order, on one thread): a render's last bounces keep few rays in flight,
    // and another scene's render fills the GPU meanwhile.
    if (result == XT_OK && !in_order) {
        std::vector<std::vector<uint32_t>> groups;   // loop items per scene, in item order
        for (uint32_t i = 0; i < n; ++i) {
            const xt_batch_item& it = items[i];
            if (uni[i] || it.params.spp == 0 || !batch_item_uses_loop(it)) continue;
            size_t gi = 0;
            while (gi < groups.size() && items[groups[gi][0]].scene != it.scene) ++gi;
            if (gi == groups.size()) groups.emplace_back();
            groups[gi].push_back(i);
        }
        const uint32_t nthreads = std::min<uint32_t>(batch_loop_threads(), (uint32_t)groups.size());
        std::vector<xt_status> tst(std::max(1u, nthreads), XT_OK);
        std::vector<std::string> terr(tst.size());
        std::atomic<size_t> next_group{0};
        auto work = [&](uint32_t t) {
            for (size_t gi = next_group++; gi < groups.size(); gi = next_group++)
                for (uint32_t i : groups[gi]) {
                    xt_batch_item& it = items[i];
                    const xt_status st = xt_render(it.scene, &it.camera, &it.params, d_films[i], it.film_is_f64, 1);
                    if (st != XT_OK) {
                        tst[t] = st;
                        terr[t] = xt_last_error();
                        next_group = groups.size();   // the other threads stop after their current item
                        return;
                    }
                    it.ms = (float)it.scene->stats.ms_total;
                }
        };
        if (!groups.empty()) {
            std::vector<std::thread> pool;
            for (uint32_t t = 1; t < nthreads; ++t) {
                try {
                    pool.emplace_back(work, t);
                } catch (...) {   // no thread: the groups it would have taken go to the others
                    break;
                }
            }
            work(0);
            for (std::thread& th : pool) th.join();
        }
        for (size_t t = 0; t < tst.size() && result == XT_OK; ++t)
            if (tst[t] != XT_OK) result = fail(tst[t], terr[t]);
    }
    for (uint32_t i = 0; i < n; ++i) {
        hipError_t e = hipStreamSynchronize(items[i].scene->stream);
        if (e != hipSuccess && result == XT_OK) result = fail(XT_ERR_DEVICE, std::string("batch render failed: ") + hipGetErrorString(e));
    }
    if (!runs.empty()) {
        hipError_t e = hipStreamSynchronize(items[0].scene->dev->multi->stream);
        if (e != hipSuccess && result == XT_OK) result = fail(XT_ERR_DEVICE, std::string("batch render failed: ") + hipGetErrorString(e));
    }
    if (result == XT_OK) {
        for (UnifiedRun& r : runs) unified_finish(items, r);
        for (uint32_t i = 0; i < n; ++i)
            if (pend[i].a) hipEventElapsedTime(&items[i].ms, pend[i].a, pend[i].b);
        for (uint32_t i = 0; i < n && result == XT_OK; ++i) {
            if (!pend[i].owned) continue;
            if (hipMemcpy(items[i].film, pend[i].d_film, pend[i].bytes, hipMemcpyDeviceToHost) != hipSuccess)
                result = fail(XT_ERR_DEVICE, "film copy failed");
        }
    }
    cleanup();
    return result;
}


// ---------------------------------------------------------------- render without allocation

// xt_render_static's carving of the static scratch: a 256-byte header (batch
// cursor at 0, queue counters at 64), at most 16 MiB of traversal spill
// columns, then 14 float4 arrays of `capacity` paths (224 bytes per path).
constexpr size_t STATIC_HEADER = 256;
constexpr size_t STATIC_SPILL_MAX = (size_t)16 << 20;
constexpr size_t STATIC_PATH_BYTES = 14 * sizeof(float4);

xt_status xt_render_static(xt_scene* s, const xt_camera* cam, const xt_render_params* p, void* film, int film_f64) {
    xt_status st = check_render_args(s, cam, p);
    if (st != XT_OK) return st;
    if (!film) return fail(XT_ERR_INVALID_ARGUMENT, "film is NULL");
    if (p->flags & XT_FLAG_GUIDE) return fail(XT_ERR_INVALID_ARGUMENT, "xt_render_static renders unguided (no XT_FLAG_GUIDE)");
    unsigned char* base = s->dev->scratch;
    if (!base) return fail(XT_ERR_DEVICE, "device has no static scratch address");
    XT_HIP(hipSetDevice(s->dev->ordinal));
    if (p->spp == 0) {
        s->stats = xt_render_stats{};
        return XT_OK;
    }
    // traversal spill columns: ovf_depth entries per lane of every block launched
    const size_t per_block = (size_t)s->ovf_depth * TRACE_BLOCK * sizeof(uint32_t);
    uint32_t tcap = s->trace_grid;
    if (per_block > 0) tcap = (uint32_t)std::min<size_t>(tcap, STATIC_SPILL_MAX / per_block);
    if (tcap == 0)
        return fail(XT_ERR_TOO_LARGE, "the scene's traversal stack (" + std::to_string(s->ovf_depth) +
                                          " spill entries per lane) exceeds the static scratch");
    const size_t spill_bytes = ((per_block * tcap + 255) / 256) * 256;
    const uint64_t npix = (uint64_t)p->width * p->height;
    const uint64_t fit = (XT_STATIC_SCRATCH_BYTES - STATIC_HEADER - spill_bytes) / STATIC_PATH_BYTES;
    const uint32_t cap = (uint32_t)std::min<uint64_t>(fit, npix * p->spp);
    BatchCursor* cursor = (BatchCursor*)base;
    BounceBufs b{};
    b.counts = (uint32_t*)(base + 64);
    b.ovf = per_block ? (uint32_t*)(base + STATIC_HEADER) : nullptr;
    b.capacity = cap;
    float4* arr = (float4*)(base + STATIC_HEADER + spill_bytes);
    auto take = [&]() {
        float4* r = arr;
        arr += cap;
        return r;
    };
    for (int k = 0; k < 2; ++k) {
        b.q[k].o = take();
        b.q[k].d = take();
        b.q[k].w = take();
        b.q[k].e = take();
        b.q[k].rec = nullptr;
    }
    b.sq.o = take();
    b.sq.d = take();
    b.sq.c = take();
    b.hits = take();
    b.radiance = take();
    b.pending = take();
    uint32_t nbatches = 0;
    const BatchPlan plan = cursor_plan(p, cap, &nbatches);
    const WaveParams wp = cursor_wave_params(cam, p, cursor);
    const uint32_t tblocks = std::min(trace_blocks(s, (size_t)plan.pix_chunk * plan.S), tcap);
    // the cursor is written by k_cursor_init and the counters by the kernels
    // (k_count_from_cursor, k_trace), so nothing needs clearing beforehand
    hipLaunchKernelGGL(k_cursor_init, dim3(1), dim3(1), 0, s->stream, cursor, plan);
    for (uint32_t k = 0; k < nbatches; ++k)
        enqueue_cursor_batch(s, p, wp, b, plan, cursor, tblocks, BounceGuide{}, film, film_f64);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s->stream);
    if (le != hipSuccess) return fail(XT_ERR_DEVICE, std::string("xt_render_static launch failed: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(XT_ERR_DEVICE, std::string("xt_render_static failed: ") + hipGetErrorString(se));
    s->stats = xt_render_stats{};
    s->stats.camera_rays = npix * p->spp;
    return XT_OK;
}


// Times the graph-replayed render of one batch for a few launch shapes and
// keeps the fastest (traversal blocks per CU x shade blocks per CU).
xt_status xt_scene_autotune(xt_scene* s, const xt_camera* cam, const xt_render_params* p, uint32_t out[2]) {
    xt_render_params q = *p;
    xt_status st = check_render_args(s, cam, &q);
    if (st != XT_OK) return st;
    XT_HIP(hipSetDevice(s->dev->ordinal));
    const uint64_t npix = (uint64_t)q.width * q.height;
    const uint32_t cap = graph_capacity(npix, std::max(1u, q.spp));
    q.spp = std::max(1u, (uint32_t)std::min<uint64_t>(std::max<uint64_t>(cap / npix, 1), std::max(1u, p->spp)));
    void* d_film = nullptr;
    XT_HIP(hipMalloc(&d_film, npix * 3 * sizeof(float)));
    const uint32_t ncu = (uint32_t)s->dev->num_cu;
    const uint32_t tcand[3] = {2, 4, 8}, scand[3] = {2, 4, 8};
    float best = INFINITY;
    uint32_t bt = s->trace_grid, bs = s->shade_grid;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    for (uint32_t ti : tcand)
        for (uint32_t si : scand) {
            s->trace_grid = std::min(ti * ncu, s->trace_grid_max);
            s->shade_grid = si * ncu;
            RenderGraph* g = nullptr;
            st = find_graph(s, cam, &q, d_film, 0, &g);
            if (st != XT_OK) break;
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && st == XT_OK; ++rep) {   // the first run warms the caches
                hipEventRecord(a, s->stream);
                st = enqueue_graph_render(s, *g, 0);
                hipEventRecord(b, s->stream);
                if (hipStreamSynchronize(s->stream) != hipSuccess) st = fail(XT_ERR_DEVICE, "autotune render failed");
                hipEventElapsedTime(&ms, a, b);
            }
            if (st != XT_OK) break;
            if (ms < best) {
                best = ms;
                bt = s->trace_grid;
                bs = s->shade_grid;
            }
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    hipFree(d_film);
    s->trace_grid = bt;
    s->shade_grid = bs;
    drop_graphs(s);
    if (st != XT_OK) return st;
    if (out) {
        out[0] = bt / ncu;
        out[1] = bs / ncu;
    }
    return XT_OK;
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

void xt_guide_config_default(xt_guide_config* c) {
    if (!c) return;
    c->guide_prob = 0.5f;
    c->learning_rate = 5e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->adam_eps = 1e-8f;
    c->minibatch = 32768;
    c->max_steps_per_batch = 256;
    c->weight_clip = 20.0f;
    c->seed = 1;
    c->max_records = 1u << 23;
    c->train_target = 1;
    c->epochs = 3;
    c->max_guided_vertex = 0;
}

xt_status xt_guide_init(xt_scene* s, const xt_guide_config* cfg) {
    if (!s) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    xt_guide_config c;
    xt_guide_config_default(&c);
    if (cfg) c = *cfg;
    if (!(c.guide_prob >= 0.0f && c.guide_prob < 1.0f))
        return fail(XT_ERR_INVALID_ARGUMENT, "guide_prob must be in [0, 1)");
    if (c.train_target > 1) return xt_fail(XT_ERR_INVALID_ARGUMENT, "train_target must be 0 or 1");
    if (c.epochs == 0) c.epochs = 1;
    if (!(c.learning_rate > 0.0f) || !(c.beta1 >= 0.0f && c.beta1 < 1.0f) || !(c.beta2 >= 0.0f && c.beta2 < 1.0f))
        return fail(XT_ERR_INVALID_ARGUMENT, "bad Adam settings");
    XT_HIP(hipSetDevice(s->dev->ordinal));
    drop_graphs(s);   // captured guided renders hold the old network's addresses
    GuideState& g = s->guide;
    g.release();
    g.cfg = c;
    for (int a = 0; a < 3; ++a) {
        float lo = s->bounds_lo[a], hi = s->bounds_hi[a];
        if (!(hi > lo)) {   // empty or flat scene
            lo = std::isfinite(lo) ? lo - 0.5f : -0.5f;
            hi = lo + 1.0f;
        }
        g.lo[a] = lo;
        g.hi[a] = hi;
    }
    g.gb.lo = mk(g.lo[0], g.lo[1], g.lo[2]);
    g.gb.inv_ext = mk(1.0f / (g.hi[0] - g.lo[0]), 1.0f / (g.hi[1] - g.lo[1]), 1.0f / (g.hi[2] - g.lo[2]));
    g.partial_rows = (uint32_t)s->dev->num_cu;
    xt_status st = XT_OK;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && st == XT_OK)
            st = fail(e == hipErrorOutOfMemory ? XT_ERR_OUT_OF_MEMORY : XT_ERR_DEVICE,
                      std::string("guide allocation failed: ") + hipGetErrorString(e));
    };
    chk(dalloc(&g.P, G_NPARAMS));
    chk(dalloc(&g.m, G_NPARAMS));
    chk(dalloc(&g.v, G_NPARAMS));
    chk(dalloc(&g.partial, (size_t)g.partial_rows * G_NPARAMS));
    chk(dalloc(&g.d_count, 1));
    chk(dalloc(&g.d_wsum, 1));
    chk(hipHostMalloc((void**)&g.host_io, 2 * sizeof(float), hipHostMallocDefault));
    // the forward and backward copies and the two published forward copies
    // in one allocation, 256-byte aligned pieces
    const size_t sizes[4] = {sizeof(GuideChain), sizeof(GuideBack), sizeof(GuideChain), sizeof(GuideChain)};
    size_t off[4], total = 0;
    for (int k = 0; k < 4; ++k) {
        off[k] = total;
        total += (sizes[k] + 255) / 256 * 256;
    }
    chk(hipMalloc(&g.net_mem, total));
    chk(hipStreamCreateWithFlags(&g.train_stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        chk(hipEventCreateWithFlags(&g.pub_ev[k], hipEventDisableTiming));
        chk(hipEventCreate(&g.t0[k]));
        chk(hipEventCreate(&g.t1[k]));
    }
    if (st != XT_OK) {
        g.release();
        return st;
    }
    char* b = (char*)g.net_mem;
    g.net.chain = (GuideChain*)(b + off[0]);
    g.net.back = (GuideBack*)(b + off[1]);
    g.pub[0] = (GuideChain*)(b + off[2]);
    g.pub[1] = (GuideChain*)(b + off[3]);
    g.ready = true;
    std::vector<float> P;
    guide_init_params(c.seed, P);
    return xt_guide_set_params(s, P.data());
}

xt_status xt_guide_get_params(xt_scene* s, float* params) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (!params) return fail(XT_ERR_INVALID_ARGUMENT, "params is NULL");
    XT_HIP(hipStreamSynchronize(s->stream));
    XT_HIP(hipMemcpy(params, s->guide.P, G_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
    return XT_OK;
}

xt_status xt_guide_set_params(xt_scene* s, const float* params) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (!params) return fail(XT_ERR_INVALID_ARGUMENT, "params is NULL");
    GuideState& g = s->guide;
    XT_HIP(hipStreamSynchronize(s->stream));
    XT_HIP(hipMemcpy(g.P, params, G_NPARAMS * sizeof(float), hipMemcpyHostToDevice));
    XT_HIP(hipMemset(g.m, 0, G_NPARAMS * sizeof(float)));
    XT_HIP(hipMemset(g.v, 0, G_NPARAMS * sizeof(float)));
    g.step = 0;
    st = guide_prepare(s);
    if (st != XT_OK) return st;
    XT_HIP(hipStreamSynchronize(s->stream));
    return XT_OK;
}

xt_status xt_guide_get_bounds(xt_scene* s, float lo[3], float hi[3]) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (!lo || !hi) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int a = 0; a < 3; ++a) {
        lo[a] = s->guide.lo[a];
        hi[a] = s->guide.hi[a];
    }
    return XT_OK;
}

xt_status xt_guide_forward(xt_scene* s, const float* queries, uint32_t n, float* outputs) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (n == 0) return XT_OK;
    if (!queries || !outputs) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    float *dq = nullptr, *dout = nullptr;
    XT_HIP(dalloc(&dq, (size_t)n * 9));
    hipError_t e = dalloc(&dout, (size_t)n * G_OUT);
    if (e == hipSuccess) e = hipMemcpyAsync(dq, queries, (size_t)n * 9 * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        const uint32_t grid = std::max(1u, std::min((n + GI_BLOCK - 1) / GI_BLOCK, (uint32_t)s->dev->num_cu * 4));
        hipLaunchKernelGGL(k_guide_forward, dim3(grid), dim3(GI_BLOCK), 0, s->stream, s->guide.net, s->guide.gb, dq,
                           n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(outputs, dout, (size_t)n * G_OUT * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(dq);
    hipFree(dout);
    if (e != hipSuccess) return fail(XT_ERR_DEVICE, std::string("xt_guide_forward: ") + hipGetErrorString(e));
    return XT_OK;
}

xt_status xt_guide_gradient(xt_scene* s, const float* queries, const float* dirs, const float* weights, uint32_t n,
                            float weight_scale, float weight_clip, float* grad) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (!grad || (n && (!queries || !dirs || !weights))) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    GuideSample* ds = nullptr;
    if (n) {
        st = guide_upload_samples(s, queries, dirs, weights, n, &ds);
        if (st != XT_OK) return st;
    }
    GuideState& g = s->guide;
    uint32_t grid = 0;
    st = guide_gradient_rows(s, s->stream, ds, n, weight_scale, weight_clip, &grid);
    if (st != XT_OK) return st;
    std::vector<float> rows((size_t)grid * G_NPARAMS);
    XT_HIP(hipMemcpyAsync(rows.data(), g.partial, rows.size() * 4, hipMemcpyDeviceToHost, s->stream));
    XT_HIP(hipStreamSynchronize(s->stream));
    for (int i = 0; i < G_NPARAMS; ++i) {
        float acc = 0.0f;
        for (uint32_t b = 0; b < grid; ++b) acc += rows[(size_t)b * G_NPARAMS + i];
        grad[i] = acc;
    }
    return XT_OK;
}

xt_status xt_guide_train_step(xt_scene* s, const float* queries, const float* dirs, const float* weights, uint32_t n) {
    xt_status st = guide_need(s);
    if (st != XT_OK) return st;
    if (n == 0) return XT_OK;
    if (!queries || !dirs || !weights) return fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    double wsum = 0.0;
    for (uint32_t i = 0; i < n; ++i) wsum += weights[i];
    if (!(wsum > 0.0) || !std::isfinite(wsum)) return fail(XT_ERR_INVALID_ARGUMENT, "weights must have a positive sum");
    GuideSample* ds = nullptr;
    st = guide_upload_samples(s, queries, dirs, weights, n, &ds);
    if (st != XT_OK) return st;
    const uint32_t saved = s->guide.cfg.minibatch;
    s->guide.cfg.minibatch = std::max(saved, n);   // one step over all n samples
    st = guide_train(s, s->stream, ds, n, (float)(n / wsum), 1, nullptr);
    s->guide.cfg.minibatch = saved;
    if (st != XT_OK) return st;
    st = guide_publish_all(s);
    if (st != XT_OK) return st;
    XT_HIP(hipStreamSynchronize(s->stream));
    return XT_OK;
}

}  // extern "C"
