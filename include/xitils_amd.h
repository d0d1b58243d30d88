/*
 * xitils_amd — MI355X (gfx950) path-tracing hot path behind a C ABI.
 *
 * The reference (mizuooon/Xitils) renders with
 *     RenderTarget<T>::render(scene, spp, [&](pFilm, sampler, color) {
 *         color += pathTracer->eval(scene, sampler, camera->generateRay(pFilm, sampler)).color; })
 * (RenderTarget.h:110-127, PathTracer.h:18-21, Sandbox/SimplePathTracer/Main.cpp:134-138).
 * A per-sample host std::function cannot cross to the GPU, so the unit this
 * library replaces is that whole composite: xt_render() below.  The scene
 * side replaces Scene::addObject / buildAccelerationStructure (Scene.h:17-28)
 * and the two ray queries Scene::intersect / intersectAny (Scene.h:30-36).
 *
 * Conventions: plain C types only; every call returns xt_status and never
 * throws; on failure xt_last_error() (thread-local) says why.  Host arrays
 * passed in are copied; the library owns all device memory behind opaque
 * handles.  A handle is used by one host thread at a time.
 */
#ifndef XITILS_AMD_H
#define XITILS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum xt_status {
    XT_OK = 0,
    XT_ERR_INVALID_ARGUMENT = 1,
    XT_ERR_UNSUPPORTED_MATERIAL = 2,
    XT_ERR_OUT_OF_MEMORY = 3,
    XT_ERR_DEVICE = 4,          /* a HIP runtime call failed */
    XT_ERR_NO_DEVICE = 5,       /* no gfx950 device at that ordinal */
    XT_ERR_TOO_LARGE = 6        /* scene or image beyond the supported limits */
} xt_status;

typedef struct xt_device xt_device;
typedef struct xt_scene xt_scene;

/* Material kinds (Material.h).  Parameters, 7 floats:
 *   XT_MAT_DIFFUSE       albedo r, g, b                         Diffuse            Material.h:50-86
 *   XT_MAT_SPEC_REFL     -                                      SpecularReflection Material.h:177-189
 *   XT_MAT_SPEC_REFR     index                                  SpecularRefraction Material.h:191-218
 *   XT_MAT_SPEC_FRESNEL  index                                  SpecularFresnel    Material.h:220-273
 *   XT_MAT_METAL         alpha (= roughness^2), f0 r, g, b      Metal (GGX)        Material.h:275-384
 *   XT_MAT_EMISSION      power r, g, b                          Emission           Material.h:386-411
 *   XT_MAT_GGX_LAMBERT   albedo r, g, b, alpha, f0, p_spec      GGX specular over Lambert, one-sample
 *                                                               lobe mixture as GlossyWithHighLight
 *                                                               (Material.h:134-175); no reference twin
 *   XT_MAT_MICROFACET_NMAP  base type, base p[0..5]             MicrofacetNormalMapping (Schuessler 2017,
 *                                                               Sandbox/MicrofacetBasedNormalMapping/
 *                                                               Main.cpp:86-340, SameMaterialExplicit)
 *                                                               over a DIFFUSE, METAL or GGX_LAMBERT base
 *                                                               whose parameters follow the type */
enum {
    XT_MAT_DIFFUSE = 0,
    XT_MAT_SPEC_REFL = 1,
    XT_MAT_SPEC_REFR = 2,
    XT_MAT_SPEC_FRESNEL = 3,
    XT_MAT_METAL = 4,
    XT_MAT_EMISSION = 5,
    XT_MAT_GGX_LAMBERT = 6,
    XT_MAT_MICROFACET_NMAP = 7,
    XT_MAT_COUNT_ = 8
};

typedef struct xt_material {   /* 32 bytes */
    uint32_t type;
    float p[7];
} xt_material;

/* One Object (Object.h:16-40): a triangle mesh (TriangleMesh::setGeometry
 * raw-array form, TriangleMesh.h:151-161), its objectToWorld Transform with
 * the explicit inverse the reference keeps (Transform.h:13-30), and one
 * material; tri_materials optionally overrides it per triangle. */
typedef struct xt_object_desc {
    const float* positions;        /* num_vertices * 3, object space */
    const float* normals;          /* num_vertices * 3, or NULL (face normals) */
    const int32_t* indices;        /* num_triangles * 3 */
    uint32_t num_vertices;
    uint32_t num_triangles;
    const float* object_to_world;  /* 16 floats row-major, or NULL = identity */
    const float* world_to_object;  /* 16 floats row-major (Transform::matInv), or NULL */
    uint32_t material;             /* index into xt_scene_desc.materials */
    const uint32_t* tri_materials; /* num_triangles material indices, or NULL */
    /* TriangleMesh texCoords / tangents / bitangents (TriangleMesh.h:151-161),
     * read where a material has a normal map: */
    const float* texcoords;        /* num_vertices * 2, or NULL (uv = 0) */
    const float* tangents;         /* num_vertices * 3, or NULL (no tangent frame) */
    const float* bitangents;       /* num_vertices * 3; required with tangents */
} xt_object_desc;

/* An RGB float texture (Texture.h): rgb(uv) is the bilinear lookup with
 * clamped addressing (Texture.h:83-103, 161-171); texel (x, y) at
 * rgb[3 * (y * width + x)], row 0 first.  Copied at scene creation. */
typedef struct xt_texture {
    uint32_t width, height;
    const float* rgb;
} xt_texture;

enum {
    XT_ENV_MAP_NONE = 0,
    XT_ENV_MAP_ANGULAR = 1,      /* SkySphereFromImage: Debevec angular map (SkySphere.h:14-34) */
    XT_ENV_MAP_LATLONG = 2       /* latitude-longitude map, importance-sampled under XT_FLAG_ENV_NEE */
};

/* SkySphereUniform (SkySphere.h:36-46) or SkySphereFromImage (SkySphere.h:
 * 14-34), plus an optional sun disk.  Without a sky the scene has
 * Scene::skySphere == nullptr (escapes add nothing).  An image sky replaces
 * the constant colour; it is looked up with Texture::rgb's bilinear filter
 * (Texture.h:83-103).  Lat-long maps: u = (atan2(d.z, d.x) + pi) / 2pi,
 * v = acos(d.y) / pi (y up, row 0 at the zenith), wrapping in u. */
typedef struct xt_env {
    int32_t has_sky;
    float sky[3];
    float sun_dir[3];       /* unit vector toward the sun */
    float sun_one_minus_cos;/* 1 - cos(half angle); 0 = no sun */
    float sun_sin2;         /* sin^2(half angle) */
    float sun_radiance[3];
    uint32_t map_kind;      /* XT_ENV_MAP_* */
    uint32_t map_width, map_height;
    const float* map_rgb;   /* map_width * map_height * 3 floats, row 0 first (copied) */
} xt_env;

typedef struct xt_scene_desc {
    const xt_object_desc* objects;
    uint32_t num_objects;
    const xt_material* materials;
    uint32_t num_materials;
    xt_env env;
    const xt_texture* textures;
    uint32_t num_textures;
    /* Material::normalmap per material: 1 + index into textures, 0 = none;
     * NULL = no normal maps.  Applied at each hit as Object.h:59-68:
     * n = rgb(texCoord) * 2 - 1, shading.n = normalize(n.b * shading.n +
     * n.r * tangent + n.g * bitangent). */
    const uint32_t* material_normal_maps;
} xt_scene_desc;

/* PinholeCamera (Camera.h:17-55): cam_to_world is Transform::mat (row-major),
 * tan_x = tanf(fov / 2), tan_y = tanf(fov / 2 * aspect) with aspect = H / W. */
typedef struct xt_camera {
    float cam_to_world[16];
    float tan_x, tan_y;
} xt_camera;

enum {
    XT_INTEGRATOR_STANDARD = 0,  /* StandardPathTracer: NEE + BSDF MIS (PathTracer.h:112-258) */
    XT_INTEGRATOR_NAIVE = 1      /* NaivePathTracer: BSDF sampling only (PathTracer.h:47-110) */
};
enum {
    XT_FLAG_UNIFORM_BARY = 1u,   /* area-uniform light points (the reference's are not, TriangleIndexed.h:184-192) */
    XT_FLAG_ENV_NEE = 2u,        /* sample sun + sky as a light, MIS with BSDF escapes */
    XT_FLAG_GUIDE = 4u,          /* standard integrator: sample with the scene's guiding network */
    XT_FLAG_GUIDE_TRAIN = 8u,    /* with XT_FLAG_GUIDE: record training samples, take Adam steps per batch */
    XT_FLAG_SORT_MATERIALS = 16u /* shade each block's queue entries grouped by material type (same film) */
};

typedef struct xt_render_params {
    uint32_t width, height;
    uint32_t spp;            /* samples per pixel in this call */
    uint32_t spp_offset;     /* index of the first sample (progressive rendering) */
    int32_t max_depth;       /* bounce cap; <= 0: unbounded, roulette only (as the reference) */
    uint32_t seed;           /* Philox key word 0 */
    uint32_t stream;         /* Philox key word 1 */
    uint32_t integrator;     /* XT_INTEGRATOR_* */
    uint32_t flags;          /* XT_FLAG_* */
    int32_t rr_start;        /* roulette when pathLength > rr_start (5, PathTracer.h:255) */
    float rr_prob;           /* 0.9 (PathTracer.h:256) */
    float ray_eps;           /* 1e-5 origin offset along d (PathTracer.h:254) */
    float shadow_margin;     /* 1e-4 (PathTracer.h:257) */
} xt_render_params;

/* Counters of the last xt_render, xt_trace_closest or xt_trace_any on a
 * scene.  After a ray query only its fields are set: extension_rays (closest)
 * or shadow_rays (any) = n, ms_extend or ms_shadow = ms_total = the traversal
 * kernel's time, and with counting on its node and triangle fetches. */
typedef struct xt_render_stats {
    uint64_t camera_rays, extension_rays, shadow_rays;  /* extension_rays includes camera rays */
    double ms_total;         /* wall time of the call's GPU work (HIP events) */
    double ms_extend, ms_shade, ms_shadow, ms_other;    /* per-stage kernel time (HIP events) */
    uint32_t extend_launches;
    /* BVH node (bytes_nodes / bvh_nodes of xt_scene_info: 64 B with the
     * default quantised 4-wide nodes) and triangle (48 B) fetches of the
     * traversal kernels; filled only after xt_scene_set_counting(scene, 1). */
    uint64_t ext_node_fetches, ext_tri_fetches, shadow_node_fetches, shadow_tri_fetches;
    /* guiding: network inference kernel time; training time = the durations
     * of the training bursts (HIP events on the guide's stream) that ended
     * since the previous render (a training render's own burst runs after it
     * returns and is counted by a later render; ms_total does not include
     * it); Adam steps taken, training samples kept, queue entries run
     * through the network and samples run through the gradient kernels (all
     * epochs) by this render's bursts */
    double ms_guide_infer, ms_guide_train;
    uint64_t guide_steps, guide_samples, guide_queries, guide_grad_samples;
} xt_render_stats;

typedef struct xt_ray {        /* 32 bytes */
    float ox, oy, oz, tmax;
    float dx, dy, dz, pad;
} xt_ray;

typedef struct xt_hit {        /* 16 bytes; miss: t = +inf, prim = 0xFFFFFFFF */
    float t;
    uint32_t prim;             /* global primitive id: object order, then triangle order */
    float b0, b1;              /* barycentrics of vertices 0 and 1 (TriangleIndexed.h:127-131) */
} xt_hit;

/* Reads the .so's version string. */
const char* xt_version(void);
/* Thread-local description of the last error on this thread. */
const char* xt_last_error(void);

/* Opens HIP device `ordinal`; fails with XT_ERR_NO_DEVICE unless it is gfx950. */
xt_status xt_device_open(int ordinal, xt_device** out);
xt_status xt_device_close(xt_device* dev);
/* No reference counterpart: a measurement helper for the kernels' issue-rate
 * figures.  Times a kernel of 16 independent fp32 FMA chains per lane (no
 * memory traffic, 8 waves per SIMD on every CU) and returns the wave-level
 * VALU instructions it issued per second. */
xt_status xt_device_valu_rate(xt_device* dev, double* wave_instructions_per_s);

/* Scene::addObject x N + Scene::buildAccelerationStructure (Scene.h:17-28):
 * flattens instances to world space, builds the BVH on the host and uploads
 * everything to HBM.  Rejects (XT_ERR_INVALID_ARGUMENT) a referenced vertex
 * whose world-space position is not finite or exceeds 1e30 in magnitude. */
xt_status xt_scene_create(xt_device* dev, const xt_scene_desc* desc, xt_scene** out);
xt_status xt_scene_destroy(xt_scene* scene);

/* Host-side facts about the built scene: num_prims, bvh_nodes, bvh_depth,
 * bytes_nodes, bytes_tris, build_ms (build_ms as double bits). */
xt_status xt_scene_info(const xt_scene* scene, uint64_t out[6]);

/* Scene::intersect (Scene.h:30-32): closest hit for each ray.  rays/hits are
 * host pointers, or device pointers when on_device != 0. */
xt_status xt_trace_closest(xt_scene* scene, const xt_ray* rays, uint32_t n, xt_hit* hits, int on_device);
/* Scene::intersectAny (Scene.h:34-36): occluded[i] = 1 when any triangle lies
 * in (0, tmax]. */
xt_status xt_trace_any(xt_scene* scene, const xt_ray* rays, uint32_t n, uint8_t* occluded, int on_device);

/* RenderTarget::render + PinholeCamera::generateRay + PathTracer::eval, for
 * params->spp samples of every pixel: film (width*height*3, row-major, y
 * down) += the sum of the path radiances, per pixel in sample order.  film is
 * float32 or, with film_is_f64 != 0, float64 (the converged-oracle pass);
 * host or device memory per film_on_device. */
xt_status xt_render(xt_scene* scene, const xt_camera* camera, const xt_render_params* params, void* film,
                    int film_is_f64, int film_on_device);
xt_status xt_render_get_stats(const xt_scene* scene, xt_render_stats* out);
/* One render of a batch: the arguments of one xt_render call. */
typedef struct xt_batch_item {
    xt_scene* scene;
    xt_camera camera;
    xt_render_params params;
    void* film;
    int32_t film_is_f64;
    int32_t film_on_device;
    float ms;                /* out: GPU time of this item (HIP events; see xt_render_batch) */
} xt_batch_item;

/* Many RenderTarget::render calls at once (no reference counterpart; the
 * closest is the tile scheduler of RenderTarget.h:84-127).  Each item gives
 * the film xt_render would give, bit for bit.  Unguided items of at most 4 Mi
 * paths, and larger unguided items that share their scene with another such
 * item, share one wavefront: their paths are numbered one after another and
 * walked in windows of the path capacity, each window one replay of a
 * captured graph that traces and shades every path in its own item's scene.
 * Other items of at most 4 Mi paths run their bounce loop as a per-scene HIP
 * graph with no host round trip, concurrently on their scenes' streams;
 * larger items render through xt_render's bounce loop while those run, items
 * of different scenes concurrently from a few host threads (a scene's items
 * in item order; XITILS_AMD_BATCH_THREADS, default 3).  Items that share a film accumulate into it in item order
 * (all of them through the wavefront when none is guided, else each through
 * xt_render in turn).  Needs a bounce cap (1 <= max_depth <= 1024) and no
 * XT_FLAG_GUIDE_TRAIN.  Graphs are cached per scene, camera, parameters and
 * film.  ms: an item's GPU time on its stream; for a wavefront item, from
 * the start of the wavefront to the end of the window holding its last path. */
xt_status xt_render_batch(xt_batch_item* items, uint32_t n);
/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): xt_render into a device film (film_device) with every buffer of
 * the bounce loop -- path and shadow queues, hits, radiance, queue counters,
 * batch cursor, traversal spill stack -- carved from one static device array
 * of XT_STATIC_SCRATCH_BYTES per GPU, which the GPU holds until the process
 * ends.  Calls no hipMalloc, hipFree or allocating library routine.  Renders
 * in batches of as many paths as the array holds (about 220 k), so any image
 * size works; unguided only, with a bounce cap (1 <= max_depth <= 1024).
 * Gives xt_render's film bit for bit.  XT_ERR_TOO_LARGE if the scene's
 * traversal stack does not fit the array's spill part (16 MiB).  Every call
 * on one GPU uses that GPU's one array: calls that overlap on one GPU race. */
#define XT_STATIC_SCRATCH_BYTES (64u << 20)
xt_status xt_render_static(xt_scene* scene, const xt_camera* camera, const xt_render_params* params, void* film_device,
                           int film_is_f64);
/* Autotuner: renders one batch of (camera, params) as a graph under 9
 * launch shapes and keeps the fastest for this scene's later renders;
 * out = {traversal blocks per CU, shade blocks per CU}. */
xt_status xt_scene_autotune(xt_scene* scene, const xt_camera* camera, const xt_render_params* params, uint32_t out[2]);

/* Measurement hook (no reference counterpart): renders and ray queries of
 * this scene use the traversal variant that counts node and triangle fetches
 * into the stats (algorithmic-traffic accounting for the roofline).  Results
 * are identical. */
xt_status xt_scene_set_counting(xt_scene* scene, int on);

/* ---- VonMisesFisherDistribution<N> (VonMisesFisherDistribution.h:9-217) -----
 * Mixtures of num_lobes (1..16) lobes in the reference's encoding
 * (encode/decode, :37-58): per lobe mu.x, mu.y, mu.z, kappa, alpha, the last
 * lobe's alpha omitted (5 * num_lobes - 1 floats); decoding sets the last
 * alpha to 1 minus the running sum.  Arrays are host pointers, or device
 * pointers when on_device != 0.  The reference draws its uniforms from a
 * Sampler (Sampler.h:17-19); here the caller passes them. */

/* VonMisesFisherDistribution::eval (:60-66): pdf[i] of direction i under
 * distribution dist_of_dir[i] (NULL: distribution 0); an index >= num_dists
 * gives NaN. */
xt_status xt_vmf_eval(xt_device* dev, uint32_t num_lobes, const float* encoded, uint32_t num_dists,
                      const uint32_t* dist_of_dir, const float* dirs, uint32_t n, float* pdf, int on_device);
/* VonMisesFisherDistribution::sample (:162-201): one direction per 4
 * uniforms (lobe choice, disk radius, disk angle, W), the reference's
 * sampler including its unit-disk tangent draw and the W = 1 rule at
 * kappa = 0. */
xt_status xt_vmf_sample(xt_device* dev, uint32_t num_lobes, const float* encoded, uint32_t num_dists,
                        const uint32_t* dist_of_sample, const float* uniforms, uint32_t n, float* dirs, int on_device);
/* VonMisesFisherDistribution::approximateBySEM (:68-160) for num_fits
 * independent sample sets, set f = samples[offsets[f] .. offsets[f+1]) (xyz
 * unit vectors), with 2 * num_lobes initialisation uniforms per set (u, v per
 * lobe, :89-98).  Writes the encoded fits and, if steps is not NULL, the EM
 * steps each took.  An empty set yields the default distribution (the
 * reference divides by zero there). */
xt_status xt_vmf_fit(xt_device* dev, uint32_t num_lobes, const float* samples, uint32_t num_samples,
                     const uint32_t* offsets, uint32_t num_fits, const float* init_uniforms, float* encoded,
                     int32_t* steps, int on_device);

/* ---- learned path guiding --------------------------------------------------
 * No reference counterpart: the reference ships only the vMF mixture
 * (VonMisesFisherDistribution.h:9-217).  A 3x64 MLP maps (position, shading
 * normal, view direction) to an 8-lobe vMF mixture over incident directions;
 * raw outputs per lobe j are y[5j..5j+4] = (mu unnormalised, log kappa,
 * alpha logit).  Weights live per scene on the device: fp32 master copy,
 * bf16 copies for the MFMA kernels, Adam state.
 *
 * Online training (xt_render with XT_FLAG_GUIDE_TRAIN): after each render
 * batch, its samples go through Adam steps (a burst) on a stream of the
 * guide's own, which then publishes the network into one of two published
 * copies; xt_render returns without waiting, so the next render runs while
 * the burst does.  A training batch samples with the network the burst
 * before last published, any other guided render with the last one (its
 * stream waits for that burst on the device).  Which copy a render reads
 * depends only on the number of bursts, not on timing.  Every xt_guide_*
 * call below first waits for the bursts in flight. */
#define XT_GUIDE_NUM_PARAMS 17648   /* W0 64x96, b0 64, W1 64x64, b1, W2 64x64, b2, W3 48x64, b3 48 */
#define XT_GUIDE_NUM_INPUTS 9       /* query: p.xyz, shading n.xyz, wo.xyz */
#define XT_GUIDE_NUM_OUTPUTS 40

typedef struct xt_guide_config {
    float guide_prob;            /* one-sample MIS probability of sampling the guide (0.5) */
    float learning_rate;         /* Adam step size (5e-3) */
    float beta1, beta2, adam_eps;/* Adam (0.9, 0.999, 1e-8) */
    uint32_t minibatch;          /* samples per Adam step (32768) */
    uint32_t max_steps_per_batch;/* Adam steps after one render batch (256) */
    float weight_clip;           /* cap on sample weights normalised to mean 1 (20) */
    uint32_t seed;               /* parameter initialisation */
    uint32_t max_records;        /* training records kept per render batch (8 Mi) */
    uint32_t train_target;       /* 0: incident radiance L_i, 1: product f cos L_i (default) */
    uint32_t epochs;             /* passes over each render batch's samples, capped by max_steps_per_batch (3; 0 = 1) */
    uint32_t max_guided_vertex;  /* guide surface vertices 1..n of each path only (0 = all) */
} xt_guide_config;

void xt_guide_config_default(xt_guide_config* cfg);
/* Allocates the scene's network and initialises its parameters from cfg->seed
 * (He-uniform hidden layers, small output layer).  Encoding bounds = the
 * scene's bounding box. */
xt_status xt_guide_init(xt_scene* scene, const xt_guide_config* cfg);
/* Flat fp32 parameters (XT_GUIDE_NUM_PARAMS, layout above, row-major [out][in]). */
xt_status xt_guide_get_params(xt_scene* scene, float* params);
/* Replaces the parameters and clears the Adam state. */
xt_status xt_guide_set_params(xt_scene* scene, const float* params);
/* Bounds used to scale positions to [0, 1] for the frequency encoding. */
xt_status xt_guide_get_bounds(xt_scene* scene, float lo[3], float hi[3]);
/* Raw outputs (n x 40) of the bf16 network for n host queries (n x 9). */
xt_status xt_guide_forward(xt_scene* scene, const float* queries, uint32_t n, float* outputs);
/* Gradient of sum_i -min(w_i * weight_scale, weight_clip) log q(dir_i | query_i)
 * with respect to the flat parameters (host arrays: queries n x 9, dirs n x 3,
 * weights n; grad XT_GUIDE_NUM_PARAMS). */
xt_status xt_guide_gradient(xt_scene* scene, const float* queries, const float* dirs, const float* weights, uint32_t n,
                            float weight_scale, float weight_clip, float* grad);
/* One Adam step on the mean loss of n host samples (weights normalised to mean
 * 1 and clipped as in training). */
xt_status xt_guide_train_step(xt_scene* scene, const float* queries, const float* dirs, const float* weights,
                              uint32_t n);

#ifdef __cplusplus
}
#endif

#endif /* XITILS_AMD_H */
