// ORACLE — test infrastructure only: extern "C" entry points over xo_core.hpp
// for the pytest suite (ctypes) and the CPU-baseline leg of bench.py.
//
// RenderTarget<T>::render (RenderTarget.h:110-127) is restated as an OpenMP
// `schedule(dynamic, 1)` loop over 16x16 tiles, then samples, then pixels.
#include "xo_core.hpp"
#include "xo_vmf.hpp"

#include <omp.h>

using namespace xo;

extern "C" {

void* orc_scene_new(void) { return new Scene(); }
void orc_scene_free(void* s) { delete (Scene*)s; }

int orc_scene_add_material(void* s, uint32_t type, const float* p7) {
    Scene* sc = (Scene*)s;
    Material m;
    m.type = type;
    for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    sc->materials.push_back(m);
    return (int)sc->materials.size() - 1;
}

// positions/normals are num_vertices*3 floats in object space; xf16/xfinv16
// are the row-major 4x4 Transform::mat / matInv (NULL = identity).
int orc_scene_add_object(void* s, const float* positions, const float* normals, int num_vertices,
                         const int32_t* indices, int num_triangles, const float* xf16, const float* xfinv16,
                         uint32_t material, const uint32_t* tri_materials) {
    Scene* sc = (Scene*)s;
    Object o;
    o.mesh.pos.resize(num_vertices);
    for (int i = 0; i < num_vertices; ++i) o.mesh.pos[i] = V3(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]);
    if (normals) {
        o.mesh.nrm.resize(num_vertices);
        for (int i = 0; i < num_vertices; ++i) o.mesh.nrm[i] = V3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    }
    o.mesh.idx.assign(indices, indices + 3 * num_triangles);
    if (xf16) {
        for (int i = 0; i < 16; ++i) o.xf.m[i / 4][i % 4] = xf16[i];
        for (int i = 0; i < 16; ++i) o.xf.mi[i / 4][i % 4] = xfinv16[i];
    }
    o.material = material;
    if (tri_materials) o.triMaterial.assign(tri_materials, tri_materials + num_triangles);
    sc->objects.push_back(std::move(o));
    return (int)sc->objects.size() - 1;
}

// Optional per-vertex texCoords (2 floats) and tangents/bitangents (3 floats).
void orc_object_set_attributes(void* s, int obj, const float* texcoords, const float* tangents,
                               const float* bitangents) {
    Mesh& m = ((Scene*)s)->objects[obj].mesh;
    const size_t nv = m.pos.size();
    if (texcoords) m.uv.assign(texcoords, texcoords + 2 * nv);
    if (tangents && bitangents) {
        m.tan.resize(nv);
        m.bitan.resize(nv);
        for (size_t i = 0; i < nv; ++i) {
            m.tan[i] = V3(tangents[3 * i], tangents[3 * i + 1], tangents[3 * i + 2]);
            m.bitan[i] = V3(bitangents[3 * i], bitangents[3 * i + 1], bitangents[3 * i + 2]);
        }
    }
}

int orc_scene_add_texture(void* s, int width, int height, const float* rgb) {
    Scene* sc = (Scene*)s;
    Texture2D t;
    t.w = width;
    t.h = height;
    t.rgb.assign(rgb, rgb + 3 * (size_t)width * height);
    sc->textures.push_back(std::move(t));
    return (int)sc->textures.size() - 1;
}

// Material::normalmap of material `mat` (texture index, -1 = none)
void orc_material_set_normal_map(void* s, int mat, int tex) {
    Scene* sc = (Scene*)s;
    if ((int)sc->matNormalMap.size() <= mat) sc->matNormalMap.resize(mat + 1, -1);
    sc->matNormalMap[mat] = tex;
}

// sky: has_sky, rgb[3]; sun: dir[3] (unit), 1 - cos(half), sin^2(half), rgb[3]
void orc_scene_set_env(void* s, int has_sky, const float* sky_rgb, const float* sun_dir, float sun_omc,
                       float sun_sin2, const float* sun_rgb) {
    Scene* sc = (Scene*)s;
    sc->env.hasSky = has_sky != 0;
    sc->env.sky = V3(sky_rgb[0], sky_rgb[1], sky_rgb[2]);
    sc->env.sunDir = V3(sun_dir[0], sun_dir[1], sun_dir[2]);
    sc->env.sunOmc = sun_omc;
    sc->env.sunSin2 = sun_sin2;
    sc->env.sunL = V3(sun_rgb[0], sun_rgb[1], sun_rgb[2]);
}

// Image sky: kind 1 = angular map (SkySphereFromImage), 2 = lat-long.
void orc_scene_set_env_map(void* s, int kind, int width, int height, const float* rgb) {
    Env& e = ((Scene*)s)->env;
    e.mapKind = kind;
    e.mapW = width;
    e.mapH = height;
    e.map.assign(rgb, rgb + 3 * (size_t)width * height);
    e.buildMapDistribution();
}

// SkySphere::getRadiance (+ sun) for n ray directions.
void orc_env_radiance(void* s, const float* dirs, int n, float* out) {
    const Env& e = ((Scene*)s)->env;
    for (int i = 0; i < n; ++i) {
        V3 L = e.radiance(V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
    }
}

// Lat-long distribution: solid-angle pdf of unit directions, and samples
// (direction + pdf) from uniform pairs.
void orc_env_map_pdf(void* s, const float* dirs, int n, float* out) {
    const Env& e = ((Scene*)s)->env;
    for (int i = 0; i < n; ++i) out[i] = e.mapPdf(V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
void orc_env_map_sample(void* s, const float* u2, int n, float* out4) {
    const Env& e = ((Scene*)s)->env;
    for (int i = 0; i < n; ++i) {
        float pdf;
        V3 d = e.sampleMap(u2[2 * i], u2[2 * i + 1], &pdf);
        out4[4 * i] = d.x; out4[4 * i + 1] = d.y; out4[4 * i + 2] = d.z; out4[4 * i + 3] = pdf;
    }
}

int orc_scene_build(void* s) {
    ((Scene*)s)->build();
    return 0;
}

uint32_t orc_scene_num_prims(void* s) { return ((Scene*)s)->nprim; }

static Ray rayFrom(const float* r8) {
    Ray r;
    r.o = V3(r8[0], r8[1], r8[2]);
    r.tMax = r8[3];
    r.d = V3(r8[4], r8[5], r8[6]);
    return r;
}

// rays: n x {ox, oy, oz, tmax, dx, dy, dz, pad}; hits: n x {t, prim(u32 bits), b0, b1}
// mode 0 = two-level reference BVH, 1 = brute force.  Miss: t = +inf, prim = 0xFFFFFFFF.
void orc_trace_closest(void* s, const float* rays, int n, float* hits, int mode, int nthreads) {
    const Scene* sc = (const Scene*)s;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int i = 0; i < n; ++i) {
        Ray r = rayFrom(rays + 8 * i);
        float t = Infinity, b0 = 0, b1 = 0;
        uint32_t prim = 0xFFFFFFFFu;
        if (mode == 0) {
            Isect is;
            if (sc->intersect(r, &is)) { t = is.t; prim = is.prim; b0 = is.b0; b1 = is.b1; }
        } else {
            sc->intersectBrute(r, &t, &prim, &b0, &b1);
        }
        hits[4 * i + 0] = t;
        std::memcpy(&hits[4 * i + 1], &prim, 4);
        hits[4 * i + 2] = b0;
        hits[4 * i + 3] = b1;
    }
}

void orc_trace_any(void* s, const float* rays, int n, uint8_t* occluded, int mode, int nthreads) {
    const Scene* sc = (const Scene*)s;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int i = 0; i < n; ++i) {
        Ray r = rayFrom(rays + 8 * i);
        occluded[i] = (mode == 0 ? sc->intersectAny(r) : sc->intersectAnyBrute(r)) ? 1 : 0;
    }
}

// cam: 16 floats camToWorld (row-major mat) then tanX, tanY.
// ip: width, height, spp, spp_offset, max_depth, seed, stream, integrator, flags, rr_start
// fp: rr_prob, eps, shadow_margin
// film32 (W*H*3, +=) and/or film64 (W*H*3, +=) may be NULL.  stats: {ext, shadow}.
void orc_render(void* s, const float* cam, const int32_t* ip, const float* fp, float* film32, double* film64,
                uint64_t* stats, int nthreads) {
    const Scene* sc = (const Scene*)s;
    Camera c;
    for (int i = 0; i < 16; ++i) c.camToWorld.m[i / 4][i % 4] = cam[i];
    c.tanX = cam[16];
    c.tanY = cam[17];
    RenderParams rp;
    rp.width = ip[0]; rp.height = ip[1]; rp.spp = ip[2]; rp.sppOffset = ip[3]; rp.maxDepth = ip[4];
    rp.seed = (uint32_t)ip[5]; rp.stream = (uint32_t)ip[6]; rp.integrator = ip[7]; rp.flags = (uint32_t)ip[8];
    rp.rrStart = ip[9];
    rp.rrProb = fp[0]; rp.eps = fp[1]; rp.shadowMargin = fp[2];
    const int W = rp.width, H = rp.height, TS = 16;
    const int tx = (W + TS - 1) / TS, ty = (H + TS - 1) / TS;
    uint64_t ext = 0, shd = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ext, shd) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int ti = 0; ti < tx * ty; ++ti) {
        const int ox = (ti % tx) * TS, oy = (ti / tx) * TS;
        PathStats st;
        for (int sIdx = 0; sIdx < rp.spp; ++sIdx) {
            for (int ly = 0; ly < TS; ++ly) {
                if (oy + ly >= H) continue;
                for (int lx = 0; lx < TS; ++lx) {
                    if (ox + lx >= W) continue;
                    const int x = ox + lx, y = oy + ly;
                    PathRng rng;
                    rng.key[0] = rp.seed; rng.key[1] = rp.stream;
                    rng.pixel = (uint32_t)(y * W + x);
                    rng.sample = (uint32_t)(rp.sppOffset + sIdx);
                    float fx, fy;
                    filmPosition(x, y, W, H, rng.uni(0, DIM_FILM_X), rng.uni(0, DIM_FILM_Y), &fx, &fy);
                    Ray ray = c.generate(fx, fy);
                    V3 L = rp.integrator == 1 ? evalNaive(*sc, rng, ray, rp, &st) : evalStandard(*sc, rng, ray, rp, &st);
                    size_t o = 3 * (size_t)(y * W + x);
                    if (film32) { film32[o] += L.x; film32[o + 1] += L.y; film32[o + 2] += L.z; }
                    if (film64) { film64[o] += L.x; film64[o + 1] += L.y; film64[o + 2] += L.z; }
                }
            }
        }
        ext += st.extRays;
        shd += st.shadowRays;
    }
    if (stats) { stats[0] = ext; stats[1] = shd; }
}

// ---- unit-level entry points for the pin tests --------------------------

// The oracle's transcendentals over n arguments: kind 0 sin, 1 cos, 2 x^5,
// 3 atan2(x[2i], x[2i+1]) (n pairs), 4 acos.
void orc_libm(int kind, const float* x, int n, float* out) {
    for (int i = 0; i < n; ++i) {
        float s, c;
        switch (kind) {
            case 0: sinCos(x[i], &s, &c); out[i] = s; break;
            case 1: sinCos(x[i], &s, &c); out[i] = c; break;
            case 2: out[i] = pow5(x[i]); break;
            case 3: out[i] = atan2F(x[2 * i], x[2 * i + 1]); break;
            default: out[i] = acosF(x[i]); break;
        }
    }
}

// RenderTargetTile::GenerateFilmPosition for n (x, y, ux, uy) rows -> (fx, fy).
void orc_film_position(const float* xyuv, int n, int W, int H, float* out) {
    for (int i = 0; i < n; ++i)
        filmPosition((int)xyuv[4 * i], (int)xyuv[4 * i + 1], W, H, xyuv[4 * i + 2], xyuv[4 * i + 3], &out[2 * i],
                     &out[2 * i + 1]);
}

void orc_philox(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4) { philox4x32_10(ctr4, key2, out4); }

float orc_uniform(uint32_t seed, uint32_t stream, uint32_t pixel, uint32_t sample, uint32_t slot, int dim) {
    PathRng r;
    r.key[0] = seed; r.key[1] = stream; r.pixel = pixel; r.sample = sample;
    return r.uni(slot, dim);
}

// isect9: p[3], n[3], wo[3] (shading normal = faceForward(n, wo)).
static Isect isectFrom(const float* i9, const Material* m) {
    Isect is;
    is.p = V3(i9[0], i9[1], i9[2]);
    is.n = V3(i9[3], i9[4], i9[5]);
    is.wo = V3(i9[6], i9[7], i9[8]);
    is.sn = faceForward(is.n, is.wo);
    is.mat = m;
    return is;
}

// out: wi[3], pdf, value[3]
void orc_bsdf_sample(uint32_t type, const float* p7, const float* isect9, const float* u3, float* out7) {
    Material m; m.type = type; for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    Isect is = isectFrom(isect9, &m);
    V3 wi; float pdf = 0;
    V3 v = evalAndSample(m, is, u3[0], u3[1], u3[2], &wi, &pdf);
    out7[0] = wi.x; out7[1] = wi.y; out7[2] = wi.z; out7[3] = pdf; out7[4] = v.x; out7[5] = v.y; out7[6] = v.z;
}

// out: bsdfCos[3], pdf
void orc_bsdf_eval(uint32_t type, const float* p7, const float* isect9, const float* wi3, float* out4) {
    Material m; m.type = type; for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    Isect is = isectFrom(isect9, &m);
    V3 wi(wi3[0], wi3[1], wi3[2]);
    V3 f = bsdfCos(m, is, wi);
    out4[0] = f.x; out4[1] = f.y; out4[2] = f.z;
    out4[3] = bsdfPdf(m, is, wi);
}

// The same with an explicit shading normal and the walk stream of the
// stochastic material: isect12 = p, n, wo, shading n; rng6 = seed, stream,
// pixel, sample, vertex slot, base dim.
static Isect isectFrom12(const float* i12, const Material* m) {
    Isect is = isectFrom(i12, m);
    is.sn = V3(i12[9], i12[10], i12[11]);
    return is;
}
static PathRng rngFrom(const uint32_t* r6) {
    PathRng r;
    r.key[0] = r6[0]; r.key[1] = r6[1]; r.pixel = r6[2]; r.sample = r6[3];
    return r;
}
void orc_bsdf_sample_walk(uint32_t type, const float* p7, const float* isect12, const uint32_t* rng6, float* out7) {
    Material m; m.type = type; for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    Isect is = isectFrom12(isect12, &m);
    PathRng r = rngFrom(rng6);
    Walk w{&r, rng6[4], (int)rng6[5]};
    V3 wi; float pdf = 0;
    V3 v = evalAndSample(m, is, r.uni(rng6[4], DIM_BSDF_SEL), r.uni(rng6[4], DIM_BSDF_U1), r.uni(rng6[4], DIM_BSDF_U2),
                         &wi, &pdf, &w);
    out7[0] = wi.x; out7[1] = wi.y; out7[2] = wi.z; out7[3] = pdf; out7[4] = v.x; out7[5] = v.y; out7[6] = v.z;
}
void orc_bsdf_eval_walk(uint32_t type, const float* p7, const float* isect12, const float* wi3, const uint32_t* rng6,
                        float* out4) {
    Material m; m.type = type; for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    Isect is = isectFrom12(isect12, &m);
    PathRng r = rngFrom(rng6);
    Walk w{&r, rng6[4], (int)rng6[5]};
    V3 wi(wi3[0], wi3[1], wi3[2]);
    V3 f = bsdfCos(m, is, wi, &w);
    out4[0] = f.x; out4[1] = f.y; out4[2] = f.z;
    out4[3] = bsdfPdf(m, is, wi);
}
// G1 and lambda_p of the two-facet model (Main.cpp:323-337) for a frame
// (wg, wp, wt) and direction w: out2 = G1(w, wm), lambda_p(w).
void orc_nmap_terms(const float* w3, const float* wm3, const float* wp3, const float* wt3, const float* wg3, float* out2) {
    V3 w(w3[0], w3[1], w3[2]), wm(wm3[0], wm3[1], wm3[2]), wp(wp3[0], wp3[1], wp3[2]);
    V3 wt(wt3[0], wt3[1], wt3[2]), wg(wg3[0], wg3[1], wg3[2]);
    out2[0] = nmapG1(w, wm, wp, wt, wg);
    out2[1] = nmapLambdaP(w, wp, wt, wg);
}

// Monte Carlo moments of a (stochastic) BSDF at one isect, n samples on
// `nthreads` threads.  mode 0: evalAndSample, x = returned weight; mode 1:
// bsdfCos at directions uniform on the sphere, x = f / (1 / 4pi).  out[0..2]
// = mean of x, out[3..5] = mean of mean(x) * wi, out[6] = variance of
// mean(x) (per sample), out[7] = fraction of pdf < 0 (delta) samples.
void orc_bsdf_moments(uint32_t type, const float* p7, const float* isect12, uint32_t seed, int n, int mode, int nthreads,
                      double* out8) {
    Material m; m.type = type; for (int i = 0; i < 7; ++i) m.p[i] = p7[i];
    Isect is = isectFrom12(isect12, &m);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        double loc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
        for (int k = 0; k < n; ++k) {
            PathRng r;
            r.key[0] = seed; r.key[1] = 77u; r.pixel = (uint32_t)k; r.sample = 0;
            V3 x, wi;
            if (mode == 0) {
                Walk w{&r, 1, DIM_WALK_SAMPLE};
                float pdf = 0;
                x = evalAndSample(m, is, r.uni(1, DIM_BSDF_SEL), r.uni(1, DIM_BSDF_U1), r.uni(1, DIM_BSDF_U2), &wi, &pdf, &w);
                if (pdf < 0) loc[7] += 1;
            } else {
                double z = 1.0 - 2.0 * r.uni(1, 0), ph = 2.0 * M_PI * r.uni(1, 1), rr = std::sqrt(std::max(0.0, 1 - z * z));
                wi = V3((float)(rr * std::cos(ph)), (float)(rr * std::sin(ph)), (float)z);
                Walk w{&r, 1, DIM_WALK_EVAL};
                x = bsdfCos(m, is, wi, &w) * (float)(4.0 * M_PI);
            }
            double xm = ((double)x.x + x.y + x.z) / 3.0;
            loc[0] += x.x; loc[1] += x.y; loc[2] += x.z;
            loc[3] += xm * wi.x; loc[4] += xm * wi.y; loc[5] += xm * wi.z;
            loc[6] += xm * xm;
        }
#pragma omp critical
        for (int j = 0; j < 8; ++j) acc[j] += loc[j];
    }
    for (int j = 0; j < 8; ++j) out8[j] = acc[j] / n;
    double mean = (out8[0] + out8[1] + out8[2]) / 3.0;
    out8[6] = out8[6] - mean * mean;
}

// Shading normal (world) at the closest hit of each ray: out 3 floats per
// ray, NaN on a miss.
void orc_shading_normals(void* s, const float* rays, int n, float* out) {
    const Scene* sc = (const Scene*)s;
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.o = V3(rays[8 * i], rays[8 * i + 1], rays[8 * i + 2]);
        r.tMax = rays[8 * i + 3];
        r.d = V3(rays[8 * i + 4], rays[8 * i + 5], rays[8 * i + 6]);
        Isect is;
        V3 sn(NAN, NAN, NAN);
        if (sc->intersect(r, &is)) sn = is.sn;
        out[3 * i] = sn.x; out[3 * i + 1] = sn.y; out[3 * i + 2] = sn.z;
    }
}

float orc_ggx_D(float alpha, const float* h3, const float* n3) {
    GGX g{alpha};
    return g.D(V3(h3[0], h3[1], h3[2]), V3(n3[0], n3[1], n3[2]));
}
float orc_ggx_G1(float alpha, const float* x3, const float* n3) {
    GGX g{alpha};
    return g.G1(V3(x3[0], x3[1], x3[2]), V3(n3[0], n3[1], n3[2]));
}

int orc_tri_test(const float* p9, const float* ray8, float* out3) {
    Ray r = rayFrom(ray8);
    float t = 0, b0 = 0, b1 = 0;
    int ok = triTest(V3(p9[0], p9[1], p9[2]), V3(p9[3], p9[4], p9[5]), V3(p9[6], p9[7], p9[8]), r, r.tMax, &t, &b0, &b1);
    out3[0] = t; out3[1] = b0; out3[2] = b1;
    return ok;
}

void orc_camera_ray(const float* cam, float fx, float fy, float* ray8) {
    Camera c;
    for (int i = 0; i < 16; ++i) c.camToWorld.m[i / 4][i % 4] = cam[i];
    c.tanX = cam[16]; c.tanY = cam[17];
    Ray r = c.generate(fx, fy);
    ray8[0] = r.o.x; ray8[1] = r.o.y; ray8[2] = r.o.z; ray8[3] = r.tMax;
    ray8[4] = r.d.x; ray8[5] = r.d.y; ray8[6] = r.d.z; ray8[7] = 0;
}

// Light sample for (seed, pixel, sample, slot): out p[3], n[3], sn[3], pdf, obj, tri
void orc_sample_light(void* s, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t slot, int uniform_bary,
                      float* out12) {
    const Scene* sc = (const Scene*)s;
    PathRng r; r.key[0] = seed; r.key[1] = 0; r.pixel = pixel; r.sample = sample;
    SampledLight l = sc->sampleLight(r, slot, uniform_bary != 0);
    float v[12] = {l.p.x, l.p.y, l.p.z, l.n.x, l.n.y, l.n.z, l.sn.x, l.sn.y, l.sn.z, l.pdf, (float)l.obj, (float)l.tri};
    std::memcpy(out12, v, sizeof v);
}

// Reference-BVH statistics: node count and max depth (structure checks).
void orc_ref_bvh_info(void* s, int obj, int32_t* out2) {
    const Scene* sc = (const Scene*)s;
    const RefBVH& b = obj < 0 ? sc->tlas : sc->objects[obj].blas;
    out2[0] = (int32_t)b.nodes.size();
    int maxd = 0;
    std::vector<std::pair<int, int>> st{{0, 0}};
    while (!st.empty()) {
        auto [n, d] = st.back(); st.pop_back();
        maxd = std::max(maxd, d);
        if (b.nodes[n].prim < 0) { st.push_back({b.nodes[n].child[0], d + 1}); st.push_back({b.nodes[n].child[1], d + 1}); }
    }
    out2[1] = maxd;
}

// ---- VonMisesFisherDistribution<N> (encoded as 5N - 1 floats)
void orc_vmf_eval(int N, const float* enc, const float* dirs, int n, float* out) {
    Vmf v(N);
    v.decode(enc);
    for (int i = 0; i < n; ++i) out[i] = v.eval(V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

void orc_vmf_sample(int N, const float* enc, const float* u4, int n, float* out) {
    Vmf v(N);
    v.decode(enc);
    for (int i = 0; i < n; ++i) {
        V3 d = v.sample(u4 + 4 * i);
        out[3 * i] = d.x;
        out[3 * i + 1] = d.y;
        out[3 * i + 2] = d.z;
    }
}

// Fits num_fits independent sample sets (offsets[f]..offsets[f+1]) in parallel.
void orc_vmf_fit(int N, const float* samples, const uint32_t* offsets, int num_fits, const float* init_u,
                 float* out_enc, int32_t* out_steps, const int32_t* fixed_steps, int nthreads) {
    const int E = 5 * N - 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int f = 0; f < num_fits; ++f) {
        std::vector<V3> smp;
        for (uint32_t i = offsets[f]; i < offsets[f + 1]; ++i)
            smp.push_back(V3(samples[3 * i], samples[3 * i + 1], samples[3 * i + 2]));
        int steps = 0;
        Vmf v = Vmf::fit(N, smp, init_u + 2 * N * f, &steps, fixed_steps ? fixed_steps[f] : 0);
        v.encode(out_enc + (size_t)E * f);
        out_steps[f] = steps;
    }
}

}  // extern "C"
