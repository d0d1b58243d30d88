// xitils.hpp — C++17 host API over the xitils_amd C ABI.
//
// Keeps the shape of the reference's scene API (Scene.h, Object.h,
// TriangleMesh.h, Material.h, Camera.h, SkySphere.h, Transform.h,
// RenderTarget.h, PathTracer.h) so a caller written against it switches by
// changing includes; the per-sample work runs on the GPU through
// xt_render().  Header-only; link with libxitils_amd.so.
//
//   auto scene = std::make_shared<xitils::Scene>();
//   scene->addObject(std::make_shared<xitils::Object>(mesh, material, xitils::transformTRS(...)));
//   scene->buildAccelerationStructure();
//   xitils::SimpleRenderTarget target(800, 800);
//   target.render(*scene, spp, camera, xitils::StandardPathTracer());   // += spp samples per pixel
//
// Errors of the C ABI surface as xitils::Error (a std::runtime_error).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../xitils_amd.h"

namespace xitils {

static const float Pi = 3.14159265358979323846f;
static const float ToRad = Pi / 180.0f;

struct Error : std::runtime_error {
    xt_status status;
    Error(xt_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
inline void check(xt_status s) {
    if (s != XT_OK) throw Error(s, xt_last_error());
}

// ------------------------------------------------------------------ math
struct Vector3f {
    float x = 0, y = 0, z = 0;
    Vector3f() {}
    explicit Vector3f(float v) : x(v), y(v), z(v) {}
    Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
    Vector3f operator*(float s) const { return Vector3f(x * s, y * s, z * s); }
};
struct Vector2f {
    float x = 0, y = 0;
};

// Transform with the explicit inverse the reference keeps (Transform.h:11-60).
struct Transform {
    float mat[4][4];
    float matInv[4][4];
    Transform() {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) mat[i][j] = matInv[i][j] = (i == j) ? 1.0f : 0.0f;
    }
    static void mul(const float a[4][4], const float b[4][4], float out[4][4]) {   // Matrix.h:184-192
        float r[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] + a[i][3] * b[3][j];
        std::memcpy(out, r, sizeof r);
    }
    // Transform(mat): inverse = transpose, valid for rotations (Transform.h:22-25)
    static Transform rotation(const float m[4][4]) {
        Transform t;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                t.mat[i][j] = m[i][j];
                t.matInv[i][j] = m[j][i];
            }
        return t;
    }
    Transform operator*(const Transform& t) const {   // Transform.h:52-54
        Transform r;
        mul(mat, t.mat, r.mat);
        mul(t.matInv, matInv, r.matInv);
        return r;
    }
    bool isIdentity() const {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (mat[i][j] != (i == j ? 1.0f : 0.0f)) return false;
        return true;
    }
};

inline Transform translate(float x, float y, float z) {   // Transform.h:227-241
    Transform t;
    t.mat[0][3] = x; t.mat[1][3] = y; t.mat[2][3] = z;
    t.matInv[0][3] = -x; t.matInv[1][3] = -y; t.matInv[2][3] = -z;
    return t;
}
inline Transform translate(const Vector3f& v) { return translate(v.x, v.y, v.z); }
inline Transform scale(float x, float y, float z) {       // Transform.h:247-261
    Transform t;
    t.mat[0][0] = x; t.mat[1][1] = y; t.mat[2][2] = z;
    t.matInv[0][0] = 1 / x; t.matInv[1][1] = 1 / y; t.matInv[2][2] = 1 / z;
    return t;
}
inline Transform rotateX(float deg) {                      // Transform.h:267-277
    float r = deg * ToRad, s = std::sin(r), c = std::cos(r);
    float m[4][4] = {{1, 0, 0, 0}, {0, c, -s, 0}, {0, s, c, 0}, {0, 0, 0, 1}};
    return Transform::rotation(m);
}
inline Transform rotateY(float deg) {                      // Transform.h:279-289
    float r = deg * ToRad, s = std::sin(r), c = std::cos(r);
    float m[4][4] = {{c, 0, s, 0}, {0, 1, 0, 0}, {-s, 0, c, 0}, {0, 0, 0, 1}};
    return Transform::rotation(m);
}
inline Transform rotateZ(float deg) {                      // Transform.h:291-301
    float r = deg * ToRad, s = std::sin(r), c = std::cos(r);
    float m[4][4] = {{c, -s, 0, 0}, {s, c, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    return Transform::rotation(m);
}
inline Transform rotateYXZ(float x, float y, float z) { return rotateY(y) * rotateX(x) * rotateZ(z); }  // Transform.h:303-308
inline Transform transformTRS(const Vector3f& t, const Vector3f& r, const Vector3f& s) {                   // Transform.h:370-374
    return translate(t.x, t.y, t.z) * rotateYXZ(r.x, r.y, r.z) * scale(s.x, s.y, s.z);
}

// ------------------------------------------------------------------ textures (Texture.h)
// An RGB float texture read with Texture::rgb's bilinear filter and clamped
// addressing (Texture.h:83-103).  The reference decodes files with stb_image
// (Texture.h:17-23); this API takes the decoded texels, row 0 first.
struct Texture {
    std::vector<float> rgb;
    int width = 0, height = 0;
    Texture(const float* texels, int w, int h) : rgb(texels, texels + 3 * (size_t)w * h), width(w), height(h) {}
};

// ------------------------------------------------------------------ materials (Material.h)
struct Material {
    bool emissive = false;
    std::shared_ptr<Texture> normalmap;   // Material::normalmap, applied as Object.h:59-68
    virtual ~Material() = default;
    virtual xt_material record() const = 0;
};
inline xt_material make_record(uint32_t type, std::initializer_list<float> p) {
    xt_material m{};
    m.type = type;
    int k = 0;
    for (float v : p) m.p[k++] = v;
    return m;
}
struct Diffuse : Material {                                 // Material.h:50-86
    Vector3f albedo;
    explicit Diffuse(const Vector3f& a) : albedo(a) {}
    xt_material record() const override { return make_record(XT_MAT_DIFFUSE, {albedo.x, albedo.y, albedo.z}); }
};
struct SpecularReflection : Material {                      // Material.h:177-189
    xt_material record() const override { return make_record(XT_MAT_SPEC_REFL, {}); }
};
struct SpecularRefraction : Material {                      // Material.h:191-218
    float index = 1.0f;
    xt_material record() const override { return make_record(XT_MAT_SPEC_REFR, {index}); }
};
struct SpecularFresnel : Material {                         // Material.h:220-273
    float index = 1.0f;
    xt_material record() const override { return make_record(XT_MAT_SPEC_FRESNEL, {index}); }
};
struct Metal : Material {                                   // Material.h:275-286
    float alpha;
    Vector3f f0;
    Metal(float roughness, const Vector3f& f) : alpha(roughness * roughness), f0(f) {}
    xt_material record() const override { return make_record(XT_MAT_METAL, {alpha, f0.x, f0.y, f0.z}); }
};
struct Emission : Material {                                // Material.h:386-395
    Vector3f power;
    explicit Emission(const Vector3f& p) : power(p) { emissive = true; }
    xt_material record() const override { return make_record(XT_MAT_EMISSION, {power.x, power.y, power.z}); }
};
// GGX specular over a Lambert base, one-sample lobe mixture (no reference twin).
struct GGXOverLambert : Material {
    Vector3f albedo;
    float alpha, f0, specularProbability;
    GGXOverLambert(const Vector3f& a, float al, float f, float ps = 0.5f) : albedo(a), alpha(al), f0(f), specularProbability(ps) {}
    xt_material record() const override {
        return make_record(XT_MAT_GGX_LAMBERT, {albedo.x, albedo.y, albedo.z, alpha, f0, specularProbability});
    }
};

// MicrofacetNormalMapping (Sandbox/MicrofacetBasedNormalMapping/Main.cpp:
// 86-340, SameMaterialExplicit) over a Diffuse, Metal or GGXOverLambert base.
struct MicrofacetNormalMapping : Material {
    std::shared_ptr<Material> material_wp;
    explicit MicrofacetNormalMapping(std::shared_ptr<Material> base) : material_wp(std::move(base)) {}
    xt_material record() const override {
        xt_material b = material_wp->record();
        xt_material m{};
        m.type = XT_MAT_MICROFACET_NMAP;
        m.p[0] = (float)b.type;
        for (int k = 0; k < 6; ++k) m.p[k + 1] = b.p[k];
        return m;
    }
};

// ------------------------------------------------------------------ geometry
struct Shape {
    virtual ~Shape() = default;
};

// TriangleMesh::setGeometry raw-array form (TriangleMesh.h:151-161).
struct TriangleMesh : Shape {
    std::vector<float> positions, normals, texCoords, tangents, bitangents;
    std::vector<int32_t> indices;
    void setGeometry(int vertexNum, const Vector3f* pos, const Vector2f* uv, const Vector3f* nrm,
                     const Vector3f* tan, const Vector3f* bitan, int indexNum, const int* indexData) {
        auto put3 = [vertexNum](std::vector<float>& dst, const Vector3f* src) {
            dst.clear();
            if (!src) return;
            dst.resize(3 * (size_t)vertexNum);
            for (int i = 0; i < vertexNum; ++i) {
                dst[3 * i] = src[i].x; dst[3 * i + 1] = src[i].y; dst[3 * i + 2] = src[i].z;
            }
        };
        put3(positions, pos);
        put3(normals, nrm);
        put3(tangents, tan);
        put3(bitangents, bitan);
        texCoords.clear();
        if (uv) {
            texCoords.resize(2 * (size_t)vertexNum);
            for (int i = 0; i < vertexNum; ++i) {
                texCoords[2 * i] = uv[i].x; texCoords[2 * i + 1] = uv[i].y;
            }
        }
        indices.assign(indexData, indexData + indexNum);
    }
    int triangleNum() const { return (int)indices.size() / 3; }
};

struct Cube : TriangleMesh {                                // TriangleMesh.h:369-397
    Cube() {
        Vector3f p[8];
        for (int i = 0; i < 8; ++i) p[i] = Vector3f(i & 1 ? -0.5f : 0.5f, i & 2 ? -0.5f : 0.5f, i & 4 ? -0.5f : 0.5f);
        const int idx[36] = {0, 1, 2, 1, 3, 2, 2, 3, 6, 3, 7, 6, 0, 4, 1, 4, 5, 1,
                             0, 2, 6, 0, 6, 4, 1, 5, 3, 5, 7, 3, 4, 6, 5, 6, 7, 5};
        setGeometry(8, p, nullptr, nullptr, nullptr, nullptr, 36, idx);
    }
};

struct Plane : TriangleMesh {                               // TriangleMesh.h:399-421
    Plane() {
        Vector3f p[4];
        for (int i = 0; i < 4; ++i) p[i] = Vector3f(i & 1 ? -0.5f : 0.5f, i & 2 ? -0.5f : 0.5f, 0);
        const int idx[6] = {0, 2, 1, 1, 2, 3};
        setGeometry(4, p, nullptr, nullptr, nullptr, nullptr, 6, idx);
    }
};

struct Object {                                             // Object.h:16-40
    std::shared_ptr<TriangleMesh> shape;
    std::shared_ptr<Material> material;
    Transform objectToWorld;
    Object(std::shared_ptr<TriangleMesh> s, std::shared_ptr<Material> m, const Transform& t)
        : shape(std::move(s)), material(std::move(m)), objectToWorld(t) {}
};

// ------------------------------------------------------------------ sky, camera
struct SkySphere {
    virtual ~SkySphere() = default;
    virtual xt_env env() const = 0;
};
struct SkySphereUniform : SkySphere {                       // SkySphere.h:36-46
    Vector3f color;
    explicit SkySphereUniform(const Vector3f& c) : color(c) {}
    xt_env env() const override {
        xt_env e{};
        e.has_sky = 1;
        e.sky[0] = color.x; e.sky[1] = color.y; e.sky[2] = color.z;
        return e;
    }
};
// Uniform sky plus a sun disk of the given half angle (degrees).
struct SkySphereWithSun : SkySphereUniform {
    Vector3f sunDirection, sunRadiance;
    double halfAngleDeg;
    SkySphereWithSun(const Vector3f& sky, const Vector3f& dir, double halfDeg, const Vector3f& L)
        : SkySphereUniform(sky), sunDirection(dir), sunRadiance(L), halfAngleDeg(halfDeg) {}
    xt_env env() const override {
        xt_env e = SkySphereUniform::env();
        double th = halfAngleDeg * 3.14159265358979323846 / 180.0;
        double n = std::sqrt((double)sunDirection.x * sunDirection.x + (double)sunDirection.y * sunDirection.y +
                             (double)sunDirection.z * sunDirection.z);
        e.sun_dir[0] = (float)(sunDirection.x / n); e.sun_dir[1] = (float)(sunDirection.y / n); e.sun_dir[2] = (float)(sunDirection.z / n);
        e.sun_one_minus_cos = (float)(2.0 * std::sin(th / 2) * std::sin(th / 2));
        e.sun_sin2 = (float)(std::sin(th) * std::sin(th));
        e.sun_radiance[0] = sunRadiance.x; e.sun_radiance[1] = sunRadiance.y; e.sun_radiance[2] = sunRadiance.z;
        return e;
    }
};

// SkySphereFromImage (SkySphere.h:14-34): an angular-map light probe.  The
// reference loads the probe with stb_image (Texture.h:17-23); this API takes
// the decoded RGB float texels (row 0 first), so any loader can feed it.
struct SkySphereFromImage : SkySphere {
    std::vector<float> rgb;
    int width = 0, height = 0;
    SkySphereFromImage(const float* texels, int w, int h) : rgb(texels, texels + 3 * (size_t)w * h), width(w), height(h) {}
    xt_env env() const override {
        xt_env e{};
        e.has_sky = 1;
        e.map_kind = XT_ENV_MAP_ANGULAR;
        e.map_width = (uint32_t)width;
        e.map_height = (uint32_t)height;
        e.map_rgb = rgb.data();
        return e;
    }
};
// Latitude-longitude HDR sky (no reference counterpart), importance-sampled
// as a light when the integrator sets XT_FLAG_ENV_NEE.
struct SkySphereLatLong : SkySphereFromImage {
    SkySphereLatLong(const float* texels, int w, int h) : SkySphereFromImage(texels, w, h) {}
    xt_env env() const override {
        xt_env e = SkySphereFromImage::env();
        e.map_kind = XT_ENV_MAP_LATLONG;
        return e;
    }
};

struct Camera {
    virtual ~Camera() = default;
    virtual xt_camera record() const = 0;
};
struct PinholeCamera : Camera {                             // Camera.h:17-55
    Transform cameraToWorld;
    float fov, aspectRatio;
    PinholeCamera(const Transform& t, float fovRad, float aspect) : cameraToWorld(t), fov(fovRad), aspectRatio(aspect) {}
    xt_camera record() const override {
        xt_camera c{};
        for (int i = 0; i < 16; ++i) c.cam_to_world[i] = cameraToWorld.mat[i / 4][i % 4];
        c.tan_x = std::tan(fov / 2);
        c.tan_y = std::tan(fov / 2 * aspectRatio);
        return c;
    }
};

// ------------------------------------------------------------------ integrators (PathTracer.h)
struct PathTracer {
    virtual ~PathTracer() = default;
    virtual uint32_t integrator() const = 0;
    int maxDepth = 0;       // 0: roulette only, as the reference
    uint32_t flags = 0;     // XT_FLAG_*
};
struct StandardPathTracer : PathTracer {
    uint32_t integrator() const override { return XT_INTEGRATOR_STANDARD; }
};
struct NaivePathTracer : PathTracer {
    uint32_t integrator() const override { return XT_INTEGRATOR_NAIVE; }
};

// ------------------------------------------------------------------ device + scene
class Device {
   public:
    static Device& get(int ordinal = 0) {
        static Device d(ordinal);
        return d;
    }
    xt_device* handle() const { return h_; }
    ~Device() { xt_device_close(h_); }

   private:
    explicit Device(int ordinal) { check(xt_device_open(ordinal, &h_)); }
    xt_device* h_ = nullptr;
};

class Scene {                                               // Scene.h:11-61
   public:
    std::shared_ptr<Camera> camera;
    std::shared_ptr<SkySphere> skySphere;

    void addObject(std::shared_ptr<Object> object) { objects_.push_back(std::move(object)); }

    void buildAccelerationStructure() {
        std::vector<xt_object_desc> od(objects_.size());
        std::vector<xt_material> mats;
        std::vector<uint32_t> nmaps;              // per material: 1 + texture index, 0 = none
        std::vector<xt_texture> texs;
        std::vector<const Texture*> texOwners;
        std::vector<std::vector<float>> xf(objects_.size()), xfi(objects_.size());
        for (size_t k = 0; k < objects_.size(); ++k) {
            const Object& o = *objects_[k];
            xt_object_desc& d = od[k];
            d = xt_object_desc{};
            d.positions = o.shape->positions.data();
            d.normals = o.shape->normals.empty() ? nullptr : o.shape->normals.data();
            d.indices = o.shape->indices.data();
            d.num_vertices = (uint32_t)(o.shape->positions.size() / 3);
            d.num_triangles = (uint32_t)o.shape->triangleNum();
            d.texcoords = o.shape->texCoords.empty() ? nullptr : o.shape->texCoords.data();
            if (!o.shape->tangents.empty() && !o.shape->bitangents.empty()) {
                d.tangents = o.shape->tangents.data();
                d.bitangents = o.shape->bitangents.data();
            }
            if (!o.objectToWorld.isIdentity()) {
                xf[k].assign(&o.objectToWorld.mat[0][0], &o.objectToWorld.mat[0][0] + 16);
                xfi[k].assign(&o.objectToWorld.matInv[0][0], &o.objectToWorld.matInv[0][0] + 16);
                d.object_to_world = xf[k].data();
                d.world_to_object = xfi[k].data();
            }
            d.material = (uint32_t)mats.size();
            mats.push_back(o.material->record());
            uint32_t nm = 0;
            if (const Texture* t = o.material->normalmap.get()) {
                size_t ti = 0;
                while (ti < texOwners.size() && texOwners[ti] != t) ++ti;
                if (ti == texOwners.size()) {
                    texOwners.push_back(t);
                    texs.push_back(xt_texture{(uint32_t)t->width, (uint32_t)t->height, t->rgb.data()});
                }
                nm = (uint32_t)ti + 1;
            }
            nmaps.push_back(nm);
        }
        if (mats.empty()) {
            mats.push_back(make_record(XT_MAT_DIFFUSE, {0, 0, 0}));
            nmaps.push_back(0);
        }
        xt_scene_desc sd{};
        sd.objects = od.data();
        sd.num_objects = (uint32_t)od.size();
        sd.materials = mats.data();
        sd.num_materials = (uint32_t)mats.size();
        if (!texs.empty()) {
            sd.textures = texs.data();
            sd.num_textures = (uint32_t)texs.size();
            sd.material_normal_maps = nmaps.data();
        }
        if (skySphere) sd.env = skySphere->env();
        xt_scene* s = nullptr;
        check(xt_scene_create(Device::get().handle(), &sd, &s));
        handle_.reset(s, [](xt_scene* p) { xt_scene_destroy(p); });
    }

    xt_scene* handle() const {
        if (!handle_) throw Error(XT_ERR_INVALID_ARGUMENT, "Scene::buildAccelerationStructure() not called");
        return handle_.get();
    }

   private:
    std::vector<std::shared_ptr<Object>> objects_;
    std::shared_ptr<xt_scene> handle_;
};

// ------------------------------------------------------------------ film (RenderTarget.h)
template <typename T>
class RenderTarget;

template <>
class RenderTarget<Vector3f> {
   public:
    std::vector<Vector3f> data;
    int width, height;
    int sampleNum = 0;   // samples accumulated so far (the caller's frameData.sampleNum)
    RenderTarget(int w, int h) : data((size_t)w * h), width(w), height(h) {}
    Vector3f& operator()(int x, int y) { return data[(size_t)x + (size_t)y * width]; }
    void clear() {
        std::fill(data.begin(), data.end(), Vector3f());
        sampleNum = 0;
    }
    // RenderTarget::render with the composite lambda of
    // Sandbox/SimplePathTracer/Main.cpp:134-138: color += eval(generateRay(pFilm)).
    void render(const Scene& scene, int samples, const Camera& camera, const PathTracer& pt, uint32_t seed = 1) {
        xt_render_params p{};
        p.width = (uint32_t)width;
        p.height = (uint32_t)height;
        p.spp = (uint32_t)samples;
        p.spp_offset = (uint32_t)sampleNum;
        p.max_depth = pt.maxDepth;
        p.seed = seed;
        p.integrator = pt.integrator();
        p.flags = pt.flags;
        p.rr_start = 5;
        p.rr_prob = 0.9f;
        p.ray_eps = 0.00001f;
        p.shadow_margin = 0.0001f;
        xt_camera c = camera.record();
        static_assert(sizeof(Vector3f) == 12, "Vector3f must be 3 packed floats");
        check(xt_render(scene.handle(), &c, &p, data.data(), 0, 0));
        sampleNum += samples;
    }
    void render(const Scene& scene, int samples, const PathTracer& pt, uint32_t seed = 1) {
        if (!scene.camera) throw Error(XT_ERR_INVALID_ARGUMENT, "scene has no camera");
        render(scene, samples, *scene.camera, pt, seed);
    }
};
using SimpleRenderTarget = RenderTarget<Vector3f>;

// render(scene, camera, spp): a fresh film holding the sum of spp samples per pixel.
inline SimpleRenderTarget render(const Scene& scene, const Camera& camera, int width, int height, int spp,
                                 const PathTracer& pt = StandardPathTracer()) {
    SimpleRenderTarget t(width, height);
    t.render(scene, spp, camera, pt);
    return t;
}

}  // namespace xitils
