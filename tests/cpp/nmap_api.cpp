// A normal-mapped surface built through the C++ host API the way
// Sandbox/MicrofacetBasedNormalMapping/Main.cpp:374-427 builds its scene: a
// mesh with texCoords and a tangent frame, a GGX (Metal) base wrapped in
// MicrofacetNormalMapping with material->normalmap set, a lat-long HDR sky.
// The arrays come from files so the Python binding can render the same
// scene.  argv: out.f32 W H spp data_dir; data_dir holds mesh.bin, tex.bin,
// sky.bin and cam.bin (little-endian int32 counts, then float32 / int32 data).
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xitils/xitils.hpp"

using namespace xitils;

template <typename T>
static std::vector<T> read_block(FILE* f, size_t n) {
    std::vector<T> v(n);
    if (fread(v.data(), sizeof(T), n, f) != n) throw Error(XT_ERR_INVALID_ARGUMENT, "short read");
    return v;
}
static int32_t read_i32(FILE* f) { return read_block<int32_t>(f, 1)[0]; }
static std::vector<Vector3f> as_v3(const std::vector<float>& a) {
    std::vector<Vector3f> v(a.size() / 3);
    for (size_t i = 0; i < v.size(); ++i) v[i] = Vector3f(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 6) return 2;
    const char* out = argv[1];
    const int W = atoi(argv[2]), H = atoi(argv[3]), spp = atoi(argv[4]);
    const std::string dir = argv[5];
    try {
        FILE* f = fopen((dir + "/mesh.bin").c_str(), "rb");
        if (!f) return 3;
        const int nv = read_i32(f), ni = read_i32(f);
        auto P = as_v3(read_block<float>(f, 3 * (size_t)nv));
        auto N = as_v3(read_block<float>(f, 3 * (size_t)nv));
        auto uvf = read_block<float>(f, 2 * (size_t)nv);
        auto T = as_v3(read_block<float>(f, 3 * (size_t)nv));
        auto B = as_v3(read_block<float>(f, 3 * (size_t)nv));
        auto idx = read_block<int32_t>(f, (size_t)ni);
        fclose(f);
        std::vector<Vector2f> uv(nv);
        for (int i = 0; i < nv; ++i) uv[i] = Vector2f{uvf[2 * i], uvf[2 * i + 1]};
        auto mesh = std::make_shared<TriangleMesh>();
        mesh->setGeometry(nv, P.data(), uv.data(), N.data(), T.data(), B.data(), ni, idx.data());

        f = fopen((dir + "/tex.bin").c_str(), "rb");
        if (!f) return 3;
        int tw = read_i32(f), th = read_i32(f);
        auto tex = read_block<float>(f, 3 * (size_t)tw * th);
        fclose(f);
        f = fopen((dir + "/sky.bin").c_str(), "rb");
        if (!f) return 3;
        int sw = read_i32(f), sh = read_i32(f);
        auto sky = read_block<float>(f, 3 * (size_t)sw * sh);
        fclose(f);
        f = fopen((dir + "/cam.bin").c_str(), "rb");
        if (!f) return 3;
        auto cam = read_block<float>(f, 18);   // cam_to_world row-major, fov, aspect
        fclose(f);

        auto scene = std::make_shared<Scene>();
        auto base = std::make_shared<Metal>(0.0f, Vector3f(0.95f, 0.9f, 0.85f));
        base->alpha = 0.2f;                                      // GGX alpha given directly
        auto mat = std::make_shared<MicrofacetNormalMapping>(base);
        mat->normalmap = std::make_shared<Texture>(tex.data(), tw, th);
        scene->addObject(std::make_shared<Object>(mesh, mat, Transform()));
        scene->skySphere = std::make_shared<SkySphereLatLong>(sky.data(), sw, sh);
        Transform c;
        for (int i = 0; i < 16; ++i) c.mat[i / 4][i % 4] = cam[i];
        scene->camera = std::make_shared<PinholeCamera>(c, cam[16], cam[17]);
        scene->buildAccelerationStructure();
        StandardPathTracer pt;
        pt.maxDepth = 12;
        pt.flags = XT_FLAG_ENV_NEE;
        SimpleRenderTarget target(W, H);
        target.render(*scene, spp, pt);
        FILE* o = fopen(out, "wb");
        if (!o) return 2;
        fwrite(target.data.data(), sizeof(float), 3 * (size_t)W * H, o);
        fclose(o);
        printf("rendered %dx%d x %d spp -> %s\n", W, H, spp, out);
    } catch (const Error& e) {
        fprintf(stderr, "xitils error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
