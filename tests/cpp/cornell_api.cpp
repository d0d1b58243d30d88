// The unit Cornell box built through the C++ host API (include/xitils/xitils.hpp)
// the way a reference sandbox builds its scene (Sandbox/SimplePathTracer/Main.cpp:66-118),
// rendered with StandardPathTracer through the GPU.  Writes the raw float32
// film (H x W x 3) to argv[1]; argv[2], argv[3] = width, height; argv[4] = spp.
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "xitils/xitils.hpp"

using namespace xitils;

static std::shared_ptr<TriangleMesh> quad(Vector3f a, Vector3f b, Vector3f c, Vector3f d) {
    auto m = std::make_shared<TriangleMesh>();
    Vector3f p[4] = {a, b, c, d};
    const int idx[6] = {0, 1, 2, 0, 2, 3};
    m->setGeometry(4, p, nullptr, nullptr, nullptr, nullptr, 6, idx);
    return m;
}

static std::shared_ptr<TriangleMesh> box(float sx, float sy, float sz, float cx, float cy, float cz, float rotDeg) {
    double hx = sx / 2.0, hy = sy / 2.0, hz = sz / 2.0;
    double c[8][3] = {{-hx, -hy, -hz}, {hx, -hy, -hz}, {hx, hy, -hz}, {-hx, hy, -hz},
                      {-hx, -hy, hz},  {hx, -hy, hz},  {hx, hy, hz},  {-hx, hy, hz}};
    double t = rotDeg * 3.14159265358979323846 / 180.0, ct = std::cos(t), st = std::sin(t);
    Vector3f v[8];
    for (int i = 0; i < 8; ++i)
        v[i] = Vector3f((float)(c[i][0] * ct + c[i][2] * st + cx), (float)(c[i][1] + cy), (float)(-c[i][0] * st + c[i][2] * ct + cz));
    const int q[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 7, 6, 2}};
    int idx[36];
    for (int f = 0; f < 6; ++f) {
        int* o = idx + 6 * f;
        o[0] = q[f][0]; o[1] = q[f][1]; o[2] = q[f][2]; o[3] = q[f][0]; o[4] = q[f][2]; o[5] = q[f][3];
    }
    auto m = std::make_shared<TriangleMesh>();
    m->setGeometry(8, v, nullptr, nullptr, nullptr, nullptr, 36, idx);
    return m;
}

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : "cornell_api.f32";
    int W = argc > 2 ? atoi(argv[2]) : 64, H = argc > 3 ? atoi(argv[3]) : 64, spp = argc > 4 ? atoi(argv[4]) : 4;
    try {
        auto scene = std::make_shared<Scene>();
        auto white = std::make_shared<Diffuse>(Vector3f(0.73f));
        auto red = std::make_shared<Diffuse>(Vector3f(0.63f, 0.065f, 0.05f));
        auto green = std::make_shared<Diffuse>(Vector3f(0.14f, 0.45f, 0.09f));
        auto light = std::make_shared<Emission>(Vector3f(17.0f, 12.0f, 4.0f));
        Transform I;
        scene->addObject(std::make_shared<Object>(quad({0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}), white, I));
        scene->addObject(std::make_shared<Object>(quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}), white, I));
        scene->addObject(std::make_shared<Object>(quad({0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}), white, I));
        scene->addObject(std::make_shared<Object>(quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}), red, I));
        scene->addObject(std::make_shared<Object>(quad({1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}), green, I));
        scene->addObject(std::make_shared<Object>(box(0.3f, 0.3f, 0.3f, 0.33f, 0.15f, 0.35f, -18.0f), white, I));
        scene->addObject(std::make_shared<Object>(box(0.3f, 0.6f, 0.3f, 0.66f, 0.3f, 0.66f, 18.0f), white, I));
        scene->addObject(std::make_shared<Object>(
            quad({0.375f, 0.999f, 0.4f}, {0.625f, 0.999f, 0.4f}, {0.625f, 0.999f, 0.6f}, {0.375f, 0.999f, 0.6f}), light, I));
        scene->camera = std::make_shared<PinholeCamera>(translate(0.5f, 0.5f, -1.35f), 40 * ToRad, (float)H / W);
        scene->buildAccelerationStructure();
        StandardPathTracer pt;
        pt.maxDepth = 8;
        SimpleRenderTarget target(W, H);
        target.render(*scene, spp, pt);
        FILE* f = fopen(out, "wb");
        if (!f) return 2;
        fwrite(target.data.data(), sizeof(float), 3 * (size_t)W * H, f);
        fclose(f);
        printf("rendered %dx%d x %d spp -> %s\n", W, H, spp, out);
    } catch (const Error& e) {
        fprintf(stderr, "xitils error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
