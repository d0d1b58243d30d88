// Sanitizer driver (host code only, g++ -fsanitize=...): exercises the CPU
// oracle (oracle/xo_capi.cpp) and the instrumented host walk of the product's
// BVH (tests/cpp/trav_count.cpp, which links the product's builder bvh.cpp)
// on small scenes: an instanced box in a room with two emitters (one
// instanced), a sun sky, renders of both estimators with and without bounce
// caps, closest/any ray queries against brute force, light samples and a vMF
// fit.  Exits non-zero on a result mismatch; the sanitizers abort on a memory
// or undefined-behaviour error.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
void* orc_scene_new(void);
void orc_scene_free(void* s);
int orc_scene_add_material(void* s, uint32_t type, const float* p7);
int orc_scene_add_object(void* s, const float* positions, const float* normals, int num_vertices,
                         const int32_t* indices, int num_triangles, const float* xf16, const float* xfinv16,
                         uint32_t material, const uint32_t* tri_materials);
void orc_scene_set_env(void* s, int has_sky, const float* sky_rgb, const float* sun_dir, float sun_omc, float sun_sin2,
                       const float* sun_rgb);
int orc_scene_build(void* s);
void orc_trace_closest(void* s, const float* rays, int n, float* hits, int mode, int nthreads);
void orc_trace_any(void* s, const float* rays, int n, uint8_t* occluded, int mode, int nthreads);
void orc_render(void* s, const float* cam, const int32_t* ip, const float* fp, float* film32, double* film64,
                uint64_t* stats, int nthreads);
void orc_sample_light(void* s, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t slot, int uniform_bary,
                      float* out12);
void orc_vmf_fit(int N, const float* samples, const uint32_t* offsets, int num_fits, const float* init_u, float* enc,
                 int32_t* steps, const int* fixed_steps, int nthreads);
void* tc_build(const float* tris9, uint32_t n, uint32_t* info);
void tc_free(void* h);
void tc_trace(void* h, const float* rays, uint32_t n, int any, float* hits, uint8_t* occluded, uint64_t* counts,
              uint32_t* per_ray);
}

static void quad(std::vector<float>& P, std::vector<int32_t>& I, const float q[4][3]) {
    const int b = (int)P.size() / 3;
    for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) P.push_back(q[k][a]);
    const int t[6] = {0, 1, 2, 0, 2, 3};
    for (int k = 0; k < 6; ++k) I.push_back(b + t[k]);
}

int main() {
    void* s = orc_scene_new();
    const float white[7] = {0.7f, 0.7f, 0.7f, 0, 0, 0, 0}, metal[7] = {0.09f, 0.9f, 0.8f, 0.6f, 0, 0, 0};
    const float le1[7] = {10, 9, 8, 0, 0, 0, 0}, le2[7] = {3, 4, 6, 0, 0, 0, 0};
    const int mw = orc_scene_add_material(s, 0, white), mm = orc_scene_add_material(s, 4, metal);
    const int ml1 = orc_scene_add_material(s, 5, le1), ml2 = orc_scene_add_material(s, 5, le2);
    // room: floor, ceiling, back, left, right (world space, identity)
    std::vector<float> P;
    std::vector<int32_t> I;
    const float room[5][4][3] = {{{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}}, {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},
                                 {{0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}}, {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}},
                                 {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}}};
    for (auto& q : room) quad(P, I, q);
    orc_scene_add_object(s, P.data(), nullptr, (int)P.size() / 3, I.data(), (int)I.size() / 3, nullptr, nullptr, mw,
                         nullptr);
    std::vector<float> world = P;   // world-space triangles for the host walk (identity objects only)
    std::vector<int32_t> worldI = I;
    // ceiling light (identity)
    std::vector<float> L;
    std::vector<int32_t> LI;
    const float lq[4][3] = {{0.4f, 0.999f, 0.4f}, {0.6f, 0.999f, 0.4f}, {0.6f, 0.999f, 0.6f}, {0.4f, 0.999f, 0.6f}};
    quad(L, LI, lq);
    orc_scene_add_object(s, L.data(), nullptr, 4, LI.data(), 2, nullptr, nullptr, ml1, nullptr);
    // unit box instanced: scale 0.3, rotation 30 deg about y, translation
    std::vector<float> B;
    std::vector<int32_t> BI;
    const float c[8][3] = {{-.5f, -.5f, -.5f}, {.5f, -.5f, -.5f}, {.5f, .5f, -.5f}, {-.5f, .5f, -.5f},
                           {-.5f, -.5f, .5f},  {.5f, -.5f, .5f},  {.5f, .5f, .5f},  {-.5f, .5f, .5f}};
    const int faces[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 7, 6, 2}};
    for (auto& f : faces) {
        float q[4][3];
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) q[k][a] = c[f[k]][a];
        quad(B, BI, q);
    }
    const double th = 0.5235987755982988, sc = 0.3;
    const double cs = std::cos(th), sn = std::sin(th);
    float M[16] = {(float)(sc * cs), 0, (float)(sc * sn), 0.4f, 0, (float)sc, 0, 0.16f, (float)(-sc * sn), 0,
                   (float)(sc * cs), 0.5f, 0, 0, 0, 1};
    float Mi[16] = {(float)(cs / sc), 0, (float)(-sn / sc), 0, 0, (float)(1 / sc), 0, 0, (float)(sn / sc), 0,
                    (float)(cs / sc), 0, 0, 0, 0, 1};
    for (int r = 0; r < 3; ++r)
        Mi[4 * r + 3] = -(Mi[4 * r] * M[3] + Mi[4 * r + 1] * M[7] + Mi[4 * r + 2] * M[11]);
    orc_scene_add_object(s, B.data(), nullptr, 24, BI.data(), 12, M, Mi, mm, nullptr);
    // second emitter: the light quad instanced at half size lower down
    float M2[16] = {0.5f, 0, 0, 0.1f, 0, 0.5f, 0, 0.3f, 0, 0, 0.5f, 0.2f, 0, 0, 0, 1};
    float M2i[16] = {2, 0, 0, -0.2f, 0, 2, 0, -0.6f, 0, 0, 2, -0.4f, 0, 0, 0, 1};
    orc_scene_add_object(s, L.data(), nullptr, 4, LI.data(), 2, M2, M2i, ml2, nullptr);
    const float sky[3] = {0.2f, 0.3f, 0.4f}, sun[3] = {0.6f, 0.8f, 0.0f}, sunL[3] = {100, 100, 100};
    orc_scene_set_env(s, 1, sky, sun, 1e-3f, 2e-3f, sunL);
    orc_scene_build(s);
    // renders: both estimators, env NEE and uniform barycentrics on and off, capped and roulette-only
    const float cam[18] = {1, 0, 0, 0.5f, 0, 1, 0, 0.5f, 0, 0, 1, -1.3f, 0, 0, 0, 1, 0.36f, 0.36f};
    const float fp[3] = {0.9f, 1e-5f, 1e-4f};
    for (int integ = 0; integ < 2; ++integ)
        for (int flags = 0; flags < 4; ++flags)
            for (int depth : {0, 3}) {
                const int32_t ip[10] = {24, 20, 3, 0, depth, 7, 0, integ, flags, 5};
                std::vector<float> f32(24 * 20 * 3, 0.0f);
                std::vector<double> f64(24 * 20 * 3, 0.0);
                uint64_t st[2];
                orc_render(s, cam, ip, fp, f32.data(), f64.data(), st, 2);
                for (size_t k = 0; k < f32.size(); ++k)
                    if (!std::isfinite(f32[k]) || f32[k] < 0) {
                        printf("FAIL render value %g\n", f32[k]);
                        return 1;
                    }
            }
    // ray queries: reference traversal vs brute force
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(0.02f, 0.98f), N(-1, 1);
    const int n = 3000;
    std::vector<float> rays(8 * n);
    for (int i = 0; i < n; ++i) {
        float d[3] = {N(rng), N(rng), N(rng)};
        float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) {
            rays[8 * i + a] = U(rng);
            rays[8 * i + 4 + a] = d[a] / l;
        }
        rays[8 * i + 3] = INFINITY;
        rays[8 * i + 7] = 0;
    }
    std::vector<float> h0(4 * n), h1(4 * n);
    std::vector<uint8_t> o0(n), o1(n);
    orc_trace_closest(s, rays.data(), n, h0.data(), 0, 2);
    orc_trace_closest(s, rays.data(), n, h1.data(), 1, 2);
    orc_trace_any(s, rays.data(), n, o0.data(), 0, 2);
    orc_trace_any(s, rays.data(), n, o1.data(), 1, 2);
    int same = 0;
    for (int i = 0; i < n; ++i) same += std::memcmp(&h0[4 * i + 1], &h1[4 * i + 1], 4) == 0;
    if (same < n * 99 / 100 || std::memcmp(o0.data(), o1.data(), n) != 0) {
        printf("FAIL oracle traversal vs brute force (%d of %d ids equal)\n", same, n);
        return 1;
    }
    float ls[12];
    for (uint32_t k = 0; k < 64; ++k) orc_sample_light(s, 3, k, 0, 1, (int)(k & 1), ls);
    // vMF fit of two clusters
    std::vector<float> smp;
    for (int k = 0; k < 200; ++k) {
        float z = k < 100 ? 1.0f : -1.0f;
        float x = 0.1f * N(rng), y = 0.1f * N(rng);
        float l = std::sqrt(x * x + y * y + z * z);
        smp.push_back(x / l);
        smp.push_back(y / l);
        smp.push_back(z / l);
    }
    const uint32_t off[3] = {0, 100, 200};
    std::vector<float> iu(2 * 2 * 2), enc(2 * 9);
    for (float& v : iu) v = U(rng);
    std::vector<int32_t> steps(2);
    orc_vmf_fit(2, smp.data(), off, 2, iu.data(), enc.data(), steps.data(), nullptr, 2);
    orc_scene_free(s);
    // host walk of the product's BVH over the identity geometry
    std::vector<float> tris9;
    for (size_t t = 0; t < worldI.size() / 3; ++t)
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) tris9.push_back(world[3 * worldI[3 * t + v] + a]);
    uint32_t info[3];
    void* h = tc_build(tris9.data(), (uint32_t)(tris9.size() / 9), info);
    std::vector<float> hh(4 * n);
    std::vector<uint8_t> oc(n);
    std::vector<uint32_t> pr(2 * n);
    uint64_t cnt[2];
    tc_trace(h, rays.data(), n, 0, hh.data(), nullptr, cnt, pr.data());
    tc_trace(h, rays.data(), n, 1, nullptr, oc.data(), cnt, nullptr);
    tc_free(h);
    printf("OK\n");
    return 0;
}
