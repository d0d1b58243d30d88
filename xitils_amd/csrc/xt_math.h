// Float3 arithmetic with the reference's rounding behaviour, usable from host
// (BVH build, light tables) and device (gfx950 kernels).  Every file that
// includes it is compiled with -ffp-contract=off, so each source-level
// operation is one IEEE rounding on both sides.
//
//   a / s      = a * (1 / s)                       Vector.h:325-335
//   dot        = (x*x' + y*y') + z*z'              Vector.h:628-632 (4-lane
//                reduce_add over a zero pad lane; SSE3 hadd order)
//   cross      = evaluated in float64, rounded     Vector.h:637-654
//   normalize  = v * (1 / sqrt(dot(v, v)))         Vector.h:357-365, 613
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define XT_HD __host__ __device__ __forceinline__
#define XT_LIBM_FN __host__ __device__ __forceinline__
#include "xt_libm.h"

namespace xt {

struct f3 {
    float x, y, z;
};

XT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
XT_HD f3 splat(float v) { return f3{v, v, v}; }
XT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
XT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
XT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
XT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
XT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
XT_HD f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
XT_HD f3 div(f3 a, float s) {
    float inv = 1.0f / s;
    return a * inv;
}
XT_HD f3& operator+=(f3& a, f3 b) {
    a = a + b;
    return a;
}
XT_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
XT_HD float length(f3 a) { return sqrtf(dot(a, a)); }
XT_HD f3 normalize(f3 a) { return div(a, length(a)); }
XT_HD f3 cross(f3 a, f3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return f3{(float)(ay * bz - az * by), (float)(az * bx - ax * bz), (float)(ax * by - ay * bx)};
}
XT_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
XT_HD float clamp_pos(float x) { return x > 0.0f ? x : 0.0f; }
XT_HD f3 clamp_pos(f3 a) { return f3{clamp_pos(a.x), clamp_pos(a.y), clamp_pos(a.z)}; }
XT_HD float safe_sqrt(float x) { return sqrtf(clamp_pos(x)); }
XT_HD f3 face_forward(f3 n, f3 v) { return dot(n, v) < 0.0f ? -n : n; }
// t1*a + t2*b + (1 - t1 - t2)*c  (Utils.h:81-83)
XT_HD f3 lerp3(f3 a, f3 b, f3 c, float t1, float t2) { return t1 * a + t2 * b + (1.0f - t1 - t2) * c; }
XT_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
XT_HD int max_dim(f3 a) { return (a.x > a.y && a.x > a.z) ? 0 : ((a.y > a.z) ? 1 : 2); }
XT_HD bool finite3(f3 a) { return __builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z); }

// Orthonormal frame around v (BasisVectors, Vector.h:781-795).
struct Frame {
    f3 e1, e2, e3;
};
XT_HD Frame make_frame(f3 v) {
    Frame f;
    f.e1 = normalize(v);
    f3 axis = fabsf(v.x) < 0.5f ? mk(1, 0, 0) : (fabsf(v.y) < 0.5f ? mk(0, 1, 0) : mk(0, 0, 1));
    f.e2 = normalize(cross(v, axis));
    f.e3 = normalize(cross(f.e1, f.e2));
    return f;
}
XT_HD f3 to_world(const Frame& f, float a, float b, float c) { return f.e1 * a + f.e2 * b + f.e3 * c; }
XT_HD f3 to_local(const Frame& f, f3 v) { return mk(dot(v, f.e1), dot(v, f.e2), dot(v, f.e3)); }

// Row-major 4x4 affine transforms (Transform.h:62-73, 105-111).
XT_HD f3 xf_point(const float* m, f3 v) {
    float xp = m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3];
    float yp = m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7];
    float zp = m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11];
    float wp = m[12] * v.x + m[13] * v.y + m[14] * v.z + m[15];
    if (wp == 1) return mk(xp, yp, zp);
    return div(mk(xp, yp, zp), wp);
}
XT_HD f3 xf_vector(const float* m, f3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

XT_HD float tri_area(f3 p0, f3 p1, f3 p2) { return length(cross(p1 - p0, p2 - p0)) / 2.0f; }

}  // namespace xt
