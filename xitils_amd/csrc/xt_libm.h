// The transcendental functions of the path (sin, cos, pow(x, 5), atan2,
// acos), evaluated the same way on the device and on the host so that a
// sampled direction has the same bits whichever side computes it.  The
// reference calls the platform's libm (std::sin, std::pow, ...), whose last
// bits differ between libraries; the functions below are deterministic in
// IEEE float / double arithmetic instead (every file that includes this
// header is compiled with -ffp-contract=off, so each operation is one
// rounding, and fused multiply-adds are written as fmaf; division and sqrt
// are the correctly rounded ones on both sides).
//
// XT_LIBM_FN is the qualifier of the functions: xt_math.h sets it to
// __host__ __device__ for the HIP build; a plain C++ compiler gets inline.
#pragma once

#include <math.h>

#ifndef XT_LIBM_FN
#define XT_LIBM_FN inline
#endif

namespace xt {

// sin x and cos x for |x| <= 256.  Cody and Waite's reduction x = k pi/2 + r
// with pi/2 split into three floats whose first two have 16 significant bits
// (k * part is exact for |k| < 2^8, so r carries only the last part's
// rounding), then the Taylor series of sin to r^9 and of cos to r^10 on
// |r| <= pi/4 (truncation below 2.3e-9 relative, under half a float ulp);
// k mod 4 selects signs and the swap.  Within 2 ulp of sin and cos.
XT_LIBM_FN void xt_sincosf(float x, float* s, float* c) {
    const float two_over_pi = 0x1.45f306p-1f;
    const float pio2_a = 0x1.921ep+0f, pio2_b = 0x1.b544p-16f, pio2_c = 0x1.0b4612p-34f;
    const float k = rintf(x * two_over_pi);
    float r = x - k * pio2_a;
    r = r - k * pio2_b;
    r = r - k * pio2_c;
    const float r2 = r * r;
    float ps = 0x1.71de3ap-19f;          // 1/9!
    ps = ps * r2 + -0x1.a01a02p-13f;     // -1/7!
    ps = ps * r2 + 0x1.111112p-7f;       // 1/5!
    ps = ps * r2 + -0x1.555556p-3f;      // -1/3!
    const float sp = r + (r * r2) * ps;
    float pc = -0x1.27e4fcp-22f;         // -1/10!
    pc = pc * r2 + 0x1.a01a02p-16f;      // 1/8!
    pc = pc * r2 + -0x1.6c16c2p-10f;     // -1/6!
    pc = pc * r2 + 0x1.555556p-5f;       // 1/4!
    pc = pc * r2 + -0.5f;                // -1/2!
    const float cp = 1.0f + r2 * pc;
    const float q = k - 4.0f * floorf(k * 0.25f);   // k mod 4, exact (NaN propagates to both results)
    if (q == 1.0f) {
        *s = cp;
        *c = -sp;
    } else if (q == 2.0f) {
        *s = -sp;
        *c = -cp;
    } else if (q == 3.0f) {
        *s = -cp;
        *c = sp;
    } else {
        *s = sp;
        *c = cp;
    }
}
XT_LIBM_FN float xt_sinf(float x) {
    float s, c;
    xt_sincosf(x, &s, &c);
    return s;
}
XT_LIBM_FN float xt_cosf(float x) {
    float s, c;
    xt_sincosf(x, &s, &c);
    return c;
}

// x^5 as std::pow(float, int) computes it (both promoted to double): the
// product in double (x^2 exact, three more roundings) rounded once to float.
XT_LIBM_FN float xt_pow5f(float x) {
    const double d = (double)x;
    return (float)(d * d * d * d * d);
}

// atan(sqrt t) / sqrt t for t in [0, 1]: a degree-8 polynomial in t fitted
// to near-minimax relative error (1.5e-8) by iteratively reweighted least
// squares on [0, 1], coefficients rounded to float, Horner with fma.
XT_LIBM_FN float xt_atan_poly(float t) {
    float p = 0x1.758192p-9f;
    p = fmaf(p, t, -0x1.073e90p-6f);
    p = fmaf(p, t, 0x1.5db3fap-5f);
    p = fmaf(p, t, -0x1.335c98p-4f);
    p = fmaf(p, t, 0x1.b3d812p-4f);
    p = fmaf(p, t, -0x1.22e3a2p-3f);
    p = fmaf(p, t, 0x1.9972d8p-3f);
    p = fmaf(p, t, -0x1.5554a6p-2f);
    return fmaf(p, t, 1.0f);
}
// atan2 in float: a = min(|x|, |y|) / max(|x|, |y|) in [0, 1], atan a = a
// P(a^2), then the octant (pi/2 - r above the diagonal, pi - r for x < 0
// or x = -0, the sign of y; zeros follow IEEE atan2).  Within 2.4 ulp of
// atan2 in float64.
XT_LIBM_FN float xt_atan2f(float y, float x) {
    const float PI = 0x1.921fb6p+1f, PI_2 = 0x1.921fb6p+0f;
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.0f ? mn / mx : 0.0f;
    float r = a * xt_atan_poly(a * a);
    if (ay > ax) r = PI_2 - r;
    if (signbit(x)) r = PI - r;
    return copysignf(r, y);
}
// acos x = atan2(sqrt((1 - x)(1 + x)), x): relative accuracy near x = +-1
// (1 - x and 1 + x are exact there).  Within 2.8 ulp of acos in float64.
XT_LIBM_FN float xt_acosf(float x) { return xt_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x); }

}  // namespace xt
