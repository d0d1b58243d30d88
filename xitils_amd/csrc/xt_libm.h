// The transcendental functions of the path (sin, cos, pow(x, 5), atan2,
// acos), evaluated the same way on the device and on the host so that a
// sampled direction has the same bits whichever side computes it.  The
// reference calls the platform's libm (std::sin, std::pow, ...), whose last
// bits differ between libraries; the functions below are deterministic in
// IEEE float / double arithmetic instead (every file that includes this
// header is compiled with -ffp-contract=off, so each operation is one
// rounding).
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

// atan2 and acos evaluated in double and rounded to float: double libraries
// within an ulp of each other round to the same float except where the value
// lies within that ulp of a float rounding boundary.
XT_LIBM_FN float xt_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
XT_LIBM_FN float xt_acosf(float x) { return (float)acos((double)x); }

}  // namespace xt
