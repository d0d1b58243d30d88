// Learned path guiding: an 8-lobe von Mises-Fisher mixture predicted per
// shading point by a 3x64 MLP over frequency-encoded position, shading normal
// and view direction.
//
// vMF lobe (VonMisesFisherDistribution.h:205-211):
//   vMF(w; mu, k) = k / (2 pi (1 - e^{-2k})) * e^{k (mu . w - 1)},  1 / (4 pi) at k = 0
// Sampling follows Jakob, "Numerically stable sampling of the von
// Mises-Fisher distribution on S^2" (the method VonMisesFisherDistribution.h:
// 162-201 cites): W = 1 + log(u + (1 - u) e^{-2k}) / k and a tangent
// direction uniform on the unit CIRCLE.  The reference draws the tangent
// vector inside the unit disk (r = sqrt(u)) and renormalises, which does not
// produce its own pdf; guiding needs pdf and sampler to agree.
//
// Network (all row-major [out][in]):
//   x (96) = 72 frequency features of the position (12 bands, sin and cos
//            per axis, position scaled to [0, 1] by the scene bounds),
//            shading normal (3), view direction wo (3), zeros (18)
//   h1 = relu(W0 x + b0)   W0: 64 x 96
//   h2 = relu(W1 h1 + b1)  W1: 64 x 64
//   h3 = relu(W2 h2 + b2)  W2: 64 x 64
//   y  = W3 h3 + b3        W3: 48 x 64 (rows 40..47 unused, kept zero)
//   lobe j (0..7): mu = normalize(y[5j..5j+2]), kappa = exp(clamp(y[5j+3], -3, 7)),
//                  alpha = softmax(y[5j+4])
#pragma once

#include "xt_math.h"

namespace xt {

constexpr int G_IN = 96, G_HID = 64, G_OUT = 40, G_OUT_PAD = 48, G_LOBES = 8, G_BANDS = 12;
// flat parameter layout (floats)
constexpr int G_OFF_W0 = 0;
constexpr int G_OFF_B0 = G_OFF_W0 + G_HID * G_IN;
constexpr int G_OFF_W1 = G_OFF_B0 + G_HID;
constexpr int G_OFF_B1 = G_OFF_W1 + G_HID * G_HID;
constexpr int G_OFF_W2 = G_OFF_B1 + G_HID;
constexpr int G_OFF_B2 = G_OFF_W2 + G_HID * G_HID;
constexpr int G_OFF_W3 = G_OFF_B2 + G_HID;
constexpr int G_OFF_B3 = G_OFF_W3 + G_OUT_PAD * G_HID;
constexpr int G_NPARAMS = G_OFF_B3 + G_OUT_PAD;   // 17,648
#ifndef XT_GUIDE_KLOG_MAX
#define XT_GUIDE_KLOG_MAX 9.0f
#endif
// log kappa range of a lobe: kappa up to e^9 ~ 8100, a lobe about 0.6
// degrees wide (1 / sqrt(kappa) rad), narrower than a 5 cm gap seen from 3 m
constexpr float G_KLOG_MIN = -3.0f, G_KLOG_MAX = XT_GUIDE_KLOG_MAX;

struct GuideBounds {
    f3 lo;
    f3 inv_ext;   // 1 / (hi - lo) per axis
};

// 96 input features of one query into x[0..95] (any element type); features
// 78..95 are zero.
template <class T>
XT_HD void guide_encode(const GuideBounds& gb, f3 p, f3 n, f3 wo, T* x) {
    const float PI = 3.14159265358979323846f;
    float u[3] = {(p.x - gb.lo.x) * gb.inv_ext.x, (p.y - gb.lo.y) * gb.inv_ext.y, (p.z - gb.lo.z) * gb.inv_ext.z};
    for (int a = 0; a < 3; ++a) {
        float ua = fminf(fmaxf(u[a], 0.0f), 1.0f);
        float s, c;
        xt_sincosf(PI * ua, &s, &c);
        for (int k = 0; k < G_BANDS; ++k) {
            x[a * 2 * G_BANDS + 2 * k] = (T)s;
            x[a * 2 * G_BANDS + 2 * k + 1] = (T)c;
            const float d = s + s;
            const float c2 = fmaf(-d, s, 1.0f);   // cos 2x = 1 - 2 sin^2 x, one rounding
            s = d * c;                            // sin 2x
            c = c2;
        }
    }
    x[72] = (T)n.x; x[73] = (T)n.y; x[74] = (T)n.z;
    x[75] = (T)wo.x; x[76] = (T)wo.y; x[77] = (T)wo.z;
    for (int k = 78; k < G_IN; ++k) x[k] = (T)0.0f;
}

struct VmfLobe {
    f3 mu;
    float kappa, alpha;
};

XT_HD float vmf_norm(float k) {   // k / (2 pi (1 - e^{-2k}))
    const float PI = 3.14159265358979323846f;
    return k / (2.0f * PI * (1.0f - expf(-2.0f * k)));
}

XT_HD float vmf_lobe_pdf(const VmfLobe& l, f3 w) {
    const float PI = 3.14159265358979323846f;
    if (l.kappa == 0.0f) return 1.0f / (4.0f * PI);
    return vmf_norm(l.kappa) * expf(l.kappa * (dot(w, l.mu) - 1.0f));
}

// Raw network outputs (40) -> lobes.
XT_HD void guide_decode(const float* y, VmfLobe* lobes) {
    float amax = -INFINITY;
    for (int j = 0; j < G_LOBES; ++j) amax = fmaxf(amax, y[5 * j + 4]);
    float asum = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) {
        f3 m = mk(y[5 * j], y[5 * j + 1], y[5 * j + 2]);
        float len = length(m);
        lobes[j].mu = len > 1e-12f ? div(m, len) : mk(0.0f, 0.0f, 1.0f);
        float kl = fminf(fmaxf(y[5 * j + 3], G_KLOG_MIN), G_KLOG_MAX);
        lobes[j].kappa = expf(kl);
        float e = expf(y[5 * j + 4] - amax);
        lobes[j].alpha = e;
        asum += e;
    }
    for (int j = 0; j < G_LOBES; ++j) lobes[j].alpha = lobes[j].alpha / asum;
}

XT_HD float guide_pdf(const VmfLobe* lobes, f3 w) {
    float p = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) p += lobes[j].alpha * vmf_lobe_pdf(lobes[j], w);
    return p;
}

// Jakob's sampler for one lobe; u1 -> W, u2 -> tangent angle.
XT_HD f3 vmf_sample(const VmfLobe& l, float u1, float u2) {
    const float PI = 3.14159265358979323846f;
    float W;
    if (l.kappa > 0.0f) {
        W = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * l.kappa)) / l.kappa;
    } else {
        W = 2.0f * u1 - 1.0f;
    }
    W = fminf(fmaxf(W, -1.0f), 1.0f);
    float phi = u2 * (2.0f * PI);
    float sq = safe_sqrt(1.0f - W * W);
    Frame fr = make_frame(l.mu);
    return normalize(to_world(fr, W, sq * cosf(phi), sq * sinf(phi)));
}

// Lobe by alpha with one uniform, then a direction.
XT_HD f3 guide_sample(const VmfLobe* lobes, float usel, float u1, float u2) {
    int j = 0;
    float acc = 0.0f;
    for (; j < G_LOBES - 1; ++j) {
        acc += lobes[j].alpha;
        if (usel < acc) break;
    }
    return vmf_sample(lobes[j], u1, u2);
}

// d(-log q(w)) / d(raw outputs), times the sample weight; store(o, value)
// receives the 40 derivatives.
template <class Store>
XT_HD void guide_loss_grad(const float* y, f3 w, float weight, Store store) {
    VmfLobe lobes[G_LOBES];
    guide_decode(y, lobes);
    float v[G_LOBES], q = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) {
        v[j] = vmf_lobe_pdf(lobes[j], w);
        q += lobes[j].alpha * v[j];
    }
    q = fmaxf(q, 1e-30f);
    for (int j = 0; j < G_LOBES; ++j) {
        float r = lobes[j].alpha * v[j] / q;          // responsibility
        float k = lobes[j].kappa;
        f3 m = mk(y[5 * j], y[5 * j + 1], y[5 * j + 2]);
        float len = fmaxf(length(m), 1e-12f);
        f3 mu = lobes[j].mu;
        float mw = dot(mu, w);
        // d log v / d m = k (w - (mu.w) mu) / |m|
        f3 gm = (w - mw * mu) * (r * k / len);
        // d log v / d k = 1/k - 2 / (e^{2k} - 1) + (mu.w - 1); times k for k = exp(raw)
        float dk = 1.0f / k - 2.0f / expm1f(2.0f * k) + (mw - 1.0f);
        float kl = y[5 * j + 3];
        float gk = (kl > G_KLOG_MIN && kl < G_KLOG_MAX) ? r * dk * k : 0.0f;
        float ga = r - lobes[j].alpha;
        store(5 * j, -weight * gm.x);
        store(5 * j + 1, -weight * gm.y);
        store(5 * j + 2, -weight * gm.z);
        store(5 * j + 3, -weight * gk);
        store(5 * j + 4, -weight * ga);
    }
}

// Mixture pdf from stored lobes: l4(j) = (mu, kappa), l2(j) = (alpha,
// alpha * vmf_norm(kappa)).
template <class L4, class L2>
XT_HD float guide_pdf_lobes(L4 l4, L2 l2, f3 w) {
    float p = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) {
        const float4 a = l4(j);
        p += l2(j).y * expf(a.w * (dot(w, mk(a.x, a.y, a.z)) - 1.0f));
    }
    return p;
}

// Lobe by alpha with usel, then Jakob's sampler with (u1, u2).
template <class L4, class L2>
XT_HD f3 guide_sample_lobes(L4 l4, L2 l2, float usel, float u1, float u2) {
    int j = 0;
    float acc = 0.0f;
    for (; j < G_LOBES - 1; ++j) {
        acc += l2(j).x;
        if (usel < acc) break;
    }
    const float4 a = l4(j);
    VmfLobe l;
    l.mu = mk(a.x, a.y, a.z);
    l.kappa = a.w;
    l.alpha = 0.0f;
    return vmf_sample(l, u1, u2);
}

}  // namespace xt
