// ORACLE — test infrastructure only.
//
// CPU restatement of VonMisesFisherDistribution<N> (VonMisesFisherDistribution.h:
// 9-217): the mixture pdf, the reference's sampler, the EM fit
// approximateBySEM and the 5N-1 float encoding.  The Sampler's randf() stream
// (Sampler.h:17-19) is replaced by caller-supplied uniforms, consumed in the
// reference's order.
#pragma once

#include "xo_core.hpp"

namespace xo {

struct Vmf {
    int N = 1;
    std::vector<V3> mu;
    std::vector<float> kappa, alpha;
    // VonMisesFisherDistribution.h:24-35
    explicit Vmf(int n) : N(n), mu(n, V3(1, 0, 0)), kappa(n, 0.0f), alpha(n, 0.0f) { alpha[0] = 1.0f; }

    // :37-45 -- 5 floats per lobe, the last lobe's alpha omitted
    void encode(float* d) const {
        for (int k = 0; k < N; ++k) {
            d[k * 5 + 0] = mu[k].x;
            d[k * 5 + 1] = mu[k].y;
            d[k * 5 + 2] = mu[k].z;
            d[k * 5 + 3] = kappa[k];
            if (k < N - 1) d[k * 5 + 4] = alpha[k];
        }
    }
    // :47-58 -- the last alpha is 1 minus the running sum
    void decode(const float* d) {
        float acc = 0.0f;
        for (int k = 0; k < N; ++k) {
            mu[k] = V3(d[k * 5 + 0], d[k * 5 + 1], d[k * 5 + 2]);
            kappa[k] = d[k * 5 + 3];
            alpha[k] = (k < N - 1) ? d[k * 5 + 4] : 1.0f - acc;
            acc += alpha[k];
        }
    }

    // :205-211
    static float lobe(const V3& m, float k, const V3& w) {
        if (k == 0.0f) return 1.0f / (4 * Pi);
        return k / (2 * Pi * (1.0f - expf(-2 * k))) * expf(k * (dot(w, m) - 1.0f));
    }
    // :60-66
    float eval(const V3& w) const {
        float v = 0.0f;
        for (int k = 0; k < N; ++k) v += alpha[k] * lobe(mu[k], kappa[k], w);
        return v;
    }

    // :162-201 with the uniforms u[0..3] in the reference's order: lobe choice
    // (Sampler.h:52-61, u <= running sum over all but the last weight), disk
    // radius, disk angle, W.
    V3 sample(const float* u) const {
        float accum = 0.0f;
        int n = 0;
        for (; n < N - 1; ++n) {
            accum += alpha[n];
            if (u[0] <= accum) break;
        }
        const V3& m = mu[n];
        const float k = kappa[n];
        if (m.isZero()) return V3(0, 0, 0);
        float r = sqrtf(u[1]);
        float theta = 2.0f * Pi * u[2] - Pi;
        float vx = r * cosf(theta), vy = r * sinf(theta);
        float W;
        if (k > 0.0f) W = 1.0f + powf(k, -1.0f) * logf(u[3] + (1.0f - u[3]) * expf(-2.0f * k));
        else W = 1.0f;
        float s = safeSqrt(1.0f - powf(W, 2.0f));
        Basis b(normalize(m));
        return b.toGlobal(W, s * vx, s * vy).normalize();
    }

    // :68-160 (Q15: responsibilities without alpha, signed kappa change in the
    // convergence test, one sample -> kappa = 1000 copies).  init_u holds 2N
    // uniforms (u, v per lobe, :89-98).  Returns the number of EM steps run.
    // fixed_steps > 0 runs exactly that many steps without the convergence
    // test (for comparing trajectories step for step).
    static Vmf fit(int N, const std::vector<V3>& samples, const float* init_u, int* steps_out, int fixed_steps = 0) {
        const float KappaLimit = 1000;
        const int StepMax = fixed_steps > 0 ? fixed_steps : 100000;
        const float DiffThreshold = 1e-6f;
        Vmf vmf(N);
        if (steps_out) *steps_out = 0;
        if (samples.empty()) return vmf;            // the reference divides by zero here
        if (samples.size() == 1) {
            for (int i = 0; i < N; ++i) {
                vmf.mu[i] = samples[0];
                vmf.kappa[i] = KappaLimit;
                vmf.alpha[i] = i == 0 ? 1.0f : 0.0f;
            }
            return vmf;
        }
        for (int j = 0; j < N; ++j) {
            float u = init_u[2 * j], v = init_u[2 * j + 1];
            float theta = asinf(u);
            float phi = 2.0f * Pi * v - Pi;
            vmf.mu[j] = V3(cosf(theta) * cosf(phi), cosf(theta) * sinf(phi), sinf(theta));
            vmf.kappa[j] = 0.5f;
            vmf.alpha[j] = 1.0f / N;
        }
        const size_t ns = samples.size();
        std::vector<float> z(ns * N);
        Vmf last(N);
        int step = 0;
        for (step = 0; step < StepMax; ++step) {
            for (size_t i = 0; i < ns; ++i) {       // E-step :109-120
                float gsum = 0.0f;
                for (int j = 0; j < N; ++j) gsum += lobe(vmf.mu[j], vmf.kappa[j], samples[i]);
                for (int j = 0; j < N; ++j)
                    z[i * N + j] = gsum > 0.0f ? lobe(vmf.mu[j], vmf.kappa[j], samples[i]) / gsum : 0.0f;
            }
            for (int j = 0; j < N; ++j) {           // M-step :123-139
                float a = 0.0f;
                V3 r(0.0f, 0.0f, 0.0f);
                for (size_t i = 0; i < ns; ++i) {
                    a += z[i * N + j];
                    r += z[i * N + j] * samples[i];
                }
                r /= a;
                a /= (float)ns;
                vmf.alpha[j] = a;
                float rl = r.length();
                vmf.kappa[j] = (3 * rl - powf(rl, 3.0f)) / (1.0f - r.lengthSq());
                vmf.mu[j] = normalize(r);
                if (vmf.kappa[j] > KappaLimit) vmf.kappa[j] = KappaLimit;
                if (vmf.kappa[j] < 0.0f) vmf.kappa[j] = 0.0f;
            }
            if (step > 0 && fixed_steps <= 0) {     // :141-148
                float diff = 0.0f;
                for (int j = 0; j < N; ++j) {
                    diff += (vmf.mu[j] - last.mu[j]).lengthSq();
                    diff += (vmf.kappa[j] - last.kappa[j]) * 0.1f;
                }
                if (diff < DiffThreshold) break;
            }
            last = vmf;
        }
        if (steps_out) *steps_out = step < StepMax ? step + 1 : StepMax;
        return vmf;
    }
};

}  // namespace xo
