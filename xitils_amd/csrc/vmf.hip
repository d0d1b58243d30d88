// VonMisesFisherDistribution<N> (VonMisesFisherDistribution.h:9-217) on the
// device: mixture pdf, the reference's sampler, and the EM fit
// approximateBySEM, batched over many independent distributions.  Mixtures
// travel in the reference's encoding (encode/decode, :37-58): 5 floats per
// lobe (mu.xyz, kappa, alpha), the last lobe's alpha omitted and recovered as
// 1 minus the running sum.
//
// The fit runs one wave64 per sample set: lanes stride over the samples, each
// EM step reduces the per-lobe sums (sum z, sum z x) with a butterfly of
// commutative adds, so every lane holds identical totals and takes the same
// branch at the convergence test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "runtime.h"
#include "xt_guide.h"

using namespace xt;

namespace {

constexpr int VMF_MAX_LOBES = 16;
constexpr float VMF_PI = 3.14159265358979323846f;

__device__ __forceinline__ void vmf_decode_lobe(const float* enc, int N, int k, float acc, f3* mu, float* kappa,
                                                float* alpha) {
    *mu = mk(enc[5 * k], enc[5 * k + 1], enc[5 * k + 2]);
    *kappa = enc[5 * k + 3];
    *alpha = (k < N - 1) ? enc[5 * k + 4] : 1.0f - acc;
}

// eval (:60-66): sum_k alpha_k vMF_k(w), alphas decoded as in :47-58.
__global__ void k_vmf_eval(int N, const float* __restrict__ enc, uint32_t num_dists, const uint32_t* __restrict__ dist,
                           const float* __restrict__ dirs, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = dist ? dist[i] : 0u;
    if (d >= num_dists) {
        out[i] = NAN;
        return;
    }
    const float* e = enc + (size_t)d * (5 * N - 1);
    const f3 w = mk(dirs[3 * (size_t)i], dirs[3 * (size_t)i + 1], dirs[3 * (size_t)i + 2]);
    float acc = 0.0f, v = 0.0f;
    for (int k = 0; k < N; ++k) {
        VmfLobe lb;
        vmf_decode_lobe(e, N, k, acc, &lb.mu, &lb.kappa, &lb.alpha);
        acc += lb.alpha;
        v += lb.alpha * vmf_lobe_pdf(lb, w);
    }
    out[i] = v;
}

// sample (:162-201) with uniforms (lobe choice, disk radius, disk angle, W).
__global__ void k_vmf_sample(int N, const float* __restrict__ enc, uint32_t num_dists, const uint32_t* __restrict__ dist,
                             const float* __restrict__ uni, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = dist ? dist[i] : 0u;
    float* o = out + 3 * (size_t)i;
    if (d >= num_dists) {
        o[0] = o[1] = o[2] = NAN;
        return;
    }
    const float* e = enc + (size_t)d * (5 * N - 1);
    const float* u = uni + 4 * (size_t)i;
    // Sampler::randiAlongNormalizedWeights (Sampler.h:52-61): u <= running sum
    float acc = 0.0f, accum = 0.0f;
    f3 mu = mk(0, 0, 0);
    float kappa = 0.0f, alpha = 0.0f;
    int sel = N - 1;
    for (int k = 0; k < N; ++k) {
        vmf_decode_lobe(e, N, k, acc, &mu, &kappa, &alpha);
        acc += alpha;
        if (k == N - 1) break;
        accum += alpha;
        if (u[0] <= accum) {
            sel = k;
            break;
        }
    }
    (void)sel;
    if (mu.x == 0.0f && mu.y == 0.0f && mu.z == 0.0f) {
        o[0] = o[1] = o[2] = 0.0f;
        return;
    }
    const float r = sqrtf(u[1]);
    const float theta = 2.0f * VMF_PI * u[2] - VMF_PI;
    const float vx = r * cosf(theta), vy = r * sinf(theta);
    float W = 1.0f;
    if (kappa > 0.0f) W = 1.0f + powf(kappa, -1.0f) * logf(u[3] + (1.0f - u[3]) * expf(-2.0f * kappa));
    const float s = safe_sqrt(1.0f - powf(W, 2.0f));
    const Frame fr = make_frame(normalize(mu));
    const f3 g = normalize(to_world(fr, W, s * vx, s * vy));
    o[0] = g.x;
    o[1] = g.y;
    o[2] = g.z;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// approximateBySEM (:68-160) for sample sets samples[off[f] .. off[f+1]).
template <int N>
__global__ __launch_bounds__(256) void k_vmf_fit(const float* __restrict__ smp, const uint32_t* __restrict__ off,
                                                 uint32_t nfits, const float* __restrict__ init_u,
                                                 float* __restrict__ enc, int32_t* __restrict__ steps_out) {
    constexpr float KappaLimit = 1000.0f;
    constexpr int StepMax = 100000;
    constexpr float DiffThreshold = 1e-6f;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    const int l = __lane_id();
    for (uint32_t f = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; f < nfits; f += nw) {
        const uint32_t b = off[f], e = off[f + 1];
        const uint32_t n = e > b ? e - b : 0u;
        f3 mu[N];
        float ka[N], al[N];
        int steps = 0;
        _Pragma("unroll") for (int j = 0; j < N; ++j) {          // default distribution (:24-35)
            mu[j] = mk(1, 0, 0);
            ka[j] = 0.0f;
            al[j] = j == 0 ? 1.0f : 0.0f;
        }
        if (n == 1) {                          // :79-86
            const f3 x = mk(smp[3 * (size_t)b], smp[3 * (size_t)b + 1], smp[3 * (size_t)b + 2]);
            _Pragma("unroll") for (int j = 0; j < N; ++j) {
                mu[j] = x;
                ka[j] = KappaLimit;
            }
        } else if (n > 1) {
            const float* iu = init_u + (size_t)2 * N * f;
            _Pragma("unroll") for (int j = 0; j < N; ++j) {      // :89-98
                const float theta = asinf(iu[2 * j]);
                const float phi = 2.0f * VMF_PI * iu[2 * j + 1] - VMF_PI;
                mu[j] = mk(cosf(theta) * cosf(phi), cosf(theta) * sinf(phi), sinf(theta));
                ka[j] = 0.5f;
                al[j] = 1.0f / N;
            }
            f3 lmu[N];
            float lka[N];
            int step = 0;
            for (step = 0; step < StepMax; ++step) {
                float sa[N];
                f3 sr[N];
                _Pragma("unroll") for (int j = 0; j < N; ++j) {
                    sa[j] = 0.0f;
                    sr[j] = mk(0, 0, 0);
                }
                for (uint32_t i = b + l; i < e; i += 64) {   // E-step (:109-120), responsibilities without alpha
                    const f3 x = mk(smp[3 * (size_t)i], smp[3 * (size_t)i + 1], smp[3 * (size_t)i + 2]);
                    float v[N];
                    float gsum = 0.0f;
                    _Pragma("unroll") for (int j = 0; j < N; ++j) {
                        VmfLobe lb{mu[j], ka[j], 0.0f};
                        v[j] = vmf_lobe_pdf(lb, x);
                        gsum += v[j];
                    }
                    _Pragma("unroll") for (int j = 0; j < N; ++j) {
                        const float z = gsum > 0.0f ? v[j] / gsum : 0.0f;
                        sa[j] += z;
                        sr[j] += z * x;
                    }
                }
                _Pragma("unroll") for (int j = 0; j < N; ++j) {          // M-step (:123-139)
                    const float a = wave_sum(sa[j]);
                    f3 r = mk(wave_sum(sr[j].x), wave_sum(sr[j].y), wave_sum(sr[j].z));
                    r = div(r, a);
                    al[j] = a / (float)n;
                    const float rl = length(r);
                    float k = (3 * rl - powf(rl, 3.0f)) / (1.0f - dot(r, r));
                    mu[j] = normalize(r);
                    if (k > KappaLimit) k = KappaLimit;
                    if (k < 0.0f) k = 0.0f;
                    ka[j] = k;
                }
                if (step > 0) {                        // :141-148
                    float diff = 0.0f;
                    _Pragma("unroll") for (int j = 0; j < N; ++j) {
                        const f3 dm = mu[j] - lmu[j];
                        diff += dot(dm, dm);
                        diff += (ka[j] - lka[j]) * 0.1f;
                    }
                    if (diff < DiffThreshold) break;
                }
                _Pragma("unroll") for (int j = 0; j < N; ++j) {
                    lmu[j] = mu[j];
                    lka[j] = ka[j];
                }
            }
            steps = step < StepMax ? step + 1 : StepMax;
        }
        if (l == 0) {                                  // encode (:37-45)
            float* o = enc + (size_t)(5 * N - 1) * f;
            _Pragma("unroll") for (int j = 0; j < N; ++j) {
                o[5 * j] = mu[j].x;
                o[5 * j + 1] = mu[j].y;
                o[5 * j + 2] = mu[j].z;
                o[5 * j + 3] = ka[j];
                if (j < N - 1) o[5 * j + 4] = al[j];
            }
            if (steps_out) steps_out[f] = steps;
        }
    }
}

template <int N>
void launch_fit(uint32_t grid, hipStream_t st, const float* smp, const uint32_t* off, uint32_t nfits, const float* iu,
                float* enc, int32_t* steps) {
    hipLaunchKernelGGL(k_vmf_fit<N>, dim3(grid), dim3(256), 0, st, smp, off, nfits, iu, enc, steps);
}

using FitLauncher = void (*)(uint32_t, hipStream_t, const float*, const uint32_t*, uint32_t, const float*, float*,
                             int32_t*);
const FitLauncher kFit[VMF_MAX_LOBES + 1] = {nullptr,         launch_fit<1>,  launch_fit<2>,  launch_fit<3>,
                                             launch_fit<4>,   launch_fit<5>,  launch_fit<6>,  launch_fit<7>,
                                             launch_fit<8>,   launch_fit<9>,  launch_fit<10>, launch_fit<11>,
                                             launch_fit<12>,  launch_fit<13>, launch_fit<14>, launch_fit<15>,
                                             launch_fit<16>};

// Device copies of host arrays (or the caller's device pointers).
struct Staged {
    std::vector<void*> owned;
    ~Staged() {
        for (void* p : owned) hipFree(p);
    }
    template <class T>
    hipError_t in(const T* src, size_t count, bool on_device, const T** dst, hipStream_t st) {
        if (on_device || count == 0 || !src) {
            *dst = src;
            return hipSuccess;
        }
        T* d = nullptr;
        hipError_t e = dalloc(&d, count);
        if (e != hipSuccess) return e;
        owned.push_back(d);
        *dst = d;
        return hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, st);
    }
    template <class T>
    hipError_t out(T* user, size_t count, bool on_device, T** dst) {
        if (on_device || count == 0) {
            *dst = user;
            return hipSuccess;
        }
        T* d = nullptr;
        hipError_t e = dalloc(&d, count);
        if (e != hipSuccess) return e;
        owned.push_back(d);
        *dst = d;
        return hipSuccess;
    }
};

xt_status vmf_check(xt_device* dev, uint32_t num_lobes) {
    if (!dev) return xt_fail(XT_ERR_INVALID_ARGUMENT, "device is NULL");
    if (num_lobes < 1 || num_lobes > (uint32_t)VMF_MAX_LOBES)
        return xt_fail(XT_ERR_INVALID_ARGUMENT, "num_lobes must be in [1, 16]");
    XT_HIP(hipSetDevice(dev->ordinal));
    return XT_OK;
}

}  // namespace

extern "C" {

xt_status xt_vmf_eval(xt_device* dev, uint32_t num_lobes, const float* encoded, uint32_t num_dists,
                      const uint32_t* dist_of_dir, const float* dirs, uint32_t n, float* pdf, int on_device) {
    xt_status s = vmf_check(dev, num_lobes);
    if (s != XT_OK) return s;
    if (n == 0) return XT_OK;
    if (!encoded || !dirs || !pdf || num_dists == 0) return xt_fail(XT_ERR_INVALID_ARGUMENT, "NULL or empty argument");
    const hipStream_t st = hipStreamPerThread;
    Staged sg;
    const float *d_enc, *d_dirs;
    const uint32_t* d_dist;
    float* d_out;
    const size_t E = 5 * num_lobes - 1;
    XT_HIP(sg.in(encoded, E * num_dists, on_device, &d_enc, st));
    XT_HIP(sg.in(dist_of_dir, dist_of_dir ? (size_t)n : 0, on_device, &d_dist, st));
    XT_HIP(sg.in(dirs, 3 * (size_t)n, on_device, &d_dirs, st));
    XT_HIP(sg.out(pdf, n, on_device, &d_out));
    hipLaunchKernelGGL(k_vmf_eval, dim3((n + 255) / 256), dim3(256), 0, st, (int)num_lobes, d_enc, num_dists, d_dist,
                       d_dirs, n, d_out);
    XT_HIP(hipGetLastError());
    if (!on_device) XT_HIP(hipMemcpyAsync(pdf, d_out, n * sizeof(float), hipMemcpyDeviceToHost, st));
    XT_HIP(hipStreamSynchronize(st));
    return XT_OK;
}

xt_status xt_vmf_sample(xt_device* dev, uint32_t num_lobes, const float* encoded, uint32_t num_dists,
                        const uint32_t* dist_of_sample, const float* uniforms, uint32_t n, float* dirs, int on_device) {
    xt_status s = vmf_check(dev, num_lobes);
    if (s != XT_OK) return s;
    if (n == 0) return XT_OK;
    if (!encoded || !uniforms || !dirs || num_dists == 0) return xt_fail(XT_ERR_INVALID_ARGUMENT, "NULL or empty argument");
    const hipStream_t st = hipStreamPerThread;
    Staged sg;
    const float *d_enc, *d_u;
    const uint32_t* d_dist;
    float* d_out;
    const size_t E = 5 * num_lobes - 1;
    XT_HIP(sg.in(encoded, E * num_dists, on_device, &d_enc, st));
    XT_HIP(sg.in(dist_of_sample, dist_of_sample ? (size_t)n : 0, on_device, &d_dist, st));
    XT_HIP(sg.in(uniforms, 4 * (size_t)n, on_device, &d_u, st));
    XT_HIP(sg.out(dirs, 3 * (size_t)n, on_device, &d_out));
    hipLaunchKernelGGL(k_vmf_sample, dim3((n + 255) / 256), dim3(256), 0, st, (int)num_lobes, d_enc, num_dists, d_dist,
                       d_u, n, d_out);
    XT_HIP(hipGetLastError());
    if (!on_device) XT_HIP(hipMemcpyAsync(dirs, d_out, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st));
    XT_HIP(hipStreamSynchronize(st));
    return XT_OK;
}

xt_status xt_vmf_fit(xt_device* dev, uint32_t num_lobes, const float* samples, uint32_t num_samples,
                     const uint32_t* offsets, uint32_t num_fits, const float* init_uniforms, float* encoded,
                     int32_t* steps, int on_device) {
    xt_status s = vmf_check(dev, num_lobes);
    if (s != XT_OK) return s;
    if (num_fits == 0) return XT_OK;
    if (!offsets || !init_uniforms || !encoded || (num_samples && !samples))
        return xt_fail(XT_ERR_INVALID_ARGUMENT, "NULL argument");
    const hipStream_t st = hipStreamPerThread;
    // the kernel reads samples[offsets[f] .. offsets[f+1]): validate the ranges
    std::vector<uint32_t> hoff(num_fits + 1);
    if (on_device) XT_HIP(hipMemcpy(hoff.data(), offsets, hoff.size() * 4, hipMemcpyDeviceToHost));
    else std::copy(offsets, offsets + num_fits + 1, hoff.begin());
    for (uint32_t f = 0; f < num_fits; ++f)
        if (hoff[f + 1] < hoff[f]) return xt_fail(XT_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    if (hoff[num_fits] > num_samples) return xt_fail(XT_ERR_INVALID_ARGUMENT, "offsets exceed num_samples");
    Staged sg;
    const float *d_smp, *d_iu;
    const uint32_t* d_off;
    float* d_enc;
    int32_t* d_steps = nullptr;
    const size_t E = 5 * num_lobes - 1;
    XT_HIP(sg.in(samples, 3 * (size_t)num_samples, on_device, &d_smp, st));
    XT_HIP(sg.in(offsets, (size_t)num_fits + 1, on_device, &d_off, st));
    XT_HIP(sg.in(init_uniforms, (size_t)2 * num_lobes * num_fits, on_device, &d_iu, st));
    XT_HIP(sg.out(encoded, E * num_fits, on_device, &d_enc));
    if (steps) XT_HIP(sg.out(steps, num_fits, on_device, &d_steps));
    const uint32_t grid = std::max(1u, std::min((num_fits + 3) / 4, 4096u));
    kFit[num_lobes](grid, st, d_smp, d_off, num_fits, d_iu, d_enc, d_steps);
    XT_HIP(hipGetLastError());
    if (!on_device) {
        XT_HIP(hipMemcpyAsync(encoded, d_enc, E * num_fits * sizeof(float), hipMemcpyDeviceToHost, st));
        if (steps) XT_HIP(hipMemcpyAsync(steps, d_steps, num_fits * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    }
    XT_HIP(hipStreamSynchronize(st));
    return XT_OK;
}

}  // extern "C"
