// Guiding network buffers, render-side guiding state and kernel declarations
// (guide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "xt_guide.h"

namespace xt {

// Device copies of the network used by the kernels: bf16 weights in natural
// [out][in] order (forward) and transposed [in][out] order (backward), fp32
// biases.
struct GuideNet {
    __bf16 *w0, *w1, *w2, *w3;
    __bf16 *w0t, *w1t, *w2t, *w3t;
    float *b0, *b1, *b2, *b3;
};

// One training sample: query (p, n, wo), sampled direction, weight.
struct GuideSample {
    float q[9];
    float dir[3];
    float weight;
};

__global__ void k_guide_prepare(const float* P, GuideNet net);
__global__ void k_guide_forward(GuideNet net, GuideBounds gb, const float* qin, uint32_t n, float* out);
__global__ void k_guide_infer(GuideNet net, GuideBounds gb, WaveParams wp, DevScene sc, PathQueue q,
                              const uint32_t* count, const float4* hits, float* lobes, uint32_t stride);
__global__ void k_guide_grad(GuideNet net, GuideBounds gb, const GuideSample* samples, uint32_t n, float weight_scale,
                             float weight_clip, float* partial);
__global__ void k_guide_adam(float* P, float* m, float* v, const float* partial, uint32_t nparts, float lr, float b1,
                             float b2, float eps, float b1t, float b2t, float grad_scale);
__global__ void k_guide_finalize(const TrainRec* recs, uint32_t n, const float4* radiance, GuideSample* out,
                                 uint32_t* count, float* weight_sum);

constexpr int GI_WAVES = 4;     // waves per block of k_guide_forward / k_guide_infer
constexpr int GG_BLOCK = 64;    // threads per block of k_guide_grad

}  // namespace xt
