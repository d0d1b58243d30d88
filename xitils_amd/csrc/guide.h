// Guiding network buffers, render-side guiding state and kernel declarations
// (guide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "xt_guide.h"

namespace xt {

// Device copies of the network used by the kernels (bf16 weights laid out as
// the kernels hold them in LDS, fp32 biases).
struct GuideNet {
    struct GuideChain* chain;   // forward copy (below)
    struct GuideBack* back;     // backward copy (below)
};

// Inference copy of the network, laid out as the kernels hold it in LDS.
// Hidden activations stay in registers between layers: the MFMA output
// fragment of one layer is the input fragment of the next once the input
// columns of the next weight matrix are permuted within each block of 32
// (unit_of_kpos); layer 0 columns follow the per-lane feature order
// (feature_of_kpos) and layer 3 rows put whole lobes in each lane
// (output_of_row).  Rows are padded (104, 72 bf16) so 16-byte fragment
// reads hit distinct LDS banks.
constexpr int GC_LD0 = 104, GC_LD = 72;
struct GuideChain {
    __bf16 w0[G_HID * GC_LD0];
    __bf16 w1[G_HID * GC_LD];
    __bf16 w2[G_HID * GC_LD];
    __bf16 w3[G_OUT_PAD * GC_LD];
    float b0[G_HID], b1[G_HID], b2[G_HID], b3[G_OUT_PAD];
};
static_assert(sizeof(GuideChain) % 16 == 0, "GuideChain is copied in 16-byte pieces");

// Backward copy: the transposed weights of layers 3, 2, 1 for the
// activation-gradient chain dH^T = W^T dZ^T, which keeps the gradients in
// registers the same way: rows are the layer's input units in natural order,
// columns k positions of its output units (unit_of_kpos; for layer 3 the
// GuideChain row order, output_of_row).
struct GuideBack {
    __bf16 w3[G_HID * GC_LD];
    __bf16 w2[G_HID * GC_LD];
    __bf16 w1[G_HID * GC_LD];
};
static_assert(sizeof(GuideBack) % 16 == 0, "GuideBack is copied in 16-byte pieces");

// One training sample (64 B, four float4): query (p, n, wo), sampled
// direction, weight.
struct GuideSample {
    float q[9];
    float dir[3];
    float weight;
    float pad[3];
};
static_assert(sizeof(GuideSample) == 64, "GuideSample is read as four float4");

__global__ void k_guide_prepare(const float* P, GuideNet net);
__global__ void k_guide_forward(GuideNet net, GuideBounds gb, const float* qin, uint32_t n, float* out);
__global__ void k_guide_infer(GuideNet net, GuideBounds gb, DevScene sc, PathQueue q, const uint32_t* count,
                              const float4* hits, uint2* lobes_mk, float2* lobes2, uint32_t stride, uint32_t max_vertex);
__global__ void k_guide_grad(GuideNet net, GuideBounds gb, const GuideSample* samples, uint32_t n, float weight_scale,
                             float weight_clip, float* partial);
__global__ void k_guide_adam(float* P, float* m, float* v, const float* partial, uint32_t nparts, float lr, float b1,
                             float b2, float eps, float b1t, float b2t, float grad_scale, GuideNet net);
__global__ void k_guide_finalize(const TrainRec* recs, uint32_t n, const float4* radiance, uint32_t target, GuideSample* out,
                                 uint32_t* count, float* weight_sum);

constexpr int GI_BLOCK = 256;   // threads per block of k_guide_forward / k_guide_infer (64 queries per wave)
#ifndef XT_GUIDE_GRAD_BLOCK
#define XT_GUIDE_GRAD_BLOCK 512
#endif
constexpr int GG_BLOCK = XT_GUIDE_GRAD_BLOCK;   // threads per block of k_guide_grad: 512 (16 samples per
                                                // wave, 2 waves per SIMD) 23.1 us per 32768-sample step,
                                                // 256 (32 per wave, 1 wave per SIMD) 27.0 us
constexpr int GG_BLOCK_SAMPLES = 128;   // a 32768-sample minibatch is one block per CU, each writing one partial row
constexpr int GF_BLOCK = 1024;  // threads per block of k_guide_finalize
constexpr int GA_BLOCK = 1024;  // threads per block of k_guide_adam (32 row groups x 32 parameters)

}  // namespace xt
