// Guiding MLP on gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16) forward and
// backward passes with FP32 accumulation, Adam, and the render-side glue
// (per-vertex inference, training-sample finalisation).
//
// Matrix products use the 16x16x32 MFMA.  Lane l of a wave64 holds
//   A[m0 + l%16][k0 + 8(l/16) + j], B[k0 + 8(l/16) + j][n0 + l%16] (j = 0..7),
//   C[m0 + 4(l/16) + i][n0 + l%16] (i = 0..3)
// (checked on the device).  Every operand is stored with its K index
// contiguous, so a fragment is one 16-byte load.  Activations stay in
// registers between layers (the output fragment of one layer is the input
// fragment of the next, see GuideChain); training writes them unit-major to
// global memory for the weight-gradient products.
#include <hip/hip_runtime.h>

#include "guide.h"

namespace xt {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ bf16x8 frag(const __bf16* M, int ld, int row0, int col0) {
    const int l = __lane_id();
    return *(const bf16x8*)(M + (size_t)(row0 + (l & 15)) * ld + col0 + 8 * (l >> 4));
}

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS written by some lanes of a wave is read by others: order the wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- parameters

// GuideChain layout maps.  An MFMA output fragment leaves lane l = 16g + r
// with units 16b + 4g + i (i = 0..3) of query r; the next layer's input
// fragment wants k positions 32c + 8g + j.  Position p of a hidden layer's
// input therefore holds unit unit_of_kpos(p).
__host__ __device__ __forceinline__ int unit_of_kpos(int p) {
    const int c = p >> 5, g = (p >> 3) & 3, j = p & 7;
    return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}
// Layer 0: lane group g < 3 computes the 24 frequency features of axis g
// (sin, cos of 2^k pi u, k = 0..11), group 3 the normal and view direction.
__host__ __device__ __forceinline__ int feature_of_kpos(int p) {   // -1: zero column
    const int c = p >> 5, g = (p >> 3) & 3, j = p & 7, s = 8 * c + j;
    if (g < 3) return 24 * g + s;
    return s < 6 ? 72 + s : -1;
}
// Layer 3: output row m lands in lane group g as slot s = 4b + i; slots
// 0..9 of group g carry lobes 2g and 2g + 1 (5 raw outputs each).
__host__ __device__ __forceinline__ int output_of_row(int m) {      // -1: unused row
    const int b = m >> 4, g = (m >> 2) & 3, i = m & 3, s = 4 * b + i;
    return s < 10 ? 10 * g + s : -1;
}

// fp32 master parameters -> the inference copy (GuideChain) and the
// backward copy (GuideBack), both bf16 with fp32 biases.
__global__ void k_guide_prepare(const float* __restrict__ P, GuideNet net) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // inference copy (GuideChain)
    GuideChain& ch = *net.chain;
    if (i < G_HID * GC_LD0) {
        const int m = i / GC_LD0, p = i % GC_LD0, f = p < G_IN ? feature_of_kpos(p) : -1;
        ch.w0[i] = (__bf16)(f >= 0 ? P[G_OFF_W0 + m * G_IN + f] : 0.0f);
    }
    if (i < G_HID * GC_LD) {
        const int m = i / GC_LD, p = i % GC_LD, u = p < G_HID ? unit_of_kpos(p) : -1;
        ch.w1[i] = (__bf16)(u >= 0 ? P[G_OFF_W1 + m * G_HID + u] : 0.0f);
        ch.w2[i] = (__bf16)(u >= 0 ? P[G_OFF_W2 + m * G_HID + u] : 0.0f);
    }
    if (i < G_OUT_PAD * GC_LD) {
        const int m = i / GC_LD, p = i % GC_LD, u = p < G_HID ? unit_of_kpos(p) : -1, o = output_of_row(m);
        ch.w3[i] = (__bf16)(u >= 0 && o >= 0 ? P[G_OFF_W3 + o * G_HID + u] : 0.0f);
    }
    if (i < G_HID) {
        ch.b0[i] = P[G_OFF_B0 + i];
        ch.b1[i] = P[G_OFF_B1 + i];
        ch.b2[i] = P[G_OFF_B2 + i];
    }
    if (i < G_OUT_PAD) {
        const int o = output_of_row(i);
        ch.b3[i] = o >= 0 ? P[G_OFF_B3 + o] : 0.0f;
    }
    // backward copy (GuideBack): bk.wL[m][p] = W_L[unit at k position p][input unit m]
    GuideBack& bk = *net.back;
    if (i < G_HID * GC_LD) {
        const int m = i / GC_LD, p = i % GC_LD, u = p < G_HID ? unit_of_kpos(p) : -1;
        const int o = u >= 0 && u < G_OUT_PAD ? output_of_row(u) : -1;
        bk.w3[i] = (__bf16)(o >= 0 ? P[G_OFF_W3 + o * G_HID + m] : 0.0f);
        bk.w2[i] = (__bf16)(u >= 0 ? P[G_OFF_W2 + u * G_HID + m] : 0.0f);
        bk.w1[i] = (__bf16)(u >= 0 ? P[G_OFF_W1 + u * G_HID + m] : 0.0f);
    }
}

// ---------------------------------------------------------------- inference

__device__ __forceinline__ void load_chain(const GuideChain* src, GuideChain& dst) {
    const uint4* s = (const uint4*)src;
    uint4* d = (uint4*)&dst;
    for (int k = threadIdx.x; k < (int)(sizeof(GuideChain) / 16); k += blockDim.x) d[k] = s[k];
    __syncthreads();
}

// (lo, hi) rounded to bf16 (nearest even) in one packed conversion
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2){lo, hi}, b2));
}
// the two bf16 halves of w clamped at zero as signed 16-bit integers: bf16
// is sign-magnitude, so every negative value (-0 included) becomes +0
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t w) {
    typedef short s2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s2, w), (s2)0));
}

// This lane's 24 input features of one query (layer-0 input fragment, k
// blocks 0..2): group g < 3 the frequency features of axis g, group 3 the
// shading normal and view direction in its first six positions.  Group 3
// also runs the band recurrence (on z) and leaves those finite values in its
// other 18 positions, whose layer-0 weights are zero, so only six values
// depend on the group.  sin and cos of pi u by xt_sincosf (within 2 ulp of
// sin and cos of the rounded argument, as a float libm gives them: the
// recurrence doubles the error per band, so the base must be that close),
// then the doubling recurrence sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x
// (one fma).
__device__ __forceinline__ void chain_features(const GuideBounds& gb, f3 p, f3 n, f3 wo, bf16x8 (&x)[3]) {
    const int g = __lane_id() >> 4;
    const float PI = 3.14159265358979323846f;
    const float pa = g == 0 ? p.x : (g == 1 ? p.y : p.z);
    const float lo = g == 0 ? gb.lo.x : (g == 1 ? gb.lo.y : gb.lo.z);
    const float inv = g == 0 ? gb.inv_ext.x : (g == 1 ? gb.inv_ext.y : gb.inv_ext.z);
    const float ua = fminf(fmaxf((pa - lo) * inv, 0.0f), 1.0f);
    float sn, cs;
    xt_sincosf(PI * ua, &sn, &cs);
    float v[24];
#pragma unroll
    for (int k = 0; k < G_BANDS; ++k) {
        v[2 * k] = sn;
        v[2 * k + 1] = cs;
        const float d = sn + sn;
        cs = __builtin_fmaf(-d, sn, 1.0f);
        sn = d * v[2 * k + 1];
    }
    if (g == 3) {
        v[0] = n.x; v[1] = n.y; v[2] = n.z;
        v[3] = wo.x; v[4] = wo.y; v[5] = wo.z;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        x[c] = __builtin_bit_cast(bf16x8, make_uint4(pack_bf16(v[8 * c], v[8 * c + 1]), pack_bf16(v[8 * c + 2], v[8 * c + 3]),
                                                     pack_bf16(v[8 * c + 4], v[8 * c + 5]),
                                                     pack_bf16(v[8 * c + 6], v[8 * c + 7])));
}

// relu(acc) of two adjacent 16-unit output tiles -> one input fragment:
// rounded to bf16 first, then clamped (equal to rounding relu(acc)), one
// packed conversion and one packed integer max per pair of units.
__device__ __forceinline__ bf16x8 relu_pack(floatx4 a, floatx4 b) {
    const uint4 w = make_uint4(relu_bf16x2(pack_bf16(a[0], a[1])), relu_bf16x2(pack_bf16(a[2], a[3])),
                               relu_bf16x2(pack_bf16(b[0], b[1])), relu_bf16x2(pack_bf16(b[2], b[3])));
    return __builtin_bit_cast(bf16x8, w);
}

// The four layers for NT tiles of 16 queries, activations in registers.
// y[t][b][i] = raw output row 16b + 4g + i (GuideChain order) of query 16t + r.
template <int NT>
__device__ __forceinline__ void chain_mlp(const GuideChain& s, const bf16x8 (&x)[NT][3], floatx4 (&y)[NT][3]) {
    const int g = __lane_id() >> 4;
    floatx4 acc[NT][4];
    bf16x8 h[NT][2];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const floatx4 bias = *(const floatx4*)&s.b0[16 * b + 4 * g];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t][b] = bias;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bf16x8 a = frag(s.w0, GC_LD0, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, x[t][c], acc[t][b]);
        }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        h[t][0] = relu_pack(acc[t][0], acc[t][1]);
        h[t][1] = relu_pack(acc[t][2], acc[t][3]);
    }
    const __bf16* wl[2] = {s.w1, s.w2};
    const float* bl[2] = {s.b1, s.b2};
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const floatx4 bias = *(const floatx4*)&bl[layer][16 * b + 4 * g];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][b] = bias;
        }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const bf16x8 a = frag(wl[layer], GC_LD, 16 * b, 32 * c);
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, h[t][c], acc[t][b]);
            }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            h[t][0] = relu_pack(acc[t][0], acc[t][1]);
            h[t][1] = relu_pack(acc[t][2], acc[t][3]);
        }
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const floatx4 bias = *(const floatx4*)&s.b3[16 * b + 4 * g];
#pragma unroll
        for (int t = 0; t < NT; ++t) y[t][b] = bias;
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bf16x8 a = frag(s.w3, GC_LD, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) y[t][b] = mfma(a, h[t][c], y[t][b]);
        }
}

// Slot k (0..9) of this lane's two lobes: raw output 10g + k of the query.
__device__ __forceinline__ float out_slot(const floatx4 (&y)[3], int k) { return y[k >> 2][k & 3]; }

// Raw network outputs for n queries (p, n, wo as 9 floats each): out[q * 40 + o].
__global__ __launch_bounds__(256) void k_guide_forward(GuideNet net, GuideBounds gb, const float* __restrict__ qin,
                                                       uint32_t n, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) GuideChain s;
    load_chain(net.chain, s);
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u; base < n; base += nwaves * 64u) {
        float qv[9];
        const uint32_t qi = base + l;
#pragma unroll
        for (int k = 0; k < 9; ++k) qv[k] = qi < n ? qin[(size_t)qi * 9 + k] : 0.0f;
        bf16x8 x[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float v[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) v[k] = __shfl(qv[k], 16 * t + r);
            chain_features(gb, mk(v[0], v[1], v[2]), mk(v[3], v[4], v[5]), mk(v[6], v[7], v[8]), x[t]);
        }
        floatx4 y[4][3];
        chain_mlp<4>(s, x, y);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t q = base + 16 * t + r;
            if (q < n)
#pragma unroll
                for (int k = 0; k < 10; ++k) out[(size_t)q * G_OUT + 10 * g + k] = out_slot(y[t], k);
        }
    }
}

// Decoded mixtures for the extension-queue entries of one bounce that the
// shade kernel may guide (hit, not emissive, not a delta material):
// lobes_mk[j * stride + i] = (mu, kappa) as halves, lobes2[j * stride + i] =
// (alpha, alpha * vmf_norm(kappa)) of lobe j, kappa as stored.
#ifndef XT_GUIDE_INFER_NT
#define XT_GUIDE_INFER_NT 2
#endif
constexpr int GI_NT = XT_GUIDE_INFER_NT;
static_assert(GI_NT == 1 || GI_NT == 2 || GI_NT == 4, "tiles per inference pass");
#ifndef XT_GUIDE_INFER_WAVES
#define XT_GUIDE_INFER_WAVES 4
#endif
constexpr int GI_WAVES = XT_GUIDE_INFER_WAVES;   // waves per SIMD the register budget allows

__global__ __launch_bounds__(256, GI_WAVES) void k_guide_infer(GuideNet net, GuideBounds gb, DevScene sc, PathQueue pq,
                                                     const uint32_t* __restrict__ count, const float4* __restrict__ hits,
                                                     uint2* __restrict__ lobes_mk, float2* __restrict__ lobes2,
                                                     uint32_t stride, uint32_t max_vertex) {
    __shared__ __attribute__((aligned(16))) GuideChain s;
    load_chain(net.chain, s);
    const uint32_t n = *count;
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    const float PI_F = 3.14159265358979323846f;
    // this lane group's two lobes (2g, 2g + 1)
    uint2* const lobe_mk[2] = {lobes_mk + (size_t)(2 * g) * stride, lobes_mk + (size_t)(2 * g + 1) * stride};
    float2* const lobe2[2] = {lobes2 + (size_t)(2 * g) * stride, lobes2 + (size_t)(2 * g + 1) * stride};
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + l;
        bool use = false;
        f3 P = mk(0, 0, 0), N = mk(0, 0, 0), WO = mk(0, 0, 0);
        if (i < n) {
            const float4 h = hits[i];
            const uint32_t slot = __float_as_uint(h.y);
            const uint32_t v = __float_as_uint(pq.w[i].w) & 0xFFFFFFu;   // vertex index (shaded at v)
            // every surface hit gets lobes: the shade kernel itself skips emitters
            // (their paths end) and delta materials, so no material lookup here
            // stands between the hit and the surface reads
            if (slot != INVALID && (max_vertex == 0 || v <= max_vertex)) {
                const Surf sf = make_surf(sc, slot, xyz(pq.d[i]), h.z, h.w);
                P = sf.p;
                N = sf.sn;
                WO = sf.wo;
                use = true;
            }
        }
        if (__ballot(use) == 0) continue;
        // the wave's 64 queries in passes of GI_NT tiles of 16 (fewer live
        // registers per pass: more waves per SIMD)
#pragma unroll
        for (int pass = 0; pass < 4 / GI_NT; ++pass) {
            bf16x8 x[GI_NT][3];
#pragma unroll
            for (int t = 0; t < GI_NT; ++t) {
                const int src = 16 * (GI_NT * pass + t) + r;
                chain_features(gb, mk(__shfl(P.x, src), __shfl(P.y, src), __shfl(P.z, src)),
                               mk(__shfl(N.x, src), __shfl(N.y, src), __shfl(N.z, src)),
                               mk(__shfl(WO.x, src), __shfl(WO.y, src), __shfl(WO.z, src)), x[t]);
            }
            floatx4 y[GI_NT][3];
            chain_mlp<GI_NT>(s, x, y);
#pragma unroll
            for (int t = 0; t < GI_NT; ++t) {
                const int src = 16 * (GI_NT * pass + t) + r;
                const bool useq = __shfl((int)use, src) != 0;
                // softmax over the 8 alpha logits held by the 4 lanes of this query
                // (hardware exp2 and reciprocals: the lobes are stored as halves
                // and alpha as a float the shade kernel uses for sampling and pdf alike)
                const float LOG2E = 1.44269504088896341f;
                const float la = out_slot(y[t], 4), lb = out_slot(y[t], 9);
                float mx = fmaxf(la, lb);
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                const float ea = __builtin_amdgcn_exp2f((la - mx) * LOG2E), eb = __builtin_amdgcn_exp2f((lb - mx) * LOG2E);
                float sum = ea + eb;
                sum += __shfl_xor(sum, 16);
                sum += __shfl_xor(sum, 32);
                const float rsum = __builtin_amdgcn_rcpf(sum);
                if (useq) {
                    const uint32_t q = base + src;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const f3 m = mk(out_slot(y[t], 5 * k), out_slot(y[t], 5 * k + 1), out_slot(y[t], 5 * k + 2));
                        const float l2 = dot(m, m);
                        const f3 mu = l2 > 1e-24f ? m * __builtin_amdgcn_rsqf(l2) : mk(0.0f, 0.0f, 1.0f);
                        const float kl = fminf(fmaxf(out_slot(y[t], 5 * k + 3), G_KLOG_MIN), G_KLOG_MAX);
                        const float kappa = __builtin_amdgcn_exp2f(kl * LOG2E);
                        const float alpha = (k == 0 ? ea : eb) * rsum;
                        const uint2 mk2 = make_uint2(pack_half2(mu.x, mu.y), pack_half2(mu.z, kappa));
                        // alpha k / (2 pi (1 - e^{-2k})) with k the stored half (>= e^-3, no cancellation)
                        const float kh = half_hi(mk2.y);
                        const float norm =
                            kh * __builtin_amdgcn_rcpf(2.0f * PI_F * (1.0f - __builtin_amdgcn_exp2f(-2.0f * LOG2E * kh)));
                        lobe_mk[k][q] = mk2;
                        lobe2[k][q] = make_float2(alpha, alpha * norm);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- training: per-sample chains

// The four layers for NT tiles of 16 samples keeping every activation (the
// input fragments of the next layer) for the backward pass.
template <int NT>
__device__ __forceinline__ void chain_forward_keep(const GuideChain& s, const bf16x8 (&x)[NT][3], bf16x8 (&h1)[NT][2],
                                                   bf16x8 (&h2)[NT][2], bf16x8 (&h3)[NT][2], floatx4 (&y)[NT][3]) {
    const int g = __lane_id() >> 4;
    floatx4 acc[NT][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const floatx4 bias = *(const floatx4*)&s.b0[16 * b + 4 * g];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t][b] = bias;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bf16x8 a = frag(s.w0, GC_LD0, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, x[t][c], acc[t][b]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        h1[t][0] = relu_pack(acc[t][0], acc[t][1]);
        h1[t][1] = relu_pack(acc[t][2], acc[t][3]);
    }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const __bf16* w = layer == 0 ? s.w1 : s.w2;
        const float* bl = layer == 0 ? s.b1 : s.b2;
        bf16x8 (&hin)[NT][2] = layer == 0 ? h1 : h2;
        bf16x8 (&hout)[NT][2] = layer == 0 ? h2 : h3;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const floatx4 bias = *(const floatx4*)&bl[16 * b + 4 * g];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][b] = bias;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const bf16x8 a = frag(w, GC_LD, 16 * b, 32 * c);
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, hin[t][c], acc[t][b]);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            hout[t][0] = relu_pack(acc[t][0], acc[t][1]);
            hout[t][1] = relu_pack(acc[t][2], acc[t][3]);
        }
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const floatx4 bias = *(const floatx4*)&s.b3[16 * b + 4 * g];
#pragma unroll
        for (int t = 0; t < NT; ++t) y[t][b] = bias;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bf16x8 a = frag(s.w3, GC_LD, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) y[t][b] = mfma(a, h3[t][c], y[t][b]);
        }
    }
}

// Two adjacent 16-unit gradient tiles -> one input fragment (no activation).
__device__ __forceinline__ bf16x8 grad_pack(floatx4 a, floatx4 b) {
    bf16x8 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h[i] = (__bf16)a[i];
        h[4 + i] = (__bf16)b[i];
    }
    return h;
}

// d = W^T z for one layer of the backward chain (weights: GuideBack rows
// [64][GC_LD], z: input fragments over the layer's output units), masked by
// the layer input's activation (h > 0) -> dz of the layer below in output
// tiles (units 16b + 4g + i in natural order).
template <int NT>
__device__ __forceinline__ void back_step(const __bf16* w, const bf16x8 (&z)[NT][2], const bf16x8 (&h)[NT][2],
                                          floatx4 (&dz)[NT][4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
        for (int t = 0; t < NT; ++t) dz[t][b] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bf16x8 a = frag(w, GC_LD, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) dz[t][b] = mfma(a, z[t][c], dz[t][b]);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (!((float)h[t][b >> 1][4 * (b & 1) + i] > 0.0f)) dz[t][b][i] = 0.0f;
    }
}

// Gradient of the loss -min(w weight_scale, weight_clip) log q(dir | query)
// with respect to the raw outputs y of NT tiles of 16 samples: lane group g
// holds lobes 2g, 2g + 1 (slots 0..9) of sample 16t + r; wgt is the clipped
// weight (0 for padding samples, whose gradient is then 0).
template <int NT>
__device__ __forceinline__ void loss_grad(const floatx4 (&y)[NT][3], const f3 (&dir)[NT], const float (&wgt)[NT],
                                          floatx4 (&dy)[NT][3]) {
    // hardware exp2 / reciprocal / square root (a few ulp, far below the bf16
    // rounding the gradient then takes); expm1 stays exact, 1/k - 2/expm1(2k)
    // cancels for small k
    const float LOG2E = 1.44269504088896341f, TWO_PI = 6.28318530717958648f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const float la = out_slot(y[t], 4), lb = out_slot(y[t], 9);
        float mx = fmaxf(la, lb);
        mx = fmaxf(mx, __shfl_xor(mx, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        const float ea = __builtin_amdgcn_exp2f((la - mx) * LOG2E), eb = __builtin_amdgcn_exp2f((lb - mx) * LOG2E);
        float esum = ea + eb;
        esum += __shfl_xor(esum, 16);
        esum += __shfl_xor(esum, 32);
        const float resum = __builtin_amdgcn_rcpf(esum);
        f3 mu[2];
        float alpha[2], kap[2], rlen[2], v[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const f3 m = mk(out_slot(y[t], 5 * k), out_slot(y[t], 5 * k + 1), out_slot(y[t], 5 * k + 2));
            const float len = __builtin_sqrtf(dot(m, m));
            rlen[k] = __builtin_amdgcn_rcpf(fmaxf(len, 1e-12f));
            mu[k] = len > 1e-12f ? m * rlen[k] : mk(0.0f, 0.0f, 1.0f);
            kap[k] = __builtin_amdgcn_exp2f(fminf(fmaxf(out_slot(y[t], 5 * k + 3), G_KLOG_MIN), G_KLOG_MAX) * LOG2E);
            alpha[k] = (k == 0 ? ea : eb) * resum;
            // vMF pdf: k / (2 pi (1 - e^{-2k})) e^{k (mu.dir - 1)}
            const float norm = kap[k] * __builtin_amdgcn_rcpf(TWO_PI * (1.0f - __builtin_amdgcn_exp2f(-2.0f * LOG2E * kap[k])));
            v[k] = norm * __builtin_amdgcn_exp2f(kap[k] * (dot(dir[t], mu[k]) - 1.0f) * LOG2E);
        }
        float q = alpha[0] * v[0] + alpha[1] * v[1];
        q += __shfl_xor(q, 16);
        q += __shfl_xor(q, 32);
        const float rq = __builtin_amdgcn_rcpf(fmaxf(q, 1e-30f));
        float d[12];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float rsp = alpha[k] * v[k] * rq;
            const float mw = dot(mu[k], dir[t]);
            const f3 gm = (dir[t] - mw * mu[k]) * (rsp * kap[k] * rlen[k]);
            const float dk = __builtin_amdgcn_rcpf(kap[k]) - 2.0f * __builtin_amdgcn_rcpf(expm1f(2.0f * kap[k])) + (mw - 1.0f);
            const float kl = out_slot(y[t], 5 * k + 3);
            const float gk = (kl > G_KLOG_MIN && kl < G_KLOG_MAX) ? rsp * dk * kap[k] : 0.0f;
            const float ga = rsp - alpha[k];
            d[5 * k] = -wgt[t] * gm.x;
            d[5 * k + 1] = -wgt[t] * gm.y;
            d[5 * k + 2] = -wgt[t] * gm.z;
            d[5 * k + 3] = -wgt[t] * gk;
            d[5 * k + 4] = -wgt[t] * ga;
        }
        d[10] = d[11] = 0.0f;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) dy[t][b][i] = d[4 * b + i];
    }
}

// ---------------------------------------------------------------- training: one fused gradient kernel

// k_guide_grad: a block of GG_BLOCK threads takes GG_SAMPLES samples (each
// wave 16 GG_NT of them), runs their forward and backward chains with the
// activations in registers, and forms the weight and bias gradients of the
// block's samples layer by layer as MFMA products with samples as K: the
// pair a layer's gradient needs (its pre-activation gradient and its input
// activations) is staged in LDS unit-major, [unit][sample] bf16, the block
// synchronises, and each wave multiplies its share of the 16 x 16 output
// tiles.  Nothing per sample leaves the chip; each block writes one partial
// row (G_NPARAMS floats), summed by k_guide_adam in a fixed order.
constexpr int GG_WAVES = GG_BLOCK / 64;
constexpr int GG_NT = GG_BLOCK_SAMPLES / (16 * GG_WAVES);   // tiles of 16 samples per wave
constexpr int GG_SAMPLES = GG_BLOCK_SAMPLES;
static_assert(GG_NT == 1 || GG_NT == 2, "one or two tiles of 16 samples per wave");
static_assert(GG_WAVES * 16 * GG_NT == GG_SAMPLES, "the block's samples");
constexpr int GG_LD = GG_SAMPLES + 8;               // bf16 per staged row: 272-byte rows, 16-byte aligned
constexpr int GG_ROWS = G_HID + G_IN;               // the largest pair: dz1 and the features
struct GradTile {
    __bf16 v[GG_ROWS * GG_LD];
};

// Column of this lane's sample 16t + r of wave w (with two tiles, the lane's
// two samples side by side: one 32-bit store per unit).  The column order is
// the same for every row, so the products over samples do not depend on it.
__device__ __forceinline__ int grad_col(int w, int r) { return w * 16 * GG_NT + GG_NT * r; }

// this lane's values of one unit, one per tile, into row `row`
__device__ __forceinline__ void put_unit(GradTile& S, int row, int col, const float (&v)[GG_NT]) {
    if constexpr (GG_NT == 2) *(uint32_t*)&S.v[row * GG_LD + col] = pack_bf16(v[0], v[1]);
    else S.v[row * GG_LD + col] = (__bf16)v[0];
}
__device__ __forceinline__ void put_unit_bf16(GradTile& S, int row, int col, __bf16 a, __bf16 b) {
    if constexpr (GG_NT == 2)
        *(uint32_t*)&S.v[row * GG_LD + col] =
            (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    else S.v[row * GG_LD + col] = a;
}
// output tiles (units 16b + 4g + i) -> rows row0 + unit
template <int NB>
__device__ __forceinline__ void stage_units(GradTile& S, int row0, int col, const floatx4 (&v)[GG_NT][NB]) {
    const int g = __lane_id() >> 4;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float u[GG_NT];
#pragma unroll
            for (int t = 0; t < GG_NT; ++t) u[t] = v[t][b][i];
            put_unit(S, row0 + 16 * b + 4 * g + i, col, u);
        }
}
// input fragments (k position 32c + 8g + j = unit unit_of_kpos) -> rows row0 + unit
__device__ __forceinline__ void stage_frags(GradTile& S, int row0, int col, const bf16x8 (&h)[GG_NT][2]) {
    const int g = __lane_id() >> 4;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 8; ++j)
            put_unit_bf16(S, row0 + unit_of_kpos(32 * c + 8 * g + j), col, h[0][c][j], h[GG_NT - 1][c][j]);
}

// This wave's share of one layer's gradient over the block's samples: the
// layer's (rows / 16) x (IN / 16) output tiles in parts of TP tiles of one
// 16-row strip, part w for wave w: dW[m][n] = sum_k D[drow + m][k] A[arow +
// n][k], and for the strip's first part db[m] = sum_k D[drow + m][k]; added
// to (first pass: written to) the block's partial row.
template <int IN, int TP>
__device__ __forceinline__ void grad_part(const GradTile& S, int strips, int drow, int arow, float* row, int off_w,
                                          int off_b, bool first) {
    constexpr int PPS = IN / 16 / TP;   // parts per strip
    static_assert(PPS * TP * 16 == IN, "parts tile the strip");
    const int w = threadIdx.x >> 6;
    if (w >= strips * PPS) return;
    const int sb = w / PPS, c0 = (w % PPS) * TP;
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    floatx4 acc[TP];
#pragma unroll
    for (int c = 0; c < TP; ++c) acc[c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
#pragma unroll
    for (int k = 0; k < GG_SAMPLES; k += 32) {
        const bf16x8 a = frag(S.v, GG_LD, drow + 16 * sb, k);
#pragma unroll
        for (int j = 0; j < 8; ++j) bsum += (float)a[j];
#pragma unroll
        for (int c = 0; c < TP; ++c) acc[c] = mfma(a, frag(S.v, GG_LD, arow + 16 * (c0 + c), k), acc[c]);
    }
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
#pragma unroll
    for (int c = 0; c < TP; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float* d = &row[off_w + (16 * sb + 4 * g + i) * IN + 16 * (c0 + c) + r];
            *d = first ? acc[c][i] : *d + acc[c][i];
        }
    if (c0 == 0 && g == 0) {
        float* d = &row[off_b + 16 * sb + r];
        *d = first ? bsum : *d + bsum;
    }
}
// tiles per part: a whole strip per wave with 4 waves; with 8 waves, half a
// strip (2 of 4 tiles, 3 of 6)
constexpr int GG_TP_HID = GG_WAVES >= 8 ? 2 : 4;
constexpr int GG_TP_IN = GG_WAVES >= 8 ? 3 : 6;

// Gradient of sum_i -min(w_i * weight_scale, weight_clip) log q(dir_i |
// query_i) over samples [0, n): block x writes partial row x, taking sample
// groups x, x + gridDim.x, ... (every block takes at least one; samples >= n
// contribute zero).
__global__ __launch_bounds__(GG_BLOCK) void k_guide_grad(GuideNet net, GuideBounds gb, const GuideSample* __restrict__ smp,
                                                         uint32_t n, float weight_scale, float weight_clip,
                                                         float* __restrict__ partial) {
    constexpr int NT = GG_NT;
    __shared__ __attribute__((aligned(16))) GuideChain s;
    __shared__ __attribute__((aligned(16))) GuideBack bk;
    __shared__ __attribute__((aligned(16))) GradTile S;
    {
        const uint4* src = (const uint4*)net.back;
        uint4* dst = (uint4*)&bk;
        for (int k = threadIdx.x; k < (int)(sizeof(GuideBack) / 16); k += blockDim.x) dst[k] = src[k];
    }
    load_chain(net.chain, s);   // includes the block barrier
    const int w = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, r = l & 15;
    const int col = grad_col(w, r);
    float* row = partial + (size_t)blockIdx.x * G_NPARAMS;
    bool first = true;
    for (uint32_t blk = blockIdx.x * GG_SAMPLES; first || blk < n; blk += gridDim.x * GG_SAMPLES) {
        const uint32_t base = blk + w * 16 * NT;
        bf16x8 x[NT][3];
        f3 dir[NT];
        float wgt[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const uint32_t si = base + 16 * t + r;
            f3 P = mk(0, 0, 0), N = mk(0, 0, 0), WO = mk(0, 0, 0);
            dir[t] = mk(0, 0, 1);
            wgt[t] = 0.0f;
            const bool valid = si < n;
            if (valid) {
                const float4* sv = (const float4*)&smp[si];
                const float4 s0 = sv[0], s1 = sv[1], s2 = sv[2], s3 = sv[3];
                P = mk(s0.x, s0.y, s0.z);
                N = mk(s0.w, s1.x, s1.y);
                WO = mk(s1.z, s1.w, s2.x);
                dir[t] = mk(s2.y, s2.z, s2.w);
                wgt[t] = fminf(s3.x * weight_scale, weight_clip);
            }
            chain_features(gb, P, N, WO, x[t]);
            if (!valid)
#pragma unroll
                for (int c = 0; c < 3; ++c) x[t][c] = bf16x8{};
        }
        bf16x8 h1[NT][2], h2[NT][2], h3[NT][2];
        floatx4 y[NT][3];
        chain_forward_keep<NT>(s, x, h1, h2, h3, y);
        floatx4 dy[NT][3];
        loss_grad<NT>(y, dir, wgt, dy);

        // layer 3: dW3 = dy^T h3 (dy rows: raw output index, 40..47 zero)
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            float u[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) u[t] = out_slot(dy[t], k);
            put_unit(S, 10 * g + k, col, u);
        }
        if (g == 0) {
            const float zero[NT] = {};
#pragma unroll
            for (int k = G_OUT; k < G_OUT_PAD; ++k) put_unit(S, k, col, zero);
        }
        stage_frags(S, G_OUT_PAD, col, h3);
        __syncthreads();
        grad_part<G_HID, GG_TP_HID>(S, 3, 0, G_OUT_PAD, row, G_OFF_W3, G_OFF_B3, first);
        __syncthreads();
        // dz3 = (W3^T dy) [h3 > 0]; dW2 = dz3^T h2
        bf16x8 z[NT][2];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            z[t][0] = grad_pack(dy[t][0], dy[t][1]);
            z[t][1] = grad_pack(dy[t][2], floatx4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        floatx4 dz[NT][4];
        back_step<NT>(bk.w3, z, h3, dz);
        stage_units<4>(S, 0, col, dz);
        stage_frags(S, G_HID, col, h2);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            z[t][0] = grad_pack(dz[t][0], dz[t][1]);
            z[t][1] = grad_pack(dz[t][2], dz[t][3]);
        }
        __syncthreads();
        grad_part<G_HID, GG_TP_HID>(S, 4, 0, G_HID, row, G_OFF_W2, G_OFF_B2, first);
        __syncthreads();
        // dz2 = (W2^T dz3) [h2 > 0]; dW1 = dz2^T h1
        back_step<NT>(bk.w2, z, h2, dz);
        stage_units<4>(S, 0, col, dz);
        stage_frags(S, G_HID, col, h1);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            z[t][0] = grad_pack(dz[t][0], dz[t][1]);
            z[t][1] = grad_pack(dz[t][2], dz[t][3]);
        }
        __syncthreads();
        grad_part<G_HID, GG_TP_HID>(S, 4, 0, G_HID, row, G_OFF_W1, G_OFF_B1, first);
        __syncthreads();
        // dz1 = (W1^T dz2) [h1 > 0]; dW0 = dz1^T x (features in natural order;
        // group 3's positions beyond its six direction features are 0)
        back_step<NT>(bk.w1, z, h1, dz);
        stage_units<4>(S, 0, col, dz);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = g < 3 ? 24 * g + 8 * c + j : 72 + 8 * c + j;
                const bool live = 8 * c + j < 6 || g < 3;
                const __bf16 z0 = (__bf16)0.0f;
                put_unit_bf16(S, G_HID + f, col, live ? x[0][c][j] : z0, live ? x[NT - 1][c][j] : z0);
            }
        __syncthreads();
        grad_part<G_IN, GG_TP_IN>(S, 4, 0, G_HID, row, G_OFF_W0, G_OFF_B0, first);
        __syncthreads();
        first = false;
    }
}

// ---------------------------------------------------------------- training: Adam

// Inverses of the GuideChain / GuideBack layout maps (k_guide_prepare
// gathers with the maps, k_guide_adam scatters with these).
__device__ __forceinline__ int kpos_of_unit(int v) {
    const int c = v >> 5, u = v & 31;
    return 32 * c + (u < 16 ? 8 * (u >> 2) + (u & 3) : 8 * ((u - 16) >> 2) + 4 + ((u - 16) & 3));
}
__device__ __forceinline__ int kpos_of_feature(int f) {   // -1: not an input of the chain
    if (f < 72) {
        const int g = f / 24, s = f % 24;
        return 32 * (s >> 3) + 8 * g + (s & 7);
    }
    return f < 78 ? 24 + (f - 72) : -1;
}
__device__ __forceinline__ int row_of_output(int o) {       // o < G_OUT
    const int g = o / 10, s = o % 10;
    return 16 * (s >> 2) + 4 * g + (s & 3);
}

// Adam (Kingma & Ba) on the fp32 master parameters: the gradient is the sum
// of nparts partial rows (GA_GROUPS interleaved row groups per parameter,
// then the groups in order: deterministic) times grad_scale; b1t, b2t =
// beta^t.  32 parameters per block, one row group per wave half, so a
// 256-row sum is 8 loads per thread with 16 waves per block in flight.  The
// updated parameter is also written, rounded, to its places in the bf16
// inference and backward copies (whose padding k_guide_prepare zeroed).
__global__ __launch_bounds__(GA_BLOCK) void k_guide_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                                                         const float* __restrict__ partial, uint32_t nparts, float lr,
                                                         float b1, float b2, float eps, float b1t, float b2t,
                                                         float grad_scale, GuideNet net) {
    constexpr int GA_GROUPS = GA_BLOCK / 32;
    __shared__ float red[GA_GROUPS][32];
    const int pl = threadIdx.x & 31, rg = threadIdx.x >> 5;
    const int i = blockIdx.x * 32 + pl;
    float gp = 0.0f;
    if (i < G_NPARAMS)
        for (uint32_t b = rg; b < nparts; b += GA_GROUPS) gp += partial[(size_t)b * G_NPARAMS + i];
    red[rg][pl] = gp;
    __syncthreads();
    if (rg != 0 || i >= G_NPARAMS) return;
    float g = 0.0f;
    for (int r = 0; r < GA_GROUPS; ++r) g += red[r][pl];
    g *= grad_scale;
    if (!isfinite(g)) g = 0.0f;
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - b1t), vh = vi / (1.0f - b2t);
    const float pn = P[i] - lr * mh / (sqrtf(vh) + eps);
    P[i] = pn;
    GuideChain& ch = *net.chain;
    GuideBack& bk = *net.back;
    const __bf16 hb = (__bf16)pn;
    auto hidden = [&](__bf16* cw, __bf16* bw, int k) {   // W1 or W2 [o][in], k = o * G_HID + in
        const int o = k / G_HID, in = k % G_HID;
        cw[o * GC_LD + kpos_of_unit(in)] = hb;
        bw[in * GC_LD + kpos_of_unit(o)] = hb;
    };
    if (i < G_OFF_B0) {
        const int o = i / G_IN, f = i % G_IN, p = kpos_of_feature(f);
        if (p >= 0) ch.w0[o * GC_LD0 + p] = hb;
    } else if (i < G_OFF_W1) {
        ch.b0[i - G_OFF_B0] = pn;
    } else if (i < G_OFF_B1) {
        hidden(ch.w1, bk.w1, i - G_OFF_W1);
    } else if (i < G_OFF_W2) {
        ch.b1[i - G_OFF_B1] = pn;
    } else if (i < G_OFF_B2) {
        hidden(ch.w2, bk.w2, i - G_OFF_W2);
    } else if (i < G_OFF_W3) {
        ch.b2[i - G_OFF_B2] = pn;
    } else if (i < G_OFF_B3) {
        const int k = i - G_OFF_W3, o = k / G_HID, in = k % G_HID;
        if (o < G_OUT) {
            const int rw = row_of_output(o);
            ch.w3[rw * GC_LD + kpos_of_unit(in)] = hb;
            bk.w3[in * GC_LD + kpos_of_unit(rw)] = hb;
        }
    } else {
        const int o = i - G_OFF_B3;
        if (o < G_OUT) ch.b3[row_of_output(o)] = pn;
    }
}

// Training records -> samples.  C = (L_path - pre) / W estimates
// f cos L_i / pdf along the sampled direction; the weight is
//   target 0 (radiance): luminance(L_i) / pdf = luminance(C / E) / pdf
//   target 1 (product):  luminance(f cos L_i) / pdf = luminance(C)
// Positive finite weights are appended to out (order not fixed) and summed.
__global__ __launch_bounds__(GF_BLOCK) void k_guide_finalize(const TrainRec* __restrict__ recs, uint32_t n,
                                                            const float4* __restrict__ radiance,
                                 uint32_t target, GuideSample* __restrict__ out, uint32_t* __restrict__ count,
                                 float* __restrict__ weight_sum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float w = 0.0f;
    bool keep = false;
    float4 e = make_float4(0, 0, 0, 0), d = e;
    if (i < n) e = recs[i].e;
    const uint32_t tag = __float_as_uint(e.w);
    if (i < n && (tag & TREC_GUIDED)) {
        d = recs[i].d;
        const float4 pre = recs[i].f;
        const float4 L = radiance[tag & ~TREC_GUIDED];
        const float Wt[3] = {d.y, d.z, d.w}, Et[3] = {e.x, e.y, e.z};
        float C[3] = {L.x - pre.x, L.y - pre.y, L.z - pre.z};
        for (int c = 0; c < 3; ++c) {
            C[c] = Wt[c] > 0.0f ? fmaxf(C[c], 0.0f) / Wt[c] : 0.0f;
            if (target == 0) C[c] = Et[c] > 0.0f ? C[c] / Et[c] : 0.0f;
        }
        const float lum = 0.2126f * C[0] + 0.7152f * C[1] + 0.0722f * C[2];
        w = target == 0 ? lum / d.x : lum;
        keep = d.x > 0.0f && w > 0.0f && isfinite(w);
    }
    // one counter atomic per block: wave counts and weight sums through LDS
    __shared__ uint32_t wcount[GF_BLOCK / 64];
    __shared__ float wsums[GF_BLOCK / 64];
    __shared__ uint32_t block_base;
    const int l = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(keep);
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << l) - 1ull));
    float wsum = keep ? w : 0.0f;
    for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off);
    if (l == 0) {
        wcount[wave] = (uint32_t)__popcll(mask);
        wsums[wave] = wsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        float ws = 0.0f;
        for (int k = 0; k < GF_BLOCK / 64; ++k) {
            const uint32_t c = wcount[k];
            wcount[k] = tot;
            tot += c;
            ws += wsums[k];
        }
        block_base = tot ? atomicAdd(count, tot) : 0u;
        if (tot) atomicAdd(weight_sum, ws);
    }
    __syncthreads();
    const uint32_t base = block_base + wcount[wave];
    if (keep) {
        float4* o = (float4*)&out[base + rank];
        o[0] = recs[i].a;
        o[1] = recs[i].b;
        o[2] = recs[i].c;
        o[3] = make_float4(w, 0.0f, 0.0f, 0.0f);
    }
}

}  // namespace xt
