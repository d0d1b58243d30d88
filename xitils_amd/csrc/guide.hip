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

// This lane's 24 input features of one query (layer-0 input fragment, k
// blocks 0..2): group g < 3 the frequency features of axis g, group 3 the
// shading normal and view direction.  Same arithmetic as guide_encode.
__device__ __forceinline__ void chain_features(const GuideBounds& gb, f3 p, f3 n, f3 wo, bf16x8 (&x)[3]) {
    const int g = __lane_id() >> 4;
    float v[24];
    if (g < 3) {
        const float PI = 3.14159265358979323846f;
        const float pa = g == 0 ? p.x : (g == 1 ? p.y : p.z);
        const float lo = g == 0 ? gb.lo.x : (g == 1 ? gb.lo.y : gb.lo.z);
        const float inv = g == 0 ? gb.inv_ext.x : (g == 1 ? gb.inv_ext.y : gb.inv_ext.z);
        const float ua = fminf(fmaxf((pa - lo) * inv, 0.0f), 1.0f);
        float sn = sinf(PI * ua), cs = cosf(PI * ua);
#pragma unroll
        for (int k = 0; k < G_BANDS; ++k) {
            v[2 * k] = sn;
            v[2 * k + 1] = cs;
            const float s2 = 2.0f * sn * cs, c2 = cs * cs - sn * sn;
            sn = s2;
            cs = c2;
        }
    } else {
        v[0] = n.x; v[1] = n.y; v[2] = n.z;
        v[3] = wo.x; v[4] = wo.y; v[5] = wo.z;
#pragma unroll
        for (int k = 6; k < 24; ++k) v[k] = 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 8; ++j) x[c][j] = (__bf16)v[8 * c + j];
}

// relu(acc) of two adjacent 16-unit output tiles -> one input fragment
__device__ __forceinline__ bf16x8 relu_pack(floatx4 a, floatx4 b) {
    bf16x8 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h[i] = (__bf16)fmaxf(a[i], 0.0f);
        h[4 + i] = (__bf16)fmaxf(b[i], 0.0f);
    }
    return h;
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
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bf16x8 a = frag(s.w0, GC_LD0, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, x[t][c], acc[t][b]);
        }
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
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const bf16x8 a = frag(wl[layer], GC_LD, 16 * b, 32 * c);
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][b] = mfma(a, h[t][c], acc[t][b]);
            }
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
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bf16x8 a = frag(s.w3, GC_LD, 16 * b, 32 * c);
#pragma unroll
            for (int t = 0; t < NT; ++t) y[t][b] = mfma(a, h[t][c], y[t][b]);
        }
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
__global__ __launch_bounds__(256) void k_guide_infer(GuideNet net, GuideBounds gb, DevScene sc, PathQueue pq,
                                                     const uint32_t* __restrict__ count, const float4* __restrict__ hits,
                                                     uint2* __restrict__ lobes_mk, float2* __restrict__ lobes2,
                                                     uint32_t stride, uint32_t max_vertex) {
    __shared__ __attribute__((aligned(16))) GuideChain s;
    load_chain(net.chain, s);
    const uint32_t n = *count;
    const int l = __lane_id(), g = l >> 4, r = l & 15;
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
        bf16x8 x[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int src = 16 * t + r;
            chain_features(gb, mk(__shfl(P.x, src), __shfl(P.y, src), __shfl(P.z, src)),
                           mk(__shfl(N.x, src), __shfl(N.y, src), __shfl(N.z, src)),
                           mk(__shfl(WO.x, src), __shfl(WO.y, src), __shfl(WO.z, src)), x[t]);
        }
        floatx4 y[4][3];
        chain_mlp<4>(s, x, y);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int src = 16 * t + r;
            const bool useq = __shfl((int)use, src) != 0;
            // softmax over the 8 alpha logits held by the 4 lanes of this query
            const float la = out_slot(y[t], 4), lb = out_slot(y[t], 9);
            float mx = fmaxf(la, lb);
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float ea = expf(la - mx), eb = expf(lb - mx);
            float sum = ea + eb;
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            if (useq) {
                const uint32_t q = base + src;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const f3 m = mk(out_slot(y[t], 5 * k), out_slot(y[t], 5 * k + 1), out_slot(y[t], 5 * k + 2));
                    const float len = length(m);
                    const f3 mu = len > 1e-12f ? div(m, len) : mk(0.0f, 0.0f, 1.0f);
                    const float kappa = expf(fminf(fmaxf(out_slot(y[t], 5 * k + 3), G_KLOG_MIN), G_KLOG_MAX));
                    const float alpha = (k == 0 ? ea : eb) / sum;
                    const size_t o = (size_t)(2 * g + k) * stride + q;
                    const uint2 mk2 = make_uint2(pack_half2(mu.x, mu.y), pack_half2(mu.z, kappa));
                    lobes_mk[o] = mk2;
                    lobes2[o] = make_float2(alpha, alpha * vmf_norm(half_hi(mk2.y)));
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

// Per-wave staging of T rows in LDS: [row][32 samples] bf16 with an 80-byte
// row stride (16-byte aligned rows; the four lane groups of a store land 16
// banks apart).  Lanes store single values at immediate offsets, then the
// wave copies whole 64-byte rows to T with 16-byte stores.
constexpr int GS_STRIDE = 40;                 // bf16 per staged row
constexpr int GS_ROWS = 96;                   // rows per flush
struct GradStage {
    __bf16 v[GS_ROWS * GS_STRIDE];
};
// this lane's units of NB output tiles: row 16b + 4g + i, sample column col (0..31)
template <int NB>
__device__ __forceinline__ void stage_tiles(GradStage& st, int col, const floatx4 (&v)[NB]) {
    const int g = __lane_id() >> 4;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) st.v[(16 * b + 4 * g + i) * GS_STRIDE + col] = (__bf16)v[b][i];
}
// one input fragment set: k position 32c + 8g + j holds unit unit_of_kpos(...)
__device__ __forceinline__ void stage_frags(GradStage& st, int col, const bf16x8 (&h)[2]) {
    const int g = __lane_id() >> 4;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 8; ++j) st.v[unit_of_kpos(32 * c + 8 * g + j) * GS_STRIDE + col] = h[c][j];
}
// rows [0, R) of the stage -> T[row0 + row][base .. base + 31]
template <int R>
__device__ __forceinline__ void flush_stage(GradStage& st, __bf16* T, uint32_t ld, int row0, uint32_t base) {
    wave_sync();
    const int l = __lane_id();
#pragma unroll
    for (int k = l; k < 4 * R; k += 64) {
        const int row = k >> 2, part = k & 3;
        const uint4 v = *(const uint4*)&st.v[row * GS_STRIDE + 8 * part];
        *(uint4*)&T[(size_t)(row0 + row) * ld + base + 8 * part] = v;
    }
    wave_sync();
}

// Per-sample part of the gradient of sum_i -min(w_i * weight_scale,
// weight_clip) log q(dir_i | query_i): forward and backward chains with
// activations in registers, 32 samples per wave pass; writes the features,
// hidden activations and pre-activation gradients of every sample to T (see
// GT_*) for the weight-gradient products (k_guide_wgrad).  Samples >= n are
// written as zeros up to the next multiple of 32.
__global__ __launch_bounds__(GB_BLOCK) void k_guide_backprop(GuideNet net, GuideBounds gb,
                                                             const GuideSample* __restrict__ smp, uint32_t n,
                                                             float weight_scale, float weight_clip, __bf16* __restrict__ T,
                                                             uint32_t ld) {
    constexpr int NT = 2;
    __shared__ __attribute__((aligned(16))) GuideChain s;
    __shared__ __attribute__((aligned(16))) GuideBack bk;
    __shared__ __attribute__((aligned(16))) GradStage stage[GB_BLOCK / 64];
    {
        const uint4* src = (const uint4*)net.back;
        uint4* dst = (uint4*)&bk;
        for (int k = threadIdx.x; k < (int)(sizeof(GuideBack) / 16); k += blockDim.x) dst[k] = src[k];
    }
    load_chain(net.chain, s);   // includes the block barrier
    GradStage& st = stage[threadIdx.x >> 6];
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    const uint32_t npad = (n + 31u) & ~31u;
    for (uint32_t base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 32u; base < npad;
         base += nwaves * 32u) {
        // compiler-only fence: the weight fragments are re-read from LDS in each
        // pass instead of being hoisted into registers for the whole loop
        asm volatile("" ::: "memory");
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
            // features: k position 32c + 8g + j holds feature 24g + 8c + j (g < 3) or 72 + 8c + j (g = 3)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int f = g < 3 ? 24 * g + 8 * c + j : 72 + 8 * c + j;
                    st.v[f * GS_STRIDE + 16 * t + r] = x[t][c][j];
                }
        }
        flush_stage<G_IN>(st, T, ld, GT_X, base);
        bf16x8 h1[NT][2], h2[NT][2], h3[NT][2];
        floatx4 y[NT][3];
        chain_forward_keep<NT>(s, x, h1, h2, h3, y);
        // loss gradient: lane group g holds lobes 2g, 2g + 1 (slots 0..9) of sample 16t + r
        floatx4 dy[NT][3];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float la = out_slot(y[t], 4), lb = out_slot(y[t], 9);
            float mx = fmaxf(la, lb);
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float ea = expf(la - mx), eb = expf(lb - mx);
            float esum = ea + eb;
            esum += __shfl_xor(esum, 16);
            esum += __shfl_xor(esum, 32);
            VmfLobe lob[2];
            float len[2], v[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const f3 m = mk(out_slot(y[t], 5 * k), out_slot(y[t], 5 * k + 1), out_slot(y[t], 5 * k + 2));
                len[k] = length(m);
                lob[k].mu = len[k] > 1e-12f ? div(m, len[k]) : mk(0.0f, 0.0f, 1.0f);
                lob[k].kappa = expf(fminf(fmaxf(out_slot(y[t], 5 * k + 3), G_KLOG_MIN), G_KLOG_MAX));
                lob[k].alpha = (k == 0 ? ea : eb) / esum;
                v[k] = vmf_lobe_pdf(lob[k], dir[t]);
            }
            float q = lob[0].alpha * v[0] + lob[1].alpha * v[1];
            q += __shfl_xor(q, 16);
            q += __shfl_xor(q, 32);
            q = fmaxf(q, 1e-30f);
            float d[12];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const float rsp = lob[k].alpha * v[k] / q;
                const float kap = lob[k].kappa;
                const f3 mu = lob[k].mu;
                const float mw = dot(mu, dir[t]);
                const f3 gm = (dir[t] - mw * mu) * (rsp * kap / fmaxf(len[k], 1e-12f));
                const float dk = 1.0f / kap - 2.0f / expm1f(2.0f * kap) + (mw - 1.0f);
                const float kl = out_slot(y[t], 5 * k + 3);
                const float gk = (kl > G_KLOG_MIN && kl < G_KLOG_MAX) ? rsp * dk * kap : 0.0f;
                const float ga = rsp - lob[k].alpha;
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
        // activations, then dy (rows GT_DY + raw output index, rounded as the backward chain reads it)
        const bf16x8 (*hs[3])[2] = {h1, h2, h3};
#pragma unroll
        for (int layer = 0; layer < 3; ++layer) {
#pragma unroll
            for (int t = 0; t < NT; ++t) stage_frags(st, 16 * t + r, hs[layer][t]);
            flush_stage<G_HID>(st, T, ld, GT_H1 + G_HID * layer, base);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int k = 0; k < 10; ++k) st.v[(10 * g + k) * GS_STRIDE + 16 * t + r] = (__bf16)out_slot(dy[t], k);
        flush_stage<G_OUT>(st, T, ld, GT_DY, base);
        // backward chain: dz3 = (W3^T dy) * [h3 > 0], dz2 = (W2^T dz3) * [h2 > 0], dz1 = (W1^T dz2) * [h1 > 0]
        bf16x8 z[NT][2];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            z[t][0] = grad_pack(dy[t][0], dy[t][1]);
            z[t][1] = grad_pack(dy[t][2], floatx4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        floatx4 dz[NT][4];
        back_step<NT>(bk.w3, z, h3, dz);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            stage_tiles<4>(st, 16 * t + r, dz[t]);
            z[t][0] = grad_pack(dz[t][0], dz[t][1]);
            z[t][1] = grad_pack(dz[t][2], dz[t][3]);
        }
        flush_stage<G_HID>(st, T, ld, GT_D3, base);
        back_step<NT>(bk.w2, z, h2, dz);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            stage_tiles<4>(st, 16 * t + r, dz[t]);
            z[t][0] = grad_pack(dz[t][0], dz[t][1]);
            z[t][1] = grad_pack(dz[t][2], dz[t][3]);
        }
        flush_stage<G_HID>(st, T, ld, GT_D2, base);
        back_step<NT>(bk.w1, z, h1, dz);
#pragma unroll
        for (int t = 0; t < NT; ++t) stage_tiles<4>(st, 16 * t + r, dz[t]);
        flush_stage<G_HID>(st, T, ld, GT_D1, base);
    }
}

// ---------------------------------------------------------------- training: weight-gradient products

// One 16-row strip of a layer's gradient over samples [k0, k1): dW[m][n] =
// sum_k D[m][k] A[n][k] (D = dz rows, A = activation rows of T), db[m] =
// sum_k D[m][k]; written to row[...] of this chunk.
template <int IN>
__device__ __forceinline__ void wgrad_strip(const __bf16* T, uint32_t ld, int drow, int arow, uint32_t k0, uint32_t k1,
                                            float* row, int off_w, int off_b, int rows_valid) {
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    floatx4 acc[IN / 16];
#pragma unroll
    for (int c = 0; c < IN / 16; ++c) acc[c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (uint32_t k = k0; k < k1; k += 32) {
        const bf16x8 a = frag(T + (size_t)drow * ld, (int)ld, 0, (int)k);
#pragma unroll
        for (int j = 0; j < 8; ++j) bsum += (float)a[j];
#pragma unroll
        for (int c = 0; c < IN / 16; ++c) acc[c] = mfma(a, frag(T + (size_t)(arow + 16 * c) * ld, (int)ld, 0, (int)k), acc[c]);
    }
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
#pragma unroll
    for (int c = 0; c < IN / 16; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 4 * g + i;
            if (m < rows_valid) row[off_w + m * IN + 16 * c + r] = acc[c][i];
        }
    if (g == 0 && r < rows_valid) row[off_b + r] = bsum;
}

// Weight and bias gradients of one chunk of `chunk` samples (block x) from
// T: 15 strips of 16 output rows (4 per hidden layer, 3 for the output
// layer) spread over the block's 4 waves; one partial row per block, summed
// by k_guide_adam.
__global__ __launch_bounds__(GW_BLOCK) void k_guide_wgrad(const __bf16* __restrict__ T, uint32_t ld, uint32_t n,
                                                          uint32_t chunk, float* __restrict__ partial) {
    const uint32_t npad = (n + 31u) & ~31u;
    const uint32_t k0 = blockIdx.x * chunk, k1 = min(npad, k0 + chunk);
    float* row = partial + (size_t)blockIdx.x * G_NPARAMS;
    const int wave = threadIdx.x >> 6;
    for (int strip = wave; strip < 15; strip += GW_BLOCK / 64) {
        const int layer = strip >> 2, sb = strip & 3, m0 = 16 * sb;
        if (layer == 0)
            wgrad_strip<G_IN>(T, ld, GT_D1 + m0, GT_X, k0, k1, row, G_OFF_W0 + m0 * G_IN, G_OFF_B0 + m0, 16);
        else if (layer == 1)
            wgrad_strip<G_HID>(T, ld, GT_D2 + m0, GT_H1, k0, k1, row, G_OFF_W1 + m0 * G_HID, G_OFF_B1 + m0, 16);
        else if (layer == 2)
            wgrad_strip<G_HID>(T, ld, GT_D3 + m0, GT_H2, k0, k1, row, G_OFF_W2 + m0 * G_HID, G_OFF_B2 + m0, 16);
        else
            wgrad_strip<G_HID>(T, ld, GT_DY + m0, GT_H3, k0, k1, row, G_OFF_W3 + m0 * G_HID, G_OFF_B3 + m0, 16);
    }
}

// Adam (Kingma & Ba) on the fp32 master parameters: the gradient is the sum
// of nparts partial rows (GA_GROUPS interleaved row groups per parameter,
// then the groups in order: deterministic) times grad_scale; b1t, b2t =
// beta^t.  32 parameters per block, one row group per wave half, so a
// 256-row sum is 8 loads per thread with 16 waves per block in flight.
__global__ __launch_bounds__(GA_BLOCK) void k_guide_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                                                         const float* __restrict__ partial, uint32_t nparts, float lr,
                                                         float b1, float b2, float eps, float b1t, float b2t,
                                                         float grad_scale) {
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
    P[i] -= lr * mh / (sqrtf(vh) + eps);
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
