// pe_guide.h — the guiding network on gfx950: input encoding, the 8-lobe von Mises-Fisher mixture (density,
// sampling, parameters from the raw outputs) and the fused MFMA inference kernel shared by pe_guide_infer and
// the renderer.
//
// Network: 160 inputs (six coordinates x 12 frequencies x sin/cos, the six raw values, zero padding) -> four
// hidden layers of 64 with ReLU -> 64 outputs, of which 5 per lobe are used (mean xyz, log concentration,
// weight logit). Parameters per layer: W[out][in] then b[out]; bf16 copies of W feed the MFMAs
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation), biases stay fp32.
//
// MFMA operand layout of v_mfma_f32_16x16x32_bf16 on a 64-lane wave (lane l, g = l / 16), checked on the
// device: A (16 x 32, row-major M x K): lane holds A[l % 16][8g .. 8g + 7]; B (32 x 16, K x N): lane holds
// B[8g .. 8g + 7][l % 16]; D (16 x 16 fp32): lane holds D[4g + r][l % 16], r = 0..3.
#pragma once

#include "photon_hip.h"
#include "pe_math.h"

#include <hip/hip_runtime.h>

namespace pe {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kGuideIn = PE_GUIDE_INPUTS;
constexpr int kGuideHid = PE_GUIDE_HIDDEN;
constexpr int kGuideOut = PE_GUIDE_OUTPUTS;
constexpr int kGuideLayers = PE_GUIDE_LAYERS;
constexpr int kGuideLobes = PE_GUIDE_LOBES;
constexpr int kGuideParams = PE_GUIDE_PARAMS;
constexpr int kGuideFreqs = 12;
constexpr float kGuideMinLogK = -4.0f, kGuideMaxLogK = 7.0f;

__host__ __device__ constexpr int guide_layer_in(int l) { return l == 0 ? kGuideIn : kGuideHid; }
__host__ __device__ constexpr int guide_layer_out(int l) { return l == kGuideLayers - 1 ? kGuideOut : kGuideHid; }
// Offset of layer l's W in the parameter vector (b follows W).
__host__ __device__ constexpr int guide_w_offset(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += guide_layer_out(i) * guide_layer_in(i) + guide_layer_out(i);
    return off;
}
__host__ __device__ constexpr int guide_b_offset(int l) { return guide_w_offset(l) + guide_layer_out(l) * guide_layer_in(l); }
// Offset of layer l's W in the bf16 weight copies (weights only).
__host__ __device__ constexpr int guide_wbf_offset(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += guide_layer_out(i) * guide_layer_in(i);
    return off;
}
constexpr int kGuideWeights = guide_wbf_offset(kGuideLayers);
static_assert(guide_w_offset(kGuideLayers) == kGuideParams, "parameter count");

// Per-query mixture as the renderer stores it: 8 x (mean xyz, concentration), the selection CDF (8 floats)
// and the density coefficients weight_j * kappa_j / (2 pi (1 - e^(-2 kappa_j))) (8 floats).
constexpr int kGuideMixFloat4 = 12;

// What the renderer and the inference kernel need to read the network.
struct GuideView {
    const __bf16* wbf;     // bf16 W of every layer, [out][in]
    const float* params;   // fp32 master parameters (biases read from here)
    V3 lo, inv_extent;     // position normalisation to [0, 1]
    float fraction;        // probability of sampling the mixture at a guided vertex (0: plain sampling)
    unsigned long long* inferred;  // optional: queries inferred (added once per wave)
};

// Input coordinate c of a query: normalised position (0..2) or outgoing direction mapped to [0, 1] (3..5).
__device__ __forceinline__ float guide_coord(const GuideView& g, V3 p, V3 d, int c) {
    switch (c) {
        case 0: return (p.x - g.lo.x) * g.inv_extent.x;
        case 1: return (p.y - g.lo.y) * g.inv_extent.y;
        case 2: return (p.z - g.lo.z) * g.inv_extent.z;
        case 3: return 0.5f * (d.x + 1.0f);
        case 4: return 0.5f * (d.y + 1.0f);
        default: return 0.5f * (d.z + 1.0f);
    }
}

// Features f0 .. f0 + 7 of the encoding (f0 a multiple of 8). Feature f < 144 is the sine (even f) or cosine
// (odd f) of 2^k pi v_c with c = f / 24 and k = (f % 24) / 2; features 144..149 are the raw coordinates; the
// rest is zero. Eight consecutive features are four consecutive frequencies of one coordinate: the first pair
// is evaluated in revolutions of the exact fractional part of 2^(k-1) v (v_sin_f32 / v_cos_f32), the next three
// by the double-angle identities. Training and inference both use this function.
__device__ __forceinline__ void guide_features8(const float v[6], int f0, float out[8]) {
    if (f0 < 6 * 2 * kGuideFreqs) {
        const int pair = f0 >> 1, c = pair / kGuideFreqs, k0 = pair % kGuideFreqs;
        const float r = v[c] * (static_cast<float>(1 << k0) * 0.5f);  // exact: a power of two
        const float fr = r - floorf(r);
        float sn = __builtin_amdgcn_sinf(fr), cs = __builtin_amdgcn_cosf(fr);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            out[2 * j] = sn;
            out[2 * j + 1] = cs;
            const float s2 = 2.0f * sn * cs, c2 = 1.0f - 2.0f * sn * sn;
            sn = s2;
            cs = c2;
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int f = f0 + e - 6 * 2 * kGuideFreqs;
        out[e] = f < 6 ? v[f < 0 ? 0 : f] : 0.0f;
    }
}

// 1 - e^(-2k) for k >= e^-4 (the smallest concentration): a short series below 2k = 0.05, exp above.
__device__ __forceinline__ float expm1_neg2k(float k) {
    const float x = 2.0f * k;
    if (x < 0.05f) return x * (1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f))));
    return 1.0f - __expf(-x);
}

// The stored mixture of one query, written by four lanes: part p holds lobes 2p and 2p + 1 (mean xyz,
// concentration), their entries of the selection CDF and their density coefficients
// weight_j * kappa_j / (2 pi (1 - e^(-2 kappa_j))).
__device__ __forceinline__ void guide_store_part(float4* mix, int part, float4 a, float4 b, float2 cdf, float2 coef) {
    mix[2 * part] = a;
    mix[2 * part + 1] = b;
    reinterpret_cast<float2*>(mix + 8)[part] = cdf;
    reinterpret_cast<float2*>(mix + 10)[part] = coef;
}

// Density of the stored mixture at unit direction d (per steradian).
__device__ __forceinline__ float guide_pdf(const float4* mix, V3 d) {
    const float4 c0 = mix[10], c1 = mix[11];
    const float coef[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    float q = 0.0f;
    for (int j = 0; j < kGuideLobes; ++j) {
        const float4 m = mix[j];
        q += coef[j] * expf(m.w * (m.x * d.x + m.y * d.y + m.z * d.z - 1.0f));
    }
    return q;
}

// A direction from the stored mixture: lobe by `us` against the selection CDF, then the vMF lobe by
// inversion of its cosine distribution (w = 1 + log(u + (1 - u) e^(-2 kappa)) / kappa) and a uniform azimuth.
__device__ __forceinline__ V3 guide_sample(const float4* mix, float us, float u0, float u1) {
    const float4 a = mix[8], b = mix[9];
    const float cdf[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    int j = 0;
    while (j < kGuideLobes - 1 && us >= cdf[j]) ++j;
    const float4 m = mix[j];
    const float kappa = m.w;
    const float w = clampf(1.0f + logf(u0 + (1.0f - u0) * expf(-2.0f * kappa)) / kappa, -1.0f, 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float phi = kTwoPi * u1;
    const Frame f = frame_from_normal(mk(m.x, m.y, m.z));
    return normalized(to_world(f, mk(s * cosf(phi), s * sinf(phi), w)));
}

// ---------------------------------------------------------------------------------------- fused inference
constexpr int kInferBlock = 256;       // 4 waves, each on tiles of 32 queries
constexpr int kActStride = 72;         // bf16 per activation row in LDS (64 + 8: conflict-free 16-byte reads)
constexpr int kOutStride = 65;         // fp32 per output row in LDS (16 rows at a time)
constexpr int kWaveScratchBytes = 32 * kActStride * 2;  // the activation tile; the output rows share it
static_assert(16 * kOutStride * 4 <= kWaveScratchBytes, "16 output rows fit the wave's scratch");
// Weights in LDS: W[out][in] rows padded by 8 bf16, so that the 16 rows a B-operand read touches start on 16
// different bank groups (a dense row of 64 or 160 bf16 would put 8 or 4 of them on the same banks).
constexpr int kWeightPad = 8;
__host__ __device__ constexpr int guide_lds_w_offset(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += guide_layer_out(i) * (guide_layer_in(i) + kWeightPad);
    return off;
}
constexpr int kGuideLdsWeights = guide_lds_w_offset(kGuideLayers);
constexpr int kInferLds = kGuideLdsWeights * 2 + (kInferBlock / 64) * kWaveScratchBytes;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One hidden or output layer for a wave's two 16-row tiles. A operands (activations) come from `act` (row
// stride kActStride bf16), B operands (W[out][in], in-major) from the block's LDS copy.
template <int K, int N>
__device__ __forceinline__ void guide_layer_mfma(const bf16x8 (&a)[2][K / 32], const __bf16* w, const float (&bias)[N / 16],
                                                 f32x4 (&acc)[2][N / 16]) {
    const int l = __lane_id(), g = l >> 4, c = l & 15;
#pragma unroll
    for (int t = 0; t < N / 16; ++t) {
        // The accumulators start at the bias of their column (every D element of a lane is in column 16t + c).
        acc[0][t] = f32x4{bias[t], bias[t], bias[t], bias[t]};
        acc[1][t] = acc[0][t];
#pragma unroll
        for (int s = 0; s < K / 32; ++s) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(w + (16 * t + c) * (K + kWeightPad) + 32 * s + 8 * g);
            acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][s], b, acc[0][t], 0, 0, 0);
            acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][s], b, acc[1][t], 0, 0, 0);
        }
    }
}

// Mixture parameters from the raw outputs o[5j .. 5j + 4]: mean direction (normalised; +z when degenerate),
// concentration exp(clamp(o3, -4, 7)), softmax weights over the eight lobes.
// Q supplies queries: uint32_t count(n) (how many of the n slots are live), bool fetch(i, V3& pos, V3& dir)
// (false: no inference needed) and store_part(i, part, lobe, lobe, cdf, coef) (see guide_store_part).
template <class Q>
__global__ void __launch_bounds__(kInferBlock) k_guide_infer(GuideView gv, Q q, uint32_t n) {
    __shared__ __align__(16) unsigned char lds_raw[kInferLds];
    __bf16* wl = reinterpret_cast<__bf16*>(lds_raw);
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; ++layer) {  // 16-byte chunks, into padded rows
        const int in = guide_layer_in(layer), chunks = guide_layer_out(layer) * in / 8;
        const bf16x8* src = reinterpret_cast<const bf16x8*>(gv.wbf + guide_wbf_offset(layer));
        for (int i = threadIdx.x; i < chunks; i += blockDim.x) {
            const int row = i / (in / 8), col = i % (in / 8);
            *reinterpret_cast<bf16x8*>(wl + guide_lds_w_offset(layer) + row * (in + kWeightPad) + 8 * col) = src[i];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, c = l & 15;
    unsigned char* scratch = lds_raw + kGuideLdsWeights * 2 + wave * kWaveScratchBytes;
    __bf16* act = reinterpret_cast<__bf16*>(scratch);
    float* out = reinterpret_cast<float*>(scratch);
    // This lane's bias of each layer and output tile (column 16 t + c).
    float bias[kGuideLayers][kGuideHid / 16];
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; ++layer)
#pragma unroll
        for (int t = 0; t < kGuideHid / 16; ++t) bias[layer][t] = gv.params[guide_b_offset(layer) + 16 * t + c];
    n = q.count(n);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t tiles = (n + 31u) >> 5;
    for (uint32_t tile = blockIdx.x * (blockDim.x >> 6) + wave; tile < tiles; tile += waves) {
        // Layer 0: this lane's 8 features of each k-step for rows c and 16 + c, built one k-step at a time
        // straight into A operands.
        float v[2][6];
        bool ok[2];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
            const uint32_t i = tile * 32u + static_cast<uint32_t>(16 * rt + c);
            V3 p = mk(0.0f, 0.0f, 0.0f), d = mk(0.0f, 0.0f, 1.0f);
            ok[rt] = i < n && q.fetch(i, p, d);
            for (int k = 0; k < 6; ++k) v[rt][k] = guide_coord(gv, p, d, k);
        }
        const uint64_t live0 = __ballot(ok[0]), live1 = __ballot(ok[1]);
        if ((live0 | live1) == 0) continue;
        if (gv.inferred && l == 0)  // each query row is held by four lanes
            atomicAdd(gv.inferred, static_cast<unsigned long long>((__popcll(live0) + __popcll(live1)) >> 2));
        f32x4 acc[2][kGuideHid / 16];
#pragma unroll
        for (int t = 0; t < kGuideHid / 16; ++t) acc[0][t] = acc[1][t] = f32x4{bias[0][t], bias[0][t], bias[0][t], bias[0][t]};
        for (int s = 0; s < kGuideIn / 32; ++s) {
            bf16x8 a0[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                float f8[8];
                guide_features8(v[rt], 32 * s + 8 * g, f8);
#pragma unroll
                for (int e = 0; e < 8; ++e) a0[rt][e] = static_cast<__bf16>(ok[rt] ? f8[e] : 0.0f);
            }
#pragma unroll
            for (int t = 0; t < kGuideHid / 16; ++t) {
                const bf16x8 bw =
                    *reinterpret_cast<const bf16x8*>(wl + (16 * t + c) * (kGuideIn + kWeightPad) + 32 * s + 8 * g);
                acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[0], bw, acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[1], bw, acc[1][t], 0, 0, 0);
            }
        }
        const __bf16* w_layer = wl;
#pragma unroll
        for (int layer = 1; layer < kGuideLayers; ++layer) {
            // Epilogue of the previous layer: bias, ReLU, bf16, into the wave's activation tile.
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int t = 0; t < kGuideHid / 16; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float h = fmaxf(0.0f, acc[rt][t][r]);
                        act[(16 * rt + 4 * g + r) * kActStride + 16 * t + c] = static_cast<__bf16>(h);
                    }
            wave_lds_sync();
            bf16x8 a[2][kGuideHid / 32];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int s = 0; s < kGuideHid / 32; ++s)
                    a[rt][s] = *reinterpret_cast<const bf16x8*>(act + (16 * rt + c) * kActStride + 32 * s + 8 * g);
            wave_lds_sync();
            w_layer = wl + guide_lds_w_offset(layer);
            guide_layer_mfma<kGuideHid, kGuideHid>(a, w_layer, bias[layer], acc);  // kGuideOut == kGuideHid
        }
        // Output layer epilogue: bias, fp32 rows in LDS (one 16-row tile at a time); then four lanes per query
        // build its mixture, two lobes each (softmax and CDF across the four lanes by shuffles).
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
#pragma unroll
            for (int t = 0; t < kGuideOut / 16; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[(4 * g + r) * kOutStride + 16 * t + c] = acc[rt][t][r];
            wave_lds_sync();
            const int row = l >> 2, part = l & 3;
            const uint32_t i = tile * 32u + static_cast<uint32_t>(16 * rt + row);
            V3 p, d;
            const bool mine = i < n && q.fetch(i, p, d);
            const float* o = out + row * kOutStride;
            float4 lobe[2];
            float logit[2], coefk[2];
            // Hardware reciprocal, reciprocal square root and exp (about 1 ulp): the stored mixture is the
            // distribution the renderer samples and evaluates, whatever its last bits.
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = 2 * part + h;
                const V3 u = mk(o[5 * j], o[5 * j + 1], o[5 * j + 2]);
                const float len2 = dot(u, u);
                const V3 mu = len2 > 1e-12f ? u * __builtin_amdgcn_rsqf(len2) : mk(0.0f, 0.0f, 1.0f);
                const float kappa = __expf(clampf(o[5 * j + 3], kGuideMinLogK, kGuideMaxLogK));
                lobe[h] = make_float4(mu.x, mu.y, mu.z, kappa);
                logit[h] = o[5 * j + 4];
                coefk[h] = kappa * __builtin_amdgcn_rcpf(kTwoPi * expm1_neg2k(kappa));
            }
            float mx = fmaxf(logit[0], logit[1]);
            mx = fmaxf(mx, __shfl_xor(mx, 1));
            mx = fmaxf(mx, __shfl_xor(mx, 2));
            const float e0 = __expf(logit[0] - mx), e1 = __expf(logit[1] - mx);
            float sum = e0 + e1;
            sum += __shfl_xor(sum, 1);
            sum += __shfl_xor(sum, 2);
            const float inv = __builtin_amdgcn_rcpf(sum);
            const float w0 = e0 * inv, w1 = e1 * inv;
            // Exclusive prefix of the parts' weight sums (parts 0..3 of this query).
            const float pw = w0 + w1;
            const float n1 = __shfl_up(pw, 1), n2 = __shfl_up(pw, 2), n3 = __shfl_up(pw, 3);
            const float before = (part >= 1 ? n1 : 0.0f) + (part >= 2 ? n2 : 0.0f) + (part >= 3 ? n3 : 0.0f);
            const float c0 = before + w0;
            const float c1 = part == 3 ? 1.0f : before + pw;
            if (mine) q.store_part(i, part, lobe[0], lobe[1], make_float2(c0, c1), make_float2(w0 * coefk[0], w1 * coefk[1]));
            wave_lds_sync();
        }
    }
}

static_assert(kGuideOut == kGuideHid, "the output layer reuses the hidden-layer tile shape");

}  // namespace pe
