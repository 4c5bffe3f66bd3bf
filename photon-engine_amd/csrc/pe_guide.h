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
#include "pe_traverse.h"

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
// The network runs transposed: D = W X^T with the weights as the A operand (neurons are MFMA rows) and the
// queries as the B operand (MFMA columns), so a lane's accumulators after a layer are four consecutive neurons
// of one query per 16-neuron tile, which is what the same lane supplies as a B operand of the next layer once
// the next layer's input columns are permuted to match. Activations never leave the registers; LDS holds the
// weights (and biases) only. A wave works on kInferTiles tiles of 16 queries, so each A operand read from LDS
// feeds kInferTiles MFMAs.
//
// Physical orders in LDS. Input column k = 32 s + 8 g + e is element e of what lane group g supplies at MFMA
// k-step s. Layer 0: each group encodes fixed coordinates, so a lane builds its features without selecting
// among the six: groups 0..2 take position coordinate g at frequencies 4s .. 4s + 3 (s < 3) and direction
// coordinate g at frequencies 4(s - 3) .. 4(s - 3) + 3 (s >= 3); group 3 takes direction coordinate s at
// frequencies 8..11 (s < 3), the six raw coordinates and two zeros (s = 3) and zeros (s = 4). Layer l >= 1:
// logical input neuron 16 (2 s + e / 4) + 4 g + e % 4, since the lane of group g holds, for the query of its
// column, neurons 16 t + 4 g + r (tile t, row r), and its k-step s uses tiles 2 s and 2 s + 1. The output layer's
// rows are permuted so that lane group g holds lobes 2 g and 2 g + 1: physical row 16 t + 4 g + r with
// e = 4 t + r is logical output 10 g + e for e < 10 (lobe 2 g + e / 5, field e % 5) and 40 + 6 g + e - 10 (unused)
// otherwise. Every other order is the identity.
constexpr int kInferBlock = 256;       // 4 waves
constexpr int kInferTiles = 4;         // query tiles of 16 per wave iteration
constexpr int kInferQueries = 16 * kInferTiles;
// Weights in LDS: physical rows padded by 8 bf16, so that the 16 rows an A-operand read touches start on 16
// different bank groups (a dense row of 64 or 160 bf16 would put 8 or 4 of them on the same banks).
constexpr int kWeightPad = 8;
__host__ __device__ constexpr int guide_lds_w_offset(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += guide_layer_out(i) * (guide_layer_in(i) + kWeightPad);
    return off;
}
constexpr int kGuideLdsWeights = guide_lds_w_offset(kGuideLayers);
constexpr int kInferLds = kGuideLdsWeights * 2 + kGuideLayers * kGuideHid * 4;  // bf16 weights, fp32 biases

// Logical input neuron of physical input column k of layer l (see above).
__host__ __device__ constexpr int guide_phys_in(int l, int k) {
    const int s = k >> 5, g = (k >> 3) & 3, e = k & 7;
    if (l != 0) return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3);
    if (g < 3) return s < 3 ? 24 * g + 8 * s + e : 24 * (3 + g) + 8 * (s - 3) + e;
    return s < 3 ? 24 * (3 + s) + 16 + e : 144 + 8 * (s - 3) + e;
}

// The eight layer-0 features lane group g supplies at k-step S (physical order above). cp and cd are position
// and direction coordinate min(g, 2); v holds all six. The same arithmetic as guide_features8 (training).
template <int S>
__device__ __forceinline__ void guide_features_phys(const float v[6], float cp, float cd, int g, float out[8]) {
    const float base = S < 3 ? (g < 3 ? cp : v[3 + (S < 3 ? S : 0)]) : cd;
    const float scale = S < 3 ? (g < 3 ? static_cast<float>(1 << (4 * S)) * 0.5f : 128.0f)
                              : static_cast<float>(1 << (4 * (S - 3))) * 0.5f;
    const float r = base * scale;  // exact: a power of two
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
    if constexpr (S >= 3) {
#pragma unroll
        for (int e = 0; e < 8; ++e) out[e] = g < 3 ? out[e] : (S == 3 && e < 6 ? v[e] : 0.0f);
    }
}
// Logical output of physical row p of layer l (see above).
__host__ __device__ constexpr int guide_phys_out(int l, int p) {
    if (l != kGuideLayers - 1) return p;
    const int t = p >> 4, g = (p >> 2) & 3, r = p & 3, e = 4 * t + r;
    return e < 10 ? 10 * g + e : 40 + 6 * g + (e - 10);
}

static_assert(kGuideOut == kGuideHid, "the output layer has the hidden layers' tile shape");
static_assert(kGuideIn % 32 == 0 && kGuideHid == 64, "k-steps of 32, four 16-neuron tiles per layer");

// Layer-0 k-step S for the wave's query tiles: features into B operands, one A operand read per neuron tile.
template <int S>
__device__ __forceinline__ void guide_layer0_step(const float (&v)[kInferTiles][6], const float (&cp)[kInferTiles],
                                                  const float (&cd)[kInferTiles], int g, const __bf16* wl, int c,
                                                  f32x4 (&acc)[kInferTiles][kGuideHid / 16]) {
    bf16x8 xb[kInferTiles];
#pragma unroll
    for (int qt = 0; qt < kInferTiles; ++qt) {
        float f8[8];
        guide_features_phys<S>(v[qt], cp[qt], cd[qt], g, f8);
#pragma unroll
        for (int e = 0; e < 8; ++e) xb[qt][e] = static_cast<__bf16>(f8[e]);
    }
#pragma unroll
    for (int t = 0; t < kGuideHid / 16; ++t) {
        const bf16x8 w = *reinterpret_cast<const bf16x8*>(wl + (16 * t + c) * (kGuideIn + kWeightPad) + 32 * S + 8 * g);
#pragma unroll
        for (int qt = 0; qt < kInferTiles; ++qt)
            acc[qt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[qt], acc[qt][t], 0, 0, 0);
    }
}

// Q supplies queries: QueueSegs segs(n) (the segments of slots to infer; a wave works on segment w % segs, so the
// grid needs at least segs waves), bool fetch(i, V3& pos, V3& dir) (false: no inference needed) and
// store_part(i, part, lobe, lobe, cdf, coef) (see guide_store_part).
// Mixture parameters from the raw outputs o[5j .. 5j + 4]: mean direction (normalised; +z when degenerate),
// concentration exp(clamp(o3, -4, 7)), softmax weights over the eight lobes.
template <class Q>
__global__ void __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) k_guide_infer(GuideView gv, Q q, uint32_t n) {
    __shared__ __align__(16) unsigned char lds_raw[kInferLds];
    __bf16* wl = reinterpret_cast<__bf16*>(lds_raw);
    float* bl = reinterpret_cast<float*>(lds_raw + kGuideLdsWeights * 2);
    // Weights into their physical order, in runs of four bf16 (8 bytes): a run of four physical columns is four
    // consecutive logical ones in every layer.
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; ++layer) {
        const int in = guide_layer_in(layer), runs = guide_layer_out(layer) * in / 4;
        const __bf16* src = gv.wbf + guide_wbf_offset(layer);
        for (int i = threadIdx.x; i < runs; i += blockDim.x) {
            const int row = i / (in / 4), k = 4 * (i % (in / 4));
            const uint2 v = *reinterpret_cast<const uint2*>(src + guide_phys_out(layer, row) * in + guide_phys_in(layer, k));
            *reinterpret_cast<uint2*>(wl + guide_lds_w_offset(layer) + row * (in + kWeightPad) + k) = v;
        }
        for (int i = threadIdx.x; i < kGuideHid; i += blockDim.x)
            bl[layer * kGuideHid + i] = gv.params[guide_b_offset(layer) + guide_phys_out(layer, i)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, c = l & 15;
    const QueueSegs qs = q.segs(n);
    const SegShare sh = seg_share(blockIdx.x * (blockDim.x >> 6) + wave, gridDim.x * (blockDim.x >> 6), qs.segs);
    // The segment's queries are items [first, count) (a source may know that the rest are never guided).
    const uint32_t first = q.begin(sh.seg), base = sh.seg * qs.seg_cap + first;
    const uint32_t nk = qs.count(sh.seg) > first ? qs.count(sh.seg) - first : 0u;
    const uint32_t tiles = (nk + kInferQueries - 1u) / kInferQueries;
    uint32_t inferred = 0;  // live queries of this wave's tiles, added to the counter once at the end
    for (uint32_t tile = sh.rank; tile < tiles; tile += sh.workers) {
        float v[kInferTiles][6];
        bool ok[kInferTiles];
        bool any_ok = false;
#pragma unroll
        for (int qt = 0; qt < kInferTiles; ++qt) {
            const uint32_t j = tile * kInferQueries + static_cast<uint32_t>(16 * qt + c);
            V3 p = mk(0.0f, 0.0f, 0.0f), d = mk(0.0f, 0.0f, 1.0f);
            ok[qt] = j < nk && q.fetch(base + j, p, d);
            for (int k = 0; k < 6; ++k) v[qt][k] = guide_coord(gv, p, d, k);
            any_ok = any_ok || ok[qt];
        }
        if (__ballot(any_ok) == 0) continue;
#pragma unroll
        for (int qt = 0; qt < kInferTiles; ++qt) inferred += __popcll(__ballot(ok[qt]) & 0xFFFFull);  // group 0's lanes
        f32x4 acc[kInferTiles][kGuideHid / 16];
        // Layer 0: the features of k-step s for this lane's queries, built straight into B operands.
#pragma unroll
        for (int t = 0; t < kGuideHid / 16; ++t) {
            const float4 b = *reinterpret_cast<const float4*>(bl + 16 * t + 4 * g);
#pragma unroll
            for (int qt = 0; qt < kInferTiles; ++qt) acc[qt][t] = f32x4{b.x, b.y, b.z, b.w};
        }
        float cp[kInferTiles], cd[kInferTiles];
#pragma unroll
        for (int qt = 0; qt < kInferTiles; ++qt) {
            cp[qt] = g == 0 ? v[qt][0] : g == 1 ? v[qt][1] : v[qt][2];
            cd[qt] = g == 0 ? v[qt][3] : g == 1 ? v[qt][4] : v[qt][5];
        }
        guide_layer0_step<0>(v, cp, cd, g, wl, c, acc);
        guide_layer0_step<1>(v, cp, cd, g, wl, c, acc);
        guide_layer0_step<2>(v, cp, cd, g, wl, c, acc);
        guide_layer0_step<3>(v, cp, cd, g, wl, c, acc);
        guide_layer0_step<4>(v, cp, cd, g, wl, c, acc);
#pragma unroll
        for (int layer = 1; layer < kGuideLayers; ++layer) {
            // ReLU and bf16 of the previous layer, in the order the permuted input columns expect. ReLU as an
            // integer max of the bit pattern (a negative float is a negative integer): one instruction, where
            // fmaxf adds a NaN-quieting one.
            bf16x8 xb[kInferTiles][kGuideHid / 32];
#pragma unroll
            for (int s = 0; s < kGuideHid / 32; ++s)
#pragma unroll
                for (int qt = 0; qt < kInferTiles; ++qt)
#pragma unroll
                    for (int e = 0; e < 8; ++e) xb[qt][s][e] = static_cast<__bf16>(
                        __int_as_float(max(__float_as_int(acc[qt][2 * s + (e >> 2)][e & 3]), 0)));
            const __bf16* w_layer = wl + guide_lds_w_offset(layer);
            const float* b_layer = bl + layer * kGuideHid;
#pragma unroll
            for (int t = 0; t < kGuideHid / 16; ++t) {
                const float4 b = *reinterpret_cast<const float4*>(b_layer + 16 * t + 4 * g);
#pragma unroll
                for (int qt = 0; qt < kInferTiles; ++qt) acc[qt][t] = f32x4{b.x, b.y, b.z, b.w};
#pragma unroll
                for (int s = 0; s < kGuideHid / 32; ++s) {
                    const bf16x8 w = *reinterpret_cast<const bf16x8*>(w_layer + (16 * t + c) * (kGuideHid + kWeightPad) + 32 * s + 8 * g);
#pragma unroll
                    for (int qt = 0; qt < kInferTiles; ++qt)
                        acc[qt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[qt][s], acc[qt][t], 0, 0, 0);
                }
            }
        }
        // Output epilogue: the four lanes of groups 0..3 in column c hold one query's lobes 2g and 2g + 1 (raw
        // outputs e = 0..9 at tile e / 4, row e % 4); softmax and CDF across the four lanes by shuffles.
        // Hardware reciprocal, reciprocal square root and exp (about 1 ulp): the stored mixture is the
        // distribution the renderer samples and evaluates, whatever its last bits.
#pragma unroll
        for (int qt = 0; qt < kInferTiles; ++qt) {
            float o[10];
#pragma unroll
            for (int e = 0; e < 10; ++e) o[e] = acc[qt][e >> 2][e & 3];
            float4 lobe[2];
            float logit[2], coefk[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const V3 u = mk(o[5 * h], o[5 * h + 1], o[5 * h + 2]);
                const float len2 = dot(u, u);
                const V3 mu = len2 > 1e-12f ? u * __builtin_amdgcn_rsqf(len2) : mk(0.0f, 0.0f, 1.0f);
                const float kappa = __expf(clampf(o[5 * h + 3], kGuideMinLogK, kGuideMaxLogK));
                lobe[h] = make_float4(mu.x, mu.y, mu.z, kappa);
                logit[h] = o[5 * h + 4];
                coefk[h] = kappa * __builtin_amdgcn_rcpf(kTwoPi * expm1_neg2k(kappa));
            }
            float mx = fmaxf(logit[0], logit[1]);
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float e0 = __expf(logit[0] - mx), e1 = __expf(logit[1] - mx);
            float sum = e0 + e1;
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            const float inv = __builtin_amdgcn_rcpf(sum);
            const float w0 = e0 * inv, w1 = e1 * inv;
            // Exclusive prefix of the groups' weight sums (groups 0..3 of this query's column).
            const float pw = w0 + w1;
            const float n1 = __shfl_up(pw, 16), n2 = __shfl_up(pw, 32), n3 = __shfl_up(pw, 48);
            const float before = (g >= 1 ? n1 : 0.0f) + (g >= 2 ? n2 : 0.0f) + (g >= 3 ? n3 : 0.0f);
            const float c0 = before + w0;
            const float c1 = g == 3 ? 1.0f : before + pw;
            const uint32_t i = base + tile * kInferQueries + static_cast<uint32_t>(16 * qt + c);
            if (ok[qt]) q.store_part(i, g, lobe[0], lobe[1], make_float2(c0, c1), make_float2(w0 * coefk[0], w1 * coefk[1]));
        }
    }
    if (gv.inferred && l == 0 && inferred) atomicAdd(gv.inferred, static_cast<unsigned long long>(inferred));
}

}  // namespace pe
