// guide.hip — training and inference of the guiding network (pe_guide_*): input encoding, MFMA GEMM kernels
// for the forward pass, the input gradients and the weight gradients, the mixture log-likelihood loss, Adam.
//
// A training step on M records (M a multiple of 256; rows past the records are zero and carry no gradient):
//   encode     X0[M][160] bf16 from (position, outgoing direction)
//   forward    H1..H4 = ReLU(H W^T + b) (bf16), O = H4 W4^T + b4 (fp32)       k_gemm_rows (MFMA)
//   loss       dO = d(-sum_i w_i log q_i / sum_i w_i) / dO                    k_mixture_loss
//   backward   dW_l = dH_l^T H_(l-1) (per-block partials, then a fixed-order sum)  k_wgrad (MFMA) + k_reduce
//              dH_(l-1) = (dH_l W_l) * [H_(l-1) > 0]                         k_gemm_rows (MFMA) on W^T
//   Adam       fp32 masters and moments; refreshes the bf16 W and W^T copies   k_adam
// The renderer's inference is the fused kernel of pe_guide.h (weights in LDS, activations never leave LDS).
#include "photon_hip.h"
#include "pe_guide_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace pe {
namespace {

constexpr int kGemmBlock = 256;
constexpr int kWgradBlocks = 256;
constexpr int kWgradChunk = 32;      // rows per MFMA k-step of the weight gradient
constexpr int kWgradPad = 8;         // bf16 padding of the transposed LDS rows
constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;

enum Epilogue { kEpiReluBf16 = 0, kEpiF32 = 1, kEpiMaskBf16 = 2 };

// Y[m][n] = epilogue(sum_k X[m][k] W[n][k]): one wave per 16-row tile, all N columns. A and B operands are
// 16-byte loads from global memory (W is a few tens of KB and stays in L1/L2).
template <int K, int N, int kEpi>
__global__ void __launch_bounds__(kGemmBlock) k_gemm_rows(const __bf16* __restrict__ X, const __bf16* __restrict__ W,
                                                          const float* __restrict__ bias, const __bf16* __restrict__ mask,
                                                          void* __restrict__ Y, uint32_t M) {
    const int l = __lane_id(), g = l >> 4, c = l & 15;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < M / 16; tile += waves) {
        const uint32_t row = tile * 16;
        bf16x8 a[K / 32];
#pragma unroll
        for (int s = 0; s < K / 32; ++s)
            a[s] = *reinterpret_cast<const bf16x8*>(X + static_cast<size_t>(row + c) * K + 32 * s + 8 * g);
#pragma unroll
        for (int t = 0; t < N / 16; ++t) {
            f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < K / 32; ++s) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (16 * t + c) * K + 32 * s + 8 * g);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], b, acc, 0, 0, 0);
            }
            const int col = 16 * t + c;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t at = static_cast<size_t>(row + 4 * g + r) * N + col;
                if constexpr (kEpi == kEpiReluBf16) {
                    static_cast<__bf16*>(Y)[at] = static_cast<__bf16>(fmaxf(0.0f, acc[r] + bias[col]));
                } else if constexpr (kEpi == kEpiF32) {
                    static_cast<float*>(Y)[at] = acc[r] + bias[col];
                } else {
                    const float keep = static_cast<float>(mask[at]) > 0.0f ? 1.0f : 0.0f;
                    static_cast<__bf16*>(Y)[at] = static_cast<__bf16>(acc[r] * keep);
                }
            }
        }
    }
}

// partial[b][n * K + k] = sum over block b's rows of dY[m][n] X[m][k]; partial[b][N * K + n] = sum dY[m][n].
// Rows are staged 32 at a time into LDS transposed (dY^T[n][m], X^T[k][m]) so that both MFMA operands are
// 16-byte LDS reads; the 16 x 16 output tiles are spread over the block's four waves.
template <int K, int N>
__global__ void __launch_bounds__(kGemmBlock) k_wgrad(const __bf16* __restrict__ dY, const __bf16* __restrict__ X,
                                                      uint32_t M, float* __restrict__ partial) {
    constexpr int kStride = kWgradChunk + kWgradPad;
    __shared__ __align__(16) __bf16 xt[K * kStride];
    __shared__ __align__(16) __bf16 yt[N * kStride];
    constexpr int kTiles = (N / 16) * (K / 16);
    constexpr int kPerWave = (kTiles + 3) / 4;
    const int wave = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, c = l & 15;
    f32x4 acc[kPerWave];
#pragma unroll
    for (int q = 0; q < kPerWave; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;  // threads n < N: bias gradient
    const uint32_t rows_per_block = (M / kWgradBlocks + kWgradChunk - 1) / kWgradChunk * kWgradChunk;
    const uint32_t begin = blockIdx.x * rows_per_block;
    const uint32_t end = std::min<uint32_t>(M, begin + rows_per_block);
    for (uint32_t m0 = begin; m0 < end; m0 += kWgradChunk) {
        // Stage: 32 rows x (K + N) bf16, 8 per thread-step, written transposed.
        for (int idx = threadIdx.x; idx < kWgradChunk * (K + N) / 8; idx += blockDim.x) {
            const int m = idx / ((K + N) / 8), kb = idx % ((K + N) / 8);
            const bool row_ok = m0 + m < end;
            if (kb < K / 8) {
                bf16x8 v{};
                if (row_ok) v = *reinterpret_cast<const bf16x8*>(X + static_cast<size_t>(m0 + m) * K + 8 * kb);
#pragma unroll
                for (int e = 0; e < 8; ++e) xt[(8 * kb + e) * kStride + m] = row_ok ? v[e] : static_cast<__bf16>(0.0f);
            } else {
                const int nb = kb - K / 8;
                bf16x8 v{};
                if (row_ok) v = *reinterpret_cast<const bf16x8*>(dY + static_cast<size_t>(m0 + m) * N + 8 * nb);
#pragma unroll
                for (int e = 0; e < 8; ++e) yt[(8 * nb + e) * kStride + m] = row_ok ? v[e] : static_cast<__bf16>(0.0f);
            }
        }
        __syncthreads();
        if (threadIdx.x < N)
            for (int m = 0; m < kWgradChunk; ++m) bsum += static_cast<float>(yt[threadIdx.x * kStride + m]);
#pragma unroll
        for (int q = 0; q < kPerWave; ++q) {
            const int tile = wave + 4 * q;
            if (tile < kTiles) {
                const int tn = tile / (K / 16), tk = tile % (K / 16);
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(yt + (16 * tn + c) * kStride + 8 * g);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(xt + (16 * tk + c) * kStride + 8 * g);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* out = partial + static_cast<size_t>(blockIdx.x) * (N * K + N);
#pragma unroll
    for (int q = 0; q < kPerWave; ++q) {
        const int tile = wave + 4 * q;
        if (tile < kTiles) {
            const int tn = tile / (K / 16), tk = tile % (K / 16);
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(16 * tn + 4 * g + r) * K + 16 * tk + c] = acc[q][r];
        }
    }
    if (threadIdx.x < N) out[N * K + threadIdx.x] = bsum;
}

// grad[i] = sum over blocks of partial[b][i], in block order (deterministic).
__global__ void k_reduce(const float* __restrict__ partial, int blocks, int size, float* __restrict__ grad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int b = 0; b < blocks; ++b) s += partial[static_cast<size_t>(b) * size + i];
        grad[i] = s;
    }
}

// Records: 3 float4 each, (position, weight), (outgoing direction, 0), (sampled direction, 0).
__global__ void k_encode(GuideView gv, const float4* __restrict__ recs, uint32_t n, uint32_t M, __bf16* __restrict__ X0) {
    const uint32_t total = M * (kGuideIn / 8);
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const uint32_t m = idx / (kGuideIn / 8), fb = idx % (kGuideIn / 8);
        bf16x8 x{};
        if (m < n) {
            const float4 p = recs[3 * m], d = recs[3 * m + 1];
            float v[6], f8[8];
            for (int k = 0; k < 6; ++k) v[k] = guide_coord(gv, mk(p.x, p.y, p.z), mk(d.x, d.y, d.z), k);
            guide_features8(v, 8 * static_cast<int>(fb), f8);
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = static_cast<__bf16>(f8[e]);
        }
        reinterpret_cast<bf16x8*>(X0)[idx] = x;
    }
}

// Sum of the record weights, per block (partial[b]).
__global__ void k_weight_sum(const float4* __restrict__ recs, uint32_t n, float* __restrict__ partial) {
    __shared__ float red[256];
    float s = 0.0f;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) s += recs[3 * m].w;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dO of the loss -sum_i w_i log q(d_i) / sum_i w_i for the vMF mixture of outputs O (as k_guide_infer builds it):
//   logits:   w (pi_j - r_j)
//   log k_j:  -w r_j (1 - 2 k / (e^(2k) - 1) + k (mu . d - 1))        (0 where the clamp is active)
//   u_j:      -w r_j k (d - (mu . d) mu) / |u_j|
// with r_j the posterior responsibility of lobe j. loss_partial[b]: the block's share of the loss.
__global__ void k_mixture_loss(const float* __restrict__ O, const float4* __restrict__ recs, uint32_t n, uint32_t M,
                               const float* __restrict__ wsum_partial, int wsum_blocks, float clamp,
                               __bf16* __restrict__ dO, float* __restrict__ loss_partial) {
    __shared__ float red[256];
    float wsum = 0.0f;
    for (int b = 0; b < wsum_blocks; ++b) wsum += wsum_partial[b];
    const float inv_wsum = wsum > 0.0f ? 1.0f / wsum : 0.0f;
    // Optional clamp of single records at a multiple of the mean weight (a few paths that find a small bright
    // region carry most of the weight; the clamp trades that variance for a slightly flattened target).
    const float wmax = clamp > 0.0f && n > 0 ? clamp * wsum / static_cast<float>(n) : 3.0e38f;
    float loss = 0.0f;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        __bf16* row = dO + static_cast<size_t>(m) * kGuideOut;
        for (int k = 5 * kGuideLobes; k < kGuideOut; ++k) row[k] = static_cast<__bf16>(0.0f);
        const float w = m < n ? fminf(recs[3 * m].w, wmax) * inv_wsum : 0.0f;
        if (!(w > 0.0f)) {
            for (int k = 0; k < 5 * kGuideLobes; ++k) row[k] = static_cast<__bf16>(0.0f);
            continue;
        }
        const float* o = O + static_cast<size_t>(m) * kGuideOut;
        const float4 di = recs[3 * m + 2];
        const V3 d = mk(di.x, di.y, di.z);
        float logit_max = -3.0e38f;
        for (int j = 0; j < kGuideLobes; ++j) logit_max = fmaxf(logit_max, o[5 * j + 4]);
        float zsum = 0.0f;
        for (int j = 0; j < kGuideLobes; ++j) zsum += expf(o[5 * j + 4] - logit_max);
        const float log_z = logf(zsum) + logit_max;
        // log of weight x density of each lobe at d, then log q by log-sum-exp.
        float lc[kGuideLobes];
        float lmax = -3.0e38f;
        for (int j = 0; j < kGuideLobes; ++j) {
            const V3 u = mk(o[5 * j], o[5 * j + 1], o[5 * j + 2]);
            const float len = length(u);
            const V3 mu = len > 1e-6f ? u / len : mk(0.0f, 0.0f, 1.0f);
            const float k = expf(clampf(o[5 * j + 3], kGuideMinLogK, kGuideMaxLogK));
            lc[j] = (o[5 * j + 4] - log_z) + logf(k) - logf(kTwoPi * expm1_neg2k(k)) + k * (dot(mu, d) - 1.0f);
            lmax = fmaxf(lmax, lc[j]);
        }
        float qs = 0.0f;
        for (int j = 0; j < kGuideLobes; ++j) qs += expf(lc[j] - lmax);
        const float log_q = logf(qs) + lmax;
        loss -= w * log_q;
        for (int j = 0; j < kGuideLobes; ++j) {
            const V3 u = mk(o[5 * j], o[5 * j + 1], o[5 * j + 2]);
            const float len = length(u);
            const V3 mu = len > 1e-6f ? u / len : mk(0.0f, 0.0f, 1.0f);
            const float lk = o[5 * j + 3];
            const float k = expf(clampf(lk, kGuideMinLogK, kGuideMaxLogK));
            const float cosd = dot(mu, d);
            const float r = expf(lc[j] - log_q);
            const float pi = expf(o[5 * j + 4] - log_z);
            V3 gu = mk(0.0f, 0.0f, 0.0f);
            if (len > 1e-6f) gu = (d - mu * cosd) * (-w * r * k / len);
            float gk = 0.0f;
            if (lk > kGuideMinLogK && lk < kGuideMaxLogK) gk = -w * r * (1.0f - 2.0f * k / expm1f(2.0f * k) + k * (cosd - 1.0f));
            row[5 * j] = static_cast<__bf16>(gu.x);
            row[5 * j + 1] = static_cast<__bf16>(gu.y);
            row[5 * j + 2] = static_cast<__bf16>(gu.z);
            row[5 * j + 3] = static_cast<__bf16>(gk);
            row[5 * j + 4] = static_cast<__bf16>(w * (pi - r));
        }
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction on every parameter; refreshes the bf16 W and W^T copies.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float c1, float c2, __bf16* __restrict__ wbf,
                       __bf16* __restrict__ wtbf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kGuideParams; i += gridDim.x * blockDim.x) {
        const float gr = grad[i];
        const float a = kAdamB1 * m1[i] + (1.0f - kAdamB1) * gr;
        const float b = kAdamB2 * m2[i] + (1.0f - kAdamB2) * gr * gr;
        m1[i] = a;
        m2[i] = b;
        const float v = p[i] - lr * (a * c1) / (sqrtf(b * c2) + kAdamEps);
        p[i] = v;
        for (int l = 0; l < kGuideLayers; ++l) {
            const int w0 = guide_w_offset(l), b0 = guide_b_offset(l);
            if (i >= w0 && i < b0) {
                const int in = guide_layer_in(l), out = guide_layer_out(l);
                const int o = (i - w0) / in, k = (i - w0) % in;
                wbf[guide_wbf_offset(l) + o * in + k] = static_cast<__bf16>(v);
                wtbf[guide_wbf_offset(l) + k * out + o] = static_cast<__bf16>(v);
            }
        }
    }
}

// bf16 copies from the masters (after set_params / create).
__global__ void k_refresh_bf16(const float* __restrict__ p, __bf16* __restrict__ wbf, __bf16* __restrict__ wtbf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kGuideParams; i += gridDim.x * blockDim.x) {
        for (int l = 0; l < kGuideLayers; ++l) {
            const int w0 = guide_w_offset(l), b0 = guide_b_offset(l);
            if (i >= w0 && i < b0) {
                const int in = guide_layer_in(l), out = guide_layer_out(l);
                const int o = (i - w0) / in, k = (i - w0) % in;
                wbf[guide_wbf_offset(l) + o * in + k] = static_cast<__bf16>(p[i]);
                wtbf[guide_wbf_offset(l) + k * out + o] = static_cast<__bf16>(p[i]);
            }
        }
    }
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Queries from host-side arrays (pe_guide_infer).
struct ArrayQueries {
    const float* pos;
    const float* dir;
    float4* mix;
    __device__ __forceinline__ bool fetch(uint32_t i, V3& p, V3& d) const {
        p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        return true;
    }
    __device__ __forceinline__ QueueSegs segs(uint32_t n) const { return QueueSegs{nullptr, n, n, 1u}; }
    __device__ __forceinline__ uint32_t begin(uint32_t) const { return 0u; }
    __device__ __forceinline__ void store_part(uint32_t i, int part, float4 a, float4 b, float2 cdf, float2 coef) const {
        guide_store_part(mix + static_cast<size_t>(i) * kGuideMixFloat4, part, a, b, cdf, coef);
    }
};

// Mixture (stored form) -> 8 x (mean xyz, concentration, weight) for the host.
__global__ void k_unpack_mix(const float4* __restrict__ mix, uint32_t n, float* __restrict__ lobes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* m = mix + static_cast<size_t>(i) * kGuideMixFloat4;
        const float cdf[8] = {m[8].x, m[8].y, m[8].z, m[8].w, m[9].x, m[9].y, m[9].z, m[9].w};
        for (int j = 0; j < kGuideLobes; ++j) {
            float* o = lobes + static_cast<size_t>(i) * 40 + 5 * j;
            o[0] = m[j].x;
            o[1] = m[j].y;
            o[2] = m[j].z;
            o[3] = m[j].w;
            o[4] = j == 0 ? cdf[0] : cdf[j] - cdf[j - 1];
        }
    }
}

int grid_of(size_t work, int block = 256, int cap = 4096) {
    return static_cast<int>(std::clamp<size_t>((work + block - 1) / block, 1, static_cast<size_t>(cap)));
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return PE_ERR_HIP;
    }
    return PE_OK;
}

// One Adam step on `n` records already in g->recs (n <= batch_rows), on stream s.
int train_step(pe_guide* g, uint32_t n, hipStream_t s) {
    const uint32_t M = g->batch_rows;
    const GuideView gv = g->view(0.0f);
    const __bf16* wbf = g->wbf.as<__bf16>();
    const __bf16* wtbf = g->wtbf.as<__bf16>();
    const float* P = g->params.as<float>();
    const int rows_grid = grid_of(M / 16 * 64, kGemmBlock, 2048);
    const float4* recs = g->recs.as<float4>();
    hipLaunchKernelGGL(k_encode, dim3(grid_of(static_cast<size_t>(M) * kGuideIn / 8)), dim3(256), 0, s, gv, recs, n, M,
                       g->x0.as<__bf16>());
    // forward
    hipLaunchKernelGGL((k_gemm_rows<kGuideIn, kGuideHid, kEpiReluBf16>), dim3(rows_grid), dim3(kGemmBlock), 0, s,
                       g->x0.as<__bf16>(), wbf + guide_wbf_offset(0), P + guide_b_offset(0), nullptr,
                       g->h[0].get(), M);
    for (int l = 1; l < kGuideLayers - 1; ++l)
        hipLaunchKernelGGL((k_gemm_rows<kGuideHid, kGuideHid, kEpiReluBf16>), dim3(rows_grid), dim3(kGemmBlock), 0, s,
                           g->h[l - 1].as<__bf16>(), wbf + guide_wbf_offset(l), P + guide_b_offset(l), nullptr,
                           g->h[l].get(), M);
    hipLaunchKernelGGL((k_gemm_rows<kGuideHid, kGuideOut, kEpiF32>), dim3(rows_grid), dim3(kGemmBlock), 0, s,
                       g->h[3].as<__bf16>(), wbf + guide_wbf_offset(4), P + guide_b_offset(4), nullptr, g->out.get(), M);
    // loss
    float* scal = g->scalars.as<float>();  // [0, 256): weight-sum partials, [256, 512): loss partials
    hipLaunchKernelGGL(k_weight_sum, dim3(256), dim3(256), 0, s, recs, n, scal);
    hipLaunchKernelGGL(k_mixture_loss, dim3(256), dim3(256), 0, s, g->out.as<float>(), recs, n, M, scal, 256,
                       g->cfg.weight_clamp, g->d_out.as<__bf16>(), scal + 256);
    // backward: layer 4 down to 0
    float* grad = g->grad.as<float>();
    float* part = g->partial.as<float>();
    const __bf16* dcur = g->d_out.as<__bf16>();
    for (int l = kGuideLayers - 1; l >= 0; --l) {
        const __bf16* xin = l == 0 ? g->x0.as<__bf16>() : g->h[l - 1].as<__bf16>();
        const int in = guide_layer_in(l), out = guide_layer_out(l);
        if (l == 0)
            hipLaunchKernelGGL((k_wgrad<kGuideIn, kGuideHid>), dim3(kWgradBlocks), dim3(kGemmBlock), 0, s, dcur, xin, M, part);
        else
            hipLaunchKernelGGL((k_wgrad<kGuideHid, kGuideHid>), dim3(kWgradBlocks), dim3(kGemmBlock), 0, s, dcur, xin, M, part);
        const int size = out * in + out;
        hipLaunchKernelGGL(k_reduce, dim3(grid_of(size)), dim3(256), 0, s, part, kWgradBlocks, size, grad + guide_w_offset(l));
        if (l > 0) {
            __bf16* dnext = g->d_h[l & 1].as<__bf16>();
            hipLaunchKernelGGL((k_gemm_rows<kGuideHid, kGuideHid, kEpiMaskBf16>), dim3(rows_grid), dim3(kGemmBlock), 0, s,
                               dcur, wtbf + guide_wbf_offset(l), nullptr, g->h[l - 1].as<__bf16>(), dnext, M);
            dcur = dnext;
        }
    }
    const float t = static_cast<float>(g->steps + 1);
    const float c1 = 1.0f / (1.0f - std::pow(kAdamB1, t)), c2 = 1.0f / (1.0f - std::pow(kAdamB2, t));
    hipLaunchKernelGGL(k_adam, dim3(grid_of(kGuideParams)), dim3(256), 0, s, g->params.as<float>(), grad,
                       g->adam_m.as<float>(), g->adam_v.as<float>(), g->cfg.learning_rate, c1, c2,
                       g->wbf.as<__bf16>(), g->wtbf.as<__bf16>());
    ++g->steps;
    return check_launch("guide training step");
}

int read_loss(pe_guide* g, hipStream_t s, float* loss) {
    float part[256];
    PE_HIP_CHECK(hipMemcpyAsync(part, g->scalars.as<float>() + 256, sizeof(part), hipMemcpyDeviceToHost, s));
    PE_HIP_CHECK(hipStreamSynchronize(s));
    float sum = 0.0f;
    for (float v : part) sum += v;
    *loss = sum;
    return PE_OK;
}

}  // namespace

int guide_reserve_paths(pe_guide* g, uint32_t paths, uint32_t slots) {
    if (paths <= g->path_capacity && slots <= g->slot_capacity) return PE_OK;
    if (!g->mix.alloc(static_cast<size_t>(slots) * kGuideMixFloat4 * 16) ||
        !g->path_records.alloc(static_cast<size_t>(paths) * 2 * 4 * 16))
        return PE_ERR_OOM;
    g->path_capacity = paths;
    g->slot_capacity = slots;
    return PE_OK;
}

int guide_clear_records(pe_guide* g, hipStream_t s) {
    const size_t n = static_cast<size_t>(g->record_slots) * 12;
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_of(n)), dim3(256), 0, s, g->records.as<float>(), n, 0.0f);
    return check_launch("guide record clear");
}

int guide_train_from_records(pe_guide* g, hipStream_t s) {
    const uint32_t B = g->batch_rows;
    const uint32_t batches = g->record_slots / B;
    for (int k = 0; k < g->cfg.steps_per_launch; ++k) {
        const uint32_t b = static_cast<uint32_t>(g->rng() % batches);
        PE_HIP_CHECK(hipMemcpyAsync(g->recs.get(), g->records.as<float4>() + static_cast<size_t>(b) * B * 3,
                                    static_cast<size_t>(B) * 3 * 16, hipMemcpyDeviceToDevice, s));
        const int rc = train_step(g, B, s);
        if (rc != PE_OK) return rc;
    }
    return PE_OK;
}

}  // namespace pe

using namespace pe;

extern "C" {

int pe_guide_create(pe_device* dev, pe_scene* scene, const pe_guide_config* config, pe_guide** out) {
    if (!dev || !scene || !out) {
        set_error("pe_guide_create: NULL argument");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    try {
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        auto g = std::make_unique<pe_guide>();
        g->dev = dev;
        pe_guide_config c = config ? *config : pe_guide_config{};
        if (c.learning_rate < 0.0f || c.guide_fraction < 0.0f || c.guide_fraction >= 1.0f || c.batch < 0 ||
            c.steps_per_launch < 0 || c.weight_clamp < 0.0f) {
            set_error("pe_guide_create: learning_rate >= 0, guide_fraction in [0, 1), batch and steps >= 0");
            return PE_ERR_INVALID;
        }
        if (c.learning_rate == 0.0f) c.learning_rate = 1e-3f;
        if (c.guide_fraction == 0.0f) c.guide_fraction = 0.5f;
        if (c.batch == 0) c.batch = 1 << 18;
        c.batch = (c.batch + 255) / 256 * 256;
        if (c.steps_per_launch == 0) c.steps_per_launch = 4;
        if (c.guided_bounces <= 0) c.guided_bounces = 3;
        g->cfg = c;
        // Position normalisation: the given box, else the scene's (union of the root's child boxes).
        V3 lo = mk(c.bounds_min[0], c.bounds_min[1], c.bounds_min[2]);
        V3 hi = mk(c.bounds_max[0], c.bounds_max[1], c.bounds_max[2]);
        if (!(hi.x > lo.x && hi.y > lo.y && hi.z > lo.z)) {
            BvhNode4 root{};
            PE_HIP_CHECK(hipMemcpy(&root, scene->nodes.get(), sizeof(root), hipMemcpyDeviceToHost));
            const float lx[4] = {root.lox.x, root.lox.y, root.lox.z, root.lox.w};
            const float hx[4] = {root.hix.x, root.hix.y, root.hix.z, root.hix.w};
            const float ly[4] = {root.loy.x, root.loy.y, root.loy.z, root.loy.w};
            const float hy[4] = {root.hiy.x, root.hiy.y, root.hiy.z, root.hiy.w};
            const float lz[4] = {root.loz.x, root.loz.y, root.loz.z, root.loz.w};
            const float hz[4] = {root.hiz.x, root.hiz.y, root.hiz.z, root.hiz.w};
            const uint32_t items[4] = {root.item.x, root.item.y, root.item.z, root.item.w};
            lo = splat(3.0e38f);
            hi = splat(-3.0e38f);
            for (int k = 0; k < 4; ++k) {
                if (items[k] == 0xFFFFFFFFu) continue;
                lo = vmin(lo, mk(lx[k], ly[k], lz[k]));
                hi = vmax(hi, mk(hx[k], hy[k], hz[k]));
            }
        }
        const V3 ext = hi - lo;
        g->lo = lo;
        g->inv_extent = mk(1.0f / std::fmax(ext.x, 1e-6f), 1.0f / std::fmax(ext.y, 1e-6f), 1.0f / std::fmax(ext.z, 1e-6f));
        // Initialisation: U(-1/sqrt(in), 1/sqrt(in)) for W and b (torch.nn.Linear's), the output layer scaled by
        // 0.1 so the first mixture is close to uniform; padded inputs and outputs zero.
        std::vector<float> p(kGuideParams, 0.0f);
        std::mt19937_64 rng(c.seed);
        std::uniform_real_distribution<float> uni(-1.0f, 1.0f);
        for (int l = 0; l < kGuideLayers; ++l) {
            const int in = guide_layer_in(l), outn = guide_layer_out(l);
            const float bound = 1.0f / std::sqrt(static_cast<float>(in)) * (l == kGuideLayers - 1 ? 0.1f : 1.0f);
            for (int o = 0; o < outn; ++o) {
                const bool used_out = l < kGuideLayers - 1 || o < 5 * kGuideLobes;
                for (int k = 0; k < in; ++k) {
                    const bool used_in = l > 0 || k < 6 * 2 * kGuideFreqs + 6;
                    const float v = uni(rng) * bound;
                    p[guide_w_offset(l) + o * in + k] = used_out && used_in ? v : 0.0f;
                }
                const float vb = uni(rng) * bound;
                p[guide_b_offset(l) + o] = used_out ? vb : 0.0f;
            }
        }
        g->rng.seed(c.seed ^ 0x9e3779b97f4a7c15ull);
        g->batch_rows = static_cast<uint32_t>(c.batch);
        const size_t M = g->batch_rows;
        g->record_slots = g->batch_rows * static_cast<uint32_t>(std::max(4, c.steps_per_launch));
        bool ok = g->params.alloc(kGuideParams * 4) && g->adam_m.alloc(kGuideParams * 4) &&
                  g->adam_v.alloc(kGuideParams * 4) && g->grad.alloc(kGuideParams * 4) &&
                  g->wbf.alloc(kGuideWeights * 2) && g->wtbf.alloc(kGuideWeights * 2) && g->recs.alloc(M * 48) &&
                  g->x0.alloc(M * kGuideIn * 2) && g->out.alloc(M * kGuideOut * 4) && g->d_out.alloc(M * kGuideOut * 2) &&
                  g->partial.alloc(static_cast<size_t>(kWgradBlocks) * (kGuideHid * kGuideIn + kGuideHid) * 4) &&
                  g->scalars.alloc(512 * 4) && g->records.alloc(static_cast<size_t>(g->record_slots) * 48);
        for (int k = 0; k < 4 && ok; ++k) ok = g->h[k].alloc(M * kGuideHid * 2);
        for (int k = 0; k < 2 && ok; ++k) ok = g->d_h[k].alloc(M * kGuideHid * 2);
        if (!ok) return PE_ERR_OOM;
        *out = g.release();
        return pe_guide_set_params(*out, p.data());
    } catch (const std::exception& e) {
        set_error(std::string("pe_guide_create: ") + e.what());
        return PE_ERR_INVALID;
    }
}

int pe_guide_destroy(pe_guide* g) {
    if (!g) return PE_OK;
    (void)hipSetDevice(g->dev->ordinal);
    (void)hipDeviceSynchronize();
    delete g;
    return PE_OK;
}

int pe_guide_get_params(pe_guide* g, float* params) {
    if (!g || !params) {
        set_error("pe_guide_get_params: NULL argument");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(g->dev->ordinal));
    PE_HIP_CHECK(hipMemcpyAsync(params, g->params.get(), kGuideParams * 4, hipMemcpyDeviceToHost, g->dev->stream));
    PE_HIP_CHECK(hipStreamSynchronize(g->dev->stream));
    return PE_OK;
}

int pe_guide_set_params(pe_guide* g, const float* params) {
    if (!g || !params) {
        set_error("pe_guide_set_params: NULL argument");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(g->dev->ordinal));
    hipStream_t s = g->dev->stream;
    PE_HIP_CHECK(hipMemcpyAsync(g->params.get(), params, kGuideParams * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_of(kGuideParams)), dim3(256), 0, s, g->adam_m.as<float>(),
                       static_cast<size_t>(kGuideParams), 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_of(kGuideParams)), dim3(256), 0, s, g->adam_v.as<float>(),
                       static_cast<size_t>(kGuideParams), 0.0f);
    hipLaunchKernelGGL(k_refresh_bf16, dim3(grid_of(kGuideParams)), dim3(256), 0, s, g->params.as<float>(),
                       g->wbf.as<__bf16>(), g->wtbf.as<__bf16>());
    int rc = check_launch("pe_guide_set_params");
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(s));
    g->steps = 0;
    return PE_OK;
}

int pe_guide_infer(pe_guide* g, size_t n, const float* pos, const float* dir, float* lobes) {
    if (!g || (n > 0 && (!pos || !dir || !lobes))) {
        set_error("pe_guide_infer: NULL argument");
        return PE_ERR_INVALID;
    }
    if (n == 0) return PE_OK;
    if (n > (size_t{1} << 26)) {
        set_error("pe_guide_infer: at most 2^26 queries per call");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(g->dev->ordinal));
    hipStream_t s = g->dev->stream;
    DevBuf dpos, ddir, dmix, dlob;
    if (!dpos.alloc(n * 12) || !ddir.alloc(n * 12) || !dmix.alloc(n * kGuideMixFloat4 * 16) || !dlob.alloc(n * 160))
        return PE_ERR_OOM;
    PE_HIP_CHECK(hipMemcpyAsync(dpos.get(), pos, n * 12, hipMemcpyHostToDevice, s));
    PE_HIP_CHECK(hipMemcpyAsync(ddir.get(), dir, n * 12, hipMemcpyHostToDevice, s));
    const ArrayQueries q{dpos.as<float>(), ddir.as<float>(), dmix.as<float4>()};
    const int blocks = std::max(1, std::min<int>(g->dev->cu_count * 2, static_cast<int>((n + 127) / 128)));
    hipLaunchKernelGGL((k_guide_infer<ArrayQueries>), dim3(blocks), dim3(kInferBlock), 0, s, g->view(0.0f), q,
                       static_cast<uint32_t>(n));
    hipLaunchKernelGGL(k_unpack_mix, dim3(grid_of(n)), dim3(256), 0, s, dmix.as<float4>(), static_cast<uint32_t>(n),
                       dlob.as<float>());
    int rc = check_launch("pe_guide_infer");
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipMemcpyAsync(lobes, dlob.get(), n * 160, hipMemcpyDeviceToHost, s));
    PE_HIP_CHECK(hipStreamSynchronize(s));
    return PE_OK;
}

int pe_guide_train_step(pe_guide* g, size_t n, const float* records, float* loss, float* grad) {
    if (!g || (n > 0 && !records)) {
        set_error("pe_guide_train_step: NULL argument");
        return PE_ERR_INVALID;
    }
    if (n > g->batch_rows) {
        set_error("pe_guide_train_step: more records than the batch (" + std::to_string(g->batch_rows) + ")");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(g->dev->ordinal));
    hipStream_t s = g->dev->stream;
    std::vector<float> packed(static_cast<size_t>(n) * 12, 0.0f);
    for (size_t i = 0; i < n; ++i) {
        const float* r = records + 10 * i;
        float* o = packed.data() + 12 * i;
        o[0] = r[0], o[1] = r[1], o[2] = r[2], o[3] = r[9];
        o[4] = r[3], o[5] = r[4], o[6] = r[5];
        o[8] = r[6], o[9] = r[7], o[10] = r[8];
    }
    if (n > 0) PE_HIP_CHECK(hipMemcpyAsync(g->recs.get(), packed.data(), packed.size() * 4, hipMemcpyHostToDevice, s));
    int rc = train_step(g, static_cast<uint32_t>(n), s);
    if (rc != PE_OK) return rc;
    float l = 0.0f;
    rc = read_loss(g, s, &l);
    if (rc != PE_OK) return rc;
    g->last_loss = l;
    if (loss) *loss = l;
    if (grad) {
        PE_HIP_CHECK(hipMemcpyAsync(grad, g->grad.get(), kGuideParams * 4, hipMemcpyDeviceToHost, s));
        PE_HIP_CHECK(hipStreamSynchronize(s));
    }
    return PE_OK;
}

int pe_guide_stats(pe_guide* g, uint64_t* steps, float* last_loss, uint64_t* records, uint64_t* inferred) {
    if (!g) {
        set_error("pe_guide_stats: NULL guide");
        return PE_ERR_INVALID;
    }
    if (last_loss && g->steps > 0) {
        float l = 0.0f;
        PE_HIP_CHECK(hipSetDevice(g->dev->ordinal));
        const int rc = read_loss(g, g->dev->stream, &l);
        if (rc != PE_OK) return rc;
        g->last_loss = l;
    }
    if (steps) *steps = g->steps;
    if (last_loss) *last_loss = g->last_loss;
    if (records) *records = g->collected;
    if (inferred) *inferred = g->inferred;
    return PE_OK;
}

}  // extern "C"
