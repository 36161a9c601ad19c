// sph_mlp.hip — BF16 MFMA kernels of the surrogate network (gfx950), FP32 accumulate.
//
// The network maps 64 per-particle features to 3 accelerations through three hidden
// layers of 128 with ReLU. Weights are stored as W[out][in] (the inner index contiguous);
// FP32 masters are updated by Adam and mirrored to BF16 for the products.
//
// One generic tile GEMM does every product: C[M x N] = A[M x K] * B[K x N] with operands
// read through element strides, so the same kernel serves the forward pass
// (X * W^T), the input gradient (dY * W) and the weight gradient (dY^T * X, reduced over
// the batch in fixed-order partial sums). Each wave computes one 32 x 32 tile with
// v_mfma_f32_32x32x16_bf16: lane l feeds A[m0 + l % 32][k0 + 8 (l / 32) + i] and
// B[k0 + 8 (l / 32) + i][n0 + l % 32], i = 0..7, and receives
// C[m0 + 8 (j / 4) + 4 (l / 32) + j % 4][n0 + l % 32] in accumulator j = 0..15.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_mlp.h"

namespace sphmlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 as_bf16(uint16_t bits) {
    __bf16 r;
    __builtin_memcpy(&r, &bits, 2);
    return r;
}

// Loads 8 BF16 values at p[0], p[s], ..., p[7 s]; one 16-byte load when contiguous.
__device__ __forceinline__ bf16x8 load8(const uint16_t* __restrict__ p, uint32_t s) {
    bf16x8 v;
    if (s == 1u) {
        const uint4 q = *reinterpret_cast<const uint4*>(p);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v[2 * i] = as_bf16((uint16_t)(w[i] & 0xFFFFu));
            v[2 * i + 1] = as_bf16((uint16_t)(w[i] >> 16));
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = as_bf16(p[(size_t)i * s]);
    }
    return v;
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t tilesN = g.N / 32u;
    const uint32_t tilesM = g.M / 32u;
    const uint32_t tile = wave % (tilesM * tilesN);
    const uint32_t chunk = wave / (tilesM * tilesN);       // batch chunk for split-K partials
    if (chunk >= g.splits) return;
    const uint32_t m0 = (tile / tilesN) * 32u, n0 = (tile % tilesN) * 32u;
    const uint32_t kBegin = chunk * g.kPerSplit;
    const uint32_t kEnd = kBegin + g.kPerSplit < g.K ? kBegin + g.kPerSplit : g.K;
    const uint32_t r = lane & 31u, half = lane >> 5;
    f32x16 acc = {};
    for (uint32_t k0 = kBegin; k0 < kEnd; k0 += 16u) {
        const uint32_t kk = k0 + half * 8u;
        const bf16x8 a = load8(g.A + (size_t)(m0 + r) * g.sAm + (size_t)kk * g.sAk, g.sAk);
        const bf16x8 b = load8(g.B + (size_t)kk * g.sBk + (size_t)(n0 + r) * g.sBn, g.sBk);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const uint32_t n = n0 + r;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t m = m0 + 8u * (uint32_t)(j / 4) + 4u * half + (uint32_t)(j % 4);
        float v = acc[j];
        if (g.partial) {
            g.partial[((size_t)chunk * g.M + m) * g.N + n] = v;
            continue;
        }
        if (g.bias) v += g.bias[n];
        if (g.reluMask) {   // backward through ReLU: the forward output was > 0
            const uint16_t hb = g.reluMask[(size_t)m * g.ldOut + n];
            v = ((hb & 0x8000u) == 0u && hb != 0u) ? v : 0.0f;
        }
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        if (g.outF32) g.outF32[(size_t)m * g.ldOut + n] = v;
        if (g.outBf16) g.outBf16[(size_t)m * g.ldOut + n] = f32_to_bf16_bits(v);
    }
}

// Sums split-K partials in chunk order (deterministic): out[i] = sum_c partial[c][i]. Eight
// loads are issued before their ordered adds, so the chain waits on memory once per eight.
__global__ void __launch_bounds__(256)
    k_reduce_partials(const float* __restrict__ partial, uint32_t splits, uint32_t count, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    uint32_t c = 0;
    for (; c + 8u <= splits; c += 8u) {
        float v[8];
#pragma unroll
        for (uint32_t e = 0; e < 8u; ++e) v[e] = partial[(size_t)(c + e) * count + i];
#pragma unroll
        for (uint32_t e = 0; e < 8u; ++e) s += v[e];
    }
    for (; c < splits; ++c) s += partial[(size_t)c * count + i];
    out[i] = s;
}

// dL/dy for L = mean over rows and the first `outs` columns of (y - t)^2; padded columns
// and rows past `rows` get zero. Also writes the per-row squared error for the loss.
__global__ void __launch_bounds__(256)
    k_mse_grad(const float* __restrict__ y, const float* __restrict__ target, uint32_t rows, uint32_t rowsPadded,
               uint32_t ld, uint32_t outs, float scale, uint16_t* __restrict__ dy, float* __restrict__ rowErr) {
    const uint32_t m = blockIdx.x * 256u + threadIdx.x;
    if (m >= rowsPadded) return;
    float e = 0.0f;
    for (uint32_t c = 0; c < ld; ++c) {
        float gval = 0.0f;
        if (m < rows && c < outs) {
            const float d = y[(size_t)m * ld + c] - target[(size_t)m * outs + c];
            e += d * d;
            gval = scale * d;
        }
        dy[(size_t)m * ld + c] = f32_to_bf16_bits(gval);
    }
    if (rowErr) rowErr[m] = e;
}

// First stage of a two-stage ordered sum of `chunks` partial rows of `cols` values:
// out[g][c] = sum over the g-th run of `per` consecutive chunks, in chunk order.
__global__ void __launch_bounds__(256)
    k_reduce_runs(const float* __restrict__ partial, uint32_t chunks, uint32_t per, uint32_t cols,
                  float* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t runs = (chunks + per - 1u) / per;
    if (i >= runs * cols) return;
    const uint32_t g = i / cols, c = i % cols;
    const uint32_t k0 = g * per, k1 = k0 + per < chunks ? k0 + per : chunks;
    float s = 0.0f;
    for (uint32_t k = k0; k < k1; ++k) s += partial[(size_t)k * cols + c];
    out[i] = s;
}

// Column sums of a row-major BF16 matrix over chunks of kColumnChunk rows: partial[chunk][col],
// rows in order inside a chunk; chunks are then summed in order by k_reduce_partials
// (deterministic). The bias gradient. Short chunks give the launch enough workgroups to
// spread over the GPU; eight rows are loaded before their ordered adds.
constexpr uint32_t kColumnChunk = 64;

__global__ void __launch_bounds__(256)
    k_column_partial_bf16(const uint16_t* __restrict__ m, uint32_t rows, uint32_t cols, float* __restrict__ partial) {
    const uint32_t r0 = blockIdx.x * kColumnChunk, r1 = r0 + kColumnChunk < rows ? r0 + kColumnChunk : rows;
    for (uint32_t c = threadIdx.x; c < cols; c += 256u) {
        float s = 0.0f;
        uint32_t r = r0;
        for (; r + 8u <= r1; r += 8u) {
            float v[8];
#pragma unroll
            for (uint32_t e = 0; e < 8u; ++e) v[e] = bf16_bits_to_f32(m[(size_t)(r + e) * cols + c]);
#pragma unroll
            for (uint32_t e = 0; e < 8u; ++e) s += v[e];
        }
        for (; r < r1; ++r) s += bf16_bits_to_f32(m[(size_t)r * cols + c]);
        partial[(size_t)blockIdx.x * cols + c] = s;
    }
}

// Adam (Kingma & Ba 2015) with host-computed bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t,
// then the BF16 mirror of the updated weight.
__global__ void __launch_bounds__(256)
    k_adam(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2, const float* __restrict__ grad,
           uint32_t count, float lr, float b1, float b2, float eps, float c1, float c2, uint16_t* __restrict__ wBf16,
           const float* __restrict__ frozen) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    if (frozen && frozen[i] != 0.0f) return;   // padding kept at zero
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float b = b2 * m2[i] + (1.0f - b2) * (g * g);
    m1[i] = a;
    m2[i] = b;
    const float mhat = a / c1;
    const float vhat = b / c2;
    const float nw = w[i] - lr * mhat / (sqrtf(vhat) + eps);
    w[i] = nw;
    if (wBf16) wBf16[i] = f32_to_bf16_bits(nw);
}

__global__ void __launch_bounds__(256)
    k_to_bf16(const float* __restrict__ src, uint32_t count, uint16_t* __restrict__ dst) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < count) dst[i] = f32_to_bf16_bits(src[i]);
}

}  // namespace

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s) {
    if (g.M % 32u || g.N % 32u || g.K % 16u || g.kPerSplit % 16u || g.splits == 0) return hipErrorInvalidValue;
    const uint32_t waves = (g.M / 32u) * (g.N / 32u) * g.splits;
    k_gemm<<<(waves + 3u) / 4u, 256, 0, s>>>(g);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* partial, uint32_t splits, uint32_t count, float* out, hipStream_t s) {
    k_reduce_partials<<<(count + 255u) / 256u, 256, 0, s>>>(partial, splits, count, out);
    return hipGetLastError();
}

hipError_t launch_mse_grad(const float* y, const float* target, uint32_t rows, uint32_t rowsPadded, uint32_t ld,
                           uint32_t outs, float scale, uint16_t* dy, float* rowErr, hipStream_t s) {
    k_mse_grad<<<(rowsPadded + 255u) / 256u, 256, 0, s>>>(y, target, rows, rowsPadded, ld, outs, scale, dy, rowErr);
    return hipGetLastError();
}

hipError_t launch_column_sum_bf16(const uint16_t* m, uint32_t rows, uint32_t cols, float* partial, float* out,
                                  hipStream_t s) {
    const uint32_t chunks = (rows + kColumnChunk - 1u) / kColumnChunk;
    if (chunks == 0) return hipErrorInvalidValue;
    k_column_partial_bf16<<<chunks, 256, 0, s>>>(m, rows, cols, partial);
    // runs of 16 chunks summed in order into the space after the partials, then the runs in order
    constexpr uint32_t kPer = 16;
    const uint32_t runs = (chunks + kPer - 1u) / kPer;
    float* stage = partial + (size_t)chunks * cols;
    k_reduce_runs<<<(runs * cols + 255u) / 256u, 256, 0, s>>>(partial, chunks, kPer, cols, stage);
    k_reduce_partials<<<(cols + 255u) / 256u, 256, 0, s>>>(stage, runs, cols, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, float* m1, float* m2, const float* grad, uint32_t count, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t* wBf16, const float* frozen, hipStream_t s) {
    k_adam<<<(count + 255u) / 256u, 256, 0, s>>>(w, m1, m2, grad, count, lr, b1, b2, eps, c1, c2, wBf16, frozen);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* src, uint32_t count, uint16_t* dst, hipStream_t s) {
    k_to_bf16<<<(count + 255u) / 256u, 256, 0, s>>>(src, count, dst);
    return hipGetLastError();
}

}  // namespace sphmlp
