// sph_mlp.h — the surrogate network's kernels (BF16 operands, FP32 accumulate) and the
// BF16 rounding shared by host and device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace sphmlp {

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet), on host and device alike.
__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_bits_to_f32(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// C[M x N] = A[M x K] * B[K x N], operands BF16 read through element strides.
// M and N multiples of 32, K and kPerSplit multiples of 16. With `partial` set, K is
// split into `splits` chunks of kPerSplit and chunk c writes partial[c][M][N]; otherwise
// the epilogue adds bias[n], applies the backward ReLU mask and/or ReLU, and writes
// FP32 and/or BF16 rows of stride ldOut.
struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    uint32_t M, N, K;
    uint32_t sAm, sAk, sBk, sBn;
    uint32_t splits, kPerSplit;
    float* partial;
    const float* bias;
    const uint16_t* reluMask;   // forward output (BF16) whose positive entries pass the gradient
    int relu;
    float* outF32;
    uint16_t* outBf16;
    uint32_t ldOut;
};

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s);
hipError_t launch_reduce_partials(const float* partial, uint32_t splits, uint32_t count, float* out, hipStream_t s);
hipError_t launch_mse_grad(const float* y, const float* target, uint32_t rows, uint32_t rowsPadded, uint32_t ld,
                           uint32_t outs, float scale, uint16_t* dy, float* rowErr, hipStream_t s);
// Column sums over `rows` rows; `partial` holds (c + ceil(c / 16)) * cols floats, c = ceil(rows / 64).
hipError_t launch_column_sum_bf16(const uint16_t* m, uint32_t rows, uint32_t cols, float* partial, float* out,
                                  hipStream_t s);
hipError_t launch_adam(float* w, float* m1, float* m2, const float* grad, uint32_t count, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t* wBf16, const float* frozen, hipStream_t s);
hipError_t launch_to_bf16(const float* src, uint32_t count, uint16_t* dst, hipStream_t s);

// Inference of the whole network in one kernel (sph_mlp_fused.hip): feature rows x
// (rows x 64 BF16) -> y[row] = (out0, out1, out2, 0). wBf16 / wF32: the parameter
// mirrors (W[out][in], b[out] per layer). blocks = 0 picks one workgroup per CU.
hipError_t launch_mlp_fused(const uint16_t* x, uint32_t rows, const uint16_t* wBf16, const float* wF32, float4* y,
                            uint32_t blocks, hipStream_t s);
// The same network pass with the rows built in the kernel from the prepared state of a
// single scene at slot offset 0: pos (density in .w) and the neighbourhood sums
// (kSumRows x total float4), for rows <= total particles.
// The training forward pass in the same kernel: rows x 64 BF16 features -> the hidden
// activations h1, h2, h3 (rows x 128 BF16, after ReLU) and the 32 outputs (rows x 32 FP32).
hipError_t launch_mlp_fused_train(const uint16_t* x, uint32_t rows, const uint16_t* wBf16, const float* wF32,
                                  uint16_t* h1, uint16_t* h2, uint16_t* h3, float* y32, hipStream_t s);
hipError_t launch_mlp_fused_sums(const float4* pos, const float4* sums, uint32_t total, uint32_t rows, float rho0,
                                 float h, float invV, const uint16_t* wBf16, const float* wF32, float4* y,
                                 uint32_t blocks, hipStream_t s);
uint32_t mlp_fused_blocks(uint32_t rows);

}  // namespace sphmlp
