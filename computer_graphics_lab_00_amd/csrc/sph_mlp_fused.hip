// sph_mlp_fused.hip — the surrogate network's inference pass in one kernel (gfx950).
//
// All four layers run back to back with the activations in registers: a workgroup
// stages the BF16 weights (97 KB, rows padded so LDS reads are conflict free) and the
// biases in LDS once, and each wave then streams groups of 64 feature rows through the
// network. HBM traffic is one read of the 128-byte feature row and one 16-byte write of
// the output per particle; the work is MFMA bound. The rollout's variant builds the rows
// itself from the density pass's neighbourhood sums (neighbourhood_features): 96 bytes
// read per particle and no feature array; its all-zero inputs 20..63 are skipped, which
// leaves every ReLU output, and so the result, unchanged.
//
// Each layer computes the transposed product out^T = W * act^T with
// v_mfma_f32_32x32x16_bf16: A = W (32 output features x 16 inputs, from LDS), B = act^T
// (16 inputs x 32 particle rows), so lane l always owns particle row l % 32. The
// accumulator of a 32 x 32 output tile holds, in lane l and element j, feature
// 8 (j / 4) + 4 (l / 32) + j % 4; the next layer's B operand needs, in lane l and
// element i of k-block kb, input 16 kb + 8 (l / 32) + i. The two orders are the same set
// of 16 features per k-block, so instead of moving data between lanes the staged
// weights of layers 1-3 are stored with their input columns permuted
// (logical column 16 kb + 8 h + i holds physical input 16 kb + 8 (i / 4) + 4 h + i % 4)
// and the accumulators are used as the next B operands in place.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_mlp.h"
#include "sph_surrogate.h"

namespace sphmlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kWaves = 8;               // waves per workgroup, 2 per SIMD
constexpr uint32_t kThreads = kWaves * 64;
constexpr uint32_t kLd0 = 64 + 8;            // padded LDS row of layer 0 (64 inputs)
constexpr uint32_t kLdH = 128 + 8;           // padded LDS row of layers 1-3 (128 inputs)
constexpr uint32_t kOff1 = 128 * kLd0;
constexpr uint32_t kOff2 = kOff1 + 128 * kLdH;
constexpr uint32_t kOff3 = kOff2 + 128 * kLdH;
constexpr uint32_t kLdsWeights = kOff3 + 32 * kLdH;   // BF16 elements
constexpr uint32_t kBiases = 128 * 3 + 32;
// parameter offsets (W[out][in] then b[out] per layer)
constexpr uint32_t kW0 = 0, kB0 = 8192, kW1 = 8320, kB1 = 24704, kW2 = 24832, kB2 = 41216, kW3 = 41344,
                   kB3 = 45440;

__device__ __forceinline__ uint32_t physical_input(uint32_t k) {
    const uint32_t r = k & 15u, h = r >> 3, i = r & 7u;
    return (k & ~15u) + ((i >> 2) << 3) + (h << 2) + (i & 3u);
}

__device__ __forceinline__ __bf16 bits_bf16(uint16_t b) {
    __bf16 r;
    __builtin_memcpy(&r, &b, 2);
    return r;
}

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

// Two values rounded to BF16 in one packed word, then ReLU on their bit patterns as signed
// 16-bit integers: every negative value (and -0) has the sign bit set, so max(bits, 0) maps
// it to +0 and keeps the others. Rounding to BF16 is monotone and keeps the sign, so ReLU
// after rounding equals rounding after ReLU.
__device__ __forceinline__ uint32_t relu_pair_bf16(float a, float b) {
    const bf16x2 c = __builtin_convertvector((f32x2){a, b}, bf16x2);
    i16x2 x;
    __builtin_memcpy(&x, &c, sizeof(x));
    x = __builtin_elementwise_max(x, (i16x2)0);
    uint32_t w;
    __builtin_memcpy(&w, &x, sizeof(w));
    return w;
}

// Bias, BF16 rounding and ReLU of one 32 x 32 accumulator tile into two B operands. With
// `store`, the tile's 16 values of this lane's row also go to store[0 .. 31] in feature order
// (the tile's 32 features): words 2q and 2q + 1 are features 8 q + 4 half .. + 3.
__device__ __forceinline__ void to_operands(const f32x16& acc, const float* bias, uint32_t half, bf16x8& lo,
                                            bf16x8& hi, uint16_t* store = nullptr) {
    uint32_t w[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int j = 2 * p;
        const uint32_t f = 8u * (uint32_t)(j / 4) + 4u * half + (uint32_t)(j % 4);   // j, j + 1: f, f + 1
        w[p] = relu_pair_bf16(acc[j] + bias[f], acc[j + 1] + bias[f + 1]);
    }
    __builtin_memcpy(&lo, &w[0], sizeof(lo));
    __builtin_memcpy(&hi, &w[4], sizeof(hi));
    if (store) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<uint2*>(store + 8u * (uint32_t)q + 4u * half) = make_uint2(w[2 * q], w[2 * q + 1]);
    }
}

// One layer over the first KB k-blocks of its (LD-strided) staged weights. The four
// 32-feature output tiles are finished two at a time (their MFMAs over every k-block, then
// their epilogue), so the epilogue of one pair can issue while the next pair's MFMAs run,
// and only four accumulators are live. Each accumulator still sums its k-blocks in order,
// so the results are those of finishing all four tiles together; at 1M rows the network
// pass ran 7 % faster this way (and 6 % one tile at a time).
template <int KB, uint32_t LD, int KIN>
__device__ __forceinline__ void hidden_layer(const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                             uint32_t lane, const bf16x8 (&in)[2][KIN], bf16x8 (&out)[2][8],
                                             uint16_t* __restrict__ hOut = nullptr, uint32_t row0 = 0,
                                             uint32_t rows = 0) {
    const uint32_t r = lane & 31u, half = lane >> 5;
    const uint32_t ld = LD;
    constexpr int kPair = 2;
#pragma unroll
    for (int m0 = 0; m0 < 4; m0 += kPair) {
        f32x16 acc[2][kPair];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int mm = 0; mm < kPair; ++mm) acc[nt][mm] = f32x16{};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
            for (int mm = 0; mm < kPair; ++mm) {
                const bf16x8 a = lds8(w + (uint32_t)((m0 + mm) * 32 + r) * ld + (uint32_t)kb * 16u + 8u * half);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    acc[nt][mm] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, in[nt][kb], acc[nt][mm], 0, 0, 0);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            // with hOut: the activations of rows row0 .. row0 + 63, 128 BF16 per row
            const uint32_t row = row0 + (uint32_t)nt * 32u + r;
            uint16_t* dst = (hOut && row < rows) ? hOut + (size_t)row * 128u : nullptr;
#pragma unroll
            for (int mm = 0; mm < kPair; ++mm) {
                const int mt = m0 + mm;
                to_operands(acc[nt][mm], bias + mt * 32, half, out[nt][2 * mt], out[nt][2 * mt + 1],
                            dst ? dst + mt * 32 : nullptr);
            }
        }
    }
}

struct FusedInput {
    const uint16_t* x;     // feature rows (rows x 64 BF16), or
    const float4* pos;     // the prepared state (density in .w) and
    const float4* sums;    // its neighbourhood sums (kSumRows x total), when SUMS
    uint32_t total;
    float rho0, h, invV;
    uint16_t* h1;          // STORE: the three hidden activations (rows x 128 BF16) and
    uint16_t* h2;
    uint16_t* h3;
    float* y32;            // the 32 outputs (rows x 32 FP32) the backward pass reads
};

// STORE (without SUMS): the training forward pass, which also keeps every layer's output.
template <bool SUMS, bool STORE = false>
__global__ void __launch_bounds__(kThreads)
    k_mlp_fused(FusedInput in, uint32_t rows, const uint16_t* __restrict__ wb, const float* __restrict__ wf,
                float4* __restrict__ y) {
    __shared__ __attribute__((aligned(16))) uint16_t sw[kLdsWeights];
    __shared__ float sb[kBiases];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < 128u * 64u; i += kThreads) sw[(i >> 6) * kLd0 + (i & 63u)] = wb[kW0 + i];
    for (uint32_t i = tid; i < 128u * 128u; i += kThreads) {
        const uint32_t o = i >> 7, k = i & 127u;
        sw[kOff1 + o * kLdH + k] = wb[kW1 + o * 128u + physical_input(k)];
        sw[kOff2 + o * kLdH + k] = wb[kW2 + o * 128u + physical_input(k)];
    }
    for (uint32_t i = tid; i < 32u * 128u; i += kThreads) {
        const uint32_t o = i >> 7, k = i & 127u;
        sw[kOff3 + o * kLdH + k] = wb[kW3 + o * 128u + physical_input(k)];
    }
    for (uint32_t i = tid; i < 128u; i += kThreads) {
        sb[i] = wf[kB0 + i];
        sb[128u + i] = wf[kB1 + i];
        sb[256u + i] = wf[kB2 + i];
    }
    if (tid < 32u) sb[384u + tid] = wf[kB3 + tid];
    __syncthreads();

    constexpr int KB0 = SUMS ? 2 : 4;   // k-blocks of layer 0 that can be non-zero
    static_assert(!SUMS || sphsur::kUsedFeatures <= 32, "the used features must fit two k-blocks");
    const uint32_t lane = tid & 63u, r = lane & 31u, half = lane >> 5;
    const uint32_t stride = gridDim.x * kWaves * 64u;
    for (uint32_t base = (blockIdx.x * kWaves + (tid >> 6)) * 64u; base < rows; base += stride) {
        // The weight fragments are loop invariant; without this the compiler hoists all
        // of them into registers and spills. They are re-read from LDS per row group.
        asm volatile("" ::: "memory");
        bf16x8 in0[2][KB0];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const uint32_t row = base + (uint32_t)nt * 32u + r;
            if (SUMS) {
                // lane (row, half) holds inputs 8 half .. 8 half + 7 of k-block 0 and
                // 16 + 8 half .. of k-block 1
                float f[sphsur::kUsedFeatures];
#pragma unroll
                for (int i = 0; i < (int)sphsur::kUsedFeatures; ++i) f[i] = 0.0f;
                if (row < rows) {
                    const size_t T = in.total;
                    sphsur::neighbourhood_features(in.pos[row].w, in.sums[row], in.sums[T + row], in.sums[2 * T + row],
                                                   in.sums[3 * T + row], in.sums[4 * T + row], in.rho0, in.h, in.invV,
                                                   f);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float a = half ? f[8 + i] : f[i];
                    const float b = (!half && 16 + i < (int)sphsur::kUsedFeatures) ? f[(16 + i) % sphsur::kUsedFeatures]
                                                                                  : 0.0f;
                    in0[nt][0][i] = bits_bf16(f32_to_bf16_bits(a));
                    in0[nt][1][i] = bits_bf16(f32_to_bf16_bits(b));
                }
            } else {
#pragma unroll
                for (int kb = 0; kb < KB0; ++kb) {
                    if (row < rows) {
                        in0[nt][kb] =
                            *reinterpret_cast<const bf16x8*>(in.x + (size_t)row * 64u + (uint32_t)kb * 16u + 8u * half);
                    } else {
#pragma unroll
                        for (int i = 0; i < 8; ++i) in0[nt][kb][i] = bits_bf16(0);
                    }
                }
            }
        }
        bf16x8 h1[2][8], h2[2][8];
        hidden_layer<KB0, kLd0, KB0>(sw, sb, lane, in0, h1, STORE ? in.h1 : nullptr, base, rows);
        hidden_layer<8, kLdH, 8>(sw + kOff1, sb + 128, lane, h1, h2, STORE ? in.h2 : nullptr, base, rows);
        hidden_layer<8, kLdH, 8>(sw + kOff2, sb + 256, lane, h2, h1, STORE ? in.h3 : nullptr, base, rows);
        // output layer: 32 (3 used) outputs, one tile per row group
        f32x16 acc[2] = {f32x16{}, f32x16{}};
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const bf16x8 a = lds8(sw + kOff3 + r * kLdH + (uint32_t)kb * 16u + 8u * half);
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h1[nt][kb], acc[nt], 0, 0, 0);
        }
        if (STORE) {   // all 32 outputs: element j of this lane is output 8 (j / 4) + 4 half + j % 4
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const uint32_t row = base + (uint32_t)nt * 32u + r;
                if (row < rows) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t f = 8u * (uint32_t)q + 4u * half;
                        *reinterpret_cast<float4*>(in.y32 + (size_t)row * 32u + f) =
                            make_float4(acc[nt][4 * q] + sb[384 + f], acc[nt][4 * q + 1] + sb[385 + f],
                                        acc[nt][4 * q + 2] + sb[386 + f], acc[nt][4 * q + 3] + sb[387 + f]);
                    }
                }
            }
        } else if (half == 0u) {   // elements 0..2 of the lower half are outputs 0..2
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const uint32_t row = base + (uint32_t)nt * 32u + r;
                if (row < rows)
                    y[row] = make_float4(acc[nt][0] + sb[384], acc[nt][1] + sb[385], acc[nt][2] + sb[386], 0.0f);
            }
        }
    }
}

}  // namespace

uint32_t mlp_fused_blocks(uint32_t rows) {
    int device = 0, cus = 256;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 256;
    const uint32_t groups = (rows + kWaves * 64u - 1u) / (kWaves * 64u);
    const uint32_t cap = (uint32_t)(cus > 0 ? cus : 256);
    return groups < cap ? (groups ? groups : 1u) : cap;
}

hipError_t launch_mlp_fused(const uint16_t* x, uint32_t rows, const uint16_t* wBf16, const float* wF32, float4* y,
                            uint32_t blocks, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (blocks == 0) blocks = mlp_fused_blocks(rows);
    FusedInput in{};
    in.x = x;
    k_mlp_fused<false><<<blocks, kThreads, 0, s>>>(in, rows, wBf16, wF32, y);
    return hipGetLastError();
}

hipError_t launch_mlp_fused_train(const uint16_t* x, uint32_t rows, const uint16_t* wBf16, const float* wF32,
                                  uint16_t* h1, uint16_t* h2, uint16_t* h3, float* y32, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (!x || !h1 || !h2 || !h3 || !y32) return hipErrorInvalidValue;
    FusedInput in{};
    in.x = x;
    in.h1 = h1;
    in.h2 = h2;
    in.h3 = h3;
    in.y32 = y32;
    k_mlp_fused<false, true><<<mlp_fused_blocks(rows), kThreads, 0, s>>>(in, rows, wBf16, wF32, nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_fused_sums(const float4* pos, const float4* sums, uint32_t total, uint32_t rows, float rho0,
                                 float h, float invV, const uint16_t* wBf16, const float* wF32, float4* y,
                                 uint32_t blocks, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (rows > total) return hipErrorInvalidValue;
    if (blocks == 0) blocks = mlp_fused_blocks(rows);
    FusedInput in{};
    in.pos = pos;
    in.sums = sums;
    in.total = total;
    in.rho0 = rho0;
    in.h = h;
    in.invV = invV;
    k_mlp_fused<true><<<blocks, kThreads, 0, s>>>(in, rows, wBf16, wF32, y);
    return hipGetLastError();
}

}  // namespace sphmlp
