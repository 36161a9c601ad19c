// sph_arith_check.hip — the GPU check behind the force pass's shared-reciprocal division and
// square root (sph_exact.h): runs the device helpers over their admitted operand ranges and
// counts results that differ from the IEEE operations (the compiler's correctly rounded
// division and square root).
//   0. recip_rn(b) == 1 / b for every float b in [kDivDenLo, kDivDenHi];
//   1. sqrt_rn(x) == sqrtf(x) for every float x in [kSqrtLo, kSqrtHi], and sqrt_rn(x) <= 1e-6
//      for every float x in [0, kSqrtLo);
//   2. div_rn(a, b, recip_rn(b)) == a / b for pseudo-random pairs with |a| in
//      [kDivNumLo, kDivNumHi], b in [kDivDenLo, kDivDenHi], plus every significand of b
//      against numerators whose significand is one or all ones.
#include <hip/hip_runtime.h>

#include <string>

#include "sph_c_api.h"
#include "sph_exact.h"
#include "sph_internal.h"

namespace {

using namespace sphdev;
using sphhost::fail;

__device__ inline void count(unsigned long long* c, bool bad) {
    const uint64_t m = __ballot(bad);
    if (m && (threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)m) - 1))
        atomicAdd(c, (unsigned long long)__popcll(m));
}

__global__ void __launch_bounds__(256) k_check_recip(uint32_t lo, uint32_t hi, unsigned long long* bad) {
    for (uint64_t u = lo + (uint64_t)blockIdx.x * 256u + threadIdx.x; u <= hi; u += (uint64_t)gridDim.x * 256u) {
        const float b = __uint_as_float((uint32_t)u);
        count(bad, __float_as_uint(recip_rn(b)) != __float_as_uint(1.0f / b));
    }
}

__global__ void __launch_bounds__(256) k_check_sqrt(uint32_t exactLo, uint32_t hi, unsigned long long* bad) {
    for (uint64_t u = (uint64_t)blockIdx.x * 256u + threadIdx.x; u <= hi; u += (uint64_t)gridDim.x * 256u) {
        const float x = __uint_as_float((uint32_t)u);
        const float s = sqrt_rn(x);
        const bool ok = u >= exactLo ? __float_as_uint(s) == __float_as_uint(sqrtf(x)) : s <= 1e-6f;
        count(bad, !ok);
    }
}

__device__ inline uint32_t mix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

// Floats with magnitude in [2^elo, 2^ehi]: exponent drawn uniformly, significand from r.
__device__ inline float draw(uint32_t r, uint32_t e, int32_t elo, int32_t ehi) {
    const uint32_t span = (uint32_t)(ehi - elo);
    const uint32_t be = (uint32_t)(127 + elo) + e % span;
    return __uint_as_float((be << 23) | (r & 0x7FFFFFu));
}

__global__ void __launch_bounds__(256)
    k_check_div(uint64_t samples, unsigned long long* bad, unsigned long long* zeroBad) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < samples; i += (uint64_t)gridDim.x * 256u) {
        const uint32_t r0 = mix(2 * i), r1 = mix(2 * i + 1);
        const uint32_t r2 = mix(i ^ 0x5bd1e995ull), r3 = mix(i * 0x9E3779B97F4A7C15ull);
        float a = draw(r0, r2, -96, 50);
        a = (r3 & 1u) ? -a : a;
        const float b = draw(r1, r3 >> 1, -20, 26);
        const float y = recip_rn(b);
        count(bad, __float_as_uint(div_rn(a, b, y)) != __float_as_uint(a / b));
        // significands of b against numerators 1.0 and 1.11...1 (the hardest quotients)
        if (i < (1ull << 23)) {
            const float bb = __uint_as_float(0x3F800000u | (uint32_t)i);
            const float yb = recip_rn(bb);
            const float a1 = __uint_as_float(0x3F800000u), a2 = __uint_as_float(0x3FFFFFFFu);
            count(bad, __float_as_uint(div_rn(a1, bb, yb)) != __float_as_uint(a1 / bb) ||
                           __float_as_uint(div_rn(a2, bb, yb)) != __float_as_uint(a2 / bb));
            count(zeroBad, div_rn(0.0f, bb, yb) != 0.0f || div_rn(-0.0f, bb, yb) != 0.0f);
        }
    }
}

}  // namespace

extern "C" int sph_check_exact_arithmetic(int32_t device, uint64_t samples, uint64_t out[8]) {
    if (!out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_check_exact_arithmetic: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(SPH_ERR_INVALID_ARGUMENT, "device index out of range");
    unsigned long long* c = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&c, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c, 0, 4 * sizeof(unsigned long long));
    const uint32_t rlo = __builtin_bit_cast(uint32_t, kDivDenLo), rhi = __builtin_bit_cast(uint32_t, kDivDenHi);
    const uint32_t slo = __builtin_bit_cast(uint32_t, kSqrtLo), shi = __builtin_bit_cast(uint32_t, kSqrtHi);
    const uint64_t divSamples = samples < (1ull << 23) ? (1ull << 23) : samples;
    if (e == hipSuccess) {
        k_check_recip<<<4096, 256>>>(rlo, rhi, c + 0);
        k_check_sqrt<<<4096, 256>>>(slo, shi, c + 1);
        k_check_div<<<4096, 256>>>(divSamples, c + 2, c + 3);
        e = hipGetLastError();
    }
    unsigned long long h[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpy(h, c, sizeof(h), hipMemcpyDeviceToHost);
    if (c) (void)hipFree(c);
    if (e != hipSuccess) return fail(SPH_ERR_DEVICE, std::string("sph_check_exact_arithmetic: ") + hipGetErrorString(e));
    out[0] = (uint64_t)rhi - rlo + 1u;   // reciprocals checked
    out[1] = h[0];
    out[2] = (uint64_t)shi + 1u;          // square roots checked
    out[3] = h[1];
    out[4] = divSamples + 2 * (1ull << 23);   // quotients checked
    out[5] = h[2];
    out[6] = 2 * (1ull << 23);                // zero numerators checked
    out[7] = h[3];
    return SPH_OK;
}
