// sph_exact.h — correctly rounded division and square root built from the hardware reciprocal
// and square root, for operands inside stated ranges (gfx950).
//
// The sequences the compiler emits for an IEEE a / b and sqrtf(x) accept every operand
// (scaling of tiny and huge values, fix-ups for zeros, infinities and NaNs): 11 instructions
// per division, 16 per square root. The force pass divides by two values per pair (the
// distance three times, the neighbour's density four times) and its operands lie in narrow
// ranges, so inside those ranges it uses:
//
//   recip_rn(b)     = RN(1 / b)   y0 = v_rcp_f32(b), e = fma(-b, y0, 1), y = fma(e, y0, y0)
//   div_rn(a, b, y) = RN(a / b)   given y = RN(1 / b):
//                                 q0 = a y, r0 = fma(-b, q0, a), q1 = fma(r0, y, q0),
//                                 r1 = fma(-b, q1, a), q = fma(r1, y, q1)
//   sqrt_rn(x)      = RN(sqrt x)  s = v_sqrt_f32(x), then the neighbour s - ulp or s + ulp when
//                                 the residual x - s' s says the root lies beyond the midpoint
//
// RN is round-to-nearest-even, u = 2^-24. Why the results are the IEEE ones:
//  - Reciprocal. v_rcp_f32 is faithful (error below one ulp), so y0 is one of the two floats
//    bracketing 1 / b. One Newton step from either gives RN(1 / b) for every significand of b
//    except all ones (1.11...1 x 2^k) started from the lower bracket, where the step lands on
//    a tie and rounds to the even neighbour (a host check runs every significand from both
//    starts). Whether the hardware ever starts there is a fact about v_rcp_f32; the GPU check
//    sph_check_exact_arithmetic runs every divisor in the admitted range through this code
//    and compares it with IEEE 1 / b.
//  - Division. With y = RN(1 / b), q0 = RN(a y) lies within (2 + u) ulp of a / b. The first
//    correction computes q0 + r0 y with r0 = RN(a - b q0), which is a / b up to
//    |q0 - a / b| (2u + u^2), a few 2^-22 ulp, so q1 is within 1/2 + 2^-21 ulp: faithful. For
//    a faithful q1 the remainder a - b q1 is a float (the fma computes it exactly), and
//    Markstein's theorem makes RN(q1 + r1 y) = RN(a / b) when y is within half an ulp of
//    1 / b. The argument needs every intermediate normal and every remainder representable:
//    |a| in [2^-96, 2^50] or a = 0, and b in [2^-20, 2^26] (then q lies in [2^-122, 2^70] and
//    the remainders are multiples of at least 2^-143). A zero numerator yields a zero whose
//    sign may differ from the IEEE quotient's (+0 for -0 / b); the force pass adds these
//    quotients' products to sums that start at +0 and never become -0, so a zero term of
//    either sign leaves them unchanged.
//  - Square root. v_sqrt_f32 is faithful; for x in [2^-96, 2^52] the correctly rounded root is
//    s or an immediate neighbour, and the exact fma residuals x - (s - ulp) s and
//    x - (s + ulp) s have the signs that pick it (the GPU check runs every such x). Below
//    2^-96 the result stays below 1e-6, where the force pass discards the pair anyway.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphdev {

// Operand ranges inside which div_rn and sqrt_rn return the IEEE results.
constexpr float kDivNumLo = 0x1p-96f;    // |a| >= this, or a == 0
constexpr float kDivNumHi = 0x1p50f;     // |a| <= this
constexpr float kDivDenLo = 0x1p-20f;    // b >= this (the force pass's divisors exceed 1e-6)
constexpr float kDivDenHi = 0x1p26f;     // b <= this
constexpr float kSqrtLo = 0x1p-96f;      // sqrt_rn(x) == sqrtf(x) for x in [kSqrtLo, kSqrtHi]
constexpr float kSqrtHi = 0x1p52f;

__device__ __forceinline__ float recip_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, y0, 1.0f);
    return fmaf(e, y0, y0);
}

__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    const float r0 = fmaf(-b, q0, a);
    const float q1 = fmaf(r0, y, q0);
    const float r1 = fmaf(-b, q1, a);
    return fmaf(r1, y, q1);
}

__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = __uint_as_float(__float_as_uint(s) - 1u);
    const float above = __uint_as_float(__float_as_uint(s) + 1u);
    const float t = fmaf(-below, s, x) <= 0.0f ? below : s;
    return fmaf(-above, s, x) > 0.0f ? above : t;
}

}  // namespace sphdev
