// sph_surrogate.h — particle-side kernels of the learned surrogate and its constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.h"

namespace sphsur {

constexpr uint32_t kFeatures = 64;   // 20 neighbourhood features, zero padded to one MFMA k-block multiple
constexpr uint32_t kHidden = 128;
constexpr uint32_t kOutPadded = 32;  // 3 outputs padded to one MFMA tile
constexpr uint32_t kOutputs = 3;
constexpr uint32_t kMacro = 4;       // substeps per surrogate macro-step
constexpr float kVelocityScale = 1.0f;   // m/s
constexpr float kAccelScale = 98.1f;     // m/s^2: targets and outputs are in units of 10 |g|
constexpr float kDensityScale = 100.0f;  // density deviations and gradients enter in percent
constexpr float kLengthScale = 0.01f;    // m: the kernel-radius feature is h / 1 cm

constexpr uint32_t kUsedFeatures = 20;

// The neighbourhood features of one particle (the first kUsedFeatures of its row; the
// rest are 0), from its density and its five rows of neighbourhood sums (NeighbourSums,
// sph_device.h):
//   0       : 100 (rho / rho0 - 1)
//   1 .. 3  : sum s^2 e            (shape of the density gradient)
//   4 .. 9  : sum s^3 e e^T        (xx, yy, zz, xy, xz, yz: the neighbourhood's anisotropy)
//   10 .. 12: sum s w / vScale     (smoothed velocity difference: the viscous pull)
//   13      : sum s^2 (w . e) / vScale   (compression rate)
//   14 .. 16: sum s^2 (w x e) / vScale   (rotation)
//   17      : in-range count / 32
//   18      : sum s / 8
//   19      : h / kLengthScale     (one network serves several resolutions)
// All are relative to the particle (translation invariant); vectors are in world axes,
// so the network also sees which way gravity points.
__device__ __forceinline__ void neighbourhood_features(float rho, float4 r0, float4 r1, float4 r2, float4 r3,
                                                       float4 r4, float rho0, float h, float invV,
                                                       float (&f)[kUsedFeatures]) {
    f[0] = (rho / rho0 - 1.0f) * kDensityScale;
    f[1] = r0.x;
    f[2] = r0.y;
    f[3] = r0.z;
    f[4] = r1.x;
    f[5] = r1.y;
    f[6] = r1.z;
    f[7] = r1.w;
    f[8] = r2.x;
    f[9] = r2.y;
    f[10] = r3.x * invV;
    f[11] = r3.y * invV;
    f[12] = r3.z * invV;
    f[13] = r2.z * invV;
    f[14] = r4.x * invV;
    f[15] = r4.y * invV;
    f[16] = r4.z * invV;
    f[17] = r0.w * (1.0f / 32.0f);
    f[18] = r2.w * (1.0f / 8.0f);
    f[19] = h / kLengthScale;
}

hipError_t launch_features(const sphdev::Launch& L, const float4* pos, const float4* sums, float vScale, uint16_t* x,
                           hipStream_t s);

// Feature kinds of a network's input rows.
constexpr uint32_t kFeaturesMoments = 0;   // neighbourhood_features from the density pass's sums
constexpr uint32_t kFeaturesKnn8 = 1;      // the 8 nearest neighbours (layout below)
constexpr uint32_t kKnn = 8;

// The 8-nearest-neighbour rows of the prepared state (density already in pos.w), slot order:
// the neighbours are the in-range candidates of the density pass's walk (27 buckets, the
// 32-slot cap, r2 < h2 computed as the pass does), distinct particles other than the
// particle itself, the 8 smallest by (r2, particle index), nearest first. For neighbour k:
//   6k .. 6k+2 : (x_i - x_j) / h
//   6k+3 .. 6k+5 : (v_j - v_i) / vScale
// (zeros for missing neighbours), then 48: 100 (rho / rho0 - 1), 49: h / kLengthScale.
hipError_t launch_knn_features(const sphdev::Launch& L, const float4* pos, const float4* vel, const uint2* buckets,
                               const uint32_t* ids, float vScale, uint16_t* x, hipStream_t s);
hipError_t launch_integrate_macro(const sphdev::Launch& L, const float4* pos, const float4* vel, const float* y,
                                  uint32_t ldY, float aScale, uint32_t substeps, const float4* sums,
                                  float smoothing, float4* posOut, float4* velOut, float4* posKeep,
                                  float4* velKeep, hipStream_t s);
hipError_t launch_deviation(const sphdev::Launch& L, const float4* pos, float2* partial, double* out, double invCount,
                            hipStream_t s);
// The second stage alone, over workgroup sums a density pass with NeighbourSums::deviation
// has written.
hipError_t launch_deviation_total(const sphdev::Launch& L, const float2* partial, double* out, double invCount,
                                  hipStream_t s);
hipError_t launch_gather_rows(const uint16_t* x, const uint32_t* ids, const uint32_t* slots, uint32_t count,
                              uint16_t* dst, uint32_t* dstIds, hipStream_t s);
hipError_t launch_accumulate_accel(const float4* force, const float4* pos, const uint32_t* ids, uint32_t n,
                                   float4* acc, hipStream_t s);
hipError_t launch_targets(const float4* acc, const uint32_t* sampleIds, uint32_t count, float scale, float* t,
                          hipStream_t s);
hipError_t launch_batch_gather(const uint16_t* x, const float* t, const uint32_t* order, uint32_t count,
                               uint32_t rowsPadded, uint16_t* xb, float* tb, hipStream_t s);
hipError_t launch_zero_f4(float4* a, uint32_t n, hipStream_t s);
hipError_t launch_unpermute_rows(const uint16_t* x, const uint32_t* ids, uint32_t n, uint16_t* out, hipStream_t s);

}  // namespace sphsur
