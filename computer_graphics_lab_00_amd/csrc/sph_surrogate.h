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

hipError_t launch_features(const sphdev::Launch& L, const float4* pos, const float4* sums, float vScale, uint16_t* x,
                           hipStream_t s);
hipError_t launch_integrate_macro(const sphdev::Launch& L, const float4* pos, const float4* vel, const float* y,
                                  uint32_t ldY, float aScale, uint32_t substeps, const float4* sums,
                                  float smoothing, float4* posOut, float4* velOut, hipStream_t s);
hipError_t launch_deviation(const sphdev::Launch& L, const float4* pos, float2* partial, hipStream_t s);
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
