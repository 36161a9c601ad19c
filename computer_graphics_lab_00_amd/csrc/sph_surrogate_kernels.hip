// sph_surrogate_kernels.hip — particle-side kernels of the learned surrogate (gfx950).
//
// The surrogate replaces the force pass of four substeps with one network evaluation:
// per particle, features summed over its neighbourhood by the density pass go through the
// MLP, the predicted SPH acceleration is held for a macro-step of 4
// substeps of symplectic Euler with the reference's speed clamp and walls
// (src/physics/fluid_sph.cpp:260-285), and the next density pass checks the result.
// Scenes with colliders are not handed to the surrogate: they run the plain path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_mlp.h"
#include "sph_surrogate.h"

namespace sphsur {

using sphdev::BoundaryConsts;
using sphdev::Launch;
using sphdev::StepConsts;
using sphdev::Where;
using sphdev::kBlock;

namespace {

__device__ inline float clampf_ref(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

// Feature rows (neighbourhood_features, zero padded to kFeatures BF16) of the prepared
// state, in slot order.
__global__ void __launch_bounds__(kBlock)
    k_features(Launch L, const float4* __restrict__ pos, const float4* __restrict__ sums, float vScale,
               uint16_t* __restrict__ x) {
    const Where w = locate(L);
    if (!w.valid) return;
    const StepConsts k = w.d->k;
    const uint32_t t = w.slot;
    float used[kUsedFeatures];
    neighbourhood_features(pos[t].w, sums[t], sums[(size_t)L.total + t], sums[2 * (size_t)L.total + t],
                           sums[3 * (size_t)L.total + t], sums[4 * (size_t)L.total + t], k.restDensity, k.h,
                           1.0f / vScale, used);
    float f[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) f[i] = i < (int)kUsedFeatures ? used[i] : 0.0f;
    uint4* row = reinterpret_cast<uint4*>(x + (size_t)t * kFeatures);
#pragma unroll
    for (int c = 0; c < kFeatures / 8; ++c) {
        uint32_t wv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            wv[e] = (uint32_t)sphmlp::f32_to_bf16_bits(f[8 * c + 2 * e]) |
                    ((uint32_t)sphmlp::f32_to_bf16_bits(f[8 * c + 2 * e + 1]) << 16);
        row[c] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
}

__device__ inline uint32_t knn_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & mask;
}

// (r2, id) before (r2', id'): nearer, or as near with the smaller particle index.
__device__ __forceinline__ bool knn_before(float r, uint32_t id, float r2, uint32_t id2) {
    return r < r2 || (r == r2 && id < id2);
}

// 8-nearest-neighbour feature rows (launch_knn_features in sph_surrogate.h), slot order.
__global__ void __launch_bounds__(kBlock)
    k_knn_features(Launch L, const float4* __restrict__ pos, const float4* __restrict__ vel,
                   const uint2* __restrict__ buckets, const uint32_t* __restrict__ ids, float invV,
                   uint16_t* __restrict__ x) {
    const Where w = locate(L);
    if (!w.valid) return;
    const StepConsts k = w.d->k;
    const uint32_t t = w.slot;
    const uint32_t cellBase = w.d->cellOffset;
    const bool capped = (k.mode & SPH_MODE_UNCAPPED) == 0;
    const float4 p = pos[t];
    const int32_t cx = (int32_t)floorf(p.x / k.h), cy = (int32_t)floorf(p.y / k.h), cz = (int32_t)floorf(p.z / k.h);
    // the 8 nearest so far, sorted by (r2, id); slot for the duplicate test
    float br[kKnn];
    uint32_t bid[kKnn], bslot[kKnn];
#pragma unroll
    for (uint32_t q = 0; q < kKnn; ++q) {
        br[q] = INFINITY;
        bid[q] = 0xFFFFFFFFu;
        bslot[q] = 0xFFFFFFFFu;
    }
    for (int32_t dz = -1; dz <= 1; ++dz)
        for (int32_t dy = -1; dy <= 1; ++dy)
            for (int32_t dx = -1; dx <= 1; ++dx) {
                const uint32_t b = cellBase + knn_hash(cx + dx, cy + dy, cz + dz, k.cellMask);
                const uint2 bd = buckets[b];
                uint32_t cnt = bd.y;
                if (cnt == 0) continue;
                if (capped && cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
                const uint32_t lo = bd.x;
                for (uint32_t j = 0; j < cnt; ++j) {
                    const uint32_t u = lo + j;
                    if (u == t) continue;
                    const float* q = reinterpret_cast<const float*>(pos + u);
                    const float ddx = p.x - q[0], ddy = p.y - q[1], ddz = p.z - q[2];
                    const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                    if (!(r2 < k.h2)) continue;
                    const uint32_t id = ids[u];
                    if (!knn_before(r2, id, br[kKnn - 1], bid[kKnn - 1])) continue;
                    bool dup = false;   // a bucket two neighbour cells share is visited twice
#pragma unroll
                    for (uint32_t e = 0; e < kKnn; ++e) dup = dup || bslot[e] == u;
                    if (dup) continue;
                    // insertion into the sorted registers, from the back
#pragma unroll
                    for (uint32_t e = kKnn - 1; e > 0; --e) {
                        const bool here = knn_before(r2, id, br[e], bid[e]);
                        const bool above = knn_before(r2, id, br[e - 1], bid[e - 1]);
                        br[e] = here ? (above ? br[e - 1] : r2) : br[e];
                        bid[e] = here ? (above ? bid[e - 1] : id) : bid[e];
                        bslot[e] = here ? (above ? bslot[e - 1] : u) : bslot[e];
                    }
                    if (knn_before(r2, id, br[0], bid[0])) {
                        br[0] = r2;
                        bid[0] = id;
                        bslot[0] = u;
                    }
                }
            }
    const float4 v = vel[t];
    const float invH = 1.0f / k.h;
    float f[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) f[i] = 0.0f;
#pragma unroll
    for (uint32_t e = 0; e < kKnn; ++e) {
        if (bslot[e] == 0xFFFFFFFFu) continue;
        const float4 q = pos[bslot[e]];
        const float4 qv = vel[bslot[e]];
        f[6 * e + 0] = (p.x - q.x) * invH;
        f[6 * e + 1] = (p.y - q.y) * invH;
        f[6 * e + 2] = (p.z - q.z) * invH;
        f[6 * e + 3] = (qv.x - v.x) * invV;
        f[6 * e + 4] = (qv.y - v.y) * invV;
        f[6 * e + 5] = (qv.z - v.z) * invV;
    }
    f[6 * kKnn] = (p.w / k.restDensity - 1.0f) * kDensityScale;
    f[6 * kKnn + 1] = k.h / kLengthScale;
    uint4* row = reinterpret_cast<uint4*>(x + (size_t)t * kFeatures);
#pragma unroll
    for (int c = 0; c < kFeatures / 8; ++c) {
        uint32_t wv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            wv[e] = (uint32_t)sphmlp::f32_to_bf16_bits(f[8 * c + 2 * e]) |
                    ((uint32_t)sphmlp::f32_to_bf16_bits(f[8 * c + 2 * e + 1]) << 16);
        row[c] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
}

__device__ void field_accel(const SphForceField& fd, const float p[3], float out[3]) {
    float off[3] = {p[0] - fd.position[0], p[1] - fd.position[1], p[2] - fd.position[2]};
    float distance = sqrtf((off[0] * off[0] + off[1] * off[1]) + off[2] * off[2]);
    float weight = 1.0f;
    if (fd.radius > 0.0f) {
        float tt = clampf_ref(1.0f - distance / fd.radius, 0.0f, 1.0f);
        weight = tt <= 0.0f ? 0.0f : powf(tt, fd.falloff);
    }
    out[0] = out[1] = out[2] = 0.0f;
    if (weight <= 0.0f) return;
    const float sw = fd.strength * weight;
    if (fd.type == SPH_FIELD_WIND) {
        out[0] = fd.axis[0] * sw; out[1] = fd.axis[1] * sw; out[2] = fd.axis[2] * sw;
    } else if (fd.type == SPH_FIELD_VORTEX) {
        float along = (off[0] * fd.axis[0] + off[1] * fd.axis[1]) + off[2] * fd.axis[2];
        float rad[3] = {off[0] - fd.axis[0] * along, off[1] - fd.axis[1] * along, off[2] - fd.axis[2] * along};
        float r = sqrtf((rad[0] * rad[0] + rad[1] * rad[1]) + rad[2] * rad[2]);
        if (r < 1.0e-4f) return;
        float u[3] = {rad[0] / r, rad[1] / r, rad[2] / r};
        out[0] = (fd.axis[1] * u[2] - u[1] * fd.axis[2]) * sw;
        out[1] = (fd.axis[2] * u[0] - u[2] * fd.axis[0]) * sw;
        out[2] = (fd.axis[0] * u[1] - u[0] * fd.axis[1]) * sw;
    } else if (fd.type == SPH_FIELD_POINT) {
        if (distance < 1.0e-4f) return;
        out[0] = (-off[0] / distance) * sw;
        out[1] = (-off[1] / distance) * sw;
        out[2] = (-off[2] / distance) * sw;
    }
}

// `substeps` substeps with the network's SPH acceleration held fixed; gravity, fields,
// the speed clamp and the walls are applied each substep as the plain step does. When
// posKeep / velKeep are set, the state the substeps start from is written there too.
__global__ void __launch_bounds__(kBlock)
    k_integrate_macro(Launch L, const float4* __restrict__ pos, const float4* __restrict__ vel,
                      const float* __restrict__ y, uint32_t ldY, float aScale, uint32_t substeps,
                      const float4* __restrict__ sums, float smoothing, float4* __restrict__ posOut,
                      float4* __restrict__ velOut, float4* __restrict__ posKeep, float4* __restrict__ velKeep) {
    const Where w = locate(L);
    if (!w.valid) return;
    const StepConsts k = w.d->k;
    const uint32_t t = w.slot;
    const float4 p0 = pos[t], v0 = vel[t];
    if (posKeep) {
        posKeep[t] = p0;
        velKeep[t] = v0;
    }
    const float asph[3] = {y[(size_t)t * ldY + 0] * aScale, y[(size_t)t * ldY + 1] * aScale,
                           y[(size_t)t * ldY + 2] * aScale};
    const BoundaryConsts& bnd = L.bounds[L.blockScene[sphdev::xcd_block(blockIdx.x, gridDim.x)]];
    float pp[3] = {p0.x, p0.y, p0.z};
    float vv[3] = {v0.x, v0.y, v0.z};
    if (smoothing > 0.0f) {
        // XSPH-style smoothing (Monaghan 1989): move the velocity toward the kernel-weighted
        // mean of its neighbours', sum s w / sum s, which damps the particle-scale velocity
        // noise that prediction errors inject.
        const float4 u = sums[3 * (size_t)L.total + t];
        const float wsum = sums[2 * (size_t)L.total + t].w;
        if (wsum > 0.0f) {
            vv[0] += smoothing * (u.x / wsum);
            vv[1] += smoothing * (u.y / wsum);
            vv[2] += smoothing * (u.z / wsum);
        }
    }
    for (uint32_t s = 0; s < substeps; ++s) {
        float fa[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t fi = 0; fi < k.fieldCount; ++fi) {
            float a[3];
            field_accel(bnd.fields[fi], pp, a);
            fa[0] += a[0]; fa[1] += a[1]; fa[2] += a[2];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) vv[a] += ((asph[a] + k.gravity[a]) + fa[a]) * k.dt;
        const float speed = sqrtf((vv[0] * vv[0] + vv[1] * vv[1]) + vv[2] * vv[2]);
        if (speed > k.maxSpeed) {
            const float sc = k.maxSpeed / speed;
            vv[0] *= sc; vv[1] *= sc; vv[2] *= sc;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) pp[a] += vv[a] * k.dt;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (pp[a] < k.low[a]) {
                pp[a] = k.low[a];
                if (vv[a] < 0.0f) vv[a] = -vv[a] * k.restitution;
            } else if (pp[a] > k.high[a]) {
                pp[a] = k.high[a];
                if (vv[a] > 0.0f) vv[a] = -vv[a] * k.restitution;
            }
        }
    }
    posOut[t] = make_float4(pp[0], pp[1], pp[2], p0.w);
    velOut[t] = make_float4(vv[0], vv[1], vv[2], v0.w);
}

// Per-block FP32 sums of max(rho / rho0 - 1, 0) and |rho / rho0 - 1| over the prepared state.
__global__ void __launch_bounds__(kBlock)
    k_deviation(Launch L, const float4* __restrict__ pos, float2* __restrict__ partial) {
    __shared__ float sa[kBlock], sb[kBlock];
    const Where w = locate(L);
    float a = 0.0f, b = 0.0f;
    if (w.valid) {
        const float d = pos[w.slot].w / w.d->k.restDensity - 1.0f;
        a = d > 0.0f ? d : 0.0f;
        b = d < 0.0f ? -d : d;
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_float2(sa[0], sb[0]);
}

// One block: the block sums added in FP64 in a fixed order, scaled by invCount, into
// out[0], out[1] (host-mapped memory, read once the launch has completed).
__global__ void __launch_bounds__(kBlock)
    k_deviation_total(const float2* __restrict__ partial, uint32_t blocks, double invCount, double* __restrict__ out) {
    __shared__ double da[kBlock], db[kBlock];
    double ta = 0.0, tb = 0.0;
    for (uint32_t i = threadIdx.x; i < blocks; i += kBlock) {
        const float2 q = partial[i];
        ta += (double)q.x;
        tb += (double)q.y;
    }
    da[threadIdx.x] = ta;
    db[threadIdx.x] = tb;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            da[threadIdx.x] += da[threadIdx.x + s];
            db[threadIdx.x] += db[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = da[0] * invCount;
        out[1] = db[0] * invCount;
        __threadfence_system();
    }
}

__global__ void __launch_bounds__(kBlock)
    k_gather_rows(const uint16_t* __restrict__ x, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ slots,
                  uint32_t count, uint16_t* __restrict__ dst, uint32_t* __restrict__ dstIds) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= count) return;
    const uint32_t s = slots[r];
    const uint4* src = reinterpret_cast<const uint4*>(x + (size_t)s * kFeatures);
    uint4* out = reinterpret_cast<uint4*>(dst + (size_t)r * kFeatures);
#pragma unroll
    for (int c = 0; c < kFeatures / 8; ++c) out[c] = src[c];
    dstIds[r] = ids[s];
}

__global__ void __launch_bounds__(kBlock)
    k_accumulate_accel(const float4* __restrict__ force, const float4* __restrict__ pos,
                       const uint32_t* __restrict__ ids, uint32_t n, float4* __restrict__ acc) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const float4 f = force[t];
    float rho = pos[t].w;
    rho = rho < 1e-6f ? 1e-6f : rho;
    float4 a = acc[ids[t]];
    a.x += f.x / rho;
    a.y += f.y / rho;
    a.z += f.z / rho;
    acc[ids[t]] = a;
}

__global__ void __launch_bounds__(kBlock)
    k_targets(const float4* __restrict__ acc, const uint32_t* __restrict__ sampleIds, uint32_t count, float scale,
              float* __restrict__ t) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= count) return;
    const float4 a = acc[sampleIds[r]];
    t[(size_t)r * 3 + 0] = a.x * scale;
    t[(size_t)r * 3 + 1] = a.y * scale;
    t[(size_t)r * 3 + 2] = a.z * scale;
}

__global__ void __launch_bounds__(kBlock)
    k_batch_gather(const uint16_t* __restrict__ x, const float* __restrict__ t, const uint32_t* __restrict__ order,
                   uint32_t count, uint32_t rowsPadded, uint16_t* __restrict__ xb, float* __restrict__ tb) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= rowsPadded) return;
    uint4* out = reinterpret_cast<uint4*>(xb + (size_t)r * kFeatures);
    if (r < count) {
        const uint32_t s = order[r];
        const uint4* src = reinterpret_cast<const uint4*>(x + (size_t)s * kFeatures);
#pragma unroll
        for (int c = 0; c < kFeatures / 8; ++c) out[c] = src[c];
        tb[(size_t)r * 3 + 0] = t[(size_t)s * 3 + 0];
        tb[(size_t)r * 3 + 1] = t[(size_t)s * 3 + 1];
        tb[(size_t)r * 3 + 2] = t[(size_t)s * 3 + 2];
    } else {
#pragma unroll
        for (int c = 0; c < kFeatures / 8; ++c) out[c] = make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ void __launch_bounds__(kBlock) k_zero_f4(float4* __restrict__ a, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void __launch_bounds__(kBlock)
    k_unpermute_rows(const uint16_t* __restrict__ x, const uint32_t* __restrict__ ids, uint32_t n,
                     uint16_t* __restrict__ out) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint4* src = reinterpret_cast<const uint4*>(x + (size_t)t * kFeatures);
    uint4* dst = reinterpret_cast<uint4*>(out + (size_t)ids[t] * kFeatures);
#pragma unroll
    for (int c = 0; c < kFeatures / 8; ++c) dst[c] = src[c];
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_features(const Launch& L, const float4* pos, const float4* sums, float vScale, uint16_t* x,
                           hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    k_features<<<L.blocks, kBlock, 0, s>>>(L, pos, sums, vScale, x);
    return hipGetLastError();
}

hipError_t launch_knn_features(const Launch& L, const float4* pos, const float4* vel, const uint2* buckets,
                               const uint32_t* ids, float vScale, uint16_t* x, hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    k_knn_features<<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, ids, 1.0f / vScale, x);
    return hipGetLastError();
}

hipError_t launch_integrate_macro(const Launch& L, const float4* pos, const float4* vel, const float* y, uint32_t ldY,
                                  float aScale, uint32_t substeps, const float4* sums, float smoothing,
                                  float4* posOut, float4* velOut, float4* posKeep, float4* velKeep, hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    k_integrate_macro<<<L.blocks, kBlock, 0, s>>>(L, pos, vel, y, ldY, aScale, substeps, sums, smoothing, posOut,
                                                  velOut, posKeep, velKeep);
    return hipGetLastError();
}

hipError_t launch_deviation(const Launch& L, const float4* pos, float2* partial, double* out, double invCount,
                            hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    k_deviation<<<L.blocks, kBlock, 0, s>>>(L, pos, partial);
    k_deviation_total<<<1, kBlock, 0, s>>>(partial, L.blocks, invCount, out);
    return hipGetLastError();
}

hipError_t launch_deviation_total(const Launch& L, const float2* partial, double* out, double invCount,
                                  hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    k_deviation_total<<<1, kBlock, 0, s>>>(partial, L.blocks, invCount, out);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const uint16_t* x, const uint32_t* ids, const uint32_t* slots, uint32_t count,
                              uint16_t* dst, uint32_t* dstIds, hipStream_t s) {
    if (count == 0) return hipSuccess;
    k_gather_rows<<<blocks_for(count), kBlock, 0, s>>>(x, ids, slots, count, dst, dstIds);
    return hipGetLastError();
}

hipError_t launch_accumulate_accel(const float4* force, const float4* pos, const uint32_t* ids, uint32_t n,
                                   float4* acc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_accumulate_accel<<<blocks_for(n), kBlock, 0, s>>>(force, pos, ids, n, acc);
    return hipGetLastError();
}

hipError_t launch_targets(const float4* acc, const uint32_t* sampleIds, uint32_t count, float scale, float* t,
                          hipStream_t s) {
    if (count == 0) return hipSuccess;
    k_targets<<<blocks_for(count), kBlock, 0, s>>>(acc, sampleIds, count, scale, t);
    return hipGetLastError();
}

hipError_t launch_batch_gather(const uint16_t* x, const float* t, const uint32_t* order, uint32_t count,
                               uint32_t rowsPadded, uint16_t* xb, float* tb, hipStream_t s) {
    if (rowsPadded == 0) return hipSuccess;
    k_batch_gather<<<blocks_for(rowsPadded), kBlock, 0, s>>>(x, t, order, count, rowsPadded, xb, tb);
    return hipGetLastError();
}

hipError_t launch_zero_f4(float4* a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_zero_f4<<<blocks_for(n), kBlock, 0, s>>>(a, n);
    return hipGetLastError();
}

hipError_t launch_unpermute_rows(const uint16_t* x, const uint32_t* ids, uint32_t n, uint16_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_unpermute_rows<<<blocks_for(n), kBlock, 0, s>>>(x, ids, n, out);
    return hipGetLastError();
}

}  // namespace sphsur
