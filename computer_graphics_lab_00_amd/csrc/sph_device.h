// sph_device.h — device-side types shared by the SPH kernels and the host driver.
//
// State layout in HBM (one solver):
//   pos4[n]  float4 (x, y, z, density)     bucket-sorted order, ping-pong pair
//   vel4[n]  float4 (vx, vy, vz, pressure) bucket-sorted order, ping-pong pair
//   ids[n]   uint32 particle index of each sorted slot, ping-pong pair
//   starts[C + 1]  first sorted slot of each hash bucket (exclusive scan of counts)
// Inside a bucket the slots are in ascending particle index, which is the order the
// reference's serial insertion produces (src/physics/fluid_sph.cpp:181-187), so the
// reference's `cellParticles[b * 32 + k]` is `ids[starts[b] + k]` for k < min(count, 32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_c_api.h"

namespace sphdev {

constexpr uint32_t kBlock = 256;          // 4 wave64 per workgroup
constexpr uint32_t kScanPerThread = 8;    // elements each scan thread owns
constexpr uint32_t kScanTile = kBlock * kScanPerThread;

// Scalars of one substep, precomputed on the host exactly as the reference computes
// them (src/physics/fluid_sph.cpp:24-46, :194-195, :264, :271-272) so device and
// oracle see identical values.
struct StepConsts {
    uint32_t n;
    uint32_t cellMask;      // cellCount - 1
    uint32_t mode;          // SphModeFlags
    uint32_t eos;           // SphEos
    float h;
    float h2;               // h * h (poly6's cut-off test)
    float r2Cut;            // smallest float x with sqrtf(x) >= h: r2 >= r2Cut <=> distance >= h
    float mass;
    float negMass;          // -mass
    float viscMass;         // viscosity * mass
    float cPoly6;           // 315 / (64 pi h^9)
    float cSpiky;           // -45 / (pi h^6)
    float cVisc;            // 45 / (pi h^6)
    float restDensity;
    float stiffness;
    float taitB;            // stiffness * restDensity / 7
    float dt;
    float maxSpeed;         // 0.4 h / dt
    float low[3];           // containerMin + radius
    float high[3];          // containerMax - radius
    float gravity[3];
    float radius;
    float restitution;
    uint32_t colliderCount;
    uint32_t fieldCount;
};

// Colliders and force fields travel in kernel arguments (at most 8 each).
struct BoundaryConsts {
    SphCollider colliders[SPH_MAX_COLLIDERS];
    SphForceField fields[SPH_MAX_FORCE_FIELDS];
};

struct GridBuffers {
    const float4* pos;      // state before the sort (any order)
    const float4* vel;
    const uint32_t* ids;
    uint32_t* keys;         // per pre-sort slot
    uint32_t* slots;        // atomic slot inside the bucket (arbitrary order)
    uint32_t* starts;       // C + 1
    uint2* pairs;           // (particle id, pre-sort slot) scattered by bucket
    uint32_t* blockSums;
    float4* posSorted;
    float4* velSorted;
    uint32_t* idsSorted;
    uint32_t* keysSorted;
};

// launchers (sph_grid.hip)
hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, uint32_t count, hipStream_t s);
hipError_t launch_keys(const StepConsts& k, const GridBuffers& g, hipStream_t s);
hipError_t launch_scan(uint32_t* data, uint32_t count, uint32_t total, uint32_t* blockSums, hipStream_t s);
hipError_t launch_scatter(const StepConsts& k, const GridBuffers& g, hipStream_t s);
hipError_t launch_rank_reorder(const StepConsts& k, const GridBuffers& g, hipStream_t s);
hipError_t launch_unpermute(const float4* src, const uint32_t* ids, float4* dst, uint32_t n, hipStream_t s);
hipError_t launch_gather_keys(const uint32_t* keysSorted, const uint32_t* ids, uint32_t* cellOf, uint32_t n,
                              hipStream_t s);
hipError_t launch_emit(const SphParams& p, uint32_t count, float4* pos, float4* vel, float4* prev, uint32_t* ids,
                       hipStream_t s);
hipError_t launch_iota(uint32_t* ids, uint32_t n, hipStream_t s);

// launchers (sph_forces.hip)
hipError_t launch_density(const StepConsts& k, float4* pos, float4* vel, const uint32_t* starts, hipStream_t s);
hipError_t launch_force(const StepConsts& k, const BoundaryConsts& b, const float4* pos, const float4* vel,
                        const uint32_t* starts, float4* posOut, float4* velOut, float4* forceOut, hipStream_t s);

}  // namespace sphdev
