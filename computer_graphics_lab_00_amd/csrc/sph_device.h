// sph_device.h — device-side types shared by the SPH kernels and the host engine.
//
// One engine advances S independent scenes (S = 1 for a plain solver) with the same
// kernels: every launch covers all scenes, and each workgroup finds its scene in a
// per-block table, so a batch of uneven scenes costs one launch per kernel.
//
// State layout in HBM, global slot space = sum of scene capacities:
//   pos4[slot]  float4 (x, y, z, density)     ping-pong pair
//   vel4[slot]  float4 (vx, vy, vz, pressure) ping-pong pair
//   ids[slot]   uint32 particle index within its scene, ping-pong pair
//   buckets[bucket] uint2 (first slot, member count): hash buckets of all scenes, scene by
//               scene; one 8-byte load gives a walk both numbers
// Scene s owns slots [offset_s, offset_s + capacity_s), its particles the first n_s.
// Inside a scene slots are ordered by space (spatial bin of the bucket), then bucket,
// then particle index. Every bucket is contiguous and ascending in index — the order of
// the reference's serial insertion (src/physics/fluid_sph.cpp:181-187) — so its
// `cellParticles[b * 32 + k]` is `ids[buckets[b].x + k]` for k < min(buckets[b].y, 32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_c_api.h"
#include "sph_exact.h"

namespace sphdev {

constexpr uint32_t kBlock = 256;          // 4 wave64 per workgroup
constexpr uint32_t kScanPerThread = 8;    // elements each scan thread owns
constexpr uint32_t kScanTile = kBlock * kScanPerThread;
constexpr uint32_t kXcds = 8;             // MI355X: 8 XCDs, each with its own L2
constexpr uint32_t kPosPad = 8;           // slots of padding after each position array: the
                                          // density pass reads candidates nine at a time
constexpr uint32_t kHalfTileBins = 32;    // spatial bins per workgroup of the tiled density pass

// Scalars of one substep of one scene, precomputed on the host exactly as the reference
// computes them (src/physics/fluid_sph.cpp:24-46, :194-195, :264, :271-272) so device and
// oracle see identical values.
struct StepConsts {
    uint32_t n;             // particles in the scene
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
    uint32_t listCap;       // neighbour-list slots per particle (0: no list)
    uint32_t shortDivide;   // 1: h, mass and viscosity * mass admit the force pass's short division
};

// Spatial bins that order a scene's buckets in memory: the container's cells (plus a
// margin), coarsened by 2^shift per axis if needed, in tiles of 2^tileLog2 cells per
// axis (64 bins per tile; Morton order inside a 4x4x4 tile, x-fastest otherwise).
struct BinGrid {
    int32_t origin[3];      // cell coordinate of bin (0, 0, 0)
    uint32_t dims[3];       // bins per axis
    uint32_t tiles[3];      // tiles per axis
    uint32_t tileLog2[3];   // sums to 6
    uint32_t shift;
    uint32_t bins;          // tiles product * 64
};

struct SceneDesc {
    StepConsts k;
    BinGrid g;
    uint32_t offset;        // first slot
    uint32_t capacity;      // slots reserved
    uint32_t cellOffset;    // first hash bucket
    uint32_t binOffset;     // first spatial bin
    uint32_t binCapacity;   // bins reserved; the filler bin binOffset + binCapacity holds capacity - n
    uint32_t blockOffset;   // first workgroup of the scene in particle-parallel launches
    uint32_t active;        // 0: this substep is past the scene's substep count; state passes through
    uint32_t tileOffset;    // first workgroup of the scene in half-tile launches (binCapacity / 32 of them)
};

// Colliders and force fields of one scene (at most 8 each).
struct BoundaryConsts {
    SphCollider colliders[SPH_MAX_COLLIDERS];
    SphForceField fields[SPH_MAX_FORCE_FIELDS];
};

// Everything a batched kernel needs to find its scene and the global arrays.
struct Launch {
    const SceneDesc* scenes;
    const BoundaryConsts* bounds;
    const uint16_t* blockScene;   // scene of each logical workgroup
    uint32_t blocks;              // particle-parallel workgroups
    uint32_t numScenes;
    uint32_t total;               // slots (sum of capacities): neighbour-list stride
    uint32_t totalCells;
    uint32_t totalBins;           // sum of (binCapacity + 1)
    uint32_t modes;               // OR of the scenes' SPH_MODE_* flags (picks kernel variants)
    const uint16_t* tileScene = nullptr;   // scene of each half-tile workgroup (32 spatial bins)
    uint32_t tiles = 0;                    // half-tile workgroups over all scenes (0: none planned)
    // When set: zeroed by the grid clear, made nonzero by the reorder and the density pass when
    // some particle's state lies outside the ranges of the force pass's short division
    // (coord_in_range and the predicates beside it); the force pass then divides with the
    // IEEE sequences.
    uint32_t* rangeGuard = nullptr;
};

// Particle state inside which every operand of the force pass's short division and square
// root (sph_exact.h) lies in the admitted ranges, for a scene whose shortDivide is set (h at
// most 2^26, mass in [2^-60, 2^15], viscosity * mass 0 or in [2^-40, 2^8]):
//  - a nonzero coordinate has magnitude at least 2^-67, so a nonzero coordinate difference
//    (the numerators x_i - x_j; the pair's distance is below h) has magnitude at least 2^-90;
//  - a nonzero velocity component has magnitude in [2^-33, 2^30], so a nonzero velocity
//    difference lies in [2^-56, 2^31] and times viscosity * mass in [2^-96, 2^39];
//  - pressure is 0 or in [2^-30, 2^30], so -mass (p_i + p_j) is 0 or in [2^-90, 2^46], and
//    its quotient by the density and that quotient's half stay normal;
//  - density (clamped below at 1e-6 > 2^-20) is at most 2^26, as is the distance (below h).
// NaN and infinities fail every test.
__device__ __forceinline__ bool coord_in_range(float c) {
    const float a = fabsf(c);
    return c == 0.0f || (a >= 0x1p-67f && a <= 0x1p60f);
}
__device__ __forceinline__ bool velocity_in_range(float v) {
    const float a = fabsf(v);
    return v == 0.0f || (a >= 0x1p-33f && a <= 0x1p30f);
}
__device__ __forceinline__ bool pressure_in_range(float p) { return p == 0.0f || (p >= 0x1p-30f && p <= 0x1p30f); }
__device__ __forceinline__ bool density_in_range(float rho) { return rho <= 0x1p26f; }

// Raise the guard once per wave that has a lane out of range (bad is rare; the first such
// lane writes).
__device__ __forceinline__ void raise_range_guard(uint32_t* guard, bool bad) {
    if (!guard) return;
    const uint64_t m = __ballot(bad);
    if (m && (threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)m) - 1)) atomicOr(guard, 1u);
}

struct GridBuffers {
    const float4* pos;      // state before the sort (previous layout)
    const float4* vel;
    const uint32_t* ids;
    uint32_t* keys;         // global bucket per pre-sort slot
    uint2* buckets;         // totalCells: (first slot, member count); the clear zeroes both, the
                            // keys pass counts, the reorder sets the first slot of non-empty ones
    uint32_t* rep;          // totalCells: smallest global bin among each bucket's cells
    uint32_t* binStart;     // totalBins + 1: counts, then their exclusive scan
    uint32_t* binOf;        // per pre-sort slot
    uint32_t* slots;        // provisional slot inside the bin (arbitrary order)
    uint2* pairKI;          // (global bucket, id) scattered by bin
    uint32_t* pairSrc;      // pre-sort slot of each scattered pair
    uint32_t* blockSums;
    float4* posSorted;
    float4* velSorted;
    uint32_t* idsSorted;
    uint32_t* keysSorted;
    uint32_t* restCount = nullptr;   // when set: zeroed with the counts (the tiled density pass's list)
    // For the tiled density pass (set when the launch has half-tiles): the list of particles
    // whose own cell lies outside the half-tile they are sorted into.
    uint2* rest = nullptr;           // (slot, scene)
    // When set: each k_bin_slots workgroup's largest bin occupancy (one per workgroup of
    // the launch), and the largest of them in mapped host memory, written by k_rank_reorder.
    uint32_t* binMax = nullptr;
    uint32_t* binMaxHost = nullptr;
};

// Logical block for a hardware workgroup id. Workgroups are dispatched round-robin over
// the XCDs; this gives XCD x a contiguous range of logical blocks (slots), so each L2
// caches one spatially compact slab of the state.
__device__ __forceinline__ uint32_t xcd_block(uint32_t hw, uint32_t blocks) {
    const uint32_t q = blocks / kXcds, r = blocks % kXcds;
    const uint32_t x = hw % kXcds, k = hw / kXcds;
    return x * q + (x < r ? x : r) + k;
}

// The scene, local index and global slot of this thread in a particle-parallel launch.
struct Where {
    const SceneDesc* d;
    uint32_t local;
    uint32_t slot;
    bool valid;
};

__device__ __forceinline__ Where locate(const Launch& L) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const SceneDesc* d = L.scenes + L.blockScene[lb];
    Where w;
    w.d = d;
    w.local = (lb - d->blockOffset) * kBlock + threadIdx.x;
    w.valid = w.local < d->k.n;
    w.slot = d->offset + w.local;
    return w;
}

// Neighbourhood sums the density pass also accumulates for the learned surrogate. For
// every in-range candidate j of particle t (r2 < h2, visiting order, duplicate visits and
// t itself included) with s = (h2 - r2) / h2, e = (x_t - x_j) / h and w = v_j - v_t:
//   sums[0 * total + t] = (sum s^2 e,                 count)
//   sums[1 * total + t] = sum s^3 (e_x e_x, e_y e_y, e_z e_z, e_x e_y)
//   sums[2 * total + t] = (sum s^3 e_x e_z, sum s^3 e_y e_z, sum s^2 (w . e), sum s)
//   sums[3 * total + t] = (sum s w,                   0)
//   sums[4 * total + t] = (sum s^2 (w x e),           0)
constexpr uint32_t kSumRows = 5;
struct NeighbourSums {
    float4* sums;                  // kSumRows x total
    float2* deviation = nullptr;   // when set: per workgroup of the pass, the sums over its particles
                                   // of max(rho / rho0 - 1, 0) and |rho / rho0 - 1| (a tree over the
                                   // workgroup's lanes in LDS), the density check's first stage
};

// launchers (sph_grid.hip); all batched launches use L.blocks workgroups
hipError_t launch_grid_clear(const Launch& L, const GridBuffers& b, hipStream_t s);
hipError_t launch_keys(const Launch& L, const GridBuffers& b, hipStream_t s);
hipError_t launch_bin_slots(const Launch& L, const GridBuffers& b, hipStream_t s);
hipError_t launch_scan(uint32_t* data, uint32_t count, uint32_t total, uint32_t* blockSums, hipStream_t s);
hipError_t launch_scatter(const Launch& L, const GridBuffers& b, hipStream_t s);
hipError_t launch_rank_reorder(const Launch& L, const GridBuffers& b, hipStream_t s);
hipError_t launch_unpermute(const float4* src, const uint32_t* ids, float4* dst, uint32_t n, hipStream_t s);
hipError_t launch_gather_keys(const uint32_t* keysSorted, const uint32_t* ids, uint32_t keyBase, uint32_t* cellOf,
                              uint32_t n, hipStream_t s);
hipError_t launch_emit(const SphParams& p, uint32_t count, float4* pos, float4* vel, float4* prev, uint32_t* ids,
                       hipStream_t s);
hipError_t launch_iota(uint32_t* ids, uint32_t n, hipStream_t s);
hipError_t launch_calibration_copy(const float4* src, float4* dst, size_t n, hipStream_t s);

// The tiled density pass (sph_forces.hip): a workgroup per half-tile of 32 spatial bins
// keeps the bucket tables of the cells around them in LDS; the particles the grid build
// listed in `rest` (bucket members whose own cell lies outside their half-tile) are walked in
// global memory by `restBlocks` further workgroups of the same launch.
struct DensityTiles {
    const uint32_t* binStart;   // the scanned bin starts of the grid build
    const uint2* rest;          // (slot, scene) per listed particle
    const uint32_t* restCount;
    uint32_t restBlocks;
};

// Single-scene launches of at most this many particle slots (capped buckets, the reference's
// arithmetic, neighbour list kept) take the small-scene density and force passes: one wave
// per particle. Measured on dam breaks (M particle-steps/s, small passes against lockstep):
// 16k 51 against 38, 32k 68 against 75, 64k 83 against 145; the crossover lies near 27k.
constexpr uint32_t kSmallDensityMax = 24576;

// launchers (sph_forces.hip)
hipError_t launch_density(const Launch& L, float4* pos, float4* vel, const uint2* buckets, uint4* nbrList,
                          uint32_t* nbrCount,
                          const NeighbourSums* sums, hipStream_t s, const DensityTiles* tiles = nullptr);
hipError_t launch_force(const Launch& L, const float4* pos, const float4* vel, const uint2* buckets,
                        const uint4* nbrList, const uint32_t* nbrCount, float4* posOut,
                        float4* velOut, float4* forceOut, hipStream_t s);

}  // namespace sphdev
