// sph_grid.hip — deterministic spatial-hash grid build for gfx950.
//
// Replaces FluidGrid::build (src/physics/fluid_sph.cpp:160-188) and the GPU's
// shaders/fluid_grid.comp:9-20. The reference inserts particles serially in ascending
// index into 32-slot buckets; its GPU twin uses atomics and so loses that order. Here:
//   1. keys:   key = spatialHash(spatialCell(p, h)) per particle; a device atomic
//              hands out a provisional slot and counts the bucket (order-free integers)
//   2. scan:   exclusive scan of the counts -> bucket starts (three passes, LDS + wave
//              shuffles)
//   3. scatter:(particle id, source slot) pairs land in their bucket's range
//   4. rank:   each pair's final slot is its rank by particle id inside its bucket, so
//              every bucket ends up in ascending index order whatever the atomics did;
//              the particle's pos / vel / id are copied to that slot (the reorder)
// The result is the reference's bucket order bit for bit, with no 32-slot cap on the
// storage (the cap is applied by the neighbour loops in reference mode).
#include "sph_device.h"

namespace sphdev {

namespace {

__device__ inline uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    // spatialHash (src/physics/spatial_hash.h:16-20): uint32 wraparound products.
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & mask;
}

__device__ inline int32_t cell_coord(float x, float h) {
    // spatialCell (src/physics/spatial_hash.h:12-14): IEEE division, floor, to int.
    return (int32_t)floorf(x / h);
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, uint32_t count) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t stride = gridDim.x * kBlock;
    for (; i < count; i += stride) dst[i] = value;
}

__global__ void __launch_bounds__(kBlock) k_iota(uint32_t* __restrict__ ids, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) ids[i] = i;
}

__global__ void __launch_bounds__(kBlock)
    k_keys(StepConsts k, const float4* __restrict__ pos, uint32_t* __restrict__ keys, uint32_t* __restrict__ slots,
           uint32_t* __restrict__ counts) {
    uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= k.n) return;
    float4 p = pos[s];
    uint32_t key = hash_cell(cell_coord(p.x, k.h), cell_coord(p.y, k.h), cell_coord(p.z, k.h), k.cellMask);
    keys[s] = key;
    slots[s] = atomicAdd(&counts[key], 1u);
}

// Inclusive scan across one wave64 with shuffles.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t up = __shfl_up(v, d, 64);
        if (lane >= d) v += up;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total in *total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t waveSums[kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    uint32_t waveOffset = 0, sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        uint32_t ws = waveSums[w];
        if (w < wave) waveOffset += ws;
        sum += ws;
    }
    __syncthreads();
    *total = sum;
    return waveOffset + inc - v;
}

// Pass 1: each block scans kScanTile elements in place (exclusive) and records its total.
__global__ void __launch_bounds__(kBlock)
    k_scan_tiles(uint32_t* __restrict__ data, uint32_t count, uint32_t* __restrict__ blockSums) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t e = 0; e < kScanPerThread; ++e) {
        uint32_t idx = base + e;
        v[e] = idx < count ? data[idx] : 0u;
        local += v[e];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(local, &total);
#pragma unroll
    for (uint32_t e = 0; e < kScanPerThread; ++e) {
        uint32_t idx = base + e;
        if (idx < count) data[idx] = run;
        run += v[e];
    }
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

// Pass 2: one block scans the tile totals (exclusive), serially per thread chunk.
__global__ void __launch_bounds__(1024) k_scan_totals(uint32_t* __restrict__ sums, uint32_t tiles) {
    __shared__ uint32_t waveSums[16];
    const uint32_t per = (tiles + 1023u) / 1024u;
    const uint32_t begin = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t e = 0; e < per; ++e) {
        uint32_t idx = begin + e;
        if (idx < tiles) local += sums[idx];
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(local);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w = 0; w < wave; ++w) off += waveSums[w];
    uint32_t run = off + inc - local;
    for (uint32_t e = 0; e < per; ++e) {
        uint32_t idx = begin + e;
        if (idx < tiles) {
            uint32_t x = sums[idx];
            sums[idx] = run;
            run += x;
        }
    }
}

// Pass 3: add each tile's offset; the element after the last bucket gets the total.
__global__ void __launch_bounds__(kBlock)
    k_scan_add(uint32_t* __restrict__ data, uint32_t count, uint32_t total, const uint32_t* __restrict__ sums) {
    const uint32_t add = sums[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanTile;
    for (uint32_t e = threadIdx.x; e < kScanTile; e += kBlock) {
        uint32_t idx = base + e;
        if (idx < count) data[idx] += add;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) data[count] = total;
}

__global__ void __launch_bounds__(kBlock)
    k_scatter(uint32_t n, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ slots,
              const uint32_t* __restrict__ ids, const uint32_t* __restrict__ starts, uint2* __restrict__ pairs) {
    uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    uint32_t dst = starts[keys[s]] + slots[s];
    pairs[dst] = make_uint2(ids[s], s);
}

// Final slot = bucket start + number of bucket members with a smaller particle id.
__global__ void __launch_bounds__(kBlock)
    k_rank_reorder(uint32_t n, const uint2* __restrict__ pairs, const uint32_t* __restrict__ keys,
                   const uint32_t* __restrict__ starts, const float4* __restrict__ pos,
                   const float4* __restrict__ vel, float4* __restrict__ posSorted, float4* __restrict__ velSorted,
                   uint32_t* __restrict__ idsSorted, uint32_t* __restrict__ keysSorted) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    uint2 pr = pairs[t];
    uint32_t key = keys[pr.y];
    uint32_t lo = starts[key], hi = starts[key + 1];
    uint32_t rank = 0;
    for (uint32_t u = lo; u < hi; ++u) rank += pairs[u].x < pr.x ? 1u : 0u;
    uint32_t dst = lo + rank;
    posSorted[dst] = pos[pr.y];
    velSorted[dst] = vel[pr.y];
    idsSorted[dst] = pr.x;
    keysSorted[dst] = key;
}

__global__ void __launch_bounds__(kBlock)
    k_unpermute(const float4* __restrict__ src, const uint32_t* __restrict__ ids, float4* __restrict__ dst, uint32_t n) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) dst[ids[t]] = src[t];
}

__global__ void __launch_bounds__(kBlock)
    k_gather_keys(const uint32_t* __restrict__ keysSorted, const uint32_t* __restrict__ ids,
                  uint32_t* __restrict__ cellOf, uint32_t n) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) cellOf[ids[t]] = keysSorted[t];
}

struct EmitConsts {
    float world[16];
    float spacing;
    float extent[3];
    float restDensity;
    uint32_t cx, cy, count;
};

// FluidSolver::emit (src/physics/fluid_sph.cpp:138-154) / shaders/fluid_emit.comp:8-23.
__global__ void __launch_bounds__(kBlock)
    k_emit(EmitConsts e, float4* __restrict__ pos, float4* __restrict__ vel, float4* __restrict__ prev,
           uint32_t* __restrict__ ids) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= e.count) return;
    uint32_t ix = i % e.cx, iy = (i / e.cx) % e.cy, iz = i / (e.cx * e.cy);
    float l[4] = {(float)ix * e.spacing - e.extent[0], (float)iy * e.spacing - e.extent[1],
                  (float)iz * e.spacing - e.extent[2], 1.0f};
    float w[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        // glm mat4 * vec4: (m0 x + m1 y) + (m2 z + m3 w)
        float a0 = e.world[0 + r] * l[0] + e.world[4 + r] * l[1];
        float a1 = e.world[8 + r] * l[2] + e.world[12 + r] * l[3];
        w[r] = a0 + a1;
    }
    pos[i] = make_float4(w[0], w[1], w[2], e.restDensity);
    vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    prev[i] = make_float4(w[0], w[1], w[2], 0.0f);
    ids[i] = i;
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    uint32_t blocks = blocks_for(count);
    if (blocks > 8192u) blocks = 8192u;
    k_fill_u32<<<blocks, kBlock, 0, s>>>(dst, value, count);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t* ids, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_iota<<<blocks_for(n), kBlock, 0, s>>>(ids, n);
    return hipGetLastError();
}

hipError_t launch_keys(const StepConsts& k, const GridBuffers& g, hipStream_t s) {
    k_keys<<<blocks_for(k.n), kBlock, 0, s>>>(k, g.pos, g.keys, g.slots, g.starts);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* data, uint32_t count, uint32_t total, uint32_t* blockSums, hipStream_t s) {
    uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    k_scan_tiles<<<tiles, kBlock, 0, s>>>(data, count, blockSums);
    k_scan_totals<<<1, 1024, 0, s>>>(blockSums, tiles);
    k_scan_add<<<tiles, kBlock, 0, s>>>(data, count, total, blockSums);
    return hipGetLastError();
}

hipError_t launch_scatter(const StepConsts& k, const GridBuffers& g, hipStream_t s) {
    k_scatter<<<blocks_for(k.n), kBlock, 0, s>>>(k.n, g.keys, g.slots, g.ids, g.starts, g.pairs);
    return hipGetLastError();
}

hipError_t launch_rank_reorder(const StepConsts& k, const GridBuffers& g, hipStream_t s) {
    k_rank_reorder<<<blocks_for(k.n), kBlock, 0, s>>>(k.n, g.pairs, g.keys, g.starts, g.pos, g.vel, g.posSorted,
                                                      g.velSorted, g.idsSorted, g.keysSorted);
    return hipGetLastError();
}

hipError_t launch_unpermute(const float4* src, const uint32_t* ids, float4* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_unpermute<<<blocks_for(n), kBlock, 0, s>>>(src, ids, dst, n);
    return hipGetLastError();
}

hipError_t launch_gather_keys(const uint32_t* keysSorted, const uint32_t* ids, uint32_t* cellOf, uint32_t n,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_gather_keys<<<blocks_for(n), kBlock, 0, s>>>(keysSorted, ids, cellOf, n);
    return hipGetLastError();
}

hipError_t launch_emit(const SphParams& p, uint32_t count, float4* pos, float4* vel, float4* prev, uint32_t* ids,
                       hipStream_t s) {
    if (count == 0) return hipSuccess;
    EmitConsts e;
    for (int i = 0; i < 16; ++i) e.world[i] = p.emitterWorld[i];
    e.spacing = p.spacing;
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = p.lattice[a] > 1u ? p.lattice[a] : 1u;
    // extent = vec3(counts - 1) * spacing * 0.5 (src/physics/fluid_sph.cpp:146)
    for (int a = 0; a < 3; ++a) e.extent[a] = ((float)(c[a] - 1u) * p.spacing) * 0.5f;
    e.restDensity = p.restDensity;
    e.cx = c[0];
    e.cy = c[1];
    e.count = count;
    k_emit<<<blocks_for(count), kBlock, 0, s>>>(e, pos, vel, prev, ids);
    return hipGetLastError();
}

}  // namespace sphdev
