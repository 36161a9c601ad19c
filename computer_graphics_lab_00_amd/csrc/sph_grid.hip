// sph_grid.hip — deterministic spatial-hash grid build for gfx950.
//
// Replaces FluidGrid::build (src/physics/fluid_sph.cpp:160-188) and the GPU's
// shaders/fluid_grid.comp:9-20. The reference inserts particles serially in ascending
// index into hash buckets of 32 slots; its GPU twin uses atomics and so loses that order.
//
// What the neighbour loops need is, for every hash bucket, its members in ascending
// particle index. Where the buckets sit in memory is free, so they are laid out by
// space: every bucket gets a representative spatial bin (the smallest bin of the cells
// it holds; bins are 4x4x4 tiles of cells in Morton order inside a tile), particles are
// counting-sorted by their bucket's bin, and inside a bin ordered by (bucket, index).
// Each bucket therefore stays contiguous and index-ascending, and neighbouring cells
// sit next to each other in HBM, which is what the 27-bucket gathers read.
// All scenes of a batch go through each step in one launch (workgroups find their scene
// in a table); the single solver is a batch of one.
//   1. clear:   buckets = (0, 0), rep = ~0, bin counts = 0
//   2. keys:    key = spatialHash(spatialCell(p, h)); buckets[key].count += 1; rep[key] min= bin
//               (wave-level run aggregation: the state is already spatially sorted, so
//               neighbouring lanes mostly share a bucket); each scene's filler bin = its
//               unused capacity, so one global scan aligns every scene to its slots
//   3. binslot: bin = rep[key]; a provisional slot in the bin by atomic
//   4. scan:    exclusive scan of binCount -> binStart (two passes: tiles scanned in LDS
//               with shuffles, then each tile adds the sum of the tile totals before it)
//   5. scatter: (key, id, source) lands in its bin's range
//   6. rank:    final slot = bin start + rank of (key, id) among the bin's members; the
//               particle's pos / vel / id are copied there, and the first member of
//               each bucket records its slot as buckets[key].first
// The atomics only ever produce counts and provisional slots; every final position is
// a pure function of the keys and indices, so the result is deterministic.
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

__device__ inline uint32_t spread2(uint32_t v) { return (v & 1u) | ((v & 2u) << 2); }

// Spatial bin of a cell: clamp into the bin grid, tiles of 2^tileLog2 bins in x-fastest
// order, 64 bins per tile in Morton order (4x4x4 tiles) or x-fastest order (others).
__device__ inline uint32_t bin_of(const BinGrid& g, int32_t cx, int32_t cy, int32_t cz) {
    const int32_t c[3] = {cx, cy, cz};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int64_t d = (int64_t)c[a] - g.origin[a];
        d = d < 0 ? 0 : d;
        d >>= g.shift;
        q[a] = d >= (int64_t)g.dims[a] ? g.dims[a] - 1u : (uint32_t)d;
    }
    const uint32_t lx = g.tileLog2[0], ly = g.tileLog2[1], lz = g.tileLog2[2];
    const uint32_t tile = ((q[2] >> lz) * g.tiles[1] + (q[1] >> ly)) * g.tiles[0] + (q[0] >> lx);
    const uint32_t ix = q[0] & ((1u << lx) - 1u), iy = q[1] & ((1u << ly) - 1u), iz = q[2] & ((1u << lz) - 1u);
    uint32_t m;
    if (lx == 2u && ly == 2u && lz == 2u) {
        m = spread2(ix) | (spread2(iy) << 1) | (spread2(iz) << 2);
    } else {
        m = (iz << (lx + ly)) | (iy << lx) | ix;
    }
    return tile * 64u + m;
}

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

// Bit mask of lanes at or below this lane.
__device__ inline uint64_t lanes_upto(uint32_t lane) { return lane == 63u ? ~0ull : ((2ull << lane) - 1ull); }

// buckets = (0, 0) and rep = ~0 over every scene's buckets, bin counts = 0 over every bin;
// the tiled pass's list and the range guard emptied.
__global__ void __launch_bounds__(kBlock)
    k_grid_clear(uint2* __restrict__ buckets, uint32_t* __restrict__ rep, uint32_t cells,
                 uint32_t* __restrict__ binCount, uint32_t bins, uint32_t* __restrict__ restCount,
                 uint32_t* __restrict__ rangeGuard) {
    const uint32_t stride = gridDim.x * kBlock;
    if (restCount && blockIdx.x == 0 && threadIdx.x == 0) *restCount = 0u;
    if (rangeGuard && blockIdx.x == 0 && threadIdx.x == 0) *rangeGuard = 0u;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cells; i += stride) {
        buckets[i] = make_uint2(0u, 0u);
        rep[i] = 0xFFFFFFFFu;
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < bins; i += stride) binCount[i] = 0u;
}

// Each scene's filler bin takes its unused capacity, so one scan over all bins puts every
// scene's slots at its capacity-aligned offset.
__global__ void __launch_bounds__(kBlock)
    k_grid_fillers(Launch L, uint32_t* __restrict__ binCount) {
    for (uint32_t s = threadIdx.x; s < L.numScenes; s += kBlock) {
        const SceneDesc& d = L.scenes[s];
        binCount[d.binOffset + d.binCapacity] = d.capacity - d.k.n;
    }
}

__global__ void __launch_bounds__(kBlock) k_iota(uint32_t* __restrict__ ids, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) ids[i] = i;
}

// Also sets each scene's filler bin (workgroup 0; the bin counts were cleared before and
// nothing else touches them until k_bin_slots).
__global__ void __launch_bounds__(kBlock)
    k_keys(Launch L, const float4* __restrict__ pos, uint32_t* __restrict__ keys, uint2* __restrict__ buckets,
           uint32_t* __restrict__ rep, uint32_t* __restrict__ binCount) {
    if (blockIdx.x == 0) {
        for (uint32_t s = threadIdx.x; s < L.numScenes; s += kBlock) {
            const SceneDesc& d = L.scenes[s];
            binCount[d.binOffset + d.binCapacity] = d.capacity - d.k.n;
        }
    }
    const Where w = locate(L);
    const StepConsts k = w.d->k;   // by value: the constants stay in scalar registers
    const uint32_t lane = lane_id();
    uint32_t key = 0xFFFFFFFFu, bin = 0xFFFFFFFFu;
    if (w.valid) {
        const float4 p = pos[w.slot];
        const int32_t cx = cell_coord(p.x, k.h), cy = cell_coord(p.y, k.h), cz = cell_coord(p.z, k.h);
        key = w.d->cellOffset + hash_cell(cx, cy, cz, k.cellMask);
        bin = w.d->binOffset + bin_of(w.d->g, cx, cy, cz);
        keys[w.slot] = key;
    }
    // Runs of equal keys across consecutive lanes add their length with one atomic.
    const uint32_t prevKey = __shfl_up(key, 1, 64);
    const uint32_t prevBin = __shfl_up(bin, 1, 64);
    const bool head = w.valid && (lane == 0u || prevKey != key);
    const uint64_t heads = __ballot(head);
    const uint64_t live = __ballot(w.valid);
    if (head) {
        const uint64_t above = heads & ~lanes_upto(lane);
        const uint32_t end = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : (uint32_t)__popcll(live);
        atomicAdd(&buckets[key].y, end - lane);
    }
    if (w.valid && (head || prevBin != bin)) atomicMin(&rep[key], bin);
}

// Maximum over the workgroup, returned to every lane (every lane of the workgroup must call it).
__device__ inline uint32_t block_max(uint32_t m) {
    __shared__ uint32_t waveMax[kBlock / 64];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63u) == 0u) waveMax[threadIdx.x >> 6] = m;
    __syncthreads();
    m = waveMax[0];
#pragma unroll
    for (uint32_t w = 1; w < kBlock / 64; ++w) m = waveMax[w] > m ? waveMax[w] : m;
    return m;
}

__global__ void __launch_bounds__(kBlock)
    k_bin_slots(Launch L, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rep,
                uint32_t* __restrict__ binCount, uint32_t* __restrict__ binOf, uint32_t* __restrict__ slots,
                uint32_t* __restrict__ binMax) {
    const Where w = locate(L);
    const uint32_t lane = lane_id();
    const uint32_t bin = w.valid ? rep[keys[w.slot]] : 0xFFFFFFFFu;
    const uint32_t prevBin = __shfl_up(bin, 1, 64);
    const bool head = w.valid && (lane == 0u || prevBin != bin);
    const uint64_t heads = __ballot(head);
    const uint64_t live = __ballot(w.valid);
    uint32_t base = 0;
    if (head) {
        const uint64_t above = heads & ~lanes_upto(lane);
        const uint32_t end = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : (uint32_t)__popcll(live);
        base = atomicAdd(&binCount[bin], end - lane);
    }
    // Every lane takes its run head's base plus its distance from the head.
    const uint64_t mine = heads & lanes_upto(lane);
    const uint32_t headLane = mine ? 63u - (uint32_t)__clzll((unsigned long long)mine) : lane;
    base = __shfl(base, headLane, 64);
    const uint32_t slot = base + (lane - headLane);
    if (w.valid) {
        binOf[w.slot] = bin;
        slots[w.slot] = slot;
    }
    if (binMax) {
        // The largest bin's last member gets slot (occupancy - 1): the maximum of slot + 1
        // over all particles is the largest bin occupancy. Each workgroup stores its own
        // maximum (no atomics: one contended address is serialised across the XCDs);
        // k_rank_reorder reduces them.
        uint32_t m = w.valid ? slot + 1u : 0u;
        m = block_max(m);
        if (threadIdx.x == 0) binMax[blockIdx.x] = m;
    }
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

// Pass 2: each tile adds the sum of the tile totals before it (every workgroup reduces that
// prefix of the totals itself: a few thousand L2-resident words); the element after the last
// one gets the total.
__global__ void __launch_bounds__(kBlock)
    k_scan_add(uint32_t* __restrict__ data, uint32_t count, uint32_t total, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t waveSums[kBlock / 64];
    uint32_t local = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kBlock) local += sums[i];
#pragma unroll
    for (uint32_t d = 32; d > 0; d >>= 1) local += __shfl_down(local, d, 64);
    if ((threadIdx.x & 63u) == 0u) waveSums[threadIdx.x >> 6] = local;
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) add += waveSums[w];
    const uint32_t base = blockIdx.x * kScanTile;
    for (uint32_t e = threadIdx.x; e < kScanTile; e += kBlock) {
        uint32_t idx = base + e;
        if (idx < count) data[idx] += add;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) data[count] = total;
}

__global__ void __launch_bounds__(kBlock)
    k_scatter(Launch L, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ binOf,
              const uint32_t* __restrict__ slots, const uint32_t* __restrict__ ids,
              const uint32_t* __restrict__ binStart, uint2* __restrict__ pairKI, uint32_t* __restrict__ pairSrc) {
    const Where w = locate(L);
    if (!w.valid) return;
    const uint32_t dst = binStart[binOf[w.slot]] + slots[w.slot];
    pairKI[dst] = make_uint2(keys[w.slot], ids[w.slot]);
    pairSrc[dst] = w.slot;
}

// Final slot = bin start + number of bin members before this one in (bucket, id) order.
// Thread `slot` handles the pair scattered to that slot; a scene's n particles fill
// exactly its first n slots.
// With `rest` (a launch with half-tiles for the tiled density pass), every particle whose own
// cell lies outside the half-tile (32 bins) it is sorted into — a member of a bucket shared
// with another cell — is listed there.
__global__ void __launch_bounds__(kBlock)
    k_rank_reorder(Launch L, const uint2* __restrict__ pairKI, const uint32_t* __restrict__ pairSrc,
                   const uint32_t* __restrict__ binOf, const uint32_t* __restrict__ binStart,
                   const float4* __restrict__ pos, const float4* __restrict__ vel, float4* __restrict__ posSorted,
                   float4* __restrict__ velSorted, uint32_t* __restrict__ idsSorted,
                   uint32_t* __restrict__ keysSorted, uint2* __restrict__ buckets, uint2* __restrict__ rest,
                   uint32_t* __restrict__ restCount, const uint32_t* __restrict__ binMax,
                   uint32_t* __restrict__ binMaxHost) {
    // Workgroup 0 reduces k_bin_slots's per-workgroup maxima (final: that launch has
    // completed) into the largest bin occupancy of this build, for the host's mapped copy.
    if (binMaxHost && blockIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) m = binMax[b] > m ? binMax[b] : m;
        m = block_max(m);
        if (threadIdx.x == 0) *binMaxHost = m;
    }
    const Where w = locate(L);
    const uint2 me = w.valid ? pairKI[w.slot] : make_uint2(0u, 0u);
    const uint32_t src = w.valid ? pairSrc[w.slot] : 0u;
    const uint32_t bin = w.valid ? binOf[src] : 0u;
    uint32_t lo = 0, hi = 0;
    if (w.valid) {
        lo = binStart[bin];
        hi = binStart[bin + 1];
    }
    uint32_t rank = 0;
    bool first = true;
    for (uint32_t u = lo; u < hi; ++u) {
        const uint2 o = pairKI[u];
        const bool sameKey = o.x == me.x;
        rank += (o.x < me.x || (sameKey && o.y < me.y)) ? 1u : 0u;
        if (sameKey && o.y < me.y) first = false;
    }
    const uint32_t dst = lo + rank;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool outOfRange = false;
    if (w.valid) {
        p = pos[src];
        const float4 v = vel[src];
        posSorted[dst] = p;
        velSorted[dst] = v;
        idsSorted[dst] = me.y;
        keysSorted[dst] = me.x;
        if (first) buckets[me.x].x = dst;
        outOfRange = !(coord_in_range(p.x) && coord_in_range(p.y) && coord_in_range(p.z) &&
                       velocity_in_range(v.x) && velocity_in_range(v.y) && velocity_in_range(v.z));
    }
    raise_range_guard(L.rangeGuard, outOfRange);
    if (!rest) return;   // the same for every lane
    // Is the particle's own cell inside the half-tile its bin belongs to? (bins are cells when
    // shift == 0; the tiled pass is planned only then)
    bool out = false;
    if (w.valid) {
        const BinGrid& g = w.d->g;
        const float h = w.d->k.h;
        const int32_t c[3] = {cell_coord(p.x, h), cell_coord(p.y, h), cell_coord(p.z, h)};
        bool inGrid = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int64_t q = (int64_t)c[a] - g.origin[a];
            inGrid = inGrid && q >= 0 && q < (int64_t)g.dims[a];
        }
        out = !inGrid ||
              (bin_of(g, c[0], c[1], c[2]) / kHalfTileBins) != ((bin - w.d->binOffset) / kHalfTileBins);
    }
    const uint64_t outs = __ballot(out);
    if (outs) {
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)outs) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(restCount, (uint32_t)__popcll(outs));
        base = __shfl(base, (int)leader, 64);
        if (out) rest[base + (uint32_t)__popcll(outs & ((1ull << lane) - 1ull))] = make_uint2(dst, w.d - L.scenes);
    }
}

__global__ void __launch_bounds__(kBlock)
    k_unpermute(const float4* __restrict__ src, const uint32_t* __restrict__ ids, float4* __restrict__ dst, uint32_t n) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) dst[ids[t]] = src[t];
}

__global__ void __launch_bounds__(kBlock)
    k_gather_keys(const uint32_t* __restrict__ keysSorted, const uint32_t* __restrict__ ids, uint32_t keyBase,
                  uint32_t* __restrict__ cellOf, uint32_t n) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) cellOf[ids[t]] = keysSorted[t] - keyBase;
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

__global__ void __launch_bounds__(kBlock)
    k_calibration_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    size_t stride = (size_t)gridDim.x * kBlock;
    for (; i < n; i += stride) dst[i] = src[i];
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_iota(uint32_t* ids, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_iota<<<blocks_for(n), kBlock, 0, s>>>(ids, n);
    return hipGetLastError();
}

hipError_t launch_grid_clear(const Launch& L, const GridBuffers& b, hipStream_t s) {
    const uint32_t span = L.totalCells > L.totalBins ? L.totalCells : L.totalBins;
    uint32_t blocks = blocks_for(span);
    if (blocks > 4096u) blocks = 4096u;
    k_grid_clear<<<blocks, kBlock, 0, s>>>(b.buckets, b.rep, L.totalCells, b.binStart, L.totalBins, b.restCount,
                                           L.rangeGuard);
    return hipGetLastError();
}

hipError_t launch_keys(const Launch& L, const GridBuffers& b, hipStream_t s) {
    if (L.blocks == 0) {   // no particle anywhere: only the filler bins
        k_grid_fillers<<<1, kBlock, 0, s>>>(L, b.binStart);
        return hipGetLastError();
    }
    k_keys<<<L.blocks, kBlock, 0, s>>>(L, b.pos, b.keys, b.buckets, b.rep, b.binStart);
    return hipGetLastError();
}

hipError_t launch_bin_slots(const Launch& L, const GridBuffers& b, hipStream_t s) {
    k_bin_slots<<<L.blocks, kBlock, 0, s>>>(L, b.keys, b.rep, b.binStart, b.binOf, b.slots, b.binMax);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* data, uint32_t count, uint32_t total, uint32_t* blockSums, hipStream_t s) {
    uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    k_scan_tiles<<<tiles, kBlock, 0, s>>>(data, count, blockSums);
    k_scan_add<<<tiles, kBlock, 0, s>>>(data, count, total, blockSums);
    return hipGetLastError();
}

hipError_t launch_scatter(const Launch& L, const GridBuffers& b, hipStream_t s) {
    k_scatter<<<L.blocks, kBlock, 0, s>>>(L, b.keys, b.binOf, b.slots, b.ids, b.binStart, b.pairKI, b.pairSrc);
    return hipGetLastError();
}

hipError_t launch_rank_reorder(const Launch& L, const GridBuffers& b, hipStream_t s) {
    const bool tiles = L.tiles > 0 && b.rest && b.restCount;
    k_rank_reorder<<<L.blocks, kBlock, 0, s>>>(L, b.pairKI, b.pairSrc, b.binOf, b.binStart, b.pos, b.vel,
                                              b.posSorted, b.velSorted, b.idsSorted, b.keysSorted, b.buckets,
                                              tiles ? b.rest : nullptr, b.restCount, b.binMax,
                                              b.binMax ? b.binMaxHost : nullptr);
    return hipGetLastError();
}

hipError_t launch_unpermute(const float4* src, const uint32_t* ids, float4* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_unpermute<<<blocks_for(n), kBlock, 0, s>>>(src, ids, dst, n);
    return hipGetLastError();
}

hipError_t launch_gather_keys(const uint32_t* keysSorted, const uint32_t* ids, uint32_t keyBase, uint32_t* cellOf,
                              uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_gather_keys<<<blocks_for(n), kBlock, 0, s>>>(keysSorted, ids, keyBase, cellOf, n);
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

hipError_t launch_calibration_copy(const float4* src, float4* dst, size_t n, hipStream_t s) {
    k_calibration_copy<<<8192, kBlock, 0, s>>>(src, dst, n);
    return hipGetLastError();
}

}  // namespace sphdev
