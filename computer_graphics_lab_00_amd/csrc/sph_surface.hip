// sph_surface.hip — the fluid's surface on gfx950: the scalar field sampled from the
// particles on the hash grid (physics::buildFluidField / sampleField,
// src/physics/marching_cubes.cpp:312-391; shaders/fluid_field.comp) and its iso-surface
// by marching cubes (physics::marchFluidField, marching_cubes.cpp:393-487;
// shaders/fluid_marching.comp).
//
// The field re-sorts the particles' current positions with the solver's own grid build,
// so each bucket is one contiguous run in ascending particle index (the reference's
// FluidGrid order, 32 per bucket), and one lane per sample point walks the 27 buckets
// around it with the reference's de-duplication, in its order and arithmetic: the field
// is the reference's bit for bit. Marching takes two passes over the cells: one counts each
// workgroup's vertices, a scan turns the counts into offsets, and the second hands each
// non-empty workgroup's vertices to its lanes one vertex per lane (a surface cell has up to
// 15 vertices, most cells none), so the vertices come out in the reference's serial order
// (cells x fastest, each cell's in table order) and the buffer is the serial loop's.
#include <math.h>
#include <string.h>

#include <vector>

#include "sph_engine.h"
#include "sph_internal.h"

using sphhost::fail;
using sphhost::hip_fail;

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kMaxResolution = 1024;

__constant__ int8_t c_table[256 * 16];
__constant__ uint8_t c_nverts[256];   // vertices of each cube case (its table row up to the first -1)

__device__ __constant__ const int8_t c_corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                                         {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__device__ __constant__ const uint8_t c_edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct FieldArgs {
    float origin[3];
    float cell[3];
    float radius;
    float radiusSq;
    uint32_t samples;     // resolution + 1 per axis
    uint32_t resolution;
    uint32_t cellMask;
};

__device__ inline uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & mask;
}

__device__ inline int32_t field_cell(float origin, float cell, uint32_t i, float radius) {
    return (int32_t)floorf((origin + (float)i * cell) / radius);
}

// sampleField (marching_cubes.cpp:312-343) at one sample point, reading the particles from
// the sorted global arrays: the 27 buckets around the point's cell in the reference's order
// (z, y, x), a bucket two neighbour cells fold into summed once (:319-326), at most 32
// particles per bucket.
__device__ float sample_global(const float4* __restrict__ pos, const uint2* __restrict__ buckets,
                               const FieldArgs& a, float px, float py, float pz,
                               int32_t bx, int32_t by, int32_t bz) {
    uint32_t bk[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) bk[o] = hash_cell(bx + o % 3 - 1, by + (o / 3) % 3 - 1, bz + o / 9 - 1, a.cellMask);
    float total = 0.0f;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        bool dup = false;
#pragma unroll
        for (int q = 0; q < o; ++q) dup |= bk[q] == bk[o];
        if (dup) continue;
        const uint2 bd = buckets[bk[o]];
        uint32_t cnt = bd.y;
        if (cnt == 0) continue;
        if (cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
        const uint32_t lo = bd.x;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float4 q = pos[lo + k];
            const float ex = q.x - px, ey = q.y - py, ez = q.z - pz;
            const float dsq = (ex * ex + ey * ey) + ez * ez;
            if (dsq >= a.radiusSq) continue;
            const float d = 1.0f - dsq / a.radiusSq;
            total += (d * d) * d;
        }
    }
    return total;
}

// The field over tiles of kTileX x kTileY x kTileZ samples, one workgroup each: the tile's
// samples need the cells of one box (their cells, one more on every side); the workgroup
// copies the particles of every cell of the box (the bucket's first 32, in bucket order)
// into LDS once, and each sample then walks its 27 cells there — same buckets, same order,
// same arithmetic as sample_global, so the same bits — instead of re-reading every particle
// from L2 for each of the ~200 samples whose neighbourhood covers it. A tile whose box has
// more than kBoxCells cells or kLdsParticles particles takes the global path. The samples on
// the grid's faces stay 0 so the surface closes there (:380-383).
constexpr uint32_t kTileX = 8, kTileY = 8, kTileZ = 4;   // 256 samples
constexpr uint32_t kBoxCells = 256;
constexpr uint32_t kLdsParticles = 1024;
constexpr uint32_t kBatch = 4;   // candidates read from LDS together
static_assert(kTileX * kTileY * kTileZ == kThreads, "one sample per thread");
static_assert(kBoxCells <= 4 * kThreads, "each thread scans at most four cells");

__global__ void __launch_bounds__(kThreads)
    k_field(const float4* __restrict__ pos, const uint2* __restrict__ buckets, FieldArgs a, uint32_t tilesX,
            uint32_t tilesY,
            float* __restrict__ field) {
    __shared__ float4 sp[kLdsParticles + kBatch];   // a batch may read past the last cell
    __shared__ uint32_t cellOff[kBoxCells + 1];
    __shared__ uint32_t cellBk[kBoxCells];
    __shared__ uint32_t scan[kThreads];
    __shared__ int32_t box[6];   // lowest cell, extent
    __shared__ uint32_t useLds;
    const uint32_t tid = threadIdx.x;
    const uint32_t s = a.samples;
    const uint32_t tx = blockIdx.x % tilesX, ty = (blockIdx.x / tilesX) % tilesY, tz = blockIdx.x / (tilesX * tilesY);
    const uint32_t x = tx * kTileX + tid % kTileX, y = ty * kTileY + (tid / kTileX) % kTileY,
                   z = tz * kTileZ + tid / (kTileX * kTileY);
    if (tid == 0) {
        const uint32_t lo[3] = {tx * kTileX, ty * kTileY, tz * kTileZ};
        const uint32_t span[3] = {kTileX, kTileY, kTileZ};
        uint64_t cells = 1u;
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            const uint32_t hi = min(lo[k] + span[k] - 1u, s - 1u);
            // with a positive cell size floor(p / h) does not decrease along an axis: the
            // tile's first and last samples bound it
            const int32_t b0 = field_cell(a.origin[k], a.cell[k], lo[k], a.radius) - 1;
            const int32_t b1 = field_cell(a.origin[k], a.cell[k], hi, a.radius) + 1;
            const int64_t ext = (int64_t)b1 - (int64_t)b0 + 1;
            if (!(a.cell[k] > 0.0f) || ext <= 0 || ext > (int64_t)kBoxCells) ok = false;
            box[k] = b0;
            box[3 + k] = ok ? (int32_t)ext : 0;
            if (ok) cells *= (uint64_t)ext;
        }
        useLds = (ok && cells <= kBoxCells) ? (uint32_t)cells : 0u;
    }
    __syncthreads();
    const uint32_t ncell = useLds;
    if (ncell) {
        // per cell: its bucket and how many of its particles are visited; offsets by a scan
        const uint32_t per = (ncell + kThreads - 1u) / kThreads;   // <= 4
        uint32_t local = 0u;
        for (uint32_t j = 0; j < per; ++j) {
            const uint32_t c = tid * per + j;
            if (c < ncell) {
                const int32_t cx = box[0] + (int32_t)(c % (uint32_t)box[3]);
                const int32_t cy = box[1] + (int32_t)((c / (uint32_t)box[3]) % (uint32_t)box[4]);
                const int32_t cz = box[2] + (int32_t)(c / ((uint32_t)box[3] * (uint32_t)box[4]));
                const uint32_t bk = hash_cell(cx, cy, cz, a.cellMask);
                uint32_t cnt = buckets[bk].y;
                if (cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
                cellBk[c] = bk;
                cellOff[c] = cnt;
                local += cnt;
            }
        }
        scan[tid] = local;
        __syncthreads();
        for (uint32_t d = 1; d < kThreads; d <<= 1) {
            const uint32_t v = tid >= d ? scan[tid - d] : 0u;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        uint32_t run = scan[tid] - local;   // exclusive prefix of this thread's cells
        for (uint32_t j = 0; j < per; ++j) {
            const uint32_t c = tid * per + j;
            if (c < ncell) {
                const uint32_t cnt = cellOff[c];
                cellOff[c] = run;
                run += cnt;
            }
        }
        if (tid == kThreads - 1u) cellOff[ncell] = scan[kThreads - 1u];
        __syncthreads();
        if (cellOff[ncell] > kLdsParticles) {
            if (tid == 0) useLds = 0u;
        } else {
            // one wave per cell at a time: lane k copies the cell's k-th particle
            const uint32_t lane = tid & 63u, wave = tid >> 6;
            for (uint32_t c = wave; c < ncell; c += kThreads / 64u) {
                const uint32_t n = cellOff[c + 1] - cellOff[c];
                if (lane < n) sp[cellOff[c] + lane] = pos[buckets[cellBk[c]].x + lane];
            }
        }
        __syncthreads();
    }
    if (x >= s || y >= s || z >= s) return;
    const uint32_t i = (z * s + y) * s + x;
    if (x == 0 || y == 0 || z == 0 || x == a.resolution || y == a.resolution || z == a.resolution) {
        field[i] = 0.0f;
        return;
    }
    const float px = a.origin[0] + (float)x * a.cell[0];
    const float py = a.origin[1] + (float)y * a.cell[1];
    const float pz = a.origin[2] + (float)z * a.cell[2];
    const int32_t bx = (int32_t)floorf(px / a.radius), by = (int32_t)floorf(py / a.radius),
                  bz = (int32_t)floorf(pz / a.radius);
    if (!useLds) {
        field[i] = sample_global(pos, buckets, a, px, py, pz, bx, by, bz);
        return;
    }
    // the 27 cells as box indices; the box holds every one of them
    uint32_t cidx[27], bk[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        const uint32_t cx = (uint32_t)(bx + o % 3 - 1 - box[0]), cy = (uint32_t)(by + (o / 3) % 3 - 1 - box[1]),
                       cz = (uint32_t)(bz + o / 9 - 1 - box[2]);
        cidx[o] = cx + (uint32_t)box[3] * (cy + (uint32_t)box[4] * cz);
        bk[o] = cellBk[cidx[o]];
    }
    float total = 0.0f;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        bool dup = false;
#pragma unroll
        for (int q = 0; q < o; ++q) dup |= bk[q] == bk[o];
        if (dup) continue;
        const uint32_t lo = cellOff[cidx[o]], hi = cellOff[cidx[o] + 1];
        for (uint32_t k = lo; k < hi; k += kBatch) {
            float4 q[kBatch];
#pragma unroll
            for (uint32_t e = 0; e < kBatch; ++e) q[e] = sp[k + e];
#pragma unroll
            for (uint32_t e = 0; e < kBatch; ++e) {
                const float ex = q[e].x - px, ey = q[e].y - py, ez = q[e].z - pz;
                const float dsq = (ex * ex + ey * ey) + ez * ez;
                if (k + e >= hi || dsq >= a.radiusSq) continue;
                const float d = 1.0f - dsq / a.radiusSq;
                total += (d * d) * d;
            }
        }
    }
    field[i] = total;
}

struct MarchArgs {
    float origin[3];
    float cellSize[3];
    float iso;
    uint32_t resolution;
    uint32_t capacity;
};

__device__ inline float field_at(const float* __restrict__ f, uint32_t s, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= (int32_t)s || y >= (int32_t)s || z >= (int32_t)s) return 0.0f;
    return f[((size_t)z * s + (size_t)y) * s + (size_t)x];
}

// asset::packUnitVector (src/asset/vertex_pack.h:40-52): octahedral map to two snorm16.
__device__ inline uint32_t pack_unit(float nx, float ny, float nz) {
    const float sum = (fabsf(nx) + fabsf(ny)) + fabsf(nz);
    if (sum <= 0.0f) return 0u;
    nx = nx / sum;
    ny = ny / sum;
    nz = nz / sum;
    float u = nx, v = ny;
    if (nz < 0.0f) {
        const float sx = nx >= 0.0f ? 1.0f : -1.0f, sy = ny >= 0.0f ? 1.0f : -1.0f;
        u = (1.0f - fabsf(ny)) * sx;
        v = (1.0f - fabsf(nx)) * sy;
    }
    auto q = [](float c) {
        c = c < -1.0f ? -1.0f : (1.0f < c ? 1.0f : c);
        return (uint32_t)(int32_t)roundf(c * 32767.0f) & 0xFFFFu;   // lround: halves away from zero
    };
    return q(u) | (q(v) << 16);
}

// Cube case of cell (bx, by, bz) and its corner values (marching_cubes.cpp:400-410).
__device__ inline uint32_t cube_case(const float* __restrict__ f, uint32_t s, int32_t bx, int32_t by, int32_t bz,
                                     float iso, float val[8]) {
    uint32_t code = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        val[c] = field_at(f, s, bx + c_corner[c][0], by + c_corner[c][1], bz + c_corner[c][2]);
        if (val[c] >= iso) code |= 1u << c;
    }
    return code;
}

// Sum of one value over the workgroup (wave shuffles, then the four wave sums).
__device__ inline uint32_t block_sum(uint32_t v, uint32_t* scratch4) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63u) == 0u) scratch4[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint32_t t = scratch4[0] + scratch4[1] + scratch4[2] + scratch4[3];
    __syncthreads();
    return t;
}

// Pass 1: the vertex count of each workgroup of kThreads cells (cells x fastest).
__global__ void __launch_bounds__(kThreads)
    k_march_count(const float* __restrict__ f, MarchArgs m, uint32_t* __restrict__ blockCount) {
    __shared__ uint32_t waves[kThreads / 64];
    const uint32_t r = m.resolution, s = r + 1;
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    uint32_t nv = 0;
    if (i < r * r * r) {
        float val[8];
        nv = c_nverts[cube_case(f, s, (int32_t)(i % r), (int32_t)((i / r) % r), (int32_t)(i / (r * r)), m.iso, val)];
    }
    const uint32_t total = block_sum(nv, waves);
    if (threadIdx.x == 0) blockCount[blockIdx.x] = total;
}

// Pass 2 (marchFluidField, marching_cubes.cpp:411-484): the workgroup's vertices, one per lane,
// at the workgroup's offset from the scan. A cell whose vertices would pass `capacity` writes
// none and records its first slot; the serial loop stops there (:467-472), and every later
// cell starts further on, so the written count is the smallest such slot.
__global__ void __launch_bounds__(kThreads)
    k_march_emit(const float* __restrict__ f, MarchArgs m, const uint32_t* __restrict__ blockStart, uint32_t blocks,
                 SphSurfaceVertex* __restrict__ out, uint32_t* __restrict__ counters) {
    __shared__ uint32_t cellOff[kThreads + 1];
    __shared__ uint8_t cellCode[kThreads];
    __shared__ uint32_t waves[kThreads / 64];
    const uint32_t b = blockIdx.x;
    const uint32_t base = blockStart[b];
    const bool last = b + 1u == blocks;
    if (!last && blockStart[b + 1u] == base) return;   // no vertex in this workgroup (all lanes agree)
    const uint32_t r = m.resolution, s = r + 1;
    const uint32_t i = b * kThreads + threadIdx.x;
    uint32_t nv = 0, code = 0;
    if (i < r * r * r) {
        float val[8];
        code = cube_case(f, s, (int32_t)(i % r), (int32_t)((i / r) % r), (int32_t)(i / (r * r)), m.iso, val);
        nv = c_nverts[code];
    }
    // exclusive scan of the cells' vertex counts in the workgroup
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = nv;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t up = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += up;
    }
    if (lane == 63u) waves[wave] = incl;
    cellCode[threadIdx.x] = (uint8_t)code;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kThreads / 64; ++w) {
        before += w < wave ? waves[w] : 0u;
        total += waves[w];
    }
    cellOff[threadIdx.x] = before + incl - nv;
    if (threadIdx.x == 0) cellOff[kThreads] = total;
    if (nv && base + before + incl > m.capacity) atomicMin(&counters[1], base + before + incl - nv);
    if (last && threadIdx.x == 0) counters[0] = base + total;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < total; k += kThreads) {
        // the cell holding vertex k: the last cell whose offset is <= k (empty cells share
        // their successor's offset, so that is the one with vertices)
        uint32_t lo = 0, hi = kThreads;   // cellOff[lo] <= k < cellOff[hi]
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cellOff[mid] <= k) lo = mid;
            else hi = mid;
        }
        const uint32_t c = lo, cc = cellCode[c];
        const uint32_t first = base + cellOff[c];
        if (first + c_nverts[cc] > m.capacity) continue;   // the cell does not fit: none of it
        const uint32_t ci = b * kThreads + c;
        const int32_t bx = (int32_t)(ci % r), by = (int32_t)((ci / r) % r), bz = (int32_t)(ci / (r * r));
        const uint32_t e = (uint32_t)c_table[cc * 16u + (k - cellOff[c])];
        const uint32_t a = c_edge[e][0], bb = c_edge[e][1];
        const int32_t ca[3] = {bx + c_corner[a][0], by + c_corner[a][1], bz + c_corner[a][2]};
        const int32_t cb[3] = {bx + c_corner[bb][0], by + c_corner[bb][1], bz + c_corner[bb][2]};
        const float va = field_at(f, s, ca[0], ca[1], ca[2]), vb = field_at(f, s, cb[0], cb[1], cb[2]);
        float t = fabsf(vb - va) > 1.0e-8f ? (m.iso - va) / (vb - va) : 0.5f;
        t = t < 0.0f ? 0.0f : (1.0f < t ? 1.0f : t);
        float pnt[3], ga[3], gb[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float pa = m.origin[q] + (float)ca[q] * m.cellSize[q];
            const float pb = m.origin[q] + (float)cb[q] * m.cellSize[q];
            pnt[q] = pa + (pb - pa) * t;
        }
        // central differences over the cell size (:446-455)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int32_t ux = q == 0, uy = q == 1, uz = q == 2;
            ga[q] = (field_at(f, s, ca[0] + ux, ca[1] + uy, ca[2] + uz) -
                     field_at(f, s, ca[0] - ux, ca[1] - uy, ca[2] - uz)) / m.cellSize[q];
            gb[q] = (field_at(f, s, cb[0] + ux, cb[1] + uy, cb[2] + uz) -
                     field_at(f, s, cb[0] - ux, cb[1] - uy, cb[2] - uz)) / m.cellSize[q];
        }
        float n[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) n[q] = -(ga[q] + (gb[q] - ga[q]) * t);   // outward: against the gradient
        const float dd = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2];
        if (!(sqrtf(dd) > 1.0e-12f)) {
            n[0] = 0.0f;
            n[1] = 1.0f;
            n[2] = 0.0f;
        }
        const float d2 = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2];
        const float inv = 1.0f / sqrtf(d2);   // glm::normalize: v * inversesqrt(dot(v, v))
        SphSurfaceVertex o;
        o.position[0] = pnt[0];
        o.position[1] = pnt[1];
        o.position[2] = pnt[2];
        o.normal = pack_unit(n[0] * inv, n[1] * inv, n[2] * inv);
        out[base + k] = o;
    }
}

int upload_table() {
    const int8_t* t = sphhost::mc_table();
    if (!t) return fail(SPH_ERR_STATE, "marching-cubes table generation failed");
    uint8_t nverts[256];
    for (uint32_t c = 0; c < 256; ++c) {
        uint32_t v = 0;
        while (v < 16u && t[16u * c + v] >= 0) ++v;
        nverts[c] = (uint8_t)v;
    }
    SPH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_table), t, 256 * 16));
    SPH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_nverts), nverts, 256));
    return SPH_OK;
}

// Workgroups of the marching passes, and the device words march_on_device works in: two
// counters, the workgroups' vertex offsets (one more for the scan) and the scan's tile sums.
uint32_t march_blocks(uint32_t resolution) {
    const size_t cells = (size_t)resolution * resolution * resolution;
    return (uint32_t)((cells + kThreads - 1) / kThreads);
}

size_t march_scratch_words(uint32_t resolution) {
    const uint32_t blocks = march_blocks(resolution);
    const uint32_t tiles = (blocks + sphdev::kScanTile - 1) / sphdev::kScanTile;
    return 2 + ((size_t)blocks + 1) + ((size_t)tiles + 1);
}

// Field of the solver's current state into the device buffer `field` ((res + 1)^3 floats).
int field_on_device(sphhost::Engine& e, const SphParams& p, uint32_t resolution, float* field) {
    const uint32_t s = resolution + 1;
    const size_t samples = (size_t)s * s * s;
    if (e.sc[0].n == 0) {
        SPH_HIP(hipMemsetAsync(field, 0, samples * sizeof(float), e.stream));
        return SPH_OK;
    }
    if (int st = e.build_grid()) return st;
    FieldArgs a;
    for (int k = 0; k < 3; ++k) {
        a.origin[k] = p.containerMin[k];
        // glm: (containerMax - containerMin) / float(resolution), per component
        a.cell[k] = (p.containerMax[k] - p.containerMin[k]) / (float)resolution;
    }
    a.radius = p.smoothingRadius;
    a.radiusSq = p.smoothingRadius * p.smoothingRadius;
    a.samples = s;
    a.resolution = resolution;
    a.cellMask = p.cellCount - 1u;
    const uint32_t tilesX = (s + kTileX - 1) / kTileX, tilesY = (s + kTileY - 1) / kTileY,
                   tilesZ = (s + kTileZ - 1) / kTileZ;
    k_field<<<tilesX * tilesY * tilesZ, kThreads, 0, e.stream>>>(e.sorted_pos(), e.bucket_table(), a, tilesX,
                                                                 tilesY, field);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

// `scratch`: march_scratch_words(resolution) device words, overwritten.
int march_on_device(const float* field, uint32_t resolution, const float origin[3], const float cellSize[3],
                    float iso, SphSurfaceVertex* out, uint32_t capacity, uint32_t* written, uint32_t* scratch,
                    hipStream_t s) {
    if (int st = upload_table()) return st;
    const uint32_t init[2] = {0u, capacity};
    uint32_t* counters = scratch;
    uint32_t* blockStart = scratch + 2;
    const uint32_t blocks = march_blocks(resolution);
    uint32_t* tileSums = blockStart + blocks + 1;
    hipError_t he = hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, s);
    MarchArgs m;
    for (int k = 0; k < 3; ++k) {
        m.origin[k] = origin[k];
        m.cellSize[k] = cellSize[k];
    }
    m.iso = iso;
    m.resolution = resolution;
    m.capacity = capacity;
    if (he == hipSuccess) {
        k_march_count<<<blocks, kThreads, 0, s>>>(field, m, blockStart);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = sphdev::launch_scan(blockStart, blocks, 0u, tileSums, s);
    if (he == hipSuccess) {
        k_march_emit<<<blocks, kThreads, 0, s>>>(field, m, blockStart, blocks, out, counters);
        he = hipGetLastError();
    }
    uint32_t got[2] = {0u, 0u};
    if (he == hipSuccess) he = hipMemcpyAsync(got, counters, sizeof(got), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return hip_fail(he, "marching cubes");
    *written = got[0] < got[1] ? got[0] : got[1];
    return SPH_OK;
}

// The solver's field, vertex and marching buffers, grown to at least these sizes (never shrunk).
int ensure_surface_buffers(SphSolver* solver, size_t fieldBytes, uint32_t vertices, size_t marchWords) {
    if (solver->surfFieldBytes < fieldBytes) {
        hipFree(solver->surfField);
        solver->surfField = nullptr;
        solver->surfFieldBytes = 0;
        SPH_HIP(hipMalloc(&solver->surfField, fieldBytes));
        solver->surfFieldBytes = fieldBytes;
    }
    if (solver->surfVertCap < vertices) {
        hipFree(solver->surfVerts);
        solver->surfVerts = nullptr;
        solver->surfVertCap = 0;
        SPH_HIP(hipMalloc(&solver->surfVerts, (size_t)vertices * sizeof(SphSurfaceVertex)));
        solver->surfVertCap = vertices;
    }
    if (solver->surfMarchWords < marchWords) {
        hipFree(solver->surfMarch);
        solver->surfMarch = nullptr;
        solver->surfMarchWords = 0;
        SPH_HIP(hipMalloc(&solver->surfMarch, marchWords * sizeof(uint32_t)));
        solver->surfMarchWords = marchWords;
    }
    return SPH_OK;
}

int check_resolution(uint32_t resolution) {
    if (resolution < 1 || resolution > kMaxResolution)
        return fail(SPH_ERR_INVALID_ARGUMENT, "resolution must be in 1..1024");
    return SPH_OK;
}

}  // namespace

// physics::buildFluidField for the solver's current particles: (resolution + 1)^3 samples,
// x fastest, into the host array `field`.
extern "C" int sph_surface_field(SphSolver* solver, const SphParams* p, uint32_t resolution, float* field) {
    sphhost::clear_error();
    if (!solver || !p || !field) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_surface_field: null argument");
    if (int st = check_resolution(resolution)) return st;
    sphhost::Engine& e = solver->e;
    if (int st = e.set_params(0, *p)) return st;
    if (int st = e.set_device()) return st;
    const uint32_t s = resolution + 1;
    const size_t bytes = (size_t)s * s * s * sizeof(float);
    if (int st = ensure_surface_buffers(solver, bytes, 0, 0)) return st;
    int st = field_on_device(e, *p, resolution, solver->surfField);
    if (!st) {
        hipError_t he = hipMemcpyAsync(field, solver->surfField, bytes, hipMemcpyDeviceToHost, e.stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e.stream);
        if (he != hipSuccess) st = hip_fail(he, "sph_surface_field");
    }
    return st;
}

// physics::marchFluidField over a host field ((resolution + 1)^3 samples, x fastest):
// at most `capacity` vertices (three per triangle) into `out`, their count in *written.
extern "C" int sph_march_field(const float* field, uint32_t resolution, const float origin[3],
                               const float cellSize[3], float iso, SphSurfaceVertex* out, uint32_t capacity,
                               uint32_t* written) {
    sphhost::clear_error();
    if (!field || !origin || !cellSize || !written || (capacity && !out))
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_march_field: null argument");
    if (int st = check_resolution(resolution)) return st;
    *written = 0;
    if (capacity < 3) return SPH_OK;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device");
    const uint32_t s = resolution + 1;
    const size_t bytes = (size_t)s * s * s * sizeof(float);
    float* d = nullptr;
    SphSurfaceVertex* v = nullptr;
    uint32_t* counters = nullptr;
    SPH_HIP(hipMalloc(&d, bytes));
    hipError_t he = hipMalloc(&v, (size_t)capacity * sizeof(SphSurfaceVertex));
    if (he == hipSuccess) he = hipMalloc(&counters, march_scratch_words(resolution) * sizeof(uint32_t));
    int st = SPH_OK;
    if (he == hipSuccess) he = hipMemcpy(d, field, bytes, hipMemcpyHostToDevice);
    if (he != hipSuccess) st = hip_fail(he, "sph_march_field");
    if (!st) st = march_on_device(d, resolution, origin, cellSize, iso, v, capacity, written, counters, nullptr);
    if (!st && *written) {
        he = hipMemcpy(out, v, (size_t)*written * sizeof(SphSurfaceVertex), hipMemcpyDeviceToHost);
        if (he != hipSuccess) st = hip_fail(he, "sph_march_field");
    }
    hipFree(d);
    hipFree(v);
    hipFree(counters);
    return st;
}

// Field and iso-surface of the solver's current particles in one go, on the device; only
// the vertices cross to the host. The grid spans the container; iso is dimensionless
// (the field sums (1 - r^2/h^2)^3 per particle).
extern "C" int sph_surface_mesh(SphSolver* solver, const SphParams* p, uint32_t resolution, float iso,
                                SphSurfaceVertex* out, uint32_t capacity, uint32_t* written) {
    sphhost::clear_error();
    if (!solver || !p || !written || (capacity && !out))
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_surface_mesh: null argument");
    if (int st = check_resolution(resolution)) return st;
    *written = 0;
    if (capacity < 3) return SPH_OK;
    sphhost::Engine& e = solver->e;
    if (int st = e.set_params(0, *p)) return st;
    if (int st = e.set_device()) return st;
    const uint32_t s = resolution + 1;
    if (int st = ensure_surface_buffers(solver, (size_t)s * s * s * sizeof(float), capacity,
                                        march_scratch_words(resolution)))
        return st;
    float* d = solver->surfField;
    SphSurfaceVertex* v = solver->surfVerts;
    int st = field_on_device(e, *p, resolution, d);
    float origin[3], cell[3];
    for (int k = 0; k < 3; ++k) {
        origin[k] = p->containerMin[k];
        cell[k] = (p->containerMax[k] - p->containerMin[k]) / (float)resolution;
    }
    if (!st) st = march_on_device(d, resolution, origin, cell, iso, v, capacity, written, solver->surfMarch, e.stream);
    if (!st && *written) {
        const hipError_t he = hipMemcpy(out, v, (size_t)*written * sizeof(SphSurfaceVertex), hipMemcpyDeviceToHost);
        if (he != hipSuccess) st = hip_fail(he, "sph_surface_mesh");
    }
    return st;
}
