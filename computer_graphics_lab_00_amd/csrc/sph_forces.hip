// sph_forces.hip — density pass and force + integrate + boundary pass for gfx950.
//
// Replace the two parallel loops of FluidSolver::substep (src/physics/fluid_sph.cpp:197-305)
// and shaders/fluid_density.comp / fluid_integrate.comp. One lane owns one particle of
// the bucket-sorted state and walks its 27 neighbour buckets in the reference order
// (dz outer, dy, dx inner), each bucket's slots in ascending particle index, so every
// floating-point sum is accumulated in exactly the reference's order. Built with
// -ffp-contract=off and correctly rounded division / square root, so each operation
// rounds as the reference's does.
//
// Consecutive lanes hold particles of the same bucket (the sort put them together), so
// a wave's candidate loads hit the same few cache lines: neighbour reads are served
// from L1/L2 and HBM sees each particle's state about once per pass.
#include <type_traits>

#include "sph_device.h"

namespace sphdev {

namespace {

__device__ inline uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & mask;
}

__device__ inline int32_t cell_coord(float x, float h) { return (int32_t)floorf(x / h); }

// The neighbour list is written once and read once per substep. Its 16-byte groups are
// written with ordinary stores, so the lines stay in the caches for the force pass that
// reads them next (at 1M particles: force pass 0.15 ms against 0.17 ms when the groups were
// written non-temporally, with the density pass unchanged); the force pass reads them as
// streaming (non-temporal) loads, since nothing reads them again.
typedef unsigned int u32x4_nt __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void store_list(uint4* p, uint4 v) { *p = v; }

__device__ __forceinline__ uint4 load_stream(const uint4* p) {
    const u32x4_nt x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt*>(p));
    return make_uint4(x.x, x.y, x.z, x.w);
}

// Whether the bucket of neighbour offset `o` (0..26) already appeared at an earlier offset.
__device__ inline bool seen_before(int32_t cx, int32_t cy, int32_t cz, uint32_t o, uint32_t b, uint32_t mask) {
    for (uint32_t e = 0; e < o; ++e) {
        int32_t dx = (int32_t)(e % 3u) - 1, dy = (int32_t)((e / 3u) % 3u) - 1, dz = (int32_t)(e / 9u) - 1;
        if (hash_cell(cx + dx, cy + dy, cz + dz, mask) == b) return true;
    }
    return false;
}

__device__ inline float eos_pressure(const StepConsts& k, float density) {
    float p;
    if (k.eos == SPH_EOS_TAIT) {
        float x = density / k.restDensity;
        float x2 = x * x;
        float x4 = x2 * x2;
        float x7 = (x4 * x2) * x;
        p = k.taitB * (x7 - 1.0f);
    } else {
        p = k.stiffness * (density - k.restDensity);   // src/physics/fluid_sph.cpp:220
    }
    return p < 0.0f ? 0.0f : p;
}

// The neighbourhood sums of NeighbourSums (sph_device.h) for one in-range candidate with
// offset (ddx, ddy, ddz) = x_t - x_j, d = h2 - r2 and velocity vj, accumulated in ax.
__device__ __forceinline__ void accumulate_sums(float ax[20], float ddx, float ddy, float ddz, float d, float4 vj,
                                                float vx, float vy, float vz, float invH2, float invH) {
    const float sw = d * invH2;
    const float ex = ddx * invH, ey = ddy * invH, ez = ddz * invH;
    const float s2 = sw * sw, s3 = s2 * sw;
    const float wx = vj.x - vx, wy = vj.y - vy, wz = vj.z - vz;
    ax[0] += s2 * ex;
    ax[1] += s2 * ey;
    ax[2] += s2 * ez;
    ax[3] += 1.0f;
    ax[4] += s3 * (ex * ex);
    ax[5] += s3 * (ey * ey);
    ax[6] += s3 * (ez * ez);
    ax[7] += s3 * (ex * ey);
    ax[8] += s3 * (ex * ez);
    ax[9] += s3 * (ey * ez);
    ax[10] += s2 * ((wx * ex + wy * ey) + wz * ez);
    ax[11] += sw;
    ax[12] += sw * wx;
    ax[13] += sw * wy;
    ax[14] += sw * wz;
    ax[16] += s2 * (wy * ez - wz * ey);
    ax[17] += s2 * (wz * ex - wx * ez);
    ax[18] += s2 * (wx * ey - wy * ex);
}

// Density pass. Besides density and pressure it records, per particle, the slots of
// every candidate with r2 < h2 in visiting order (duplicates included). The force
// pass's candidates are a subset of these (its cut-off is r2 < r2Cut <= h2 on the same
// positions), so it walks ~33 entries instead of the 27 buckets. Entries are stored as
// 16-byte groups of four; group g of particle t sits at groups[g * n + t], so the force
// pass reads a whole wave's group g in one coalesced access. A particle with more than
// `listCap` entries stores count = listCap + 1 and the force pass scans its buckets
// instead.
// One lane per particle walks its 27 buckets. Candidates are read nine at a time (four with
// SUMS) from one base address (a typical bucket is one batch, one wait for memory); the last
// batch of a bucket may read up to eight slots past it (the position arrays carry kPosPad
// slots of padding) and ignores them. Nine rather than eight because buckets of the dam
// breaks' lattices (two particles per cell edge, jittered) hold 7 to 10 members: at 1M
// particles the pass took 259 us with batches of nine against 266 us with eight (6: 261,
// 7: 277, 10: 260, 12: 288), and the 64-scene batch 7.59 against 7.83 ms per step. Per candidate the work is the reference's arithmetic, a
// select and one bit of the batch's in-range mask. After the batch, its in-range slots
// (~15 % of the candidates) go into the lane's ring of 16 pending entries in LDS, one
// write each, and every completed group of four goes out to the list: no branch per
// candidate, no list work for candidates out of range. (Measured at 1M particles: 0.30 ms
// per pass against 0.34 ms when each candidate shifted a four-entry register group and
// stored it under a branch.)
// With SUMS the pass also accumulates the neighbourhood sums of NeighbourSums
// (sph_device.h) over the same candidates, in the same order, and builds no list. The
// in-range slots go into the lane's ring of 16 in LDS instead, and whenever a lane of the
// wave has more than 12 pending, every lane of the wave sums its pending candidates (their
// offsets recomputed from the positions with the visit's operations, so the sums are the
// ones the visit would give). Summing inside the visit would run the ~45 instructions of
// the sums for every candidate that any lane of the wave has in range; at 1M particles in a
// developed flow this pass took 0.54 ms that way against 0.31 ms for the pass without sums.
// With `deviation` it also writes each workgroup's sums of max(rho / rho0 - 1, 0) and
// |rho / rho0 - 1| (the density check's first stage), in the tree the check's own reduction
// uses, so the check needs no pass of its own.
// DEDUPE = false is the variant for launches where no scene asked for de-duplicated
// buckets: it keeps no bucket hashes live. With the bucket descriptors fetched one dy row
// (three buckets) at a time, the reference-mode pass fits 62 VGPRs and the pass with sums,
// reading four candidates per batch, 64: 8 waves per SIMD for both (16 KB of LDS per
// workgroup). Occupancy is what hides these passes' load latency: at 1M particles the pass
// with sums ran 6 % faster rollouts this way than at 5 waves (nine descriptors per z layer,
// eight candidates per batch, a ring of 32), and each wave fewer than 8 measured slower.
// FAST (instantiated without SUMS only): scenes with SPH_MODE_FAST_MATH evaluate r2 with
// fused multiply-adds and poly6 with the mass folded into its constant; other scenes of the
// launch keep the reference's arithmetic.
template <bool SUMS, bool DEDUPE, bool FAST = false>
__global__ void __launch_bounds__(kBlock, !DEDUPE ? (FAST ? 6 : 8) : 1)
    k_density(Launch L, float4* __restrict__ pos, float4* __restrict__ vel, const uint2* __restrict__ buckets,
              uint4* __restrict__ nbrList, uint32_t* __restrict__ nbrCount,
              float4* __restrict__ sums, float2* __restrict__ deviation) {
    const Where w = locate(L);
    const bool live = w.valid && w.d->active;
    // With a deviation output every lane stays to the end, where the workgroup reduces it.
    if (!live && !(SUMS && deviation)) return;
    const StepConsts k = w.d->k;   // by value: the constants stay in scalar registers
    const uint32_t t = w.slot;
    const uint32_t cellBase = w.d->cellOffset;
    // Only xyz of any slot is read here; .w of this lane's own slot is written below
    // (the reference writes positions[i].w in place the same way, fluid_sph.cpp:197-199).
    const float* self = reinterpret_cast<const float*>(pos + (live ? t : w.d->offset));
    const float px = self[0], py = self[1], pz = self[2];
    const int32_t cx = cell_coord(px, k.h), cy = cell_coord(py, k.h), cz = cell_coord(pz, k.h);
    const bool capped = (k.mode & SPH_MODE_UNCAPPED) == 0;
    const bool dedupe = DEDUPE && (k.mode & SPH_MODE_DEDUPE) != 0;
    const uint32_t cap = (!SUMS && nbrList) ? k.listCap : 0u;
    float density = 0.0f;
    uint32_t listed = 0, flushed = 0;
    // candidates per batch (SUMS: four, which keeps the pass at 8 waves per SIMD), and the
    // ring of pending entries per lane: list entries, or with SUMS in-range slots to be summed
    constexpr uint32_t kB = SUMS ? 4u : 9u;
    constexpr uint32_t kSumRing = 16;
    static_assert(kB - 1u <= kPosPad, "a batch reads at most kB - 1 slots past a bucket, into the padding");
    static_assert(kB + 3u <= 16u, "the list ring holds a batch and the up to 3 entries pending before it");
    __shared__ uint32_t ring[16][kBlock];
    uint32_t pushed = 0, summed = 0;   // SUMS: in-range slots staged in the ring / summed
    // neighbourhood sums (SUMS only)
    float ax[20];
#pragma unroll
    for (int q = 0; q < 20; ++q) ax[q] = 0.0f;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f, invH2 = 0.0f, invH = 0.0f;
    if (SUMS) {
        // Lanes past the scene's particles stay alive only for the workgroup reduction:
        // they read nothing (the state arrays end at the last slot).
        const float4 v = live ? vel[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        vx = v.x;
        vy = v.y;
        vz = v.z;
        invH2 = 1.0f / k.h2;
        invH = 1.0f / k.h;
    }
    const bool fast = FAST && (k.mode & SPH_MODE_FAST_MATH) != 0;
    const float massPoly6 = k.mass * k.cPoly6;
    // One candidate at slot u, in visiting order; `valid` is false for the padding reads.
    // Returns whether it is in range.
    auto visit = [&](bool valid, uint32_t u, float qx, float qy, float qz) -> bool {
        const float ddx = px - qx, ddy = py - qy, ddz = pz - qz;
        if (FAST && fast) {
            const float r2f = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
            const bool inrf = valid && r2f < k.h2;
            const float df = k.h2 - r2f;
            density = inrf ? fmaf(massPoly6, (df * df) * df, density) : density;
            return inrf;
        }
        const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
        // poly6 is 0 at r2 >= h2 and density + mass * 0 == density: such candidates leave
        // the sum as it is.
        const bool inr = valid && r2 < k.h2;
        const float d = k.h2 - r2;
        const float term = k.mass * (((k.cPoly6 * d) * d) * d);
        density = inr ? density + term : density;
        return inr;
    };
    // SUMS: the staged in-range candidates, in order; their offsets are recomputed from the
    // positions with the visit's operations.
    auto sum_pending = [&]() {
        for (; summed < pushed; ++summed) {
            const uint32_t u = ring[summed & (kSumRing - 1u)][threadIdx.x];
            const float* q = reinterpret_cast<const float*>(pos + u);
            const float ddx = px - q[0], ddy = py - q[1], ddz = pz - q[2];
            const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
            accumulate_sums(ax, ddx, ddy, ddz, k.h2 - r2, vel[u], vx, vy, vz, invH2, invH);
        }
    };
    for (int32_t dz = -1; dz <= 1; ++dz)
    for (uint32_t g0 = 0; g0 < 9u; g0 += 3u) {
        // The three buckets of one dy row: their descriptors are fetched before any is used.
        uint32_t bk[3], cn[3], st[3];
#pragma unroll
        for (uint32_t oo = 0; oo < 3u; ++oo) {
            const uint32_t o = g0 + oo;
            bk[oo] = hash_cell(cx + (int32_t)(o % 3u) - 1, cy + (int32_t)(o / 3u) - 1, cz + dz, k.cellMask);
            const uint2 bd = buckets[cellBase + bk[oo]];
            st[oo] = bd.x;
            cn[oo] = bd.y;
        }
#pragma unroll
        for (uint32_t oo = 0; oo < 3u; ++oo) {
            const uint32_t o = g0 + oo;
            if (dedupe && seen_before(cx, cy, cz, (uint32_t)(dz + 1) * 9u + o, bk[oo], k.cellMask)) continue;
            uint32_t cnt = live ? cn[oo] : 0u;
            if (cnt == 0) continue;
            const uint32_t lo = st[oo];
            if (capped && cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
            const float4* base = pos + lo;
            for (uint32_t j = 0; j < cnt; j += kB) {
                float qx[kB], qy[kB], qz[kB];
#pragma unroll
                for (uint32_t e = 0; e < kB; ++e) {
                    const float* q = reinterpret_cast<const float*>(base + j + e);
                    qx[e] = q[0];
                    qy[e] = q[1];
                    qz[e] = q[2];
                }
                uint32_t m = 0;
#pragma unroll
                for (uint32_t e = 0; e < kB; ++e)
                    m |= visit(j + e < cnt, lo + j + e, qx[e], qy[e], qz[e]) ? (1u << e) : 0u;
                if (SUMS) {
                    for (uint32_t mm = m; mm != 0u; mm &= mm - 1u) {
                        ring[pushed & (kSumRing - 1u)][threadIdx.x] = lo + j + (uint32_t)__builtin_ctz(mm);
                        ++pushed;
                    }
                    if (__any(pushed - summed > kSumRing - kB)) sum_pending();   // every lane here empties its ring
                }
                // The batch's in-range slots join the list in order, one per set bit, through
                // this lane's ring of 16 pending entries in LDS; completed groups of four go out.
                uint32_t mm = cap ? m : 0u;
                for (; mm != 0u && listed < cap; mm &= mm - 1u) {
                    ring[listed & 15u][threadIdx.x] = lo + j + (uint32_t)__builtin_ctz(mm);
                    ++listed;
                }
                listed += (uint32_t)__builtin_popcount(mm);   // past the cap: counted only
                const uint32_t done = listed < cap ? listed : cap;
                for (; flushed + 4u <= done; flushed += 4u) {
                    const uint32_t r0 = flushed & 15u;
                    store_list(nbrList + (size_t)(flushed >> 2) * L.total + t,
                                 make_uint4(ring[r0][threadIdx.x], ring[r0 + 1u][threadIdx.x],
                                            ring[r0 + 2u][threadIdx.x], ring[r0 + 3u][threadIdx.x]));
                }
            }
        }
    }
    if (SUMS) sum_pending();
    density = density < 1e-6f ? 1e-6f : density;
    if (SUMS && deviation) {
        // the density check's workgroup sums, in the same tree as a separate reduction pass
        const float dv = density / k.restDensity - 1.0f;
        __syncthreads();   // every lane is done with its ring
        float* sa = reinterpret_cast<float*>(&ring[0][0]);
        float* sb = sa + kBlock;
        sa[threadIdx.x] = live ? (dv > 0.0f ? dv : 0.0f) : 0.0f;
        sb[threadIdx.x] = live ? (dv < 0.0f ? -dv : dv) : 0.0f;
        __syncthreads();
        for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) {
                sa[threadIdx.x] += sa[threadIdx.x + h];
                sb[threadIdx.x] += sb[threadIdx.x + h];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) deviation[blockIdx.x] = make_float2(sa[0], sb[0]);
        if (!live) return;
    }
    const float pressure = eos_pressure(k, density);
    reinterpret_cast<float*>(pos + t)[3] = density;
    reinterpret_cast<float*>(vel + t)[3] = pressure;
    raise_range_guard(L.rangeGuard, !(density_in_range(density) && pressure_in_range(pressure)));
    if (nbrCount) {
        if (listed <= cap && flushed < listed) {
            // the last, partial group
            const uint32_t r0 = flushed & 15u, r = listed - flushed;
            store_list(nbrList + (size_t)(flushed >> 2) * L.total + t,
                         make_uint4(ring[r0][threadIdx.x], r > 1u ? ring[r0 + 1u][threadIdx.x] : 0u,
                                    r > 2u ? ring[r0 + 2u][threadIdx.x] : 0u, 0u));
        }
        // cap == 0: this scene keeps no list (its force cut-off is not a subset of h2)
        nbrCount[t] = cap && listed <= cap ? listed : cap + 1u;
    }
    if (SUMS) {
#pragma unroll
        for (uint32_t r = 0; r < kSumRows; ++r)
            sums[(size_t)r * L.total + t] = make_float4(ax[4 * r], ax[4 * r + 1], ax[4 * r + 2], ax[4 * r + 3]);
    }
}

// ---- tiled density pass ----
//
// The plain step's density pass (fluid_sph.cpp:197-222) for launches whose scenes all use
// one spatial bin per cell. One workgroup per half-tile: 32 spatial bins, i.e. a block of
// 32 cells (4 x 4 x 2 for the default Morton tile) and the particles sorted into them. The
// workgroup first loads, for every cell of the block and of the cells around it
// (6 x 6 x 4 = 144 cells), its bucket's first slot, member count (capped at 32 unless
// uncapped) and the bounding box of the members the walk visits into LDS: the 27 lookups
// per particle of the plain walk become about one per particle, and lanes of one cell read
// the same LDS entries.
//
// Culling: a bucket whose bounding box is at least h away from the particle cannot hold a
// candidate in range, and is skipped. This is exact, not a tolerance: with g the gap between
// the particle and the box per axis, computed as px - hi (or lo - px), and
// G2 = (gx gx + gy gy) + gz gz in the visit's order, every member q has
// |fl(px - qx)| >= gx per axis (rounding is monotone) and therefore a computed r2 >= G2;
// G2 >= h2 means every member's r2 >= h2, and a candidate out of range adds nothing to the
// sum (poly6 is 0 there and density + 0 == density). The density and the neighbour list
// are therefore the ones the full walk gives, bit for bit. At rest about 40 % of the
// candidates survive.
//
// Each lane then walks its kept buckets in the reference's order at its own pace, four
// candidates per iteration read from the sorted state in global memory, so a lane does not
// wait at every bucket for the wave-mate with the largest one.
//
// Particles whose own cell lies outside the block they are sorted into (members of a
// bucket shared with a cell elsewhere: a bucket's members sit in the bin of its smallest
// cell) were listed by the grid build; the launch's first `restBlocks` workgroups walk them
// as the plain pass does, while the half-tiles run.
constexpr uint32_t kBoxCells = 256;    // cells of the neighbourhood table, at most
constexpr uint32_t kRing = 16;         // pending neighbour-list entries per lane (batches of 8)

// The in-range slots of one batch of B candidates (bit e of m: slot base + e) join the list
// in order through the lane's ring of kRing rows (row kRing takes the writes of the slots that
// do not join); completed groups of four are written out. At most 3 entries are pending before
// a batch, so kRing >= B + 3 rows hold them. Past `cap` entries are only counted (the force
// pass then walks the buckets itself).
template <uint32_t B>
__device__ __forceinline__ void list_push(uint32_t (*ring)[kBlock], uint32_t m, uint32_t base, uint32_t cap,
                                          uint32_t& listed, uint32_t& flushed, uint4* __restrict__ nbrList,
                                          size_t stride, uint32_t t) {
    static_assert(B + 3 <= kRing, "ring too small for the batch");
#pragma unroll
    for (uint32_t e = 0; e < B; ++e) {
        const uint32_t at = listed + (uint32_t)__builtin_popcount(m & ((1u << e) - 1u));
        const bool put = ((m >> e) & 1u) != 0u && at < cap;
        ring[put ? (at & (kRing - 1u)) : kRing][threadIdx.x] = base + e;
    }
    listed += (uint32_t)__builtin_popcount(m);
    const uint32_t done = listed < cap ? listed : cap;
#pragma unroll
    for (uint32_t g = 0; g < (B + 3) / 4; ++g) {
        if (flushed + 4u <= done) {
            const uint32_t r0 = flushed & (kRing - 1u);
            store_list(nbrList + (size_t)(flushed >> 2) * stride + t,
                         make_uint4(ring[r0][threadIdx.x], ring[r0 + 1u][threadIdx.x], ring[r0 + 2u][threadIdx.x],
                                    ring[r0 + 3u][threadIdx.x]));
            flushed += 4u;
        }
    }
}

// The density, pressure and list count of particle t once its walk is done.
__device__ __forceinline__ void density_finish(const StepConsts& k, float density, float4* __restrict__ pos,
                                               float4* __restrict__ vel, uint32_t t, uint32_t cap, uint32_t listed,
                                               uint32_t flushed, uint32_t (*ring)[kBlock],
                                               uint4* __restrict__ nbrList, uint32_t* __restrict__ nbrCount,
                                               size_t stride, uint32_t* rangeGuard) {
    density = density < 1e-6f ? 1e-6f : density;
    const float pressure = eos_pressure(k, density);
    reinterpret_cast<float*>(pos + t)[3] = density;
    reinterpret_cast<float*>(vel + t)[3] = pressure;
    raise_range_guard(rangeGuard, !(density_in_range(density) && pressure_in_range(pressure)));
    if (nbrCount) {
        if (listed <= cap && flushed < listed) {
            const uint32_t r0 = flushed & (kRing - 1u), r = listed - flushed;
            store_list(nbrList + (size_t)(flushed >> 2) * stride + t,
                         make_uint4(ring[r0][threadIdx.x], r > 1u ? ring[r0 + 1u][threadIdx.x] : 0u,
                                    r > 2u ? ring[r0 + 2u][threadIdx.x] : 0u, 0u));
        }
        nbrCount[t] = cap && listed <= cap ? listed : cap + 1u;
    }
}

// Candidates base[0..B-1] (valid: index < validCount) of particle (px, py, pz): the
// reference's density term for each in range, in order; returns the in-range mask. The reads
// may run up to B - 1 slots past the bucket (the position arrays are padded by kPosPad).
template <uint32_t B>
__device__ __forceinline__ uint32_t density_batch(const StepConsts& k, float px, float py, float pz,
                                                  const float4* __restrict__ base, uint32_t validCount,
                                                  float& density) {
    static_assert(B <= kPosPad, "reads past a bucket stay inside the padding");
    float qx[B], qy[B], qz[B];
#pragma unroll
    for (uint32_t e = 0; e < B; ++e) {
        const float* q = reinterpret_cast<const float*>(base + e);
        qx[e] = q[0];
        qy[e] = q[1];
        qz[e] = q[2];
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t e = 0; e < B; ++e) {
        const float ddx = px - qx[e], ddy = py - qy[e], ddz = pz - qz[e];
        const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
        const bool inr = e < validCount && r2 < k.h2;
        const float d = k.h2 - r2;
        const float term = k.mass * (((k.cPoly6 * d) * d) * d);
        density = inr ? density + term : density;
        m |= inr ? (1u << e) : 0u;
    }
    return m;
}

// The plain walk of one listed particle (slot t of scene d), in global memory.
__device__ __forceinline__ void density_walk_global(const SceneDesc* d, uint32_t t, float4* __restrict__ pos,
                                                    float4* __restrict__ vel, const uint2* __restrict__ buckets,
                                                    uint4* __restrict__ nbrList,
                                                    uint32_t* __restrict__ nbrCount, uint32_t (*ring)[kBlock],
                                                    size_t stride, uint32_t* rangeGuard) {
    const StepConsts k = d->k;
    const uint32_t cellBase = d->cellOffset;
    const bool capped = (k.mode & SPH_MODE_UNCAPPED) == 0;
    const uint32_t cap = nbrList ? k.listCap : 0u;
    const float* self = reinterpret_cast<const float*>(pos + t);
    const float px = self[0], py = self[1], pz = self[2];
    const int32_t cx = cell_coord(px, k.h), cy = cell_coord(py, k.h), cz = cell_coord(pz, k.h);
    float density = 0.0f;
    uint32_t listed = 0, flushed = 0;
    for (int32_t dz = -1; dz <= 1; ++dz) {
        uint32_t cn[9], st[9];
#pragma unroll
        for (uint32_t o = 0; o < 9; ++o) {
            const uint32_t b = cellBase + hash_cell(cx + (int32_t)(o % 3u) - 1, cy + (int32_t)(o / 3u) - 1, cz + dz,
                                                    k.cellMask);
            const uint2 bd = buckets[b];
            st[o] = bd.x;
            cn[o] = bd.y;
        }
        for (uint32_t o = 0; o < 9; ++o) {
            uint32_t cnt = cn[o];
            if (capped && cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
            for (uint32_t j = 0; j < cnt; j += 8) {
                const uint32_t m = density_batch<8>(k, px, py, pz, pos + st[o] + j, cnt - j, density);
                list_push<8>(ring, m, st[o] + j, cap, listed, flushed, nbrList, stride, t);
            }
        }
    }
    density_finish(k, density, pos, vel, t, cap, listed, flushed, ring, nbrList, nbrCount, stride, rangeGuard);
}

__global__ void __launch_bounds__(kBlock, 8)
    k_density_tiled(Launch L, const uint32_t* __restrict__ binStart, float4* __restrict__ pos, float4* __restrict__ vel, const uint2* __restrict__ buckets,
                    uint4* __restrict__ nbrList, uint32_t* __restrict__ nbrCount,
                    const uint2* __restrict__ rest, const uint32_t* __restrict__ restCount) {
    __shared__ float4 boxLo[kBoxCells];   // per neighbourhood cell: bucket box low corner, .w = first slot
    __shared__ float4 boxHi[kBoxCells];   // bucket box high corner, .w = member count (capped)
    __shared__ uint32_t ring[kRing + 1][kBlock];

    // The first restBlocks workgroups (dispatched first, so they run beside the half-tiles)
    // walk the listed particles, one per lane; the others are the half-tiles.
    const uint32_t restBlocks = gridDim.x - L.tiles;
    if (blockIdx.x < restBlocks) {
        const uint32_t n = *restCount;
        for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += restBlocks * kBlock) {
            const uint2 r = rest[i];
            const SceneDesc* d = L.scenes + r.y;
            if (d->active)
                density_walk_global(d, r.x, pos, vel, buckets, nbrList, nbrCount, ring, L.total, L.rangeGuard);
        }
        return;
    }
    const uint32_t lb = xcd_block(blockIdx.x - restBlocks, L.tiles);
    const uint32_t sIdx = L.tileScene[lb];
    const SceneDesc* d = L.scenes + sIdx;
    if (!d->active) return;
    const uint32_t ht = lb - d->tileOffset;
    const uint32_t b0 = d->binOffset + ht * kHalfTileBins;
    const uint32_t s0 = binStart[b0], s1 = binStart[b0 + kHalfTileBins];
    if (s0 >= s1) return;   // an empty half-tile (the same for every lane)
    const StepConsts k = d->k;
    const uint32_t cellBase = d->cellOffset;
    const bool capped = (k.mode & SPH_MODE_UNCAPPED) == 0;
    const uint32_t cap = nbrList ? k.listCap : 0u;

    // The half-tile's cells: tile (tx, ty, tz) of 2^tileLog2 bins per axis, halved along the
    // axis that holds bit 5 of the in-tile bin number (z for the Morton tile).
    const uint32_t lx = d->g.tileLog2[0], ly = d->g.tileLog2[1], lz = d->g.tileLog2[2];
    const uint32_t tile = ht >> 1, half = ht & 1u;
    int32_t dim[3] = {1 << lx, 1 << ly, 1 << lz};
    int32_t org[3] = {(int32_t)((tile % d->g.tiles[0]) << lx), (int32_t)(((tile / d->g.tiles[0]) % d->g.tiles[1]) << ly),
                      (int32_t)((tile / (d->g.tiles[0] * d->g.tiles[1])) << lz)};
    const int ha = lz ? 2 : (ly ? 1 : 0);
    dim[ha] >>= 1;
    org[ha] += (int32_t)half * dim[ha];
    // the neighbourhood: the block's cells and one more on each side
    const int32_t bx0 = d->g.origin[0] + org[0] - 1, by0 = d->g.origin[1] + org[1] - 1, bz0 = d->g.origin[2] + org[2] - 1;
    const uint32_t BX = (uint32_t)dim[0] + 2u, BY = (uint32_t)dim[1] + 2u, BZ = (uint32_t)dim[2] + 2u;
    const uint32_t BXY = BX * BY, nbox = BXY * BZ;   // <= kBoxCells: the planner checks the tile shape

    // Per neighbourhood cell, one thread (striding, so every entry is written whatever the
    // workgroup size): its bucket's first slot, member count and the bounding box of the
    // members the walk visits (read from the sorted state, which also brings them into the
    // caches the lanes read them from next).
    for (uint32_t kc = threadIdx.x; kc < nbox; kc += kBlock) {
        const int32_t cx = bx0 + (int32_t)(kc % BX), cy = by0 + (int32_t)((kc / BX) % BY), cz = bz0 + (int32_t)(kc / BXY);
        const uint32_t b = cellBase + hash_cell(cx, cy, cz, k.cellMask);
        const uint2 bd = buckets[b];
        uint32_t cnt = bd.y;
        if (capped && cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t first = 0;
        if (cnt) {
            first = bd.x;
            for (uint32_t m = 0; m < cnt; m += 4) {
                float q[4][3];
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) {
                    const float* c = reinterpret_cast<const float*>(pos + first + m + e);   // padded
                    q[e][0] = c[0];
                    q[e][1] = c[1];
                    q[e][2] = c[2];
                }
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) {
                    if (m + e < cnt) {
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            lo[a] = fminf(lo[a], q[e][a]);
                            hi[a] = fmaxf(hi[a], q[e][a]);
                        }
                    }
                }
            }
        }
        boxLo[kc] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(first));
        boxHi[kc] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(cnt));
    }
    __syncthreads();

    for (uint32_t t = s0 + threadIdx.x; t < s1; t += kBlock) {
        const float* self = reinterpret_cast<const float*>(pos + t);
        const float px = self[0], py = self[1], pz = self[2];
        const int32_t rx = cell_coord(px, k.h) - bx0 - 1, ry = cell_coord(py, k.h) - by0 - 1,
                      rz = cell_coord(pz, k.h) - bz0 - 1;
        // a particle of another cell's bucket: listed by the grid build, walked elsewhere
        if (!(rx >= 0 && rx < dim[0] && ry >= 0 && ry < dim[1] && rz >= 0 && rz < dim[2])) continue;
        // the 27 neighbour cells in the reference's order: the one at offset (dx, dy, dz),
        // each in 0..2, is table entry kb0 + dz BXY + dy BX + dx
        const uint32_t kb0 = (uint32_t)rz * BXY + (uint32_t)ry * BX + (uint32_t)rx;
        uint32_t keep = 0;
#pragma unroll
        for (uint32_t o = 0; o < 27; ++o) {
            const uint32_t kb = kb0 + (o / 9u) * BXY + ((o / 3u) % 3u) * BX + (o % 3u);
            const float4 lo = boxLo[kb], hi = boxHi[kb];
            const float gx = fmaxf(fmaxf(lo.x - px, px - hi.x), 0.0f);
            const float gy = fmaxf(fmaxf(lo.y - py, py - hi.y), 0.0f);
            const float gz = fmaxf(fmaxf(lo.z - pz, pz - hi.z), 0.0f);
            const float g2 = (gx * gx + gy * gy) + gz * gz;
            // an empty cell has an infinite box: g2 = inf
            keep |= (uint32_t)(g2 < k.h2) << o;
        }
        float density = 0.0f;
        uint32_t listed = 0, flushed = 0;
        // cursor: the kept bucket being walked (first slot, member count) and the position in it
        uint32_t cnt = 0, start = 0, j = 0;
        auto open = [&](uint32_t o) {
            const uint32_t dz = (o * 57u) >> 9, r = o - 9u * dz, dy = (r * 11u) >> 5, dx = r - 3u * dy;
            const uint32_t kb = kb0 + dz * BXY + dy * BX + dx;
            start = __float_as_uint(boxLo[kb].w);
            cnt = __float_as_uint(boxHi[kb].w);
            j = 0;
        };
        if (keep) open((uint32_t)__builtin_ctz(keep));
        while (keep) {
            const uint32_t m = density_batch<8>(k, px, py, pz, pos + start + j, cnt - j, density);
            list_push<8>(ring, m, start + j, cap, listed, flushed, nbrList, L.total, t);
            j += 8u;
            if (j >= cnt) {
                keep &= keep - 1u;
                if (keep) open((uint32_t)__builtin_ctz(keep));
            }
        }
        density_finish(k, density, pos, vel, t, cap, listed, flushed, ring, nbrList, nbrCount, L.total, L.rangeGuard);
    }
}

// ---- boundary pieces (src/physics/force_field.h:31-70, collider_shapes.h:33-131) ----

struct Surface {
    float p[3];
    float n[3];
    float dist;
};

__device__ inline void sphere_local(float radius, const float p[3], Surface& s) {
    float d = sqrtf((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
    if (d > 1.0e-6f) {
        s.n[0] = p[0] / d; s.n[1] = p[1] / d; s.n[2] = p[2] / d;
    } else {
        s.n[0] = 0.0f; s.n[1] = 1.0f; s.n[2] = 0.0f;
    }
    s.p[0] = s.n[0] * radius; s.p[1] = s.n[1] * radius; s.p[2] = s.n[2] * radius;
    s.dist = d - radius;
}

__device__ inline float clampf_ref(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

__device__ void closest_local(uint32_t shape, float radius, const float* half, const float p[3], Surface& s) {
    switch (shape) {
    case SPH_SHAPE_SPHERE: sphere_local(radius, p, s); return;
    case SPH_SHAPE_BOX: {
        float cl[3], dl[3];
        for (int a = 0; a < 3; ++a) {
            float lo = -half[a];
            float mx = p[a] < lo ? lo : p[a];
            cl[a] = half[a] < mx ? half[a] : mx;
            dl[a] = p[a] - cl[a];
        }
        float dsq = (dl[0] * dl[0] + dl[1] * dl[1]) + dl[2] * dl[2];
        if (dsq > 1.0e-10f) {
            float dist = sqrtf(dsq);
            s.n[0] = dl[0] / dist; s.n[1] = dl[1] / dist; s.n[2] = dl[2] / dist;
            s.p[0] = cl[0]; s.p[1] = cl[1]; s.p[2] = cl[2];
            s.dist = dist;
            return;
        }
        float gap[3];
        for (int a = 0; a < 3; ++a) gap[a] = half[a] - (p[a] < 0.0f ? -p[a] : p[a]);
        int axis = gap[0] < gap[1] ? (gap[0] < gap[2] ? 0 : 2) : (gap[1] < gap[2] ? 1 : 2);
        s.n[0] = s.n[1] = s.n[2] = 0.0f;
        s.n[axis] = p[axis] >= 0.0f ? 1.0f : -1.0f;
        s.p[0] = p[0]; s.p[1] = p[1]; s.p[2] = p[2];
        s.p[axis] = s.n[axis] * half[axis];
        s.dist = -gap[axis];
        return;
    }
    case SPH_SHAPE_PLANE:
        s.p[0] = p[0]; s.p[1] = 0.0f; s.p[2] = p[2];
        s.n[0] = 0.0f; s.n[1] = 1.0f; s.n[2] = 0.0f;
        s.dist = p[1];
        return;
    case SPH_SHAPE_CYLINDER: {
        float hh = half[1];
        float d = sqrtf((p[0] * p[0] + 0.0f * 0.0f) + p[2] * p[2]);
        float rd[3];
        if (d > 1.0e-6f) {
            rd[0] = p[0] / d; rd[1] = 0.0f / d; rd[2] = p[2] / d;
        } else {
            rd[0] = 1.0f; rd[1] = 0.0f; rd[2] = 0.0f;
        }
        float ay = p[1] < 0.0f ? -p[1] : p[1];
        if (ay <= hh && d <= radius) {
            float gapY = hh - ay, gapR = radius - d;
            if (gapY < gapR) {
                float side = p[1] >= 0.0f ? 1.0f : -1.0f;
                s.n[0] = 0.0f; s.n[1] = side; s.n[2] = 0.0f;
                s.p[0] = p[0]; s.p[1] = side * hh; s.p[2] = p[2];
                s.dist = -gapY;
            } else {
                s.n[0] = rd[0]; s.n[1] = rd[1]; s.n[2] = rd[2];
                s.p[0] = rd[0] * radius + 0.0f; s.p[1] = rd[1] * radius + p[1]; s.p[2] = rd[2] * radius + 0.0f;
                s.dist = -gapR;
            }
            return;
        }
        float m = d < radius ? d : radius;
        float cl[3] = {rd[0] * m, rd[1] * m, rd[2] * m};
        cl[1] = clampf_ref(p[1], -hh, hh);
        float dl[3] = {p[0] - cl[0], p[1] - cl[1], p[2] - cl[2]};
        float dist = sqrtf((dl[0] * dl[0] + dl[1] * dl[1]) + dl[2] * dl[2]);
        s.p[0] = cl[0]; s.p[1] = cl[1]; s.p[2] = cl[2];
        if (dist > 1.0e-6f) {
            s.n[0] = dl[0] / dist; s.n[1] = dl[1] / dist; s.n[2] = dl[2] / dist;
        } else {
            s.n[0] = 0.0f; s.n[1] = 1.0f; s.n[2] = 0.0f;
        }
        s.dist = dist;
        return;
    }
    case SPH_SHAPE_CAPSULE: {
        float hh = half[1];
        float seg = clampf_ref(p[1], -hh, hh);
        float q[3] = {p[0] - 0.0f, p[1] - seg, p[2] - 0.0f};
        sphere_local(radius, q, s);
        s.p[0] += 0.0f; s.p[1] += seg; s.p[2] += 0.0f;
        return;
    }
    default: {
        sphere_local(0.0f, p, s);
        return;
    }
    }
}

__device__ void collide(const SphCollider& c, float radius, float restitution, float pp[3], float vv[3]) {
    float loc[3];
    for (int r = 0; r < 3; ++r) {
        float a0 = c.inverseWorld[0 + r] * pp[0] + c.inverseWorld[4 + r] * pp[1];
        float a1 = c.inverseWorld[8 + r] * pp[2] + c.inverseWorld[12 + r] * 1.0f;
        loc[r] = a0 + a1;
    }
    Surface s;
    closest_local(c.shape, c.radius, c.halfExtents, loc, s);
    if (!(s.dist < radius)) return;
    float nw[3];
    for (int r = 0; r < 3; ++r) nw[r] = (c.world[0 + r] * s.n[0] + c.world[4 + r] * s.n[1]) + c.world[8 + r] * s.n[2];
    float inv = 1.0f / sqrtf((nw[0] * nw[0] + nw[1] * nw[1]) + nw[2] * nw[2]);
    float n[3] = {nw[0] * inv, nw[1] * inv, nw[2] * inv};
    for (int r = 0; r < 3; ++r) {
        float a0 = c.world[0 + r] * s.p[0] + c.world[4 + r] * s.p[1];
        float a1 = c.world[8 + r] * s.p[2] + c.world[12 + r] * 1.0f;
        pp[r] = (a0 + a1) + n[r] * radius;
    }
    float along = (vv[0] * n[0] + vv[1] * n[1]) + vv[2] * n[2];
    if (along < 0.0f) {
        float sc = (1.0f + restitution) * along;
        vv[0] -= sc * n[0]; vv[1] -= sc * n[1]; vv[2] -= sc * n[2];
    }
}

__device__ void field_one(const SphForceField& f, const float p[3], float out[3]) {
    float off[3] = {p[0] - f.position[0], p[1] - f.position[1], p[2] - f.position[2]};
    float distance = sqrtf((off[0] * off[0] + off[1] * off[1]) + off[2] * off[2]);
    float weight = 1.0f;
    if (f.radius > 0.0f) {
        float t = clampf_ref(1.0f - distance / f.radius, 0.0f, 1.0f);
        weight = t <= 0.0f ? 0.0f : powf(t, f.falloff);
    }
    out[0] = out[1] = out[2] = 0.0f;
    if (weight <= 0.0f) return;
    float sw = f.strength * weight;
    if (f.type == SPH_FIELD_WIND) {
        out[0] = f.axis[0] * sw; out[1] = f.axis[1] * sw; out[2] = f.axis[2] * sw;
    } else if (f.type == SPH_FIELD_VORTEX) {
        float along = (off[0] * f.axis[0] + off[1] * f.axis[1]) + off[2] * f.axis[2];
        float rad[3] = {off[0] - f.axis[0] * along, off[1] - f.axis[1] * along, off[2] - f.axis[2] * along};
        float r = sqrtf((rad[0] * rad[0] + rad[1] * rad[1]) + rad[2] * rad[2]);
        if (r < 1.0e-4f) return;
        float u[3] = {rad[0] / r, rad[1] / r, rad[2] / r};
        out[0] = (f.axis[1] * u[2] - u[1] * f.axis[2]) * sw;
        out[1] = (f.axis[2] * u[0] - u[2] * f.axis[0]) * sw;
        out[2] = (f.axis[0] * u[1] - u[0] * f.axis[1]) * sw;
    } else if (f.type == SPH_FIELD_POINT) {
        if (distance < 1.0e-4f) return;
        out[0] = (-off[0] / distance) * sw;
        out[1] = (-off[1] / distance) * sw;
        out[2] = (-off[2] / distance) * sw;
    }
}

// Pressure + viscosity contribution of neighbour slot u to particle (p, v, pressure),
// accumulated into f in the reference's order (fluid_sph.cpp:244-254).
// SHORT: the square root and the seven divisions by sqrt_rn / recip_rn / div_rn (sph_exact.h),
// one reciprocal per divisor, for launches whose operands all lie in the ranges where those
// give the IEEE results (the caller checks the scene's shortDivide and the range guard); the
// results are the same bits, from ~100 instead of ~146 instructions per pair.
// One pair's contribution to the force on particle (px, py, pz): the pressure term a and the
// viscosity term b, each per component; false when the pair is out of range. The reference adds
// them to the force as f += a, then f += b (pair_force_from).
template <bool SHORT>
__device__ __forceinline__ bool pair_terms(const StepConsts& k, const float4 q, const float4 ov, float px, float py,
                                           float pz, float vx, float vy, float vz, float pressure, float a[3],
                                           float b[3]) {
    const float ddx = px - q.x, ddy = py - q.y, ddz = pz - q.z;
    const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
    // r2 >= r2Cut <=> sqrtf(r2) >= h (both correctly rounded, monotone): the
    // reference's `distance >= h` test without the square root.
    if (r2 >= k.r2Cut) return false;
    const float distance = SHORT ? sqrt_rn(r2) : sqrtf(r2);
    if (distance >= k.h || distance <= 1e-6f) return false;
    const float otherDensity = q.w < 1e-6f ? 1e-6f : q.w;
    // spikyGradient (fluid_sph.cpp:33-39): ((c d) d) (delta / length)
    const float d = k.h - distance;
    const float g = (k.cSpiky * d) * d;
    const float lap = k.cVisc * (k.h - distance);
    if (SHORT) {
        // a / (2 rho_j) is RN(a / rho_j) / 2: halving is exact for the normal quotients here
        const float yd = recip_rn(distance), yr = recip_rn(otherDensity);
        const float gx = g * div_rn(ddx, distance, yd), gy = g * div_rn(ddy, distance, yd),
                    gz = g * div_rn(ddz, distance, yd);
        const float sp = div_rn(k.negMass * (pressure + ov.w), otherDensity, yr) * 0.5f;
        a[0] = sp * gx;
        a[1] = sp * gy;
        a[2] = sp * gz;
        b[0] = div_rn(k.viscMass * (ov.x - vx), otherDensity, yr) * lap;
        b[1] = div_rn(k.viscMass * (ov.y - vy), otherDensity, yr) * lap;
        b[2] = div_rn(k.viscMass * (ov.z - vz), otherDensity, yr) * lap;
        return true;
    }
    const float gx = g * (ddx / distance), gy = g * (ddy / distance), gz = g * (ddz / distance);
    // -mass (p_i + p_j) / (2 rho_j) * grad (fluid_sph.cpp:251-252)
    const float sp = (k.negMass * (pressure + ov.w)) / (2.0f * otherDensity);
    a[0] = sp * gx;
    a[1] = sp * gy;
    a[2] = sp * gz;
    // viscosity mass (v_j - v_i) / rho_j * laplacian (fluid_sph.cpp:253-254)
    b[0] = ((k.viscMass * (ov.x - vx)) / otherDensity) * lap;
    b[1] = ((k.viscMass * (ov.y - vy)) / otherDensity) * lap;
    b[2] = ((k.viscMass * (ov.z - vz)) / otherDensity) * lap;
    return true;
}

template <bool SHORT>
__device__ __forceinline__ void pair_force_from(const StepConsts& k, const float4 q, const float4 ov, float px,
                                                float py, float pz, float vx, float vy, float vz, float pressure,
                                                float f[3]) {
    float a[3], b[3];
    if (!pair_terms<SHORT>(k, q, ov, px, py, pz, vx, vy, vz, pressure, a, b)) return;
    f[0] += a[0];
    f[1] += a[1];
    f[2] += a[2];
    f[0] += b[0];
    f[1] += b[1];
    f[2] += b[2];
}

// The same contribution for scenes with SPH_MODE_FAST_MATH: r2 by fused multiply-adds (the
// expression the fast density pass uses, so its neighbour list stays a superset), 1 / r from
// the hardware reciprocal square root, 1 / rho_j from the hardware reciprocal, products
// accumulated with fused multiply-adds. Not bit-exact: within the 1e-5 tolerance (relative to
// the largest force) of the reference after one substep.
__device__ __forceinline__ void pair_force_fast(const StepConsts& k, const float4 q, const float4 ov, float px,
                                                float py, float pz, float vx, float vy, float vz, float pressure,
                                                float f[3]) {
    const float ddx = px - q.x, ddy = py - q.y, ddz = pz - q.z;
    const float r2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
    if (r2 >= k.r2Cut || !(r2 > 0.0f)) return;
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float distance = r2 * inv;
    if (distance >= k.h || distance <= 1e-6f) return;
    const float otherDensity = q.w < 1e-6f ? 1e-6f : q.w;
    const float irho = __builtin_amdgcn_rcpf(otherDensity);
    const float d = k.h - distance;
    // -m (p_i + p_j) / (2 rho_j) * cSpiky d^2 / r * delta
    const float sp = ((k.negMass * (pressure + ov.w)) * (0.5f * irho)) * (((k.cSpiky * d) * d) * inv);
    f[0] = fmaf(sp, ddx, f[0]);
    f[1] = fmaf(sp, ddy, f[1]);
    f[2] = fmaf(sp, ddz, f[2]);
    // mu m (v_j - v_i) / rho_j * cVisc (h - r)
    const float vl = (k.viscMass * irho) * (k.cVisc * d);
    f[0] = fmaf(vl, ov.x - vx, f[0]);
    f[1] = fmaf(vl, ov.y - vy, f[1]);
    f[2] = fmaf(vl, ov.z - vz, f[2]);
}

template <bool FAST>
__device__ __forceinline__ void pair_force(const StepConsts& k, bool fast, const float4* __restrict__ pos,
                                           const float4* __restrict__ vel, uint32_t u, float px, float py, float pz,
                                           float vx, float vy, float vz, float pressure, float f[3]) {
    if (FAST && fast)
        pair_force_fast(k, pos[u], vel[u], px, py, pz, vx, vy, vz, pressure, f);
    else
        pair_force_from<false>(k, pos[u], vel[u], px, py, pz, vx, vy, vz, pressure, f);
}

// 64 VGPRs at most (8 waves per SIMD): this VALU-bound pass runs ~4 % faster at 8 waves
// than at the 6 its natural register count allows, with no spills.
// No list (or it overflowed): the force on particle t from its 27 buckets, walked as the
// reference walks them (fluid_sph.cpp:234-258).
template <bool FAST>
__device__ __forceinline__ void bucket_walk_force(const StepConsts& k, bool fast, uint32_t cellBase,
                                                  const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                  const uint2* __restrict__ buckets, uint32_t t, float px, float py,
                                                  float pz, float vx, float vy, float vz, float pressure,
                                                  float f[3]) {
    const int32_t cx = cell_coord(px, k.h), cy = cell_coord(py, k.h), cz = cell_coord(pz, k.h);
    const bool capped = (k.mode & SPH_MODE_UNCAPPED) == 0;
    const bool dedupe = (k.mode & SPH_MODE_DEDUPE) != 0;
    uint32_t o = 0;
    for (int32_t dz = -1; dz <= 1; ++dz) {
        for (int32_t dy = -1; dy <= 1; ++dy) {
            for (int32_t dx = -1; dx <= 1; ++dx, ++o) {
                const uint32_t b = hash_cell(cx + dx, cy + dy, cz + dz, k.cellMask);
                if (dedupe && seen_before(cx, cy, cz, o, b, k.cellMask)) continue;
                const uint2 bd = buckets[cellBase + b];
                uint32_t cnt = bd.y;
                if (cnt == 0) continue;
                const uint32_t lo = bd.x;
                if (capped && cnt > SPH_CELL_CAPACITY) cnt = SPH_CELL_CAPACITY;
                for (uint32_t j = 0; j < cnt; ++j) {
                    const uint32_t u = lo + j;
                    if (u == t) continue;
                    pair_force<FAST>(k, fast, pos, vel, u, px, py, pz, vx, vy, vz, pressure, f);
                }
            }
        }
    }
}

// The integration of particle t from its pressure + viscosity force f: acceleration, symplectic
// Euler, speed clamp, container walls and colliders (fluid_sph.cpp:260-300); writes its new
// state (and, when asked, the force).
__device__ __forceinline__ void integrate_write(const StepConsts& k, const BoundaryConsts& bnd, uint32_t t, float px,
                                                float py, float pz, float vx, float vy, float vz, float density,
                                                float pressure, const float f[3], float4* __restrict__ posOut,
                                                float4* __restrict__ velOut, float4* __restrict__ forceOut) {
    const float fx = f[0], fy = f[1], fz = f[2];
    if (forceOut) forceOut[t] = make_float4(fx, fy, fz, 0.0f);

    // acceleration = force / density + gravity + fields (fluid_sph.cpp:260-261)
    float fa[3] = {0.0f, 0.0f, 0.0f};
    const float here[3] = {px, py, pz};
    for (uint32_t f = 0; f < k.fieldCount; ++f) {
        float a[3];
        field_one(bnd.fields[f], here, a);
        fa[0] += a[0]; fa[1] += a[1]; fa[2] += a[2];
    }
    const float ax = (fx / density + k.gravity[0]) + fa[0];
    const float ay = (fy / density + k.gravity[1]) + fa[1];
    const float az = (fz / density + k.gravity[2]) + fa[2];
    vx += ax * k.dt;
    vy += ay * k.dt;
    vz += az * k.dt;
    // speed clamp (fluid_sph.cpp:264-268)
    const float speed = sqrtf((vx * vx + vy * vy) + vz * vz);
    if (speed > k.maxSpeed) {
        const float sc = k.maxSpeed / speed;
        vx *= sc;
        vy *= sc;
        vz *= sc;
    }
    px += vx * k.dt;
    py += vy * k.dt;
    pz += vz * k.dt;
    // container walls (fluid_sph.cpp:271-285)
    float pp[3] = {px, py, pz};
    float vv[3] = {vx, vy, vz};
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
    // colliders (fluid_sph.cpp:289-300)
    for (uint32_t c = 0; c < k.colliderCount; ++c) collide(bnd.colliders[c], k.radius, k.restitution, pp, vv);

    posOut[t] = make_float4(pp[0], pp[1], pp[2], density);
    velOut[t] = make_float4(vv[0], vv[1], vv[2], pressure);
}

template <bool FAST>
__global__ void __launch_bounds__(kBlock, FAST ? 5 : 8)
    k_force(Launch L, const float4* __restrict__ pos, const float4* __restrict__ vel,
            const uint2* __restrict__ buckets, const uint4* __restrict__ nbrList, const uint32_t* __restrict__ nbrCount, float4* __restrict__ posOut,
            float4* __restrict__ velOut, float4* __restrict__ forceOut) {
    const Where w = locate(L);
    if (!w.valid) return;
    const uint32_t t = w.slot;
    if (!w.d->active) {
        // Past this scene's substep count: the state passes through unchanged.
        posOut[t] = pos[t];
        velOut[t] = vel[t];
        return;
    }
    const StepConsts k = w.d->k;   // by value: the constants stay in scalar registers
    const uint32_t cellBase = w.d->cellOffset;
    const float4 self = pos[t];
    const float4 selfVel = vel[t];
    float px = self.x, py = self.y, pz = self.z;
    const float density = self.w < 1e-6f ? 1e-6f : self.w;
    float vx = selfVel.x, vy = selfVel.y, vz = selfVel.z;
    const float pressure = selfVel.w;
    const bool fast = FAST && (k.mode & SPH_MODE_FAST_MATH) != 0;
    float f[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t listed = nbrCount ? nbrCount[t] : k.listCap + 1u;
    if (FAST && fast && listed <= k.listCap) {
        // Fast arithmetic leaves this loop bound by memory latency: eight pairs in flight.
        for (uint32_t e = 0; e < listed; e += 8) {
            const uint4 g0 = load_stream(nbrList + (size_t)(e >> 2) * L.total + t);
            const uint4 g1 = e + 4 < listed ? load_stream(nbrList + (size_t)((e >> 2) + 1) * L.total + t) : g0;
            const uint32_t us[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
            float4 q[8], ov[8];
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r) {
                const uint32_t u = e + r < listed ? us[r] : us[0];
                q[r] = pos[u];
                ov[r] = vel[u];
            }
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r) {
                if (e + r < listed && us[r] != t)
                    pair_force_fast(k, q[r], ov[r], px, py, pz, vx, vy, vz, pressure, f);
            }
        }
    } else if (listed <= k.listCap) {
        // The in-range candidates recorded by the density pass, in visiting order.
        auto walk = [&](auto shortTag) {
            constexpr bool SHORT = decltype(shortTag)::value;
            for (uint32_t e = 0; e < listed; e += 4) {
                const uint4 g = load_stream(nbrList + (size_t)(e >> 2) * L.total + t);
                const uint32_t us[4] = {g.x, g.y, g.z, g.w};
                // All four neighbours' state in flight before the first is consumed; slots
                // past the end of the list re-read the last valid one and are not used.
                float4 q[4], ov[4];
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) {
                    const uint32_t u = e + r < listed ? us[r] : us[0];
                    q[r] = pos[u];
                    ov[r] = vel[u];
                }
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) {
                    if (e + r < listed && us[r] != t)
                        pair_force_from<SHORT>(k, q[r], ov[r], px, py, pz, vx, vy, vz, pressure, f);
                }
            }
        };
        // The short division when the scene's constants and every particle's state of this
        // substep lie in its exact ranges (the same for the whole launch).
        if (k.shortDivide && L.rangeGuard && *L.rangeGuard == 0u)
            walk(std::true_type{});
        else
            walk(std::false_type{});
    } else {
        // No list (or it overflowed): walk the 27 buckets as the reference does.
        bucket_walk_force<FAST>(k, fast, cellBase, pos, vel, buckets, t, px, py, pz, vx, vy, vz, pressure, f);
    }
    const BoundaryConsts& bnd = L.bounds[L.blockScene[xcd_block(blockIdx.x, gridDim.x)]];
    integrate_write(k, bnd, t, px, py, pz, vx, vy, vz, density, pressure, f, posOut, velOut, forceOut);
}

// ---- small-scene density pass ----
//
// A scene of a few thousand particles gives the lockstep pass one wave per CU or fewer, and
// each lane's walk over its 27 buckets is then pure memory latency (16,384 particles: 46 us
// for 64 workgroups). Here one wave takes one particle: lane o < 27 walks neighbour bucket o
// (its members in order, with the lockstep pass's operations) and stages its in-range terms
// and slots in LDS; then lane 0 adds the terms bucket by bucket in the reference's visiting
// order (out-of-range candidates, which add +0 there, are simply absent: the sum is the same
// bits) and applies the EOS, and every lane writes its slots to their places in the
// neighbour list.
constexpr uint32_t kSmallWave = 64;   // threads per workgroup: one particle

__global__ void __launch_bounds__(kSmallWave)
    k_density_small(Launch L, float4* __restrict__ pos, float4* __restrict__ vel, const uint2* __restrict__ buckets,
                    uint4* __restrict__ nbrList, uint32_t* __restrict__ nbrCount) {
    __shared__ float sTerm[27 * SPH_CELL_CAPACITY];
    __shared__ uint32_t sSlot[27 * SPH_CELL_CAPACITY];
    const SceneDesc* d = L.scenes;   // single-scene launches only
    const uint32_t local = blockIdx.x;
    if (!d->active || local >= d->k.n) return;   // the same for the whole workgroup
    const StepConsts& k = d->k;
    const uint32_t t = d->offset + local, o = threadIdx.x;
    const float* self = reinterpret_cast<const float*>(pos + t);
    const float px = self[0], py = self[1], pz = self[2];
    uint32_t m = 0;
    if (o < 27u) {
        const int32_t cx = cell_coord(px, k.h), cy = cell_coord(py, k.h), cz = cell_coord(pz, k.h);
        // visiting order: dz outer, then dy, dx inner (fluid_sph.cpp:205-216)
        const int32_t dx = (int32_t)(o % 3u) - 1, dy = (int32_t)((o / 3u) % 3u) - 1, dz = (int32_t)(o / 9u) - 1;
        const uint2 bd = buckets[d->cellOffset + hash_cell(cx + dx, cy + dy, cz + dz, k.cellMask)];
        const uint32_t cnt = bd.y < SPH_CELL_CAPACITY ? bd.y : SPH_CELL_CAPACITY;
        // eight candidates per batch from one base address (the last batch may read up to
        // seven slots past the bucket, into the next bucket or the padding, and ignores them)
        static_assert(7u <= kPosPad, "reads past a bucket stay inside the padding");
        for (uint32_t j = 0; j < cnt; j += 8u) {
            float qx[8], qy[8], qz[8];
#pragma unroll
            for (uint32_t e = 0; e < 8u; ++e) {
                const float* q = reinterpret_cast<const float*>(pos + bd.x + j + e);
                qx[e] = q[0];
                qy[e] = q[1];
                qz[e] = q[2];
            }
#pragma unroll
            for (uint32_t e = 0; e < 8u; ++e) {
                const float ddx = px - qx[e], ddy = py - qy[e], ddz = pz - qz[e];
                const float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                if (j + e < cnt && r2 < k.h2) {
                    const float dd = k.h2 - r2;
                    sTerm[o * SPH_CELL_CAPACITY + m] = k.mass * (((k.cPoly6 * dd) * dd) * dd);
                    sSlot[o * SPH_CELL_CAPACITY + m] = bd.x + j + e;
                    ++m;
                }
            }
        }
    }
    // each bucket's place in the particle's in-range sequence: exclusive scan of m over lanes
    uint32_t before = m;
#pragma unroll
    for (uint32_t w = 1; w < 32u; w <<= 1) {
        const uint32_t up = __shfl_up(before, w, 64);
        if (o >= w) before += up;
    }
    before -= m;
    const uint32_t listed = __shfl(before + m, 26, 64);
    __syncthreads();
    const uint32_t cap = nbrList ? k.listCap : 0u;
    if (nbrCount && listed <= cap) {
        // this bucket's slots, then zeros after the last entry up to its group's end
        uint32_t* list = reinterpret_cast<uint32_t*>(nbrList);
        for (uint32_t e = 0; e < m; ++e) {
            const uint32_t at = before + e;
            list[((size_t)(at >> 2) * L.total + t) * 4u + (at & 3u)] = sSlot[o * SPH_CELL_CAPACITY + e];
        }
        if (o < ((4u - (listed & 3u)) & 3u)) {
            const uint32_t at = listed + o;
            list[((size_t)(at >> 2) * L.total + t) * 4u + (at & 3u)] = 0u;
        }
    } else if (nbrCount && cap) {
        // past the cap: the complete groups of four within the first cap entries are kept, as
        // the lockstep pass keeps them (the force pass then walks the buckets itself)
        uint32_t* list = reinterpret_cast<uint32_t*>(nbrList);
        for (uint32_t e = 0; e < m && before + e < (cap & ~3u); ++e) {
            const uint32_t at = before + e;
            list[((size_t)(at >> 2) * L.total + t) * 4u + (at & 3u)] = sSlot[o * SPH_CELL_CAPACITY + e];
        }
    }
    float density = 0.0f;
    if (o == 0u) {
        for (uint32_t b = 0; b < 27u; ++b) {
            const uint32_t c = __builtin_amdgcn_readlane(m, b);
            for (uint32_t e = 0; e < c; ++e) density = density + sTerm[b * SPH_CELL_CAPACITY + e];
        }
    }
    density = density < 1e-6f ? 1e-6f : density;
    const float pressure = eos_pressure(k, density);
    raise_range_guard(L.rangeGuard, o == 0u && !(density_in_range(density) && pressure_in_range(pressure)));
    if (o == 0u) {
        reinterpret_cast<float*>(pos + t)[3] = density;
        reinterpret_cast<float*>(vel + t)[3] = pressure;
        if (nbrCount) nbrCount[t] = cap && listed <= cap ? listed : cap + 1u;
    }
}

// The force pass for the same small single-scene launches: one wave per particle, lane e < 64
// computing the contributions of list entry e, lane 0 adding them in list order (each pair's
// pressure term, then its viscosity term, as pair_force_from adds them) and integrating.
// Particles whose list overflowed walk their buckets on lane 0 as in k_force.
__global__ void __launch_bounds__(kSmallWave)
    k_force_small(Launch L, const float4* __restrict__ pos, const float4* __restrict__ vel,
                  const uint2* __restrict__ buckets, const uint4* __restrict__ nbrList,
                  const uint32_t* __restrict__ nbrCount, float4* __restrict__ posOut, float4* __restrict__ velOut,
                  float4* __restrict__ forceOut) {
    static_assert(kSmallWave >= 64u, "one lane per list entry of the 64-entry list");
    __shared__ float sA[64][3], sB[64][3];
    __shared__ uint32_t sIn[64];
    const SceneDesc* d = L.scenes;   // single-scene launches only
    const uint32_t local = blockIdx.x;
    if (local >= d->k.n) return;     // the same for the whole workgroup
    const uint32_t t = d->offset + local, e = threadIdx.x;
    if (!d->active) {
        if (e == 0u) {
            posOut[t] = pos[t];
            velOut[t] = vel[t];
        }
        return;
    }
    const StepConsts& k = d->k;
    const float4 self = pos[t];
    const float4 selfVel = vel[t];
    const float px = self.x, py = self.y, pz = self.z;
    const float density = self.w < 1e-6f ? 1e-6f : self.w;
    const float vx = selfVel.x, vy = selfVel.y, vz = selfVel.z;
    const float pressure = selfVel.w;
    const uint32_t listed = nbrCount[t];
    float f[3] = {0.0f, 0.0f, 0.0f};
    if (listed <= k.listCap && listed <= 64u) {
        const bool shortDiv = k.shortDivide && L.rangeGuard && *L.rangeGuard == 0u;
        uint32_t in = 0u;
        if (e < listed) {
            const uint32_t u = reinterpret_cast<const uint32_t*>(nbrList)[((size_t)(e >> 2) * L.total + t) * 4u + (e & 3u)];
            if (u != t) {
                float a[3], b[3];
                const float4 q = pos[u], ov = vel[u];
                in = shortDiv ? pair_terms<true>(k, q, ov, px, py, pz, vx, vy, vz, pressure, a, b)
                              : pair_terms<false>(k, q, ov, px, py, pz, vx, vy, vz, pressure, a, b);
                if (in) {
                    sA[e][0] = a[0]; sA[e][1] = a[1]; sA[e][2] = a[2];
                    sB[e][0] = b[0]; sB[e][1] = b[1]; sB[e][2] = b[2];
                }
            }
        }
        sIn[e] = in;
        __syncthreads();
        if (e != 0u) return;
        for (uint32_t r = 0; r < listed; ++r) {
            if (!sIn[r]) continue;
            f[0] += sA[r][0];
            f[1] += sA[r][1];
            f[2] += sA[r][2];
            f[0] += sB[r][0];
            f[1] += sB[r][1];
            f[2] += sB[r][2];
        }
    } else {
        if (e != 0u) return;
        bucket_walk_force<false>(k, false, d->cellOffset, pos, vel, buckets, t, px, py, pz, vx, vy, vz, pressure, f);
    }
    integrate_write(k, L.bounds[0], t, px, py, pz, vx, vy, vz, density, pressure, f, posOut, velOut, forceOut);
}

}  // namespace

hipError_t launch_density(const Launch& L, float4* pos, float4* vel, const uint2* buckets, uint4* nbrList,
                          uint32_t* nbrCount,
                          const NeighbourSums* sums, hipStream_t s, const DensityTiles* tiles) {
    if (L.blocks == 0) return hipSuccess;
    const bool dedupe = (L.modes & SPH_MODE_DEDUPE) != 0;
    const uint32_t plainModes = SPH_MODE_DEDUPE | SPH_MODE_UNCAPPED | SPH_MODE_FAST_MATH;
    if (!sums && !tiles && nbrList && nbrCount && L.numScenes == 1 && !(L.modes & plainModes) &&
        L.blocks * kBlock <= kSmallDensityMax) {
        // one workgroup per particle slot of the scene's workgroups (the kernel skips those past n)
        k_density_small<<<L.blocks * kBlock, kSmallWave, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount);
        return hipGetLastError();
    }
    if (sums) {
        if (dedupe)
            k_density<true, true><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nullptr, nullptr,
                                                              sums->sums, sums->deviation);
        else
            k_density<true, false><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nullptr, nullptr,
                                                               sums->sums, sums->deviation);
    } else if (tiles && !dedupe && !(L.modes & SPH_MODE_FAST_MATH) && L.tiles > 0 && L.tileScene) {
        k_density_tiled<<<L.tiles + tiles->restBlocks, kBlock, 0, s>>>(L, tiles->binStart, pos, vel,
                                                                       buckets, nbrList, nbrCount,
                                                                       tiles->rest, tiles->restCount);
    } else {
        const bool fast = (L.modes & SPH_MODE_FAST_MATH) != 0;
        if (dedupe && fast)
            k_density<false, true, true><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList,
                                                                     nbrCount, nullptr, nullptr);
        else if (dedupe)
            k_density<false, true><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount,
                                                               nullptr, nullptr);
        else if (fast)
            k_density<false, false, true><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList,
                                                                      nbrCount, nullptr, nullptr);
        else
            k_density<false, false><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount,
                                                                nullptr, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_force(const Launch& L, const float4* pos, const float4* vel, const uint2* buckets,
                        const uint4* nbrList, const uint32_t* nbrCount, float4* posOut,
                        float4* velOut, float4* forceOut, hipStream_t s) {
    if (L.blocks == 0) return hipSuccess;
    if (nbrList && nbrCount && L.numScenes == 1 && !(L.modes & SPH_MODE_FAST_MATH) &&
        L.blocks * kBlock <= kSmallDensityMax) {
        k_force_small<<<L.blocks * kBlock, kSmallWave, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount, posOut,
                                                               velOut, forceOut);
        return hipGetLastError();
    }
    if (L.modes & SPH_MODE_FAST_MATH)
        k_force<true><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount, posOut, velOut,
                                                  forceOut);
    else
        k_force<false><<<L.blocks, kBlock, 0, s>>>(L, pos, vel, buckets, nbrList, nbrCount, posOut, velOut,
                                                   forceOut);
    return hipGetLastError();
}

}  // namespace sphdev
