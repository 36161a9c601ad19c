// sph_engine.hip — host engine: device arrays for S scenes, per-substep scene tables,
// the launch sequence of one substep (src/physics/fluid_sph.cpp:190-309), and hipGraph
// capture of a whole step (src/physics/fluid_sph.cpp:311-321).
//
// Per substep, for all scenes at once:
//   clear + fillers -> keys -> bin slots -> scan -> scatter -> rank + reorder  (FluidGrid::build)
//   density (in place .w, neighbour list)                               (fluid_sph.cpp:197-222)
//   force + integrate + walls + colliders -> back buffer                (fluid_sph.cpp:224-305)
// pos/vel return to buffer `pc` after every substep; the id buffer flips, so a step is
// captured twice, once per starting id parity.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "sph_engine.h"
#include "sph_internal.h"

namespace sphhost {

namespace {

constexpr float kPi = 3.14159265358979323846f;   // std::numbers::pi_v<float>

const char* const kKernelNames[KC_COUNT] = {"grid_clear", "keys", "bin_slots", "scan", "scatter", "rank_reorder",
                                            "density", "force_integrate"};

bool is_pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

// The smallest float x with sqrtf(x) >= h. sqrtf is correctly rounded and monotone on
// host and device alike, so `r2 >= cut` decides `sqrtf(r2) >= h` exactly.
float sqrt_cut(float h) {
    float x = h * h;
    if (sqrtf(x) >= h) {
        for (;;) {
            float down = nextafterf(x, 0.0f);
            if (!(sqrtf(down) >= h)) break;
            x = down;
        }
        return x;
    }
    while (!(sqrtf(x) >= h)) x = nextafterf(x, INFINITY);
    return x;
}

// Keep libm's powf out of reach of constant folding so host and oracle share glibc's value.
float libm_powf(float x, float y) {
    float (*volatile fn)(float, float) = powf;
    return fn(x, y);
}

uint32_t blocks_for(uint32_t n) { return (n + sphdev::kBlock - 1) / sphdev::kBlock; }

// Workgroups of the tiled density launch that walk the particles the grid build listed
// (a multiple of the 8 XCDs, so the half-tiles after them keep their XCD mapping).
constexpr uint32_t kRestBlocks = 256;

}  // namespace

int hip_fail(hipError_t e, const char* what) {
    return fail(SPH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

sphdev::StepConsts make_consts(const SceneHost& s, const SphParams& p, float dt, bool hasList) {
    sphdev::StepConsts k;
    memset(&k, 0, sizeof(k));
    const float h = p.smoothingRadius;
    k.n = s.n;
    k.cellMask = p.cellCount - 1u;
    k.mode = s.mode;
    k.eos = s.eos;
    k.h = h;
    k.h2 = h * h;
    k.r2Cut = sqrt_cut(h);
    k.mass = p.particleMass;
    k.negMass = -p.particleMass;
    k.viscMass = p.viscosity * p.particleMass;
    // Kernel constants in the reference's expression order (fluid_sph.cpp:30, :38, :45).
    k.cPoly6 = 315.0f / (64.0f * kPi * libm_powf(h, 9.0f));
    k.cSpiky = -45.0f / (kPi * libm_powf(h, 6.0f));
    k.cVisc = 45.0f / (kPi * libm_powf(h, 6.0f));
    k.restDensity = p.restDensity;
    k.stiffness = p.stiffness;
    k.taitB = (p.stiffness * p.restDensity) / 7.0f;
    k.dt = dt;
    k.maxSpeed = 0.4f * h / dt;
    for (int a = 0; a < 3; ++a) {
        k.low[a] = p.containerMin[a] + p.particleRadius;
        k.high[a] = p.containerMax[a] - p.particleRadius;
        k.gravity[a] = p.gravity[a];
    }
    k.radius = p.particleRadius;
    k.restitution = p.wallRestitution;
    k.colliderCount = p.colliderCount < SPH_MAX_COLLIDERS ? p.colliderCount : SPH_MAX_COLLIDERS;
    k.fieldCount = p.fieldCount < SPH_MAX_FORCE_FIELDS ? p.fieldCount : SPH_MAX_FORCE_FIELDS;
    // The force pass may reuse the density pass's in-range list only if its own cut-off
    // (r2 < r2Cut) selects a subset of the density's (r2 < h2).
    k.listCap = (hasList && s.list && k.r2Cut <= k.h2) ? 64u : 0u;
    // The force pass's short division is exact for these constants (sph_exact.h, and the
    // particle-state ranges of coord_in_range and the predicates beside it in sph_device.h).
    const float vm = k.viscMass;
    k.shortDivide = (h > 0.0f && h <= 0x1p26f && k.mass >= 0x1p-60f && k.mass <= 0x1p15f &&
                     (vm == 0.0f || (vm >= 0x1p-40f && vm <= 0x1p8f)))
                        ? 1u
                        : 0u;
    return k;
}

// Spatial bins over the container (plus one cell of margin each side). Any deterministic
// mapping is correct here; this one keeps neighbouring cells close in memory. If the
// container holds more cells than the scene's bins, bins are coarsened by powers of two.
sphdev::BinGrid make_bins(const SceneHost& s, const SphParams& p) {
    sphdev::BinGrid g;
    memset(&g, 0, sizeof(g));
    const double h = p.smoothingRadius;
    int64_t lo[3], span[3];
    for (int a = 0; a < 3; ++a) {
        double mn = p.containerMin[a] < p.containerMax[a] ? p.containerMin[a] : p.containerMax[a];
        double mx = p.containerMin[a] < p.containerMax[a] ? p.containerMax[a] : p.containerMin[a];
        double l = floor(mn / h) - 1.0, u = floor(mx / h) + 1.0;
        if (!(l > -1e9)) l = -1e9;
        if (!(u < 1e9)) u = 1e9;
        if (u < l) u = l;
        lo[a] = (int64_t)l;
        span[a] = (int64_t)u - (int64_t)l + 1;
    }
    for (int a = 0; a < 3; ++a) g.tileLog2[a] = s.tileLog2[a];
    uint32_t shift = 0;
    for (;;) {
        uint64_t tiles = 1;
        for (int a = 0; a < 3; ++a) {
            const uint64_t d = (uint64_t)((span[a] + (1ll << shift) - 1) >> shift);
            g.dims[a] = (uint32_t)d;
            g.tiles[a] = (uint32_t)((d + (1u << g.tileLog2[a]) - 1) >> g.tileLog2[a]);
            tiles *= g.tiles[a];
        }
        if (tiles * 64ull <= s.binCap || shift >= 31) {
            g.bins = (uint32_t)(tiles * 64ull);
            break;
        }
        ++shift;
    }
    g.shift = shift;
    for (int a = 0; a < 3; ++a) g.origin[a] = (int32_t)lo[a];
    return g;
}

hipEvent_t Engine::take_event() {
    if (!eventPool.empty()) {
        hipEvent_t e = eventPool.back();
        eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one kernel-class launch with events when profiling is on (eager mode only).
struct Timer {
    Engine* s;
    uint32_t cls;
    hipEvent_t b = nullptr;
    Timer(Engine* e, uint32_t c) : s(e), cls(c) {
        if (s->profiling) {
            b = s->take_event();
            if (b) hipEventRecord(b, s->stream);
        }
    }
    ~Timer() {
        if (s->profiling && b) {
            hipEvent_t e = s->take_event();
            if (e) {
                hipEventRecord(e, s->stream);
                s->record(cls, b, e);
            }
        }
    }
};

int Engine::set_device() const {
    SPH_HIP(hipSetDevice(device));
    return SPH_OK;
}

void Engine::free_all() {
    for (int i = 0; i < 2; ++i) {
        hipFree(pos[i]);
        hipFree(vel[i]);
        hipFree(ids[i]);
        if (graph[i]) hipGraphExecDestroy(graph[i]);
        graph[i] = nullptr;
    }
    hipFree(keys);
    hipFree(buckets);
    hipFree(rep);
    hipFree(binStart);
    hipFree(binOf);
    hipFree(slots);
    hipFree(blockSums);
    hipFree(keysSorted);
    hipFree(pairKI);
    hipFree(pairSrc);
    hipFree(nbrList);
    hipFree(nbrCount);
    hipFree(prev);
    hipFree(force);
    hipFree(staging);
    hipFree(stagingU);
    hipFree(dTables);
    hipFree(dBounds);
    hipFree(dBlockScene);
    hipFree(dTileScene);
    hipFree(restList);
    hipFree(restCount);
    hipFree(binMax);
    binMax = nullptr;
    hipFree(rangeGuard);
    rangeGuard = nullptr;
    if (binMaxHost) hipHostFree(binMaxHost);
    binMaxHost = binMaxHostDev = nullptr;
    if (hostStage) hipHostFree(hostStage);
    hostStage = nullptr;
    for (auto& t : timed) {
        hipEventDestroy(t.begin);
        hipEventDestroy(t.end);
    }
    timed.clear();
    for (hipEvent_t e : eventPool) hipEventDestroy(e);
    eventPool.clear();
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
}

Engine::~Engine() {
    if (stream) {
        hipSetDevice(device);
        hipStreamSynchronize(stream);
    }
    free_all();
}

int Engine::create(const std::vector<std::pair<uint32_t, uint32_t>>& caps, int dev, uint32_t fl) {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device is visible");
    if (dev < 0 || dev >= devices) return fail(SPH_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    if (caps.empty() || caps.size() > 65535) return fail(SPH_ERR_INVALID_ARGUMENT, "scene count must be 1..65535");
    device = dev;
    flags = fl;
    uint64_t tn = 0, tc = 0, tb = 0, blk = 0, halfTiles = 0;
    sc.resize(caps.size());
    for (size_t i = 0; i < caps.size(); ++i) {
        const uint32_t n = caps[i].first, c = caps[i].second;
        if (n == 0) return fail(SPH_ERR_INVALID_ARGUMENT, "maxParticles must be > 0");
        if (!is_pow2(c)) return fail(SPH_ERR_NOT_POWER_OF_TWO, "maxCellCount must be a power of two");
        SceneHost& s = sc[i];
        s.maxN = n;
        s.maxC = c;
        uint64_t nb = 1u << 16;   // spatial bins: about 4 per particle of capacity
        while (nb < 4ull * n && nb < (1u << 23)) nb <<= 1;
        s.binCap = (uint32_t)nb;
        s.offset = (uint32_t)tn;
        s.cellOffset = (uint32_t)tc;
        s.binOffset = (uint32_t)tb;
        tn += n;
        tc += c;
        tb += nb + 1;
        blk += blocks_for(n);
        halfTiles += nb / sphdev::kHalfTileBins;
    }
    if (tn > 0x7FFFFFFFull || tc > 0x7FFFFFFFull || tb > 0x7FFFFFFFull)
        return fail(SPH_ERR_CAPACITY, "total capacity above 2^31 slots");
    totalN = (uint32_t)tn;
    totalC = (uint32_t)tc;
    totalBins = (uint32_t)tb;
    maxBlocks = (uint32_t)blk;
    maxTiles = (uint32_t)halfTiles;
    const size_t N = totalN, C = totalC, B = totalBins, S = sc.size();
    const size_t tiles = (B + sphdev::kScanTile - 1) / sphdev::kScanTile;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        if (e == hipSuccess) e = hipMalloc(&pos[i], (N + sphdev::kPosPad) * sizeof(float4));
        if (e == hipSuccess) e = hipMemset(pos[i] + N, 0, sphdev::kPosPad * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&vel[i], N * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&ids[i], N * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipMalloc(&keys, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&buckets, C * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&rep, C * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&binStart, (B + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&binOf, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&slots, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&blockSums, (tiles + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&keysSorted, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&pairKI, N * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&pairSrc, N * sizeof(uint32_t));
    if (e == hipSuccess && !(flags & SPH_CFG_BUCKET_WALK)) {
        e = hipMalloc(&nbrList, N * 16 * sizeof(uint4));
        if (e == hipSuccess) e = hipMalloc(&nbrCount, N * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipMalloc(&prev, N * sizeof(float4));
    if (e == hipSuccess && (flags & SPH_CFG_KEEP_FORCES)) e = hipMalloc(&force, N * sizeof(float4));
    size_t stagingWords = std::max(N, C) + 4;
    if (e == hipSuccess) e = hipMalloc(&staging, N * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&stagingU, stagingWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&dTables, kMaxSubsteps * S * sizeof(sphdev::SceneDesc));
    if (e == hipSuccess) e = hipMalloc(&dBounds, S * sizeof(sphdev::BoundaryConsts));
    if (e == hipSuccess) e = hipMalloc(&dBlockScene, (maxBlocks + 1) * sizeof(uint16_t));
    if (e == hipSuccess && (flags & SPH_CFG_TILED_DENSITY)) {
        e = hipMalloc(&dTileScene, ((size_t)maxTiles + 1) * sizeof(uint16_t));
        if (e == hipSuccess) e = hipMalloc(&restList, N * sizeof(uint2));
        if (e == hipSuccess) e = hipMalloc(&restCount, sizeof(uint32_t));
    }
    hostStageBytes = kMaxSubsteps * S * sizeof(sphdev::SceneDesc) + S * sizeof(sphdev::BoundaryConsts) +
                     (maxBlocks + 1) * sizeof(uint16_t) + ((size_t)maxTiles + 1) * sizeof(uint16_t) + 512;
    if (e == hipSuccess) e = hipHostMalloc(&hostStage, hostStageBytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&binMax, (maxBlocks + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&rangeGuard, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(rangeGuard, 0xFF, sizeof(uint32_t));   // set by every grid build
    if (e == hipSuccess) e = hipHostMalloc(&binMaxHost, sizeof(uint32_t), hipHostMallocMapped);
    if (e == hipSuccess) {
        *binMaxHost = 0u;
        e = hipHostGetDevicePointer((void**)&binMaxHostDev, binMaxHost, 0);
    }
    if (e != hipSuccess) {
        int st = e == hipErrorOutOfMemory ? SPH_ERR_OUT_OF_MEMORY : SPH_ERR_DEVICE;
        std::string msg = std::string("engine create: ") + hipGetErrorString(e);
        free_all();
        return fail(st, msg);
    }
    return SPH_OK;
}

namespace {
std::atomic<uint64_t> g_epoch{0};
}

void Engine::touch() { epoch_ = ++g_epoch; }

void Engine::invalidate() {
    planKey.clear();
    // A replay of either graph may still be running on the stream: wait for it before the
    // executable graphs are destroyed.
    if ((graph[0] || graph[1]) && stream) hipStreamSynchronize(stream);
    for (int i = 0; i < 2; ++i) {
        if (graph[i]) hipGraphExecDestroy(graph[i]);
        graph[i] = nullptr;
    }
}

int Engine::set_schedule(const uint32_t* order, uint32_t count) {
    if (count == 0) {
        schedule.clear();
        return SPH_OK;
    }
    if (!order || count != sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "schedule: one entry per scene");
    std::vector<uint8_t> seen(sc.size(), 0);
    for (uint32_t q = 0; q < count; ++q) {
        if (order[q] >= sc.size() || seen[order[q]]) return fail(SPH_ERR_INVALID_ARGUMENT, "schedule: not a permutation");
        seen[order[q]] = 1;
    }
    schedule.assign(order, order + count);
    touch();
    return SPH_OK;
}

int Engine::set_params(uint32_t scene, const SphParams& p) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (!is_pow2(p.cellCount)) return fail(SPH_ERR_NOT_POWER_OF_TWO, "cellCount must be a power of two");
    if (p.cellCount > sc[scene].maxC) return fail(SPH_ERR_CAPACITY, "cellCount exceeds the scene's maxCellCount");
    if (!(p.smoothingRadius > 0.0f)) return fail(SPH_ERR_INVALID_ARGUMENT, "smoothingRadius must be > 0");
    if (!sc[scene].hasParams || memcmp(&sc[scene].params, &p, sizeof(SphParams)) != 0) touch();
    sc[scene].params = p;
    sc[scene].hasParams = true;
    return SPH_OK;
}

int Engine::emit(uint32_t scene, const SphParams& p, uint32_t count) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    SceneHost& s = sc[scene];
    if (count > s.maxN) return fail(SPH_ERR_CAPACITY, "count exceeds the scene's maxParticles");
    if (int st = set_device()) return st;
    s.n = count;
    hasGrid = false;
    touch();
    reset_occupancy();
    SPH_HIP(sphdev::launch_emit(p, count, pos[pc] + s.offset, vel[pc] + s.offset, prev + s.offset,
                                ids[ic] + s.offset, stream));
    return SPH_OK;
}

int Engine::load_state(uint32_t scene, const float* pos4, const float* vel4, uint32_t count) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    SceneHost& s = sc[scene];
    if (count && !pos4) return fail(SPH_ERR_INVALID_ARGUMENT, "load_state: null positions");
    if (count > s.maxN) return fail(SPH_ERR_CAPACITY, "count exceeds the scene's maxParticles");
    if (int st = set_device()) return st;
    s.n = count;
    hasGrid = false;
    touch();
    reset_occupancy();
    if (count == 0) return SPH_OK;
    const size_t bytes = (size_t)count * sizeof(float4);
    SPH_HIP(hipMemcpyAsync(pos[pc] + s.offset, pos4, bytes, hipMemcpyHostToDevice, stream));
    SPH_HIP(hipMemcpyAsync(prev + s.offset, pos4, bytes, hipMemcpyHostToDevice, stream));
    if (vel4) {
        SPH_HIP(hipMemcpyAsync(vel[pc] + s.offset, vel4, bytes, hipMemcpyHostToDevice, stream));
    } else {
        SPH_HIP(hipMemsetAsync(vel[pc] + s.offset, 0, bytes, stream));
    }
    SPH_HIP(sphdev::launch_iota(ids[ic] + s.offset, count, stream));
    SPH_HIP(hipStreamSynchronize(stream));
    return SPH_OK;
}

sphdev::Launch Engine::launch_for(uint32_t j) const {
    sphdev::Launch L;
    L.scenes = dTables + (size_t)j * sc.size();
    L.bounds = dBounds;
    L.blockScene = dBlockScene;
    L.blocks = planBlocks;
    L.numScenes = (uint32_t)sc.size();
    L.total = totalN;
    L.totalCells = totalC;
    L.totalBins = planBins;
    L.modes = 0u;
    for (const SceneHost& s : sc) L.modes |= s.mode;
    L.tileScene = dTileScene;
    L.tiles = dTileScene ? planTiles : 0u;
    L.rangeGuard = rangeGuard;
    return L;
}

// Builds (when anything changed) the per-substep scene tables and the block table and
// uploads them; returns the number of substeps the step runs.
int Engine::ensure_plan(float dt, bool singleSubstep, uint32_t* maxSubOut) {
    const size_t S = sc.size();
    std::vector<uint32_t> sub(S, 0);
    std::vector<float> dts(S, 0.0f);
    uint32_t maxSub = 0;
    for (size_t i = 0; i < S; ++i) {
        const SceneHost& s = sc[i];
        if (s.n == 0) continue;
        if (!s.hasParams) return fail(SPH_ERR_STATE, "a scene with particles has no params");
        if (singleSubstep) {
            sub[i] = 1;
            dts[i] = dt;
        } else {
            // FluidSolver::step (fluid_sph.cpp:311-321)
            sub[i] = sph_substeps(&s.params, dt);
            const float frameStep = SPH_MAX_FRAME_STEP < dt ? SPH_MAX_FRAME_STEP : dt;
            dts[i] = frameStep / (float)sub[i];
        }
        maxSub = std::max(maxSub, sub[i]);
    }
    *maxSubOut = maxSub;
    if (maxSub == 0) return SPH_OK;
    planDt = dt;
    planSingle = singleSubstep;
    // Key: everything the tables depend on.
    std::vector<uint8_t> key;
    auto put = [&key](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        key.insert(key.end(), b, b + n);
    };
    put(&dt, sizeof(dt));
    put(&singleSubstep, sizeof(singleSubstep));
    for (size_t i = 0; i < S; ++i) {
        const SceneHost& s = sc[i];
        put(&s.n, sizeof(s.n));
        put(&s.mode, sizeof(s.mode));
        put(&s.eos, sizeof(s.eos));
        put(s.tileLog2, sizeof(s.tileLog2));
        put(&s.list, sizeof(s.list));
        if (s.n) put(&s.params, sizeof(s.params));
    }
    put(schedule.data(), schedule.size() * sizeof(uint32_t));
    if (key == planKey) return SPH_OK;
    invalidate();
    // Tables into pinned memory, then one copy each (the previous upload has finished:
    // the plan changes only between steps, after a synchronise below).
    SPH_HIP(hipStreamSynchronize(stream));
    uint8_t* h = static_cast<uint8_t*>(hostStage);
    sphdev::SceneDesc* tables = reinterpret_cast<sphdev::SceneDesc*>(h);
    sphdev::BoundaryConsts* bounds = reinterpret_cast<sphdev::BoundaryConsts*>(h + kMaxSubsteps * S * sizeof(sphdev::SceneDesc));
    uint16_t* blockScene = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(bounds) + S * sizeof(sphdev::BoundaryConsts));
    uint16_t* tileScene = blockScene + (maxBlocks + 1);
    // Workgroups of particle-parallel (and half-tile) launches are laid out scene by scene
    // in the order of `schedule` (scene order unless set): offsets first, tables after.
    std::vector<uint32_t> blockOff(S, 0), tileOff(S, 0);
    uint32_t blocks = 0, tiles = 0;
    for (size_t q = 0; q < S; ++q) {
        const uint32_t i = schedule.empty() ? (uint32_t)q : schedule[q];
        const SceneHost& s = sc[i];
        const uint32_t b = blocks_for(s.n);
        const uint32_t sceneTiles = s.n ? make_bins(s, s.params).bins / sphdev::kHalfTileBins : 0u;
        blockOff[i] = blocks;
        tileOff[i] = tiles;
        for (uint32_t r = 0; r < b; ++r) blockScene[blocks + r] = (uint16_t)i;
        for (uint32_t r = 0; r < sceneTiles; ++r) tileScene[tiles + r] = (uint16_t)i;
        blocks += b;
        tiles += sceneTiles;
    }
    uint32_t binBase = 0;   // bins are packed by the bins each scene actually uses
    for (size_t i = 0; i < S; ++i) {
        const SceneHost& s = sc[i];
        const sphdev::BinGrid grid = s.n ? make_bins(s, s.params) : sphdev::BinGrid{};
        const uint32_t sceneBins = s.n ? grid.bins : 0u;
        sphdev::BoundaryConsts& bc = bounds[i];
        memset(&bc, 0, sizeof(bc));
        if (s.n) {
            memcpy(bc.colliders, s.params.colliders, sizeof(bc.colliders));
            memcpy(bc.fields, s.params.fields, sizeof(bc.fields));
        }
        for (uint32_t j = 0; j < maxSub; ++j) {
            sphdev::SceneDesc& d = tables[(size_t)j * S + i];
            memset(&d, 0, sizeof(d));
            d.offset = s.offset;
            d.capacity = s.maxN;
            d.cellOffset = s.cellOffset;
            d.binOffset = binBase;
            d.binCapacity = sceneBins;
            d.blockOffset = blockOff[i];
            d.tileOffset = tileOff[i];
            d.active = j < sub[i] ? 1u : 0u;
            if (s.n) {
                d.k = make_consts(s, s.params, dts[i], nbrList != nullptr);
                d.g = grid;
            } else {
                d.k.cellMask = 0;
                d.g.bins = 0;
                d.g.dims[0] = d.g.dims[1] = d.g.dims[2] = 1;
                d.g.tiles[0] = d.g.tiles[1] = d.g.tiles[2] = 1;
                d.g.tileLog2[0] = d.g.tileLog2[1] = d.g.tileLog2[2] = 2;
            }
            d.k.n = s.n;
        }
        binBase += sceneBins + 1u;
    }
    planBlocks = blocks;
    // The tiled density pass needs one bin per cell and a neighbourhood table of at most 256
    // cells in every scene; scenes with far fewer buckets than particles would send most
    // particles to its global walk. Otherwise the whole launch uses the plain walk.
    bool tileable = dTileScene != nullptr;
    for (size_t i = 0; i < S && tileable; ++i) {
        const SceneHost& sh = sc[i];
        if (sh.n == 0) continue;
        const sphdev::BinGrid& gg = tables[i].g;
        uint32_t dm[3] = {1u << gg.tileLog2[0], 1u << gg.tileLog2[1], 1u << gg.tileLog2[2]};
        dm[gg.tileLog2[2] ? 2 : (gg.tileLog2[1] ? 1 : 0)] >>= 1;
        const uint32_t box = (dm[0] + 2u) * (dm[1] + 2u) * (dm[2] + 2u);
        if (gg.shift != 0u || box > 256u || sh.params.cellCount < sh.n) tileable = false;
    }
    planTiles = tileable ? tiles : 0u;
    planBins = binBase;
    planSub = maxSub;
    SPH_HIP(hipMemcpyAsync(dTables, tables, (size_t)maxSub * S * sizeof(sphdev::SceneDesc), hipMemcpyHostToDevice, stream));
    SPH_HIP(hipMemcpyAsync(dBounds, bounds, S * sizeof(sphdev::BoundaryConsts), hipMemcpyHostToDevice, stream));
    SPH_HIP(hipMemcpyAsync(dBlockScene, blockScene, (size_t)(blocks + 1) * sizeof(uint16_t), hipMemcpyHostToDevice,
                           stream));
    if (planTiles)
        SPH_HIP(hipMemcpyAsync(dTileScene, tileScene, (size_t)planTiles * sizeof(uint16_t), hipMemcpyHostToDevice,
                               stream));
    SPH_HIP(hipStreamSynchronize(stream));
    planKey.swap(key);
    return SPH_OK;
}

int Engine::run_substep(uint32_t j, bool prepareOnly, const sphdev::NeighbourSums* sums, bool gridOnly) {
    const sphdev::Launch L = launch_for(j);
    sphdev::GridBuffers g;
    g.pos = pos[pc];
    g.vel = vel[pc];
    g.ids = ids[ic];
    g.keys = keys;
    g.buckets = buckets;
    g.rep = rep;
    g.binStart = binStart;
    g.binOf = binOf;
    g.slots = slots;
    g.pairKI = pairKI;
    g.pairSrc = pairSrc;
    g.blockSums = blockSums;
    g.posSorted = pos[1 - pc];
    g.velSorted = vel[1 - pc];
    g.idsSorted = ids[1 - ic];
    g.keysSorted = keysSorted;
    g.restCount = restCount;
    g.rest = restList;
    g.binMax = binMax;
    g.binMaxHost = binMaxHostDev;
    {
        Timer t(this, KC_FILL);
        SPH_HIP(sphdev::launch_grid_clear(L, g, stream));
    }
    {
        Timer t(this, KC_KEYS);
        SPH_HIP(sphdev::launch_keys(L, g, stream));
    }
    {
        Timer t(this, KC_BINSLOT);
        SPH_HIP(sphdev::launch_bin_slots(L, g, stream));
    }
    {
        Timer t(this, KC_SCAN);
        SPH_HIP(sphdev::launch_scan(binStart, planBins, totalN, blockSums, stream));
    }
    {
        Timer t(this, KC_SCATTER);
        SPH_HIP(sphdev::launch_scatter(L, g, stream));
    }
    {
        Timer t(this, KC_RANK);
        SPH_HIP(sphdev::launch_rank_reorder(L, g, stream));
    }
    if (gridOnly) return SPH_OK;
    {
        Timer t(this, KC_DENSITY);
        // A prepare-only pass is followed by no force pass here: no neighbour list.
        const bool list = !prepareOnly;
        sphdev::DensityTiles tiles{binStart, restList, restCount, kRestBlocks};
        SPH_HIP(sphdev::launch_density(L, g.posSorted, g.velSorted, buckets, list ? nbrList : nullptr,
                                       list ? nbrCount : nullptr, sums, stream,
                                       (restList && !sums) ? &tiles : nullptr));
    }
    if (prepareOnly) return SPH_OK;
    {
        Timer t(this, KC_FORCE);
        SPH_HIP(sphdev::launch_force(L, g.posSorted, g.velSorted, buckets, nbrList, nbrCount, pos[pc],
                                     vel[pc], force, stream));
    }
    ic = 1 - ic;
    return SPH_OK;
}

int Engine::enqueue_step(uint32_t maxSub) {
    for (uint32_t j = 0; j < maxSub; ++j) {
        if (int st = run_substep(j)) return st;
    }
    return SPH_OK;
}

// A new state: the occupancy of grids built from the old one no longer applies. Grid builds
// still in flight write their copy when they finish, so they are waited for first.
void Engine::reset_occupancy() {
    if (!binMaxHost) return;
    hipStreamSynchronize(stream);
    __atomic_store_n(binMaxHost, 0u, __ATOMIC_RELAXED);
}

// The grid build ranks each particle against every member of its spatial bin, so its cost
// grows with the square of the largest bin occupancy. A state that piled SPH_MAX_BIN_OCCUPANCY
// or more particles into one bin (an unstable run collapsing onto a wall or a corner) is
// refused rather than stepped at that cost. The figure is the last finished build's.
int Engine::check_occupancy() {
    const uint32_t m = last_bin_max();
    if (m < SPH_MAX_BIN_OCCUPANCY) return SPH_OK;
    return fail(SPH_ERR_DEGENERATE, "a grid bin holds " + std::to_string(m) +
                                        " particles (SPH_MAX_BIN_OCCUPANCY is " +
                                        std::to_string(SPH_MAX_BIN_OCCUPANCY) + "): the state has piled up");
}

int Engine::step(float dt, bool singleSubstep) {
    if (int st = set_device()) return st;
    if (int st = check_occupancy()) return st;
    touch();
    uint32_t maxSub = 0;
    if (int st = ensure_plan(dt, singleSubstep, &maxSub)) return st;
    if (maxSub == 0) return SPH_OK;
    const bool useGraph = !(flags & SPH_CFG_NO_GRAPH) && !profiling;
    if (!useGraph) {
        if (int st = enqueue_step(maxSub)) return st;
    } else {
        const int parity = ic;
        if (!graph[parity]) {
            hipGraph_t g = nullptr;
            SPH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int st = enqueue_step(maxSub);
            hipError_t ce = hipStreamEndCapture(stream, &g);
            if (st) {
                if (g) hipGraphDestroy(g);
                return st;
            }
            if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
            hipError_t ie = hipGraphInstantiate(&graph[parity], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
            ic = parity;   // capture only recorded the launches; the flips happen at replay
        }
        SPH_HIP(hipGraphLaunch(graph[parity], stream));
        ic = (maxSub & 1u) ? 1 - parity : parity;
    }
    hasGrid = true;
    return SPH_OK;
}

int Engine::prepare(float dt, const sphdev::NeighbourSums* sums) {
    if (int st = set_device()) return st;
    if (int st = check_occupancy()) return st;
    touch();
    uint32_t maxSub = 0;
    if (int st = ensure_plan(dt, true, &maxSub)) return st;
    if (maxSub == 0) return SPH_OK;
    return run_substep(0, true, sums);
}

int Engine::build_grid() {
    if (int st = set_device()) return st;
    if (int st = check_occupancy()) return st;
    touch();
    // The grid does not depend on dt: re-validate the last plan (its tables are rebuilt if
    // counts or params changed since), or make one for a 1/600 s step if there is none.
    uint32_t maxSub = 0;
    if (int st = ensure_plan(planDt, planSingle, &maxSub)) return st;
    if (maxSub == 0) return SPH_OK;
    return run_substep(0, true, nullptr, true);
}

int Engine::ensure_force_buffer() {
    if (force) return SPH_OK;
    if (int st = set_device()) return st;
    SPH_HIP(hipMalloc(&force, (size_t)totalN * sizeof(float4)));
    invalidate();   // captured graphs wrote no forces
    return SPH_OK;
}

int Engine::synchronize() {
    if (int st = set_device()) return st;
    SPH_HIP(hipStreamSynchronize(stream));
    return SPH_OK;
}

int Engine::read_particles(uint32_t scene, float* pos4, float* vel4) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    const SceneHost& s = sc[scene];
    if (int st = set_device()) return st;
    if (s.n == 0) return SPH_OK;
    const size_t bytes = (size_t)s.n * sizeof(float4);
    if (pos4) {
        SPH_HIP(sphdev::launch_unpermute(pos[pc] + s.offset, ids[ic] + s.offset, staging, s.n, stream));
        SPH_HIP(hipMemcpyAsync(pos4, staging, bytes, hipMemcpyDeviceToHost, stream));
        SPH_HIP(hipStreamSynchronize(stream));
    }
    if (vel4) {
        SPH_HIP(sphdev::launch_unpermute(vel[pc] + s.offset, ids[ic] + s.offset, staging, s.n, stream));
        SPH_HIP(hipMemcpyAsync(vel4, staging, bytes, hipMemcpyDeviceToHost, stream));
        SPH_HIP(hipStreamSynchronize(stream));
    }
    return SPH_OK;
}

int Engine::read_previous(uint32_t scene, float* pos4) {
    if (scene >= sc.size() || !pos4) return fail(SPH_ERR_INVALID_ARGUMENT, "read_previous: bad argument");
    const SceneHost& s = sc[scene];
    if (s.n == 0) return SPH_OK;
    if (int st = set_device()) return st;
    SPH_HIP(hipMemcpyAsync(pos4, prev + s.offset, (size_t)s.n * sizeof(float4), hipMemcpyDeviceToHost, stream));
    SPH_HIP(hipStreamSynchronize(stream));
    return SPH_OK;
}

int Engine::keep_rendered(uint32_t scene) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    const SceneHost& s = sc[scene];
    if (s.n == 0) return SPH_OK;
    if (int st = set_device()) return st;
    SPH_HIP(sphdev::launch_unpermute(pos[pc] + s.offset, ids[ic] + s.offset, prev + s.offset, s.n, stream));
    return SPH_OK;
}

int Engine::read_grid(uint32_t scene, uint32_t* cellOf, uint32_t* cellCounts, uint32_t* cellParticles) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (!hasGrid) return fail(SPH_ERR_STATE, "no grid: run a substep first");
    const SceneHost& s = sc[scene];
    if (int st = set_device()) return st;
    const uint32_t n = s.n, C = s.params.cellCount;
    if (cellOf && n) {
        SPH_HIP(sphdev::launch_gather_keys(keysSorted + s.offset, ids[ic] + s.offset, s.cellOffset, stagingU, n, stream));
        SPH_HIP(hipMemcpyAsync(cellOf, stagingU, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
        SPH_HIP(hipStreamSynchronize(stream));
    }
    if (cellCounts || cellParticles) {
        std::vector<uint2> bt(C);
        std::vector<uint32_t> idv(n);
        SPH_HIP(hipMemcpyAsync(bt.data(), buckets + s.cellOffset, (size_t)C * sizeof(uint2), hipMemcpyDeviceToHost,
                               stream));
        if (n) SPH_HIP(hipMemcpyAsync(idv.data(), ids[ic] + s.offset, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
        SPH_HIP(hipStreamSynchronize(stream));
        for (uint32_t b = 0; b < C; ++b) {
            const uint32_t c = bt[b].y;
            if (cellCounts) cellCounts[b] = c;
            if (cellParticles) {
                for (uint32_t k = 0; k < SPH_CELL_CAPACITY; ++k)
                    cellParticles[(size_t)b * SPH_CELL_CAPACITY + k] =
                        k < c ? idv[bt[b].x - s.offset + k] : 0xFFFFFFFFu;
            }
        }
    }
    return SPH_OK;
}

int Engine::read_sorted(uint32_t scene, uint32_t* keysOut, uint32_t* idsOut, uint32_t* bucketStarts) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (!hasGrid) return fail(SPH_ERR_STATE, "no grid: run a substep first");
    const SceneHost& s = sc[scene];
    if (int st = set_device()) return st;
    const uint32_t n = s.n, C = s.params.cellCount;
    std::vector<uint32_t> k(n);
    if ((keysOut || idsOut) && n) {
        SPH_HIP(hipMemcpyAsync(k.data(), keysSorted + s.offset, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
        if (idsOut) SPH_HIP(hipMemcpyAsync(idsOut, ids[ic] + s.offset, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    }
    std::vector<uint2> bt(bucketStarts ? C : 0);
    if (bucketStarts)
        SPH_HIP(hipMemcpyAsync(bt.data(), buckets + s.cellOffset, (size_t)C * sizeof(uint2), hipMemcpyDeviceToHost,
                               stream));
    SPH_HIP(hipStreamSynchronize(stream));
    if (keysOut)
        for (uint32_t i = 0; i < n; ++i) keysOut[i] = k[i] - s.cellOffset;
    if (bucketStarts)
        for (uint32_t b = 0; b < C; ++b) bucketStarts[b] = bt[b].y ? bt[b].x - s.offset : 0xFFFFFFFFu;
    return SPH_OK;
}

int Engine::short_division(uint32_t scene, uint32_t* out) {
    if (scene >= sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (int st = set_device()) return st;
    uint32_t guard = 1u;
    SPH_HIP(hipStreamSynchronize(stream));
    SPH_HIP(hipMemcpy(&guard, rangeGuard, sizeof(guard), hipMemcpyDeviceToHost));
    const SceneHost& s = sc[scene];
    const sphdev::StepConsts k = make_consts(s, s.params, 1.0f, false);
    *out = hasGrid && s.n && guard == 0u && k.shortDivide ? 1u : 0u;
    return SPH_OK;
}

int Engine::read_forces(uint32_t scene, float* force4) {
    if (scene >= sc.size() || !force4) return fail(SPH_ERR_INVALID_ARGUMENT, "read_forces: bad argument");
    if (!force) return fail(SPH_ERR_STATE, "created without SPH_CFG_KEEP_FORCES");
    if (!hasGrid) return fail(SPH_ERR_STATE, "no forces: run a substep first");
    const SceneHost& s = sc[scene];
    if (s.n == 0) return SPH_OK;
    if (int st = set_device()) return st;
    SPH_HIP(sphdev::launch_unpermute(force + s.offset, ids[ic] + s.offset, staging, s.n, stream));
    SPH_HIP(hipMemcpyAsync(force4, staging, (size_t)s.n * sizeof(float4), hipMemcpyDeviceToHost, stream));
    SPH_HIP(hipStreamSynchronize(stream));
    return SPH_OK;
}

int Engine::profile_read(const char** names, double* ms, uint32_t* launches, uint32_t capacity) {
    if (int st = set_device()) return st;
    SPH_HIP(hipStreamSynchronize(stream));
    for (const TimedLaunch& t : timed) {
        float el = 0.0f;
        if (hipEventElapsedTime(&el, t.begin, t.end) == hipSuccess) {
            accMs[t.cls] += el;
            accLaunches[t.cls] += 1;
        }
        eventPool.push_back(t.begin);
        eventPool.push_back(t.end);
    }
    timed.clear();
    uint32_t w = 0;
    for (uint32_t c = 0; c < KC_COUNT && w < capacity; ++c, ++w) {
        if (names) names[w] = kKernelNames[c];
        if (ms) ms[w] = accMs[c];
        if (launches) launches[w] = accLaunches[c];
    }
    for (uint32_t c = 0; c < KC_COUNT; ++c) {
        accMs[c] = 0.0;
        accLaunches[c] = 0;
    }
    return (int)w;
}

}  // namespace sphhost
