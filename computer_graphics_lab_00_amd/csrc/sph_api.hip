// sph_api.hip — the solver handle behind the C ABI (physics::FluidSolver,
// src/physics/fluid_sph.h:101-127), driving the gfx950 kernels on one HIP stream.
//
// Per substep (src/physics/fluid_sph.cpp:190-309):
//   fill(counts = 0) -> keys -> scan -> scatter -> rank+reorder   (FluidGrid::build)
//   density (in place .w)                                         (fluid_sph.cpp:197-222)
//   force + integrate + walls + colliders -> back buffer          (fluid_sph.cpp:224-305)
// The state stays bucket-sorted between substeps; ids carry each slot's particle index
// and every read un-permutes to index order.
#include <math.h>
#include <string.h>

#include <cstdio>
#include <string>
#include <vector>

#include "sph_device.h"
#include "sph_internal.h"

using sphdev::StepConsts;

namespace {

constexpr float kPi = 3.14159265358979323846f;   // std::numbers::pi_v<float>

enum KernelClass : uint32_t {
    KC_FILL = 0,
    KC_KEYS,
    KC_BINSLOT,
    KC_SCAN,
    KC_SCATTER,
    KC_RANK,
    KC_DENSITY,
    KC_FORCE,
    KC_COUNT
};

const char* const kKernelNames[KC_COUNT] = {"grid_clear", "keys", "bin_slots", "scan", "scatter", "rank_reorder",
                                            "density", "force_integrate"};

struct TimedLaunch {
    uint32_t cls;
    hipEvent_t begin;
    hipEvent_t end;
};

}  // namespace

struct SphSolver {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t maxN = 0, maxC = 0, flags = 0;
    uint32_t n = 0;
    uint32_t mode = SPH_MODE_REFERENCE;
    uint32_t eos = SPH_EOS_LINEAR;
    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    uint32_t* ids[2] = {nullptr, nullptr};
    int pc = 0;   // pos/vel buffer holding the state
    int ic = 0;   // ids buffer matching the state's order
    uint32_t maxBins = 0;
    uint32_t* keys = nullptr;
    uint32_t* counts = nullptr;
    uint32_t* rep = nullptr;
    uint32_t* bucketStart = nullptr;
    uint32_t* binStart = nullptr;
    uint32_t* binOf = nullptr;
    uint32_t* slots = nullptr;
    uint32_t* blockSums = nullptr;
    uint32_t* keysSorted = nullptr;
    uint2* pairKI = nullptr;
    uint32_t* pairSrc = nullptr;
    uint4* nbrList = nullptr;       // kNeighbourCap / 4 groups x maxN
    uint32_t* nbrCount = nullptr;
    float4* prev = nullptr;
    float4* force = nullptr;
    float4* staging = nullptr;
    uint32_t* stagingU = nullptr;
    uint32_t gridC = 0;
    bool hasGrid = false;
    bool profiling = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> eventPool;
    double accMs[KC_COUNT] = {};
    uint32_t accLaunches[KC_COUNT] = {};
};

namespace {

int hip_fail(hipError_t e, const char* what) {
    return sphhost::fail(SPH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define SPH_HIP(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #call);    \
    } while (0)

hipEvent_t take_event(SphSolver* s) {
    if (!s->eventPool.empty()) {
        hipEvent_t e = s->eventPool.back();
        s->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one kernel-class launch with events when profiling is on.
struct Timer {
    SphSolver* s;
    uint32_t cls;
    hipEvent_t b = nullptr;
    Timer(SphSolver* solver, uint32_t c) : s(solver), cls(c) {
        if (s->profiling) {
            b = take_event(s);
            if (b) hipEventRecord(b, s->stream);
        }
    }
    ~Timer() {
        if (s->profiling && b) {
            hipEvent_t e = take_event(s);
            if (e) {
                hipEventRecord(e, s->stream);
                s->timed.push_back({cls, b, e});
            }
        }
    }
};

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

StepConsts make_consts(const SphSolver* s, const SphParams& p, float dt) {
    StepConsts k;
    memset(&k, 0, sizeof(k));
    const float h = p.smoothingRadius;
    k.n = s->n;
    k.cellMask = p.cellCount - 1u;
    k.mode = s->mode;
    k.eos = s->eos;
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
    k.listCap = (s->nbrList && k.r2Cut <= k.h2) ? sphdev::kNeighbourCap : 0u;
    return k;
}

// Spatial bins over the container (plus one cell of margin each side). Any deterministic
// mapping is correct here; this one keeps neighbouring cells close in memory. If the
// container holds more cells than the bin arrays, bins are coarsened by powers of two.
sphdev::BinGrid make_bins(const SphSolver* s, const SphParams& p) {
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
    uint32_t shift = 0;
    for (;;) {
        uint64_t tiles = 1;
        for (int a = 0; a < 3; ++a) {
            const uint64_t d = (uint64_t)((span[a] + (1ll << shift) - 1) >> shift);
            g.dims[a] = (uint32_t)d;
            g.tiles[a] = (uint32_t)((d + 3) / 4);
            tiles *= g.tiles[a];
        }
        if (tiles * 64ull <= s->maxBins || shift >= 31) {
            g.bins = (uint32_t)(tiles * 64ull);
            break;
        }
        ++shift;
    }
    g.shift = shift;
    for (int a = 0; a < 3; ++a) g.origin[a] = (int32_t)lo[a];
    return g;
}

void free_all(SphSolver* s) {
    for (int i = 0; i < 2; ++i) {
        hipFree(s->pos[i]);
        hipFree(s->vel[i]);
        hipFree(s->ids[i]);
    }
    hipFree(s->keys);
    hipFree(s->counts);
    hipFree(s->rep);
    hipFree(s->bucketStart);
    hipFree(s->binStart);
    hipFree(s->binOf);
    hipFree(s->slots);
    hipFree(s->blockSums);
    hipFree(s->keysSorted);
    hipFree(s->pairKI);
    hipFree(s->pairSrc);
    hipFree(s->nbrList);
    hipFree(s->nbrCount);
    hipFree(s->prev);
    hipFree(s->force);
    hipFree(s->staging);
    hipFree(s->stagingU);
    for (auto& t : s->timed) {
        hipEventDestroy(t.begin);
        hipEventDestroy(t.end);
    }
    for (hipEvent_t e : s->eventPool) hipEventDestroy(e);
    if (s->stream) hipStreamDestroy(s->stream);
}

int check_params(const SphSolver* s, const SphParams* p) {
    if (!p) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "params is null");
    if (!is_pow2(p->cellCount)) return sphhost::fail(SPH_ERR_NOT_POWER_OF_TWO, "cellCount must be a power of two");
    if (p->cellCount > s->maxC) return sphhost::fail(SPH_ERR_CAPACITY, "cellCount exceeds the solver's maxCellCount");
    if (!(p->smoothingRadius > 0.0f)) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "smoothingRadius must be > 0");
    return SPH_OK;
}

int run_substep(SphSolver* s, const SphParams& p, float dt) {
    const StepConsts k = make_consts(s, p, dt);
    const sphdev::BinGrid bins = make_bins(s, p);
    const uint32_t C = p.cellCount;
    sphdev::GridBuffers g;
    g.pos = s->pos[s->pc];
    g.vel = s->vel[s->pc];
    g.ids = s->ids[s->ic];
    g.keys = s->keys;
    g.counts = s->counts;
    g.rep = s->rep;
    g.binStart = s->binStart;
    g.binOf = s->binOf;
    g.slots = s->slots;
    g.pairKI = s->pairKI;
    g.pairSrc = s->pairSrc;
    g.blockSums = s->blockSums;
    g.posSorted = s->pos[1 - s->pc];
    g.velSorted = s->vel[1 - s->pc];
    g.idsSorted = s->ids[1 - s->ic];
    g.keysSorted = s->keysSorted;
    g.bucketStart = s->bucketStart;
    {
        Timer t(s, KC_FILL);
        SPH_HIP(sphdev::launch_grid_clear(k, bins, g, s->stream));
    }
    {
        Timer t(s, KC_KEYS);
        SPH_HIP(sphdev::launch_keys(k, bins, g, s->stream));
    }
    {
        Timer t(s, KC_BINSLOT);
        SPH_HIP(sphdev::launch_bin_slots(k, g, s->stream));
    }
    {
        Timer t(s, KC_SCAN);
        SPH_HIP(sphdev::launch_scan(s->binStart, bins.bins, s->n, s->blockSums, s->stream));
    }
    {
        Timer t(s, KC_SCATTER);
        SPH_HIP(sphdev::launch_scatter(k, g, s->stream));
    }
    {
        Timer t(s, KC_RANK);
        SPH_HIP(sphdev::launch_rank_reorder(k, g, s->stream));
    }
    {
        Timer t(s, KC_DENSITY);
        SPH_HIP(sphdev::launch_density(k, g.posSorted, g.velSorted, s->bucketStart, s->counts,
                                       k.listCap ? s->nbrList : nullptr,
                                       k.listCap ? s->nbrCount : nullptr, s->stream));
    }
    {
        sphdev::BoundaryConsts b;
        memset(&b, 0, sizeof(b));
        memcpy(b.colliders, p.colliders, sizeof(b.colliders));
        memcpy(b.fields, p.fields, sizeof(b.fields));
        Timer t(s, KC_FORCE);
        SPH_HIP(sphdev::launch_force(k, b, g.posSorted, g.velSorted, s->bucketStart, s->counts,
                                     k.listCap ? s->nbrList : nullptr,
                                     k.listCap ? s->nbrCount : nullptr, s->pos[s->pc], s->vel[s->pc],
                                     (s->flags & SPH_CFG_KEEP_FORCES) ? s->force : nullptr, s->stream));
    }
    s->ic = 1 - s->ic;
    s->gridC = C;
    s->hasGrid = true;
    return SPH_OK;
}

int set_device(const SphSolver* s) {
    SPH_HIP(hipSetDevice(s->device));
    return SPH_OK;
}

}  // namespace

extern "C" int sph_create(const SphSolverConfig* cfg, SphSolver** out) {
    sphhost::clear_error();
    if (!cfg || !out) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_create: null argument");
    *out = nullptr;
    if (cfg->maxParticles == 0) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "maxParticles must be > 0");
    if (!is_pow2(cfg->maxCellCount)) return sphhost::fail(SPH_ERR_NOT_POWER_OF_TWO, "maxCellCount must be a power of two");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0)
        return sphhost::fail(SPH_ERR_NO_DEVICE, "no HIP device is visible");
    if (cfg->device < 0 || cfg->device >= devices)
        return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    SphSolver* s = new SphSolver();
    s->device = cfg->device;
    s->maxN = cfg->maxParticles;
    s->maxC = cfg->maxCellCount;
    s->flags = cfg->flags;
    const size_t N = cfg->maxParticles, C = cfg->maxCellCount;
    // Spatial bins: about 4 per particle of capacity, between 2^16 and 2^23.
    size_t nb = 1u << 16;
    while (nb < 4 * N && nb < (1u << 23)) nb <<= 1;
    s->maxBins = (uint32_t)nb;
    const size_t tiles = (nb + sphdev::kScanTile - 1) / sphdev::kScanTile;
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        if (e == hipSuccess) e = hipMalloc(&s->pos[i], N * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&s->vel[i], N * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&s->ids[i], N * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipMalloc(&s->keys, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->counts, C * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->rep, C * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->bucketStart, C * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->binStart, (nb + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->binOf, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->slots, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->blockSums, (tiles + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->keysSorted, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->pairKI, N * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&s->pairSrc, N * sizeof(uint32_t));
    if (e == hipSuccess && !(cfg->flags & SPH_CFG_BUCKET_WALK)) {
        e = hipMalloc(&s->nbrList, N * (sphdev::kNeighbourCap / 4) * sizeof(uint4));
        if (e == hipSuccess) e = hipMalloc(&s->nbrCount, N * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipMalloc(&s->prev, N * sizeof(float4));
    if (e == hipSuccess && (cfg->flags & SPH_CFG_KEEP_FORCES)) e = hipMalloc(&s->force, N * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&s->staging, N * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&s->stagingU, (N > C ? N : C) * sizeof(uint32_t) + 4);
    if (e != hipSuccess) {
        int st = e == hipErrorOutOfMemory ? SPH_ERR_OUT_OF_MEMORY : SPH_ERR_DEVICE;
        std::string msg = std::string("sph_create: ") + hipGetErrorString(e);
        free_all(s);
        delete s;
        return sphhost::fail(st, msg);
    }
    *out = s;
    return SPH_OK;
}

extern "C" int sph_destroy(SphSolver* s) {
    if (!s) return SPH_OK;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return SPH_OK;
}

extern "C" int sph_emit(SphSolver* s, const SphParams* p, uint32_t count) {
    sphhost::clear_error();
    if (!s || !p) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_emit: null argument");
    if (count > s->maxN) return sphhost::fail(SPH_ERR_CAPACITY, "count exceeds the solver's maxParticles");
    if (int st = set_device(s)) return st;
    s->n = count;
    s->pc = 0;
    s->ic = 0;
    s->hasGrid = false;
    SPH_HIP(sphdev::launch_emit(*p, count, s->pos[0], s->vel[0], s->prev, s->ids[0], s->stream));
    return SPH_OK;
}

extern "C" int sph_load_state(SphSolver* s, const float* pos4, const float* vel4, uint32_t count) {
    sphhost::clear_error();
    if (!s || (count && !pos4)) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_load_state: null argument");
    if (count > s->maxN) return sphhost::fail(SPH_ERR_CAPACITY, "count exceeds the solver's maxParticles");
    if (int st = set_device(s)) return st;
    s->n = count;
    s->pc = 0;
    s->ic = 0;
    s->hasGrid = false;
    if (count == 0) return SPH_OK;
    const size_t bytes = (size_t)count * sizeof(float4);
    SPH_HIP(hipMemcpyAsync(s->pos[0], pos4, bytes, hipMemcpyHostToDevice, s->stream));
    SPH_HIP(hipMemcpyAsync(s->prev, pos4, bytes, hipMemcpyHostToDevice, s->stream));
    if (vel4) {
        SPH_HIP(hipMemcpyAsync(s->vel[0], vel4, bytes, hipMemcpyHostToDevice, s->stream));
    } else {
        SPH_HIP(hipMemsetAsync(s->vel[0], 0, bytes, s->stream));
    }
    SPH_HIP(sphdev::launch_iota(s->ids[0], count, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

extern "C" int sph_substep(SphSolver* s, const SphParams* p, float dt) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep: null solver");
    if (int st = check_params(s, p)) return st;
    if (s->n == 0) return SPH_OK;
    if (int st = set_device(s)) return st;
    return run_substep(s, *p, dt);
}

extern "C" int sph_step(SphSolver* s, const SphParams* p, float deltaSeconds) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_step: null solver");
    // FluidSolver::step (src/physics/fluid_sph.cpp:311-321): silent no-op when empty or dt <= 0.
    if (s->n == 0 || deltaSeconds <= 0.0f) return SPH_OK;
    if (int st = check_params(s, p)) return st;
    if (int st = set_device(s)) return st;
    const uint32_t substeps = sph_substeps(p, deltaSeconds);
    const float frameStep = SPH_MAX_FRAME_STEP < deltaSeconds ? SPH_MAX_FRAME_STEP : deltaSeconds;
    const float dt = frameStep / (float)substeps;
    for (uint32_t i = 0; i < substeps; ++i) {
        if (int st = run_substep(s, *p, dt)) return st;
    }
    return SPH_OK;
}

extern "C" int sph_synchronize(SphSolver* s) {
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_synchronize: null solver");
    if (int st = set_device(s)) return st;
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

extern "C" uint32_t sph_particle_count(const SphSolver* s) { return s ? s->n : 0u; }

extern "C" int sph_read_particles(SphSolver* s, float* pos4, float* vel4) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_particles: null solver");
    if (int st = set_device(s)) return st;
    const size_t bytes = (size_t)s->n * sizeof(float4);
    if (s->n == 0) return SPH_OK;
    if (pos4) {
        SPH_HIP(sphdev::launch_unpermute(s->pos[s->pc], s->ids[s->ic], s->staging, s->n, s->stream));
        SPH_HIP(hipMemcpyAsync(pos4, s->staging, bytes, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
    }
    if (vel4) {
        SPH_HIP(sphdev::launch_unpermute(s->vel[s->pc], s->ids[s->ic], s->staging, s->n, s->stream));
        SPH_HIP(hipMemcpyAsync(vel4, s->staging, bytes, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
    }
    return SPH_OK;
}

extern "C" int sph_read_previous(SphSolver* s, float* pos4) {
    sphhost::clear_error();
    if (!s || !pos4) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_previous: null argument");
    if (s->n == 0) return SPH_OK;
    if (int st = set_device(s)) return st;
    SPH_HIP(hipMemcpyAsync(pos4, s->prev, (size_t)s->n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

extern "C" int sph_keep_rendered(SphSolver* s) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_keep_rendered: null solver");
    if (s->n == 0) return SPH_OK;
    if (int st = set_device(s)) return st;
    SPH_HIP(sphdev::launch_unpermute(s->pos[s->pc], s->ids[s->ic], s->prev, s->n, s->stream));
    return SPH_OK;
}

extern "C" int sph_read_grid(SphSolver* s, uint32_t* cellOf, uint32_t* cellCounts, uint32_t* cellParticles) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_grid: null solver");
    if (!s->hasGrid) return sphhost::fail(SPH_ERR_STATE, "no grid: run a substep first");
    if (int st = set_device(s)) return st;
    const uint32_t n = s->n, C = s->gridC;
    if (cellOf) {
        SPH_HIP(sphdev::launch_gather_keys(s->keysSorted, s->ids[s->ic], s->stagingU, n, s->stream));
        SPH_HIP(hipMemcpyAsync(cellOf, s->stagingU, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
    }
    if (cellCounts || cellParticles) {
        std::vector<uint32_t> counts(C), starts(C);
        std::vector<uint32_t> ids(n);
        SPH_HIP(hipMemcpyAsync(counts.data(), s->counts, (size_t)C * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipMemcpyAsync(starts.data(), s->bucketStart, (size_t)C * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipMemcpyAsync(ids.data(), s->ids[s->ic], (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
        for (uint32_t b = 0; b < C; ++b) {
            const uint32_t cnt = counts[b];
            if (cellCounts) cellCounts[b] = cnt;
            if (cellParticles) {
                for (uint32_t k = 0; k < SPH_CELL_CAPACITY; ++k)
                    cellParticles[(size_t)b * SPH_CELL_CAPACITY + k] = k < cnt ? ids[starts[b] + k] : 0xFFFFFFFFu;
            }
        }
    }
    return SPH_OK;
}

extern "C" int sph_read_sorted(SphSolver* s, uint32_t* keys, uint32_t* ids, uint32_t* bucketStarts) {
    sphhost::clear_error();
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_sorted: null solver");
    if (!s->hasGrid) return sphhost::fail(SPH_ERR_STATE, "no grid: run a substep first");
    if (int st = set_device(s)) return st;
    if (keys) SPH_HIP(hipMemcpyAsync(keys, s->keysSorted, (size_t)s->n * 4, hipMemcpyDeviceToHost, s->stream));
    if (ids) SPH_HIP(hipMemcpyAsync(ids, s->ids[s->ic], (size_t)s->n * 4, hipMemcpyDeviceToHost, s->stream));
    if (bucketStarts) {
        const uint32_t C = s->gridC;
        std::vector<uint32_t> counts(C);
        SPH_HIP(hipMemcpyAsync(counts.data(), s->counts, (size_t)C * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipMemcpyAsync(bucketStarts, s->bucketStart, (size_t)C * 4, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
        for (uint32_t b = 0; b < C; ++b)
            if (counts[b] == 0) bucketStarts[b] = 0xFFFFFFFFu;
    }
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

extern "C" int sph_read_forces(SphSolver* s, float* force4) {
    sphhost::clear_error();
    if (!s || !force4) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_forces: null argument");
    if (!(s->flags & SPH_CFG_KEEP_FORCES)) return sphhost::fail(SPH_ERR_STATE, "solver created without SPH_CFG_KEEP_FORCES");
    if (!s->hasGrid) return sphhost::fail(SPH_ERR_STATE, "no forces: run a substep first");
    if (int st = set_device(s)) return st;
    SPH_HIP(sphdev::launch_unpermute(s->force, s->ids[s->ic], s->staging, s->n, s->stream));
    SPH_HIP(hipMemcpyAsync(force4, s->staging, (size_t)s->n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

extern "C" int sph_set_mode(SphSolver* s, uint32_t modeFlags) {
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_set_mode: null solver");
    if (modeFlags & ~(uint32_t)(SPH_MODE_UNCAPPED | SPH_MODE_DEDUPE))
        return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "unknown mode flag");
    s->mode = modeFlags;
    return SPH_OK;
}

extern "C" int sph_set_eos(SphSolver* s, uint32_t eos) {
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_set_eos: null solver");
    if (eos != SPH_EOS_LINEAR && eos != SPH_EOS_TAIT) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "unknown eos");
    s->eos = eos;
    return SPH_OK;
}

extern "C" int sph_conservation(SphSolver* s, const SphParams* p, double out[3]) {
    sphhost::clear_error();
    if (!s || !p || !out) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_conservation: null argument");
    std::vector<float> pos((size_t)s->n * 4), vel((size_t)s->n * 4);
    if (int st = sph_read_particles(s, pos.data(), vel.data())) return st;
    const double m = (double)p->particleMass;
    double ke = 0.0, pe = 0.0;
    for (uint32_t i = 0; i < s->n; ++i) {
        const float* x = &pos[4 * (size_t)i];
        const float* v = &vel[4 * (size_t)i];
        ke += 0.5 * m * ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
        pe -= m * ((double)p->gravity[0] * x[0] + (double)p->gravity[1] * x[1] + (double)p->gravity[2] * x[2]);
    }
    out[0] = m * (double)s->n;
    out[1] = ke;
    out[2] = pe;
    return SPH_OK;
}

extern "C" int sph_profile_enable(SphSolver* s, int enable) {
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_profile_enable: null solver");
    s->profiling = enable != 0;
    return SPH_OK;
}

extern "C" int sph_profile_read(SphSolver* s, const char** names, double* ms, uint32_t* launches, uint32_t capacity) {
    if (!s) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_profile_read: null solver");
    if (int st = set_device(s)) return st;
    SPH_HIP(hipStreamSynchronize(s->stream));
    for (const TimedLaunch& t : s->timed) {
        float el = 0.0f;
        if (hipEventElapsedTime(&el, t.begin, t.end) == hipSuccess) {
            s->accMs[t.cls] += el;
            s->accLaunches[t.cls] += 1;
        }
        s->eventPool.push_back(t.begin);
        s->eventPool.push_back(t.end);
    }
    s->timed.clear();
    uint32_t w = 0;
    for (uint32_t c = 0; c < KC_COUNT && w < capacity; ++c, ++w) {
        if (names) names[w] = kKernelNames[c];
        if (ms) ms[w] = s->accMs[c];
        if (launches) launches[w] = s->accLaunches[c];
    }
    for (uint32_t c = 0; c < KC_COUNT; ++c) {
        s->accMs[c] = 0.0;
        s->accLaunches[c] = 0;
    }
    return (int)w;
}

// Test hook: the per-substep constants the kernels receive, in a fixed order
// {h2, r2Cut, cPoly6, cSpiky, cVisc, maxSpeed, negMass, viscMass}.
extern "C" int sph_debug_step_constants(const SphParams* p, float dt, float out[8]) {
    if (!p || !out) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "null argument");
    SphSolver tmp;
    tmp.n = 0;
    const StepConsts k = make_consts(&tmp, *p, dt);
    out[0] = k.h2;
    out[1] = k.r2Cut;
    out[2] = k.cPoly6;
    out[3] = k.cSpiky;
    out[4] = k.cVisc;
    out[5] = k.maxSpeed;
    out[6] = k.negMass;
    out[7] = k.viscMass;
    return SPH_OK;
}

// Profiling aid: one plain device-to-device copy of `bytes` bytes by a kernel, so that
// hardware byte counters can be scaled against a transfer of known size.
extern "C" int sph_debug_copy(uint64_t bytes) {
    float4* a = nullptr;
    float4* b = nullptr;
    const size_t n = (size_t)(bytes / sizeof(float4));
    SPH_HIP(hipMalloc(&a, n * sizeof(float4)));
    SPH_HIP(hipMalloc(&b, n * sizeof(float4)));
    hipMemset(a, 0, n * sizeof(float4));
    hipError_t e = sphdev::launch_calibration_copy(a, b, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "sph_debug_copy");
    return SPH_OK;
}
