// sph_engine.h — the host engine behind both C-ABI handles: S independent scenes in one
// set of device arrays, advanced together by batched launches on one HIP stream, with
// the per-step launch sequence captured once as a hipGraph and replayed.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "sph_device.h"

namespace sphhost {

struct SceneHost {
    uint32_t maxN = 0, maxC = 0, binCap = 0;
    uint32_t offset = 0, cellOffset = 0, binOffset = 0;   // fixed at creation (capacity based)
    uint32_t n = 0;
    SphParams params{};
    bool hasParams = false;
    uint32_t mode = SPH_MODE_REFERENCE;
    uint32_t eos = SPH_EOS_LINEAR;
    uint32_t tileLog2[3] = {2, 2, 2};
    bool list = true;
};

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

struct TimedLaunch {
    uint32_t cls;
    hipEvent_t begin;
    hipEvent_t end;
};

constexpr uint32_t kMaxSubsteps = 8;

class Engine {
public:
    ~Engine();
    // capacities: {maxParticles, maxCellCount} per scene
    int create(const std::vector<std::pair<uint32_t, uint32_t>>& capacities, int device, uint32_t flags);

    int emit(uint32_t scene, const SphParams& p, uint32_t count);
    int load_state(uint32_t scene, const float* pos4, const float* vel4, uint32_t count);
    int set_params(uint32_t scene, const SphParams& p);
    // Order of the scenes' workgroups in every batched launch (a permutation of the scenes;
    // count 0 restores scene order). Results do not depend on it.
    int set_schedule(const uint32_t* order, uint32_t count);
    // Every scene advances by dt with its own substep rule (FluidSolver::step); when
    // `singleSubstep` the scenes advance one substep of exactly dt.
    int step(float dt, bool singleSubstep);
    int synchronize();

    int read_particles(uint32_t scene, float* pos4, float* vel4);
    int read_previous(uint32_t scene, float* pos4);
    int keep_rendered(uint32_t scene);
    int read_grid(uint32_t scene, uint32_t* cellOf, uint32_t* cellCounts, uint32_t* cellParticles);
    int read_sorted(uint32_t scene, uint32_t* keys, uint32_t* ids, uint32_t* bucketStarts);
    int read_forces(uint32_t scene, float* force4);
    // 1 if the last substep of `scene` ran the force pass's short division, else 0.
    int short_division(uint32_t scene, uint32_t* out);

    int profile_read(const char** names, double* ms, uint32_t* launches, uint32_t capacity);

    int set_device() const;

    // ---- hooks of the learned surrogate ----
    // Grid build + density pass of one substep of length dt on the current state (with
    // the neighbourhood sums when `sums` is set). The sorted state is left in the back
    // buffers; nothing is committed.
    int prepare(float dt, const sphdev::NeighbourSums* sums);
    // Grid build only (keys, sort, bucket starts) on the current state: the sorted state
    // and the buckets describe the current positions. Nothing is committed.
    int build_grid();
    const uint2* bucket_table() const { return buckets; }   // (first slot, member count) per bucket
    float4* sorted_pos() { return pos[1 - pc]; }
    float4* sorted_vel() { return vel[1 - pc]; }
    uint32_t* sorted_ids() { return ids[1 - ic]; }
    float4* state_pos() { return pos[pc]; }
    float4* state_vel() { return vel[pc]; }
    uint32_t* state_ids() { return ids[ic]; }
    float4* force_buffer() { return force; }
    const float4* previous_pos() const { return prev; }   // the kept rendered positions, by index
    // After a kernel wrote the state buffers from the prepared (sorted) ones.
    void commit_sorted() {
        ic = 1 - ic;
        hasGrid = true;
        touch();
    }
    // A value unique in the process that changes whenever the state or the prepared
    // (sorted) buffers change: equal epochs mean nothing was written in between.
    uint64_t epoch() const { return epoch_; }
    void touch();
    sphdev::Launch plan_launch() const { return launch_for(0); }
    uint32_t total_slots() const { return totalN; }
    int ensure_force_buffer();
    uint32_t planned_blocks() const { return planBlocks; }
    // Largest bin occupancy of the last grid build the device has finished (0 before any).
    uint32_t last_bin_max() const { return binMaxHost ? __atomic_load_n(binMaxHost, __ATOMIC_RELAXED) : 0u; }

    std::vector<SceneHost> sc;
    std::vector<uint32_t> schedule;   // empty: scene order
    int device = 0;
    uint32_t flags = 0;
    bool profiling = false;
    hipStream_t stream = nullptr;

private:
    int ensure_plan(float dt, bool singleSubstep, uint32_t* maxSub);
    int run_substep(uint32_t j, bool prepareOnly = false, const sphdev::NeighbourSums* sums = nullptr,
                    bool gridOnly = false);
    int enqueue_step(uint32_t maxSub);
    int check_occupancy();
    void reset_occupancy();
    void invalidate();
    sphdev::Launch launch_for(uint32_t j) const;

    uint64_t epoch_ = 0;
    uint32_t totalN = 0, totalC = 0, totalBins = 0, maxBlocks = 0;
    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    uint32_t* ids[2] = {nullptr, nullptr};
    int pc = 0, ic = 0;
    uint32_t *keys = nullptr, *rep = nullptr;
    uint2* buckets = nullptr;   // (first slot, member count) per bucket
    uint32_t *binStart = nullptr, *binOf = nullptr, *slots = nullptr, *blockSums = nullptr, *keysSorted = nullptr;
    uint2* pairKI = nullptr;
    uint32_t* pairSrc = nullptr;
    uint4* nbrList = nullptr;
    uint32_t* nbrCount = nullptr;
    float4 *prev = nullptr, *force = nullptr, *staging = nullptr;
    uint32_t* stagingU = nullptr;
    // plan: per-substep scene tables, boundaries, block -> scene
    sphdev::SceneDesc* dTables = nullptr;
    sphdev::BoundaryConsts* dBounds = nullptr;
    uint16_t* dBlockScene = nullptr;
    uint16_t* dTileScene = nullptr;   // scene of each half-tile workgroup of the tiled density pass
    uint32_t maxTiles = 0, planTiles = 0;
    uint2* restList = nullptr;        // particles the tiled density pass leaves to its global walk
    uint32_t* restCount = nullptr;
    uint32_t* binMax = nullptr;         // per bin-slot workgroup: its largest bin occupancy
    uint32_t* binMaxHost = nullptr;     // its copy in mapped host memory (host address)
    uint32_t* binMaxHostDev = nullptr;  // the same memory's device address
    uint32_t* rangeGuard = nullptr;     // nonzero: some particle's state is outside the short division's ranges
    void* hostStage = nullptr;
    size_t hostStageBytes = 0;
    std::vector<uint8_t> planKey;
    uint32_t planSub = 0, planBlocks = 0, planBins = 0;
    float planDt = 1.0f / 600.0f;   // arguments of the last plan, re-used by build_grid
    bool planSingle = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool hasGrid = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> eventPool;
    double accMs[KC_COUNT] = {};
    uint32_t accLaunches[KC_COUNT] = {};

    friend struct Timer;
    hipEvent_t take_event();
    void record(uint32_t cls, hipEvent_t b, hipEvent_t e) { timed.push_back({cls, b, e}); }
    void free_all();
};

// Host-side helpers shared with the batch and surrogate code.
sphdev::StepConsts make_consts(const SceneHost& s, const SphParams& p, float dt, bool hasList);
sphdev::BinGrid make_bins(const SceneHost& s, const SphParams& p);
int hip_fail(hipError_t e, const char* what);

}  // namespace sphhost

// The C-ABI handles are engines: a solver has one scene, a batch many.
struct SphSolver {
    sphhost::Engine e;
    // device buffers of the surface extraction, grown on demand and kept with the solver
    float* surfField = nullptr;
    size_t surfFieldBytes = 0;
    SphSurfaceVertex* surfVerts = nullptr;
    uint32_t surfVertCap = 0;
    uint32_t* surfMarch = nullptr;   // the marching passes' counters, offsets and scan sums
    size_t surfMarchWords = 0;
    // device staging of the renderer export, grown on demand
    SphGpuInstance* instStage = nullptr;
    SphTlasInstance* tlasStage = nullptr;
    uint32_t instCap = 0, tlasCap = 0;

    ~SphSolver() {
        if (surfField || surfVerts || surfMarch || instStage || tlasStage) {
            hipSetDevice(e.device);
            hipDeviceSynchronize();
            hipFree(surfField);
            hipFree(surfVerts);
            hipFree(surfMarch);
            hipFree(instStage);
            hipFree(tlasStage);
        }
    }
};

struct SphBatch {
    sphhost::Engine e;
};

#define SPH_HIP(call)                                                     \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return ::sphhost::hip_fail(e_, #call);      \
    } while (0)
