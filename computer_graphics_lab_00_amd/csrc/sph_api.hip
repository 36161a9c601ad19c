// sph_api.hip — the C ABI: the solver handle (physics::FluidSolver,
// src/physics/fluid_sph.h:101-127) is an engine with one scene; the batch handle is an
// engine with many. Both advance on one HIP stream with hand-written gfx950 kernels.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "sph_engine.h"
#include "sph_internal.h"

using sphhost::Engine;
using sphhost::fail;

namespace {

int conservation_of(Engine& e, uint32_t scene, const SphParams* p, double out[3]) {
    const uint32_t n = e.sc[scene].n;
    std::vector<float> pos((size_t)n * 4), vel((size_t)n * 4);
    if (int st = e.read_particles(scene, pos.data(), vel.data())) return st;
    const double m = (double)p->particleMass;
    double ke = 0.0, pe = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const float* x = &pos[4 * (size_t)i];
        const float* v = &vel[4 * (size_t)i];
        ke += 0.5 * m * ((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
        pe -= m * ((double)p->gravity[0] * x[0] + (double)p->gravity[1] * x[1] + (double)p->gravity[2] * x[2]);
    }
    out[0] = m * (double)n;
    out[1] = ke;
    out[2] = pe;
    return SPH_OK;
}

}  // namespace

// ---- solver ----

extern "C" int sph_create(const SphSolverConfig* cfg, SphSolver** out) {
    sphhost::clear_error();
    if (!cfg || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_create: null argument");
    *out = nullptr;
    SphSolver* s = new SphSolver();
    if (int st = s->e.create({{cfg->maxParticles, cfg->maxCellCount}}, cfg->device, cfg->flags)) {
        delete s;
        return st;
    }
    *out = s;
    return SPH_OK;
}

extern "C" int sph_destroy(SphSolver* s) {
    delete s;
    return SPH_OK;
}

extern "C" int sph_emit(SphSolver* s, const SphParams* p, uint32_t count) {
    sphhost::clear_error();
    if (!s || !p) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_emit: null argument");
    return s->e.emit(0, *p, count);
}

extern "C" int sph_load_state(SphSolver* s, const float* pos4, const float* vel4, uint32_t count) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_load_state: null solver");
    return s->e.load_state(0, pos4, vel4, count);
}

extern "C" int sph_step(SphSolver* s, const SphParams* p, float deltaSeconds) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_step: null solver");
    // FluidSolver::step (src/physics/fluid_sph.cpp:311-321): silent no-op when empty or dt <= 0.
    if (s->e.sc[0].n == 0 || deltaSeconds <= 0.0f) return SPH_OK;
    if (!p) return fail(SPH_ERR_INVALID_ARGUMENT, "params is null");
    if (int st = s->e.set_params(0, *p)) return st;
    return s->e.step(deltaSeconds, false);
}

extern "C" int sph_substep(SphSolver* s, const SphParams* p, float dt) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep: null solver");
    if (!p) return fail(SPH_ERR_INVALID_ARGUMENT, "params is null");
    if (int st = s->e.set_params(0, *p)) return st;
    if (s->e.sc[0].n == 0) return SPH_OK;
    return s->e.step(dt, true);
}

extern "C" int sph_synchronize(SphSolver* s) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_synchronize: null solver");
    return s->e.synchronize();
}

extern "C" int sph_max_bin_occupancy(SphSolver* s, uint32_t* out) {
    sphhost::clear_error();
    if (!s || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_max_bin_occupancy: null argument");
    if (int st = s->e.synchronize()) return st;
    *out = s->e.last_bin_max();
    return SPH_OK;
}

extern "C" int sph_short_division(SphSolver* s, uint32_t* out) {
    sphhost::clear_error();
    if (!s || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_short_division: null argument");
    return s->e.short_division(0, out);
}

extern "C" uint32_t sph_particle_count(const SphSolver* s) { return s ? s->e.sc[0].n : 0u; }

extern "C" int sph_read_particles(SphSolver* s, float* pos4, float* vel4) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_particles: null solver");
    return s->e.read_particles(0, pos4, vel4);
}

extern "C" int sph_read_previous(SphSolver* s, float* pos4) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_previous: null solver");
    return s->e.read_previous(0, pos4);
}

extern "C" int sph_keep_rendered(SphSolver* s) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_keep_rendered: null solver");
    return s->e.keep_rendered(0);
}

extern "C" int sph_read_grid(SphSolver* s, uint32_t* cellOf, uint32_t* cellCounts, uint32_t* cellParticles) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_grid: null solver");
    return s->e.read_grid(0, cellOf, cellCounts, cellParticles);
}

extern "C" int sph_read_sorted(SphSolver* s, uint32_t* keys, uint32_t* ids, uint32_t* bucketStarts) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_sorted: null solver");
    return s->e.read_sorted(0, keys, ids, bucketStarts);
}

extern "C" int sph_read_forces(SphSolver* s, float* force4) {
    sphhost::clear_error();
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_read_forces: null solver");
    return s->e.read_forces(0, force4);
}

extern "C" int sph_set_mode(SphSolver* s, uint32_t modeFlags) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_set_mode: null solver");
    if (modeFlags & ~(uint32_t)(SPH_MODE_UNCAPPED | SPH_MODE_DEDUPE | SPH_MODE_FAST_MATH)) return fail(SPH_ERR_INVALID_ARGUMENT, "unknown mode flag");
    s->e.sc[0].mode = modeFlags;
    return SPH_OK;
}

extern "C" int sph_set_eos(SphSolver* s, uint32_t eos) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_set_eos: null solver");
    if (eos != SPH_EOS_LINEAR && eos != SPH_EOS_TAIT) return fail(SPH_ERR_INVALID_ARGUMENT, "unknown eos");
    s->e.sc[0].eos = eos;
    return SPH_OK;
}

extern "C" int sph_conservation(SphSolver* s, const SphParams* p, double out[3]) {
    sphhost::clear_error();
    if (!s || !p || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_conservation: null argument");
    return conservation_of(s->e, 0, p, out);
}

extern "C" int sph_profile_enable(SphSolver* s, int enable) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_profile_enable: null solver");
    s->e.profiling = enable != 0;
    return SPH_OK;
}

extern "C" int sph_profile_read(SphSolver* s, const char** names, double* ms, uint32_t* launches, uint32_t capacity) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_profile_read: null solver");
    return s->e.profile_read(names, ms, launches, capacity);
}

// ---- batch ----

extern "C" int sph_batch_create(const SphSceneCapacity* scenes, uint32_t sceneCount, int32_t device, uint32_t flags,
                                SphBatch** out) {
    sphhost::clear_error();
    if (!scenes || !out || sceneCount == 0) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_create: bad argument");
    *out = nullptr;
    std::vector<std::pair<uint32_t, uint32_t>> caps;
    for (uint32_t i = 0; i < sceneCount; ++i) caps.push_back({scenes[i].maxParticles, scenes[i].maxCellCount});
    SphBatch* b = new SphBatch();
    if (int st = b->e.create(caps, device, flags)) {
        delete b;
        return st;
    }
    *out = b;
    return SPH_OK;
}

extern "C" int sph_batch_destroy(SphBatch* b) {
    delete b;
    return SPH_OK;
}

extern "C" uint32_t sph_batch_scene_count(const SphBatch* b) { return b ? (uint32_t)b->e.sc.size() : 0u; }

extern "C" int sph_batch_set_params(SphBatch* b, uint32_t scene, const SphParams* p) {
    sphhost::clear_error();
    if (!b || !p) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_set_params: null argument");
    return b->e.set_params(scene, *p);
}

extern "C" int sph_batch_set_layout(SphBatch* b, uint32_t scene, const uint32_t tileLog2[3], uint32_t neighbourList) {
    sphhost::clear_error();
    if (!b || !tileLog2) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_set_layout: null argument");
    if (scene >= b->e.sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (tileLog2[0] + tileLog2[1] + tileLog2[2] != 6u || tileLog2[0] > 6 || tileLog2[1] > 6 || tileLog2[2] > 6)
        return fail(SPH_ERR_INVALID_ARGUMENT, "tile must hold 64 bins (log2 sizes summing to 6)");
    for (int a = 0; a < 3; ++a) b->e.sc[scene].tileLog2[a] = tileLog2[a];
    b->e.sc[scene].list = neighbourList != 0;
    return SPH_OK;
}

extern "C" int sph_batch_set_schedule(SphBatch* b, const uint32_t* order, uint32_t count) {
    sphhost::clear_error();
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_set_schedule: null batch");
    return b->e.set_schedule(order, count);
}

extern "C" int sph_batch_set_mode(SphBatch* b, uint32_t scene, uint32_t modeFlags, uint32_t eos) {
    sphhost::clear_error();
    if (!b || scene >= b->e.sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_set_mode: bad argument");
    if (modeFlags & ~(uint32_t)(SPH_MODE_UNCAPPED | SPH_MODE_DEDUPE | SPH_MODE_FAST_MATH)) return fail(SPH_ERR_INVALID_ARGUMENT, "unknown mode flag");
    if (eos != SPH_EOS_LINEAR && eos != SPH_EOS_TAIT) return fail(SPH_ERR_INVALID_ARGUMENT, "unknown eos");
    b->e.sc[scene].mode = modeFlags;
    b->e.sc[scene].eos = eos;
    return SPH_OK;
}

extern "C" int sph_batch_emit(SphBatch* b, uint32_t scene, uint32_t count) {
    sphhost::clear_error();
    if (!b || scene >= b->e.sc.size()) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_emit: bad argument");
    if (!b->e.sc[scene].hasParams) return fail(SPH_ERR_STATE, "set the scene's params first");
    return b->e.emit(scene, b->e.sc[scene].params, count);
}

extern "C" int sph_batch_load_state(SphBatch* b, uint32_t scene, const float* pos4, const float* vel4, uint32_t count) {
    sphhost::clear_error();
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_load_state: null batch");
    return b->e.load_state(scene, pos4, vel4, count);
}

extern "C" int sph_batch_step(SphBatch* b, float deltaSeconds) {
    sphhost::clear_error();
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_step: null batch");
    if (deltaSeconds <= 0.0f) return SPH_OK;
    return b->e.step(deltaSeconds, false);
}

extern "C" int sph_batch_synchronize(SphBatch* b) {
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_synchronize: null batch");
    return b->e.synchronize();
}

extern "C" int sph_batch_max_bin_occupancy(SphBatch* b, uint32_t* out) {
    sphhost::clear_error();
    if (!b || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_max_bin_occupancy: null argument");
    if (int st = b->e.synchronize()) return st;
    *out = b->e.last_bin_max();
    return SPH_OK;
}

extern "C" uint32_t sph_batch_particle_count(const SphBatch* b, uint32_t scene) {
    return (b && scene < b->e.sc.size()) ? b->e.sc[scene].n : 0u;
}

extern "C" int sph_batch_read_particles(SphBatch* b, uint32_t scene, float* pos4, float* vel4) {
    sphhost::clear_error();
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_read_particles: null batch");
    return b->e.read_particles(scene, pos4, vel4);
}

extern "C" int sph_batch_profile_enable(SphBatch* b, int enable) {
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_profile_enable: null batch");
    b->e.profiling = enable != 0;
    return SPH_OK;
}

extern "C" int sph_batch_profile_read(SphBatch* b, const char** names, double* ms, uint32_t* launches,
                                      uint32_t capacity) {
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_batch_profile_read: null batch");
    return b->e.profile_read(names, ms, launches, capacity);
}

// ---- test and profiling hooks ----

// The per-substep constants the kernels receive, in a fixed order
// {h2, r2Cut, cPoly6, cSpiky, cVisc, maxSpeed, negMass, viscMass}.
extern "C" int sph_debug_step_constants(const SphParams* p, float dt, float out[8]) {
    if (!p || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "null argument");
    sphhost::SceneHost tmp;
    const sphdev::StepConsts k = sphhost::make_consts(tmp, *p, dt, false);
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

// One plain device-to-device copy of `bytes` bytes by a kernel, so hardware byte counters
// can be scaled against a transfer of known size.
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
    if (e != hipSuccess) return sphhost::hip_fail(e, "sph_debug_copy");
    return SPH_OK;
}
