// sph_host.cpp — stateless host-side pieces of the C ABI: parameter derivation, the
// substep rule, the cell-count rule, the hash, seeded jitter and the last-error string.
#include <math.h>
#include <string.h>

#include <string>

#include "sph_c_api.h"
#include "sph_internal.h"

namespace {
thread_local std::string g_lastError;
}  // namespace

namespace sphhost {

int fail(int status, const std::string& message) {
    g_lastError = message;
    return status;
}

void clear_error() { g_lastError.clear(); }

}  // namespace sphhost

extern "C" const char* sph_last_error(void) { return g_lastError.c_str(); }

// physics::deriveFluidParams (src/physics/fluid_sph.cpp:65-100).
extern "C" int sph_derive_params(const SphSettings* s, const float emitterWorld[16], uint32_t particleCount,
                                 uint32_t cellCount, const SphCollider* colliders, uint32_t colliderCount,
                                 const SphForceField* fields, uint32_t fieldCount, SphParams* out) {
    if (!s || !emitterWorld || !out) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_derive_params: null argument");
    memset(out, 0, sizeof(*out));
    const float spacing = s->particleRadius * 2.0f;
    memcpy(out->emitterWorld, emitterWorld, sizeof(out->emitterWorld));
    memcpy(out->emitterHalfExtents, s->emitterHalfExtents, sizeof(out->emitterHalfExtents));
    out->spacing = spacing;
    memcpy(out->containerMin, s->containerMin, sizeof(out->containerMin));
    memcpy(out->containerMax, s->containerMax, sizeof(out->containerMax));
    out->particleRadius = s->particleRadius;
    out->wallRestitution = 0.3f;
    memcpy(out->gravity, s->gravity, sizeof(out->gravity));
    out->particleMass = s->restDensity * spacing * spacing * spacing;
    out->smoothingRadius = s->smoothingRadius < spacing ? spacing : s->smoothingRadius;
    out->restDensity = s->restDensity;
    out->stiffness = s->stiffness;
    out->viscosity = s->viscosity;
    out->cellCount = cellCount;
    uint32_t along[2];
    for (int a = 0; a < 2; ++a) {
        uint32_t v = (uint32_t)floorf(s->emitterHalfExtents[a] * 2.0f / spacing);
        along[a] = v > 1u ? v : 1u;
    }
    const uint32_t layer = along[0] * along[1];
    const uint32_t nz = (particleCount + layer - 1u) / layer;
    out->lattice[0] = along[0];
    out->lattice[1] = along[1];
    out->lattice[2] = nz > 1u ? nz : 1u;
    // collectShapeColliders skips mesh colliders and stops at 8 (fluid_sph.cpp:102-126).
    for (uint32_t i = 0; colliders && i < colliderCount && out->colliderCount < SPH_MAX_COLLIDERS; ++i) {
        if (colliders[i].shape == SPH_SHAPE_MESH) continue;
        out->colliders[out->colliderCount++] = colliders[i];
    }
    for (uint32_t i = 0; fields && i < fieldCount && out->fieldCount < SPH_MAX_FORCE_FIELDS; ++i) {
        out->fields[out->fieldCount++] = fields[i];
    }
    return SPH_OK;
}

// physics::fluidSubsteps (src/physics/fluid_sph.cpp:128-136).
extern "C" uint32_t sph_substeps(const SphParams* p, float deltaSeconds) {
    if (!p) return 1u;
    const float h = p->smoothingRadius;
    const float glen = sqrtf((p->gravity[0] * p->gravity[0] + p->gravity[1] * p->gravity[1]) + p->gravity[2] * p->gravity[2]);
    const float gravity = glen < 0.1f ? 0.1f : glen;
    const float k = p->stiffness < 1.0f ? 1.0f : p->stiffness;
    const float byPressure = 0.4f * h / sqrtf(k);
    const float byGravity = 0.25f * sqrtf(h / gravity);
    const float stable = byGravity < byPressure ? byGravity : byPressure;
    const float dt = deltaSeconds < 0.0f ? 0.0f : deltaSeconds;
    const float frameStep = SPH_MAX_FRAME_STEP < dt ? SPH_MAX_FRAME_STEP : dt;
    const float denom = stable < 1e-5f ? 1e-5f : stable;
    uint32_t steps = (uint32_t)ceilf(frameStep / denom);
    return steps < 1u ? 1u : (steps > 8u ? 8u : steps);
}

// gfx::cellCountFor (src/gfx/fluid.cpp:50-52).
extern "C" uint32_t sph_cell_count_for(uint32_t particleCount) {
    const uint32_t v = (particleCount > 1u ? particleCount : 1u) * 2u;
    uint32_t c = 1u;
    while (c < v && c != 0x80000000u) c <<= 1;
    return c < 1024u ? 1024u : (c > 65536u ? 65536u : c);
}

// physics::spatialHash (src/physics/spatial_hash.h:16-20).
extern "C" uint32_t sph_hash_cell(int32_t cx, int32_t cy, int32_t cz, uint32_t cellCount) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & (cellCount - 1u);
}

// physics::spatialCell (src/physics/spatial_hash.h:12-14).
extern "C" void sph_spatial_cell(const float position[3], float cellSize, int32_t outCell[3]) {
    for (int a = 0; a < 3; ++a) outCell[a] = (int32_t)floorf(position[a] / cellSize);
}

extern "C" void sph_lattice_jitter(float* pos4, uint32_t count, uint32_t seed, float amplitude) {
    if (!pos4) return;
    uint32_t state = seed;
    const float span = 2.0f * amplitude;
    for (uint32_t i = 0; i < count; ++i) {
        for (int a = 0; a < 3; ++a) {
            state = state * 1664525u + 1013904223u;
            const float u = (float)(state >> 8) / 16777216.0f;
            pos4[4 * (size_t)i + a] += (u - 0.5f) * span;
        }
    }
}
