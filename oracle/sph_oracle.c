/*
 * sph_oracle.c — CPU oracle of the SPH fluid step. TEST INFRASTRUCTURE ONLY: loaded by
 * the test suite, the smoke check and the benchmark's CPU-baseline leg, never by the
 * product library.
 *
 * Restates src/physics/fluid_sph.cpp, src/physics/spatial_hash.h,
 * src/physics/collider_shapes.h:33-131, src/physics/force_field.h:31-70 and
 * src/gfx/fluid.cpp:50-52 in C. Build: -O2 -ffp-contract=off, no fast-math (x86-64 SSE,
 * so every float operation rounds to binary32 like the reference's).
 *
 * glm (header-only, not under the reference tree; version resolved by vcpkg at
 * configure time) fixes the evaluation order of the vector expressions. This file
 * follows glm's scalar code path as published:
 *   dot(a, b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   length(v)   = sqrt(dot(v, v))
 *   v / s       = (v.x/s, v.y/s, v.z/s)        (IEEE division per component)
 *   normalize(v)= v * (1 / sqrt(dot(v, v)))
 *   mat4 * vec4 = (m[0]*x + m[1]*y) + (m[2]*z + m[3]*w)
 *   mat3 * vec3 = (m[0][r]*x + m[1][r]*y) + m[2][r]*z
 *   cross(a, b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *
 * What pins each function (tests/test_oracle.py):
 *   orc_hash, orc_cell          known answers from the integer closed form; brute force
 *   grid build                  invariants (sum of counts, ascending slots, 32-cap) and a
 *                               brute-force re-bucketing
 *   poly6 / spiky / viscosity   closed forms (W(0), zero at r >= h, FP64 quadrature of
 *                               the normalisation)
 *   density pass                O(N^2) brute force with a large cellCount; the C = 1
 *                               duplicate-bucket closed form (27 x the sum); the 32-cap
 *   force pass                  two-particle closed form; zero viscosity for equal
 *                               velocities; FP64 momentum sum ~ 0
 *   integrate / walls           free fall closed form, speed clamp, -0.3 v wall bounce
 *   step                        tests/fluid_test.cpp:53-207 invariants, serial == threaded
 *   colliders                   per-shape closest-point closed forms, fluid_test collider case
 *   force fields                per-type closed forms
 *   derive / substeps / emit    tests/fluid_test.cpp:53-75, 181-207 known answers
 * Parity unpinned at the bit level (no reference binary can be built here): glm's
 * expression order for the mat4 paths (emit with a rotated emitter, colliders) and
 * libm powf (kernel constants, force-field falloff), which the product shares through
 * the same host libm.
 */
#define _GNU_SOURCE
#include "sph_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* std::numbers::pi_v<float> (fluid_sph.cpp:20) */
static const float ORC_PI_F = 3.14159265358979323846f;

/* ------------------------------------------------------------------------------------
 * Thread pool: the JobSystem::parallelFor shape (src/core/job_system.cpp:80-119):
 * chunks of 256 particles, the caller participates, serial when count <= 256 or no
 * pool (fluid_sph.cpp:55-61). Results never depend on the thread count.
 * ---------------------------------------------------------------------------------- */

typedef void (*orc_body_fn)(void* ctx, uint32_t begin, uint32_t end);

#define ORC_GRANULARITY 256u

typedef struct {
    pthread_t* threads;
    uint32_t workers;
    pthread_mutex_t lock;
    pthread_cond_t wake;
    pthread_cond_t done;
    uint64_t generation;
    int quit;
    orc_body_fn body;
    void* ctx;
    uint32_t count;
    atomic_uint next;
    uint32_t busy;
} OrcPool;

static OrcPool g_pool;
static pthread_mutex_t g_pool_guard = PTHREAD_MUTEX_INITIALIZER;

static void orc_pool_drain(OrcPool* pool) {
    for (;;) {
        uint32_t chunk = atomic_fetch_add(&pool->next, 1u);
        uint64_t begin = (uint64_t)chunk * ORC_GRANULARITY;
        if (begin >= pool->count) break;
        uint64_t end = begin + ORC_GRANULARITY;
        if (end > pool->count) end = pool->count;
        pool->body(pool->ctx, (uint32_t)begin, (uint32_t)end);
    }
}

static void* orc_pool_worker(void* arg) {
    OrcPool* pool = (OrcPool*)arg;
    uint64_t seen = 0;
    for (;;) {
        pthread_mutex_lock(&pool->lock);
        while (pool->generation == seen && !pool->quit) pthread_cond_wait(&pool->wake, &pool->lock);
        if (pool->quit) {
            pthread_mutex_unlock(&pool->lock);
            return NULL;
        }
        seen = pool->generation;
        pthread_mutex_unlock(&pool->lock);
        orc_pool_drain(pool);
        pthread_mutex_lock(&pool->lock);
        if (--pool->busy == 0) pthread_cond_signal(&pool->done);
        pthread_mutex_unlock(&pool->lock);
    }
}

static void orc_pool_resize(uint32_t workers) {
    OrcPool* pool = &g_pool;
    if (pool->threads && pool->workers == workers) return;
    if (pool->threads) {
        pthread_mutex_lock(&pool->lock);
        pool->quit = 1;
        pthread_cond_broadcast(&pool->wake);
        pthread_mutex_unlock(&pool->lock);
        for (uint32_t i = 0; i < pool->workers; ++i) pthread_join(pool->threads[i], NULL);
        free(pool->threads);
        pthread_mutex_destroy(&pool->lock);
        pthread_cond_destroy(&pool->wake);
        pthread_cond_destroy(&pool->done);
        memset(pool, 0, sizeof(*pool));
    }
    if (workers == 0) return;
    pthread_mutex_init(&pool->lock, NULL);
    pthread_cond_init(&pool->wake, NULL);
    pthread_cond_init(&pool->done, NULL);
    pool->workers = workers;
    pool->threads = (pthread_t*)calloc(workers, sizeof(pthread_t));
    for (uint32_t i = 0; i < workers; ++i) pthread_create(&pool->threads[i], NULL, orc_pool_worker, pool);
}

/* threads = total threads including the caller; 0 or 1 runs serially. */
static void orc_parallel_for(uint32_t threads, uint32_t count, orc_body_fn body, void* ctx) {
    if (count == 0) return;
    if (threads <= 1 || count <= ORC_GRANULARITY) {
        body(ctx, 0, count);
        return;
    }
    pthread_mutex_lock(&g_pool_guard);
    orc_pool_resize(threads - 1);
    OrcPool* pool = &g_pool;
    pthread_mutex_lock(&pool->lock);
    pool->body = body;
    pool->ctx = ctx;
    pool->count = count;
    atomic_store(&pool->next, 0u);
    pool->busy = pool->workers;
    pool->generation++;
    pthread_cond_broadcast(&pool->wake);
    pthread_mutex_unlock(&pool->lock);
    orc_pool_drain(pool);
    pthread_mutex_lock(&pool->lock);
    while (pool->busy != 0) pthread_cond_wait(&pool->done, &pool->lock);
    pthread_mutex_unlock(&pool->lock);
    pthread_mutex_unlock(&g_pool_guard);
}

uint32_t orc_hardware_threads(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (uint32_t)n : 1u;
}

/* ------------------------------------------------------------------------------------
 * Stateless pieces
 * ---------------------------------------------------------------------------------- */

/* spatialHash (spatial_hash.h:16-20): uint32 wraparound products, xor, mask. */
uint32_t orc_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t cellCount) {
    uint32_t hashed = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return hashed & (cellCount - 1u);
}

/* spatialCell (spatial_hash.h:12-14): ivec3(floor(p / h)), IEEE division per component. */
void orc_cell(const float p[3], float h, int32_t out[3]) {
    for (int a = 0; a < 3; ++a) out[a] = (int32_t)floorf(p[a] / h);
}

/* fluidSubsteps (fluid_sph.cpp:128-136). */
uint32_t orc_substeps(const SphParams* params, float dt) {
    float h = params->smoothingRadius;
    const float* g = params->gravity;
    float glen = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    float gravity = glen < 0.1f ? 0.1f : glen;                              /* std::max(len, 0.1) */
    float k = params->stiffness < 1.0f ? 1.0f : params->stiffness;         /* std::max(k, 1) */
    float a = 0.4f * h / sqrtf(k);
    float b = 0.25f * sqrtf(h / gravity);
    float stable = b < a ? b : a;                                          /* std::min(a, b) */
    float clampedDt = dt < 0.0f ? 0.0f : dt;                               /* std::max(dt, 0) */
    float frameStep = SPH_MAX_FRAME_STEP < clampedDt ? SPH_MAX_FRAME_STEP : clampedDt;
    float denom = stable < 1e-5f ? 1e-5f : stable;
    uint32_t steps = (uint32_t)ceilf(frameStep / denom);
    if (steps < 1u) steps = 1u;
    if (steps > 8u) steps = 8u;
    return steps;
}

/* cellCountFor (src/gfx/fluid.cpp:50-52). */
uint32_t orc_cell_count_for(uint32_t n) {
    uint32_t v = (n > 1u ? n : 1u) * 2u;
    uint32_t c = 1u;
    while (c < v && c != 0x80000000u) c <<= 1;
    if (c < 1024u) c = 1024u;
    if (c > 65536u) c = 65536u;
    return c;
}

/* deriveFluidParams (fluid_sph.cpp:65-100). */
void orc_derive_params(const SphSettings* st, const float world[16], uint32_t n, uint32_t cellCount,
                       const SphCollider* colliders, uint32_t nc, const SphForceField* fields, uint32_t nf,
                       SphParams* out) {
    memset(out, 0, sizeof(*out));
    float spacing = st->particleRadius * 2.0f;
    memcpy(out->emitterWorld, world, 16 * sizeof(float));
    memcpy(out->emitterHalfExtents, st->emitterHalfExtents, 3 * sizeof(float));
    out->spacing = spacing;
    memcpy(out->containerMin, st->containerMin, 3 * sizeof(float));
    memcpy(out->containerMax, st->containerMax, 3 * sizeof(float));
    out->particleRadius = st->particleRadius;
    out->wallRestitution = 0.3f;
    memcpy(out->gravity, st->gravity, 3 * sizeof(float));
    out->particleMass = st->restDensity * spacing * spacing * spacing;
    out->smoothingRadius = st->smoothingRadius < spacing ? spacing : st->smoothingRadius;
    out->restDensity = st->restDensity;
    out->stiffness = st->stiffness;
    out->viscosity = st->viscosity;
    out->cellCount = cellCount;
    uint32_t lat[2];
    for (int a = 0; a < 2; ++a) {
        uint32_t v = (uint32_t)floorf(st->emitterHalfExtents[a] * 2.0f / spacing);
        lat[a] = v > 1u ? v : 1u;
    }
    uint32_t nz = (n + lat[0] * lat[1] - 1u) / (lat[0] * lat[1]);
    out->lattice[0] = lat[0];
    out->lattice[1] = lat[1];
    out->lattice[2] = nz > 1u ? nz : 1u;
    out->colliderCount = 0;
    for (uint32_t i = 0; colliders && i < nc && out->colliderCount < SPH_MAX_COLLIDERS; ++i) {
        if (colliders[i].shape == SPH_SHAPE_MESH) continue; /* fluid_sph.cpp:110-114 */
        out->colliders[out->colliderCount++] = colliders[i];
    }
    out->fieldCount = 0;
    for (uint32_t i = 0; fields && i < nf && out->fieldCount < SPH_MAX_FORCE_FIELDS; ++i) {
        out->fields[out->fieldCount++] = fields[i];
    }
}

void orc_lattice_jitter(float* pos4, uint32_t count, uint32_t seed, float amplitude) {
    uint32_t s = seed;
    for (uint32_t i = 0; i < count; ++i) {
        for (int a = 0; a < 3; ++a) {
            s = s * 1664525u + 1013904223u;
            float u = (float)(s >> 8) / 16777216.0f;
            pos4[4 * (size_t)i + a] += (u - 0.5f) * (2.0f * amplitude);
        }
    }
}

/* mat4 * vec4 in glm's pairwise order, column-major m. */
static void orc_mat4_mul_vec4(const float m[16], const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r) {
        float a0 = m[0 + r] * v[0] + m[4 + r] * v[1];
        float a1 = m[8 + r] * v[2] + m[12 + r] * v[3];
        out[r] = a0 + a1;
    }
}

/* ---- kernel functions, FP32 (fluid_sph.cpp:24-46) ---- */

float orc_poly6_f32(float r2, float h) {
    float h2 = h * h;
    if (r2 >= h2) return 0.0f;
    float d = h2 - r2;
    return 315.0f / (64.0f * ORC_PI_F * powf(h, 9.0f)) * d * d * d;
}

double orc_poly6_f64(double r2, double h) {
    double h2 = h * h;
    if (r2 >= h2) return 0.0;
    double d = h2 - r2;
    return 315.0 / (64.0 * M_PI * pow(h, 9.0)) * d * d * d;
}

void orc_spiky_gradient_f32(const float delta[3], float len, float h, float out[3]) {
    if (len >= h || len <= 1e-6f) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    float d = h - len;
    float s = -45.0f / (ORC_PI_F * powf(h, 6.0f)) * d * d;
    for (int a = 0; a < 3; ++a) out[a] = s * (delta[a] / len);
}

float orc_visc_laplacian_f32(float len, float h) {
    if (len >= h) return 0.0f;
    return 45.0f / (ORC_PI_F * powf(h, 6.0f)) * (h - len);
}

/* ---- helpers the templated solver uses ---- */

static inline float orc32_poly6_const(float h) { return 315.0f / (64.0f * ORC_PI_F * powf(h, 9.0f)); }
static inline float orc32_spiky_const(float h) { return -45.0f / (ORC_PI_F * powf(h, 6.0f)); }
static inline float orc32_visc_const(float h) { return 45.0f / (ORC_PI_F * powf(h, 6.0f)); }
static inline float orc32_sqrt(float x) { return sqrtf(x); }
static inline double orc64_poly6_const(double h) { return 315.0 / (64.0 * M_PI * pow(h, 9.0)); }
static inline double orc64_spiky_const(double h) { return -45.0 / (M_PI * pow(h, 6.0)); }
static inline double orc64_visc_const(double h) { return 45.0 / (M_PI * pow(h, 6.0)); }
static inline double orc64_sqrt(double x) { return sqrt(x); }

#define ORC_DEFINE_GEOMETRY(R, SFX, SQRT, POW)                                                            \
    /* forceFieldAcceleration (force_field.h:31-60) */                                                     \
    static void orc##SFX##_field_one(const SphForceField* f, const R p[3], R out[3]) {                     \
        R off[3] = {p[0] - (R)f->position[0], p[1] - (R)f->position[1], p[2] - (R)f->position[2]};         \
        R distance = SQRT((off[0] * off[0] + off[1] * off[1]) + off[2] * off[2]);                          \
        R weight = (R)1;                                                                                   \
        if (f->radius > 0.0f) {                                                                            \
            R t = (R)1 - distance / (R)f->radius;                                                          \
            t = t < (R)0 ? (R)0 : ((R)1 < t ? (R)1 : t);                                                   \
            weight = t <= (R)0 ? (R)0 : POW(t, (R)f->falloff);                                             \
        }                                                                                                  \
        out[0] = out[1] = out[2] = (R)0;                                                                   \
        if (weight <= (R)0) return;                                                                        \
        R ax[3] = {(R)f->axis[0], (R)f->axis[1], (R)f->axis[2]};                                           \
        R sw = (R)f->strength * weight;                                                                    \
        if (f->type == SPH_FIELD_WIND) {                                                                   \
            out[0] = ax[0] * sw; out[1] = ax[1] * sw; out[2] = ax[2] * sw;                                 \
        } else if (f->type == SPH_FIELD_VORTEX) {                                                          \
            R along = (off[0] * ax[0] + off[1] * ax[1]) + off[2] * ax[2];                                  \
            R rad[3] = {off[0] - ax[0] * along, off[1] - ax[1] * along, off[2] - ax[2] * along};           \
            R r = SQRT((rad[0] * rad[0] + rad[1] * rad[1]) + rad[2] * rad[2]);                             \
            if (r < (R)1.0e-4f) return;                                                                    \
            R u[3] = {rad[0] / r, rad[1] / r, rad[2] / r};                                                 \
            R cr[3] = {ax[1] * u[2] - u[1] * ax[2], ax[2] * u[0] - u[2] * ax[0], ax[0] * u[1] - u[0] * ax[1]}; \
            out[0] = cr[0] * sw; out[1] = cr[1] * sw; out[2] = cr[2] * sw;                                 \
        } else if (f->type == SPH_FIELD_POINT) {                                                           \
            if (distance < (R)1.0e-4f) return;                                                             \
            out[0] = (-off[0] / distance) * sw;                                                            \
            out[1] = (-off[1] / distance) * sw;                                                            \
            out[2] = (-off[2] / distance) * sw;                                                            \
        }                                                                                                  \
    }                                                                                                      \
    /* forceFieldsAcceleration (force_field.h:62-70): total = 0 + a_0 + a_1 ... */                        \
    static void orc##SFX##_fields_accel(const SphParams* p, R x, R y, R z, R out[3]) {                     \
        R pos[3] = {x, y, z};                                                                              \
        out[0] = out[1] = out[2] = (R)0;                                                                   \
        for (uint32_t i = 0; i < p->fieldCount && i < SPH_MAX_FORCE_FIELDS; ++i) {                         \
            R a[3];                                                                                        \
            orc##SFX##_field_one(&p->fields[i], pos, a);                                                   \
            out[0] += a[0]; out[1] += a[1]; out[2] += a[2];                                                \
        }                                                                                                  \
    }                                                                                                      \
    /* closestOn*Local (collider_shapes.h:33-131); out = point[3], normal[3], distance */                  \
    static void orc##SFX##_sphere_local(R radius, const R p[3], R out[7]) {                                \
        R d = SQRT((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);                                             \
        if (d > (R)1.0e-6f) { out[3] = p[0] / d; out[4] = p[1] / d; out[5] = p[2] / d; }                   \
        else { out[3] = 0; out[4] = 1; out[5] = 0; }                                                       \
        out[0] = out[3] * radius; out[1] = out[4] * radius; out[2] = out[5] * radius;                      \
        out[6] = d - radius;                                                                               \
    }                                                                                                      \
    static void orc##SFX##_closest_local(uint32_t shape, R radius, const R half[3], const R p[3], R out[7]) { \
        switch (shape) {                                                                                   \
        case SPH_SHAPE_SPHERE: orc##SFX##_sphere_local(radius, p, out); return;                            \
        case SPH_SHAPE_BOX: {                                                                              \
            R cl[3], dl[3];                                                                                \
            for (int a = 0; a < 3; ++a) {                                                                  \
                R lo = -half[a], v = p[a];                                                                 \
                R mx = v < lo ? lo : v;              /* glm::max(v, lo) */                                 \
                cl[a] = half[a] < mx ? half[a] : mx; /* glm::min(mx, hi) */                                \
                dl[a] = p[a] - cl[a];                                                                      \
            }                                                                                              \
            R dsq = (dl[0] * dl[0] + dl[1] * dl[1]) + dl[2] * dl[2];                                       \
            if (dsq > (R)1.0e-10f) {                                                                       \
                R dist = SQRT(dsq);                                                                        \
                out[3] = dl[0] / dist; out[4] = dl[1] / dist; out[5] = dl[2] / dist;                       \
                out[0] = cl[0]; out[1] = cl[1]; out[2] = cl[2];                                            \
                out[6] = dist;                                                                             \
                return;                                                                                    \
            }                                                                                              \
            R gap[3] = {half[0] - (p[0] < 0 ? -p[0] : p[0]), half[1] - (p[1] < 0 ? -p[1] : p[1]),          \
                        half[2] - (p[2] < 0 ? -p[2] : p[2])};                                              \
            int axis = gap[0] < gap[1] ? (gap[0] < gap[2] ? 0 : 2) : (gap[1] < gap[2] ? 1 : 2);            \
            R n[3] = {0, 0, 0};                                                                            \
            n[axis] = p[axis] >= (R)0 ? (R)1 : (R)-1;                                                      \
            out[0] = p[0]; out[1] = p[1]; out[2] = p[2];                                                   \
            out[axis] = n[axis] * half[axis];                                                              \
            out[3] = n[0]; out[4] = n[1]; out[5] = n[2];                                                   \
            out[6] = -gap[axis];                                                                           \
            return;                                                                                        \
        }                                                                                                  \
        case SPH_SHAPE_PLANE:                                                                              \
            out[0] = p[0]; out[1] = 0; out[2] = p[2];                                                      \
            out[3] = 0; out[4] = 1; out[5] = 0;                                                            \
            out[6] = p[1];                                                                                 \
            return;                                                                                        \
        case SPH_SHAPE_CYLINDER: {                                                                         \
            R hh = half[1];                                                                                \
            R d = SQRT((p[0] * p[0] + (R)0 * (R)0) + p[2] * p[2]);                                         \
            R rd[3];                                                                                       \
            if (d > (R)1.0e-6f) { rd[0] = p[0] / d; rd[1] = (R)0 / d; rd[2] = p[2] / d; }                  \
            else { rd[0] = 1; rd[1] = 0; rd[2] = 0; }                                                      \
            R ay = p[1] < 0 ? -p[1] : p[1];                                                                \
            if (ay <= hh && d <= radius) {                                                                 \
                R gapY = hh - ay, gapR = radius - d;                                                       \
                if (gapY < gapR) {                                                                         \
                    R side = p[1] >= (R)0 ? (R)1 : (R)-1;                                                  \
                    out[3] = 0; out[4] = side; out[5] = 0;                                                 \
                    out[0] = p[0]; out[1] = side * hh; out[2] = p[2];                                      \
                    out[6] = -gapY;                                                                        \
                } else {                                                                                   \
                    out[3] = rd[0]; out[4] = rd[1]; out[5] = rd[2];                                        \
                    out[0] = rd[0] * radius + (R)0; out[1] = rd[1] * radius + p[1];                        \
                    out[2] = rd[2] * radius + (R)0;                                                        \
                    out[6] = -gapR;                                                                        \
                }                                                                                          \
                return;                                                                                    \
            }                                                                                              \
            R m = d < radius ? d : radius;           /* std::min(d, radius) */                             \
            R cl[3] = {rd[0] * m, rd[1] * m, rd[2] * m};                                                   \
            cl[1] = p[1] < -hh ? -hh : (hh < p[1] ? hh : p[1]);                                            \
            R dl[3] = {p[0] - cl[0], p[1] - cl[1], p[2] - cl[2]};                                          \
            R dist = SQRT((dl[0] * dl[0] + dl[1] * dl[1]) + dl[2] * dl[2]);                                \
            out[0] = cl[0]; out[1] = cl[1]; out[2] = cl[2];                                                \
            if (dist > (R)1.0e-6f) { out[3] = dl[0] / dist; out[4] = dl[1] / dist; out[5] = dl[2] / dist; } \
            else { out[3] = 0; out[4] = 1; out[5] = 0; }                                                   \
            out[6] = dist;                                                                                 \
            return;                                                                                        \
        }                                                                                                  \
        case SPH_SHAPE_CAPSULE: {                                                                          \
            R hh = half[1];                                                                                \
            R seg = p[1] < -hh ? -hh : (hh < p[1] ? hh : p[1]);                                            \
            R q[3] = {p[0] - (R)0, p[1] - seg, p[2] - (R)0};                                               \
            orc##SFX##_sphere_local(radius, q, out);                                                       \
            out[0] += (R)0; out[1] += seg; out[2] += (R)0;                                                 \
            return;                                                                                        \
        }                                                                                                  \
        default: orc##SFX##_sphere_local((R)0, p, out); return;                                            \
        }                                                                                                  \
    }                                                                                                      \
    /* one collider of the integrate pass (fluid_sph.cpp:289-300) and bounce (:48-53) */                   \
    static void orc##SFX##_collide(const SphCollider* c, R radius, R restitution, R pp[3], R vv[3]) {      \
        const float* iw = c->inverseWorld;                                                                 \
        const float* w = c->world;                                                                         \
        R loc[3];                                                                                          \
        for (int r = 0; r < 3; ++r) {                                                                      \
            R a0 = (R)iw[0 + r] * pp[0] + (R)iw[4 + r] * pp[1];                                            \
            R a1 = (R)iw[8 + r] * pp[2] + (R)iw[12 + r] * (R)1;                                            \
            loc[r] = a0 + a1;                                                                              \
        }                                                                                                  \
        R half[3] = {(R)c->halfExtents[0], (R)c->halfExtents[1], (R)c->halfExtents[2]};                    \
        R sp[7];                                                                                           \
        orc##SFX##_closest_local(c->shape, (R)c->radius, half, loc, sp);                                   \
        if (!(sp[6] < radius)) return;                                                                     \
        R nw[3];                                                                                           \
        for (int r = 0; r < 3; ++r) nw[r] = ((R)w[0 + r] * sp[3] + (R)w[4 + r] * sp[4]) + (R)w[8 + r] * sp[5]; \
        R inv = (R)1 / SQRT((nw[0] * nw[0] + nw[1] * nw[1]) + nw[2] * nw[2]);                              \
        R n[3] = {nw[0] * inv, nw[1] * inv, nw[2] * inv};                                                  \
        for (int r = 0; r < 3; ++r) {                                                                      \
            R a0 = (R)w[0 + r] * sp[0] + (R)w[4 + r] * sp[1];                                              \
            R a1 = (R)w[8 + r] * sp[2] + (R)w[12 + r] * (R)1;                                              \
            pp[r] = (a0 + a1) + n[r] * radius;                                                             \
        }                                                                                                  \
        R along = (vv[0] * n[0] + vv[1] * n[1]) + vv[2] * n[2];                                            \
        if (along < (R)0) {                                                                                \
            R s = ((R)1 + restitution) * along;                                                            \
            vv[0] -= s * n[0]; vv[1] -= s * n[1]; vv[2] -= s * n[2];                                       \
        }                                                                                                  \
    }

ORC_DEFINE_GEOMETRY(float, 32, sqrtf, powf)
ORC_DEFINE_GEOMETRY(double, 64, sqrt, pow)

void orc_field_accel_f32(const SphForceField* fields, uint32_t count, const float p[3], float out[3]) {
    SphParams tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.fieldCount = count < SPH_MAX_FORCE_FIELDS ? count : SPH_MAX_FORCE_FIELDS;
    memcpy(tmp.fields, fields, tmp.fieldCount * sizeof(SphForceField));
    orc32_fields_accel(&tmp, p[0], p[1], p[2], out);
}

void orc_closest_on_collider_f32(uint32_t shape, float radius, const float half[3], const float p[3], float out[7]) {
    orc32_closest_local(shape, radius, half, p, out);
}

void orc_conservation(const float* pos4, const float* vel4, uint32_t count, const SphParams* p, double out[3]) {
    double m = (double)p->particleMass;
    double ke = 0.0, pe = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        const float* x = pos4 + 4 * (size_t)i;
        const float* v = vel4 + 4 * (size_t)i;
        double v2 = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
        ke += 0.5 * m * v2;
        pe -= m * ((double)p->gravity[0] * x[0] + (double)p->gravity[1] * x[1] + (double)p->gravity[2] * x[2]);
    }
    out[0] = m * (double)count;
    out[1] = ke;
    out[2] = pe;
}

/* ---- the two solver instantiations ---- */

#define R float
#define S(x) orc32_##x
#define SOLVER OrcSolver32
#include "sph_oracle_impl.inc"
#undef R
#undef S
#undef SOLVER

#define R double
#define S(x) orc64_##x
#define SOLVER OrcSolver64
#include "sph_oracle_impl.inc"
#undef R
#undef S
#undef SOLVER

void orc32_read(const OrcSolver32* s, float* pos4, float* vel4) {
    if (pos4) memcpy(pos4, s->pos, (size_t)s->n * 4 * sizeof(float));
    if (vel4) memcpy(vel4, s->vel, (size_t)s->n * 4 * sizeof(float));
}

void orc32_read_previous(const OrcSolver32* s, float* pos4) {
    memcpy(pos4, s->prev, (size_t)s->n * 4 * sizeof(float));
}

void orc32_read_forces(const OrcSolver32* s, float* force4) {
    memcpy(force4, s->force, (size_t)s->n * 4 * sizeof(float));
}

void orc64_read(const OrcSolver64* s, double* pos4, double* vel4) {
    if (pos4) memcpy(pos4, s->pos, (size_t)s->n * 4 * sizeof(double));
    if (vel4) memcpy(vel4, s->vel, (size_t)s->n * 4 * sizeof(double));
}

void orc64_read_forces(const OrcSolver64* s, double* force4) {
    memcpy(force4, s->force, (size_t)s->n * 4 * sizeof(double));
}
