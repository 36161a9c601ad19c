/*
 * sph_c_api.h — C ABI of the MI355X-native SPH fluid step.
 *
 * This is the drop-in boundary for the reference's `src/physics` fluid interface
 * (namespace physics, src/physics/fluid_sph.h:42-127). Every entry point below names
 * the reference function or member it replaces. Plain C types only: opaque handles,
 * caller-owned host buffers, status-code returns, no exceptions, no global state
 * other than the last-error string of the calling thread.
 *
 * Layout conventions shared with the reference:
 *   - particle arrays crossing the ABI are AoS vec4 in particle-index order:
 *     positions (x, y, z, density), velocities (vx, vy, vz, pressure)
 *     — the layout of FluidSolver::particles() (src/physics/fluid_sph.h:108-123);
 *   - 4x4 matrices are column-major float[16], the glm::mat4 memory layout.
 *
 * Inside the library the state lives in HBM in bucket-sorted order with carried
 * particle ids; reads un-permute it back to index order.
 */
#ifndef SPH_C_API_H
#define SPH_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/physics/fluid_sph.h:22-23 (FLUID_MAX_COLLIDERS, FLUID_CELL_CAPACITY),
 * src/physics/force_field.h:17 (MAX_FORCE_FIELDS). */
#define SPH_MAX_COLLIDERS 8u
#define SPH_MAX_FORCE_FIELDS 8u
#define SPH_CELL_CAPACITY 32u
/* src/physics/fluid_sph.h:76 (FLUID_MAX_FRAME_STEP). */
#define SPH_MAX_FRAME_STEP (1.0f / 30.0f)
/* No counterpart in the reference, whose grid stores at most FLUID_CELL_CAPACITY particles
 * per cell (src/physics/fluid_sph.cpp:160-188). This grid keeps every particle and orders
 * each spatial bin's members by rank, at a cost quadratic in the bin's occupancy; a step
 * from a state whose last grid held this many particles in one bin returns
 * SPH_ERR_DEGENERATE. */
#define SPH_MAX_BIN_OCCUPANCY 16384u

/* Status codes. */
typedef enum SphStatus {
    SPH_OK = 0,
    SPH_ERR_INVALID_ARGUMENT = 1,
    SPH_ERR_NO_DEVICE = 2,
    SPH_ERR_OUT_OF_MEMORY = 3,
    SPH_ERR_DEVICE = 4,          /* a HIP runtime call failed */
    SPH_ERR_CAPACITY = 5,        /* count or cellCount above what the handle was created for */
    SPH_ERR_NOT_POWER_OF_TWO = 6,
    SPH_ERR_STATE = 7,           /* call not valid in the current state (e.g. read before emit) */
    SPH_ERR_DEGENERATE = 8       /* the state has piled SPH_MAX_BIN_OCCUPANCY particles into one bin */
} SphStatus;

/* scene::ColliderShape (src/scene/scene.h:119-131). */
typedef enum SphColliderShape {
    SPH_SHAPE_SPHERE = 0,
    SPH_SHAPE_BOX = 1,
    SPH_SHAPE_PLANE = 2,
    SPH_SHAPE_CYLINDER = 3,
    SPH_SHAPE_CAPSULE = 4,
    SPH_SHAPE_MESH = 5
} SphColliderShape;

/* scene::ForceFieldType (src/scene/scene.h:287-291). */
typedef enum SphForceFieldType {
    SPH_FIELD_WIND = 0,
    SPH_FIELD_VORTEX = 1,
    SPH_FIELD_POINT = 2
} SphForceFieldType;

/* physics::FluidCollider (src/physics/fluid_sph.h:28-35). world / inverseWorld are the
 * rigid transform and its inverse, as collectShapeColliders builds them
 * (src/physics/fluid_sph.cpp:115-123). */
typedef struct SphCollider {
    uint32_t shape;          /* SphColliderShape */
    float radius;
    float halfExtents[3];    /* box half size; y is the half height of cylinder / capsule */
    float pad0;
    float world[16];
    float inverseWorld[16];
} SphCollider;

/* physics::ForceFieldSample (src/physics/force_field.h:20-28). */
typedef struct SphForceField {
    float position[3];
    float axis[3];
    float strength;
    float radius;
    float falloff;
    uint32_t type;           /* SphForceFieldType */
} SphForceField;

/* physics::FluidParams (src/physics/fluid_sph.h:42-67), field for field. */
typedef struct SphParams {
    float emitterWorld[16];
    float emitterHalfExtents[3];
    float spacing;
    float containerMin[3];
    float particleRadius;
    float containerMax[3];
    float wallRestitution;
    float gravity[3];
    float particleMass;
    float smoothingRadius;
    float restDensity;
    float stiffness;
    float viscosity;
    uint32_t lattice[3];
    uint32_t cellCount;      /* power of two */
    uint32_t colliderCount;
    SphCollider colliders[SPH_MAX_COLLIDERS];
    uint32_t fieldCount;
    SphForceField fields[SPH_MAX_FORCE_FIELDS];
} SphParams;

/* The simulation part of scene::Fluid (src/scene/scene.h:186-203) that
 * deriveFluidParams reads. Rendering-only settings are not mirrored. */
typedef struct SphSettings {
    float emitterHalfExtents[3];
    uint32_t particleCount;
    float particleRadius;
    float smoothingRadius;
    float restDensity;
    float stiffness;
    float viscosity;
    float containerMin[3];
    float containerMax[3];
    float gravity[3];
} SphSettings;

/* Neighbour-search semantics. SPH_MODE_REFERENCE reproduces the reference exactly:
 * at most 32 particles per bucket, lowest indices first (src/physics/fluid_sph.cpp:179-187),
 * and all 27 neighbour buckets visited even when two of them are the same bucket
 * (src/physics/fluid_sph.cpp:205-216). The flags switch those quirks off. */
typedef enum SphModeFlags {
    SPH_MODE_REFERENCE = 0,
    SPH_MODE_UNCAPPED = 1,    /* every particle of a bucket is a candidate */
    SPH_MODE_DEDUPE = 2,      /* a bucket reached from two neighbour cells is visited once */
    SPH_MODE_FAST_MATH = 4    /* density and force by fused multiply-adds and the hardware reciprocal
                               * and reciprocal square root: the same neighbours and semantics, results
                               * within 1e-5 of the reference after one substep instead of bit-exact */
} SphModeFlags;

/* Equation of state. LINEAR is the reference's p = max(k (rho - rho0), 0)
 * (src/physics/fluid_sph.cpp:220). TAIT has no reference counterpart:
 * p = max(B ((rho/rho0)^7 - 1), 0) with B = k rho0 / 7, which has the same slope
 * dp/drho = k at rest density. */
typedef enum SphEos {
    SPH_EOS_LINEAR = 0,
    SPH_EOS_TAIT = 1
} SphEos;

typedef struct SphSolver SphSolver;

/* Creation options. maxParticles / maxCellCount size the device buffers once. */
typedef struct SphSolverConfig {
    uint32_t maxParticles;
    uint32_t maxCellCount;   /* power of two */
    int32_t device;          /* HIP device ordinal */
    uint32_t flags;          /* SPH_CFG_* */
} SphSolverConfig;

#define SPH_CFG_KEEP_FORCES 1u  /* keep each particle's pressure+viscosity force of the last substep */
#define SPH_CFG_NO_GRAPH 2u     /* launch substeps eagerly instead of replaying a captured hipGraph */
#define SPH_CFG_BUCKET_WALK 4u  /* force pass walks the 27 buckets instead of the density pass's
                                 * neighbour list (saves 256 B of device memory per particle) */
#define SPH_CFG_TILED_DENSITY 8u  /* density pass by blocks of 32 cells: the block's neighbourhood
                                   * bucket tables in LDS, buckets culled per particle by their
                                   * bounding boxes (same results; at 1M particles it measured
                                   * 483 us per pass against 262 us for the default pass) */

/* ---- stateless helpers ---- */

/* Human-readable text of the last error on the calling thread ("" if none). */
const char* sph_last_error(void);

/* physics::deriveFluidParams (src/physics/fluid_sph.cpp:65-100). The scene walk of
 * collectShapeColliders / collectForceFields is replaced by the caller passing the
 * already collected colliders and fields (either pointer may be NULL with count 0;
 * counts above 8 are truncated as the reference's collectors stop at 8). */
int sph_derive_params(const SphSettings* settings,
                      const float emitterWorld[16],
                      uint32_t particleCount,
                      uint32_t cellCount,
                      const SphCollider* colliders,
                      uint32_t colliderCount,
                      const SphForceField* fields,
                      uint32_t fieldCount,
                      SphParams* out);

/* physics::fluidSubsteps (src/physics/fluid_sph.cpp:128-136). */
uint32_t sph_substeps(const SphParams* params, float deltaSeconds);

/* gfx::cellCountFor (src/gfx/fluid.cpp:50-52): clamp(bit_ceil(2n), 1024, 65536). */
uint32_t sph_cell_count_for(uint32_t particleCount);

/* physics::spatialCell + physics::spatialHash (src/physics/spatial_hash.h:12-20)
 * evaluated on the host, for tests and tools. */
uint32_t sph_hash_cell(int32_t cx, int32_t cy, int32_t cz, uint32_t cellCount);
void sph_spatial_cell(const float position[3], float cellSize, int32_t outCell[3]);

/* Seeded lattice jitter used by the benchmark workloads (no reference counterpart:
 * FluidSolver::emit places an exact lattice). For each particle i in order and each
 * component x, y, z: seed = seed * 1664525 + 1013904223 (mod 2^32),
 * u = (seed >> 8) / 2^24, pos += (u - 0.5) * 2 * amplitude. The LCG form is the one
 * tests/marching_cubes_test.cpp:265-269 uses. pos4 is AoS vec4, w untouched. */
void sph_lattice_jitter(float* pos4, uint32_t count, uint32_t seed, float amplitude);

/* ---- solver handle: physics::FluidSolver (src/physics/fluid_sph.h:101-127) ---- */

int sph_create(const SphSolverConfig* config, SphSolver** out);
int sph_destroy(SphSolver* solver);

/* FluidSolver::emit (src/physics/fluid_sph.cpp:138-154): lattice positions, w = rest
 * density, zero velocity, previous = current. Runs on the device. */
int sph_emit(SphSolver* solver, const SphParams* params, uint32_t count);

/* FluidSolver::step (src/physics/fluid_sph.cpp:311-321): fluidSubsteps(params, dt)
 * substeps of min(dt, 1/30) / n each; no-op on an empty solver or dt <= 0.
 * Enqueued on the solver's stream; returns before the device finishes. */
int sph_step(SphSolver* solver, const SphParams* params, float deltaSeconds);

/* One FluidSolver::substep (src/physics/fluid_sph.cpp:190-309) of length dt. */
int sph_substep(SphSolver* solver, const SphParams* params, float dt);

/* Block until the solver's stream is idle. */
int sph_synchronize(SphSolver* solver);

/* No counterpart in the reference: the largest number of particles in one spatial bin of
 * the last grid build (0 before the first, and after sph_emit / sph_load_state).
 * Synchronous. sph_step, sph_substep and the grid builds return SPH_ERR_DEGENERATE while
 * it is SPH_MAX_BIN_OCCUPANCY or more; sph_load_state or sph_emit starts over. */
int sph_max_bin_occupancy(SphSolver* solver, uint32_t* out);

/* No counterpart in the reference: whether the last substep's force pass divided and took
 * square roots with its short sequences (1) or the IEEE ones (0). Both give the same bits;
 * the short ones are used when the scene's h, mass and viscosity and every particle's
 * position, velocity, density and pressure lie inside the ranges where they are exact
 * (sph_check_exact_arithmetic). Synchronous; 0 before the first substep. */
int sph_short_division(SphSolver* solver, uint32_t* out);

/* FluidSolver::particleCount (src/physics/fluid_sph.h:108). */
uint32_t sph_particle_count(const SphSolver* solver);

/* FluidSolver::particles (src/physics/fluid_sph.h:110) plus the velocities the reference
 * keeps privately (src/physics/fluid_sph.h:121-122). Either pointer may be NULL.
 * Synchronous. Buffers hold count * 4 floats. */
int sph_read_particles(SphSolver* solver, float* pos4, float* vel4);

/* FluidSolver::previousParticles / keepRendered (src/physics/fluid_sph.h:112-114). */
int sph_read_previous(SphSolver* solver, float* pos4);
int sph_keep_rendered(SphSolver* solver);

/* ---- extensions the reference has no counterpart for ---- */

/* Replace the state with caller data (index order, AoS vec4). previous = positions. */
int sph_load_state(SphSolver* solver, const float* pos4, const float* vel4, uint32_t count);

/* The FluidGrid of the last substep in the reference's format
 * (src/physics/fluid_sph.h:84-97): cellOf[count] by particle index, cellCounts[cellCount]
 * uncapped, cellParticles[cellCount * 32] with the first min(count, 32) slots of each
 * bucket valid (the rest are 0xFFFFFFFF). Any pointer may be NULL. Synchronous. */
int sph_read_grid(SphSolver* solver, uint32_t* cellOf, uint32_t* cellCounts, uint32_t* cellParticles);

/* The device layout behind the grid: keys[count] and ids[count] per slot in memory
 * order, and bucketStarts[cellCount], the first slot of each bucket (0xFFFFFFFF for empty
 * buckets). Every bucket occupies bucketStarts[b] .. + cellCounts[b] in ascending
 * particle index; buckets are ordered by space, not by key. Any pointer may be NULL.
 * Synchronous. */
int sph_read_sorted(SphSolver* solver, uint32_t* keys, uint32_t* ids, uint32_t* bucketStarts);

/* Pressure + viscosity force of the last substep, per particle in index order, as
 * (fx, fy, fz, 0) — the `force` accumulator of src/physics/fluid_sph.cpp:233-258, which
 * the reference never stores. Requires SPH_CFG_KEEP_FORCES. Synchronous. */
int sph_read_forces(SphSolver* solver, float* force4);

/* SphModeFlags / SphEos. Default: SPH_MODE_REFERENCE, SPH_EOS_LINEAR. */
int sph_set_mode(SphSolver* solver, uint32_t modeFlags);
int sph_set_eos(SphSolver* solver, uint32_t eos);

/* Per-step conservation sums in FP64 over the current state: total mass
 * count * particleMass and E = sum(0.5 m |v|^2 - m g.x) (kinetic plus gravitational
 * potential). out[0] = mass, out[1] = kinetic, out[2] = potential. Synchronous. */
int sph_conservation(SphSolver* solver, const SphParams* params, double out[3]);

/* Device time, in milliseconds, of each kernel class over the substeps run since the
 * last reset, measured with HIP events on the solver's stream when profiling is on
 * (profiling launches substeps one by one instead of replaying the step's hipGraph).
 * names/ms/launches hold up to `capacity` entries; returns the number written. */
int sph_profile_enable(SphSolver* solver, int enable);
int sph_profile_read(SphSolver* solver, const char** names, double* ms, uint32_t* launches, uint32_t capacity);

/* ---- batches of independent scenes (no reference counterpart: the reference keeps one
 * physics::FluidSolver per fluid, gfx::FluidSimulator::State, src/gfx/fluid.h:169-320) ----
 *
 * All scenes of a batch advance together: every kernel of a substep is one launch over
 * all scenes, and a step (each scene with its own fluidSubsteps rule; a scene that needs
 * fewer substeps than the batch's maximum passes through the extra ones unchanged) is
 * captured once as a hipGraph and replayed. Each scene's results are bit-identical to
 * running it alone on a solver with the same params and layout. */
typedef struct SphBatch SphBatch;

typedef struct SphSceneCapacity {
    uint32_t maxParticles;
    uint32_t maxCellCount;   /* power of two */
} SphSceneCapacity;

int sph_batch_create(const SphSceneCapacity* scenes, uint32_t sceneCount, int32_t device, uint32_t flags,
                     SphBatch** out);
int sph_batch_destroy(SphBatch* batch);
uint32_t sph_batch_scene_count(const SphBatch* batch);
/* The scene's FluidParams; required before emit / step. */
int sph_batch_set_params(SphBatch* batch, uint32_t scene, const SphParams* params);
/* Layout knobs (results do not depend on them): spatial-bin tile of 2^tileLog2 cells per
 * axis (the three sum to 6) and whether the force pass uses the density pass's
 * neighbour list. */
int sph_batch_set_layout(SphBatch* batch, uint32_t scene, const uint32_t tileLog2[3], uint32_t neighbourList);
int sph_batch_set_mode(SphBatch* batch, uint32_t scene, uint32_t modeFlags, uint32_t eos);
/* Block schedule: the order in which the scenes' workgroups are laid out in every batched
 * launch (`order` is a permutation of the scene indices; count 0 restores scene order).
 * A layout knob for the autotuner: results do not depend on it. */
int sph_batch_set_schedule(SphBatch* batch, const uint32_t* order, uint32_t count);
int sph_batch_emit(SphBatch* batch, uint32_t scene, uint32_t count);
int sph_batch_load_state(SphBatch* batch, uint32_t scene, const float* pos4, const float* vel4, uint32_t count);
/* FluidSolver::step for every scene; dt <= 0 is a no-op. Enqueued, returns early. */
int sph_batch_step(SphBatch* batch, float deltaSeconds);
int sph_batch_synchronize(SphBatch* batch);
/* No counterpart in the reference: sph_max_bin_occupancy over every scene of the batch (a
 * batch step returns SPH_ERR_DEGENERATE the same way). */
int sph_batch_max_bin_occupancy(SphBatch* batch, uint32_t* out);
uint32_t sph_batch_particle_count(const SphBatch* batch, uint32_t scene);
int sph_batch_read_particles(SphBatch* batch, uint32_t scene, float* pos4, float* vel4);
int sph_batch_profile_enable(SphBatch* batch, int enable);
int sph_batch_profile_read(SphBatch* batch, const char** names, double* ms, uint32_t* launches, uint32_t capacity);

/* ---- Renderer export ----
 * gfx::FluidSimulator::writeCpuInstances (src/gfx/fluid.cpp:377-469) and
 * shaders/fluid_instances.comp: per particle, a draw instance (gfx::GpuInstance,
 * src/gfx/geometry.h:98-118) with the sphere's model matrix (translate to the position,
 * scale by the particle radius) and the previous frame's, and optionally a ray-tracing
 * instance in VkAccelerationStructureInstanceKHR layout. */
typedef struct SphGpuInstance {
    float model[16];            /* column-major */
    float previousModel[16];
    float normalMatrix[16];
    uint32_t meshIndex;
    uint32_t bucket;
    uint32_t bucketBase;
    uint32_t jointOffset;
    uint32_t skinnedVertexOffset;
    uint32_t previousSkinnedVertexOffset;
    uint32_t skinnedMeshletOffset;
    uint32_t visibilityBase;
} SphGpuInstance;   /* 224 bytes */

typedef struct SphTlasInstance {
    float transform[3][4];      /* row-major 3 x 4 */
    uint32_t customIndexMask;   /* instanceCustomIndex (24 bits) | mask << 24 */
    uint32_t sbtOffsetFlags;    /* instanceShaderBindingTableRecordOffset (24 bits) | flags << 24 */
    uint64_t accelerationStructureReference;
} SphTlasInstance;   /* 64 bytes */

typedef struct SphInstanceConfig {
    uint32_t instanceBase;      /* first slot of the fluid in the instance array */
    uint32_t tlasBase;          /* first slot in the ray-tracing instance array */
    uint32_t sphereMesh;        /* mesh index of the particle sphere */
    uint32_t fluidIndex;        /* fluid number carried by the surface instance */
    uint64_t sphereBlas;        /* sphere structure address: one RT instance per particle */
    uint64_t surfaceBlas;       /* non-zero: one RT instance for the surface, the rest inactive */
    uint32_t resetHistory;      /* previous model = current model (after emit, teleports) */
    uint32_t pad;
} SphInstanceConfig;

/* Writes instances[instanceBase + i] for every particle i and, when tlas is not NULL,
 * tlas[tlasBase + i]; then keeps the rendered positions as the next previous positions
 * (FluidSolver::keepRendered). */
int sph_write_instances(SphSolver* solver, const SphParams* params, const SphInstanceConfig* config,
                        SphGpuInstance* instances, SphTlasInstance* tlas);

/* ---- Fluid surface ----
 * physics::buildFluidField / physics::marchFluidField (src/physics/marching_cubes.h:44-60,
 * marching_cubes.cpp:312-487). The field has (resolution + 1)^3 samples spanning the
 * container, x fastest; each interior sample sums (1 - r^2/h^2)^3 over the particles
 * within h found through the hash grid (a bucket reached from two cells counted once,
 * at most 32 per bucket), the samples on the grid's faces are 0. */
typedef struct SphSurfaceVertex {   /* physics::SurfaceVertex (marching_cubes.h:31-36) */
    float position[3];
    uint32_t normal;                /* octahedral snorm16 x 2, asset::packUnitVector */
} SphSurfaceVertex;

/* The marching-cubes case table, 256 x 16 edge indices, -1 terminated (the reference's
 * MC_TABLE role, marching_cubes.h:24; generated from the same rule, see sph_mc_table.cpp). */
int sph_mc_table(int8_t* out);
/* buildFluidField for the solver's current particles into the host array `field`. */
int sph_surface_field(SphSolver* solver, const SphParams* params, uint32_t resolution, float* field);
/* marchFluidField over a host field: at most `capacity` vertices (three per triangle) into
 * `out`, their count in *written (the reference's capacity rule). Vertices come out in
 * the reference's serial loop order (cells x fastest, each cell's triangles in table
 * order), so the buffer equals the reference's array entry for entry; a cell whose
 * vertices would pass `capacity` writes none and ends the list, as the serial loop does. */
int sph_march_field(const float* field, uint32_t resolution, const float origin[3], const float cellSize[3],
                    float iso, SphSurfaceVertex* out, uint32_t capacity, uint32_t* written);
/* Both on the device for the solver's current particles (grid over the container). */
int sph_surface_mesh(SphSolver* solver, const SphParams* params, uint32_t resolution, float iso,
                     SphSurfaceVertex* out, uint32_t capacity, uint32_t* written);

/* ---- Learned surrogate ----
 * No counterpart in the reference: an addition on top of FluidSolver::step
 * (src/physics/fluid_sph.cpp:151-303). A network predicts each particle's SPH
 * acceleration (pressure + viscosity force / density) from sums over its neighbourhood and
 * holds it for 4 substeps; the density pass of the next substep checks the result and
 * the plain path redoes any macro-step whose mean density excess is above a threshold.
 *
 * Network: 64 -> 128 -> 128 -> 128 -> 3, ReLU, BF16 operands with FP32 accumulation,
 * FP32 master weights, Adam. Parameters are W[out][in] then b[out] per layer (the last
 * layer padded to 32 outputs, rows 3..31 held at zero); sph_net_param_count() floats.
 * Feature row (64 BF16, bits as uint16), from the density pass's own candidates (r2 < h2,
 * in visiting order, duplicate visits and the particle itself included), with
 * s = (h2 - r2) / h2, e = (x_t - x_j) / h and w = (v_j - v_t) / (1 m/s): at 0 the
 * particle's 100 (rho / rho0 - 1); 1..3 sum s^2 e; 4..9 sum s^3 (ex ex, ey ey, ez ez,
 * ex ey, ex ez, ey ez); 10..12 sum s w; 13 sum s^2 (w . e); 14..16 sum s^2 (w x e);
 * 17 the candidate count / 32; 18 sum s / 8; 19 h / (1 cm); zero at 20..63. Outputs
 * and targets are accelerations in units of 98.1 m/s^2 (10 g). */
typedef struct SphNet SphNet;
typedef struct SphDataset SphDataset;

typedef struct SphRolloutStats {
    uint64_t macroSteps;      /* network macro-steps attempted */
    uint64_t fallbacks;       /* of those, redone with the plain path */
    uint64_t plainSubsteps;   /* substeps run by the plain path (fallbacks, back-off and leftovers) */
    uint64_t backedOff;       /* macro-steps run plain while backing off after a failed check */
    float maxDeviation;       /* largest mean density excess seen by the check */
    float lastDeviation;
    float maxRarefactionGrowth;   /* largest growth of the mean density deficit over a macro-step */
    uint32_t pad0;
    uint64_t rarefactionFallbacks;   /* fallbacks caused by the rarefaction side alone */
} SphRolloutStats;

/* The two sides of the rollout's density check. COMPRESSION fails a macro-step whose
 * resulting mean density excess mean(max(rho / rho0 - 1, 0)) is above the threshold;
 * RAREFACTION fails one that grew the mean density deficit mean(max(1 - rho / rho0, 0)) by
 * more than the threshold (free surfaces carry a deficit of their own, so its growth over
 * the macro-step is what is checked). */
#define SPH_CHECK_COMPRESSION 1u
#define SPH_CHECK_RAREFACTION 2u

uint32_t sph_net_param_count(void);
/* He-uniform weights from `seed` (splitmix64), zero biases. */
int sph_net_create(int32_t device, uint64_t seed, SphNet** out);
int sph_net_destroy(SphNet* net);
int sph_net_get_params(SphNet* net, float* params);
/* Sets the FP32 weights (frozen padding forced to zero) and resets Adam's state. */
int sph_net_set_params(SphNet* net, const float* params);
int sph_net_save(SphNet* net, const char* path);
int sph_net_load(SphNet* net, const char* path);
/* y[rows x 3] for feature rows x[rows x 64]. */
int sph_net_forward(SphNet* net, const uint16_t* x, uint32_t rows, float* y);
/* Gradient of the mean squared error (mean over rows x 3 outputs) w.r.t. the
 * parameters, and the loss; no update. */
int sph_net_gradients(SphNet* net, const uint16_t* x, const float* t, uint32_t rows, float* grad, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on host rows. */
int sph_net_train_batch(SphNet* net, const uint16_t* x, const float* t, uint32_t rows, float lr, float* loss);

/* Device-resident training set of (feature row, target) pairs. */
int sph_dataset_create(int32_t device, uint32_t capacity, SphDataset** out);
int sph_dataset_destroy(SphDataset* data);
uint32_t sph_dataset_size(const SphDataset* data);
/* Records one macro-step of `solver`'s plain path: the features of `samples` distinct
 * particles (seeded choice) at its start and, as targets, their mean SPH acceleration
 * over the 4 following substeps of length dtSub. Advances the solver by 4 substeps. */
int sph_dataset_record(SphDataset* data, SphSolver* solver, const SphParams* params, float dtSub, uint32_t samples,
                       uint64_t seed);
/* Rows [begin, begin + count): features x, targets t and, in t0, each particle's SPH
 * acceleration of the first substep alone; any pointer may be NULL. */
int sph_dataset_read(SphDataset* data, uint32_t begin, uint32_t count, uint16_t* x, float* t, float* t0);
/* `steps` Adam steps on batches of batchRows rows drawn with replacement (seeded) from
 * dataset rows [begin, end); losses[i] = loss of step i when losses is not NULL. */
int sph_net_train(SphNet* net, SphDataset* data, uint32_t begin, uint32_t end, uint32_t batchRows, uint32_t steps,
                  float lr, uint64_t seed, float* losses);
/* Mean squared errors on rows [begin, end): out[0] of the network, out[1] of the
 * no-network prediction (zero SPH acceleration: gravity alone), out[2] of holding the
 * first substep's SPH acceleration for the macro-step (which needs a plain force pass). */
int sph_net_evaluate(SphNet* net, SphDataset* data, uint32_t begin, uint32_t end, double out[3]);

/* Feature rows of every particle of the solver's current state, in particle order. */
int sph_features(SphSolver* solver, const SphParams* params, float dtSub, uint16_t* x);

/* The input a network reads. MOMENTS: 20 permutation-invariant sums over the neighbourhood
 * the density pass accumulates (the default). KNN8: the relative positions (/ h) and
 * velocities of the particle's 8 nearest in-range neighbours (distinct particles, by
 * (r2, index), nearest first; zeros where fewer), its density deviation (x 100) and h / 1 cm
 * — the 8-nearest-neighbour input of the surrogate-training configuration, computed by an
 * extra pass over the same 27 buckets. Weight files keep the kind. */
#define SPH_FEATURES_MOMENTS 0u
#define SPH_FEATURES_KNN8 1u
int sph_net_set_features(SphNet* net, uint32_t kind);
uint32_t sph_net_features(const SphNet* net);
int sph_dataset_set_features(SphDataset* data, uint32_t kind);
/* KNN8 rows of every particle of the solver's current state, in particle order. */
int sph_knn_features(SphSolver* solver, const SphParams* params, float dtSub, uint16_t* x);
/* `steps` calls of FluidSolver::step(deltaSeconds) with the surrogate: every 4 substeps
 * are one network macro-step checked by the next density pass; a macro-step that fails the
 * check (see SPH_CHECK_*; mean excess above `threshold`, or mean deficit grown by more than
 * `threshold`) is redone by the plain path. Leftover substeps, and every substep of scenes
 * with colliders, run plain. */
int sph_rollout(SphSolver* solver, SphNet* net, const SphParams* params, float deltaSeconds, uint32_t steps,
                float threshold, SphRolloutStats* stats);
/* Optional XSPH smoothing (Monaghan 1989) in sph_rollout: at the start of a macro-step
 * each velocity moves by `smoothing` (0..1, default 0) toward the s-weighted mean velocity
 * of its neighbourhood (sum s v_j / sum s), damping the particle-scale noise that
 * prediction errors inject. */
int sph_net_set_smoothing(SphNet* net, float smoothing);
/* The sides of the density check sph_rollout applies: SPH_CHECK_COMPRESSION,
 * SPH_CHECK_RAREFACTION or both (the default). */
int sph_net_set_check(SphNet* net, uint32_t sides);
/* Back-off in sph_rollout: after a failed check the next 1, 2, 4, ... (at most
 * maxBackoff) macro-steps run plain before the network is tried again; a passed check
 * resets it. 0 (the default) tries the network on every macro-step. */
int sph_net_set_backoff(SphNet* net, uint32_t maxBackoff);
/* Milliseconds spent in the timed network forward passes inside sph_rollout since the last
 * call (one pass in eight is timed), and the rows they covered; resets both. */
int sph_net_inference_time(SphNet* net, double* ms, uint64_t* rows);
/* out[0] = mean max(rho / rho0 - 1, 0), out[1] = mean |rho / rho0 - 1| of the current
 * state, from a density pass of a substep of length dtSub. */
int sph_density_deviation(SphSolver* solver, SphNet* net, const SphParams* params, float dtSub, double out[2]);

/* ---- arithmetic self-check ----
 * No counterpart in the reference. The force pass divides and takes square roots with
 * shorter sequences than the compiler's IEEE ones, inside operand ranges where they return
 * the same bits (computational core in computer_graphics_lab_00_amd/csrc/sph_exact.h). This
 * runs those sequences on `device` over the admitted ranges and compares them with IEEE
 * division and square root: out[0] reciprocals checked (every divisor in range), out[1] of
 * them different; out[2] square roots checked (every float below the upper bound), out[3]
 * different; out[4] quotients checked (`samples` pseudo-random pairs plus every divisor
 * significand against two numerators), out[5] different; out[6] zero numerators checked,
 * out[7] not zero. Blocks until done. */
int sph_check_exact_arithmetic(int32_t device, uint64_t samples, uint64_t out[8]);

/* ---- allocation-free substep ----
 * No counterpart in the reference. For callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. One
 * FluidSolver::substep (src/physics/fluid_sph.cpp:190-309) of length dt — the results of
 * sph_substep on a solver holding the same state, mode and equation of state, bit for
 * bit — on caller-owned DEVICE buffers: posIn / velIn hold `count` AoS vec4 in particle
 * order (the particles() layout), posOut / velOut receive the next state in the same
 * layout (they may be posIn / velIn). Enqueued on `stream` (a hipStream_t; NULL is the
 * default stream) and returns before the device finishes; allocates nothing and copies
 * nothing from host memory. The buckets, bins, sort scratch and sorted state live in one
 * static device array of SPH_NOALLOC_SCRATCH_BYTES per device (held until the process
 * ends), so count is at most SPH_NOALLOC_MAX_PARTICLES and params->cellCount at most
 * SPH_NOALLOC_MAX_CELLS (larger values return SPH_ERR_CAPACITY). Every call on a device
 * shares that array: calls that overlap on one device race and must be serialised. */
#define SPH_NOALLOC_MAX_PARTICLES 65536u
#define SPH_NOALLOC_MAX_CELLS 65536u
#define SPH_NOALLOC_SCRATCH_BYTES 8392448u
int sph_substep_device(const SphParams* params, float dt, uint32_t modeFlags, uint32_t eos, const float* posIn,
                       const float* velIn, float* posOut, float* velOut, uint32_t count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SPH_C_API_H */
