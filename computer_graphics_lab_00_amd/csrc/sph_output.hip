// sph_output.hip — the renderer-facing particle export: one draw instance and one
// ray-tracing instance per particle, as gfx::FluidSimulator::writeCpuInstances
// (src/gfx/fluid.cpp:377-469) and shaders/fluid_instances.comp fill them, written by one
// lane per particle on the GPU and handed over as host-visible arrays of structs.
#include "sph_engine.h"
#include "sph_internal.h"

using sphhost::fail;
using sphhost::hip_fail;

static_assert(sizeof(SphGpuInstance) == 224, "gfx::GpuInstance is 3 mat4 + 8 u32");
static_assert(sizeof(SphTlasInstance) == 64, "VkAccelerationStructureInstanceKHR is 64 bytes");

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kNoJoints = 0xFFFFFFFFu;            // gfx::NO_JOINTS (src/gfx/geometry.h:87)
constexpr uint32_t kNoSkinnedVertices = 0xFFFFFFFFu;   // gfx::NO_SKINNED_VERTICES (geometry.h:93)
constexpr uint32_t kSurfaceCustomIndex = 0x800000u;    // gfx::FLUID_SURFACE_CUSTOM_INDEX (src/gfx/fluid.h:151)
constexpr uint32_t kSurfaceRayMask = 0x01u;            // gfx::FLUID_SURFACE_RAY_MASK (fluid.h:153)
constexpr uint32_t kCullDisable = 0x1u;                // VK_GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT_KHR

struct InstanceArgs {
    float radius;
    uint32_t n;
    uint32_t sphereMesh;
    uint32_t instanceBase;
    uint32_t fluidIndex;
    uint32_t history;    // 0: previous model = current model
    uint32_t tlasMode;   // 0: none, 1: one sphere per particle, 2: the surface in the first entry
    uint32_t pad;
    uint64_t sphereBlas;
    uint64_t surfaceBlas;
};

// glm::scale(glm::translate(mat4(1), p), vec3(r)), column-major.
__device__ inline void sphere_model(float* m, float x, float y, float z, float r) {
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = 0.0f;
    m[0] = r;
    m[5] = r;
    m[10] = r;
    m[12] = x;
    m[13] = y;
    m[14] = z;
    m[15] = 1.0f;
}

__global__ void __launch_bounds__(kThreads)
    k_instances(const float4* __restrict__ pos, const uint32_t* __restrict__ ids, const float4* __restrict__ prev,
                InstanceArgs a, SphGpuInstance* __restrict__ inst, SphTlasInstance* __restrict__ tlas) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t i = ids[t];
    const float4 p = pos[t];
    const float4 b = a.history ? prev[i] : p;
    SphGpuInstance o;
    sphere_model(o.model, p.x, p.y, p.z, a.radius);
    sphere_model(o.previousModel, b.x, b.y, b.z, a.radius);
#pragma unroll
    for (int k = 0; k < 16; ++k) o.normalMatrix[k] = (k % 5 == 0) ? 1.0f : 0.0f;   // uniform scale: identity
    o.meshIndex = a.sphereMesh;
    o.bucket = 0u;
    o.bucketBase = 0u;
    o.jointOffset = kNoJoints;
    o.skinnedVertexOffset = kNoSkinnedVertices;
    o.previousSkinnedVertexOffset = kNoSkinnedVertices;
    o.skinnedMeshletOffset = 0u;
    o.visibilityBase = 0u;
    inst[i] = o;
    if (a.tlasMode == 0u) return;
    SphTlasInstance e;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) e.transform[r][c] = 0.0f;
    e.customIndexMask = 0u;
    e.sbtOffsetFlags = 0u;
    e.accelerationStructureReference = 0u;
    if (a.tlasMode == 1u) {
        // the 3 x 4 row-major form of the model matrix
        e.transform[0][0] = a.radius;
        e.transform[1][1] = a.radius;
        e.transform[2][2] = a.radius;
        e.transform[0][3] = p.x;
        e.transform[1][3] = p.y;
        e.transform[2][3] = p.z;
        e.customIndexMask = ((a.instanceBase + i) & 0xFFFFFFu) | (0xFFu << 24);
        e.accelerationStructureReference = a.sphereBlas;
    } else if (i == 0u) {
        // surface mode: the first entry holds the surface structure, the others are inactive
        e.transform[0][0] = 1.0f;
        e.transform[1][1] = 1.0f;
        e.transform[2][2] = 1.0f;
        e.customIndexMask = ((kSurfaceCustomIndex | a.fluidIndex) & 0xFFFFFFu) | (kSurfaceRayMask << 24);
        e.sbtOffsetFlags = kCullDisable << 24;
        e.accelerationStructureReference = a.surfaceBlas;
    }
    tlas[i] = e;
}

}  // namespace

// gfx::FluidSimulator::writeCpuInstances for the solver's particles: instances[base + i]
// for particle i, and, when `tlas` is set, either one sphere instance per particle
// (sphereBlas) or the surface instance in the first entry and inactive entries after it
// (surfaceBlas != 0), at tlas[tlasBase + i]. Then the rendered positions are kept as the
// next call's previous positions (FluidSolver::keepRendered).
extern "C" int sph_write_instances(SphSolver* solver, const SphParams* p, const SphInstanceConfig* cfg,
                                   SphGpuInstance* instances, SphTlasInstance* tlas) {
    sphhost::clear_error();
    if (!solver || !p || !cfg || !instances) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_write_instances: null argument");
    sphhost::Engine& e = solver->e;
    const uint32_t n = e.sc[0].n;
    if (n == 0) return SPH_OK;
    if (int st = e.set_device()) return st;
    InstanceArgs a;
    a.radius = p->particleRadius;
    a.n = n;
    a.sphereMesh = cfg->sphereMesh;
    a.instanceBase = cfg->instanceBase;
    a.fluidIndex = cfg->fluidIndex;
    a.history = cfg->resetHistory ? 0u : 1u;
    a.tlasMode = !tlas ? 0u : (cfg->surfaceBlas != 0u ? 2u : (cfg->sphereBlas != 0u ? 1u : 0u));
    a.pad = 0u;
    a.sphereBlas = cfg->sphereBlas;
    a.surfaceBlas = cfg->surfaceBlas;
    if (solver->instCap < n) {
        hipFree(solver->instStage);
        solver->instStage = nullptr;
        solver->instCap = 0;
        SPH_HIP(hipMalloc(&solver->instStage, (size_t)n * sizeof(SphGpuInstance)));
        solver->instCap = n;
    }
    if (a.tlasMode && solver->tlasCap < n) {
        hipFree(solver->tlasStage);
        solver->tlasStage = nullptr;
        solver->tlasCap = 0;
        SPH_HIP(hipMalloc(&solver->tlasStage, (size_t)n * sizeof(SphTlasInstance)));
        solver->tlasCap = n;
    }
    SphGpuInstance* di = solver->instStage;
    SphTlasInstance* dt = a.tlasMode ? solver->tlasStage : nullptr;
    k_instances<<<(n + kThreads - 1) / kThreads, kThreads, 0, e.stream>>>(e.state_pos(), e.state_ids(),
                                                                          e.previous_pos(), a, di, dt);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess)
        he = hipMemcpyAsync(instances + cfg->instanceBase, di, (size_t)n * sizeof(SphGpuInstance),
                            hipMemcpyDeviceToHost, e.stream);
    if (he == hipSuccess && a.tlasMode)
        he = hipMemcpyAsync(tlas + cfg->tlasBase, dt, (size_t)n * sizeof(SphTlasInstance), hipMemcpyDeviceToHost,
                            e.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e.stream);
    if (he != hipSuccess) return hip_fail(he, "sph_write_instances");
    return e.keep_rendered(0);
}
