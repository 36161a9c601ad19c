// sph_fluid.hpp — header-only C++ adapter over sph_c_api.h with the shape of the
// reference's physics::FluidSolver (src/physics/fluid_sph.h:101-127), for C++ callers
// such as gfx::FluidSimulator's CPU-backend path (src/gfx/fluid.cpp:377-483).
//
//   sphmi::FluidSolver solver(maxParticles, maxCellCount);
//   solver.emit(params, count);            // FluidSolver::emit
//   solver.step(params, dt);               // FluidSolver::step (enqueued on the GPU)
//   const auto& p = solver.particles();    // FluidSolver::particles (synchronises)
//
// The reference's step takes a core::JobSystem*; the GPU solver has no CPU pool, so that
// argument is dropped. Errors surface as sphmi::Error (the C ABI itself never throws).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "sph_c_api.h"

namespace sphmi {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
    if (status != SPH_OK) throw Error(status, sph_last_error());
}

struct Vec4 {
    float x, y, z, w;
};

inline SphParams deriveFluidParams(const SphSettings& settings, const float emitterWorld[16], uint32_t particleCount,
                                   uint32_t cellCount, const SphCollider* colliders = nullptr,
                                   uint32_t colliderCount = 0, const SphForceField* fields = nullptr,
                                   uint32_t fieldCount = 0) {
    SphParams out;
    check(sph_derive_params(&settings, emitterWorld, particleCount, cellCount, colliders, colliderCount, fields,
                            fieldCount, &out));
    return out;
}

inline uint32_t fluidSubsteps(const SphParams& params, float dt) { return sph_substeps(&params, dt); }

class FluidSolver {
public:
    FluidSolver(uint32_t maxParticles, uint32_t maxCellCount, int device = 0, uint32_t flags = 0) {
        SphSolverConfig cfg{maxParticles, maxCellCount, device, flags};
        check(sph_create(&cfg, &handle_));
    }
    ~FluidSolver() { sph_destroy(handle_); }
    FluidSolver(const FluidSolver&) = delete;
    FluidSolver& operator=(const FluidSolver&) = delete;

    void emit(const SphParams& params, uint32_t count) {
        check(sph_emit(handle_, &params, count));
        dirty_ = true;
    }
    void step(const SphParams& params, float deltaSeconds) {
        check(sph_step(handle_, &params, deltaSeconds));
        dirty_ = true;
    }
    uint32_t particleCount() const { return sph_particle_count(handle_); }
    // xyz position, w density; valid until the next emit / step.
    const std::vector<Vec4>& particles() {
        if (dirty_) {
            positions_.resize(particleCount());
            check(sph_read_particles(handle_, reinterpret_cast<float*>(positions_.data()), nullptr));
            dirty_ = false;
        }
        return positions_;
    }
    std::vector<Vec4> previousParticles() const {
        std::vector<Vec4> out(particleCount());
        check(sph_read_previous(handle_, reinterpret_cast<float*>(out.data())));
        return out;
    }
    void keepRendered() { check(sph_keep_rendered(handle_)); }
    SphSolver* handle() const { return handle_; }

private:
    SphSolver* handle_ = nullptr;
    std::vector<Vec4> positions_;
    bool dirty_ = true;
};

}  // namespace sphmi
