// Uses the C++ adapter the way gfx::FluidSimulator uses physics::FluidSolver:
// derive params, emit, step a few frames, read particles back.
// Exit codes: 0 ok, 2 no device visible, 1 anything else.
#include <cmath>
#include <cstdio>

#include "sph_fluid.hpp"

int main(int argc, char** argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 10;
    SphSettings s{};
    s.emitterHalfExtents[0] = s.emitterHalfExtents[1] = s.emitterHalfExtents[2] = 0.5f;
    s.particleCount = 512;
    s.particleRadius = 0.025f;
    s.smoothingRadius = 0.1f;
    s.restDensity = 1000.0f;
    s.stiffness = 50.0f;
    s.viscosity = 0.5f;
    s.containerMin[0] = -1.0f; s.containerMin[1] = 0.0f; s.containerMin[2] = -1.0f;
    s.containerMax[0] = 1.0f; s.containerMax[1] = 2.0f; s.containerMax[2] = 1.0f;
    s.gravity[1] = -9.81f;
    const float world[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1};
    try {
        SphParams p = sphmi::deriveFluidParams(s, world, 512, 1024);
        sphmi::FluidSolver solver(512, 1024);
        solver.emit(p, 512);
        for (int f = 0; f < frames; ++f) solver.step(p, 1.0f / 60.0f);
        const auto& pts = solver.particles();
        float highest = -1e9f;
        for (const auto& q : pts) {
            if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) return 1;
            highest = q.y > highest ? q.y : highest;
        }
        std::printf("particles %u highest %.4f\n", solver.particleCount(), highest);
        return pts.size() == 512 ? 0 : 1;
    } catch (const sphmi::Error& e) {
        std::printf("sph error %d: %s\n", e.status, e.what());
        return e.status == SPH_ERR_NO_DEVICE ? 2 : 1;
    }
}
