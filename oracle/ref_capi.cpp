// ref_capi.cpp - extern "C" shim over vapor's own CPU MAC 3D solver.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the reference
// sources where they lie (the vapor tree's src/), into oracle/_ref/. Nothing
// from the reference is copied into this repository: this file only calls the
// reference's public classes (MacGridFluidSolver3D, CpuAdvection,
// CpuPressureSolver3D, MacGridSampler3D, SOR, StencilLinearSystem) and
// converts their exceptions into return codes. It writes the golden fixtures
// (tests/golden/gen_golden.py) and is timed as the CPU baseline (bench.py).
#include "solvers/mac3d/CpuAdvection.hpp"
#include "solvers/mac3d/CpuPressureSolver3D.hpp"
#include "solvers/mac3d/MacGridFluidSolver3D.hpp"
#include "solvers/mac3d/MacGridOperators3D.hpp"
#include "solvers/mac3d/MacGridSampler3D.hpp"
#include "solvers/mac3d/SOR.hpp"

#include "../include/vapor_mac3d.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

namespace {
// 0 ok, 1 invalid_argument, 2 length_error, 3 other.
template <class F> int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::length_error&) {
        return 2;
    } catch (...) {
        return 3;
    }
}

void apply(MacGridParameters3D& p, const vgx_params& q) {
    p.emitterRadius = q.emitter_radius;
    p.emitterCenterX = q.emitter_center_x;
    p.emitterCenterY = q.emitter_center_y;
    p.emitterCenterZ = q.emitter_center_z;
    p.emitterKickMultiplier = q.emitter_kick_multiplier;
    p.emitterKickBase = q.emitter_kick_base;
    p.emitterKickScale = q.emitter_kick_scale;
    p.stirStrength = q.stir_strength;
    p.randomSeed = q.random_seed;
    p.sorOmega = q.sor_omega;
    p.buoyancySplit = q.buoyancy_split;
    p.cflMc = q.cfl_mc;
    p.macCormackSmoke = q.maccormack_smoke != 0;
    p.macCormackVel = q.maccormack_vel != 0;
    p.smokeDecay = q.smoke_decay;
    p.reflection = q.reflection != 0;
    p.buoyancy = q.buoyancy;
    p.sourceStrength = q.source_strength;
    p.projectIterations = q.project_iterations;
}

void export_params(const MacGridParameters3D& p, vgx_params& q) {
    std::memset(&q, 0, sizeof(q));
    q.emitter_radius = p.emitterRadius;
    q.emitter_center_x = p.emitterCenterX;
    q.emitter_center_y = p.emitterCenterY;
    q.emitter_center_z = p.emitterCenterZ;
    q.emitter_kick_multiplier = p.emitterKickMultiplier;
    q.emitter_kick_base = p.emitterKickBase;
    q.emitter_kick_scale = p.emitterKickScale;
    q.stir_strength = p.stirStrength;
    q.random_seed = p.randomSeed;
    q.sor_omega = p.sorOmega;
    q.buoyancy_split = p.buoyancySplit;
    q.cfl_mc = p.cflMc;
    q.maccormack_smoke = p.macCormackSmoke;
    q.maccormack_vel = p.macCormackVel;
    q.smoke_decay = p.smokeDecay;
    q.reflection = p.reflection;
    q.buoyancy = p.buoyancy;
    q.source_strength = p.sourceStrength;
    q.project_iterations = p.projectIterations;
}

void export_result(const ProjectionResult& r, vgx_projection_result* out) {
    std::memset(out, 0, sizeof(*out));
    out->iterations = r.linearSolve.iterations;
    out->converged = r.linearSolve.converged;
    out->initial_residual = r.linearSolve.initialResidual;
    out->final_residual = r.linearSolve.finalResidual;
    out->residual_available = r.residualAvailable;
    out->divergence_before = r.divergenceBefore;
    out->divergence_after = r.divergenceAfter;
    out->removed_rhs_mean = r.removedRhsMean;
}

void fill(MacGridState3D& s, const float* rho, const float* u, const float* v, const float* w) {
    if (rho) std::copy(rho, rho + s.density().size(), s.density().begin());
    if (u) std::copy(u, u + s.velocityX().size(), s.velocityX().begin());
    if (v) std::copy(v, v + s.velocityY().size(), s.velocityY().begin());
    if (w) std::copy(w, w + s.velocityZ().size(), s.velocityZ().begin());
}

struct RefSim {
    std::unique_ptr<MacGridFluidSolver3D> solver;
};
} // namespace

extern "C" {

// MacGridParameters3D() and MacGridParameters3D(MacAlgorithm).
int ref_params_default(vgx_params* out) {
    export_params(MacGridParameters3D{}, *out);
    return 0;
}
int ref_params_preset(int algorithm, vgx_params* out) {
    return guarded([&] { export_params(MacGridParameters3D(static_cast<MacAlgorithm>(algorithm)), *out); });
}

// MacGridFluidSolver3D(MacGridState3D(n, box)) with params applied, then
// resetState() so the seed takes effect (as the reference tests do).
int ref_sim_create(int n, float box, const vgx_params* params, void** out) {
    return guarded([&] {
        auto sim = std::make_unique<RefSim>();
        sim->solver = std::make_unique<MacGridFluidSolver3D>(MacGridState3D(n, box));
        if (params) apply(sim->solver->parameters(), *params);
        sim->solver->resetState();
        *out = sim.release();
    });
}

// MacGridFluidSolver3D(MacGridState3D) adopting prefilled fields: no reset,
// so the RNG keeps the default seed (MacGridFluidSolver3D.hpp:111).
int ref_sim_create_from(int n, float box, const vgx_params* params, const float* rho, const float* u,
                        const float* v, const float* w, void** out) {
    return guarded([&] {
        MacGridState3D state(n, box);
        fill(state, rho, u, v, w);
        auto sim = std::make_unique<RefSim>();
        sim->solver = std::make_unique<MacGridFluidSolver3D>(std::move(state));
        if (params) apply(sim->solver->parameters(), *params);
        *out = sim.release();
    });
}

int ref_sim_create_algorithm(int algorithm, void** out) {
    return guarded([&] {
        auto sim = std::make_unique<RefSim>();
        sim->solver = std::make_unique<MacGridFluidSolver3D>(static_cast<MacAlgorithm>(algorithm));
        *out = sim.release();
    });
}

void ref_sim_destroy(void* h) { delete static_cast<RefSim*>(h); }

int ref_sim_get_params(void* h, vgx_params* out) {
    export_params(static_cast<RefSim*>(h)->solver->parameters(), *out);
    return 0;
}
int ref_sim_set_params(void* h, const vgx_params* p) {
    apply(static_cast<RefSim*>(h)->solver->parameters(), *p);
    return 0;
}
int ref_sim_advance(void* h, float dt) {
    return guarded([&] { static_cast<RefSim*>(h)->solver->advance(dt); });
}
int ref_sim_reset(void* h) {
    return guarded([&] { static_cast<RefSim*>(h)->solver->resetState(); });
}
int ref_sim_set_box(void* h, float box) {
    return guarded([&] { static_cast<RefSim*>(h)->solver->setBoxSize(box); });
}
int ref_sim_resolution(void* h) { return static_cast<RefSim*>(h)->solver->resolution(); }
float ref_sim_box(void* h) { return static_cast<RefSim*>(h)->solver->boxSize(); }

// which: 0 rho, 1 u, 2 v, 3 w.
int ref_sim_download(void* h, int which, float* out) {
    const auto& s = static_cast<RefSim*>(h)->solver->state();
    std::span<const float> f = which == 0 ? s.density() : which == 1 ? s.velocityX() : which == 2 ? s.velocityY()
                                                                                                  : s.velocityZ();
    std::copy(f.begin(), f.end(), out);
    return 0;
}
int ref_sim_last_projection(void* h, vgx_projection_result* out) {
    export_result(static_cast<RefSim*>(h)->solver->lastProjection(), out);
    return 0;
}
int ref_sim_diagnostics(void* h, vgx_diagnostics* out) {
    return guarded([&] {
        const auto d = static_cast<RefSim*>(h)->solver->diagnostics();
        out->kinetic_energy = d.kineticEnergy;
        out->max_velocity = d.maxVelocity;
        out->max_divergence = d.maxDivergence;
    });
}

// MacGridSampler3D on a grid of resolution n; axis -1 cell, 0..2 faces.
float ref_sample(int n, float box, int axis, const float* f, float x, float y, float z) {
    MacGridState3D grid(n, box);
    MacGridSampler3D sampler(grid);
    const std::size_t count = axis < 0 ? grid.cellCount() : grid.velocityX().size();
    std::span<const float> field(f, count);
    const Vec3 p{x, y, z};
    switch (axis) {
    case 0: return sampler.sampleVxField(field, p);
    case 1: return sampler.sampleVyField(field, p);
    case 2: return sampler.sampleVzField(field, p);
    default: return sampler.sampleCell(field, p);
    }
}

// CpuAdvection::advectFace / advectScalar with a flow (u, v, w).
int ref_advect_face(int n, float box, int axis, int method, float dt, float cfl, const float* u, const float* v,
                    const float* w, const float* src, float* out) {
    return guarded([&] {
        MacGridState3D grid(n, box);
        const std::size_t faces = grid.velocityX().size();
        AdvectionContext ctx{grid, {{u, faces}, {v, faces}, {w, faces}}, dt, cfl};
        AdvectionWorkspace ws;
        CpuAdvection().advectFace(ctx, static_cast<FaceAxis>(axis), {src, faces}, {out, faces}, ws,
                                  method ? AdvectionMethod::MacCormack : AdvectionMethod::SemiLagrangian);
    });
}
int ref_advect_scalar(int n, float box, int method, float dt, float cfl, const float* u, const float* v,
                      const float* w, float* field) {
    return guarded([&] {
        MacGridState3D grid(n, box);
        const std::size_t faces = grid.velocityX().size();
        AdvectionContext ctx{grid, {{u, faces}, {v, faces}, {w, faces}}, dt, cfl};
        AdvectionWorkspace ws;
        CpuAdvection().advectScalar(ctx, {field, grid.cellCount()}, ws,
                                    method ? AdvectionMethod::MacCormack : AdvectionMethod::SemiLagrangian);
    });
}

// CpuPressureSolver3D(floor).project on host fields, in place. pressure and
// divergence (n^3, may be NULL) receive the solver's workspaces afterwards.
int ref_project(int n, float box, const vgx_projection_options* o, float* u, float* v, float* w, float* pressure,
                float* divergence, vgx_projection_result* result) {
    return guarded([&] {
        MacGridState3D state(n, box);
        fill(state, nullptr, u, v, w);
        CpuPressureSolver3D solver(o->diagonal_floor);
        ProjectionOptions options;
        options.dt = o->dt;
        options.relaxation = o->relaxation;
        options.linearSolve.maxIterations = o->max_iterations;
        options.linearSolve.residualCheckInterval = o->residual_check_interval;
        options.linearSolve.absoluteTolerance = o->absolute_tolerance;
        options.linearSolve.relativeTolerance = o->relative_tolerance;
        options.linearSolve.fixedIterations = o->fixed_iterations != 0;
        const auto r = solver.project(state, options);
        export_result(r, result);
        std::copy(state.velocityX().begin(), state.velocityX().end(), u);
        std::copy(state.velocityY().begin(), state.velocityY().end(), v);
        std::copy(state.velocityZ().begin(), state.velocityZ().end(), w);
        if (pressure) std::copy(solver.pressure().begin(), solver.pressure().end(), pressure);
        if (divergence) std::copy(solver.inputDivergence().begin(), solver.inputDivergence().end(), divergence);
    });
}

// MacGridOperators3D on host fields.
int ref_divergence(int n, float box, const float* u, const float* v, const float* w, float* div) {
    return guarded([&] {
        MacGridState3D state(n, box);
        fill(state, nullptr, u, v, w);
        MacGridOperators3D::computeDivergence(state, {div, state.cellCount()});
    });
}
int ref_max_divergence(int n, float box, const float* u, const float* v, const float* w, float* out) {
    return guarded([&] {
        MacGridState3D state(n, box);
        fill(state, nullptr, u, v, w);
        *out = MacGridOperators3D::maxDivergence(state);
    });
}
int ref_apply_gradient(int n, float box, const float* p, float dt, float* u, float* v, float* w) {
    return guarded([&] {
        MacGridState3D state(n, box);
        fill(state, nullptr, u, v, w);
        MacGridOperators3D::applyPressureGradient(state, {p, state.cellCount()}, dt);
        std::copy(state.velocityX().begin(), state.velocityX().end(), u);
        std::copy(state.velocityY().begin(), state.velocityY().end(), v);
        std::copy(state.velocityZ().begin(), state.velocityZ().end(), w);
    });
}

// SOR::solve on the MAC Neumann system with right-hand side b, x in/out.
int ref_sor(int n, float omega, float floor, const float* b, float* x, const vgx_projection_options* o,
            vgx_projection_result* result) {
    return guarded([&] {
        StencilLinearSystem s;
        s.resize(n);
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int i = 0; i < n; ++i) {
                    auto& a = s.matrix[s.index(i, y, z)];
                    a.diagonal = float((i > 0) + (i + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
                    a.right = i + 1 < n ? -1.f : 0.f;
                    a.up = y + 1 < n ? -1.f : 0.f;
                    a.front = z + 1 < n ? -1.f : 0.f;
                }
        std::copy(b, b + s.b.size(), s.b.begin());
        std::copy(x, x + s.x.size(), s.x.begin());
        LinearSolveOptions options;
        options.maxIterations = o->max_iterations;
        options.residualCheckInterval = o->residual_check_interval;
        options.absoluteTolerance = o->absolute_tolerance;
        options.relativeTolerance = o->relative_tolerance;
        options.fixedIterations = o->fixed_iterations != 0;
        const auto r = SOR(omega, floor).solve(s, options);
        std::memset(result, 0, sizeof(*result));
        result->iterations = r.iterations;
        result->converged = r.converged;
        result->initial_residual = r.initialResidual;
        result->final_residual = r.finalResidual;
        std::copy(s.x.begin(), s.x.end(), x);
    });
}

} // extern "C"
