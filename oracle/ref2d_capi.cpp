// ref2d_capi.cpp - extern "C" shim over vapor's own MAC 2D solvers.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the reference
// sources where they lie (the vapor tree's src/), into oracle/_ref/. It only
// calls the reference's public classes (MacGridFluidSolver2D, SemiLagrangian2D,
// CpuPressureSolver2D, ReflectionMacFluidSolver2D) and turns their exceptions
// into return codes. tests/golden/gen_golden.py writes the 2D fixtures with it.
#include "solvers/mac2d/CpuPressureSolver2D.hpp"
#include "solvers/mac2d/MacGridFluidSolver2D.hpp"
#include "solvers/mac2d/MacGridOperators2D.hpp"
#include "solvers/mac2d/ReflectionMacFluidSolver2D.hpp"
#include "solvers/mac2d/SemiLagrangian2D.hpp"

#include "../include/vapor_mac2d.h"

#include <algorithm>
#include <cstring>
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

void apply(MacGridParameters2D& p, const vgx2_params& q) {
    p.reflection = q.reflection != 0;
    p.sourceStrength = q.source_strength;
    p.emitterKickMultiplier = q.emitter_kick_multiplier;
    p.emitterRadius = q.emitter_radius;
    p.emitterCenterX = q.emitter_center_x;
    p.emitterCenterY = q.emitter_center_y;
    p.buoyancy = q.buoyancy;
    p.smokeDecay = q.smoke_decay;
    p.sorOmega = q.sor_omega;
    p.projectIterations = q.project_iterations;
}

void export_params(const MacGridParameters2D& p, vgx2_params& q) {
    std::memset(&q, 0, sizeof q);
    q.reflection = p.reflection;
    q.source_strength = p.sourceStrength;
    q.emitter_kick_multiplier = p.emitterKickMultiplier;
    q.emitter_radius = p.emitterRadius;
    q.emitter_center_x = p.emitterCenterX;
    q.emitter_center_y = p.emitterCenterY;
    q.buoyancy = p.buoyancy;
    q.smoke_decay = p.smokeDecay;
    q.sor_omega = p.sorOmega;
    q.project_iterations = p.projectIterations;
}

void to_result(const ProjectionResult& r, vgx_projection_result& out) {
    std::memset(&out, 0, sizeof out);
    out.iterations = r.linearSolve.iterations;
    out.converged = r.linearSolve.converged;
    out.initial_residual = r.linearSolve.initialResidual;
    out.final_residual = r.linearSolve.finalResidual;
    out.residual_available = r.residualAvailable;
    out.divergence_before = r.divergenceBefore;
    out.divergence_after = r.divergenceAfter;
    out.removed_rhs_mean = r.removedRhsMean;
}

void apply_hist(ReflectionMacFluidSolver2D& s, const vgx2r_params& q) {
    s.macCormackSmoke = q.maccormack_smoke != 0;
    s.sourceStrength = q.source_strength;
    s.buoyancy = q.buoyancy;
    s.smokeDecay = q.smoke_decay;
    s.projectIterations = q.project_iterations;
    s.emitterRadius = q.emitter_radius;
    s.emitterFalloff = q.emitter_falloff;
    s.emitterCenterX = q.emitter_center_x;
    s.emitterCenterY = q.emitter_center_y;
    s.emitterSmokeScale = q.emitter_smoke_scale;
    s.emitterKickMultiplier = q.emitter_kick_multiplier;
    s.emitterKickBase = q.emitter_kick_base;
    s.emitterKickScale = q.emitter_kick_scale;
    s.emitterHeight = q.emitter_height;
    s.sorOmega = q.sor_omega;
    s.buoyancySplit = q.buoyancy_split;
    s.buoyancyExp = q.buoyancy_exp;
    s.stirStrength = q.stir_strength;
    s.randomSeed = q.random_seed;
}

void export_hist(const ReflectionMacFluidSolver2D& s, vgx2r_params& q) {
    std::memset(&q, 0, sizeof q);
    q.maccormack_smoke = s.macCormackSmoke;
    q.source_strength = s.sourceStrength;
    q.buoyancy = s.buoyancy;
    q.smoke_decay = s.smokeDecay;
    q.project_iterations = s.projectIterations;
    q.emitter_radius = s.emitterRadius;
    q.emitter_falloff = s.emitterFalloff;
    q.emitter_center_x = s.emitterCenterX;
    q.emitter_center_y = s.emitterCenterY;
    q.emitter_smoke_scale = s.emitterSmokeScale;
    q.emitter_kick_multiplier = s.emitterKickMultiplier;
    q.emitter_kick_base = s.emitterKickBase;
    q.emitter_kick_scale = s.emitterKickScale;
    q.emitter_height = s.emitterHeight;
    q.sor_omega = s.sorOmega;
    q.buoyancy_split = s.buoyancySplit;
    q.buoyancy_exp = s.buoyancyExp;
    q.stir_strength = s.stirStrength;
    q.random_seed = s.randomSeed;
}

std::vector<float>& hist_field(ReflectionMacFluidSolver2D& s, int which) {
    switch (which) {
        case 0: return s.smoke;
        case 1: return s.vx;
        case 2: return s.vy;
        case 3: return s.pressure;
        default: return s.divergence;
    }
}
}  // namespace

extern "C" {

// MacGridParameters2D{} defaults and the MacGridFluidSolver2D() plume.
int r2_params_default(vgx2_params* out) {
    return guarded([&] { export_params(MacGridParameters2D{}, *out); });
}
int r2_params_plume(vgx2_params* out) {
    return guarded([&] {
        MacGridFluidSolver2D s;
        export_params(s.parameters(), *out);
    });
}

int r2_sim_create(int n, float box, const vgx2_params* params, const float* rho, const float* u, const float* v,
                  void** out) {
    return guarded([&] {
        MacGridState2D state(n, box);
        if (rho) std::copy(rho, rho + state.density().size(), state.density().begin());
        if (u) std::copy(u, u + state.velocityX().size(), state.velocityX().begin());
        if (v) std::copy(v, v + state.velocityY().size(), state.velocityY().begin());
        auto* s = new MacGridFluidSolver2D(std::move(state));
        if (params) apply(s->parameters(), *params);
        *out = s;
    });
}
void r2_sim_destroy(void* h) { delete static_cast<MacGridFluidSolver2D*>(h); }
int r2_sim_set_params(void* h, const vgx2_params* p) {
    return guarded([&] { apply(static_cast<MacGridFluidSolver2D*>(h)->parameters(), *p); });
}
int r2_sim_advance(void* h, float dt) {
    return guarded([&] { static_cast<MacGridFluidSolver2D*>(h)->advance(dt); });
}
int r2_sim_reset(void* h) {
    return guarded([&] { static_cast<MacGridFluidSolver2D*>(h)->resetState(); });
}
int r2_sim_set_box(void* h, float box) {
    return guarded([&] { static_cast<MacGridFluidSolver2D*>(h)->setBoxSize(box); });
}
int r2_sim_download(void* h, float* rho, float* u, float* v) {
    return guarded([&] {
        const auto& s = static_cast<MacGridFluidSolver2D*>(h)->state();
        if (rho) std::copy(s.density().begin(), s.density().end(), rho);
        if (u) std::copy(s.velocityX().begin(), s.velocityX().end(), u);
        if (v) std::copy(s.velocityY().begin(), s.velocityY().end(), v);
    });
}
int r2_sim_projection(void* h, int midpoint, vgx_projection_result* out) {
    return guarded([&] {
        auto* s = static_cast<MacGridFluidSolver2D*>(h);
        to_result(midpoint ? s->midpointProjection() : s->lastProjection(), *out);
    });
}
int r2_sim_diagnostics(void* h, vgx_diagnostics* out) {
    return guarded([&] {
        const auto d = static_cast<MacGridFluidSolver2D*>(h)->diagnostics();
        out->kinetic_energy = d.kineticEnergy;
        out->max_velocity = d.maxVelocity;
        out->max_divergence = d.maxDivergence;
    });
}

// SemiLagrangian2D::advect on caller arrays.
int r2_advect(int n, float box, int field, float dt, const float* u, const float* v, const float* src, float* out) {
    return guarded([&] {
        MacGridDomain2D domain(n, box);
        const int w = n + (field == 1), hgt = n + (field == 2);
        const size_t F = (size_t)n * (n + 1), count = (size_t)w * hgt;
        SemiLagrangian2D sl;
        sl.advect({domain, std::span<const float>(u, F), std::span<const float>(v, F),
                   std::span<const float>(src, count), std::span<float>(out, count), dt,
                   static_cast<MacField2D>(field)});
    });
}

// CpuPressureSolver2D::project on caller faces; pressure (n^2 doubles) out.
int r2_project(int n, float box, const vgx_projection_options* o, float* u, float* v, double* pressure,
               vgx_projection_result* result) {
    return guarded([&] {
        MacGridState2D s(n, box);
        std::copy(u, u + s.velocityX().size(), s.velocityX().begin());
        std::copy(v, v + s.velocityY().size(), s.velocityY().begin());
        ProjectionOptions options;
        options.dt = o->dt;
        options.relaxation = o->relaxation;
        options.linearSolve.maxIterations = o->max_iterations;
        options.linearSolve.residualCheckInterval = o->residual_check_interval;
        options.linearSolve.absoluteTolerance = o->absolute_tolerance;
        options.linearSolve.relativeTolerance = o->relative_tolerance;
        options.linearSolve.fixedIterations = o->fixed_iterations != 0;
        CpuPressureSolver2D solver;
        ProjectionResult r;
        try {
            r = solver.project(s, options);
        } catch (...) {
            std::copy(s.velocityX().begin(), s.velocityX().end(), u);
            std::copy(s.velocityY().begin(), s.velocityY().end(), v);
            throw;
        }
        std::copy(s.velocityX().begin(), s.velocityX().end(), u);
        std::copy(s.velocityY().begin(), s.velocityY().end(), v);
        if (pressure) std::copy(solver.pressure().begin(), solver.pressure().end(), pressure);
        if (result) to_result(r, *result);
    });
}

// ---------------------------------------------------------- historical solver
int r2h_params(int n, int smoke_method, vgx2r_params* out) {
    return guarded([&] {
        if (smoke_method < 0) {
            ReflectionMacFluidSolver2D s(n);
            export_hist(s, *out);
        } else {
            ReflectionMacFluidSolver2D s(static_cast<AdvectionMethod>(smoke_method));
            export_hist(s, *out);
        }
    });
}
int r2h_create(int n, const vgx2r_params* params, void** out) {
    return guarded([&] {
        auto* s = new ReflectionMacFluidSolver2D(n);
        if (params) apply_hist(*s, *params);
        *out = s;
    });
}
void r2h_destroy(void* h) { delete static_cast<ReflectionMacFluidSolver2D*>(h); }
int r2h_set_params(void* h, const vgx2r_params* p) {
    return guarded([&] { apply_hist(*static_cast<ReflectionMacFluidSolver2D*>(h), *p); });
}
int r2h_advance(void* h, float dt) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->advance(dt); });
}
int r2h_step(void* h, float dt, float buoyancy, float dissipation, float strength, int iterations) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->step(dt, buoyancy, dissipation, strength, iterations); });
}
int r2h_emit(void* h, float strength, float dt) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->emit(strength, dt); });
}
int r2h_apply_boundary(void* h) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->applyBoundary(); });
}
int r2h_advect_smoke(void* h, float dt) {
    return guarded([&] {
        auto* s = static_cast<ReflectionMacFluidSolver2D*>(h);
        s->advectScalar(s->smoke, dt);
    });
}
int r2h_project(void* h, float dt, int iterations) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->project(dt, iterations); });
}
int r2h_reset(void* h) {
    return guarded([&] { static_cast<ReflectionMacFluidSolver2D*>(h)->resetState(); });
}
int r2h_get_field(void* h, int which, float* out) {
    return guarded([&] {
        auto& f = hist_field(*static_cast<ReflectionMacFluidSolver2D*>(h), which);
        std::copy(f.begin(), f.end(), out);
    });
}
int r2h_set_field(void* h, int which, const float* in) {
    return guarded([&] {
        auto& f = hist_field(*static_cast<ReflectionMacFluidSolver2D*>(h), which);
        std::copy(in, in + f.size(), f.begin());
    });
}
uint32_t r2h_random_state(void* h) { return static_cast<ReflectionMacFluidSolver2D*>(h)->randomState; }
int r2h_diagnostics(void* h, vgx_diagnostics* out) {
    return guarded([&] {
        const auto d = static_cast<ReflectionMacFluidSolver2D*>(h)->diagnostics();
        out->kinetic_energy = d.kineticEnergy;
        out->max_velocity = d.maxVelocity;
        out->max_divergence = d.maxDivergence;
    });
}

}  // extern "C"
