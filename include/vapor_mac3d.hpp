// vapor_mac3d.hpp - C++ drop-in for vapor's MAC 3D solver, over the C ABI.
//
// Header-only. Mirrors the reference interface so a caller of
// MacGridFluidSolver3D (src/solvers/mac3d/MacGridFluidSolver3D.hpp:39-73) or
// GpuMacSimulation3D (src/solvers/mac3d/gpu/GpuMacSimulation3D.hpp:9-27) can
// switch by changing the include and the class's namespace:
//
//   vapor::hip::MacGridFluidSolver3D solver(64);     // was MacGridFluidSolver3D
//   solver.parameters().emitterRadius = .08f;
//   solver.advance(1.f / 60.f);
//   auto rho = solver.state().density();             // std::span<const float>
//
// Same names, same defaults, same exception types (std::invalid_argument,
// std::length_error, std::runtime_error for device failures). The fields live
// in GPU memory; state() downloads a host snapshot on demand and caches it
// until the next advance. Link with vapor_amd/_lib/libvapor_mac3d.so.
#pragma once

#include "vapor_mac3d.h"

#include <cmath>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vapor::hip {

// MacAlgorithm (src/common/MacAlgorithm.hpp:3-9).
enum class MacAlgorithm {
    SimpleSemiLagrangian = VGX_SIMPLE_SEMI_LAGRANGIAN,
    SimpleMacCormack = VGX_SIMPLE_MACCORMACK,
    ReflectionSemiLagrangian = VGX_REFLECTION_SEMI_LAGRANGIAN,
    ReflectionMacCormack = VGX_REFLECTION_MACCORMACK
};

inline void throw_status(int status) {
    if (status == VGX_OK) return;
    const std::string message = vgx_last_error();
    if (status == VGX_EINVAL) throw std::invalid_argument(message);
    if (status == VGX_ERANGE) throw std::length_error(message);
    throw std::runtime_error(message);
}

// MacGridParameters3D (MacGridFluidSolver3D.hpp:14-36) with the same field
// names and defaults, plus the extension controls of vgx_params.
struct MacGridParameters3D {
    MacGridParameters3D() {
        vgx_params p;
        vgx_params_default(&p);
        load(p);
    }
    explicit MacGridParameters3D(MacAlgorithm algorithm) {
        vgx_params p;
        throw_status(vgx_params_preset(static_cast<int>(algorithm), &p));
        load(p);
    }
    float emitterRadius, emitterCenterX, emitterCenterY, emitterCenterZ;
    float emitterKickMultiplier, emitterKickBase, emitterKickScale;
    float stirStrength;
    std::uint32_t randomSeed;
    float sorOmega, buoyancySplit, cflMc;
    bool macCormackSmoke, macCormackVel;
    float smokeDecay;
    bool reflection;
    float buoyancy, sourceStrength;
    int projectIterations;
    // Extensions (no reference counterpart): see vgx_params.
    vgx_params extension{};

    vgx_params to_c() const {
        vgx_params p = extension;
        p.emitter_radius = emitterRadius;
        p.emitter_center_x = emitterCenterX;
        p.emitter_center_y = emitterCenterY;
        p.emitter_center_z = emitterCenterZ;
        p.emitter_kick_multiplier = emitterKickMultiplier;
        p.emitter_kick_base = emitterKickBase;
        p.emitter_kick_scale = emitterKickScale;
        p.stir_strength = stirStrength;
        p.random_seed = randomSeed;
        p.sor_omega = sorOmega;
        p.buoyancy_split = buoyancySplit;
        p.cfl_mc = cflMc;
        p.maccormack_smoke = macCormackSmoke;
        p.maccormack_vel = macCormackVel;
        p.smoke_decay = smokeDecay;
        p.reflection = reflection;
        p.buoyancy = buoyancy;
        p.source_strength = sourceStrength;
        p.project_iterations = projectIterations;
        return p;
    }
    void load(const vgx_params& p) {
        extension = p;
        emitterRadius = p.emitter_radius;
        emitterCenterX = p.emitter_center_x;
        emitterCenterY = p.emitter_center_y;
        emitterCenterZ = p.emitter_center_z;
        emitterKickMultiplier = p.emitter_kick_multiplier;
        emitterKickBase = p.emitter_kick_base;
        emitterKickScale = p.emitter_kick_scale;
        stirStrength = p.stir_strength;
        randomSeed = p.random_seed;
        sorOmega = p.sor_omega;
        buoyancySplit = p.buoyancy_split;
        cflMc = p.cfl_mc;
        macCormackSmoke = p.maccormack_smoke != 0;
        macCormackVel = p.maccormack_vel != 0;
        smokeDecay = p.smoke_decay;
        reflection = p.reflection != 0;
        buoyancy = p.buoyancy;
        sourceStrength = p.source_strength;
        projectIterations = p.project_iterations;
    }
};

// ProjectionResult / LinearSolveResult (ProjectionOptions.hpp:11-17).
struct LinearSolveResult {
    int iterations = 0;
    double initialResidual = 0.0;
    double finalResidual = 0.0;
    bool converged = false;
};
struct ProjectionResult {
    LinearSolveResult linearSolve;
    bool residualAvailable = false;
    float divergenceBefore = 0.f;
    float divergenceAfter = 0.f;
    double removedRhsMean = 0.0;
    vgx_projection_result extension{};  // L2 residuals, checks, warm start
};

// FluidDiagnostics (FluidDiagnostics.hpp:3-7).
struct FluidDiagnostics {
    float kineticEnergy = 0.f;
    float maxVelocity = 0.f;
    float maxDivergence = 0.f;
};

// Host snapshot of the physical fields with MacGridState3D's accessors
// (MacGridState3D.hpp:11-51).
class MacGridState3D {
public:
    explicit MacGridState3D(int resolution, float boxSize = 4.5f) : n_(resolution), box_(boxSize) {
        if (resolution <= 0) throw std::invalid_argument("MAC grid resolution must be positive");
        const std::size_t n = static_cast<std::size_t>(resolution), limit = 2147483647u;
        if (n > limit / n || n * n > limit / (n + 1))
            throw std::length_error("MAC grid exceeds the supported index range");
        if (!std::isfinite(boxSize) || boxSize <= 0.f || boxSize / resolution <= 0.f)
            throw std::invalid_argument("MAC grid requires a finite positive box size and spacing");
        density_.assign(n * n * n, 0.f);
        vx_.assign((n + 1) * n * n, 0.f);
        vy_.assign((n + 1) * n * n, 0.f);
        vz_.assign((n + 1) * n * n, 0.f);
    }
    int resolution() const noexcept { return n_; }
    float boxSize() const noexcept { return box_; }
    float cellSize() const noexcept { return box_ / n_; }
    std::size_t cellCount() const noexcept { return density_.size(); }
    std::size_t idC(int x, int y, int z) const noexcept { return x + std::size_t(n_) * (y + std::size_t(n_) * z); }
    std::size_t idX(int x, int y, int z) const noexcept { return x + std::size_t(n_ + 1) * (y + std::size_t(n_) * z); }
    std::size_t idY(int x, int y, int z) const noexcept { return x + std::size_t(n_) * (y + std::size_t(n_ + 1) * z); }
    std::size_t idZ(int x, int y, int z) const noexcept { return idC(x, y, z); }
    std::span<float> density() noexcept { return density_; }
    std::span<const float> density() const noexcept { return density_; }
    std::span<float> velocityX() noexcept { return vx_; }
    std::span<const float> velocityX() const noexcept { return vx_; }
    std::span<float> velocityY() noexcept { return vy_; }
    std::span<const float> velocityY() const noexcept { return vy_; }
    std::span<float> velocityZ() noexcept { return vz_; }
    std::span<const float> velocityZ() const noexcept { return vz_; }

private:
    int n_;
    float box_;
    std::vector<float> density_, vx_, vy_, vz_;
};

// The solver. Construction mirrors MacGridFluidSolver3D: (MacAlgorithm) gives
// 32^3, box 4.5 and the preset; (int res) gives the defaults with
// projectIterations = 2 res; (int res, float box, parameters) is the
// GpuMacSimulation3D form.
class MacGridFluidSolver3D {
public:
    explicit MacGridFluidSolver3D(MacAlgorithm algorithm)
        : MacGridFluidSolver3D(32, 4.5f, MacGridParameters3D(algorithm)) {}
    explicit MacGridFluidSolver3D(int res) : n_(res) {
        throw_status(vgx_sim_create(res, 4.5f, nullptr, &sim_));
        vgx_params p;
        vgx_sim_get_params(sim_, &p);
        parameters_.load(p);
    }
    MacGridFluidSolver3D(int res, float boxSize, const MacGridParameters3D& parameters)
        : n_(res), parameters_(parameters) {
        const vgx_params p = parameters.to_c();
        throw_status(vgx_sim_create(res, boxSize, &p, &sim_));
    }
    // Adopts prefilled fields without resetting, as MacGridFluidSolver3D(MacGridState3D).
    explicit MacGridFluidSolver3D(const MacGridState3D& state) : n_(state.resolution()) {
        throw_status(vgx_sim_create(state.resolution(), state.boxSize(), nullptr, &sim_));
        vgx_params p;
        vgx_sim_get_params(sim_, &p);
        parameters_.load(p);
        throw_status(vgx_sim_upload(sim_, state.density().data(), state.velocityX().data(),
                                    state.velocityY().data(), state.velocityZ().data()));
    }
    ~MacGridFluidSolver3D() { vgx_sim_destroy(sim_); }
    MacGridFluidSolver3D(const MacGridFluidSolver3D&) = delete;
    MacGridFluidSolver3D& operator=(const MacGridFluidSolver3D&) = delete;

    MacGridParameters3D& parameters() noexcept { return parameters_; }
    const MacGridParameters3D& parameters() const noexcept { return parameters_; }

    void advance(float dt) {
        push();
        throw_status(vgx_sim_advance(sim_, dt));
        snapshot_valid_ = false;
    }
    void step(float dt, float buoyancy, float sourceStrength, int projectIterations) {
        with_overrides(true, buoyancy, sourceStrength, projectIterations, dt);
    }
    void stepNoReflect(float dt, float buoyancy, float sourceStrength, int projectIterations) {
        with_overrides(false, buoyancy, sourceStrength, projectIterations, dt);
    }
    void resetState() {
        push();
        throw_status(vgx_sim_reset(sim_));
        snapshot_valid_ = false;
    }
    void setBoxSize(float size) {
        push();
        throw_status(vgx_sim_set_box(sim_, size));
        snapshot_valid_ = false;
    }
    int resolution() const { return n_; }
    float boxSize() const noexcept { return vgx_sim_box_size(sim_); }
    const ProjectionResult& lastProjection() const {
        vgx_projection_result r;
        throw_status(vgx_sim_last_projection(sim_, 0, &r));
        last_.linearSolve = {r.iterations, r.initial_residual, r.final_residual, r.converged != 0};
        last_.residualAvailable = r.residual_available != 0;
        last_.divergenceBefore = r.divergence_before;
        last_.divergenceAfter = r.divergence_after;
        last_.removedRhsMean = r.removed_rhs_mean;
        last_.extension = r;
        return last_;
    }
    FluidDiagnostics diagnostics() const {
        vgx_diagnostics d;
        throw_status(vgx_sim_diagnostics(sim_, &d));
        return {d.kinetic_energy, d.max_velocity, d.max_divergence};
    }
    float kineticEnergy() const { return diagnostics().kineticEnergy; }
    float maxVelocity() const { return diagnostics().maxVelocity; }
    float computeDivergenceNorm() const { return diagnostics().maxDivergence; }

    // Host snapshot of the fields (downloaded on first use after a change).
    const MacGridState3D& state() const {
        if (!snapshot_valid_ || !snapshot_ || snapshot_->boxSize() != boxSize()) {
            snapshot_ = std::make_unique<MacGridState3D>(n_, boxSize());
            throw_status(vgx_sim_download(sim_, snapshot_->density().data(), snapshot_->velocityX().data(),
                                          snapshot_->velocityY().data(), snapshot_->velocityZ().data()));
            snapshot_valid_ = true;
        }
        return *snapshot_;
    }
    // Extension: sigma * N(0, 1) added to every face velocity (see
    // vgx_sim_add_velocity_noise).
    void addVelocityNoise(float sigma, std::uint64_t seed) {
        throw_status(vgx_sim_add_velocity_noise(sim_, sigma, seed));
        snapshot_valid_ = false;
    }
    void upload(const MacGridState3D& state) {
        if (state.resolution() != n_ || state.boxSize() != boxSize())
            throw std::invalid_argument("GPU snapshot box mismatch");
        throw_status(vgx_sim_upload(sim_, state.density().data(), state.velocityX().data(),
                                    state.velocityY().data(), state.velocityZ().data()));
        snapshot_valid_ = false;
    }
    void download(MacGridState3D& state) const {
        if (state.resolution() != n_ || state.boxSize() != boxSize())
            throw std::invalid_argument("GPU snapshot box mismatch");
        throw_status(vgx_sim_download(sim_, state.density().data(), state.velocityX().data(),
                                      state.velocityY().data(), state.velocityZ().data()));
    }
    std::uint64_t uploadedBytes() const {
        std::uint64_t up = 0;
        vgx_sim_transfer_bytes(sim_, &up, nullptr);
        return up;
    }
    std::uint64_t downloadedBytes() const {
        std::uint64_t down = 0;
        vgx_sim_transfer_bytes(sim_, nullptr, &down);
        return down;
    }
    vgx_sim* handle() noexcept { return sim_; }

private:
    void push() {
        const vgx_params p = parameters_.to_c();
        throw_status(vgx_sim_set_params(sim_, &p));
    }
    void with_overrides(bool reflection, float buoyancy, float source, int iterations, float dt) {
        MacGridParameters3D saved = parameters_;
        parameters_.reflection = reflection;
        parameters_.buoyancy = buoyancy;
        parameters_.sourceStrength = source;
        parameters_.projectIterations = iterations;
        try {
            advance(dt);
        } catch (...) {
            parameters_ = saved;
            throw;
        }
        parameters_ = saved;
    }

    int n_;
    vgx_sim* sim_ = nullptr;
    MacGridParameters3D parameters_;
    mutable ProjectionResult last_;
    mutable std::unique_ptr<MacGridState3D> snapshot_;
    mutable bool snapshot_valid_ = false;
};

}  // namespace vapor::hip
