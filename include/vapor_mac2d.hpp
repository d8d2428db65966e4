// vapor_mac2d.hpp - C++ drop-in for vapor's MAC 2D solvers, over the C ABI.
//
// Header-only, like vapor_mac3d.hpp, whose common types (ProjectionOptions,
// ProjectionResult, FluidDiagnostics, Dimension, DomainBounds,
// AdvectionMethod) it shares, as vapor keeps them in src/common. A caller of
// vapor's 2D solvers switches by changing the include and the namespace:
//
//   vapor::hip::MacGridFluidSolver2D solver;          // was MacGridFluidSolver2D
//   solver.parameters().reflection = true;
//   solver.advance(1.f / 60.f);
//   auto rho = solver.state().density();              // std::span<const float>
//
// Classes (src/solvers/mac2d):
//   MacGridDomain2D, MacGridState2D ....... MacGridDomain2D.hpp:7-24, MacGridState2D.hpp:8-31 (host)
//   MacGridParameters2D, MacGridFluidSolver2D  MacGridFluidSolver2D.hpp:6-51 (fields on the GPU)
//   SemiLagrangian2D ....................... SemiLagrangian2D.hpp:4-20 (one GPU call)
//   GpuPressureSolver2D .................... CpuPressureSolver2D.hpp:8-18 (one GPU call)
//   ReflectionMacFluidSolver2D ............. ReflectionMacFluidSolver2D.hpp:19-425 (fields on the GPU)
//
// Same names, defaults and exception types. MacGridFluidSolver2D::state()
// downloads a host snapshot on demand and caches it until the next step.
// vapor's ReflectionMacFluidSolver2D exposes its fields as public vectors
// that callers read and write; here those vectors are host mirrors: pull()
// refreshes them from the device, push() sends edits back, and every method
// leaves the device state authoritative. Link with libvapor_mac3d.so.
#pragma once

#include "vapor_mac2d.h"
#include "vapor_mac3d.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <span>
#include <stdexcept>
#include <vector>

namespace vapor::hip {

// MacField2D (SemiLagrangian2D.hpp:4).
enum class MacField2D { Density = VGX2_DENSITY, VelocityX = VGX2_VELOCITY_X, VelocityY = VGX2_VELOCITY_Y };

// Square grid in the xy plane with MacGridDomain2D's checks and messages.
class MacGridDomain2D {
public:
    explicit MacGridDomain2D(int resolution, float boxSize = 4.f) : n_(resolution) {
        if (resolution <= 0) throw std::invalid_argument("2D MAC resolution must be positive");
        const auto n = static_cast<std::size_t>(resolution);
        if (n > static_cast<std::size_t>(2147483647) / (n + 1))
            throw std::length_error("2D MAC grid exceeds the supported index range");
        setBoxSize(boxSize);
    }
    Dimension dimension() const noexcept { return Dimension::D2; }
    DomainBounds bounds() const noexcept { return {{0.f, 0.f, 0.f}, {box_, box_, 0.f}}; }
    int resolution() const noexcept { return n_; }
    float boxSize() const noexcept { return box_; }
    float cellSize() const noexcept { return box_ / n_; }
    std::size_t idC(int x, int y) const noexcept { return x + std::size_t(n_) * y; }
    std::size_t idX(int x, int y) const noexcept { return x + std::size_t(n_ + 1) * y; }
    std::size_t idY(int x, int y) const noexcept { return idC(x, y); }
    void setBoxSize(float size) {
        if (!std::isfinite(size) || size <= 0.f || size / n_ <= 0.f || !std::isfinite(1.f / (size / n_)))
            throw std::invalid_argument("2D MAC requires finite positive size and spacing");
        box_ = size;
    }

private:
    int n_;
    float box_ = 4.f;
};

// Host fields: n^2 density, (n+1)n x faces, n(n+1) y faces.
class MacGridState2D {
public:
    explicit MacGridState2D(int resolution, float boxSize = 4.f) : MacGridState2D(MacGridDomain2D(resolution, boxSize)) {}
    explicit MacGridState2D(MacGridDomain2D domain) : domain_(domain) {
        const auto n = static_cast<std::size_t>(domain_.resolution());
        density_.assign(n * n, 0.f);
        u_.assign((n + 1) * n, 0.f);
        v_.assign(n * (n + 1), 0.f);
    }
    const MacGridDomain2D& domain() const noexcept { return domain_; }
    int resolution() const noexcept { return domain_.resolution(); }
    float boxSize() const noexcept { return domain_.boxSize(); }
    float cellSize() const noexcept { return domain_.cellSize(); }
    std::size_t cellCount() const noexcept { return density_.size(); }
    std::size_t idC(int x, int y) const noexcept { return domain_.idC(x, y); }
    std::size_t idX(int x, int y) const noexcept { return domain_.idX(x, y); }
    std::size_t idY(int x, int y) const noexcept { return domain_.idY(x, y); }
    std::span<float> density() noexcept { return density_; }
    std::span<const float> density() const noexcept { return density_; }
    std::span<float> velocityX() noexcept { return u_; }
    std::span<const float> velocityX() const noexcept { return u_; }
    std::span<float> velocityY() noexcept { return v_; }
    std::span<const float> velocityY() const noexcept { return v_; }
    void reset() noexcept {
        for (auto* f : {&density_, &u_, &v_}) std::fill(f->begin(), f->end(), 0.f);
    }
    void setBoxSize(float size) {
        const float previous = boxSize();
        domain_.setBoxSize(size);
        if (size != previous) reset();
    }

private:
    MacGridDomain2D domain_;
    std::vector<float> density_, u_, v_;
};

// MacGridParameters2D (MacGridFluidSolver2D.hpp:6-16).
struct MacGridParameters2D {
    bool reflection = false;
    float sourceStrength = 4.f;
    float emitterKickMultiplier = 1.f;
    float emitterRadius = .08f;
    float emitterCenterX = .5f, emitterCenterY = .2f;
    float buoyancy = 3.f;
    float smokeDecay = .08f;
    float sorOmega = 1.9f;
    int projectIterations = 600;

    vgx2_params to_c() const {
        vgx2_params p{};
        p.reflection = reflection;
        p.source_strength = sourceStrength;
        p.emitter_kick_multiplier = emitterKickMultiplier;
        p.emitter_radius = emitterRadius;
        p.emitter_center_x = emitterCenterX;
        p.emitter_center_y = emitterCenterY;
        p.buoyancy = buoyancy;
        p.smoke_decay = smokeDecay;
        p.sor_omega = sorOmega;
        p.project_iterations = projectIterations;
        return p;
    }
};

// MacGridFluidSolver2D (MacGridFluidSolver2D.hpp:19-51) on the GPU.
class MacGridFluidSolver2D {
public:
    // The ready-to-run plume: 128^2, box 1 (MacGridFluidSolver2D.cpp:8-14).
    MacGridFluidSolver2D() : MacGridFluidSolver2D(128, 1.f) {
        parameters_.emitterRadius = .04f;
        parameters_.sourceStrength = 1.f;
        parameters_.buoyancy = .8f;
        parameters_.smokeDecay = .3007525f;
        parameters_.projectIterations = 120;
    }
    explicit MacGridFluidSolver2D(int resolution, float boxSize = 4.f) : MacGridFluidSolver2D(MacGridState2D(resolution, boxSize)) {}
    explicit MacGridFluidSolver2D(MacGridState2D state) : host_(std::move(state)) {
        const vgx2_params p = parameters_.to_c();
        throw_status(vgx2_sim_create(host_.resolution(), host_.boxSize(), &p, &sim_));
        throw_status(vgx2_sim_upload(sim_, host_.density().data(), host_.velocityX().data(), host_.velocityY().data()));
        fresh_ = true;
    }
    MacGridFluidSolver2D(const MacGridFluidSolver2D&) = delete;
    MacGridFluidSolver2D& operator=(const MacGridFluidSolver2D&) = delete;
    MacGridFluidSolver2D(MacGridFluidSolver2D&& o) noexcept { *this = std::move(o); }
    MacGridFluidSolver2D& operator=(MacGridFluidSolver2D&& o) noexcept {
        std::swap(sim_, o.sim_);
        std::swap(host_, o.host_);
        std::swap(parameters_, o.parameters_);
        std::swap(fresh_, o.fresh_);
        return *this;
    }
    ~MacGridFluidSolver2D() { vgx2_sim_destroy(sim_); }

    void advance(float dt) {
        const vgx2_params p = parameters_.to_c();
        throw_status(vgx2_sim_set_params(sim_, &p));
        fresh_ = false;
        throw_status(vgx2_sim_advance(sim_, dt));
    }
    void resetState() {
        throw_status(vgx2_sim_reset(sim_));
        fresh_ = false;
    }
    const MacGridDomain2D& domain() const noexcept { return host_.domain(); }
    FluidDiagnostics diagnostics() const {
        vgx_diagnostics d{};
        throw_status(vgx2_sim_diagnostics(sim_, &d));
        return {d.kinetic_energy, d.max_velocity, d.max_divergence};
    }
    // Host snapshot of the device fields (downloaded once per step).
    const MacGridState2D& state() const {
        if (!fresh_) {
            throw_status(vgx2_sim_download(sim_, host_.density().data(), host_.velocityX().data(),
                                           host_.velocityY().data()));
            fresh_ = true;
        }
        return host_;
    }
    MacGridParameters2D& parameters() noexcept { return parameters_; }
    const MacGridParameters2D& parameters() const noexcept { return parameters_; }
    ProjectionResult lastProjection() const {
        vgx_projection_result r{};
        throw_status(vgx2_sim_last_projection(sim_, &r));
        return projection_result(r);
    }
    ProjectionResult midpointProjection() const {
        vgx_projection_result r{};
        throw_status(vgx2_sim_midpoint_projection(sim_, &r));
        return projection_result(r);
    }
    float boxSize() const noexcept { return host_.boxSize(); }
    void setBoxSize(float size) {
        throw_status(vgx2_sim_set_box(sim_, size));
        host_.setBoxSize(size);
        fresh_ = false;
    }
    // The device density (imageData's borrowed buffer, MacRenderData2D).
    const float* deviceDensity() const {
        const float* rho = nullptr;
        throw_status(vgx2_sim_device_fields(sim_, &rho, nullptr, nullptr));
        return rho;
    }

private:
    vgx2_sim* sim_ = nullptr;
    mutable MacGridState2D host_{1, 1.f};
    MacGridParameters2D parameters_;
    mutable bool fresh_ = false;
};

// MacAdvectionOperation2D (SemiLagrangian2D.hpp:8-15).
struct MacAdvectionOperation2D {
    const MacGridDomain2D& domain;
    std::span<const float> velocityX, velocityY;
    std::span<const float> source;
    std::span<float> output;
    float dt;
    MacField2D field;
};

// SemiLagrangian2D::advect on the GPU (host spans in and out).
class SemiLagrangian2D {
public:
    void advect(const MacAdvectionOperation2D& op) const {
        const int n = op.domain.resolution();
        const auto f = static_cast<std::size_t>(n) * (n + 1);
        const int w = n + (op.field == MacField2D::VelocityX), h = n + (op.field == MacField2D::VelocityY);
        if (op.velocityX.size() != f || op.velocityY.size() != f || op.source.size() != std::size_t(w) * h ||
            op.output.size() != op.source.size())
            throw std::invalid_argument("Invalid 2D advection timestep, extent, or overlapping output");
        throw_status(vgx2_backend_advect(n, op.domain.boxSize(), static_cast<int>(op.field), op.dt,
                                         op.velocityX.data(), op.velocityY.data(), op.source.data(),
                                         op.output.data()));
    }
};

// CpuPressureSolver2D's interface with the solve on the GPU; pressure() is
// the FP64 workspace of the last successful projection.
class GpuPressureSolver2D {
public:
    ProjectionResult project(MacGridState2D& state, const ProjectionOptions& options) {
        vgx_projection_options o;
        vgx_projection_options_default(&o);
        o.dt = options.dt;
        o.relaxation = options.relaxation;
        o.max_iterations = options.linearSolve.maxIterations;
        o.residual_check_interval = options.linearSolve.residualCheckInterval;
        o.absolute_tolerance = options.linearSolve.absoluteTolerance;
        o.relative_tolerance = options.linearSolve.relativeTolerance;
        o.fixed_iterations = options.linearSolve.fixedIterations;
        std::vector<double> p(state.cellCount());
        vgx_projection_result r{};
        throw_status(vgx2_backend_project(state.resolution(), state.boxSize(), &o, state.velocityX().data(),
                                          state.velocityY().data(), p.data(), &r));
        pressure_ = std::move(p);
        return projection_result(r);
    }
    void reset() noexcept { std::fill(pressure_.begin(), pressure_.end(), 0.0); }
    std::span<const double> pressure() const noexcept { return pressure_; }

private:
    std::vector<double> pressure_;
};

// ReflectionMacFluidSolver2D (ReflectionMacFluidSolver2D.hpp:19-425) on the
// GPU. Parameters are public members with the reference's names; the field
// vectors are host mirrors (pull / push).
class ReflectionMacFluidSolver2D {
public:
    bool macCormackSmoke = true;
    float sourceStrength = 1.f, buoyancy = .8f;
    float smokeDecay = .3007525f;
    int projectIterations = 60;
    int n = 256;
    float emitterRadius = 0.08f;
    float emitterFalloff = 3.5f;
    float emitterCenterX = 0.5f, emitterCenterY = 0.2f;
    float emitterSmokeScale = 3.0f;
    float emitterKickMultiplier = 1.f;
    float emitterKickBase = 0.6f;
    float emitterKickScale = 1.0f;
    int emitterHeight = 4;
    float sorOmega = 1.9f;
    float buoyancySplit = 0.5f;
    float buoyancyExp = 0.25f;
    float stirStrength = 0.f;
    std::uint32_t randomSeed = 0x9E3779B9u;
    std::vector<float> smoke, vx, vy, pressure, divergence;

    // The historical plume: 128^2, radius 0.04, 120 iterations (:78-83).
    explicit ReflectionMacFluidSolver2D(AdvectionMethod smokeMethod = AdvectionMethod::SemiLagrangian)
        : ReflectionMacFluidSolver2D(128) {
        macCormackSmoke = smokeMethod == AdvectionMethod::MacCormack;
        emitterRadius = .04f;
        projectIterations = 120;
    }
    explicit ReflectionMacFluidSolver2D(int res) {
        if (res < 2 || res > 512) throw std::invalid_argument("2D reflection resolution must be 2..512");
        n = res;
        const vgx2r_params p = to_c();
        throw_status(vgx2r_sim_create(n, &p, &sim_));
        pull();
    }
    ReflectionMacFluidSolver2D(const ReflectionMacFluidSolver2D&) = delete;
    ReflectionMacFluidSolver2D& operator=(const ReflectionMacFluidSolver2D&) = delete;
    ~ReflectionMacFluidSolver2D() { vgx2r_sim_destroy(sim_); }

    ReflectionMacFluidSolver2D& parameters() noexcept { return *this; }
    float boxSize() const noexcept { return 1.f; }
    void setBoxSize(float size) { throw_status(vgx2r_sim_set_box(sim_, size)); }
    float h() const { return 1.0f / n; }
    int idC(int x, int y) const { return x + n * y; }
    int idX(int x, int y) const { return x + (n + 1) * y; }
    int idY(int x, int y) const { return x + n * y; }

    void resetState() {
        sync_params();
        throw_status(vgx2r_sim_reset(sim_));
        pull();
    }
    void advance(float dt) {
        sync_params();
        throw_status(vgx2r_sim_advance(sim_, dt));
    }
    void step(float dt, float b, float dissipation, float strength, int iterations) {
        sync_params();
        throw_status(vgx2r_sim_step(sim_, dt, b, dissipation, strength, iterations));
    }
    void emit(float strength, float dt) {
        sync_params();
        throw_status(vgx2r_sim_emit(sim_, strength, dt));
    }
    void applyBoundary() { throw_status(vgx2r_sim_apply_boundary(sim_)); }
    // advectScalar on the solver's own smoke (the reference takes any field).
    void advectSmoke(float dt) {
        sync_params();
        throw_status(vgx2r_sim_advect_smoke(sim_, dt));
    }
    void project(float dt, int iterations) {
        sync_params();
        throw_status(vgx2r_sim_project(sim_, dt, iterations));
    }
    FluidDiagnostics diagnostics() const {
        vgx_diagnostics d{};
        throw_status(vgx2r_sim_diagnostics(sim_, &d));
        return {d.kinetic_energy, d.max_velocity, d.max_divergence};
    }
    std::uint32_t randomState() const {
        std::uint32_t s = 0;
        throw_status(vgx2r_sim_random_state(sim_, &s));
        return s;
    }
    // Device -> host mirrors, and edited mirrors -> device.
    void pull() {
        const std::size_t N = std::size_t(n) * n, F = std::size_t(n) * (n + 1);
        smoke.resize(N);
        vx.resize(F);
        vy.resize(F);
        pressure.resize(N);
        divergence.resize(N);
        throw_status(vgx2r_sim_download(sim_, smoke.data(), vx.data(), vy.data(), pressure.data(), divergence.data()));
    }
    void push() {
        throw_status(vgx2r_sim_upload(sim_, smoke.data(), vx.data(), vy.data(), pressure.data(), divergence.data()));
    }

private:
    vgx2r_params to_c() const {
        vgx2r_params p{};
        p.maccormack_smoke = macCormackSmoke;
        p.source_strength = sourceStrength;
        p.buoyancy = buoyancy;
        p.smoke_decay = smokeDecay;
        p.project_iterations = projectIterations;
        p.emitter_radius = emitterRadius;
        p.emitter_falloff = emitterFalloff;
        p.emitter_center_x = emitterCenterX;
        p.emitter_center_y = emitterCenterY;
        p.emitter_smoke_scale = emitterSmokeScale;
        p.emitter_kick_multiplier = emitterKickMultiplier;
        p.emitter_kick_base = emitterKickBase;
        p.emitter_kick_scale = emitterKickScale;
        p.emitter_height = emitterHeight;
        p.sor_omega = sorOmega;
        p.buoyancy_split = buoyancySplit;
        p.buoyancy_exp = buoyancyExp;
        p.stir_strength = stirStrength;
        p.random_seed = randomSeed;
        return p;
    }
    void sync_params() {
        const vgx2r_params p = to_c();
        throw_status(vgx2r_sim_set_params(sim_, &p));
    }
    vgx2r_sim* sim_ = nullptr;
};

}  // namespace vapor::hip
