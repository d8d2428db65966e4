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
//
// Stages install as in vapor (MacGridFluidSolver3D.hpp:54-57): a Projection3D
// or Advection3D passed to setPressureSolver / setAdvection replaces that
// stage of every later step and receives host spans, while the rest of the
// step stays on the device; useDefault* returns to the device stage. The
// staged GPU stages GpuPressureSolver / GpuAdvection (over a GpuMacBackend,
// as vapor's gpu/ classes) are such stages, usable inside vapor's own solver
// too.
#pragma once

#include "vapor_mac3d.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <functional>
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

// AdvectionMethod (src/common/MacAlgorithm.hpp:9), FaceAxis
// (AdvectionContext.hpp:9), Dimension / DomainBounds (common/DomainInfo.hpp).
enum class AdvectionMethod { SemiLagrangian = VGX_SEMI_LAGRANGIAN, MacCormack = VGX_MACCORMACK };
enum class FaceAxis { X = VGX_AXIS_X, Y = VGX_AXIS_Y, Z = VGX_AXIS_Z };
enum class Dimension { D2 = 2, D3 = 3 };
struct Vec3 {
    float x, y, z;
};
struct DomainBounds {
    Vec3 min;
    Vec3 max;
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

// ProjectionOptions / LinearSolveOptions (ProjectionOptions.hpp:5-9,
// LinearSolveOptions.hpp:4-10).
struct LinearSolveOptions {
    int maxIterations = 200;
    int residualCheckInterval = 8;
    double absoluteTolerance = 1e-6;
    double relativeTolerance = 1e-5;
    bool fixedIterations = false;
};
struct ProjectionOptions {
    float dt = 1.f / 60.f;
    float relaxation = 1.95f;
    LinearSolveOptions linearSolve;
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
inline ProjectionResult projection_result(const vgx_projection_result& r) {
    ProjectionResult out;
    out.linearSolve = {r.iterations, r.initial_residual, r.final_residual, r.converged != 0};
    out.residualAvailable = r.residual_available != 0;
    out.divergenceBefore = r.divergence_before;
    out.divergenceAfter = r.divergence_after;
    out.removedRhsMean = r.removed_rhs_mean;
    out.extension = r;
    return out;
}
inline vgx_projection_result projection_record(const ProjectionResult& p) {
    vgx_projection_result r = p.extension;
    r.iterations = p.linearSolve.iterations;
    r.converged = p.linearSolve.converged;
    r.initial_residual = p.linearSolve.initialResidual;
    r.final_residual = p.linearSolve.finalResidual;
    r.residual_available = p.residualAvailable;
    r.divergence_before = p.divergenceBefore;
    r.divergence_after = p.divergenceAfter;
    r.removed_rhs_mean = p.removedRhsMean;
    return r;
}

// FluidDiagnostics (FluidDiagnostics.hpp:3-7).
struct FluidDiagnostics {
    float kineticEnergy = 0.f;
    float maxVelocity = 0.f;
    float maxDivergence = 0.f;
};

// Geometry and indexing of the cubic grid (MacGridDomain3D.hpp:9-37,
// MacGridDomain3D.cpp:7-23), with the same validation and messages.
class MacGridDomain3D {
public:
    explicit MacGridDomain3D(int resolution, float boxSize = 4.5f) : n_(resolution) {
        if (resolution <= 0) throw std::invalid_argument("MAC grid resolution must be positive");
        const std::size_t n = static_cast<std::size_t>(resolution), limit = 2147483647u;
        if (n > limit / n || n * n > limit / (n + 1))
            throw std::length_error("MAC grid exceeds the supported index range");
        setBoxSize(boxSize);
    }
    Dimension dimension() const noexcept { return Dimension::D3; }
    DomainBounds bounds() const noexcept { return {{0.f, 0.f, 0.f}, {box_, box_, box_}}; }
    int resolution() const noexcept { return n_; }
    float boxSize() const noexcept { return box_; }
    float cellSize() const noexcept { return box_ / n_; }
    std::size_t cellCount() const noexcept { return std::size_t(n_) * n_ * n_; }
    std::size_t faceCount() const noexcept { return std::size_t(n_ + 1) * n_ * n_; }
    std::size_t idC(int x, int y, int z) const noexcept { return x + std::size_t(n_) * (y + std::size_t(n_) * z); }
    std::size_t idX(int x, int y, int z) const noexcept { return x + std::size_t(n_ + 1) * (y + std::size_t(n_) * z); }
    std::size_t idY(int x, int y, int z) const noexcept { return x + std::size_t(n_) * (y + std::size_t(n_ + 1) * z); }
    std::size_t idZ(int x, int y, int z) const noexcept { return idC(x, y, z); }
    void setBoxSize(float size) {
        if (!std::isfinite(size) || size <= 0.f || size / n_ <= 0.f)
            throw std::invalid_argument("MAC grid requires a finite positive box size and spacing");
        box_ = size;
    }

private:
    int n_;
    float box_ = 4.5f;
};

// Host snapshot of the physical fields with MacGridState3D's accessors
// (MacGridState3D.hpp:11-51).
class MacGridState3D {
public:
    explicit MacGridState3D(int resolution, float boxSize = 4.5f) : MacGridState3D(MacGridDomain3D(resolution, boxSize)) {}
    explicit MacGridState3D(MacGridDomain3D domain) : domain_(domain) {
        density_.assign(domain_.cellCount(), 0.f);
        vx_.assign(domain_.faceCount(), 0.f);
        vy_.assign(domain_.faceCount(), 0.f);
        vz_.assign(domain_.faceCount(), 0.f);
    }
    const MacGridDomain3D& domain() const noexcept { return domain_; }
    int resolution() const noexcept { return domain_.resolution(); }
    float boxSize() const noexcept { return domain_.boxSize(); }
    float cellSize() const noexcept { return domain_.cellSize(); }
    std::size_t cellCount() const noexcept { return density_.size(); }
    std::size_t idC(int x, int y, int z) const noexcept { return domain_.idC(x, y, z); }
    std::size_t idX(int x, int y, int z) const noexcept { return domain_.idX(x, y, z); }
    std::size_t idY(int x, int y, int z) const noexcept { return domain_.idY(x, y, z); }
    std::size_t idZ(int x, int y, int z) const noexcept { return domain_.idZ(x, y, z); }
    std::span<float> density() noexcept { return density_; }
    std::span<const float> density() const noexcept { return density_; }
    std::span<float> velocityX() noexcept { return vx_; }
    std::span<const float> velocityX() const noexcept { return vx_; }
    std::span<float> velocityY() noexcept { return vy_; }
    std::span<const float> velocityY() const noexcept { return vy_; }
    std::span<float> velocityZ() noexcept { return vz_; }
    std::span<const float> velocityZ() const noexcept { return vz_; }
    void reset() {
        for (auto* f : {&density_, &vx_, &vy_, &vz_}) std::fill(f->begin(), f->end(), 0.f);
    }
    // A changed size resets the fields (MacGridState3D.cpp:23-29).
    void setBoxSize(float size) {
        const float previous = domain_.boxSize();
        domain_.setBoxSize(size);
        if (size != previous) reset();
    }

private:
    MacGridDomain3D domain_;
    std::vector<float> density_, vx_, vy_, vz_;
};

// ---- stage interfaces (Advection3D.hpp:6-14, Projection3D.hpp:6-11,
// AdvectionContext.hpp:11-29, VelocityFieldView.hpp) ----
struct VelocityFieldView {
    std::span<const float> x, y, z;
};
inline VelocityFieldView velocityView(const MacGridState3D& state) {
    return {state.velocityX(), state.velocityY(), state.velocityZ()};
}
struct AdvectionContext {
    const MacGridState3D& grid;
    VelocityFieldView flow;
    float dt;
    float cflLimit = 3.f;
};
struct AdvectionWorkspace {
    std::vector<float> forward, backward;
    void resize(const MacGridState3D& grid) {
        forward.resize(grid.velocityX().size());
        backward.resize(grid.velocityX().size());
    }
    void reset() {
        std::fill(forward.begin(), forward.end(), 0.f);
        std::fill(backward.begin(), backward.end(), 0.f);
    }
};

// The advection contract (AdvectionContext.cpp:15-67): finite dt >= 0, a
// finite CFL limit > 0, tracer sizes of the grid's faces, field sizes of the
// grid, and no output aliasing its source or a tracer velocity.
namespace detail {
inline bool overlaps(std::span<const float> a, std::span<const float> b) {
    if (a.empty() || b.empty()) return false;
    const std::less<const float*> less;
    return less(a.data(), b.data() + b.size()) && less(b.data(), a.data() + a.size());
}
inline void validate(const AdvectionContext& c) {
    const std::size_t faces = c.grid.domain().faceCount();
    if (!std::isfinite(c.dt) || c.dt < 0.f || !std::isfinite(c.cflLimit) || c.cflLimit <= 0.f ||
        c.flow.x.size() != faces || c.flow.y.size() != faces || c.flow.z.size() != faces)
        throw std::invalid_argument("Invalid advection time step, CFL limit, or tracer field sizes");
}
inline void reject_tracer_overlap(std::span<const float> out, const VelocityFieldView& flow) {
    if (overlaps(out, flow.x) || overlaps(out, flow.y) || overlaps(out, flow.z))
        throw std::invalid_argument("Advection output must not alias tracer velocities");
}
}  // namespace detail
inline void validateScalarAdvection(const AdvectionContext& context, std::span<float> field) {
    detail::validate(context);
    if (field.size() != context.grid.cellCount())
        throw std::invalid_argument("Scalar field must match the grid cell count");
    detail::reject_tracer_overlap(field, context.flow);
}
inline void validateFaceAdvection(const AdvectionContext& context, FaceAxis axis, std::span<const float> source,
                                  std::span<float> output) {
    detail::validate(context);
    if (axis != FaceAxis::X && axis != FaceAxis::Y && axis != FaceAxis::Z)
        throw std::invalid_argument("Invalid staggered field axis");
    const std::size_t count = context.grid.domain().faceCount();
    if (source.size() != count || output.size() != count || detail::overlaps(source, output))
        throw std::invalid_argument("Velocity fields must match their grid layout and not overlap");
    detail::reject_tracer_overlap(output, context.flow);
}

class Advection3D {
public:
    virtual ~Advection3D() = default;
    virtual void advectScalar(const AdvectionContext& context, std::span<float> field, AdvectionWorkspace& workspace,
                              AdvectionMethod method) const = 0;
    virtual void advectFace(const AdvectionContext& context, FaceAxis axis, std::span<const float> source,
                            std::span<float> output, AdvectionWorkspace& workspace, AdvectionMethod method) const = 0;
};
class Projection3D {
public:
    virtual ~Projection3D() = default;
    virtual ProjectionResult project(MacGridState3D& state, const ProjectionOptions& options) = 0;
    virtual void reset() noexcept = 0;
};

// ---- staged GPU stages (gpu/GpuMacBackend.hpp, GpuAdvection.hpp,
// GpuPressureSolver.hpp): each call uploads its host spans, runs on the
// device and downloads the result ----
class GpuMacBackend {
public:
    GpuMacBackend() = default;
    ~GpuMacBackend() { vgx_backend_destroy(backend_); }
    GpuMacBackend(const GpuMacBackend&) = delete;
    GpuMacBackend& operator=(const GpuMacBackend&) = delete;
    void init(const MacGridState3D& state) {
        vgx_backend* b = nullptr;
        throw_status(vgx_backend_create(state.resolution(), state.boxSize(), &b));
        vgx_backend_destroy(backend_);
        backend_ = b;
        n_ = state.resolution();
        box_ = state.boxSize();
    }
    bool initialized() const noexcept { return backend_ != nullptr; }
    void validateGrid(const MacGridState3D& state) const {
        if (!backend_ || state.resolution() != n_) throw std::invalid_argument("GPU backend grid mismatch");
        if (state.boxSize() != box_) throw std::invalid_argument("GPU backend box mismatch");
    }
    // The backend for this grid, created on first use.
    vgx_backend* for_grid(const MacGridState3D& state) {
        if (!backend_) init(state);
        validateGrid(state);
        return backend_;
    }

private:
    vgx_backend* backend_ = nullptr;
    int n_ = 0;
    float box_ = 0.f;
};

class GpuAdvection final : public Advection3D {
public:
    explicit GpuAdvection(std::shared_ptr<GpuMacBackend> backend) : backend_(std::move(backend)) {
        if (!backend_) throw std::invalid_argument("GPU backend cannot be null");
    }
    void advectScalar(const AdvectionContext& c, std::span<float> field, AdvectionWorkspace&,
                      AdvectionMethod method) const override {
        validateScalarAdvection(c, field);
        throw_status(vgx_backend_advect_scalar(backend_->for_grid(c.grid), static_cast<int>(method), c.dt, c.cflLimit,
                                               c.flow.x.data(), c.flow.y.data(), c.flow.z.data(), field.data()));
    }
    void advectFace(const AdvectionContext& c, FaceAxis axis, std::span<const float> source, std::span<float> output,
                    AdvectionWorkspace&, AdvectionMethod method) const override {
        validateFaceAdvection(c, axis, source, output);
        throw_status(vgx_backend_advect_face(backend_->for_grid(c.grid), static_cast<int>(axis),
                                             static_cast<int>(method), c.dt, c.cflLimit, c.flow.x.data(),
                                             c.flow.y.data(), c.flow.z.data(), source.data(), output.data()));
    }

private:
    std::shared_ptr<GpuMacBackend> backend_;
};

// The reference's fixed-budget solve on the device, with the ProjectionResult
// fields the CPU solver fills (residualAvailable = true: the FP64 residual is
// computed on the device).
class GpuPressureSolver final : public Projection3D {
public:
    explicit GpuPressureSolver(std::shared_ptr<GpuMacBackend> backend) : backend_(std::move(backend)) {
        if (!backend_) throw std::invalid_argument("GPU backend cannot be null");
    }
    ProjectionResult project(MacGridState3D& state, const ProjectionOptions& o) override {
        vgx_projection_options c;
        vgx_projection_options_default(&c);
        c.dt = o.dt;
        c.relaxation = o.relaxation;
        c.max_iterations = o.linearSolve.maxIterations;
        c.residual_check_interval = o.linearSolve.residualCheckInterval;
        c.absolute_tolerance = o.linearSolve.absoluteTolerance;
        c.relative_tolerance = o.linearSolve.relativeTolerance;
        c.fixed_iterations = o.linearSolve.fixedIterations;
        vgx_projection_result r;
        throw_status(vgx_backend_project(backend_->for_grid(state), &c, state.velocityX().data(),
                                         state.velocityY().data(), state.velocityZ().data(), nullptr, &r));
        return projection_result(r);
    }
    // The device pressure starts from zero in every projection.
    void reset() noexcept override {}

private:
    std::shared_ptr<GpuMacBackend> backend_;
};

// Device-resident density for a renderer (RenderData.hpp:5-14 with a device
// pointer in place of the GL buffer; GpuMacSimulation3D.cpp:107-109): valid
// until the next advance, reset or box change.
struct DeviceVolume {
    const float* density = nullptr;
    int width = 0, height = 0, depth = 0;
    float boxSize = 1.f;
};

// The solver. Construction mirrors MacGridFluidSolver3D: (MacAlgorithm) gives
// 32^3, box 4.5 and the preset; (int res) gives the defaults with
// projectIterations = 2 res; (int res, float box, parameters) is the
// GpuMacSimulation3D form.
class MacGridFluidSolver3D {
public:
    explicit MacGridFluidSolver3D(MacAlgorithm algorithm)
        : MacGridFluidSolver3D(32, 4.5f, MacGridParameters3D(algorithm)) {}
    explicit MacGridFluidSolver3D(int res) : domain_(res, 4.5f) {
        throw_status(vgx_sim_create(res, 4.5f, nullptr, &sim_));
        vgx_params p;
        vgx_sim_get_params(sim_, &p);
        parameters_.load(p);
    }
    MacGridFluidSolver3D(int res, float boxSize, const MacGridParameters3D& parameters)
        : domain_(res, boxSize), parameters_(parameters) {
        const vgx_params p = parameters.to_c();
        throw_status(vgx_sim_create(res, boxSize, &p, &sim_));
    }
    // Adopts prefilled fields without resetting, as MacGridFluidSolver3D(MacGridState3D).
    explicit MacGridFluidSolver3D(const MacGridState3D& state) : domain_(state.domain()) {
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
    const MacGridDomain3D& domain() const noexcept { return domain_; }

    // Ownership transfers to the solver (MacGridFluidSolver3D.cpp:305-321):
    // null is rejected, the fields are kept, and an installed pressure stage
    // clears lastProjection. The stage then runs inside every step on host
    // spans; the rest of the step stays on the device.
    void setPressureSolver(std::unique_ptr<Projection3D> solver) {
        if (!solver) throw std::invalid_argument("Pressure backend cannot be null");
        throw_status(vgx_sim_set_projection_hook(sim_, &MacGridFluidSolver3D::project_hook, this));
        pressure_ = std::move(solver);
    }
    void setAdvection(std::unique_ptr<Advection3D> method) {
        if (!method) throw std::invalid_argument("Advection backend cannot be null");
        throw_status(vgx_sim_set_advection_hook(sim_, &MacGridFluidSolver3D::advect_hook, this));
        advection_ = std::move(method);
    }
    void useDefaultPressureSolver() {
        throw_status(vgx_sim_set_projection_hook(sim_, nullptr, nullptr));
        pressure_.reset();
    }
    void useDefaultAdvection() {
        throw_status(vgx_sim_set_advection_hook(sim_, nullptr, nullptr));
        advection_.reset();
    }

    void advance(float dt) {
        push();
        snapshot_valid_ = false;
        run(vgx_sim_advance(sim_, dt));
    }
    void step(float dt, float buoyancy, float sourceStrength, int projectIterations) {
        with_overrides(true, buoyancy, sourceStrength, projectIterations, dt);
    }
    void stepNoReflect(float dt, float buoyancy, float sourceStrength, int projectIterations) {
        with_overrides(false, buoyancy, sourceStrength, projectIterations, dt);
    }
    // resetState (MacGridFluidSolver3D.cpp:271-290): zero fields, reseed, and
    // reset the installed pressure stage.
    void resetState() {
        push();
        throw_status(vgx_sim_reset(sim_));
        if (pressure_) pressure_->reset();
        advection_workspace_.reset();
        snapshot_valid_ = false;
    }
    void setBoxSize(float size) {
        if (size == domain_.boxSize()) return;
        MacGridDomain3D next = domain_;
        next.setBoxSize(size);  // the reference's validation and message
        push();
        throw_status(vgx_sim_set_box(sim_, size));
        domain_ = next;
        if (pressure_) pressure_->reset();
        snapshot_valid_ = false;
    }
    int resolution() const { return domain_.resolution(); }
    float boxSize() const noexcept { return domain_.boxSize(); }
    const ProjectionResult& lastProjection() const {
        vgx_projection_result r;
        throw_status(vgx_sim_last_projection(sim_, 0, &r));
        last_ = projection_result(r);
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
            snapshot_ = std::make_unique<MacGridState3D>(domain_);
            throw_status(vgx_sim_download(sim_, snapshot_->density().data(), snapshot_->velocityX().data(),
                                          snapshot_->velocityY().data(), snapshot_->velocityZ().data()));
            snapshot_valid_ = true;
        }
        return *snapshot_;
    }
    // The density where it lives, for a renderer: no download.
    DeviceVolume deviceVolume() const {
        DeviceVolume v;
        throw_status(vgx_sim_synchronize(sim_));
        throw_status(vgx_sim_device_fields(sim_, &v.density, nullptr, nullptr, nullptr, nullptr));
        v.width = v.height = v.depth = domain_.resolution();
        v.boxSize = domain_.boxSize();
        return v;
    }
    // Extension: sigma * N(0, 1) added to every face velocity (see
    // vgx_sim_add_velocity_noise).
    void addVelocityNoise(float sigma, std::uint64_t seed) {
        throw_status(vgx_sim_add_velocity_noise(sim_, sigma, seed));
        snapshot_valid_ = false;
    }
    void upload(const MacGridState3D& state) {
        if (state.resolution() != resolution() || state.boxSize() != boxSize())
            throw std::invalid_argument("GPU snapshot box mismatch");
        throw_status(vgx_sim_upload(sim_, state.density().data(), state.velocityX().data(),
                                    state.velocityY().data(), state.velocityZ().data()));
        snapshot_valid_ = false;
    }
    void download(MacGridState3D& state) const {
        if (state.resolution() != resolution() || state.boxSize() != boxSize())
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
    // A step's status: an exception an installed stage threw is rethrown as
    // it was thrown (it cannot cross the C ABI, so the hook kept it).
    void run(int status) {
        if (stage_error_) {
            std::exception_ptr error = stage_error_;
            stage_error_ = nullptr;
            std::rethrow_exception(error);
        }
        throw_status(status);
    }
    static int stage_status(std::exception_ptr& slot) {
        slot = std::current_exception();
        try {
            throw;
        } catch (const std::invalid_argument&) {
            return VGX_EINVAL;
        } catch (const std::length_error&) {
            return VGX_ERANGE;
        } catch (...) {
            return VGX_EDEVICE;
        }
    }
    // Geometry-only grid of the stage contexts (its fields are not the step's).
    const MacGridState3D& stage_grid() {
        if (!stage_state_ || stage_state_->resolution() != resolution() || stage_state_->boxSize() != boxSize())
            stage_state_ = std::make_unique<MacGridState3D>(domain_);
        return *stage_state_;
    }
    static int advect_hook(void* user, int target, int method, float dt, float cfl, const float* u, const float* v,
                           const float* w, const float* source, float* out) {
        auto* self = static_cast<MacGridFluidSolver3D*>(user);
        try {
            const std::size_t faces = self->domain_.faceCount(), cells = self->domain_.cellCount();
            const AdvectionContext ctx{self->stage_grid(), {{u, faces}, {v, faces}, {w, faces}}, dt, cfl};
            self->advection_workspace_.resize(ctx.grid);
            const auto m = static_cast<AdvectionMethod>(method);
            if (target == VGX_TARGET_DENSITY)
                self->advection_->advectScalar(ctx, {out, cells}, self->advection_workspace_, m);
            else
                self->advection_->advectFace(ctx, static_cast<FaceAxis>(target), {source, faces}, {out, faces},
                                             self->advection_workspace_, m);
            return VGX_OK;
        } catch (...) {
            return stage_status(self->stage_error_);
        }
    }
    static int project_hook(void* user, const vgx_projection_options* o, float* u, float* v, float* w,
                            const float* rho, vgx_projection_result* result) {
        auto* self = static_cast<MacGridFluidSolver3D*>(user);
        try {
            self->stage_grid();
            MacGridState3D& s = *self->stage_state_;
            const std::size_t faces = self->domain_.faceCount();
            std::copy(rho, rho + s.cellCount(), s.density().begin());
            std::copy(u, u + faces, s.velocityX().begin());
            std::copy(v, v + faces, s.velocityY().begin());
            std::copy(w, w + faces, s.velocityZ().begin());
            ProjectionOptions options;
            options.dt = o->dt;
            options.relaxation = o->relaxation;
            options.linearSolve.maxIterations = o->max_iterations;
            options.linearSolve.residualCheckInterval = o->residual_check_interval;
            options.linearSolve.absoluteTolerance = o->absolute_tolerance;
            options.linearSolve.relativeTolerance = o->relative_tolerance;
            options.linearSolve.fixedIterations = o->fixed_iterations != 0;
            *result = projection_record(self->pressure_->project(s, options));
            std::copy(s.velocityX().begin(), s.velocityX().end(), u);
            std::copy(s.velocityY().begin(), s.velocityY().end(), v);
            std::copy(s.velocityZ().begin(), s.velocityZ().end(), w);
            return VGX_OK;
        } catch (...) {
            return stage_status(self->stage_error_);
        }
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

    MacGridDomain3D domain_;
    vgx_sim* sim_ = nullptr;
    MacGridParameters3D parameters_;
    std::unique_ptr<Projection3D> pressure_;
    std::unique_ptr<Advection3D> advection_;
    AdvectionWorkspace advection_workspace_;
    std::unique_ptr<MacGridState3D> stage_state_;
    std::exception_ptr stage_error_;
    mutable ProjectionResult last_;
    mutable std::unique_ptr<MacGridState3D> snapshot_;
    mutable bool snapshot_valid_ = false;
};

}  // namespace vapor::hip
