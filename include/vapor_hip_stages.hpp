// vapor_hip_stages.hpp - the reference-side binding: vapor's own stage
// interfaces implemented on the MI355X library, for installation into vapor's
// CPU MacGridFluidSolver3D with setPressureSolver / setAdvection
// (src/solvers/mac3d/MacGridFluidSolver3D.hpp:54-57, .cpp:305-321).
//
// This header compiles inside vapor's source tree: it includes vapor's
// headers (solvers/mac3d/...) and the C ABI (vapor_mac3d.h). A vapor
// maintainer adds it as, e.g., src/solvers/mac3d/HipStages3D.hpp and links
// libvapor_mac3d.so.
//
//   solver.setPressureSolver(std::make_unique<vapor_hip::HipProjection3D>(backend));
//   solver.setAdvection(std::make_unique<vapor_hip::HipAdvection3D>(backend));
//
// Each call uploads the borrowed spans, runs the stage on the device and
// downloads the result (the staged path of vapor's gpu/GpuAdvection and
// gpu/GpuPressureSolver): suited to parity work and plug-in wiring; the
// resident whole-step path (vapor_mac3d.hpp) is the fast one.
#pragma once

#include "solvers/mac3d/Advection3D.hpp"
#include "solvers/mac3d/MacGridState3D.hpp"
#include "solvers/mac3d/Projection3D.hpp"
#include "vapor_mac3d.h"

#include <memory>
#include <stdexcept>
#include <string>

namespace vapor_hip {

// Status code -> vapor's exception types (GpuMacBackend.cpp:17,22,40 throw
// runtime_error for device failures; invalid input is invalid_argument).
inline void hip_check(int status) {
    if (status == VGX_OK) return;
    const std::string message = vgx_last_error();
    if (status == VGX_EINVAL) throw std::invalid_argument(message);
    if (status == VGX_ERANGE) throw std::length_error(message);
    throw std::runtime_error(message);
}

// One device workspace for a grid, shared by the stages of one solver (as
// vapor shares its GpuMacBackend between GpuAdvection and GpuPressureSolver).
// Created for the first grid it sees; a different grid is rejected.
class HipBackend {
public:
    HipBackend() = default;
    ~HipBackend() { vgx_backend_destroy(backend_); }
    HipBackend(const HipBackend&) = delete;
    HipBackend& operator=(const HipBackend&) = delete;
    vgx_backend* for_grid(const MacGridState3D& grid) {
        if (!backend_) {
            hip_check(vgx_backend_create(grid.resolution(), grid.boxSize(), &backend_));
            n_ = grid.resolution();
            box_ = grid.boxSize();
        }
        if (grid.resolution() != n_ || grid.boxSize() != box_)
            throw std::invalid_argument("GPU backend grid mismatch");
        return backend_;
    }

private:
    vgx_backend* backend_ = nullptr;
    int n_ = 0;
    float box_ = 0.f;
};

// Projection3D (Projection3D.hpp:6-11) on the GPU: vapor's projection
// (CpuPressureSolver3D.cpp:25-60) with the options vapor's solver passes.
class HipProjection3D final : public Projection3D {
public:
    explicit HipProjection3D(std::shared_ptr<HipBackend> backend) : backend_(std::move(backend)) {
        if (!backend_) throw std::invalid_argument("GPU backend cannot be null");
    }
    ProjectionResult project(MacGridState3D& s, const ProjectionOptions& o) override {
        vgx_projection_options c;
        vgx_projection_options_default(&c);  // diagonal floor 6, as vapor's MAC solver builds its SOR
        c.dt = o.dt;
        c.relaxation = o.relaxation;
        c.max_iterations = o.linearSolve.maxIterations;
        c.residual_check_interval = o.linearSolve.residualCheckInterval;
        c.absolute_tolerance = o.linearSolve.absoluteTolerance;
        c.relative_tolerance = o.linearSolve.relativeTolerance;
        c.fixed_iterations = o.linearSolve.fixedIterations;
        vgx_projection_result r;
        hip_check(vgx_backend_project(backend_->for_grid(s), &c, s.velocityX().data(), s.velocityY().data(),
                                      s.velocityZ().data(), nullptr, &r));
        ProjectionResult out;
        out.linearSolve.iterations = r.iterations;
        out.linearSolve.initialResidual = r.initial_residual;
        out.linearSolve.finalResidual = r.final_residual;
        out.linearSolve.converged = r.converged != 0;
        out.residualAvailable = r.residual_available != 0;
        out.divergenceBefore = r.divergence_before;
        out.divergenceAfter = r.divergence_after;
        out.removedRhsMean = r.removed_rhs_mean;
        return out;
    }
    // The device pressure starts from zero in every projection.
    void reset() noexcept override {}

private:
    std::shared_ptr<HipBackend> backend_;
};

// Advection3D (Advection3D.hpp:6-14) on the GPU. vapor's own contract checks
// run first (AdvectionContext.cpp); the workspace is unused, the device keeps
// its forward/backward traces.
class HipAdvection3D final : public Advection3D {
public:
    explicit HipAdvection3D(std::shared_ptr<HipBackend> backend) : backend_(std::move(backend)) {
        if (!backend_) throw std::invalid_argument("GPU backend cannot be null");
    }
    void advectScalar(const AdvectionContext& ctx, std::span<float> field, AdvectionWorkspace&,
                      AdvectionMethod method) const override {
        validateScalarAdvection(ctx, field);
        hip_check(vgx_backend_advect_scalar(backend_->for_grid(ctx.grid), static_cast<int>(method), ctx.dt,
                                            ctx.cflLimit, ctx.flow.x.data(), ctx.flow.y.data(), ctx.flow.z.data(),
                                            field.data()));
    }
    void advectFace(const AdvectionContext& ctx, FaceAxis axis, std::span<const float> source,
                    std::span<float> output, AdvectionWorkspace&, AdvectionMethod method) const override {
        validateFaceAdvection(ctx, axis, source, output);
        hip_check(vgx_backend_advect_face(backend_->for_grid(ctx.grid), static_cast<int>(axis),
                                          static_cast<int>(method), ctx.dt, ctx.cflLimit, ctx.flow.x.data(),
                                          ctx.flow.y.data(), ctx.flow.z.data(), source.data(), output.data()));
    }

private:
    std::shared_ptr<HipBackend> backend_;
};

}  // namespace vapor_hip
