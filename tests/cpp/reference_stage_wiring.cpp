// reference_stage_wiring.cpp - the reference-side binding
// (include/vapor_hip_stages.hpp) installed into vapor's own CPU
// MacGridFluidSolver3D, compiled against vapor's headers and sources where
// the reference tree exists (oracle/Makefile target `wiring`, output in
// oracle/_ref/), run on the GPU.
//
// What vapor's solver must hand an installed stage (BackendTests.cpp:62-143
// states the same contract for its own stages): per step one projection
// (two under reflection, each with dt/2 and K/2, fixed budget), three face
// advections per pass and one scalar advection; the stage's results must
// agree with vapor's CPU stages within 2e-5; the solver owns, resets and
// destroys what is installed.
//
// Prints "ok <check>" / "FAIL <check>: detail"; exit status = failures.
#include "solvers/mac3d/MacGridFluidSolver3D.hpp"
#include "vapor_hip_stages.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const std::string& name, const std::string& detail = "") {
    std::printf("%s %s%s%s\n", ok ? "ok" : "FAIL", name.c_str(), ok ? "" : ": ", ok ? "" : detail.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

double worst(std::span<const float> a, std::span<const float> b) {
    double w = a.size() == b.size() ? 0.0 : INFINITY;
    for (std::size_t i = 0; i < std::min(a.size(), b.size()); ++i)
        w = std::max(w, std::fabs(double(a[i]) - double(b[i])));
    return w;
}

struct Tally {
    std::vector<float> dts;
    std::vector<int> budgets;
    std::vector<bool> fixed;
    int faces = 0, scalars = 0, resets = 0, destroyed = 0;
    AdvectionMethod faceMethod{}, scalarMethod{};
};

// Counting wrappers around the binding's stages.
class CountedProjection final : public Projection3D {
public:
    CountedProjection(std::shared_ptr<Tally> t, std::shared_ptr<vapor_hip::HipBackend> b)
        : tally_(std::move(t)), inner_(std::move(b)) {}
    ProjectionResult project(MacGridState3D& s, const ProjectionOptions& o) override {
        tally_->dts.push_back(o.dt);
        tally_->budgets.push_back(o.linearSolve.maxIterations);
        tally_->fixed.push_back(o.linearSolve.fixedIterations);
        return inner_.project(s, o);
    }
    void reset() noexcept override { ++tally_->resets; }

private:
    std::shared_ptr<Tally> tally_;
    vapor_hip::HipProjection3D inner_;
};

class CountedAdvection final : public Advection3D {
public:
    CountedAdvection(std::shared_ptr<Tally> t, std::shared_ptr<vapor_hip::HipBackend> b)
        : tally_(std::move(t)), inner_(std::move(b)) {}
    ~CountedAdvection() override { ++tally_->destroyed; }
    void advectScalar(const AdvectionContext& c, std::span<float> f, AdvectionWorkspace& w,
                      AdvectionMethod m) const override {
        ++tally_->scalars;
        tally_->scalarMethod = m;
        inner_.advectScalar(c, f, w, m);
    }
    void advectFace(const AdvectionContext& c, FaceAxis a, std::span<const float> s, std::span<float> o,
                    AdvectionWorkspace& w, AdvectionMethod m) const override {
        ++tally_->faces;
        tally_->faceMethod = m;
        inner_.advectFace(c, a, s, o, w, m);
    }

private:
    std::shared_ptr<Tally> tally_;
    vapor_hip::HipAdvection3D inner_;
};

void configure(MacGridFluidSolver3D& s, bool reflect, bool velocityMc, bool smokeMc) {
    auto& p = s.parameters();
    p.stirStrength = 0.f;
    p.emitterRadius = .18f;
    p.emitterCenterZ = .3f;
    p.reflection = reflect;
    p.projectIterations = 16;
    p.macCormackVel = velocityMc;
    p.macCormackSmoke = smokeMc;
}

double field_error(const MacGridFluidSolver3D& a, const MacGridFluidSolver3D& b) {
    const auto& x = a.state();
    const auto& y = b.state();
    return std::max({worst(x.density(), y.density()), worst(x.velocityX(), y.velocityX()),
                     worst(x.velocityY(), y.velocityY()), worst(x.velocityZ(), y.velocityZ())});
}

void wiring(bool reflect) {
    const std::string tag = reflect ? "reflection: " : "simple: ";
    const float dt = 1.f / 60.f;
    const int n = 8;
    MacGridFluidSolver3D gpu(n), cpu(n);
    configure(gpu, reflect, false, false);
    configure(cpu, reflect, false, false);
    auto tally = std::make_shared<Tally>();
    auto backend = std::make_shared<vapor_hip::HipBackend>();
    gpu.setPressureSolver(std::make_unique<CountedProjection>(tally, backend));
    gpu.setAdvection(std::make_unique<CountedAdvection>(tally, backend));
    for (int f = 0; f < 3; ++f) {
        gpu.advance(dt);
        cpu.advance(dt);
    }
    const int per = reflect ? 2 : 1;
    check(int(tally->dts.size()) == 3 * per, tag + "one projection per pass");
    bool dts = true, budgets = true, fixed = true;
    for (float v : tally->dts) dts = dts && v == dt / per;
    for (int v : tally->budgets) budgets = budgets && v == 16 / per;
    for (bool v : tally->fixed) fixed = fixed && v;
    check(dts, tag + "projection dt = dt / passes");
    check(budgets, tag + "projection budget = K / passes");
    check(fixed, tag + "projection budget is fixed");
    check(tally->faces == 9 * per && tally->scalars == 3, tag + "advection call counts",
          std::to_string(tally->faces) + " faces, " + std::to_string(tally->scalars) + " scalars");
    const double e = field_error(gpu, cpu);
    check(e <= 2e-5, tag + "GPU stages agree with vapor's CPU stages within 2e-5", std::to_string(e));
    std::printf("   %smax |GPU stages - CPU stages| = %.3g\n", tag.c_str(), e);
    check(gpu.lastProjection().residualAvailable, tag + "projection reports its residual");
    const double ke = std::fabs(double(gpu.kineticEnergy()) - double(cpu.kineticEnergy()));
    check(ke <= 2e-5 * std::max(1.0, double(cpu.kineticEnergy())), tag + "kinetic energy within 2e-5");

    // MacCormack for velocity only: the stage sees the per-field methods
    gpu.parameters().macCormackVel = cpu.parameters().macCormackVel = true;
    gpu.advance(dt);
    cpu.advance(dt);
    check(tally->faceMethod == AdvectionMethod::MacCormack && tally->scalarMethod == AdvectionMethod::SemiLagrangian,
          tag + "per-field methods reach the stage");
    check(field_error(gpu, cpu) <= 2e-5, tag + "MacCormack velocity within 2e-5", std::to_string(field_error(gpu, cpu)));

    gpu.resetState();
    check(tally->resets == 1, tag + "reset reaches the installed projection");
    gpu.useDefaultPressureSolver();
    gpu.useDefaultAdvection();
    check(tally->destroyed == 1, tag + "useDefaultAdvection destroys the stage");
    bool rejected = false;
    try {
        gpu.setPressureSolver(nullptr);
    } catch (const std::invalid_argument&) {
        rejected = true;
    }
    check(rejected, tag + "null stage rejected");
}

// vapor's own contract reaches the binding's callers unchanged
// (AdvectionTests.cpp:159-177): aliasing and dt errors are invalid_argument.
void contract() {
    MacGridState3D grid(3);
    auto backend = std::make_shared<vapor_hip::HipBackend>();
    vapor_hip::HipAdvection3D method(backend);
    AdvectionWorkspace workspace;
    AdvectionContext context{grid, velocityView(grid), .1f};
    std::vector<float> source(grid.velocityX().size(), 1.f), output(source.size());
    auto rejects = [](const std::function<void()>& f) {
        try {
            f();
        } catch (const std::invalid_argument&) {
            return true;
        }
        return false;
    };
    check(rejects([&] { method.advectFace(context, FaceAxis::X, source, source, workspace, AdvectionMethod::SemiLagrangian); }),
          "contract: output aliasing the source");
    check(rejects([&] { method.advectFace(context, FaceAxis::X, source, grid.velocityX(), workspace, AdvectionMethod::SemiLagrangian); }),
          "contract: output aliasing the tracer");
    context.dt = -1.f;
    check(rejects([&] { method.advectScalar(context, grid.density(), workspace, AdvectionMethod::SemiLagrangian); }),
          "contract: negative dt");
    context.dt = .1f;
    method.advectFace(context, FaceAxis::Y, source, output, workspace, AdvectionMethod::MacCormack);
    bool still = true;
    for (float v : output) still = still && v == 1.f;
    check(still, "contract: stationary flow keeps the field");
}

}  // namespace

int main() {
    try {
        wiring(false);
        wiring(true);
        contract();
    } catch (const std::exception& e) {
        check(false, "unexpected exception", e.what());
    }
    std::printf("%d failed\n", failures);
    return std::min(failures, 100);
}
