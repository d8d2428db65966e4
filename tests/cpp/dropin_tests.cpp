// dropin_tests.cpp - checks of the C++ drop-in (include/vapor_mac3d.hpp) on
// the GPU, driven by data files and run as separate groups:
//
//   dropin_tests baselines <baselines.json>  every published whole-step row
//                                             the file holds (setup + rows)
//   dropin_tests replay                       seeded streams, reset replay
//   dropin_tests plume                        long-run stability bounds
//   dropin_tests stages                       installed stages: wiring,
//                                             ownership and agreement with
//                                             the device-resident stages
//
// Each check prints one line "ok <name>" or "FAIL <name>: <detail>"; the
// exit status is the number of failed checks (capped at 100).
#include "vapor_mac3d.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace hip = vapor::hip;

namespace {

// ------------------------------------------------------------ reporting

struct Report {
    int failures = 0;
    void check(bool ok, const std::string& name, const std::string& detail = "") {
        if (ok) {
            std::printf("ok %s\n", name.c_str());
        } else {
            ++failures;
            std::printf("FAIL %s: %s\n", name.c_str(), detail.c_str());
        }
        std::fflush(stdout);
    }
    // |actual - expected| <= rel * max(1, |expected|), the published bound.
    void close(double actual, double expected, double rel, const std::string& name) {
        const double bound = rel * std::max(1.0, std::fabs(expected));
        std::ostringstream s;
        s.precision(12);
        s << "actual " << actual << " expected " << expected << " bound " << bound;
        check(std::isfinite(actual) && std::fabs(actual - expected) <= bound, name, s.str());
    }
    void guard(const std::string& name, const std::function<void()>& body) {
        try {
            body();
        } catch (const std::exception& e) {
            check(false, name, std::string("exception: ") + e.what());
        }
    }
};

bool same_bits(std::span<const float> a, std::span<const float> b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

double max_diff(std::span<const float> a, std::span<const float> b) {
    double worst = a.size() == b.size() ? 0.0 : INFINITY;
    for (std::size_t i = 0; i < std::min(a.size(), b.size()); ++i)
        worst = std::max(worst, std::fabs(double(a[i]) - double(b[i])));
    return worst;
}

// ------------------------------------------------------ the data file

// The numbers of one flat JSON object: "key": number pairs (strings and
// nested objects are skipped). Enough for the baseline file's setup block
// and rows, which hold numbers only where the test reads them.
std::map<std::string, double> numbers_of(const std::string& object) {
    std::map<std::string, double> out;
    std::size_t at = 0;
    while ((at = object.find('"', at)) != std::string::npos) {
        const std::size_t end = object.find('"', at + 1);
        if (end == std::string::npos) break;
        const std::string key = object.substr(at + 1, end - at - 1);
        std::size_t colon = object.find_first_not_of(" \t\r\n", end + 1);
        at = end + 1;
        if (colon == std::string::npos || object[colon] != ':') continue;
        const std::size_t value = object.find_first_not_of(" \t\r\n", colon + 1);
        if (value == std::string::npos) break;
        const char c = object[value];
        if (c == '-' || c == '.' || (c >= '0' && c <= '9')) {
            out[key] = std::strtod(object.c_str() + value, nullptr);
            at = value;
        }
    }
    return out;
}

// The text of the object or array that opens at `open` (matching brackets).
std::string bracketed(const std::string& text, std::size_t open) {
    const char o = text[open], c = o == '{' ? '}' : ']';
    int depth = 0;
    for (std::size_t i = open; i < text.size(); ++i) {
        if (text[i] == o) ++depth;
        if (text[i] == c && --depth == 0) return text.substr(open, i - open + 1);
    }
    throw std::runtime_error("unbalanced JSON");
}

struct BaselineFile {
    std::map<std::string, double> setup;
    std::vector<std::map<std::string, double>> rows;
};

BaselineFile read_baselines(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream buffer;
    buffer << in.rdbuf();
    const std::string text = buffer.str();
    BaselineFile file;
    file.setup = numbers_of(bracketed(text, text.find('{', text.find("\"setup\""))));
    const std::string rows = bracketed(text, text.find('[', text.find("\"rows\"")));
    for (std::size_t at = rows.find('{'); at != std::string::npos; at = rows.find('{', at + 1)) {
        const std::string row = bracketed(rows, at);
        file.rows.push_back(numbers_of(row));
        at += row.size() - 1;
    }
    return file;
}

// ------------------------------------------------------------ groups

// The published whole-step baselines: for each row, an n^3 solver with the
// file's setup (emitter radius and height, stir, sweep budget, box) and the
// row's variant, stepped `frames` frames; five diagnostics against the row.
void baselines(Report& report, const std::string& path) {
    const BaselineFile file = read_baselines(path);
    report.check(file.rows.size() == 16, "baseline file holds 16 rows",
                 std::to_string(file.rows.size()) + " rows");
    const auto& s = file.setup;
    const int frames = int(s.at("frames"));
    const float dt = float(s.at("dt"));
    for (const auto& row : file.rows) {
        const int n = int(row.at("n")), flags = int(row.at("flags"));
        const bool reflect = row.at("reflect") != 0.0;
        const std::string tag = "n" + std::to_string(n) + (reflect ? " reflection" : " simple") + " flags " +
                                std::to_string(flags);
        report.guard(tag, [&] {
            hip::MacGridFluidSolver3D solver(n);
            if (solver.boxSize() != float(s.at("box"))) solver.setBoxSize(float(s.at("box")));
            auto& p = solver.parameters();
            p.stirStrength = float(s.at("stir_strength"));
            p.emitterRadius = float(s.at("emitter_radius"));
            p.emitterCenterZ = float(s.at("emitter_center_z"));
            p.projectIterations = int(s.at("project_iterations"));
            p.reflection = reflect;
            p.macCormackSmoke = (flags & 1) != 0;
            p.macCormackVel = (flags & 2) != 0;
            for (int f = 0; f < frames; ++f) solver.advance(dt);
            const hip::FluidDiagnostics d = solver.diagnostics();
            const auto rho = solver.state().density();
            double total = 0.0, weighted = 0.0;
            for (std::size_t i = 0; i < rho.size(); ++i) {
                total += rho[i];
                weighted += rho[i] * double(1 + i % 31);
            }
            const double rel = 2e-5;
            report.close(d.kineticEnergy, row.at("energy"), rel, tag + " kinetic energy");
            report.close(d.maxVelocity, row.at("max_velocity"), rel, tag + " max velocity");
            report.close(d.maxDivergence, row.at("divergence"), rel, tag + " max divergence");
            report.close(total, row.at("density_sum"), rel, tag + " density sum");
            report.close(weighted, row.at("weighted_density"), rel, tag + " weighted density");
            solver.resetState();
            const hip::FluidDiagnostics z = solver.diagnostics();
            report.check(z.kineticEnergy == 0.f && z.maxVelocity == 0.f, tag + " reset clears the flow");
        });
    }
}

hip::MacGridParameters3D replay_parameters(std::uint32_t seed) {
    hip::MacGridParameters3D p;
    p.emitterRadius = .18f;
    p.emitterCenterZ = .3f;
    p.projectIterations = 16;
    p.randomSeed = seed;
    return p;
}

// Seeded stirring: every solver owns its stream. Two solvers with one seed
// agree bit for bit whatever else runs between their steps; another seed
// differs; a reset replays the trajectory; a solver created later with the
// first seed reproduces it too.
void replay(Report& report) {
    report.guard("replay", [&] {
        const float dt = 1.f / 60.f;
        hip::MacGridFluidSolver3D a(8, 4.5f, replay_parameters(12345)), b(8, 4.5f, replay_parameters(12345)),
            c(8, 4.5f, replay_parameters(54321));
        for (int f = 0; f < 6; ++f) {
            a.advance(dt);
            for (int k = 0; k < 2; ++k) c.advance(dt);  // interleaved, twice as often
            b.advance(dt);
        }
        const auto fields = [](const hip::MacGridFluidSolver3D& s) {
            const auto& st = s.state();
            return std::vector<std::vector<float>>{
                {st.density().begin(), st.density().end()}, {st.velocityX().begin(), st.velocityX().end()},
                {st.velocityY().begin(), st.velocityY().end()}, {st.velocityZ().begin(), st.velocityZ().end()}};
        };
        const auto fa = fields(a), fb = fields(b), fc = fields(c);
        bool ab = true;
        for (int k = 0; k < 4; ++k) ab = ab && same_bits(fa[k], fb[k]);
        report.check(ab, "same seed, interleaved with another solver: identical fields");
        report.check(!same_bits(fa[1], fc[1]), "another seed: different stirring");
        a.resetState();
        for (int f = 0; f < 6; ++f) a.advance(dt);
        const auto fr = fields(a);
        bool replayed = true;
        for (int k = 0; k < 4; ++k) replayed = replayed && same_bits(fr[k], fb[k]);
        report.check(replayed, "reset replays the seeded trajectory");
        hip::MacGridFluidSolver3D late(8, 4.5f, replay_parameters(12345));
        for (int f = 0; f < 6; ++f) late.advance(dt);
        report.check(same_bits(fields(late)[2], fb[2]), "a later solver with the seed reproduces it");
    });
}

// Long-run stability of a 24^3 plume (the reference's published bounds):
// 180 frames, 300 sweeps at omega 1.8, radius .08; max |div u| < .005 and
// density within [0, 1.00001] after every frame, with and without reflection.
void plume(Report& report) {
    for (bool reflection : {false, true}) {
        const std::string tag = std::string("plume 24^3 ") + (reflection ? "reflection" : "simple");
        report.guard(tag, [&] {
            hip::MacGridFluidSolver3D solver(24);
            auto& p = solver.parameters();
            p.reflection = reflection;
            p.projectIterations = 300;
            p.emitterRadius = .08f;
            p.sorOmega = 1.8f;
            float worst_div = 0.f, lo = 0.f, hi = 0.f;
            bool finite = true;
            for (int f = 0; f < 180; ++f) {
                solver.advance(1.f / 60.f);
                const auto d = solver.diagnostics();
                finite = finite && std::isfinite(d.kineticEnergy) && std::isfinite(d.maxDivergence);
                worst_div = std::max(worst_div, d.maxDivergence);
                for (float r : solver.state().density()) {
                    finite = finite && std::isfinite(r);
                    lo = std::min(lo, r);
                    hi = std::max(hi, r);
                }
            }
            report.check(finite, tag + " stays finite");
            report.check(worst_div < .005f, tag + " max divergence < .005", std::to_string(worst_div));
            report.check(lo >= 0.f && hi <= 1.00001f, tag + " density within [0, 1.00001]",
                         std::to_string(lo) + " .. " + std::to_string(hi));
            std::printf("   %s: energy %.6g, max divergence %.3g\n", tag.c_str(), solver.kineticEnergy(),
                        solver.computeDivergenceNorm());
        });
    }
}

// ---- installed stages ----

struct StageLog {
    std::vector<float> pressure_dt;
    std::vector<int> pressure_budget;
    std::vector<bool> pressure_fixed;
    int face_calls = 0, scalar_calls = 0, resets = 0, destroyed = 0;
    double face_time = 0.0, scalar_time = 0.0;
    hip::AdvectionMethod face_method{}, scalar_method{};
};

// A pressure stage that records what the solver hands it and forwards to
// the staged GPU solve.
class LoggedPressure final : public hip::Projection3D {
public:
    LoggedPressure(std::shared_ptr<StageLog> log, std::shared_ptr<hip::GpuMacBackend> backend)
        : log_(std::move(log)), inner_(std::move(backend)) {}
    hip::ProjectionResult project(hip::MacGridState3D& state, const hip::ProjectionOptions& o) override {
        log_->pressure_dt.push_back(o.dt);
        log_->pressure_budget.push_back(o.linearSolve.maxIterations);
        log_->pressure_fixed.push_back(o.linearSolve.fixedIterations);
        return inner_.project(state, o);
    }
    void reset() noexcept override { ++log_->resets; }

private:
    std::shared_ptr<StageLog> log_;
    hip::GpuPressureSolver inner_;
};

class LoggedAdvection final : public hip::Advection3D {
public:
    LoggedAdvection(std::shared_ptr<StageLog> log, std::shared_ptr<hip::GpuMacBackend> backend)
        : log_(std::move(log)), inner_(std::move(backend)) {}
    ~LoggedAdvection() override { ++log_->destroyed; }
    void advectScalar(const hip::AdvectionContext& c, std::span<float> field, hip::AdvectionWorkspace& w,
                      hip::AdvectionMethod m) const override {
        ++log_->scalar_calls;
        log_->scalar_time += c.dt;
        log_->scalar_method = m;
        inner_.advectScalar(c, field, w, m);
    }
    void advectFace(const hip::AdvectionContext& c, hip::FaceAxis axis, std::span<const float> source,
                    std::span<float> output, hip::AdvectionWorkspace& w, hip::AdvectionMethod m) const override {
        ++log_->face_calls;
        log_->face_time += c.dt;
        log_->face_method = m;
        inner_.advectFace(c, axis, source, output, w, m);
    }

private:
    std::shared_ptr<StageLog> log_;
    hip::GpuAdvection inner_;
};

// A stage that refuses its input, to check that its exception reaches the
// caller unchanged through the C boundary.
class RefusingPressure final : public hip::Projection3D {
public:
    hip::ProjectionResult project(hip::MacGridState3D&, const hip::ProjectionOptions&) override {
        throw std::domain_error("refused by the stage");
    }
    void reset() noexcept override {}
};

hip::MacGridParameters3D wiring_parameters(bool reflect) {
    hip::MacGridParameters3D p;
    p.stirStrength = 0.f;
    p.emitterRadius = .18f;
    p.emitterCenterZ = .3f;
    p.reflection = reflect;
    p.projectIterations = 16;
    p.macCormackVel = p.macCormackSmoke = false;
    return p;
}

void stages(Report& report, bool reflect) {
    const std::string tag = reflect ? "stages reflection: " : "stages simple: ";
    report.guard(tag + "wiring", [&] {
        const float dt = 1.f / 60.f;
        const int n = 8;
        hip::MacGridFluidSolver3D staged(n, 4.5f, wiring_parameters(reflect)),
            resident(n, 4.5f, wiring_parameters(reflect));
        auto log = std::make_shared<StageLog>();
        auto backend = std::make_shared<hip::GpuMacBackend>();
        staged.setPressureSolver(std::make_unique<LoggedPressure>(log, backend));
        staged.setAdvection(std::make_unique<LoggedAdvection>(log, backend));

        const hip::MacGridDomain3D& domain = staged.domain();
        report.check(domain.dimension() == hip::Dimension::D3, tag + "domain is 3D");
        report.check(domain.bounds().min.x == 0.f && domain.bounds().max.z == 4.5f, tag + "domain bounds [0, 4.5]^3");
        report.check(domain.resolution() == n && domain.cellSize() == 4.5f / n, tag + "domain resolution and spacing");

        for (int f = 0; f < 3; ++f) {
            staged.advance(dt);
            resident.advance(dt);
        }
        const int per_step = reflect ? 2 : 1;
        report.check(int(log->pressure_dt.size()) == 3 * per_step, tag + "pressure stage called per projection",
                     std::to_string(log->pressure_dt.size()));
        bool dts = true, budgets = true, fixed = true;
        for (float v : log->pressure_dt) dts = dts && v == dt / per_step;
        for (int v : log->pressure_budget) budgets = budgets && v == 16 / per_step;
        for (bool v : log->pressure_fixed) fixed = fixed && v;
        report.check(dts, tag + "pressure stage gets dt / projections per step");
        report.check(budgets, tag + "pressure stage gets K / projections per step");
        report.check(fixed, tag + "pressure stage gets the fixed sweep budget");
        report.check(log->face_calls == 9 * per_step && log->scalar_calls == 3, tag + "advection stage call counts",
                     std::to_string(log->face_calls) + " faces, " + std::to_string(log->scalar_calls) + " scalars");
        report.close(log->face_time, 9 * dt, 1e-6, tag + "face advection time");
        report.close(log->scalar_time, 3 * dt, 1e-6, tag + "scalar advection time");
        report.check(staged.lastProjection().residualAvailable, tag + "stage's residual is reported");

        // the staged GPU stages run the same kernels as the resident step
        const auto& a = staged.state();
        const auto& b = resident.state();
        const double err = std::max({max_diff(a.density(), b.density()), max_diff(a.velocityX(), b.velocityX()),
                                     max_diff(a.velocityY(), b.velocityY()), max_diff(a.velocityZ(), b.velocityZ())});
        report.check(err <= 2e-5, tag + "installed GPU stages agree with the resident step within 2e-5",
                     "max |diff| " + std::to_string(err));
        std::printf("   %sinstalled vs resident max |diff| %.3g (%s)\n", tag.c_str(), err,
                    same_bits(a.velocityX(), b.velocityX()) && same_bits(a.density(), b.density()) ? "bit-identical"
                                                                                                  : "not bitwise");
        report.close(staged.kineticEnergy(), resident.kineticEnergy(), 2e-5, tag + "kinetic energy");

        // per-field method flags still select the methods of the installed stage
        staged.parameters().macCormackVel = resident.parameters().macCormackVel = true;
        staged.advance(dt);
        resident.advance(dt);
        report.check(log->face_method == hip::AdvectionMethod::MacCormack &&
                         log->scalar_method == hip::AdvectionMethod::SemiLagrangian,
                     tag + "per-field method flags reach the stage");
        report.close(staged.kineticEnergy(), resident.kineticEnergy(), 2e-5, tag + "kinetic energy with MacCormack");

        staged.resetState();
        report.check(log->resets == 1, tag + "reset reaches the installed pressure stage");
        bool zero = true;
        for (float r : staged.state().density()) zero = zero && r == 0.f;
        report.check(zero, tag + "reset clears the density");

        staged.useDefaultPressureSolver();
        staged.useDefaultAdvection();
        report.check(log->destroyed == 1, tag + "the solver destroys a detached stage");
        const int before = log->face_calls;
        staged.advance(dt);
        report.check(log->face_calls == before, tag + "a detached stage is not called");

        bool rejected = false;
        try {
            staged.setAdvection(nullptr);
        } catch (const std::invalid_argument&) {
            rejected = true;
        }
        report.check(rejected, tag + "a null stage is rejected");
        staged.setBoxSize(3.f);
        report.check(domain.bounds().max.x == 3.f, tag + "setBoxSize moves the domain bounds");
        report.check(staged.deviceVolume().boxSize == 3.f && staged.deviceVolume().width == n &&
                         staged.deviceVolume().density != nullptr,
                     tag + "device volume follows the grid");

        staged.setPressureSolver(std::make_unique<RefusingPressure>());
        bool passed_through = false;
        try {
            staged.advance(dt);
        } catch (const std::domain_error& e) {
            passed_through = std::string(e.what()) == "refused by the stage";
        }
        report.check(passed_through, tag + "a stage's exception reaches the caller unchanged");
        staged.useDefaultPressureSolver();
        staged.advance(dt);
        report.check(std::isfinite(staged.kineticEnergy()), tag + "the solver steps again after a failed stage");
    });
}

}  // namespace

int main(int argc, char** argv) {
    const std::string group = argc > 1 ? argv[1] : "";
    Report report;
    if (group == "baselines" && argc > 2) {
        baselines(report, argv[2]);
    } else if (group == "replay") {
        replay(report);
    } else if (group == "plume") {
        plume(report);
    } else if (group == "stages") {
        stages(report, false);
        stages(report, true);
    } else {
        std::fprintf(stderr, "usage: dropin_tests baselines <baselines.json> | replay | plume | stages\n");
        return 100;
    }
    std::printf("%d failed\n", report.failures);
    return std::min(report.failures, 100);
}
