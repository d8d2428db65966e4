// simulation_tests.cpp - vapor's whole-step tests (tests/SimulationTests.cpp of
// the reference) run unchanged in structure against the GPU drop-in
// vapor::hip::MacGridFluidSolver3D (include/vapor_mac3d.hpp).
//
//   simulation_tests            16 published whole-step baselines
//   simulation_tests seeded_replay
//   simulation_tests plume      24^3, 180 frames, stability bounds
#include "vapor_mac3d.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>

using vapor::hip::MacGridFluidSolver3D;

namespace {
struct Baseline {
    int n, reflect, flags;
    double energy, maxVelocity, divergence, densitySum, weightedDensity;
};
// Published baselines, tests/SimulationTests.cpp:17-34 of the reference: six
// 1/60 s frames, 16 projection iterations, radius .18, centreZ .3, no stir.
const Baseline baselines[] = {
    {8, 0, 0, 5.32227563858, 1.79323649406, 0.00224227714352, 12.5633865102, 132.161955374},
    {8, 0, 1, 5.31826114655, 1.79280233383, 0.00225179060362, 12.3054398399, 130.187893328},
    {8, 0, 2, 5.63376235962, 1.80886173248, 0.00317340437323, 12.5620898544, 132.073204243},
    {8, 0, 3, 5.62872838974, 1.8083230257, 0.0031795501709, 12.3062464306, 130.151977868},
    {8, 1, 0, 5.3048157692, 1.79005801678, 0.0247040055692, 12.560423065, 132.109861564},
    {8, 1, 1, 5.30073595047, 1.78961789608, 0.0246352888644, 12.3041853862, 130.164586014},
    {8, 1, 2, 5.63356494904, 1.80607974529, 0.0130067132413, 12.5588216419, 132.010692464},
    {8, 1, 3, 5.62848520279, 1.80551791191, 0.0129552418366, 12.3049406262, 130.123926043},
    {12, 0, 0, 6.23518705368, 1.70936644077, 0.0195950269699, 47.985539313, 721.707724965},
    {12, 0, 1, 6.22765159607, 1.70904874802, 0.0196563154459, 46.408280425, 697.377404871},
    {12, 0, 2, 6.55779933929, 1.77157914639, 0.0131081836298, 47.9547437223, 721.24064451},
    {12, 0, 3, 6.55442667007, 1.76948368549, 0.0132885873318, 46.4151888065, 697.523262066},
    {12, 1, 0, 6.25715303421, 1.71856784821, 0.115855634212, 48.0296325346, 722.386514655},
    {12, 1, 1, 6.24975967407, 1.71815788746, 0.11610905081, 46.4359233264, 697.809387388},
    {12, 1, 2, 6.60593128204, 1.801410079, 0.0750236585736, 47.997528489, 721.888705373},
    {12, 1, 3, 6.60220718384, 1.79932284355, 0.0749910548329, 46.446071623, 698.00428824},
};
void near(double actual, double expected, const char* name) {
    const double tolerance = 2e-5 * std::max(1.0, std::abs(expected));
    if (!std::isfinite(actual) || std::abs(actual - expected) > tolerance)
        throw std::runtime_error(std::string(name) + ": actual=" + std::to_string(actual) +
                                 ", baseline=" + std::to_string(expected));
}
bool equal(std::span<const float> a, std::span<const float> b) {
    return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin());
}

void seededReplay() {
    MacGridFluidSolver3D first(8), second(8), other(8);
    for (auto* solver : {&first, &second, &other}) {
        solver->parameters().emitterRadius = .18f;
        solver->parameters().emitterCenterZ = .3f;
        solver->parameters().projectIterations = 16;
        solver->parameters().randomSeed = 12345;
        solver->resetState();
    }
    other.parameters().randomSeed = 54321;
    other.resetState();
    for (int frame = 0; frame < 6; ++frame) {
        first.advance(1.f / 60.f);
        other.advance(1.f / 60.f);
        second.advance(1.f / 60.f);
    }
    if (!equal(first.state().velocityX(), second.state().velocityX()) ||
        !equal(first.state().density(), second.state().density()))
        throw std::runtime_error("Seeded solvers depend on each other's random stream");
    if (equal(first.state().velocityX(), other.state().velocityX()))
        throw std::runtime_error("Different stirring seeds produced identical velocity");
    first.resetState();
    for (int frame = 0; frame < 6; ++frame) first.advance(1.f / 60.f);
    if (!equal(first.state().velocityX(), second.state().velocityX()) ||
        !equal(first.state().velocityY(), second.state().velocityY()) ||
        !equal(first.state().velocityZ(), second.state().velocityZ()) ||
        !equal(first.state().density(), second.state().density()))
        throw std::runtime_error("Reset did not reproduce the seeded simulation");
}
}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc > 1 && std::string(argv[1]) == "plume") {
            for (bool reflection : {false, true}) {
                MacGridFluidSolver3D solver(24);
                solver.parameters().reflection = reflection;
                solver.parameters().projectIterations = 300;
                solver.parameters().emitterRadius = .08f;
                solver.parameters().sorOmega = 1.8f;
                for (int frame = 0; frame < 180; ++frame) {
                    solver.advance(1.f / 60.f);
                    const auto d = solver.diagnostics();
                    if (!std::isfinite(d.kineticEnergy) || !std::isfinite(d.maxDivergence) || d.maxDivergence > .005f)
                        throw std::runtime_error("3D plume became non-finite or lost incompressibility");
                    for (float value : solver.state().density())
                        if (!std::isfinite(value) || value < 0.f || value > 1.00001f)
                            throw std::runtime_error("3D plume density outside limiter bounds");
                }
                std::cout << "3D plume reflection=" << reflection << " energy=" << solver.kineticEnergy()
                          << " divergence=" << solver.computeDivergenceNorm() << '\n';
            }
            return 0;
        }
        if (argc > 1 && std::string(argv[1]) == "seeded_replay") {
            seededReplay();
            std::cout << "PASS independent random streams and reset replay\n";
            return 0;
        }
        for (const auto& baseline : baselines) {
            MacGridFluidSolver3D solver(baseline.n);
            solver.parameters().stirStrength = 0.f;
            solver.parameters().emitterRadius = .18f;
            solver.parameters().emitterCenterZ = .3f;
            solver.parameters().macCormackSmoke = (baseline.flags & 1) != 0;
            solver.parameters().macCormackVel = (baseline.flags & 2) != 0;
            solver.parameters().reflection = baseline.reflect != 0;
            solver.parameters().projectIterations = 16;
            for (int frame = 0; frame < 6; ++frame) solver.advance(1.f / 60.f);
            double sum = 0, weighted = 0;
            const auto density = solver.state().density();
            for (std::size_t i = 0; i < density.size(); ++i) {
                sum += density[i];
                weighted += density[i] * double(1 + i % 31);
            }
            near(solver.kineticEnergy(), baseline.energy, "Kinetic energy");
            near(solver.maxVelocity(), baseline.maxVelocity, "Maximum velocity");
            near(solver.computeDivergenceNorm(), baseline.divergence, "Divergence");
            near(sum, baseline.densitySum, "Density sum");
            near(weighted, baseline.weightedDensity, "Weighted density");
            solver.resetState();
            near(solver.kineticEnergy(), 0, "Reset energy");
            for (float value : solver.state().density()) near(value, 0, "Reset density");
            solver.setBoxSize(3.f);
            near(solver.boxSize(), 3, "Physical domain update");
            std::cout << "PASS n=" << baseline.n << " reflect=" << baseline.reflect << " MC flags=" << baseline.flags
                      << '\n';
        }
    } catch (const std::exception& error) {
        std::cerr << "FAIL simulation regression: " << error.what() << '\n';
        return 1;
    }
    return 0;
}
