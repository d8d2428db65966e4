// reference_simplicial_tests.cpp - vapor's simplicial 3D solver with the HIP
// backend installed, checked against the same solver on its CPU path.
//
// Built by oracle/Makefile from vapor's own simplicial sources
// (src/solvers/simplicial3d/*.cpp) and include/vapor_gpu_simplicial3d.hpp,
// linked against libvapor_mac3d.so. The modes follow vapor's own GPU test
// (tests/SimplicialGpu3DTests.cpp:18-91) and add what the HIP backend
// promises beyond it:
//
// Kinds: box, domain, nobins (the general domain without spatial bins).
//
//   exact  <mesh.tet> box|domain <steps>   reference recovery order: the
//          seeded circulation and every step's flux, potential, vorticity and
//          smoke are bit-identical to the CPU solver's.
//   parity <mesh.tet> box|domain <steps>   fast recovery order: vapor's own
//          bars (circulation 1e-8, flux 1e-5, density 1e-3, boundary flux
//          1e-10, divergence 1e-8, resize 1e-5, replay 1e-10).
//   motion <mesh.tet> <frames>             vapor's motion check: smoke rises.
//   errors <mesh.tet> box|domain           the reference's exceptions.
//   bench  <mesh.tet> box|domain <steps>   per-stage times, CPU and GPU.
//
// Exit status 0 on success, 1 on a failed check (the message says which).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "vapor_gpu_simplicial3d.hpp"

using namespace simplicial3d;

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

void require(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error("check failed: " + what);
}

// kind "box" and "domain" load as vapor's app does; "nobins" rebuilds the
// general domain through the public Mesh(vertices, tets) constructor, which
// builds no spatial bins (Mesh::locate and clipSegment then scan every tet and
// every boundary face).
Mesh load(const std::string& path, const std::string& kind) {
    if (kind == "box") return Mesh::loadBox(path, 3);
    Mesh m = Mesh::loadDomain(path, 2);
    if (kind != "nobins") return m;
    std::vector<std::array<int, 4>> tets;
    for (const auto& t : m.tets) tets.push_back(t.vertices);
    return Mesh(m.vertices, tets);
}

template <class T> double max_diff(const std::vector<T>& a, const std::vector<T>& b) {
    require(a.size() == b.size(), "sizes differ");
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::abs(double(a[i]) - double(b[i])));
    return m;
}

template <class T> size_t bit_mismatches(const std::vector<T>& a, const std::vector<T>& b) {
    require(a.size() == b.size(), "sizes differ");
    size_t n = 0;
    for (size_t i = 0; i < a.size(); ++i) n += std::memcmp(&a[i], &b[i], sizeof(T)) != 0;
    return n;
}

// The seeded potential of tests/SimplicialGpu3DTests.cpp:44-48.
std::vector<double> seeded_potential(const Mesh& mesh) {
    std::vector<double> potential(mesh.edges.size());
    for (size_t e = 0; e < potential.size(); ++e) {
        auto ids = mesh.edges[e].vertices;
        auto a = mesh.vertices[ids[0]], b = mesh.vertices[ids[1]], p = (a + b) * .5;
        potential[e] = std::sin(3.141592653589793 * p.x / mesh.size) * std::sin(3.141592653589793 * p.y / mesh.size) *
                       std::sin(3.141592653589793 * p.z / mesh.size) * (b.z - a.z);
    }
    return potential;
}

// Smoke mass and height (tests/SimplicialGpu3DTests.cpp:7-16).
std::pair<double, double> smoke_center(const SimplicialFluidSolver3D& solver) {
    double mass = 0, height = 0;
    for (size_t i = 0; i < solver.domain().tets.size(); ++i) {
        const auto& tet = solver.domain().tets[i];
        const double amount = solver.density()[i] * tet.volume;
        mass += amount;
        for (int v : tet.vertices) height += .25 * amount * solver.domain().vertices[v].z;
    }
    return {mass > 0 ? height / mass : 0, mass};
}

int exact(const std::string& path, const std::string& kind, int steps) {
    require(vsx_set_default_recovery(VSX_RECOVER_REFERENCE) == VGX_OK, "reference recovery order");
    Mesh mesh = load(path, kind);
    const bool box = mesh.boxDomain;
    SimplicialFluidSolver3D cpu(mesh), gpu(mesh);
    GpuSimplicial3D backend;
    backend.initialize(gpu);
    gpu.setComputeBackend(&backend);
    std::cout << "exact " << path << " on " << backend.device() << ": tets=" << mesh.tets.size()
              << " edges=" << mesh.edges.size() << " dual=" << cpu.dual().vertices.size() << std::endl;
    if (box) {
        auto potential = seeded_potential(mesh);
        cpu.setPotential(potential);
        gpu.setPotential(potential);
        cpu.advectVorticity(.017);
        gpu.advectVorticity(.017);
        const size_t bad = bit_mismatches(cpu.vorticity(), gpu.vorticity());
        std::cout << "seeded circulation: " << bad << " of " << cpu.vorticity().size()
                  << " values differ, max " << max_diff(cpu.vorticity(), gpu.vorticity()) << std::endl;
        require(bad == 0, "seeded circulation bit-identical");
        cpu.resetState();
        gpu.resetState();
    }
    for (int i = 0; i < steps; ++i) {
        cpu.advance(1.f / 30);
        gpu.advance(1.f / 30);
        const size_t flux = bit_mismatches(cpu.flux(), gpu.flux()), phi = bit_mismatches(cpu.potential(), gpu.potential()),
                     omega = bit_mismatches(cpu.vorticity(), gpu.vorticity()),
                     smoke = bit_mismatches(cpu.vertexDensity(), gpu.vertexDensity()),
                     cells = bit_mismatches(cpu.density(), gpu.density());
        const bool same_solve = cpu.lastRecovery().iterations == gpu.lastRecovery().iterations &&
                                cpu.lastRecovery().finalResidual == gpu.lastRecovery().finalResidual &&
                                cpu.lastRecovery().initialResidual == gpu.lastRecovery().initialResidual;
        if (i % 10 == 0 || i + 1 == steps || flux || phi || omega || smoke || cells || !same_solve)
            std::cout << "step " << i << " differing values: flux=" << flux << " potential=" << phi
                      << " vorticity=" << omega << " vertexDensity=" << smoke << " density=" << cells
                      << " cg=" << gpu.lastRecovery().iterations << "/" << cpu.lastRecovery().iterations
                      << " maxFluxDiff=" << max_diff(cpu.flux(), gpu.flux()) << std::endl;
        require(flux == 0 && phi == 0 && omega == 0 && smoke == 0 && cells == 0 && same_solve,
                "step " + std::to_string(i) + " bit-identical");
    }
    std::cout << "PASS exact: " << steps << " steps bit-identical" << std::endl;
    return 0;
}

int parity(const std::string& path, const std::string& kind, int steps) {
    require(vsx_set_default_recovery(VSX_RECOVER_FAST) == VGX_OK, "fast recovery order");
    Mesh mesh = load(path, kind);
    const bool box = mesh.boxDomain;
    SimplicialFluidSolver3D cpu(mesh), gpu(mesh);
    GpuSimplicial3D backend;
    backend.initialize(gpu);
    gpu.setComputeBackend(&backend);
    std::cout << "parity " << path << " on " << backend.device() << std::endl;
    if (box) {
        auto potential = seeded_potential(mesh);
        cpu.setPotential(potential);
        gpu.setPotential(potential);
        cpu.advectVorticity(.017);
        gpu.advectVorticity(.017);
        const double error = max_diff(cpu.vorticity(), gpu.vorticity());
        std::cout << "seeded circulation error=" << error << std::endl;
        require(error <= 1e-8, "seeded circulation within 1e-8");
        cpu.resetState();
        gpu.resetState();
    }
    double cpu_ms = 0, gpu_ms = 0;
    for (int i = 0; i < steps; ++i) {
        auto a = Clock::now();
        cpu.advance(1.f / 30);
        cpu_ms += ms_since(a);
        auto b = Clock::now();
        gpu.advance(1.f / 30);
        gpu_ms += ms_since(b);
        const double flux_error = max_diff(cpu.flux(), gpu.flux());
        const double density_error = max_diff(cpu.vertexDensity(), gpu.vertexDensity());
        const auto d = gpu.diagnostics();
        if (i % 10 == 0 || i + 1 == steps)
            std::cout << i << " fluxError=" << flux_error << " densityError=" << density_error
                      << " cg=" << gpu.lastRecovery().iterations << "/" << cpu.lastRecovery().iterations
                      << " substeps=" << gpu.timings().substeps << " gpuMs=" << gpu_ms / (i + 1)
                      << " cpuMs=" << cpu_ms / (i + 1) << std::endl;
        require(std::isfinite(d.maxSpeed) && d.maxBoundaryFlux <= 1e-10 && d.maxDivergence <= 1e-8 &&
                    flux_error <= 1e-5 && density_error <= 1e-3,
                "step " + std::to_string(i) + " within vapor's GPU bars");
    }
    const auto& t = gpu.timings();
    std::cout << "last GPU-path step ms: circulation=" << t.advectionMs << " forces=" << t.forcesMs
              << " recovery=" << t.recoveryMs << " smoke=" << t.smokeMs << std::endl;
    cpu.setBoxSize(2.5f);
    gpu.setBoxSize(2.5f);
    cpu.advance(.01f);
    gpu.advance(.01f);
    const double resize_error = max_diff(cpu.flux(), gpu.flux());
    require(resize_error <= 1e-5 && gpu.lastRecovery().converged, "resize parity and recovery");
    auto replay = gpu.flux();
    gpu.resetState();
    gpu.advance(.01f);
    const double replay_error = max_diff(replay, gpu.flux());
    require(replay_error <= 1e-10, "reset replay");
    std::cout << "PASS resize and reset; resize flux error=" << resize_error << " replay error=" << replay_error
              << std::endl;
    return 0;
}

int motion(const std::string& path, int frames) {
    SimplicialFluidSolver3D solver(Mesh::loadBox(path, 3));
    GpuSimplicial3D backend;
    backend.initialize(solver);
    solver.setComputeBackend(&backend);
    solver.parameters().smokeDecay = 0;
    double first_height = 0, first_mass = 0, elapsed = 0;
    for (int i = 0; i < frames; ++i) {
        if (i == 30) {
            auto center = smoke_center(solver);
            first_height = center.first;
            first_mass = center.second;
            solver.parameters().emitterEnabled = false;
        }
        auto start = Clock::now();
        solver.advance(1.f / 60);
        elapsed += ms_since(start);
    }
    const auto [height, mass] = smoke_center(solver);
    std::cout << "GPU motion tets=" << solver.domain().tets.size() << " frames=" << frames << " center z=" << first_height
              << " -> " << height << " mass=" << first_mass << " -> " << mass
              << " speed=" << solver.diagnostics().maxSpeed << " ms/frame=" << elapsed / frames << std::endl;
    if (frames >= 90) require(height > first_height + .03, "GPU smoke did not rise");
    return 0;
}

template <class E, class F> void expect_throw(F&& f, const std::string& message, const char* what) {
    try {
        f();
    } catch (const E& e) {
        require(message.empty() || message == e.what(), std::string(what) + ": message '" + e.what() + "'");
        std::cout << what << ": " << e.what() << std::endl;
        return;
    }
    throw std::runtime_error(std::string("expected an exception: ") + what);
}

int errors(const std::string& path, const std::string& kind) {
    Mesh mesh = load(path, kind);
    const bool box = mesh.boxDomain;
    SimplicialFluidSolver3D solver(mesh);
    GpuSimplicial3D backend;
    backend.initialize(solver);
    solver.setComputeBackend(&backend);
    solver.advance(1.f / 30);
    const int budget = solver.parameters().cgIterations;
    std::vector<double> rhs(solver.vorticity()), x;
    rhs[rhs.size() / 2] = std::nan("");
    expect_throw<std::runtime_error>([&] { backend.recover(solver, rhs, x); }, "Non-finite CG RHS", "non-finite rhs");
    rhs[rhs.size() / 2] = INFINITY;
    expect_throw<std::runtime_error>([&] { backend.recover(solver, rhs, x); }, "Non-finite CG RHS", "infinite rhs");
    std::vector<double> omega;
    expect_throw<std::invalid_argument>([&] { backend.advectCirculation(solver, -1, omega); }, "", "negative dt");
    std::vector<float> density;
    expect_throw<std::invalid_argument>([&] { backend.advectDensity(solver, NAN, density); }, "", "non-finite dt");
    solver.parameters().cgIterations = -1;
    expect_throw<std::invalid_argument>([&] { backend.recover(solver, solver.vorticity(), x); }, "Invalid CG parameters",
                                        "negative iterations");
    solver.parameters().cgIterations = 100;
    solver.parameters().cgRelativeTolerance = -1;
    expect_throw<std::invalid_argument>([&] { backend.recover(solver, solver.vorticity(), x); }, "Invalid CG parameters",
                                        "negative tolerance");
    solver.parameters().cgRelativeTolerance = 1e-9;
    // A zero budget reports the initial residual unconverged, as the CPU solve does.
    solver.parameters().cgIterations = 0;
    auto r = backend.recover(solver, solver.vorticity(), x);
    std::cout << "zero budget: converged=" << r.converged << " iterations=" << r.iterations << std::endl;
    require(r.iterations == 0, "zero budget takes no iteration");
    // The backend recovers after an error: the next step runs normally.
    solver.parameters().cgIterations = budget;
    solver.advance(1.f / 30);
    require(solver.lastRecovery().converged, "step after errors");
    std::cout << "PASS errors" << std::endl;
    return 0;
}

int bench(const std::string& path, const std::string& kind, int steps) {
    require(vsx_set_default_recovery(VSX_RECOVER_FAST) == VGX_OK, "fast recovery order");
    Mesh mesh = load(path, kind);
    SimplicialFluidSolver3D cpu(mesh), gpu(mesh);
    GpuSimplicial3D backend;
    backend.initialize(gpu);
    gpu.setComputeBackend(&backend);
    // Warm both up to a state with smoke and motion.
    for (int i = 0; i < 3; ++i) {
        cpu.advance(1.f / 30);
        gpu.advance(1.f / 30);
    }
    double cpu_total = 0, gpu_total = 0, ca = 0, cr = 0, cs = 0, ga = 0, gr = 0, gs = 0;
    int substeps = 0;
    for (int i = 0; i < steps; ++i) {
        auto a = Clock::now();
        cpu.advance(1.f / 30);
        cpu_total += ms_since(a);
        ca += cpu.timings().advectionMs;
        cr += cpu.timings().recoveryMs;
        cs += cpu.timings().smokeMs;
        auto b = Clock::now();
        gpu.advance(1.f / 30);
        gpu_total += ms_since(b);
        ga += gpu.timings().advectionMs;
        gr += gpu.timings().recoveryMs;
        gs += gpu.timings().smokeMs;
        substeps += gpu.timings().substeps;
    }
    // The backend's stage calls on the current state, device time apart.
    const double dt = 1. / 60;
    std::vector<double> omega, x;
    std::vector<float> density;
    double wall[3] = {0, 0, 0}, kernel[3] = {0, 0, 0};
    const int reps = 20;
    for (int k = 0; k < reps; ++k) {
        auto t = Clock::now();
        backend.advectCirculation(gpu, dt, omega);
        wall[0] += ms_since(t);
        kernel[0] += vsx_last_kernel_ms();
        t = Clock::now();
        backend.advectDensity(gpu, dt, density);
        wall[1] += ms_since(t);
        kernel[1] += vsx_last_kernel_ms();
        t = Clock::now();
        auto r = backend.recover(gpu, gpu.vorticity(), x);
        wall[2] += ms_since(t);
        kernel[2] += vsx_last_kernel_ms();
        require(r.converged, "bench recovery converged");
    }
    std::printf("bench %s tets=%zu edges=%zu dual=%zu steps=%d substeps/step=%.2f\n", path.c_str(), mesh.tets.size(),
                mesh.edges.size(), cpu.dual().vertices.size(), steps, double(substeps) / steps);
    std::printf("step ms: cpu=%.3f gpu=%.3f speedup=%.2f\n", cpu_total / steps, gpu_total / steps, cpu_total / gpu_total);
    std::printf("stage ms per step (cpu/gpu): circulation=%.3f/%.3f recovery=%.3f/%.3f smoke=%.3f/%.3f\n", ca / steps,
                ga / steps, cr / steps, gr / steps, cs / steps, gs / steps);
    std::printf("backend call ms (wall/kernel): circulation=%.4f/%.4f density=%.4f/%.4f recover=%.4f/%.4f cg=%d\n",
                wall[0] / reps, kernel[0] / reps, wall[1] / reps, kernel[1] / reps, wall[2] / reps, kernel[2] / reps,
                gpu.lastRecovery().iterations);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: simplicial_tests exact|parity|motion|errors|bench <mesh.tet> [box|domain|nobins] [steps]\n";
        return 2;
    }
    const std::string mode = argv[1], path = argv[2];
    const std::string kind = argc > 3 ? argv[3] : "box";
    const int count = argc > 4 ? std::stoi(argv[4]) : 10;
    try {
        if (mode == "exact") return exact(path, kind, count);
        if (mode == "parity") return parity(path, kind, count);
        if (mode == "motion") return motion(path, argc > 3 ? std::stoi(argv[3]) : 90);
        if (mode == "errors") return errors(path, kind);
        if (mode == "bench") return bench(path, kind, count);
    } catch (const std::exception& e) {
        std::cerr << "FAIL: " << e.what() << std::endl;
        return 1;
    }
    std::cerr << "unknown mode " << mode << "\n";
    return 2;
}
