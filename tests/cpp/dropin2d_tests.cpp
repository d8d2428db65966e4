// dropin2d_tests.cpp - checks of the C++ 2D drop-in (include/vapor_mac2d.hpp)
// on the GPU, by group:
//   grid        index maps, extents, box resets and rejections
//   advection   exact translation of a linear field, rejections
//   projection  a manufactured pressure is recovered; a rotational field is
//               kept; option handling and fixed budgets
//   plume       the 24^2 plume stays divergence-free, bounded and rises
//   reflection  midpoint reflection: converged, replayable, not simple stepping
//   historical  the historical solver's translation, walls, bounds, replay
// Bounds are the ones vapor's own 2D tests assert (Mac2DTests.cpp,
// Historical2DTests.cpp). Prints "ok <name>" per check, "FAIL ..." on failure.
#include "vapor_mac2d.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numbers>
#include <numeric>
#include <string>
#include <vector>

using namespace vapor::hip;

namespace {

int failures = 0;

void check(bool condition, const std::string& name) {
    std::printf("%s %s\n", condition ? "ok" : "FAIL", name.c_str());
    if (!condition) ++failures;
}

template <class E = std::invalid_argument> bool throws(const std::function<void()>& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

double max_abs_diff(std::span<const float> a, std::span<const float> b) {
    double m = a.size() == b.size() ? 0.0 : INFINITY;
    for (std::size_t i = 0; i < a.size() && i < b.size(); ++i) m = std::max(m, double(std::abs(a[i] - b[i])));
    return m;
}

bool walls_closed(const MacGridState2D& s) {
    const int n = s.resolution();
    for (int i = 0; i < n; ++i)
        if (s.velocityX()[s.idX(0, i)] != 0.f || s.velocityX()[s.idX(n, i)] != 0.f || s.velocityY()[s.idY(i, 0)] != 0.f ||
            s.velocityY()[s.idY(i, n)] != 0.f)
            return false;
    return true;
}

ProjectionOptions tight(float dt) {
    ProjectionOptions o;
    o.dt = dt;
    o.relaxation = 1.6f;
    o.linearSolve.maxIterations = 4000;
    o.linearSolve.absoluteTolerance = 3e-6;
    o.linearSolve.relativeTolerance = 1e-6;
    return o;
}

void grid() {
    for (int n : {1, 2, 7}) {
        MacGridState2D s(n, 3.f);
        check(s.domain().dimension() == Dimension::D2 && s.domain().bounds().max.z == 0.f, "planar domain n=" + std::to_string(n));
        check(s.cellCount() == std::size_t(n * n) && s.velocityX().size() == std::size_t((n + 1) * n) &&
                  s.velocityY().size() == std::size_t(n * (n + 1)),
              "staggered extents n=" + std::to_string(n));
        std::vector<int> seen(s.velocityX().size(), 0);
        for (int y = 0; y < n; ++y)
            for (int x = 0; x <= n; ++x) ++seen.at(s.idX(x, y));
        check(std::all_of(seen.begin(), seen.end(), [](int c) { return c == 1; }), "x-face indices unique n=" + std::to_string(n));
        std::fill(s.density().begin(), s.density().end(), 1.f);
        s.setBoxSize(2.f);
        check(std::all_of(s.density().begin(), s.density().end(), [](float v) { return v == 0.f; }), "resize resets");
    }
    check(throws([] { MacGridState2D s(0); }), "rejects n = 0");
    check(throws<std::length_error>([] { MacGridState2D s(2147483647); }), "rejects an index overflow");
    check(throws([] { MacGridState2D s(4, 0.f); }), "rejects box 0");
    check(throws([] { MacGridFluidSolver2D s(0); }), "solver rejects n = 0");
}

void advection() {
    const int n = 8;
    MacGridState2D s(n, 1.f);
    std::fill(s.velocityX().begin(), s.velocityX().end(), .5f);
    std::fill(s.velocityY().begin(), s.velocityY().end(), -.25f);
    const SemiLagrangian2D sl;
    const float dt = .125f, h = s.cellSize();
    for (auto field : {MacField2D::Density, MacField2D::VelocityX, MacField2D::VelocityY}) {
        const int w = n + (field == MacField2D::VelocityX), hh = n + (field == MacField2D::VelocityY);
        const float ox = field == MacField2D::VelocityX ? 0.f : .5f, oy = field == MacField2D::VelocityY ? 0.f : .5f;
        std::vector<float> src(w * hh), out(w * hh);
        for (int y = 0; y < hh; ++y)
            for (int x = 0; x < w; ++x) src[x + w * y] = 1.f + .2f * (x + ox) * h + .3f * (y + oy) * h;
        sl.advect({s.domain(), s.velocityX(), s.velocityY(), src, out, 0.f, field});
        check(max_abs_diff(src, out) <= 1e-6, "dt = 0 is the identity");
        sl.advect({s.domain(), s.velocityX(), s.velocityY(), src, out, dt, field});
        double err = 0;
        for (int y = 1; y < hh - 1; ++y)
            for (int x = 1; x < w - 1; ++x)
                err = std::max(err, std::abs(double(out[x + w * y]) -
                                             (1. + .2 * ((x + ox) * h - .5 * dt) + .3 * ((y + oy) * h + .25 * dt))));
        check(err <= 3e-7, "linear field translates exactly, field " + std::to_string(int(field)));
        check(throws([&] { sl.advect({s.domain(), s.velocityX(), s.velocityY(), src, src, dt, field}); }), "rejects aliasing");
        check(throws([&] { sl.advect({s.domain(), s.velocityX(), s.velocityY(), src, out, -dt, field}); }), "rejects dt < 0");
    }
}

void projection() {
    GpuPressureSolver2D solver;
    for (int n : {1, 8, 15})
        for (float box : {1.f, 3.f})
            for (float dt : {.01f, .1f}) {
                MacGridState2D s(n, box);
                std::vector<float> exact(n * n);
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x)
                        exact[s.idC(x, y)] = std::cos(std::numbers::pi_v<float> * (x + .5f) / n) *
                                             std::cos(2.f * std::numbers::pi_v<float> * (y + .5f) / n);
                const double mean = std::accumulate(exact.begin(), exact.end(), 0.0) / exact.size();
                for (auto& p : exact) p -= float(mean);
                for (int y = 0; y < n; ++y)
                    for (int x = 1; x < n; ++x)
                        s.velocityX()[s.idX(x, y)] = dt / s.cellSize() * (exact[s.idC(x, y)] - exact[s.idC(x - 1, y)]);
                for (int y = 1; y < n; ++y)
                    for (int x = 0; x < n; ++x)
                        s.velocityY()[s.idY(x, y)] = dt / s.cellSize() * (exact[s.idC(x, y)] - exact[s.idC(x, y - 1)]);
                const auto r = solver.project(s, tight(dt));
                double perr = 0, verr = 0;
                for (std::size_t i = 0; i < exact.size(); ++i) perr = std::max(perr, std::abs(solver.pressure()[i] - exact[i]));
                for (float u : s.velocityX()) verr = std::max(verr, double(std::abs(u)));
                for (float v : s.velocityY()) verr = std::max(verr, double(std::abs(v)));
                const std::string tag = " n=" + std::to_string(n);
                check(r.residualAvailable && r.linearSolve.converged, "manufactured solve converged" + tag);
                check(perr <= 8e-5 && verr <= 3e-5 && r.divergenceAfter < 2e-4f && walls_closed(s), "gradient removed, p recovered" + tag);
            }
    MacGridState2D rest(8);
    auto options = tight(.02f);
    check(solver.project(rest, options).linearSolve.iterations == 0, "rest exits early");
    options.linearSolve.fixedIterations = true;
    options.linearSolve.maxIterations = 17;
    check(solver.project(rest, options).linearSolve.iterations == 17, "fixed budget");
    options = tight(.02f);
    options.relaxation = 2.f;
    check(throws([&] { solver.project(rest, options); }), "rejects relaxation 2");
    options = tight(0.f);
    check(throws([&] { solver.project(rest, options); }), "rejects dt 0");
    rest.velocityY()[3] = NAN;
    check(throws([&] { solver.project(rest, tight(.02f)); }), "rejects a NaN face");
}

double center_y(const MacGridState2D& s) {
    double mass = 0, moment = 0;
    for (int y = 0; y < s.resolution(); ++y)
        for (int x = 0; x < s.resolution(); ++x) {
            const float d = s.density()[s.idC(x, y)];
            mass += d;
            moment += d * (y + .5) * s.cellSize();
        }
    return mass > 0 ? moment / mass : 0;
}

void plume() {
    MacGridFluidSolver2D solver(24);
    solver.parameters().sourceStrength = 0.f;
    for (int i = 0; i < 5; ++i) solver.advance(1.f / 60.f);
    check(solver.diagnostics().kineticEnergy == 0.f, "rest stays at rest");
    solver.parameters().sourceStrength = 4.f;
    bool bounded = true, divergence_free = true, closed = true;
    for (int i = 0; i < 120; ++i) {
        solver.advance(1.f / 60.f);
        divergence_free = divergence_free && solver.diagnostics().maxDivergence < .005f;
        closed = closed && walls_closed(solver.state());
        for (float d : solver.state().density()) bounded = bounded && std::isfinite(d) && d >= 0.f && d <= 8.01f;
    }
    check(divergence_free, "plume divergence < 0.005 every frame");
    check(closed && bounded, "walls closed, density bounded");
    check(solver.diagnostics().kineticEnergy > .001f, "buoyancy produces motion");
    const double before = center_y(solver.state());
    solver.parameters().sourceStrength = 0.f;
    for (int i = 0; i < 30; ++i) solver.advance(1.f / 60.f);
    check(center_y(solver.state()) > before + .02, "smoke rises after emission stops");
    const auto snapshot = solver.state().density();
    const std::vector<float> kept(snapshot.begin(), snapshot.end());
    check(throws([&] { solver.advance(0.f); }), "rejects dt = 0");
    check(max_abs_diff(solver.state().density(), kept) == 0.0, "a rejected step changes nothing");
    solver.resetState();
    check(solver.diagnostics().kineticEnergy == 0.f && center_y(solver.state()) == 0, "reset clears the state");
    MacGridFluidSolver2D deflt;
    deflt.advance(1.f / 60.f);
    check(deflt.domain().resolution() == 128 && deflt.lastProjection().residualAvailable && deflt.deviceDensity(),
          "the ready-to-run plume steps");
}

void reflection() {
    MacGridState2D initial(20, 4.f);
    const auto psi = [](int x, int y) { return .8 * std::sin(std::numbers::pi * x / 20.) * std::sin(2. * std::numbers::pi * y / 20.); };
    for (int y = 0; y < 20; ++y)
        for (int x = 0; x <= 20; ++x) initial.velocityX()[initial.idX(x, y)] = float((psi(x, y + 1) - psi(x, y)) / initial.cellSize());
    for (int y = 0; y <= 20; ++y)
        for (int x = 0; x < 20; ++x) initial.velocityY()[initial.idY(x, y)] = float(-(psi(x + 1, y) - psi(x, y)) / initial.cellSize());
    MacGridFluidSolver2D reflected(initial), simple(initial), replay(initial);
    for (auto* s : {&reflected, &simple, &replay}) {
        s->parameters().sourceStrength = 0.f;
        s->parameters().buoyancy = 0.f;
        s->parameters().reflection = s != &simple;
    }
    bool converged = true, quiet = true;
    for (int frame = 0; frame < 40; ++frame) {
        for (auto* s : {&reflected, &simple, &replay}) s->advance(.02f);
        converged = converged && reflected.midpointProjection().linearSolve.converged && reflected.lastProjection().linearSolve.converged;
        quiet = quiet && reflected.diagnostics().maxDivergence < 1e-4f && walls_closed(reflected.state());
    }
    check(converged, "midpoint and final projections converge");
    check(quiet, "reflection stays divergence-free with closed walls");
    check(max_abs_diff(reflected.state().velocityX(), replay.state().velocityX()) == 0.0, "replay is bit-identical");
    check(max_abs_diff(reflected.state().velocityX(), simple.state().velocityX()) > 1e-4, "reflection differs from simple stepping");
    reflected.resetState();
    check(!reflected.midpointProjection().residualAvailable, "reset clears the midpoint record");
}

void historical() {
    ReflectionMacFluidSolver2D sl(16), mc(16);
    sl.macCormackSmoke = false;
    for (auto* s : {&sl, &mc}) {
        std::fill(s->vx.begin(), s->vx.end(), 1.f);
        s->smoke[s->idC(7, 8)] = 1.f;
        s->push();
        s->advectSmoke(.25f * s->h());
        s->pull();
    }
    double err = 0;
    for (int y = 0; y < 16; ++y)
        for (int x = 0; x < 16; ++x) {
            const float expected = y == 8 ? (x == 7 ? .75f : x == 8 ? .25f : 0.f) : 0.f;
            err = std::max(err, double(std::abs(sl.smoke[sl.idC(x, y)] - expected)));
        }
    check(err <= 1e-6, "semi-Lagrangian moves the spike a quarter cell");
    check(sl.smoke != mc.smoke, "MacCormack differs from semi-Lagrangian");
    sl.resetState();
    bool bounded = true, closed = true;
    for (int f = 0; f < 120; ++f) {
        sl.advance(1.f / 60.f);
        sl.pull();
        for (float d : sl.smoke) bounded = bounded && std::isfinite(d) && d >= 0.f && d <= 1.f;
        for (int i = 0; i < sl.n; ++i)
            closed = closed && sl.vx[sl.idX(0, i)] == 0.f && sl.vx[sl.idX(sl.n, i)] == 0.f && sl.vy[sl.idY(i, 0)] == 0.f &&
                     sl.vy[sl.idY(i, sl.n)] == 0.f;
    }
    check(bounded && closed, "smoke bounded, walls closed for 120 frames");
    const auto first = sl.smoke;
    sl.resetState();
    for (int f = 0; f < 120; ++f) sl.advance(1.f / 60.f);
    sl.pull();
    check(sl.smoke == first, "reset and replay give the same smoke");
    check(throws([&] { sl.setBoxSize(2.f); }), "box size other than 1 rejected");
    check(throws([] { ReflectionMacFluidSolver2D bad(1); }), "resolution 1 rejected");
    ReflectionMacFluidSolver2D plume;
    plume.advance(1.f / 60.f);
    check(plume.n == 128 && !plume.macCormackSmoke && plume.diagnostics().maxVelocity > 0.f, "the historical plume steps");
}

}  // namespace

int main(int argc, char** argv) {
    const std::string group = argc > 1 ? argv[1] : "";
    try {
        if (group == "grid") grid();
        else if (group == "advection") advection();
        else if (group == "projection") projection();
        else if (group == "plume") plume();
        else if (group == "reflection") reflection();
        else if (group == "historical") historical();
        else {
            std::printf("FAIL unknown group %s\n", group.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("FAIL %s: exception %s\n", group.c_str(), e.what());
        return 1;
    }
    return failures ? 1 : 0;
}
