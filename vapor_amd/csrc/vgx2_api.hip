// vgx2_api.hip - C ABI of the MAC 2D solvers (include/vapor_mac2d.h).
//
// Host orchestration of vapor's MacGridFluidSolver2D and
// ReflectionMacFluidSolver2D on one HIP stream per handle, every field
// resident in device memory. A MacGridFluidSolver2D step enqueues its kernels
// and reads one small status record at its end, which carries the
// projections' results and any failure the reference would have thrown.
#include "vgx2_device.hpp"
#include "vgx2_kernels.hpp"
#include "vgx_host.hpp"

#include "../../include/vapor_mac2d.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <utility>

using vgx_host::check;
using vgx_host::device_alloc;
using vgx_host::Failure;
using vgx_host::guarded;
using vgx_host::invalid;
using namespace vgx2;

namespace {

bool finite(float v) { return std::isfinite(v); }

// MacGridDomain2D's checks (MacGridDomain2D.cpp:6-19).
void check_domain(int n, float box) {
    if (n <= 0) invalid("2D MAC resolution must be positive");
    const auto nn = static_cast<size_t>(n);
    if (nn > static_cast<size_t>(std::numeric_limits<int>::max()) / (nn + 1))
        throw Failure{VGX_ERANGE, "2D MAC grid exceeds the supported index range"};
    if (!std::isfinite(box) || box <= 0.f || box / n <= 0.f || !std::isfinite(1.f / (box / n)))
        invalid("2D MAC requires finite positive size and spacing");
}

void free_all(std::initializer_list<void*> ptrs) {
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// Error bits of a step or stage -> the reference's exception (status + message).
void raise_errors(uint32_t bits) {
    if (!bits) return;
    const uint32_t first = bits & 0xffu, second = bits >> 8;
    for (uint32_t b : {first, second}) {
        if (b & kErrAdvectionInput) invalid("Non-finite 2D advection input");
        if (b & kErrBacktrace) invalid("Non-finite 2D backtrace");
        if (b & kErrVelocity) invalid("Non-finite 2D velocity");
        if (b & kErrDivergence) invalid("Non-finite 2D MAC divergence");
        if (b & kErrRhs) invalid("2D pressure RHS overflow");
        if (b & kErrPressure) throw Failure{VGX_EDEVICE, "2D SOR produced non-finite pressure"};
    }
    invalid("2D step failed");
}

vgx_projection_result to_result(const ProjRecord2& r) {
    vgx_projection_result out{};
    out.iterations = r.iterations;
    out.converged = r.converged;
    out.initial_residual = r.initial_residual;
    out.final_residual = r.final_residual;
    out.residual_available = r.residual_available;
    out.divergence_before = r.divergence_before;
    out.divergence_after = r.divergence_after;
    out.removed_rhs_mean = r.removed_rhs_mean;
    out.residual_checks = -1;
    return out;
}

// CpuPressureSolver2D::project's option checks (CpuPressureSolver2D.cpp:11-16).
ProjOpts2 make_options(float dt, float relaxation, int max_iterations, int check_interval, double abs_tol,
                       double rel_tol, int fixed, float h) {
    const float scale = h * h / dt;
    if (!std::isfinite(dt) || dt <= 0.f || !std::isfinite(scale) || scale <= 0.f || !std::isfinite(relaxation) ||
        relaxation <= 0.f || relaxation >= 2.f || max_iterations < 0 || check_interval <= 0 ||
        !std::isfinite(abs_tol) || abs_tol < 0.0 || !std::isfinite(rel_tol) || rel_tol < 0.0)
        invalid("Invalid 2D SOR projection options");
    ProjOpts2 o{};
    o.dt = dt;
    o.relaxation = relaxation;
    o.scale = scale;
    o.h = h;
    o.max_iterations = max_iterations;
    o.check_interval = check_interval;
    o.absolute_tolerance = abs_tol;
    o.relative_tolerance = rel_tol;
    o.fixed_iterations = fixed;
    return o;
}

struct Status {
    Status2* dev = nullptr;
    Status2* host = nullptr;
    void allocate() {
        dev = device_alloc<Status2>(1, "hipMalloc status");
        check(hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(Status2), hipHostMallocDefault), "hipHostMalloc");
        std::memset(host, 0, sizeof(Status2));
        check(hipMemcpy(dev, host, sizeof(Status2), hipMemcpyHostToDevice), "hipMemcpy status");
    }
    void release() {
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        dev = nullptr;
        host = nullptr;
    }
    // Clears the error bits and both projection records before a step.
    void clear(hipStream_t s) {
        check(hipMemsetAsync(&dev->error, 0, sizeof(uint32_t), s), "hipMemsetAsync");
        check(hipMemsetAsync(&dev->proj[0], 0, sizeof(dev->proj), s), "hipMemsetAsync");
    }
    // Reads the record back; throws the reference's exception for raised bits
    // (and clears them on the device so the handle stays usable).
    void finish(hipStream_t s) {
        check(hipGetLastError(), "kernel launch");
        check(hipMemcpyAsync(host, dev, sizeof(Status2), hipMemcpyDeviceToHost, s), "hipMemcpyAsync status");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (host->error) {
            const uint32_t bits = host->error;
            check(hipMemsetAsync(&dev->error, 0, sizeof(uint32_t), s), "hipMemsetAsync");
            check(hipStreamSynchronize(s), "hipStreamSynchronize");
            host->error = 0;
            raise_errors(bits);
        }
    }
};

}  // namespace

// ===================================================== MacGridFluidSolver2D

struct vgx2_sim {
    int n = 0;
    float box = 0.f;
    vgx2_params params{};
    hipStream_t stream = nullptr;
    size_t N = 0, F = 0;
    float *rho = nullptr, *u = nullptr, *v = nullptr;
    float *un = nullptr, *vn = nullptr, *rhon = nullptr, *uh = nullptr, *vh = nullptr, *rhs = nullptr;
    double* p = nullptr;
    Status status;
    vgx_projection_result last{}, mid{};

    float h() const { return box / (float)n; }

    void allocate() {
        N = (size_t)n * n;
        F = (size_t)n * (n + 1);
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        rho = device_alloc<float>(N, "hipMalloc");
        u = device_alloc<float>(F, "hipMalloc");
        v = device_alloc<float>(F, "hipMalloc");
        un = device_alloc<float>(F, "hipMalloc");
        vn = device_alloc<float>(F, "hipMalloc");
        rhon = device_alloc<float>(N, "hipMalloc");
        uh = device_alloc<float>(F, "hipMalloc");
        vh = device_alloc<float>(F, "hipMalloc");
        rhs = device_alloc<float>(N, "hipMalloc");
        p = device_alloc<double>(N, "hipMalloc");
        status.allocate();
        reset();
    }
    void release() {
        free_all({rho, u, v, un, vn, rhon, uh, vh, rhs, p});
        status.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void reset() {
        for (float* f : {rho, rhon, rhs}) check(hipMemsetAsync(f, 0, N * sizeof(float), stream), "hipMemsetAsync");
        for (float* f : {u, v, un, vn, uh, vh})
            check(hipMemsetAsync(f, 0, F * sizeof(float), stream), "hipMemsetAsync");
        check(hipMemsetAsync(p, 0, N * sizeof(double), stream), "hipMemsetAsync");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        last = vgx_projection_result{};
        mid = vgx_projection_result{};
    }

    // MacGridFluidSolver2D::advance's checks (MacGridFluidSolver2D.cpp:22-32).
    void validate(float dt) const {
        const auto& q = params;
        if (!finite(q.emitter_kick_multiplier) || q.emitter_kick_multiplier < 0.f)
            invalid("Invalid emitter kick multiplier");
        if (!finite(dt) || dt <= 0.f || !finite(q.source_strength) || q.source_strength < 0.f ||
            !finite(q.emitter_radius) || q.emitter_radius <= 0.f || q.emitter_radius > 1.f ||
            !finite(q.emitter_center_x) || q.emitter_center_x < 0.f || q.emitter_center_x > 1.f ||
            !finite(q.emitter_center_y) || q.emitter_center_y < 0.f || q.emitter_center_y > 1.f ||
            !finite(q.buoyancy) || !finite(q.smoke_decay) || q.smoke_decay < 0.f || !finite(q.sor_omega) ||
            q.sor_omega <= 0.f || q.sor_omega >= 2.f || q.project_iterations < 0)
            invalid("Invalid 2D simulation time step or parameters");
    }

    void advect_velocity(const float* su, const float* sv, float dt) {
        launch_advect_velocity(stream, u, v, su, sv, un, vn, n, dt / h(), status.dev);
    }
    void advect_density(float dt) {
        launch_advect(stream, kDensity, u, v, rho, rhon, n, dt / h(), status.dev);
        const float decay = std::exp(-dt * params.smoke_decay);
        launch_scale(stream, rhon, rho, N, decay, status.dev);
    }
    void emit() {
        const auto& q = params;
        if (q.source_strength <= 0.f) return;
        launch_emit(stream, rho, v, n, q.emitter_center_x, q.emitter_center_y, q.emitter_radius * q.emitter_radius,
                    q.source_strength, q.emitter_kick_multiplier, status.dev);
    }
    // project(dt) (:94-101): the projection options the solver builds. Option
    // failures (a step so small that h^2/dt overflows) are raised after the
    // stages before them ran, as vapor's exception is.
    bool project(float dt, int slot, std::string& deferred) {
        ProjOpts2 o;
        try {
            o = make_options(dt, params.sor_omega, params.project_iterations, 8, 1e-6, 1e-5, 0, h());
        } catch (const Failure& f) {
            deferred = f.message;
            return false;
        }
        launch_project(stream, u, v, rhs, p, n, o, status.dev, slot);
        return true;
    }

    void advance(float dt) {
        validate(dt);
        status.clear(stream);
        launch_hist_walls(stream, u, v, n);
        mid = vgx_projection_result{};
        std::string deferred;
        bool ok = true;
        if (params.reflection) {
            const float half = .5f * dt;
            advect_density(half);
            advect_velocity(u, v, half);
            launch_commit_velocity(stream, un, vn, u, v, rho, n, half * params.buoyancy, true, status.dev);
            emit();
            ok = project(half, 1, deferred);
            if (ok) {
                launch_hat(stream, u, v, un, vn, uh, vh, n, status.dev);
                advect_density(half);
                advect_velocity(uh, vh, half);
                launch_commit_velocity(stream, un, vn, u, v, rho, n, 0.f, false, status.dev);
                ok = project(half, 0, deferred);
            }
        } else {
            advect_velocity(u, v, dt);
            launch_commit_velocity(stream, un, vn, u, v, rho, n, dt * params.buoyancy, true, status.dev);
            emit();
            ok = project(dt, 0, deferred);
            if (ok) advect_density(dt);
        }
        try {
            status.finish(stream);
        } catch (...) {
            collect(params.reflection != 0);
            throw;
        }
        collect(params.reflection != 0);
        if (!ok) invalid(deferred);
    }
    // Results of the projections that completed (a record is written only
    // when its projection returned).
    void collect(bool reflection) {
        const Status2& s = *status.host;
        if (reflection && s.proj[1].residual_available) mid = to_result(s.proj[1]);
        if (s.proj[0].residual_available) last = to_result(s.proj[0]);
    }
};

extern "C" {

size_t vgx2_struct_size(int which) {
    switch (which) {
        case 0: return sizeof(vgx2_params);
        case 1: return sizeof(vgx2r_params);
        default: return 0;
    }
}

void vgx2_params_default(vgx2_params* p) {
    if (!p) return;
    *p = vgx2_params{};
    p->reflection = 0;
    p->source_strength = 4.f;
    p->emitter_kick_multiplier = 1.f;
    p->emitter_radius = .08f;
    p->emitter_center_x = .5f;
    p->emitter_center_y = .2f;
    p->buoyancy = 3.f;
    p->smoke_decay = .08f;
    p->sor_omega = 1.9f;
    p->project_iterations = 600;
}

void vgx2_params_plume(vgx2_params* p) {
    if (!p) return;
    vgx2_params_default(p);
    p->emitter_radius = .04f;
    p->source_strength = 1.f;
    p->buoyancy = .8f;
    p->smoke_decay = .3007525f;
    p->project_iterations = 120;
}

int vgx2_sim_create(int n, float box, const vgx2_params* params, vgx2_sim** out) {
    return guarded([&] {
        if (!out) invalid("null output handle");
        *out = nullptr;
        check_domain(n, box);
        auto* sim = new vgx2_sim();
        sim->n = n;
        sim->box = box;
        if (params) sim->params = *params;
        else vgx2_params_default(&sim->params);
        try {
            sim->allocate();
        } catch (...) {
            sim->release();
            delete sim;
            throw;
        }
        *out = sim;
    });
}

int vgx2_sim_create_default(vgx2_sim** out) {
    vgx2_params p;
    vgx2_params_plume(&p);
    return vgx2_sim_create(128, 1.f, &p, out);
}

void vgx2_sim_destroy(vgx2_sim* sim) {
    if (!sim) return;
    sim->release();
    delete sim;
}

int vgx2_sim_get_params(const vgx2_sim* sim, vgx2_params* params) {
    return guarded([&] {
        if (!sim || !params) invalid("null argument");
        *params = sim->params;
    });
}

int vgx2_sim_set_params(vgx2_sim* sim, const vgx2_params* params) {
    return guarded([&] {
        if (!sim || !params) invalid("null argument");
        sim->params = *params;
    });
}

int vgx2_sim_resolution(const vgx2_sim* sim) { return sim ? sim->n : 0; }
float vgx2_sim_box_size(const vgx2_sim* sim) { return sim ? sim->box : 0.f; }

int vgx2_sim_advance(vgx2_sim* sim, float dt) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->advance(dt);
    });
}

int vgx2_sim_reset(vgx2_sim* sim) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->reset();
    });
}

int vgx2_sim_set_box(vgx2_sim* sim, float box) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        const float previous = sim->box;
        check_domain(sim->n, box);
        sim->box = box;
        if (box != previous) sim->reset();
    });
}

int vgx2_sim_upload(vgx2_sim* sim, const float* rho, const float* u, const float* v) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        if (rho) check(hipMemcpyAsync(sim->rho, rho, sim->N * 4, hipMemcpyHostToDevice, sim->stream), "upload");
        if (u) check(hipMemcpyAsync(sim->u, u, sim->F * 4, hipMemcpyHostToDevice, sim->stream), "upload");
        if (v) check(hipMemcpyAsync(sim->v, v, sim->F * 4, hipMemcpyHostToDevice, sim->stream), "upload");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx2_sim_download(vgx2_sim* sim, float* rho, float* u, float* v) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        if (rho) check(hipMemcpyAsync(rho, sim->rho, sim->N * 4, hipMemcpyDeviceToHost, sim->stream), "download");
        if (u) check(hipMemcpyAsync(u, sim->u, sim->F * 4, hipMemcpyDeviceToHost, sim->stream), "download");
        if (v) check(hipMemcpyAsync(v, sim->v, sim->F * 4, hipMemcpyDeviceToHost, sim->stream), "download");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx2_sim_last_projection(const vgx2_sim* sim, vgx_projection_result* result) {
    return guarded([&] {
        if (!sim || !result) invalid("null argument");
        *result = sim->last;
    });
}

int vgx2_sim_midpoint_projection(const vgx2_sim* sim, vgx_projection_result* result) {
    return guarded([&] {
        if (!sim || !result) invalid("null argument");
        *result = sim->mid;
    });
}

int vgx2_sim_diagnostics(vgx2_sim* sim, vgx_diagnostics* out) {
    return guarded([&] {
        if (!sim || !out) invalid("null argument");
        launch_diagnostics(sim->stream, sim->u, sim->v, sim->n, sim->h(), false, sim->status.dev);
        sim->status.finish(sim->stream);
        out->kinetic_energy = sim->status.host->kinetic_energy;
        out->max_velocity = sim->status.host->max_velocity;
        out->max_divergence = sim->status.host->max_divergence;
    });
}

int vgx2_sim_device_fields(vgx2_sim* sim, const float** rho, const float** u, const float** v) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        if (rho) *rho = sim->rho;
        if (u) *u = sim->u;
        if (v) *v = sim->v;
    });
}

int vgx2_sim_synchronize(vgx2_sim* sim) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

}  // extern "C"

// ================================================================== stages

namespace {

// Device buffers of the staged entry points, grown on demand, one set per
// host thread (the stages are called from one thread per handle in vapor).
struct Stage2 {
    hipStream_t stream = nullptr;
    size_t cap_faces = 0;
    float *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *rhs = nullptr;
    double* p = nullptr;
    Status status;

    void ensure(size_t faces) {
        if (!stream) {
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
            status.allocate();
        }
        if (faces <= cap_faces) return;
        free_all({a, b, c, d, rhs, p});
        a = device_alloc<float>(faces, "hipMalloc");
        b = device_alloc<float>(faces, "hipMalloc");
        c = device_alloc<float>(faces, "hipMalloc");
        d = device_alloc<float>(faces, "hipMalloc");
        rhs = device_alloc<float>(faces, "hipMalloc");
        p = device_alloc<double>(faces, "hipMalloc");
        cap_faces = faces;
    }
    ~Stage2() {
        free_all({a, b, c, d, rhs, p});
        status.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

Stage2& stage2() {
    static thread_local Stage2 s;
    return s;
}

bool overlaps(const float* a, size_t na, const float* b, size_t nb) {
    const std::less<const float*> less;
    return na && nb && less(a, b + nb) && less(b, a + na);
}

}  // namespace

extern "C" {

int vgx2_backend_advect(int n, float box, int field, float dt, const float* u, const float* v, const float* src,
                        float* out) {
    return guarded([&] {
        check_domain(n, box);
        if (field != VGX2_DENSITY && field != VGX2_VELOCITY_X && field != VGX2_VELOCITY_Y)
            invalid("Unknown 2D MAC field");
        if (!u || !v || !src || !out) invalid("null field pointer");
        const int w = n + (field == VGX2_VELOCITY_X), h = n + (field == VGX2_VELOCITY_Y);
        const size_t F = (size_t)n * (n + 1), count = (size_t)w * h;
        if (!std::isfinite(dt) || dt < 0.f || overlaps(src, count, out, count) || overlaps(u, F, out, count) ||
            overlaps(v, F, out, count))
            invalid("Invalid 2D advection timestep, extent, or overlapping output");
        Stage2& s = stage2();
        s.ensure(F);
        check(hipMemcpyAsync(s.a, u, F * 4, hipMemcpyHostToDevice, s.stream), "upload");
        check(hipMemcpyAsync(s.b, v, F * 4, hipMemcpyHostToDevice, s.stream), "upload");
        check(hipMemcpyAsync(s.c, src, count * 4, hipMemcpyHostToDevice, s.stream), "upload");
        s.status.clear(s.stream);
        launch_advect(s.stream, field, s.a, s.b, s.c, s.d, n, dt / (box / (float)n), s.status.dev);
        s.status.finish(s.stream);
        check(hipMemcpy(out, s.d, count * 4, hipMemcpyDeviceToHost), "download");
    });
}

int vgx2_backend_project(int n, float box, const vgx_projection_options* options, float* u, float* v,
                         double* pressure_out, vgx_projection_result* result) {
    return guarded([&] {
        check_domain(n, box);
        if (!options || !u || !v) invalid("null argument");
        const float h = box / (float)n;
        const auto& q = *options;
        const ProjOpts2 o = make_options(q.dt, q.relaxation, q.max_iterations, q.residual_check_interval,
                                         q.absolute_tolerance, q.relative_tolerance, q.fixed_iterations, h);
        const size_t F = (size_t)n * (n + 1), N = (size_t)n * n;
        Stage2& s = stage2();
        s.ensure(F);
        check(hipMemcpyAsync(s.a, u, F * 4, hipMemcpyHostToDevice, s.stream), "upload");
        check(hipMemcpyAsync(s.b, v, F * 4, hipMemcpyHostToDevice, s.stream), "upload");
        s.status.clear(s.stream);
        launch_project(s.stream, s.a, s.b, s.rhs, s.p, n, o, s.status.dev, 0);
        std::string failure;
        int code = VGX_OK;
        try {
            s.status.finish(s.stream);
        } catch (const Failure& f) {
            code = f.code;
            failure = f.message;
        }
        // The faces hold what vapor's solver leaves in them, also on failure.
        check(hipMemcpy(u, s.a, F * 4, hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(v, s.b, F * 4, hipMemcpyDeviceToHost), "download");
        if (code != VGX_OK) throw Failure{code, failure};
        if (pressure_out) check(hipMemcpy(pressure_out, s.p, N * 8, hipMemcpyDeviceToHost), "download");
        if (result) *result = to_result(s.status.host->proj[0]);
    });
}

}  // extern "C"

// =============================================== ReflectionMacFluidSolver2D

struct vgx2r_sim {
    int n = 0;
    vgx2r_params params{};
    hipStream_t stream = nullptr;
    size_t N = 0, F = 0;
    float *smoke = nullptr, *vx = nullptr, *vy = nullptr, *pressure = nullptr, *divergence = nullptr;
    float *fwd = nullptr, *corr = nullptr, *smoke_next = nullptr;
    float *vxt = nullptr, *vyt = nullptr, *vxh = nullptr, *vyh = nullptr, *vxn = nullptr, *vyn = nullptr;
    int* ranks = nullptr;
    Status status;

    void allocate() {
        N = (size_t)n * n;
        F = (size_t)n * (n + 1);
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (float** f : {&smoke, &pressure, &divergence, &fwd, &corr, &smoke_next}) *f = device_alloc<float>(N, "hipMalloc");
        for (float** f : {&vx, &vy, &vxt, &vyt, &vxh, &vyh, &vxn, &vyn}) *f = device_alloc<float>(F, "hipMalloc");
        ranks = device_alloc<int>(N, "hipMalloc");
        status.allocate();
        zero_fields();
        set_random(0x9E3779B9u);  // the member initialiser (ReflectionMacFluidSolver2D.hpp:72)
    }
    void release() {
        free_all({smoke, vx, vy, pressure, divergence, fwd, corr, smoke_next, vxt, vyt, vxh, vyh, vxn, vyn, ranks});
        status.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void zero_fields() {
        for (float* f : {smoke, pressure, divergence}) check(hipMemsetAsync(f, 0, N * 4, stream), "hipMemsetAsync");
        for (float* f : {vx, vy}) check(hipMemsetAsync(f, 0, F * 4, stream), "hipMemsetAsync");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    void set_random(uint32_t s) {
        check(hipMemcpyAsync(&status.dev->random_state, &s, 4, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void advect_smoke(float dt, float factor) {
        if (params.maccormack_smoke) {
            launch_hist_advect_scalar(stream, smoke, fwd, corr, vx, vy, n, dt, true, factor, nullptr);
        } else {
            launch_hist_advect_scalar(stream, smoke, fwd, corr, vx, vy, n, dt, false, factor, smoke_next);
            std::swap(smoke, smoke_next);
        }
    }
    void emit(float strength) {
        if (strength <= 0.f) return;
        const auto& q = params;
        HistEmit e{};
        e.strength = strength;
        e.radius = q.emitter_radius;
        e.falloff = q.emitter_falloff;
        e.cx = q.emitter_center_x;
        e.cy = q.emitter_center_y;
        e.smoke_scale = q.emitter_smoke_scale;
        e.kick = q.emitter_kick_multiplier;
        e.kick_base = q.emitter_kick_base;
        e.kick_scale = q.emitter_kick_scale;
        e.stir = q.stir_strength;
        const float fn = (float)n;
        e.y_end = std::min((int)(q.emitter_center_y * fn + q.emitter_radius * fn + 2.f), n);
        e.y_begin = std::max(0, (int)(q.emitter_center_y * fn - q.emitter_radius * fn - 1.f));
        launch_hist_emit(stream, smoke, vx, vy, ranks, n, e, status.dev);
    }
    void project(float dt, int iterations) {
        launch_hist_project(stream, vx, vy, pressure, divergence, n, dt, iterations, params.sor_omega);
    }
    // step (ReflectionMacFluidSolver2D.hpp:316-408) for the closed box.
    void step(float dt, float buoyancy, float dissipation, float strength, int iterations) {
        const float half = 0.5f * dt;
        const int half_iters = std::max(1, iterations / 2);
        const float half_decay = std::pow(dissipation, half * 60.0f);
        advect_smoke(half, half_decay);
        launch_hist_advect_velocity(stream, vx, vy, vx, vy, vxt, vyt, n, half);
        launch_hist_forces(stream, vxt, vyt, vx, vy, smoke, n, buoyancy, params.buoyancy_exp, half,
                           params.buoyancy_split);
        emit(strength);
        project(half, half_iters);
        launch_hist_hat(stream, vx, vy, vxt, vyt, vxh, vyh, n);
        advect_smoke(half, half_decay);
        launch_hist_advect_velocity(stream, vxh, vyh, vx, vy, vxn, vyn, n, half);
        std::swap(vx, vxn);
        std::swap(vy, vyn);
        project(half, half_iters);
        check(hipGetLastError(), "kernel launch");
    }
    // advance's checks (:33-42).
    void validate(float dt) const {
        const auto& q = params;
        if (!finite(q.emitter_kick_multiplier) || q.emitter_kick_multiplier < 0.f)
            invalid("Invalid emitter kick multiplier");
        if (!finite(dt) || dt <= 0.f || !finite(q.buoyancy) || !finite(q.smoke_decay) || q.smoke_decay < 0.f ||
            !finite(q.source_strength) || q.source_strength < 0.f || !finite(q.emitter_radius) ||
            q.emitter_radius <= 0.f || q.emitter_radius > 1.f || !finite(q.stir_strength) || q.stir_strength < 0.f ||
            !finite(q.sor_omega) || q.sor_omega <= 0.f || q.sor_omega >= 2.f || q.project_iterations < 0)
            invalid("Invalid 2D reflection step or parameters");
    }
    void advance(float dt) {
        validate(dt);
        launch_hist_walls(stream, vx, vy, n);
        step(dt, params.buoyancy, std::exp(-params.smoke_decay / 60.f), params.source_strength,
             params.project_iterations);
    }
};

extern "C" {

void vgx2r_params_default(vgx2r_params* p) {
    if (!p) return;
    *p = vgx2r_params{};
    p->maccormack_smoke = 1;
    p->source_strength = 1.f;
    p->buoyancy = .8f;
    p->smoke_decay = .3007525f;
    p->project_iterations = 60;
    p->emitter_radius = 0.08f;
    p->emitter_falloff = 3.5f;
    p->emitter_center_x = 0.5f;
    p->emitter_center_y = 0.2f;
    p->emitter_smoke_scale = 3.0f;
    p->emitter_kick_multiplier = 1.f;
    p->emitter_kick_base = 0.6f;
    p->emitter_kick_scale = 1.0f;
    p->emitter_height = 4;
    p->sor_omega = 1.9f;
    p->buoyancy_split = 0.5f;
    p->buoyancy_exp = 0.25f;
    p->stir_strength = 0.f;
    p->random_seed = 0x9E3779B9u;
}

int vgx2r_params_plume(int smoke_method, vgx2r_params* p) {
    return guarded([&] {
        if (!p) invalid("null argument");
        if (smoke_method != VGX_SEMI_LAGRANGIAN && smoke_method != VGX_MACCORMACK)
            invalid("Unknown smoke advection method");
        vgx2r_params_default(p);
        p->maccormack_smoke = smoke_method == VGX_MACCORMACK;
        p->emitter_radius = .04f;
        p->project_iterations = 120;
    });
}

int vgx2r_sim_create(int n, const vgx2r_params* params, vgx2r_sim** out) {
    return guarded([&] {
        if (!out) invalid("null output handle");
        *out = nullptr;
        if (n < 2 || n > 512) invalid("2D reflection resolution must be 2..512");
        auto* sim = new vgx2r_sim();
        sim->n = n;
        if (params) sim->params = *params;
        else vgx2r_params_default(&sim->params);
        try {
            sim->allocate();
        } catch (...) {
            sim->release();
            delete sim;
            throw;
        }
        *out = sim;
    });
}

void vgx2r_sim_destroy(vgx2r_sim* sim) {
    if (!sim) return;
    sim->release();
    delete sim;
}

int vgx2r_sim_get_params(const vgx2r_sim* sim, vgx2r_params* params) {
    return guarded([&] {
        if (!sim || !params) invalid("null argument");
        *params = sim->params;
    });
}

int vgx2r_sim_set_params(vgx2r_sim* sim, const vgx2r_params* params) {
    return guarded([&] {
        if (!sim || !params) invalid("null argument");
        sim->params = *params;
    });
}

int vgx2r_sim_resolution(const vgx2r_sim* sim) { return sim ? sim->n : 0; }

int vgx2r_sim_advance(vgx2r_sim* sim, float dt) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->advance(dt);
    });
}

int vgx2r_sim_step(vgx2r_sim* sim, float dt, float buoyancy, float dissipation, float source_strength,
                   int project_iterations) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->step(dt, buoyancy, dissipation, source_strength, project_iterations);
    });
}

int vgx2r_sim_emit(vgx2r_sim* sim, float strength, float /*dt*/) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->emit(strength);
        check(hipGetLastError(), "kernel launch");
    });
}

int vgx2r_sim_apply_boundary(vgx2r_sim* sim) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        launch_hist_walls(sim->stream, sim->vx, sim->vy, sim->n);
        check(hipGetLastError(), "kernel launch");
    });
}

int vgx2r_sim_advect_smoke(vgx2r_sim* sim, float dt) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->advect_smoke(dt, 1.f);
        check(hipGetLastError(), "kernel launch");
    });
}

int vgx2r_sim_project(vgx2r_sim* sim, float dt, int iterations) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->project(dt, iterations);
        check(hipGetLastError(), "kernel launch");
    });
}

int vgx2r_sim_reset(vgx2r_sim* sim) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->zero_fields();
        sim->set_random(sim->params.random_seed);
    });
}

int vgx2r_sim_set_box(vgx2r_sim* sim, float box) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        if (box != 1.f) invalid("2D reflection requires box_size=1");
    });
}

int vgx2r_sim_upload(vgx2r_sim* sim, const float* smoke, const float* vx, const float* vy, const float* pressure,
                     const float* divergence) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        const hipStream_t s = sim->stream;
        if (smoke) check(hipMemcpyAsync(sim->smoke, smoke, sim->N * 4, hipMemcpyHostToDevice, s), "upload");
        if (vx) check(hipMemcpyAsync(sim->vx, vx, sim->F * 4, hipMemcpyHostToDevice, s), "upload");
        if (vy) check(hipMemcpyAsync(sim->vy, vy, sim->F * 4, hipMemcpyHostToDevice, s), "upload");
        if (pressure) check(hipMemcpyAsync(sim->pressure, pressure, sim->N * 4, hipMemcpyHostToDevice, s), "upload");
        if (divergence)
            check(hipMemcpyAsync(sim->divergence, divergence, sim->N * 4, hipMemcpyHostToDevice, s), "upload");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

int vgx2r_sim_download(vgx2r_sim* sim, float* smoke, float* vx, float* vy, float* pressure, float* divergence) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        const hipStream_t s = sim->stream;
        if (smoke) check(hipMemcpyAsync(smoke, sim->smoke, sim->N * 4, hipMemcpyDeviceToHost, s), "download");
        if (vx) check(hipMemcpyAsync(vx, sim->vx, sim->F * 4, hipMemcpyDeviceToHost, s), "download");
        if (vy) check(hipMemcpyAsync(vy, sim->vy, sim->F * 4, hipMemcpyDeviceToHost, s), "download");
        if (pressure) check(hipMemcpyAsync(pressure, sim->pressure, sim->N * 4, hipMemcpyDeviceToHost, s), "download");
        if (divergence)
            check(hipMemcpyAsync(divergence, sim->divergence, sim->N * 4, hipMemcpyDeviceToHost, s), "download");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

int vgx2r_sim_random_state(vgx2r_sim* sim, uint32_t* state) {
    return guarded([&] {
        if (!sim || !state) invalid("null argument");
        check(hipMemcpyAsync(state, &sim->status.dev->random_state, 4, hipMemcpyDeviceToHost, sim->stream),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx2r_sim_set_random_state(vgx2r_sim* sim, uint32_t state) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        sim->set_random(state);
    });
}

int vgx2r_sim_diagnostics(vgx2r_sim* sim, vgx_diagnostics* out) {
    return guarded([&] {
        if (!sim || !out) invalid("null argument");
        launch_diagnostics(sim->stream, sim->vx, sim->vy, sim->n, 1.0f / (float)sim->n, true, sim->status.dev);
        sim->status.finish(sim->stream);
        out->kinetic_energy = sim->status.host->kinetic_energy;
        out->max_velocity = sim->status.host->max_velocity;
        out->max_divergence = sim->status.host->max_divergence;
    });
}

int vgx2r_sim_device_fields(vgx2r_sim* sim, const float** smoke, const float** vx, const float** vy) {
    return guarded([&] {
        if (!sim) invalid("null handle");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
        if (smoke) *smoke = sim->smoke;
        if (vx) *vx = sim->vx;
        if (vy) *vy = sim->vy;
    });
}

}  // extern "C"
