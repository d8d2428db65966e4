// vsx_api.hip - C ABI of the simplicial 3D backend (include/vapor_simplicial3d.h).
//
// vsx_backend_create checks and rearranges the mesh (vsx_prepare.hpp) and
// uploads everything once. The stage calls upload
// the state that changes per step, run the kernels on the backend's stream,
// download the result and turn the kernels' error bits into the reference's
// exceptions (as status codes and messages).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/vapor_simplicial3d.h"
#include "vgx_host.hpp"
#include "vsx_kernels.hpp"
#include "vsx_prepare.hpp"

using namespace vgx_host;

namespace {
int g_default_recovery = VSX_RECOVER_FAST;
thread_local double t_last_ms = 0;
}  // namespace

struct vsx_backend {
    vsx::View V{};
    vsx::CgOperator A{};
    vsx::CgVectors W{};
    std::vector<void*> allocations;
    int nv = 0, ne = 0, nd = 0, nn = 0;
    bool box = false;
    const int* loop_off = nullptr;
    const int* loop_edges = nullptr;
    const int* loop_signs = nullptr;
    const int* dual_edges = nullptr;
    double* velocity = nullptr;
    double* samples = nullptr;
    float* density = nullptr;
    float* density_out = nullptr;
    double* traced = nullptr;
    double* traced_velocity = nullptr;
    double* omega = nullptr;
    int* error = nullptr;
    vsx::CgStatus* status = nullptr;
    vsx::GridPartials partials{};
    int grid_capacity = 0;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    int recovery = VSX_RECOVER_FAST;
    std::string device;

    template <class T> T* alloc(size_t count, const char* what) {
        T* p = device_alloc<T>(count, what);
        allocations.push_back(p);
        return p;
    }
    // A device copy of `host`; an empty vector becomes a null pointer.
    template <class T> const T* upload(const std::vector<T>& host, const char* what) {
        if (host.empty()) return nullptr;
        T* p = alloc<T>(host.size(), what);
        check(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), what);
        return p;
    }
    void release() {
        for (void* p : allocations) (void)hipFree(p);
        allocations.clear();
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
        if (stream) (void)hipStreamDestroy(stream);
        start = stop = nullptr;
        stream = nullptr;
    }
};

namespace {

void build(vsx_backend* b, const vsx_mesh& m) {
    const vsx::Prepared P = vsx::prepare(m);
    b->nv = P.nv;
    b->ne = P.ne;
    b->nd = P.nd;
    b->nn = P.nn;
    b->box = P.box != 0;
    auto map = [&](const auto& v) { return b->upload(v, "simplicial mesh"); };
    vsx::make_views(P, map, b->V, b->A);
    b->dual_edges = map(P.dual_edges);
    b->loop_off = map(P.loop_off);
    b->loop_edges = map(P.loop_edges);
    b->loop_signs = map(P.loop_signs);

    // Per-call state.
    const long nv = P.nv, ne = P.ne, nd = P.nd, nf = P.nf;
    b->velocity = b->alloc<double>(3 * nd, "dual velocity");
    if (!b->box) b->samples = b->alloc<double>(3L * b->nn, "node samples");
    b->density = b->alloc<float>(nv, "vertex density");
    b->density_out = b->alloc<float>(nv, "advected density");
    b->traced = b->alloc<double>(3 * nd, "traced points");
    b->traced_velocity = b->alloc<double>(3 * nd, "traced velocities");
    b->omega = b->alloc<double>(ne, "circulation");
    b->error = b->alloc<int>(1, "error flags");
    b->status = b->alloc<vsx::CgStatus>(1, "solve status");
    vsx::CgVectors& W = b->W;
    W.x = b->alloc<double>(ne, "cg x");
    W.r = b->alloc<double>(ne, "cg r");
    W.z = b->alloc<double>(ne, "cg z");
    W.p = b->alloc<double>(ne, "cg p");
    W.ap = b->alloc<double>(ne, "cg Ap");
    W.rhs = b->alloc<double>(ne, "cg rhs");
    W.phi = b->alloc<double>(ne, "cg phi");
    W.prod = b->alloc<double>(ne, "cg products");
    W.c = b->alloc<double>(nf, "cg face values");
    W.div = b->alloc<double>(nv, "cg divergence");

    check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
    check(hipEventCreate(&b->start), "event");
    check(hipEventCreate(&b->stop), "event");
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    b->device = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    b->grid_capacity = vsx::cg_grid_capacity(dev);
    if (b->grid_capacity < 1) throw Failure{VGX_EDEVICE, "the device cannot run the cooperative CG kernel"};
    const size_t g = (size_t)b->grid_capacity;
    b->partials = {b->alloc<double>(g, "cg partials"), b->alloc<double>(g, "cg partials"),
                   b->alloc<double>(g, "cg partials"), b->alloc<double>(g, "cg partials")};
}

void check_dt(double dt) {
    if (!std::isfinite(dt) || dt < 0) invalid("Timestep must be finite and nonnegative");
}

// Download the error bits (after the stream's work) and raise the
// reference's exception for the first one set.
void raise_errors(vsx_backend* b) {
    int bits = 0;
    check(hipMemcpyAsync(&bits, b->error, sizeof(int), hipMemcpyDeviceToHost, b->stream), "error flags");
    check(hipStreamSynchronize(b->stream), "stage");
    float ms = 0;
    check(hipEventElapsedTime(&ms, b->start, b->stop), "event time");
    t_last_ms = ms;
    if (bits & vsx::kErrWeights) throw Failure{VGX_EDEVICE, "Invalid generalized barycentric weights"};
    if (bits & vsx::kErrSubdivision) throw Failure{VGX_EDEVICE, "Point outside circumcentric subdivision"};
}

void upload_state(vsx_backend* b, const double* dual_velocity, const double* node_samples) {
    if (!dual_velocity) invalid("null dual velocity");
    check(hipMemcpyAsync(b->velocity, dual_velocity, 3 * sizeof(double) * b->nd, hipMemcpyHostToDevice, b->stream),
          "dual velocity");
    if (!b->box) {
        if (!node_samples) invalid("general domains need the subdivision samples");
        check(hipMemcpyAsync(b->samples, node_samples, 3 * sizeof(double) * b->nn, hipMemcpyHostToDevice, b->stream),
              "node samples");
    }
    check(hipMemsetAsync(b->error, 0, sizeof(int), b->stream), "error flags");
}

}  // namespace

extern "C" {

int vsx_backend_create(const vsx_mesh* mesh, vsx_backend** out) {
    return guarded([&] {
        if (!out || !mesh) invalid("null argument");
        *out = nullptr;
        auto* b = new vsx_backend();
        b->recovery = g_default_recovery;
        try {
            build(b, *mesh);
        } catch (...) {
            b->release();
            delete b;
            throw;
        }
        *out = b;
    });
}

void vsx_backend_destroy(vsx_backend* backend) {
    if (!backend) return;
    backend->release();
    delete backend;
}

const char* vsx_backend_device(const vsx_backend* backend) { return backend ? backend->device.c_str() : ""; }

int vsx_backend_set_recovery(vsx_backend* backend, int order) {
    return guarded([&] {
        if (!backend) invalid("null backend");
        if (order != VSX_RECOVER_FAST && order != VSX_RECOVER_REFERENCE) invalid("unknown recovery order");
        backend->recovery = order;
    });
}

int vsx_backend_recovery(const vsx_backend* backend) { return backend ? backend->recovery : -1; }

int vsx_set_default_recovery(int order) {
    return guarded([&] {
        if (order != VSX_RECOVER_FAST && order != VSX_RECOVER_REFERENCE) invalid("unknown recovery order");
        g_default_recovery = order;
    });
}

int vsx_default_recovery(void) { return g_default_recovery; }

double vsx_last_kernel_ms(void) { return t_last_ms; }

int vsx_advect_circulation(vsx_backend* b, const double* dual_velocity, const double* node_samples, double dt,
                           double* omega) {
    return guarded([&] {
        if (!b || !omega) invalid("null argument");
        check_dt(dt);
        upload_state(b, dual_velocity, node_samples);
        check(hipEventRecord(b->start, b->stream), "event");
        vsx::launch_circulation(b->V, b->velocity, b->samples, dt, b->traced, b->traced_velocity, b->loop_off,
                                b->loop_edges, b->loop_signs, b->dual_edges, b->omega, b->error, b->stream);
        check(hipGetLastError(), "circulation kernels");
        check(hipEventRecord(b->stop, b->stream), "event");
        check(hipMemcpyAsync(omega, b->omega, sizeof(double) * b->ne, hipMemcpyDeviceToHost, b->stream), "omega");
        raise_errors(b);
    });
}

int vsx_advect_density(vsx_backend* b, const double* dual_velocity, const double* node_samples,
                       const float* vertex_density, double dt, double decay, float* out) {
    return guarded([&] {
        if (!b || !out || !vertex_density) invalid("null argument");
        check_dt(dt);
        if (!std::isfinite(decay)) invalid("non-finite decay");
        upload_state(b, dual_velocity, node_samples);
        check(hipMemcpyAsync(b->density, vertex_density, sizeof(float) * b->nv, hipMemcpyHostToDevice, b->stream),
              "vertex density");
        check(hipEventRecord(b->start, b->stream), "event");
        vsx::launch_density(b->V, b->velocity, b->samples, b->density, dt, decay, b->density_out, b->error, b->stream);
        check(hipGetLastError(), "density kernel");
        check(hipEventRecord(b->stop, b->stream), "event");
        check(hipMemcpyAsync(out, b->density_out, sizeof(float) * b->nv, hipMemcpyDeviceToHost, b->stream), "density");
        raise_errors(b);
    });
}

int vsx_recover(vsx_backend* b, const double* rhs, const double* phi, int32_t max_iterations,
                double absolute_tolerance, double relative_tolerance, double* x, vsx_solve_result* result) {
    return guarded([&] {
        if (!b || !rhs || !x || !result) invalid("null argument");
        // SimplicialFluidSolver3D.cpp:136-139.
        if (max_iterations < 0 || !std::isfinite(absolute_tolerance) || absolute_tolerance < 0 ||
            !std::isfinite(relative_tolerance) || relative_tolerance < 0)
            invalid("Invalid CG parameters");
        if (!b->box && !phi) invalid("general domains start from phi");
        check(hipMemcpyAsync(b->W.rhs, rhs, sizeof(double) * b->ne, hipMemcpyHostToDevice, b->stream), "rhs");
        if (!b->box)
            check(hipMemcpyAsync(b->W.phi, phi, sizeof(double) * b->ne, hipMemcpyHostToDevice, b->stream), "phi");
        vsx::CgControl control{max_iterations, b->box ? 0 : 1, absolute_tolerance, relative_tolerance};
        check(hipEventRecord(b->start, b->stream), "event");
        if (b->recovery == VSX_RECOVER_REFERENCE) {
            vsx::launch_cg(b->A, b->W, control, b->status, b->stream);
            check(hipGetLastError(), "cg kernel");
        } else {
            const int groups = vsx::cg_grid_groups(b->A, b->grid_capacity);
            check(vsx::launch_cg_grid(b->A, b->W, control, b->partials, groups, b->status, b->stream), "cg kernel");
        }
        check(hipEventRecord(b->stop, b->stream), "event");
        vsx::CgStatus status{};
        check(hipMemcpyAsync(&status, b->status, sizeof(status), hipMemcpyDeviceToHost, b->stream), "solve status");
        check(hipStreamSynchronize(b->stream), "cg");
        float ms = 0;
        check(hipEventElapsedTime(&ms, b->start, b->stop), "event time");
        t_last_ms = ms;
        if (status.error & vsx::kErrNonFiniteRhs) throw Failure{VGX_EDEVICE, "Non-finite CG RHS"};
        if (status.error & vsx::kErrDefinite) throw Failure{VGX_EDEVICE, "CG lost positive definiteness"};
        check(hipMemcpy(x, b->W.x, sizeof(double) * b->ne, hipMemcpyDeviceToHost), "x");
        result->iterations = status.iterations;
        result->converged = status.converged;
        result->initial_residual = status.initial;
        result->final_residual = status.final;
    });
}

}  // extern "C"
