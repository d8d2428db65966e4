// simplicial_host_abi.cpp - the simplicial backend's C ABI on the CPU (test
// infrastructure).
//
// Implements the vsx_* entry points of include/vapor_simplicial3d.h on the
// host with the device code's own functions: the mesh goes through
// vsx::prepare (the same checks and rearrangement the library uploads), each
// traced point through box_trace / box_density_point / general_trace_point /
// general_density_point / loop_circulation, and the CG through the operator's
// row functions in the kernel's order (reference-order sums, the IC factor's
// level schedule). Linked with vapor's solver and the C++ binding in place of
// libvapor_mac3d.so, it lets the binding and the device arithmetic be checked
// bit for bit against vapor's CPU path, under AddressSanitizer, before
// anything runs on a GPU.
#include <cmath>
#include <string>
#include <vector>

#include "../../vapor_amd/csrc/vsx_prepare.hpp"

using namespace vgx_host;

namespace vgx_host {
std::string& error_slot() {
    thread_local std::string slot;
    return slot;
}
}  // namespace vgx_host

struct vsx_backend {
    vsx::Prepared P;
    vsx::View V{};
    vsx::CgOperator A{};
    int recovery = VSX_RECOVER_FAST;
    std::string device = "CPU (device code on the host)";
};

namespace {

int g_default_recovery = VSX_RECOVER_FAST;

void check_dt(double dt) {
    if (!std::isfinite(dt) || dt < 0) invalid("Timestep must be finite and nonnegative");
}

void raise(int bits) {
    if (bits & vsx::kErrWeights) throw Failure{VGX_EDEVICE, "Invalid generalized barycentric weights"};
    if (bits & vsx::kErrSubdivision) throw Failure{VGX_EDEVICE, "Point outside circumcentric subdivision"};
}

int nearest(const vsx::View& V, vsx::P3 p) {
    double best = 1.7976931348623157e308;
    int site = 0;
    for (int i = 0; i < V.nv; ++i) vsx::nearer(p, vsx::ld3(V.vert, i), i, best, site);
    return site;
}

struct Cg {
    const vsx::CgOperator& A;
    std::vector<double> c, div;
    Cg(const vsx::CgOperator& a) : A(a), c(a.nf), div(a.nv) {}
    void apply(const std::vector<double>& u, std::vector<double>& y) {
        const vsx::Entries ue{u.data()};
        for (int f = 0; f < A.nf; ++f) c[f] = vsx::face_value(A, ue, f);
        for (int v = 0; v < A.nv; ++v) div[v] = vsx::vertex_divergence(A, ue, v);
        for (int e = 0; e < A.ne; ++e) y[e] = vsx::edge_value(A, u[e], c.data(), div.data(), e);
    }
    double inner(const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0;
        for (int e = 0; e < A.ne; ++e) s = s + a[e] * b[e];
        return s;
    }
    double norm(const std::vector<double>& a) {
        double s = 0;
        for (int e = 0; e < A.ne; ++e) {
            const double v = std::fabs(a[e]);
            s = (s < v) ? v : s;
        }
        return s;
    }
    void precondition(bool ic, const std::vector<double>& r, std::vector<double>& z) {
        if (!ic) {
            for (int e = 0; e < A.ne; ++e) z[e] = r[e] / A.diagonal[e];
            return;
        }
        for (int level = 0; level < A.forward_levels; ++level)
            for (int k = A.forward_off[level]; k < A.forward_off[level + 1]; ++k)
                z[A.forward_rows[k]] = vsx::ic_forward_row(A, r.data(), z.data(), A.forward_rows[k]);
        for (int level = 0; level < A.backward_levels; ++level)
            for (int k = A.backward_off[level]; k < A.backward_off[level + 1]; ++k)
                z[A.backward_rows[k]] = vsx::ic_backward_row(A, z.data(), A.backward_rows[k]);
    }
};

}  // namespace

extern "C" {

const char* vgx_last_error(void) { return error_slot().c_str(); }

int vsx_backend_create(const vsx_mesh* mesh, vsx_backend** out) {
    return guarded([&] {
        if (!out || !mesh) invalid("null argument");
        *out = nullptr;
        auto* b = new vsx_backend();
        try {
            b->P = vsx::prepare(*mesh);
        } catch (...) {
            delete b;
            throw;
        }
        b->recovery = g_default_recovery;
        vsx::make_views(b->P, [](const auto& v) { return v.empty() ? nullptr : v.data(); }, b->V, b->A);
        *out = b;
    });
}

void vsx_backend_destroy(vsx_backend* backend) { delete backend; }
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
double vsx_last_kernel_ms(void) { return 0.0; }

int vsx_advect_circulation(vsx_backend* b, const double* dual_velocity, const double* node_samples, double dt,
                           double* omega) {
    return guarded([&] {
        if (!b || !omega || !dual_velocity) invalid("null argument");
        if (!b->V.box && !node_samples) invalid("general domains need the subdivision samples");
        check_dt(dt);
        const vsx::View& V = b->V;
        std::vector<double> traced(3L * V.nd), velocity(3L * V.nd);
        int err = 0;
        auto near = [&](vsx::P3 p) { return nearest(V, p); };
        for (int i = 0; i < V.nd; ++i) {
            vsx::P3 q, v;
            if (V.box) {
                q = vsx::box_trace(V, near, vsx::ld3(V.dual_pos, i), dual_velocity, dt, true, err);
                v = vsx::warren(V, near(q), q, dual_velocity, err);
            } else {
                vsx::general_trace_point(V, i, dual_velocity, node_samples, dt, q, v, err);
            }
            vsx::st3(traced.data(), i, q);
            vsx::st3(velocity.data(), i, v);
        }
        raise(err);
        const auto& P = b->P;
        for (int e = 0; e < V.ne; ++e)
            omega[e] = vsx::loop_circulation(P.loop_edges.data(), P.loop_signs.data(), P.loop_off[e], P.loop_off[e + 1],
                                             P.dual_edges.data(), traced.data(), velocity.data());
    });
}

int vsx_advect_density(vsx_backend* b, const double* dual_velocity, const double* node_samples,
                       const float* vertex_density, double dt, double decay, float* out) {
    return guarded([&] {
        if (!b || !out || !vertex_density || !dual_velocity) invalid("null argument");
        if (!b->V.box && !node_samples) invalid("general domains need the subdivision samples");
        check_dt(dt);
        if (!std::isfinite(decay)) invalid("non-finite decay");
        const vsx::View& V = b->V;
        int err = 0;
        auto near = [&](vsx::P3 p) { return nearest(V, p); };
        std::vector<float> next(V.nv);
        for (int i = 0; i < V.nv; ++i)
            next[i] = V.box ? vsx::box_density_point(V, near, i, dual_velocity, vertex_density, dt, decay, true, err)
                            : vsx::general_density_point(V, i, node_samples, vertex_density, dt, decay, err);
        raise(err);
        std::copy(next.begin(), next.end(), out);
    });
}

int vsx_recover(vsx_backend* b, const double* rhs_in, const double* phi, int32_t max_iterations,
                double absolute_tolerance, double relative_tolerance, double* x_out, vsx_solve_result* result) {
    return guarded([&] {
        if (!b || !rhs_in || !x_out || !result) invalid("null argument");
        if (max_iterations < 0 || !std::isfinite(absolute_tolerance) || absolute_tolerance < 0 ||
            !std::isfinite(relative_tolerance) || relative_tolerance < 0)
            invalid("Invalid CG parameters");
        const bool warm = !b->V.box;
        if (warm && !phi) invalid("general domains start from phi");
        const vsx::CgOperator& A = b->A;
        const int ne = A.ne;
        const bool ic = b->recovery == VSX_RECOVER_REFERENCE && A.forward_levels > 0;
        Cg cg(A);
        std::vector<double> rhs(rhs_in, rhs_in + ne), x(ne, 0), r(ne), z(ne), p(ne), ap(ne);
        if (warm) {
            x.assign(phi, phi + ne);
            cg.apply(x, ap);
            for (int e = 0; e < ne; ++e) r[e] = rhs[e] - ap[e];
        } else {
            r = rhs;
        }
        for (int e = 0; e < ne; ++e) {
            if (!std::isfinite(r[e])) throw Failure{VGX_EDEVICE, "Non-finite CG RHS"};
            if (A.edge_boundary[e]) r[e] = rhs[e] = 0;
        }
        const double initial = cg.norm(r);
        const double target = vsx::smax(absolute_tolerance, relative_tolerance * (warm ? cg.norm(rhs) : initial));
        double final_residual = initial;
        int iterations = 0;
        if (initial <= target) {
            std::copy(x.begin(), x.end(), x_out);
            *result = {0, 1, initial, initial};
            return;
        }
        cg.precondition(ic, r, z);
        p = z;
        double rz = cg.inner(r, z);
        for (int it = 0; it < max_iterations; ++it) {
            cg.apply(p, ap);
            const double pap = cg.inner(p, ap);
            if (!std::isfinite(pap) || pap <= 0) throw Failure{VGX_EDEVICE, "CG lost positive definiteness"};
            const double alpha = rz / pap;
            for (int e = 0; e < ne; ++e) {
                x[e] += alpha * p[e];
                r[e] -= alpha * ap[e];
            }
            iterations = it + 1;
            final_residual = cg.norm(r);
            if (final_residual <= target) break;
            cg.precondition(ic, r, z);
            const double next = cg.inner(r, z), beta = next / rz;
            for (int e = 0; e < ne; ++e) p[e] = z[e] + beta * p[e];
            rz = next;
        }
        cg.apply(x, ap);
        for (int e = 0; e < ne; ++e) r[e] = A.edge_boundary[e] ? 0 : rhs[e] - ap[e];
        final_residual = cg.norm(r);
        std::copy(x.begin(), x.end(), x_out);
        *result = {iterations, final_residual <= target ? 1 : 0, initial, final_residual};
    });
}

}  // extern "C"
