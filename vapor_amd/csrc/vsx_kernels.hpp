// vsx_kernels.hpp - launch interface of the simplicial 3D backend's kernels.
#pragma once

#include "vsx_device.hpp"

namespace vsx {

constexpr int kTraceBlock = 256; // box traces: G lanes per point, vertex tiles of 256 in LDS
constexpr int kGeneralBlock = 64; // general-domain traces: one lane per point, one wave per workgroup
constexpr int kCgBlock = 1024;   // the reference-order CG runs in one workgroup
constexpr int kGridBlock = 1024; // the fast CG: one cooperative grid of these, about one item per lane

struct CgVectors {
    double *x, *r, *z, *p, *ap, *rhs, *phi, *prod;
    double *c, *div;
};

struct CgControl {
    int max_iterations;
    int warm_start;  // general domains start from phi
    double absolute_tolerance, relative_tolerance;
};

// One partial per workgroup for each of the fast CG's reductions.
struct GridPartials {
    double *a, *b, *c, *d;
};

struct CgStatus {
    int error, iterations, converged, pad;
    double initial, final;
};

void launch_circulation(const View& V, const double* velocity, const double* samples, double dt, double* traced,
                        double* traced_velocity, const int* loop_off, const int* loop_edges, const int* loop_signs,
                        const int* dual_edges, double* omega, int* error, hipStream_t stream);
void launch_density(const View& V, const double* velocity, const double* samples, const float* density, double dt,
                    double decay, float* out, int* error, hipStream_t stream);
// The reference-order CG (one workgroup; the IC(0) factor where there is one).
void launch_cg(const CgOperator& A, const CgVectors& W, const CgControl& control, CgStatus* status,
               hipStream_t stream);
// The fast CG over `groups` co-resident workgroups (at most max_groups, at
// most cg_grid_capacity(device); 0 if the device cannot run it).
int cg_grid_capacity(int device);
int cg_grid_groups(const CgOperator& A, int max_groups);
hipError_t launch_cg_grid(const CgOperator& A, const CgVectors& W, const CgControl& control, const GridPartials& G,
                          int groups, CgStatus* status, hipStream_t stream);

}  // namespace vsx
