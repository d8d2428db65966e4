// vsx_kernels.hpp - launch interface of the simplicial 3D backend's kernels.
#pragma once

#include "vsx_device.hpp"

namespace vsx {

constexpr int kTraceBlock = 64;  // one wave per workgroup: small meshes still spread over the CUs
constexpr int kCgBlock = 1024;   // the whole CG solve runs in one workgroup

struct CgVectors {
    double *x, *r, *z, *p, *ap, *rhs, *phi, *prod;
    double *c, *div;
};

struct CgControl {
    int max_iterations;
    int warm_start;  // general domains start from phi
    double absolute_tolerance, relative_tolerance;
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
void launch_cg(const CgOperator& A, const CgVectors& W, const CgControl& control, bool reference, CgStatus* status,
               hipStream_t stream);

}  // namespace vsx
