// vgx2_kernels.hpp - kernel interfaces of the MAC 2D solvers (vgx2_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgx2 {

// Error bits a kernel raises on the device; later kernels of the same step see
// them and return without writing, so the state stays where the reference's
// exception leaves it. The host turns the bits into a status and message.
enum : uint32_t {
    kErrAdvectionInput = 1u,   // SemiLagrangian2D.cpp:36-37 "Non-finite 2D advection input"
    kErrBacktrace = 2u,        // SemiLagrangian2D.cpp:14 "Non-finite 2D backtrace"
    kErrVelocity = 4u,         // CpuPressureSolver2D.cpp:17-18 "Non-finite 2D velocity"
    kErrDivergence = 8u,       // MacGridOperators2D.cpp:10 "Non-finite 2D MAC divergence"
    kErrRhs = 16u,             // CpuPressureSolver2D.cpp:28 "2D pressure RHS overflow"
    kErrPressure = 32u,        // CpuPressureSolver2D.cpp:73 "2D SOR produced non-finite pressure"
};

// ProjectionResult + LinearSolveResult as the projection kernel writes them.
struct ProjRecord2 {
    int32_t iterations;
    int32_t converged;
    double initial_residual;
    double final_residual;
    int32_t residual_available;
    float divergence_before;
    float divergence_after;
    int32_t pad;
    double removed_rhs_mean;
};

// Device-side status of one handle: error bits, the two projection records
// (0 = last, 1 = midpoint), the historical solver's LCG state and the
// diagnostics.
struct Status2 {
    uint32_t error;
    uint32_t random_state;
    ProjRecord2 proj[2];
    float kinetic_energy;
    float max_velocity;
    float max_divergence;
    uint32_t pad;
};

// ProjectionOptions / LinearSolveOptions of one projection.
struct ProjOpts2 {
    float dt;
    float relaxation;
    float scale;       // h^2 / dt in float, CpuPressureSolver2D.cpp:11
    float h;
    int32_t max_iterations;
    int32_t check_interval;
    double absolute_tolerance;
    double relative_tolerance;
    int32_t fixed_iterations;
    int32_t pad;
};

// ReflectionMacFluidSolver2D::emit arguments (ReflectionMacFluidSolver2D.hpp:291-313).
struct HistEmit {
    float strength;
    float radius, falloff, cx, cy, smoke_scale, kick, kick_base, kick_scale, stir;
    int32_t y_begin, y_end;  // scanned rows, computed on the host with the reference's float expressions
};

// ---------------- MacGridFluidSolver2D ----------------
void launch_advect(hipStream_t s, int field, const float* u, const float* v, const float* src, float* out, int n,
                   float dt_over_h, Status2* st);
void launch_advect_velocity(hipStream_t s, const float* u, const float* v, const float* su, const float* sv,
                            float* out_u, float* out_v, int n, float dt_over_h, Status2* st);
void launch_commit_velocity(hipStream_t s, const float* un, const float* vn, float* u, float* v, const float* rho,
                            int n, float dt_buoyancy, bool buoyancy, Status2* st);
void launch_emit(hipStream_t s, float* rho, float* v, int n, float cx, float cy, float r2, float strength,
                 float kick, Status2* st);
void launch_project(hipStream_t s, float* u, float* v, float* rhs, double* p, int n, const ProjOpts2& o,
                    Status2* st, int slot);
void launch_scale(hipStream_t s, const float* src, float* dst, size_t count, float factor, Status2* st);
void launch_hat(hipStream_t s, const float* u, const float* v, const float* un, const float* vn, float* uh,
                float* vh, int n, Status2* st);
void launch_diagnostics(hipStream_t s, const float* u, const float* v, int n, float h, bool historical,
                        Status2* st);

// ---------------- ReflectionMacFluidSolver2D ----------------
void launch_hist_walls(hipStream_t s, float* vx, float* vy, int n);
void launch_hist_advect_scalar(hipStream_t s, float* smoke, float* fwd, float* corr, const float* vx,
                               const float* vy, int n, float dt, bool maccormack, float decay, float* out);
void launch_hist_advect_velocity(hipStream_t s, const float* srcx, const float* srcy, const float* fu,
                                 const float* fv, float* outx, float* outy, int n, float dt);
void launch_hist_forces(hipStream_t s, const float* vxt, const float* vyt, float* vx, float* vy,
                        const float* smoke, int n, float buoyancy, float exponent, float half_dt, float split);
void launch_hist_emit(hipStream_t s, float* smoke, float* vx, float* vy, int* ranks, int n, const HistEmit& e,
                      Status2* st);
void launch_hist_project(hipStream_t s, float* vx, float* vy, float* p, float* div, int n, float dt,
                         int iterations, float omega);
void launch_hist_hat(hipStream_t s, const float* mid_x, const float* mid_y, const float* tx, const float* ty,
                     float* hx, float* hy, int n);

}  // namespace vgx2
