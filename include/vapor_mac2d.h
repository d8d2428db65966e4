/*
 * vapor_mac2d.h - C ABI of the MI355X (gfx950) MAC 2D smoke solvers.
 *
 * The drop-in boundary for vapor's two 2D MAC solvers. Plain pointers and
 * sizes, status codes and vgx_last_error() as in vapor_mac3d.h, whose
 * ProjectionOptions / ProjectionResult / FluidDiagnostics mirrors
 * (vgx_projection_options, vgx_projection_result, vgx_diagnostics) these
 * entry points share: vapor keeps those types in src/common for every
 * discretisation.
 *
 * Reference interfaces replaced (paths relative to the vapor source tree):
 *   MacGridParameters2D ........ src/solvers/mac2d/MacGridFluidSolver2D.hpp:6-16
 *   MacGridFluidSolver2D ....... MacGridFluidSolver2D.hpp:19-51
 *   MacGridDomain2D / State2D .. MacGridDomain2D.hpp:7-24, MacGridState2D.hpp:8-31
 *   SemiLagrangian2D ........... SemiLagrangian2D.hpp:4-20
 *   CpuPressureSolver2D ........ CpuPressureSolver2D.hpp:8-18
 *   ReflectionMacFluidSolver2D . ReflectionMacFluidSolver2D.hpp:19-425
 *
 * Field layouts are the reference's (MacGridDomain2D.hpp:17-19), FP32:
 *   density  rho[x + n*y]            n*n values
 *   u        u[x + (n+1)*y]          (n+1)*n values
 *   v        v[x + n*y], y in [0,n]  n*(n+1) values
 *   pressure p[x + n*y]              n*n values (FP64 for MacGridFluidSolver2D,
 *                                    FP32 for ReflectionMacFluidSolver2D)
 *
 * Transcendentals: vapor evaluates e^v (emitter weight) and x^(1/4) (the
 * buoyancy's nonlinear smoke term) with the C library's expf / powf. Here they
 * are fixed sequences of IEEE double operations rounded once to float (within
 * one unit in the last place of the C library, the same bits on every host
 * and device). Every other operation rounds as vapor's does.
 */
#ifndef VAPOR_MAC2D_H
#define VAPOR_MAC2D_H

#include "vapor_mac3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MacField2D (SemiLagrangian2D.hpp:4). */
#define VGX2_DENSITY 0
#define VGX2_VELOCITY_X 1
#define VGX2_VELOCITY_Y 2

/* Mirror of MacGridParameters2D (MacGridFluidSolver2D.hpp:6-16). */
typedef struct vgx2_params {
    int32_t reflection;            /* 0 */
    float source_strength;         /* 4 */
    float emitter_kick_multiplier; /* 1 */
    float emitter_radius;          /* 0.08, fraction of the box */
    float emitter_center_x;        /* 0.5 */
    float emitter_center_y;        /* 0.2 */
    float buoyancy;                /* 3 */
    float smoke_decay;             /* 0.08 */
    float sor_omega;               /* 1.9 */
    int32_t project_iterations;    /* 600 */
} vgx2_params;

/* sizeof of the 2D structs, so bindings can check their layouts:
 * 0 vgx2_params, 1 vgx2r_params. */
size_t vgx2_struct_size(int which);

/* MacGridParameters2D{} defaults. */
void vgx2_params_default(vgx2_params* params);
/* The parameters MacGridFluidSolver2D() sets (MacGridFluidSolver2D.cpp:8-14). */
void vgx2_params_plume(vgx2_params* params);

/* ---------------- MacGridFluidSolver2D, GPU-resident ---------------- */
typedef struct vgx2_sim vgx2_sim;

/* MacGridFluidSolver2D(int resolution, float boxSize) (.cpp:15-16); params
 * NULL gives MacGridParameters2D{}. Rejects n <= 0 and bad boxes with
 * VGX_EINVAL, an index range beyond int with VGX_ERANGE
 * (MacGridDomain2D.cpp:6-19). */
int vgx2_sim_create(int n, float box, const vgx2_params* params, vgx2_sim** out);
/* MacGridFluidSolver2D() (.cpp:8-14): 128^2, box 1, the plume parameters. */
int vgx2_sim_create_default(vgx2_sim** out);
void vgx2_sim_destroy(vgx2_sim* sim);

int vgx2_sim_get_params(const vgx2_sim* sim, vgx2_params* params);
int vgx2_sim_set_params(vgx2_sim* sim, const vgx2_params* params);
int vgx2_sim_resolution(const vgx2_sim* sim);
float vgx2_sim_box_size(const vgx2_sim* sim);

/* advance(dt) (.cpp:22-45): parameter checks (VGX_EINVAL), walls, then the
 * simple step (velocity advection, buoyancy, emitter, projection, density
 * advection with decay) or the midpoint-reflection step (.cpp:114-142). A
 * failure inside the step (non-finite velocity, divergence or backtrace)
 * leaves the state where vapor's exception leaves it. */
int vgx2_sim_advance(vgx2_sim* sim, float dt);
/* resetState() (.cpp:144-151). */
int vgx2_sim_reset(vgx2_sim* sim);
/* setBoxSize() (.cpp:152-156): a changed size resets the state. */
int vgx2_sim_set_box(vgx2_sim* sim, float box);

/* MacGridFluidSolver2D(MacGridState2D) adopts fields: upload replaces them
 * (caller buffers of n^2, (n+1)n, n(n+1) floats; NULL skips a field). */
int vgx2_sim_upload(vgx2_sim* sim, const float* rho, const float* u, const float* v);
int vgx2_sim_download(vgx2_sim* sim, float* rho, float* u, float* v);

/* lastProjection() / midpointProjection() (.hpp:37-38). */
int vgx2_sim_last_projection(const vgx2_sim* sim, vgx_projection_result* result);
int vgx2_sim_midpoint_projection(const vgx2_sim* sim, vgx_projection_result* result);
/* diagnostics() (.cpp:157-167); VGX_EINVAL on a non-finite divergence. */
int vgx2_sim_diagnostics(vgx2_sim* sim, vgx_diagnostics* out);

/* Device-resident fields (imageData's borrowed density, MacRenderData2D),
 * valid until the next advance/reset/set_box. */
int vgx2_sim_device_fields(vgx2_sim* sim, const float** rho, const float** u, const float** v);
int vgx2_sim_synchronize(vgx2_sim* sim);

/* ---------------- stages on host buffers ---------------- */
/* SemiLagrangian2D::advect (SemiLagrangian2D.cpp:24-50) for field
 * VGX2_DENSITY / _VELOCITY_X / _VELOCITY_Y of an n x n grid with cell size
 * box / n: u ((n+1)n), v (n(n+1)) the tracing velocity, src and out of the
 * field's extent. Rejects (VGX_EINVAL) a negative or non-finite dt, output
 * overlapping an input, non-finite inputs and non-finite backtraces. */
int vgx2_backend_advect(int n, float box, int field, float dt, const float* u, const float* v, const float* src,
                        float* out);
/* CpuPressureSolver2D::project (CpuPressureSolver2D.cpp:9-81) on the faces
 * (u, v) in place; options->diagonal_floor and the extension fields are
 * ignored. pressure_out (n^2 doubles, may be NULL) receives pressure(). */
int vgx2_backend_project(int n, float box, const vgx_projection_options* options, float* u, float* v,
                         double* pressure_out, vgx_projection_result* result);

/* ---------------- ReflectionMacFluidSolver2D, GPU-resident ---------------- */
/* Mirror of the historical solver's public parameters
 * (ReflectionMacFluidSolver2D.hpp:20-23,59-77). */
typedef struct vgx2r_params {
    int32_t maccormack_smoke;      /* 1 */
    float source_strength;         /* 1 */
    float buoyancy;                /* 0.8 */
    float smoke_decay;             /* 0.3007525 */
    int32_t project_iterations;    /* 60 */
    float emitter_radius;          /* 0.08 */
    float emitter_falloff;         /* 3.5 */
    float emitter_center_x;        /* 0.5 */
    float emitter_center_y;        /* 0.2 */
    float emitter_smoke_scale;     /* 3 */
    float emitter_kick_multiplier; /* 1 */
    float emitter_kick_base;       /* 0.6 */
    float emitter_kick_scale;      /* 1 */
    int32_t emitter_height;        /* 4 (unused by the step, as in vapor) */
    float sor_omega;               /* 1.9 */
    float buoyancy_split;          /* 0.5 */
    float buoyancy_exp;            /* 0.25 */
    float stir_strength;           /* 0 */
    uint32_t random_seed;          /* 0x9E3779B9 */
} vgx2r_params;

typedef struct vgx2r_sim vgx2r_sim;

/* Member defaults: ReflectionMacFluidSolver2D(int res) (:84-92). */
void vgx2r_params_default(vgx2r_params* params);
/* ReflectionMacFluidSolver2D(AdvectionMethod) (:76-83): 128^2, radius 0.04,
 * 120 iterations, the given smoke method (VGX_SEMI_LAGRANGIAN / VGX_MACCORMACK). */
int vgx2r_params_plume(int smoke_method, vgx2r_params* params);
/* ReflectionMacFluidSolver2D(int res): 2 <= res <= 512, else VGX_EINVAL. */
int vgx2r_sim_create(int n, const vgx2r_params* params, vgx2r_sim** out);
void vgx2r_sim_destroy(vgx2r_sim* sim);
int vgx2r_sim_get_params(const vgx2r_sim* sim, vgx2r_params* params);
int vgx2r_sim_set_params(vgx2r_sim* sim, const vgx2r_params* params);
int vgx2r_sim_resolution(const vgx2r_sim* sim);

/* advance(dt) (:33-43): checks, applyBoundary, step(dt, buoyancy,
 * exp(-smokeDecay/60), sourceStrength, projectIterations). */
int vgx2r_sim_advance(vgx2r_sim* sim, float dt);
/* The public steps of the reference struct: step (:316-408), emit (:291-314),
 * applyBoundary (:229-238), advectScalar on the smoke (:162-205) and project
 * (:240-289). */
int vgx2r_sim_step(vgx2r_sim* sim, float dt, float buoyancy, float dissipation, float source_strength,
                   int project_iterations);
int vgx2r_sim_emit(vgx2r_sim* sim, float strength, float dt);
int vgx2r_sim_apply_boundary(vgx2r_sim* sim);
int vgx2r_sim_advect_smoke(vgx2r_sim* sim, float dt);
int vgx2r_sim_project(vgx2r_sim* sim, float dt, int iterations);
/* resetState() (:28-31): zero fields, reseed the LCG. */
int vgx2r_sim_reset(vgx2r_sim* sim);
/* setBoxSize (:26-27): only 1 is accepted. */
int vgx2r_sim_set_box(vgx2r_sim* sim, float box);

/* The public vectors smoke (n^2), vx ((n+1)n), vy (n(n+1)), pressure (n^2)
 * and divergence (n^2); NULL skips one. */
int vgx2r_sim_upload(vgx2r_sim* sim, const float* smoke, const float* vx, const float* vy, const float* pressure,
                     const float* divergence);
int vgx2r_sim_download(vgx2r_sim* sim, float* smoke, float* vx, float* vy, float* pressure, float* divergence);
int vgx2r_sim_random_state(vgx2r_sim* sim, uint32_t* state);
int vgx2r_sim_set_random_state(vgx2r_sim* sim, uint32_t state);
/* diagnostics() (:44-57). */
int vgx2r_sim_diagnostics(vgx2r_sim* sim, vgx_diagnostics* out);
int vgx2r_sim_device_fields(vgx2r_sim* sim, const float** smoke, const float** vx, const float** vy);

#ifdef __cplusplus
}
#endif

#endif /* VAPOR_MAC2D_H */
