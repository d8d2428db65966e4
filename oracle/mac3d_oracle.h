/*
 * mac3d_oracle.h - CPU restatement of vapor's MAC 3D smoke step.
 *
 * TEST INFRASTRUCTURE ONLY. This oracle is the checker the GPU path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product (vapor_amd) never calls it and
 * has no CPU fallback.
 *
 * Every function restates one reference routine in the same floating-point
 * operation order, so that an FP32 build without contraction
 * (-ffp-contract=off) reproduces the reference bit for bit. Paths are relative
 * to the vapor source tree.
 */
#ifndef VAPOR_MAC3D_ORACLE_H
#define VAPOR_MAC3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/vapor_mac3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reference defaults of MacGridParameters3D and ProjectionOptions. */
void orc_params_default(vgx_params* params);
void orc_projection_options_default(vgx_projection_options* options);

/* Trilinear samplers, MacGridSampler3D.cpp:5-75. axis -1 = cell centred,
 * 0/1/2 = x/y/z faces. */
float orc_sample(int n, float box, int axis, const float* f, float px, float py, float pz);
/* MacGridSampler3D::velocity, MacGridSampler3D.cpp:73-75. */
void orc_velocity(int n, float box, const float* u, const float* v, const float* w,
                  float px, float py, float pz, float out[3]);

/* CpuAdvectionKernels::advectScalar, CpuAdvectionKernels.cpp:18-80.
 * scratch holds 2 * n^3 floats. */
int orc_advect_scalar(int n, float box, const float* u, const float* v, const float* w,
                      float dt, float cfl, int maccormack, float* field, float* scratch);
/* CpuAdvectionKernels::advectVx/Vy/Vz, CpuAdvectionKernels.cpp:82-269.
 * scratch holds 2 * (n+1) n^2 floats. */
int orc_advect_face(int n, float box, int axis, const float* src, const float* u, const float* v,
                    const float* w, float dt, float cfl, int maccormack, float* out, float* scratch);

/* MacGridOperators3D, MacGridOperators3D.cpp:8-67. */
void orc_enforce_walls(int n, float* u, float* v, float* w);
void orc_divergence(int n, float box, const float* u, const float* v, const float* w, float* div);
/* Returns 0 and writes the max; returns -1 on a non-finite value. */
int orc_max_divergence(int n, float box, const float* u, const float* v, const float* w, float* out);
void orc_apply_gradient(int n, float box, const float* p, float dt, float* u, float* v, float* w);

/* StencilLinearSystem::residualInfinityNorm on the MAC Neumann matrix,
 * StencilLinearSystem.cpp:53-70, CpuPressureSolver3D.cpp:9-23. */
double orc_residual_inf(int n, const float* b, const float* x);
/* Extension: ||r - mean(r)||_2 of r = b - Ax in FP64, and ||b - mean(b)||_2:
 * the compatible parts (constants are the Neumann matrix's null space). */
double orc_residual_l2(int n, const float* b, const float* x);
double orc_norm_l2(size_t count, const float* b);

/* SOR::solve on the Neumann matrix, SOR.cpp:22-64. Returns iterations. */
int orc_sor_solve(int n, float omega, float floor, const float* b, float* x,
                  const vgx_projection_options* options, vgx_projection_result* result);
/* One red-black sweep (both colours), SOR.cpp:36-49. */
void orc_sor_sweep(int n, float omega, float floor, const float* b, float* x);

/* CpuPressureSolver3D::project, CpuPressureSolver3D.cpp:25-60, with the
 * extension solve policies. pressure (n^3) is in/out (warm start). div and b
 * are n^3 workspaces. Returns 0, or -1 for invalid input (reference throws). */
int orc_project(int n, float box, const vgx_projection_options* options, float* u, float* v,
                float* w, float* pressure, float* div, float* b, vgx_projection_result* result);

/* Whole simulation, MacGridFluidSolver3D.cpp. */
typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(int n, float box, const vgx_params* params);
void orc_sim_destroy(orc_sim* sim);
void orc_sim_set_params(orc_sim* sim, const vgx_params* params);
void orc_sim_reset(orc_sim* sim);
/* Returns 0, or -1 for invalid parameters / non-finite state. */
int orc_sim_advance(orc_sim* sim, float dt);
/* which: 0 rho, 1 u, 2 v, 3 w, 4 pressure. */
float* orc_sim_field(orc_sim* sim, int which);
uint32_t orc_sim_random_state(const orc_sim* sim);
void orc_sim_last_projection(const orc_sim* sim, int index, vgx_projection_result* out);
void orc_sim_diagnostics(const orc_sim* sim, vgx_diagnostics* out);
float orc_kinetic_energy(int n, float box, const float* u, const float* v, const float* w);
float orc_max_velocity(int n, const float* u, const float* v, const float* w);

/* Extension: emitter pass alone (MacGridFluidSolver3D.cpp:140-165 and the
 * capped-Gaussian / second-emitter variants). Returns the new RNG state. */
uint32_t orc_emit(int n, float box, const vgx_params* params, float dt, uint32_t random_state,
                  float* rho, float* u, float* v, float* w);

/* Extension (no reference code): counter-based initial velocity noise.
 * Philox4x32-10 (Salmon et al., SC'11): 10 rounds of the 4x32 Philox round
 * with multipliers 0xD2511F53, 0xCD9E8D57 and key increments 0x9E3779B9,
 * 0xBB67AE85. out receives the 4 output words for counter ctr and key key. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* ln(x) and cos(2 pi t) by fixed polynomials in IEEE float arithmetic (the
 * generator's Box-Muller step; exported for accuracy tests). */
float orc_noise_log(float x);
float orc_noise_cos2pi(float t);
/* e^v of the capped-Gaussian emitter profile by fixed IEEE float operations
 * (valid for -87 <= v <= 88; the profile uses [-3.5, 0]). */
float orc_profile_exp(float v);
/* Standard normal variate of entry i of stream `stream`: Philox at counter
 * (i, stream, 0, 0) and key (seed low, seed high); Box-Muller of
 * u0 = ((w0 >> 8) + 1) 2^-24 and u1 = (w1 >> 8) 2^-24. */
float orc_noise_normal(uint64_t seed, uint32_t stream, uint32_t i);
/* Adds sigma * N(0,1) to every face of u (stream 0), v (1), w (2). */
void orc_add_velocity_noise(int n, float sigma, uint64_t seed, float* u, float* v, float* w);
void orc_sim_add_velocity_noise(orc_sim* sim, float sigma, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif
