/*
 * vapor_mac3d.h - C ABI of the MI355X (gfx950) MAC 3D smoke step.
 *
 * This is the drop-in boundary for vapor's 3D MAC solver. Every entry point
 * takes plain pointers and sizes; no C++ or torch types cross it, and no
 * exception escapes it. Errors come back as status codes, with a
 * thread-local message from vgx_last_error().
 *
 * Reference interfaces replaced (paths relative to the vapor source tree):
 *   MacGridParameters3D ........ src/solvers/mac3d/MacGridFluidSolver3D.hpp:14-36
 *   MacAlgorithm presets ....... src/solvers/mac3d/MacGridFluidSolver3D.cpp:11-26
 *   MacGridFluidSolver3D /
 *   GpuMacSimulation3D ......... MacGridFluidSolver3D.hpp:39-73,
 *                                gpu/GpuMacSimulation3D.hpp:9-27
 *   ProjectionOptions/Result ... src/common/ProjectionOptions.hpp:5-17,
 *                                src/common/LinearSolveOptions.hpp:4-17
 *   FluidDiagnostics ........... src/common/FluidDiagnostics.hpp:3-7
 *   Advection3D plug-in ........ src/solvers/mac3d/Advection3D.hpp:6-14
 *   Projection3D plug-in ....... src/solvers/mac3d/Projection3D.hpp:6-11
 *   GpuMacBackend (staged) ..... gpu/GpuMacBackend.hpp:7-48
 *
 * Field layouts are the reference's (MacGridDomain3D.hpp:22-31), FP32 SoA:
 *   density  rho[x + n*(y + n*z)]             n*n*n values
 *   u        u[x + (n+1)*(y + n*z)]           (n+1)*n*n values
 *   v        v[x + n*(y + (n+1)*z)]           n*(n+1)*n values
 *   w        w[x + n*(y + n*z)], z in [0, n]  n*n*(n+1) values
 *   pressure p[x + n*(y + n*z)]               n*n*n values
 */
#ifndef VAPOR_MAC3D_H
#define VAPOR_MAC3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. They map onto the reference's exception types:
 * invalid_argument -> VGX_EINVAL, length_error -> VGX_ERANGE,
 * runtime_error (device failure) -> VGX_EDEVICE. */
#define VGX_OK 0
#define VGX_EINVAL 1
#define VGX_ERANGE 2
#define VGX_EDEVICE 3

/* MacAlgorithm (src/common/MacAlgorithm.hpp:3-9). */
#define VGX_SIMPLE_SEMI_LAGRANGIAN 0
#define VGX_SIMPLE_MACCORMACK 1
#define VGX_REFLECTION_SEMI_LAGRANGIAN 2
#define VGX_REFLECTION_MACCORMACK 3

/* AdvectionMethod (src/common/MacAlgorithm.hpp:9). */
#define VGX_SEMI_LAGRANGIAN 0
#define VGX_MACCORMACK 1

/* FaceAxis (src/solvers/mac3d/AdvectionContext.hpp:9). */
#define VGX_AXIS_X 0
#define VGX_AXIS_Y 1
#define VGX_AXIS_Z 2

/* Extension enums (no counterpart in the reference). */
#define VGX_EMITTER_SPHERE 0   /* reference: density 1 inside the sphere */
#define VGX_EMITTER_GAUSSIAN 1 /* min(1, 3 s exp(-3.5 (d/r)^2)), capped */
#define VGX_SOLVE_FIXED 0      /* reference: a fixed sweep budget */
#define VGX_SOLVE_TOLERANCE 1  /* stop at relative L2 residual <= tolerance */
#define VGX_WARM_ZERO 0        /* reference: x = 0 before every solve */
#define VGX_WARM_PREVIOUS 1    /* start each projection from its own pressure of the previous step */
#define VGX_LAUNCH_GRAPH 0     /* tolerance chunks replay captured HIP graphs */
#define VGX_LAUNCH_STREAM 1    /* every kernel launched on the stream (same results) */

/* Mirror of MacGridParameters3D (MacGridFluidSolver3D.hpp:14-36) followed by
 * the extension fields. vgx_params_default() gives the reference defaults and
 * extension fields that reproduce the reference behaviour. */
typedef struct vgx_params {
    float emitter_radius;          /* 0.035, fraction of the box */
    float emitter_center_x;        /* 0.5 */
    float emitter_center_y;        /* 0.5 */
    float emitter_center_z;        /* 0.1 (z is up) */
    float emitter_kick_multiplier; /* 1 */
    float emitter_kick_base;       /* 0.8 */
    float emitter_kick_scale;      /* 0.8 */
    float stir_strength;           /* 0.5 */
    uint32_t random_seed;          /* 0x9E3779B9, applied at creation and reset */
    float sor_omega;               /* 1.95 */
    float buoyancy_split;          /* 0.5 */
    float cfl_mc;                  /* 3 */
    int32_t maccormack_smoke;      /* 1 */
    int32_t maccormack_vel;        /* 1 */
    float smoke_decay;             /* 0.06 */
    int32_t reflection;            /* 1 */
    float buoyancy;                /* 2.5 */
    float source_strength;         /* 1 */
    int32_t project_iterations;    /* 64 */
    /* ---- extensions ---- */
    int32_t emitter_profile;       /* VGX_EMITTER_SPHERE */
    int32_t emitter2_enabled;      /* 0; second emitter after the first */
    float emitter2_radius;
    float emitter2_center_x, emitter2_center_y, emitter2_center_z;
    int32_t solve_mode;            /* VGX_SOLVE_FIXED */
    int32_t warm_start;            /* VGX_WARM_ZERO */
    int32_t launch_mode;           /* VGX_LAUNCH_GRAPH */
    double tolerance;              /* 1e-5, relative L2 residual (tolerance mode) */
    int32_t max_sweeps;            /* 20000, sweep cap per projection (tolerance mode) */
    int32_t check_interval;        /* 16, sweeps between FP64 residual checks */
    int32_t refinement_steps;      /* 0, FP64 iterative-refinement corrections */
    int32_t final_check_interval;  /* 0; k > 0: once the residual's trend says the next
                                      check_interval sweeps reach the tolerance, they are
                                      checked every k (single-scene tolerance solves) */
} vgx_params;

/* Mirror of ProjectionOptions + LinearSolveOptions, plus the diagonal floor
 * the MAC solver builds its SOR with (MacGridFluidSolver3D.hpp:118 uses 6;
 * CpuPressureSolver3D's own default is 0). */
typedef struct vgx_projection_options {
    float dt;                      /* 1/60 */
    float relaxation;              /* 1.95 */
    int32_t max_iterations;        /* 200 */
    int32_t residual_check_interval; /* 8 */
    double absolute_tolerance;     /* 1e-6 (infinity norm) */
    double relative_tolerance;     /* 1e-5 (infinity norm, relative to initial) */
    int32_t fixed_iterations;      /* 0 */
    float diagonal_floor;          /* 6 */
    /* ---- extensions ---- */
    int32_t use_l2;                /* 0: reference inf-norm rule; 1: ||r||2/||b||2 <= l2_tolerance */
    int32_t warm_start;            /* 0: x = 0 (reference); 1: start from pressure_inout */
    double l2_tolerance;           /* 1e-5 */
    int32_t refinement_steps;      /* 0 */
    int32_t reserved0;
} vgx_projection_options;

/* Mirror of ProjectionResult + LinearSolveResult, with extension fields. */
typedef struct vgx_projection_result {
    int32_t iterations;            /* SOR sweeps performed */
    int32_t converged;
    double initial_residual;       /* infinity norm of b - Ax, FP64 */
    double final_residual;
    int32_t residual_available;
    float divergence_before;       /* max |div u| before */
    float divergence_after;        /* max |div u| after */
    int32_t learned_start;         /* extension: 0 not tried, 1 learned start used, 2 rejected after 10 sweeps */
    double removed_rhs_mean;
    /* ---- extensions ---- */
    double initial_residual_l2;    /* ||b - Ax||2 / ||b||2 at the start */
    double final_residual_l2;      /* ... at the end */
    int32_t residual_checks;       /* FP64 residual evaluations */
    int32_t refinement_steps;      /* FP64 corrections applied */
    int32_t warm_started;          /* 1 if the solve started from a nonzero guess */
    int32_t reserved1;
} vgx_projection_result;

/* Mirror of FluidDiagnostics (FluidDiagnostics.hpp:3-7). */
typedef struct vgx_diagnostics {
    float kinetic_energy;
    float max_velocity;
    float max_divergence;
} vgx_diagnostics;

/* ---------------- library ---------------- */

/* Thread-local message of the last failing call on this thread. */
const char* vgx_last_error(void);
/* sizeof of each struct, so bindings can check their layouts. */
size_t vgx_struct_size(int which); /* 0 params, 1 options, 2 result, 3 diagnostics, 4 batch tuning, 5 field error */
/* Device name of the current HIP device, or an error. */
int vgx_device_name(char* buffer, size_t size);

/* MacGridParameters3D() defaults (MacGridFluidSolver3D.hpp:14-36). */
void vgx_params_default(vgx_params* params);
/* MacGridParameters3D(MacAlgorithm) presets (MacGridFluidSolver3D.cpp:11-26). */
int vgx_params_preset(int algorithm, vgx_params* params);
/* ProjectionOptions{} defaults with diagonal floor 6. */
void vgx_projection_options_default(vgx_projection_options* options);

/* ---------------- GPU-resident simulation ---------------- */
/* Mirrors GpuMacSimulation3D / MacGridFluidSolver3D. One host thread per
 * handle; each handle owns a HIP stream and its device buffers. */
typedef struct vgx_sim vgx_sim;

/* GpuMacSimulation3D(int n, float box, MacGridParameters3D)
 * (GpuMacSimulation3D.hpp:12). params may be NULL: reference defaults with
 * project_iterations = 2n, as MacGridFluidSolver3D(int) sets (.cpp:43). */
int vgx_sim_create(int n, float box, const vgx_params* params, vgx_sim** out);
/* GpuMacSimulation3D(MacAlgorithm): 32^3, box 4.5, preset parameters. */
int vgx_sim_create_algorithm(int algorithm, vgx_sim** out);
void vgx_sim_destroy(vgx_sim* sim);

int vgx_sim_get_params(const vgx_sim* sim, vgx_params* params);
int vgx_sim_set_params(vgx_sim* sim, const vgx_params* params);
int vgx_sim_resolution(const vgx_sim* sim);
float vgx_sim_box_size(const vgx_sim* sim);

/* advance(dt) (MacGridFluidSolver3D.cpp:323-330): emit, buoyancy, walls,
 * velocity advection, projection(s), smoke advection, decay. */
int vgx_sim_advance(vgx_sim* sim, float dt);
/* resetState(): zero fields and pressure, reseed the emitter RNG. */
int vgx_sim_reset(vgx_sim* sim);
/* setBoxSize(): a changed size resets the state. */
int vgx_sim_set_box(vgx_sim* sim, float box);

/* upload/download (GpuMacSimulation3D.cpp:110-141). Caller-owned host buffers
 * of n^3, (n+1)n^2, (n+1)n^2, (n+1)n^2 floats. Any pointer may be NULL to skip
 * that field. Upload rejects non-finite values. */
int vgx_sim_upload(vgx_sim* sim, const float* rho, const float* u, const float* v, const float* w);
int vgx_sim_download(vgx_sim* sim, float* rho, float* u, float* v, float* w);
/* Extension (no counterpart in the reference; the initial state of the 256^3
 * configuration): adds sigma * N(0, 1) to every face velocity, the variate of
 * face i of axis a being Philox4x32-10 at counter (i, a, 0, 0) and key seed,
 * through Box-Muller, the same bits on any host or device. */
int vgx_sim_add_velocity_noise(vgx_sim* sim, float sigma, uint64_t seed);
/* Pressure of the last projection of the step (zero-mean gauge), n^3 floats. */
int vgx_sim_download_pressure(vgx_sim* sim, float* p);
/* Replaces the pressure both projection slots start from under
 * VGX_WARM_PREVIOUS (each projection of a reflection step keeps its own
 * history, because the two solve for pressures of roughly opposite sign). */
int vgx_sim_upload_pressure(vgx_sim* sim, const float* p);
/* uploadedBytes/downloadedBytes (GpuMacSimulation3D.hpp:25-26). */
int vgx_sim_transfer_bytes(const vgx_sim* sim, uint64_t* uploaded, uint64_t* downloaded);

/* lastProjection(); index 0 = the step's last projection, 1 = its first
 * (reflection steps run two). */
int vgx_sim_last_projection(const vgx_sim* sim, int index, vgx_projection_result* result);
/* diagnostics(): FP64 kinetic energy, max |velocity|, max |div u| computed on
 * the device (MacGridFluidSolver3D.cpp:247-269,332-334). */
int vgx_sim_diagnostics(vgx_sim* sim, vgx_diagnostics* out);

/* Device-resident field pointers (RenderData hand-off without a download,
 * GpuMacSimulation3D.cpp:107-109). Valid until the next advance/reset. */
int vgx_sim_device_fields(vgx_sim* sim, const float** rho, const float** u,
                          const float** v, const float** w, const float** p);
/* Blocks until the handle's stream is idle. */
int vgx_sim_synchronize(vgx_sim* sim);

/* Snapshots: the state on disk as raw little-endian FP32 files in the
 * reference layouts plus a JSON header. Files: <path>.json (format
 * "vapor-mac3d-snapshot", version 1, resolution, box_size, random_state and
 * the list below), <path>.density (n^3; the bytes of vapor's raw density dump,
 * ReflectionComparison.hpp:74-75), <path>.velx/.vely/.velz ((n+1)n^2 each),
 * <path>.pressure and <path>.pressure1 (n^3: the pressure histories the
 * warm start continues from). Loading a snapshot into a simulation of the same
 * grid (GpuMacSimulation3D::upload's checks, GpuMacSimulation3D.cpp:110-113)
 * restores the fields, pressures and emitter RNG state, so the trajectory
 * continues bit for bit. Loading validates every file before changing state. */
int vgx_sim_save_snapshot(vgx_sim* sim, const char* path);
int vgx_sim_load_snapshot(vgx_sim* sim, const char* path);

/* Installed stages: MacGridFluidSolver3D::setAdvection / setPressureSolver /
 * useDefaultAdvection / useDefaultPressureSolver (MacGridFluidSolver3D.hpp:54-57,
 * .cpp:305-321). A hook replaces that stage of every later step; NULL restores
 * the device stage. Its arguments are pinned host buffers of the handle,
 * valid for the call; the rest of the step stays on the device.
 *   advection hook: Advection3D::advectFace (Advection3D.hpp:10-12) for
 *     target VGX_AXIS_X/Y/Z, source -> out ((n+1)n^2 floats each), and
 *     Advection3D::advectScalar (.hpp:8-9) for target VGX_TARGET_DENSITY,
 *     source NULL and out the n^3 density advected in place; flow (u, v, w)
 *     and dt, cfl_limit, method (VGX_SEMI_LAGRANGIAN / VGX_MACCORMACK) as the
 *     reference's AdvectionContext carries them. Calls per step: x, y, z
 *     (twice under reflection, the second time with source 2u - u~), then
 *     the density; decay is applied after the density call.
 *   projection hook: Projection3D::project (Projection3D.hpp:8) on the faces
 *     (u, v, w) in place, with the options MacGridFluidSolver3D::project
 *     builds (dt, relaxation, max_iterations, fixed_iterations = 1; dt/2 and
 *     max(1, K/2) under reflection); rho is the current density (read only);
 *     *result becomes lastProjection().
 * A hook returns VGX_OK, or a status that the step returns. */
#define VGX_TARGET_DENSITY 3
typedef int (*vgx_advection_hook)(void* user, int target, int method, float dt, float cfl_limit, const float* u,
                                  const float* v, const float* w, const float* source, float* out);
typedef int (*vgx_projection_hook)(void* user, const vgx_projection_options* options, float* u, float* v, float* w,
                                   const float* rho, vgx_projection_result* result);
int vgx_sim_set_advection_hook(vgx_sim* sim, vgx_advection_hook hook, void* user);
int vgx_sim_set_projection_hook(vgx_sim* sim, vgx_projection_hook hook, void* user);

/* Per-kernel timing with HIP events on the handle's stream. */
int vgx_sim_set_timing(vgx_sim* sim, int enabled);
int vgx_sim_reset_timing(vgx_sim* sim);
/* Number of kernel kinds that vgx_sim_kernel_timing can report. */
int vgx_kernel_kinds(void);
const char* vgx_kernel_name(int kind);
/* Total device milliseconds and launch count of one kernel kind. */
int vgx_sim_kernel_timing(vgx_sim* sim, int kind, double* total_ms, uint64_t* launches);
/* Sweeps performed by all projections since creation or reset_timing. */
int vgx_sim_sweep_count(const vgx_sim* sim, uint64_t* sweeps);

/* ---------------- staged plug-in backend ---------------- */
/* Mirrors GpuMacBackend + GpuAdvection + GpuPressureSolver: each call uploads
 * its host inputs, runs on the device and downloads the output. */
typedef struct vgx_backend vgx_backend;

int vgx_backend_create(int n, float box, vgx_backend** out);
void vgx_backend_destroy(vgx_backend* backend);

/* Advection3D::advectScalar (Advection3D.hpp:8-9): field is advected in place
 * by flow (u, v, w) over dt. Output must not alias the flow. */
int vgx_backend_advect_scalar(vgx_backend* backend, int method, float dt, float cfl_limit,
                              const float* u, const float* v, const float* w, float* field);
/* Advection3D::advectFace (Advection3D.hpp:10-12). out must not alias source
 * or the flow. */
int vgx_backend_advect_face(vgx_backend* backend, int axis, int method, float dt, float cfl_limit,
                            const float* u, const float* v, const float* w,
                            const float* source, float* out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vgx_backend_advect_face's computation on caller-owned
 * DEVICE buffers of an n^3 grid ((n+1)n^2 floats each), enqueued on `stream`
 * (a hipStream_t; NULL is the default stream) without waiting, with no
 * allocation of any kind. The MacCormack forward trace lives in a fixed
 * library-owned device array of 8,454,144 bytes per GPU (the faces of one
 * axis at n = 128), so n > 128 is refused with VGX_ERANGE. Calls that overlap
 * on one GPU (other streams or threads) share that array and race; the
 * caller orders them. Same checks and results as vgx_backend_advect_face. */
int vgx_advect_face_noalloc(int n, float box, int axis, int method, float dt, float cfl_limit, const float* u,
                            const float* v, const float* w, const float* source, float* out, void* stream);
/* Projection3D::project (Projection3D.hpp:8) on host fields, in place.
 * pressure_inout (n^3, may be NULL) receives the zero-mean pressure; with
 * options->warm_start it also supplies the starting guess. */
int vgx_backend_project(vgx_backend* backend, const vgx_projection_options* options,
                        float* u, float* v, float* w, float* pressure_inout,
                        vgx_projection_result* result);

/* ---------------- learned pressure warm start (extension) ---------------- */
/* No counterpart in the reference. A 4-level 3D U-Net (16/32/48/64 channels,
 * two 3^3 convolutions with ReLU per level, average pooling, nearest
 * upsampling, concatenated skips, no biases, 1x1x1 head) in BF16 with FP32
 * accumulation on the matrix cores. Input: a scalar field on an n^3 grid
 * (n % 8 == 0) scaled by in_scale, plus a boundary mask; output: one scalar
 * field scaled by out_scale. Parameters are FP32 masters, ordered layer by
 * layer as [cout][27 taps][cin] (tap = (dz+1) 9 + (dy+1) 3 + (dx+1)), then
 * the 16 head weights. */
typedef struct vgx_unet vgx_unet;
int vgx_unet_create(uint64_t seed, vgx_unet** out); /* He-normal initial weights */
void vgx_unet_destroy(vgx_unet* unet);
size_t vgx_unet_param_count(void);
int vgx_unet_get_params(vgx_unet* unet, float* params);
int vgx_unet_set_params(vgx_unet* unet, const float* params);
/* The grid the network runs on inside a simulation's learned start: 0 (the
 * default) for the simulation's own grid; m > 0 (a multiple of 8) for the
 * residual restricted to m^3 by cell means and the prediction prolonged back
 * by cell-centred trilinear interpolation, on simulations whose n is a
 * multiple of m. The restriction keeps the smooth modes red-black SOR is
 * slow on and gives the network's four levels the whole domain. */
int vgx_unet_set_grid(vgx_unet* unet, int m);
int vgx_unet_get_grid(const vgx_unet* unet);
int vgx_unet_get_grads(vgx_unet* unet, float* grads);
int vgx_unet_set_grads(vgx_unet* unet, const float* grads);
/* out = out_scale * f(in_scale * field), host buffers of n^3 floats. */
int vgx_unet_predict(vgx_unet* unet, int n, const float* field, float in_scale, float out_scale, float* out);
/* Mean squared error of f(in_scale * field) against target and its gradient
 * into the gradient buffer (added to it with accumulate). */
int vgx_unet_loss_grad(vgx_unet* unet, int n, const float* field, float in_scale, const float* target,
                       int accumulate, double* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on gradient * grad_scale. */
int vgx_unet_adam_step(vgx_unet* unet, float lr, float grad_scale);

/* Training pairs generated by the solver itself, kept on the device: each
 * warm tolerance projection of a recording simulation adds (residual of its
 * starting pressure, gauged correction the solve applied to that pressure),
 * both divided by the residual's RMS. */
typedef struct vgx_dataset vgx_dataset;
int vgx_dataset_create(int n, int capacity, vgx_dataset** out);
void vgx_dataset_destroy(vgx_dataset* data);
int vgx_dataset_size(const vgx_dataset* data);
int vgx_dataset_clear(vgx_dataset* data);
/* Copies pair `index` (n^3 floats each; either pointer may be NULL) to the host. */
int vgx_dataset_get(const vgx_dataset* data, int index, float* input, float* target);
/* Adam on mini-batches of `batch` pairs in a seeded random order, `epochs`
 * passes; mean_loss receives the last epoch's mean squared error. */
int vgx_unet_train(vgx_unet* unet, const vgx_dataset* data, int epochs, int batch, float lr, uint64_t seed,
                   double* mean_loss);
/* Record training pairs from this simulation's projections into data (NULL
 * stops); the dataset's n equals the simulation's, or divides it and the
 * pairs are restricted to the dataset's grid by cell means. */
int vgx_sim_record(vgx_sim* sim, vgx_dataset* data);
/* Learned warm start (NULL disables; the network is borrowed and must
 * outlive its use): warm tolerance solves start from the previous pressure
 * plus the network's correction for its residual when, after 10 sweeps from
 * each start, that start's residual is the smaller; otherwise the plain
 * previous-pressure start continues (vgx_projection_result.learned_start). */
int vgx_sim_set_unet(vgx_sim* sim, vgx_unet* unet);

/* ---------------- batches of scenes (extension) ---------------- */
/* No counterpart in the reference. `count` independent scenes, each an n[s]^3
 * grid on a box[s] box with its own parameters, stepped together on one
 * stream: every stage is one launch over all scenes, and a tolerance solve
 * masks converged scenes out of the shared sweeps. Each scene's fields are
 * those of the same scene stepped alone (vgx_sim_advance). The scenes must
 * share the step variant (reflection, MacCormack flags) and the solver
 * parameters (solve_mode, warm_start, tolerance, max_sweeps, check_interval,
 * sor_omega, project_iterations, launch_mode); refinement_steps must be 0.
 * Emitters, buoyancy, decay, stir and seeds may differ per scene. Creation
 * runs the launch-shape autotuner (vgx_batch_get_tuning). */
typedef struct vgx_batch vgx_batch;
typedef struct vgx_batch_tuning {
    int elementwise_per_thread_log2; /* advection and forces: 2^k faces per thread */
    int advect_grid;                 /* blocks of the advection launches (0: one per item) */
    int sweep_grid;                  /* blocks of the red-black sweep launches (0: one per tile) */
    int stencil_grid;                /* blocks of the divergence / residual / gradient launches */
    int elementwise_items, tile_items, stencil_items; /* work items over all scenes */
    int sweep_register;   /* sweep staging: 1 right-hand side in registers, 0 in LDS (bit-identical) */
    double advect_ms[12]; /* MacCormack x-face pass: [per-thread log2 0..2][grid 0, 4096, 2048, 1024] */
    double sweep_ms[4];   /* one sweep, register staging: grid 0, 2048, 1024, 512 (0: not available) */
    double stencil_ms[4]; /* residual check: grid 0, 4096, 2048, 1024 */
    double sweep_lds_ms[4]; /* one sweep, LDS staging: grid 0, 2048, 1024, 512 */
} vgx_batch_tuning;

int vgx_batch_create(int count, const int* n, const float* box, const vgx_params* params, vgx_batch** out);
void vgx_batch_destroy(vgx_batch* batch);
int vgx_batch_size(const vgx_batch* batch);
/* Borrowed handle of scene `index` for the per-scene calls (download, upload,
 * diagnostics, last_projection, device_fields, ...). Advance, reset,
 * set_params, set_box, set_unet and record are refused on it, and destroy
 * does nothing: the batch owns the scene. */
int vgx_batch_scene(vgx_batch* batch, int index, vgx_sim** scene);
int vgx_batch_advance(vgx_batch* batch, float dt);
int vgx_batch_reset(vgx_batch* batch);
/* Per scene (count doubles): milliseconds from the start of the last step
 * until that scene's last projection reached its tolerance (-1 in fixed
 * mode or when the sweep budget ran out first). */
int vgx_batch_scene_latency(const vgx_batch* batch, double* ms);
int vgx_batch_get_tuning(const vgx_batch* batch, vgx_batch_tuning* out);
int vgx_batch_set_timing(vgx_batch* batch, int enabled);
int vgx_batch_reset_timing(vgx_batch* batch);
int vgx_batch_kernel_timing(vgx_batch* batch, int kind, double* total_ms, uint64_t* launches);
int vgx_batch_synchronize(vgx_batch* batch);

/* ---------------- shadow runs in a chosen precision (extension) ---------------- */
/* No counterpart in the reference. A shadow is a copy of a simulation that
 * steps the same scene with plain kernels, one per operation of vapor's CPU
 * step (MacGridFluidSolver3D.cpp:192-245, CpuPressureSolver3D.cpp:25-60) in
 * the reference's operation order, with every field, intermediate and solve
 * in the chosen precision:
 *   VGX_PRECISION_FP64: the all-FP64 run the FP32 path's error is measured
 *     against; its SOR solves to the tolerance in FP64 (no refinement);
 *   VGX_PRECISION_FP32: vapor's FP32 step by a second route.
 * The copy takes the source's grid, parameters, fields, emitter RNG state and
 * pressure histories (FP32 values convert exactly). It follows the source's
 * solve mode: fixed sweeps, or the relative L2 tolerance checked every
 * check_interval sweeps, with or without the previous-pressure warm start;
 * refinement_steps is ignored. One host thread per handle. */
#define VGX_PRECISION_FP32 32
#define VGX_PRECISION_FP64 64
typedef struct vgx_shadow vgx_shadow;
/* Error of one field against the shadow's: max |a - ref|, RMS of a - ref,
 * max |ref|, RMS of ref, all in FP64. */
typedef struct vgx_field_error {
    double max_abs;
    double rms;
    double ref_max;
    double ref_rms;
} vgx_field_error;
int vgx_shadow_create(vgx_sim* source, int precision, vgx_shadow** out);
void vgx_shadow_destroy(vgx_shadow* shadow);
/* One step, as vgx_sim_advance. */
int vgx_shadow_advance(vgx_shadow* shadow, float dt);
/* Fields widened to double (any pointer may be NULL): n^3 density, three
 * (n+1)n^2 face arrays, and the last projection's zero-mean pressure (n^3). */
int vgx_shadow_download(vgx_shadow* shadow, double* rho, double* u, double* v, double* w, double* p);
/* As vgx_sim_last_projection (0 the step's last projection, 1 its first). */
int vgx_shadow_last_projection(const vgx_shadow* shadow, int index, vgx_projection_result* result);
/* SOR sweeps of all the shadow's projections since creation. */
int vgx_shadow_sweep_count(const vgx_shadow* shadow, uint64_t* sweeps);
/* errors[0..4]: the simulation's density, u, v, w and last projection's
 * pressure against the shadow's (the grids must match). */
int vgx_shadow_compare(vgx_shadow* shadow, vgx_sim* sim, vgx_field_error errors[5]);

#ifdef __cplusplus
}
#endif

#endif /* VAPOR_MAC3D_H */
