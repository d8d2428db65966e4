// vgx_internal.hpp - host-side declarations shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "vgx_device.hpp"  // RbGeom

namespace vgx {

// Device-resident record of one projection, filled by the finalize kernels
// and read back by the host once per projection.
struct ProjState {
    double rhs_sum;       // sum of b = -div h^2/dt before the mean is removed
    double rhs_mean;      // removedRhsMean (CpuPressureSolver3D.cpp:44)
    double initial_inf;   // inf-norm of b - A x at x = 0
    double res2;          // last FP64 residual check: ||b - A x||_2^2
    double res_inf;       // ... its inf-norm
    double bnorm2;        // ||b'||_2^2 of the last full residual check (chunk checks keep it)
    double x_sum;         // sum of x at the last check / last sweep
    double final_res2;    // after the zero-mean gauge
    double final_inf;
    double final_bnorm2;
    double res_sum;       // sum of the residual at the last check (its mean is out of reach)
    double b_sum;         // sum of b', as bnorm2
    double final_res_sum;
    double final_b_sum;
    double checked_rel;   // tolerance chunks: last checked relative L2 residual
    int done;             // tolerance chunks: the checked residual met the tolerance
    int sweeps_run;       // tolerance chunks: sweeps of the chunks that ran
    float rhs_meanf;      // float(removedRhsMean), subtracted from every b
    float div_before;     // max |div u| before
    float x_meanf;        // float(mean x), subtracted from every x
    float div_after;      // max |div u| after
    int nonfinite_in;
    int nonfinite_x;
    int nonfinite_out;
    int chunks;  // tolerance chunks: checks run since the last reset (selects the ring slot)
    int zero_start;        // batched solves: the warm start was replaced by x = 0
    long long done_clock;  // device wall clock (wall_clock64) when `done` was raised
    double start_rel;      // batched solves: relative L2 residual of the starting iterate
    double start_inf;      // ... and its inf-norm
    double learned_alpha;  // learned start decided on the device: the line-search step
    int learned;           // ... and the judge's decision (1 kept, 2 rejected)
};

struct DiagState {
    double ke_sum;  // sum of squared face velocities
    double max_vel;
    double max_div;
    double nonfinite;
};

// Kernel timing: a launch made while a ProbeSource is current on the calling
// thread carries its own start/stop events (hipExtLaunchKernelGGL), so the
// measured time is the kernel's execution, not the gaps around it.
struct Probe {
    hipEvent_t start = nullptr, stop = nullptr;
};
struct ProbeSource {
    virtual Probe take() = 0;
    virtual ~ProbeSource() = default;
};
ProbeSource*& current_probe_source();

// Kinds of kernels, for the timing table.
// Doubles per block in a reduction-partials buffer.
constexpr int kPartialStride = 8;

enum Kind : int {
    K_EMIT = 0,
    K_FORCES,
    K_ADVECT_SL_FACES,
    K_ADVECT_MC_FACES,
    K_ADVECT_SL_SCALAR,
    K_ADVECT_MC_SCALAR,
    K_DIVERGENCE,
    K_SOR,
    K_RESIDUAL,
    K_GRADIENT,
    K_REDUCE,
    K_DIAGNOSTICS,
    K_FILL,
    K_SOR_CHECK,  // the sweep that ends a tolerance chunk (with its residual check)
    K_LEARNED,    // the learned warm start's network inference
    K_COUNT
};
extern const char* const kKindNames[K_COUNT];

// Emitter sphere (MacGridFluidSolver3D.cpp:140-165), with the patch bounds and
// every value the in-sphere test uses precomputed on the host exactly as the
// reference computes them, in the field precision T.
template <class T> struct EmitArgsT {
    int n;
    int lo[3];
    int ext[3];
    int profile;  // VGX_EMITTER_SPHERE / VGX_EMITTER_GAUSSIAN
    int kick_on;  // emitterKickMultiplier > 0
    T dx, ecx, ecy, ecz, erad;
    T kick;      // mult * (base + scale * strength)
    T stir;      // stirStrength
    T dt_scale;  // dt * 60
    T strength;  // sourceStrength
};
using EmitArgs = EmitArgsT<float>;

// One advection launch: up to 4 fields (faces x, y, z, then the scalar).
template <class T> struct AdvFieldT {
    const T* src;   // source S
    const T* tsrc;  // T for a reflected source 2 S - T (pass 2), else null
    const T* fwd;   // forward (SL) result, MacCormack pass only
    T* out;
};
template <class T> struct AdvArgsT {
    AdvFieldT<T> f[4];
    const T* u;
    const T* v;
    const T* w;
    int n;
    int first_field;  // blockIdx.y + first_field selects the field (0..2 faces, 3 cell)
    T h, box, dt, cfl;
    T out_scale;    // density decay factor, scalar only (1 = none)
    int scale_out;  // apply out_scale to the scalar output
};
using AdvField = AdvFieldT<float>;
using AdvArgs = AdvArgsT<float>;

// ---- precision-generic reference step (vgx_kernels.hip), T = float or double ----
// One plain kernel per operation of vapor's CPU step; reductions leave FP64
// per-block partials (stride kPartialStride) that the host folds in order.
namespace wide {
template <class T>
hipError_t emit(hipStream_t s, const EmitArgsT<T>& a, uint32_t* rng, T* noise, T* rho, T* u, T* v, T* w);
template <class T> hipError_t forces(hipStream_t s, int n, const T* rho, T* u, T* v, T* w, T buoyancy, T dt, T split);
// field 0..2 faces, 3 the scalar; a.f[field] names the arrays.
template <class T> hipError_t advect(hipStream_t s, const AdvArgsT<T>& a, int field, bool mc, bool hat);
template <class T> hipError_t walls(hipStream_t s, int n, T* u, T* v, T* w);
// partials: sum b, max |div|, non-finite count (b may be null: divergence only).
template <class T>
hipError_t divergence(hipStream_t s, int n, T h, T scale, const T* u, const T* v, const T* w, T* b, double* partials,
                      int* nblocks);
template <class T> hipError_t shift(hipStream_t s, T* x, T m, size_t count);
template <class T> hipError_t sor_color(hipStream_t s, int n, T omega, T floor, const T* b, T* x, int parity);
// partials: sum r, sum r^2, max |r|, sum b, sum b^2, non-finite x count.
template <class T> hipError_t residual(hipStream_t s, int n, const T* b, const T* x, double* partials, int* nblocks);
template <class T> hipError_t sum(hipStream_t s, const T* x, size_t count, double* partials, int* nblocks);
template <class T> hipError_t gradient(hipStream_t s, int n, T hInv, T dt, const T* p, T* u, T* v, T* w);
// partials: max |a - ref|, sum (a - ref)^2, max |ref|, sum ref^2.
template <class A, class B>
hipError_t compare(hipStream_t s, const A* a, const B* ref, size_t count, double* partials, int* nblocks);
template <class T> hipError_t convert(hipStream_t s, T* out, const float* in, size_t count);
}  // namespace wide

// ---- launchers (vgx_kernels.hip); all enqueue on `s` ----
hipError_t launch_fill_f32(hipStream_t s, float* p, size_t count, float value);
hipError_t launch_add_f32(hipStream_t s, const float* a, const float* b, float* out, size_t count);
hipError_t launch_scale_f32(hipStream_t s, const float* a, float scale, float* out, size_t count);
// Partials (stride kPartialStride, 2 fields) of r . e and e . (A e).
hipError_t launch_energy_dots(hipStream_t s, int n, const float* r, const float* e, double* partials, int* nblocks);
hipError_t launch_add_scaled_f32(hipStream_t s, const float* a, float scale, const float* b, float* out, size_t count);
// out = ((x - ps->x_meanf) - prev) * scale: the gauged correction of a solve.
hipError_t launch_correction(hipStream_t s, const float* x, const float* prev, const ProjState* ps, float scale,
                             float* out, size_t count);
// The learned start decided on the device (no host round trip):
// line search alpha = (r . e) / (e . A e) from launch_energy_dots' partials
// into ps->learned_alpha (0 unless e . A e > 0 and the ratio is finite);
hipError_t launch_line_search(hipStream_t s, const double* partials, int nblocks, ProjState* ps);
// out = x + float(ps->learned_alpha) * e (out may be x);
hipError_t launch_axpy_alpha(hipStream_t s, const float* x, const ProjState* ps, const float* e, float* out,
                             size_t count);
// the judge: the learned record wins when its iterate is finite and its
// compatible relative L2 residual is below the plain record's; then it is
// copied over the plain one; plain->learned records the decision;
hipError_t launch_judge(hipStream_t s, ProjState* plain, const ProjState* learned, double count);
// The judge's two residual checks (launch_residual + Finalize::Residual for
// each record, the same sums bit for bit) and launch_judge, in two launches;
// the partial buffers must differ.
hipError_t launch_judge_pair(hipStream_t s, int n, const float* b, const float* x_plain, ProjState* plain,
                             double* partials_plain, const float* x_learned, ProjState* learned,
                             double* partials_learned);
// dst = src where ps->learned == 1.
hipError_t launch_copy_if_learned(hipStream_t s, const float* src, float* dst, const ProjState* ps, size_t count);
// field[i] += sigma * noise_normal(seed, stream, i) for i < count.
hipError_t launch_add_noise(hipStream_t s, float* field, size_t count, float sigma, uint64_t seed, uint32_t stream);
hipError_t launch_fill_f64(hipStream_t s, double* p, size_t count, double value);
hipError_t launch_set_u32(hipStream_t s, uint32_t* p, uint32_t value);
hipError_t launch_emit(hipStream_t s, const EmitArgs& a, uint32_t* rng, float* noise, float* rho, float* u,
                       float* v, float* w);
hipError_t launch_forces(hipStream_t s, int n, const float* rho, float* u, float* v, float* w, float buoyancy,
                         float dt, float split);
// fields = number of consecutive fields starting at a.first_field.
hipError_t launch_advect_sl(hipStream_t s, const AdvArgs& a, int fields, bool hat);
hipError_t launch_advect_mc(hipStream_t s, const AdvArgs& a, int fields, bool hat);
// The reducing launchers write one partial record per block into `partials`
// and report the block count; launch_finalize folds them into ps / ds.
enum class Finalize { Rhs, XSum, Residual, Final, Diag };

// Divergence of (u, v, w) with zero wall faces; b = -div * scale.
// Partials -> Finalize::Rhs (sum b, max |div|, max b, min b, non-finite).
hipError_t launch_divergence_rhs(hipStream_t s, int n, float h, float scale, const float* u, const float* v,
                                 const float* w, float* b, double* partials, int* nblocks);
// One full red-black sweep (both colours) from xin into xout; mode 1 adds the
// FP64 residual of xin (partials as launch_residual), mode 2 the sum of xout
// (Finalize::XSum), mode 3 the partials of launch_check (the residual of xin
// without the right-hand side's sums, and the sum of xout). A set *done (may
// be null) skips the launch.
hipError_t launch_sor_rb(hipStream_t s, int n, float omega, float floor, const float* b, const float* xin,
                         float* xout, const ProjState* ps, const int* done, int mode, double* partials, int* nblocks);
// End of a tolerance chunk: fold the residual, count `chunk` sweeps, raise
// done, and copy the record to ring[chunks % ring_size] (device-visible
// pinned host memory), where `chunks` counts the record's checks since its
// last reset.
hipError_t launch_check(hipStream_t s, const double* partials, int nblocks, int n, ProjState* ps, double tol,
                        int chunk, ProjState* ring, int ring_size);
hipError_t launch_reset_status(hipStream_t s, ProjState* ps);
// FP64 residual check of b' - A x -> Finalize::Residual (res2, res_inf,
// bnorm2, x_sum, non-finite x).
hipError_t launch_residual(hipStream_t s, int n, const float* b, const float* x, const ProjState* ps,
                           double* partials, int* nblocks);
// FP64 residual of an FP64 iterate (iterative refinement): writes float(r) to
// r32 -> Finalize::Residual.
hipError_t launch_residual64(hipStream_t s, int n, const float* b, const double* x, const ProjState* ps, float* r32,
                             double* partials, int* nblocks);
// The same for an FP32 iterate: float(r) to r32, and partials equal to
// launch_residual's bit for bit -> Finalize::Residual; with xcopy (not x)
// also a copy of x.
hipError_t launch_residual_field(hipStream_t s, int n, const float* b, const float* x, const ProjState* ps, float* r32,
                                 double* partials, int* nblocks, float* xcopy = nullptr);
hipError_t launch_axpy_f64(hipStream_t s, double* x, const float* e, size_t count);  // x += e
hipError_t launch_widen(hipStream_t s, double* x, const float* e, size_t count);     // x = e
hipError_t launch_narrow(hipStream_t s, float* x, const double* e, size_t count);    // x = float(e)
// Zero-mean gauge + gradient subtraction (in -> out faces) + walls, fused
// with the final residual and the max divergence afterwards -> Finalize::Final.
hipError_t launch_gradient(hipStream_t s, int n, float h, float dt, const float* uin, const float* vin,
                           const float* win, const float* p, float* p_out, const float* b, float* uout,
                           float* vout, float* wout, const ProjState* ps, double* partials, int* nblocks);
// Kinetic energy, max |velocity|, max |div| -> Finalize::Diag.
hipError_t launch_diagnostics(hipStream_t s, int n, float h, const float* u, const float* v, const float* w,
                              double* partials, int* nblocks);
hipError_t launch_finalize(hipStream_t s, Finalize mode, const double* partials, int nblocks, int n, ProjState* ps,
                           DiagState* ds);

// Upper bound of partial records any reduction above writes (per element of
// `partials`, 8 doubles per block).
size_t partials_capacity(int n);

// ---- batches of scenes (one launch per stage over every scene) ----

// The work list of a batched launch: scene s owns items [prefix[s],
// prefix[s + 1]) (device arrays of scenes + 1 and items entries).
struct BatchList {
    const int* prefix;
    const uint16_t* owner;  // owner[w]: the scene of item w (one load instead of a search)
    int scenes;
    int items;
};
// Work-list families: element-wise over faces (forces, advection, emitter
// patch points use their own counts), over cells (fills), the cell-stencil
// decomposition of the reducing kernels (fixed: it defines the reduction
// partials), and the red-black sweep's tiles (fixed).
enum class BatchFamily { Elementwise, Cells, Stencil, Tiles };
// The sweep items of the scenes still solving (device arrays): item w is tile
// local[w] of scene owner[w], for w < *count. Rebuilt after every check, so a
// converged scene costs its sweeps nothing.
struct ActiveList {
    uint16_t* owner;
    int* local;
    int* count;
};
int batch_items(BatchFamily f, int n, int per_thread);

struct EmitSlot {
    EmitArgs a;
    uint32_t* rng;
    float* noise;
    float *rho, *u, *v, *w;
    int active;
};
struct ForceArgs {
    int n;
    const float* rho;
    float *u, *v, *w;
    float buoyancy, dt, split;
};
// Arguments of the cell-stencil kernels: divergence (u, v, w -> b), residual
// (b, p, ps), gradient (u, v, w, p -> p_out, uo, vo, wo).
struct CellArgs {
    int n;
    float h, scale, dt;
    const float *u, *v, *w;
    const float* p;
    float* p_out;
    float* b;
    float *uo, *vo, *wo;
    ProjState* ps;
    double* partials;
};
enum class CellOp { Divergence, Residual, Gradient };
struct RbSlot {
    int n;
    const float* b;
    const float* xin;
    float* xout;
    ProjState* ps;
    double* partials;
    RbGeom geo;  // rb_geom(n)
};
struct FinArgs {
    const double* partials;
    int nblocks;
    double count;
    ProjState* ps;
};
struct FillSlot {
    float* p;
    size_t count;
    const ProjState* ps;
};

// `grid` = blocks of the launch (<= 0 or >= items: one block per item).
hipError_t batch_emit(hipStream_t s, const EmitSlot* t, int scenes, BatchList apply, int grid);
hipError_t batch_forces(hipStream_t s, const ForceArgs* t, BatchList L, int grid);
hipError_t batch_advect(hipStream_t s, bool mc, int field, bool hat, const AdvArgs* t, BatchList L, int grid);
hipError_t batch_cell(hipStream_t s, CellOp op, const CellArgs* t, BatchList L, int grid);
// Sweeps of every scene of L, or (act non-null) of the active list only.
hipError_t batch_sor(hipStream_t s, float omega, float floor, bool vec, bool reg, int mode, const ActiveList* act,
                     const RbSlot* t, BatchList L, int grid);
// Builds the active list of the tile list L from the records' done flags.
hipError_t batch_active(hipStream_t s, const FinArgs* t, BatchList L, ActiveList act);
hipError_t batch_finalize(hipStream_t s, Finalize mode, const FinArgs* t, int scenes);
// Chunk check per scene; rows go to ring[(chunks % ring_size) * scenes + s].
hipError_t batch_check(hipStream_t s, const FinArgs* t, int scenes, double tol, int chunk, ProjState* ring,
                       int ring_size);
hipError_t batch_start(hipStream_t s, const FinArgs* t, int scenes, double tol, int budget, bool warm, ProjState* ring,
                       int ring_size);
hipError_t batch_zero(hipStream_t s, const FillSlot* t, BatchList L, int grid, bool only_zero_start);
hipError_t batch_stamp(hipStream_t s, long long* dst);
hipError_t batch_gather(hipStream_t s, ProjState* const* records, int scenes, int per, ProjState* out);

}  // namespace vgx
