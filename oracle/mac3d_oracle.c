/*
 * mac3d_oracle.c - CPU restatement of vapor's MAC 3D smoke step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path (see the header).
 * Nothing in the product calls into this file. Built twice: serially (the
 * checker), and with -fopenmp for the benchmark's all-cores CPU leg, where the
 * sample loops of advection and the colour passes of SOR split their planes
 * over threads (every such update is independent, so both builds give the
 * same bits; sums stay sequential).
 *
 * Build with -O2 -ffp-contract=off on x86-64 (SSE float, no FMA) so every
 * float expression rounds exactly as the reference's does. Each routine cites
 * the reference routine it restates (paths relative to vapor's src/); the
 * operation order inside every expression follows the reference.
 *
 * Parity status of each part (how it is pinned; see tests/):
 *   samplers, SL/MC advection, walls, divergence, gradient, SOR, projection,
 *   emitter (sphere), buoyancy, decay, whole step: pinned bit-exactly against
 *   golden fixtures written by the reference itself (tests/golden), plus the
 *   reference's published step baselines (SimulationTests.cpp:17-34) and its
 *   closed-form tests (AdvectionTests.cpp, NumericsTests.cpp).
 *   Extensions with no reference code (capped-Gaussian emitter, second
 *   emitter, L2 tolerance stop, warm start, FP64 iterative refinement) are
 *   pinned by closed forms and invariants in tests/test_oracle_extensions.py.
 */
#include "mac3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* std::clamp / std::min / std::max with the reference's comparison order. */
static inline float clampf_(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
static inline int clampi_(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
static inline float minf_(float a, float b) { return (b < a) ? b : a; }
static inline float maxf_(float a, float b) { return (a < b) ? b : a; }
static inline int mini_(int a, int b) { return (b < a) ? b : a; }
static inline int maxi_(int a, int b) { return (a < b) ? b : a; }
static inline double maxd_(double a, double b) { return (a < b) ? b : a; }

/* Index maps, MacGridDomain3D.hpp:22-31. */
static inline size_t idC(int n, int x, int y, int z) { return (size_t)x + (size_t)n * ((size_t)y + (size_t)n * (size_t)z); }
static inline size_t idX(int n, int x, int y, int z) { return (size_t)x + (size_t)(n + 1) * ((size_t)y + (size_t)n * (size_t)z); }
static inline size_t idY(int n, int x, int y, int z) { return (size_t)x + (size_t)n * ((size_t)y + (size_t)(n + 1) * (size_t)z); }
static inline size_t idZ(int n, int x, int y, int z) { return idC(n, x, y, z); }
static inline size_t idA(int n, int axis, int x, int y, int z) {
    return axis == 0 ? idX(n, x, y, z) : axis == 1 ? idY(n, x, y, z) : idZ(n, x, y, z);
}
/* MacGridDomain3D::cellSize, MacGridDomain3D.hpp:18. */
static inline float cell_size(int n, float box) { return box / (float)n; }

/* ---------------------------------------------------------------- sampling */

/* MacGridSampler3D::sampleCell / sampleV{x,y,z}Field, MacGridSampler3D.cpp:5-71.
 * On the staggered axis the coordinate is p/h clamped to [0, n] and the lower
 * index is capped at n-1; on the others it is p/h - 0.5 clamped to [0, n-1]
 * and the upper index is capped at n-1. */
float orc_sample(int n, float box, int axis, const float* f, float px, float py, float pz) {
    const float h = cell_size(n, box);
    const float p[3] = {px, py, pz};
    float g[3];
    int i0[3], i1[3];
    for (int a = 0; a < 3; ++a) {
        if (a == axis) {
            g[a] = clampf_(p[a] / h, 0.f, (float)n);
            i0[a] = mini_((int)g[a], n - 1);
            i1[a] = i0[a] + 1;
        } else {
            g[a] = clampf_(p[a] / h - 0.5f, 0.f, (float)(n - 1));
            i0[a] = (int)g[a];
            i1[a] = mini_(i0[a] + 1, n - 1);
        }
    }
    const float tx = g[0] - (float)i0[0], ty = g[1] - (float)i0[1], tz = g[2] - (float)i0[2];
#define AT(X, Y, Z) f[axis < 0 ? idC(n, X, Y, Z) : idA(n, axis, X, Y, Z)]
    return AT(i0[0], i0[1], i0[2]) * (1 - tx) * (1 - ty) * (1 - tz) +
           AT(i1[0], i0[1], i0[2]) * tx * (1 - ty) * (1 - tz) +
           AT(i0[0], i1[1], i0[2]) * (1 - tx) * ty * (1 - tz) +
           AT(i1[0], i1[1], i0[2]) * tx * ty * (1 - tz) +
           AT(i0[0], i0[1], i1[2]) * (1 - tx) * (1 - ty) * tz +
           AT(i1[0], i0[1], i1[2]) * tx * (1 - ty) * tz +
           AT(i0[0], i1[1], i1[2]) * (1 - tx) * ty * tz +
           AT(i1[0], i1[1], i1[2]) * tx * ty * tz;
#undef AT
}

/* MacGridSampler3D::velocity, MacGridSampler3D.cpp:73-75. */
void orc_velocity(int n, float box, const float* u, const float* v, const float* w,
                  float px, float py, float pz, float out[3]) {
    out[0] = orc_sample(n, box, 0, u, px, py, pz);
    out[1] = orc_sample(n, box, 1, v, px, py, pz);
    out[2] = orc_sample(n, box, 2, w, px, py, pz);
}

/* --------------------------------------------------------------- advection */

/* Sample position of entry (x, y, z) of a field: cell centre
 * (CpuAdvectionKernels.cpp:25) or face centre (:91, :154, :217). */
static inline void sample_point(int axis, float dx, int x, int y, int z, float p[3]) {
    p[0] = axis == 0 ? (float)x * dx : ((float)x + .5f) * dx;
    p[1] = axis == 1 ? (float)y * dx : ((float)y + .5f) * dx;
    p[2] = axis == 2 ? (float)z * dx : ((float)z + .5f) * dx;
}

/* One advected field (cell axis -1 or face axis 0..2) with the reference's
 * three passes: forward trace (SL), backward trace of the forward field, then
 * the MacCormack correction with its fallback and limiter.
 * CpuAdvectionKernels.cpp:18-80 (scalar) and :82-269 (faces). */
static void advect_field(int n, float box, int axis, const float* src, const float* u,
                         const float* v, const float* w, float dt, float cflMc, int mc,
                         float* out, float* fwd, float* back) {
    const float dx = cell_size(n, box);
    const int ex = n + (axis == 0), ey = n + (axis == 1), ez = n + (axis == 2);
    const size_t count = (size_t)ex * ey * ez;
#define IDX(X, Y, Z) (axis < 0 ? idC(n, X, Y, Z) : idA(n, axis, X, Y, Z))
    /* Every sample is independent: the OpenMP build splits the planes. */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < ez; ++z)
        for (int y = 0; y < ey; ++y)
            for (int x = 0; x < ex; ++x) {
                float p[3], vel[3];
                sample_point(axis, dx, x, y, z, p);
                orc_velocity(n, box, u, v, w, p[0], p[1], p[2], vel);
                fwd[IDX(x, y, z)] = orc_sample(n, box, axis, src, p[0] - vel[0] * dt,
                                               p[1] - vel[1] * dt, p[2] - vel[2] * dt);
            }
    if (!mc) {
        memcpy(out, fwd, count * sizeof(float));
        return;
    }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < ez; ++z)
        for (int y = 0; y < ey; ++y)
            for (int x = 0; x < ex; ++x) {
                float p[3], vel[3];
                sample_point(axis, dx, x, y, z, p);
                orc_velocity(n, box, u, v, w, p[0], p[1], p[2], vel);
                back[IDX(x, y, z)] = orc_sample(n, box, axis, fwd, p[0] + vel[0] * dt,
                                                p[1] + vel[1] * dt, p[2] + vel[2] * dt);
            }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < ez; ++z)
        for (int y = 0; y < ey; ++y)
            for (int x = 0; x < ex; ++x) {
                float p[3], vel[3];
                const size_t i = IDX(x, y, z);
                sample_point(axis, dx, x, y, z, p);
                orc_velocity(n, box, u, v, w, p[0], p[1], p[2], vel);
                const float pf[3] = {p[0] - vel[0] * dt, p[1] - vel[1] * dt, p[2] - vel[2] * dt};
                const float pb[3] = {p[0] + vel[0] * dt, p[1] + vel[1] * dt, p[2] + vel[2] * dt};
                const float cfl = sqrtf(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) * dt / dx;
                if (pf[0] < 0 || pf[0] > box || pf[1] < 0 || pf[1] > box || pf[2] < 0 || pf[2] > box ||
                    pb[0] < 0 || pb[0] > box || pb[1] < 0 || pb[1] > box || pb[2] < 0 || pb[2] > box ||
                    cfl > cflMc) {
                    out[i] = fwd[i];
                    continue;
                }
                const float corrected = fwd[i] + 0.5f * (src[i] - back[i]);
                int g[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    hi[a] = a == axis ? n : n - 1;
                    g[a] = clampi_(a == axis ? (int)(pf[a] / dx) : (int)(pf[a] / dx - 0.5f), 0, hi[a]);
                }
                float fMin = src[IDX(g[0], g[1], g[2])], fMax = fMin;
                for (int dz = 0; dz <= 1; ++dz)
                    for (int dy = 0; dy <= 1; ++dy)
                        for (int dx2 = 0; dx2 <= 1; ++dx2) {
                            const float val = src[IDX(clampi_(g[0] + dx2, 0, hi[0]), clampi_(g[1] + dy, 0, hi[1]),
                                                      clampi_(g[2] + dz, 0, hi[2]))];
                            fMin = minf_(fMin, val);
                            fMax = maxf_(fMax, val);
                        }
                out[i] = clampf_(corrected, fMin, fMax);
            }
#undef IDX
}

/* Advection contract, AdvectionContext.cpp:15-30 (sizes are implied here). */
static int valid_advection(float dt, float cfl) {
    return isfinite(dt) && dt >= 0.f && isfinite(cfl) && cfl > 0.f;
}

int orc_advect_scalar(int n, float box, const float* u, const float* v, const float* w,
                      float dt, float cfl, int maccormack, float* field, float* scratch) {
    if (!valid_advection(dt, cfl)) return -1;
    const size_t cells = (size_t)n * n * n;
    float* fwd = scratch;
    float* back = scratch + cells;
    float* result = (float*)malloc(cells * sizeof(float));
    advect_field(n, box, -1, field, u, v, w, dt, cfl, maccormack, result, fwd, back);
    memcpy(field, result, cells * sizeof(float));
    free(result);
    return 0;
}

int orc_advect_face(int n, float box, int axis, const float* src, const float* u, const float* v,
                    const float* w, float dt, float cfl, int maccormack, float* out, float* scratch) {
    if (!valid_advection(dt, cfl) || axis < 0 || axis > 2) return -1;
    const size_t faces = (size_t)(n + 1) * n * n;
    advect_field(n, box, axis, src, u, v, w, dt, cfl, maccormack, out, scratch, scratch + faces);
    return 0;
}

/* --------------------------------------------------------------- operators */

/* MacGridOperators3D::enforceClosedBoundaries, MacGridOperators3D.cpp:16-28. */
void orc_enforce_walls(int n, float* u, float* v, float* w) {
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y) u[idX(n, 0, y, z)] = u[idX(n, n, y, z)] = 0.f;
    for (int z = 0; z < n; ++z)
        for (int x = 0; x < n; ++x) v[idY(n, x, 0, z)] = v[idY(n, x, n, z)] = 0.f;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) w[idZ(n, x, y, 0)] = w[idZ(n, x, y, n)] = 0.f;
}

/* divergenceAt, MacGridOperators3D.cpp:8-13. */
static inline float divergence_at(int n, float h, const float* u, const float* v, const float* w,
                                  int x, int y, int z) {
    return (u[idX(n, x + 1, y, z)] - u[idX(n, x, y, z)] + v[idY(n, x, y + 1, z)] - v[idY(n, x, y, z)] +
            w[idZ(n, x, y, z + 1)] - w[idZ(n, x, y, z)]) / h;
}

/* MacGridOperators3D::computeDivergence, MacGridOperators3D.cpp:30-38. */
void orc_divergence(int n, float box, const float* u, const float* v, const float* w, float* div) {
    const float h = cell_size(n, box);
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) div[idC(n, x, y, z)] = divergence_at(n, h, u, v, w, x, y, z);
}

/* MacGridOperators3D::maxDivergence, MacGridOperators3D.cpp:40-49. */
int orc_max_divergence(int n, float box, const float* u, const float* v, const float* w, float* out) {
    const float h = cell_size(n, box);
    float result = 0.f;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float value = divergence_at(n, h, u, v, w, x, y, z);
                if (!isfinite(value)) return -1;
                result = maxf_(result, fabsf(value));
            }
    *out = result;
    return 0;
}

/* MacGridOperators3D::applyPressureGradient, MacGridOperators3D.cpp:51-67. */
void orc_apply_gradient(int n, float box, const float* p, float dt, float* u, float* v, float* w) {
    const float hInv = 1.f / cell_size(n, box);
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 1; x < n; ++x)
                u[idX(n, x, y, z)] -= dt * (p[idC(n, x, y, z)] - p[idC(n, x - 1, y, z)]) * hInv;
    for (int z = 0; z < n; ++z)
        for (int y = 1; y < n; ++y)
            for (int x = 0; x < n; ++x)
                v[idY(n, x, y, z)] -= dt * (p[idC(n, x, y, z)] - p[idC(n, x, y - 1, z)]) * hInv;
    for (int z = 1; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                w[idZ(n, x, y, z)] -= dt * (p[idC(n, x, y, z)] - p[idC(n, x, y, z - 1)]) * hInv;
}

/* ------------------------------------------------------------- linear solve */

/* Neumann matrix of CpuPressureSolver3D::prepareSystem (CpuPressureSolver3D.cpp:
 * 9-23): diagonal = number of in-domain neighbours, off-diagonals -1. */
static inline float diag_at(int n, int x, int y, int z) {
    return (float)((x > 0) + (x + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
}

/* A x at one row in FP64, StencilLinearSystem.cpp:60-67 (same term order). */
static inline double ax_row(int n, const float* x, int i, int j, int k) {
    const size_t idx = idC(n, i, j, k);
    const size_t nn = (size_t)n * n;
    double ax = (double)diag_at(n, i, j, k) * x[idx];
    if (i > 0) ax += (double)(-1.f) * x[idx - 1];
    if (i + 1 < n) ax += (double)(-1.f) * x[idx + 1];
    if (j > 0) ax += (double)(-1.f) * x[idx - n];
    if (j + 1 < n) ax += (double)(-1.f) * x[idx + n];
    if (k > 0) ax += (double)(-1.f) * x[idx - nn];
    if (k + 1 < n) ax += (double)(-1.f) * x[idx + nn];
    return ax;
}

static int all_finite(size_t count, const float* a) {
    for (size_t i = 0; i < count; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* StencilLinearSystem::residualInfinityNorm, StencilLinearSystem.cpp:53-70. */
double orc_residual_inf(int n, const float* b, const float* x) {
    double result = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const size_t idx = idC(n, i, j, k);
                result = maxd_(result, fabs((double)b[idx] - ax_row(n, x, i, j, k)));
            }
    return result;
}

/* Extension: L2 norm of the compatible part of b - Ax (the residual minus its
 * mean). Constants are the Neumann matrix's null space, so the residual's
 * mean is out of reach of any iterate; the solve is measured on the rest. */
double orc_residual_l2(int n, const float* b, const float* x) {
    double sum = 0.0, sum2 = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double r = (double)b[idC(n, i, j, k)] - ax_row(n, x, i, j, k);
                sum += r;
                sum2 += r * r;
            }
    const double count = (double)n * n * n;
    const double v = sum2 - sum * sum / count;
    return sqrt(v > 0.0 ? v : 0.0);
}

/* L2 norm of the compatible (mean-free) part of b. */
double orc_norm_l2(size_t count, const float* b) {
    double sum = 0.0, sum2 = 0.0;
    for (size_t i = 0; i < count; ++i) {
        sum += (double)b[i];
        sum2 += (double)b[i] * b[i];
    }
    const double v = sum2 - sum * sum / (double)count;
    return sqrt(v > 0.0 ? v : 0.0);
}

/* One red-black sweep, SOR.cpp:36-49: parity 0 updates (i+j+k) even first. */
void orc_sor_sweep(int n, float omega, float floor, const float* b, float* x) {
    const size_t nn = (size_t)n * n;
    for (int parity = 0; parity < 2; ++parity)
        /* cells of one colour read only the other colour: the OpenMP build
           splits the planes of a colour pass with the same results */
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = (j + k + parity) & 1; i < n; i += 2) {
                    const size_t idx = idC(n, i, j, k);
                    const float d = diag_at(n, i, j, k);
                    if (d == 0.f) continue;
                    float offDiagonal = 0.f;
                    if (i > 0) offDiagonal += -1.f * x[idx - 1];
                    if (i + 1 < n) offDiagonal += -1.f * x[idx + 1];
                    if (j > 0) offDiagonal += -1.f * x[idx - n];
                    if (j + 1 < n) offDiagonal += -1.f * x[idx + n];
                    if (k > 0) offDiagonal += -1.f * x[idx - nn];
                    if (k + 1 < n) offDiagonal += -1.f * x[idx + nn];
                    const float diagonal = maxf_(d, floor);
                    x[idx] += omega * (b[idx] - offDiagonal - d * x[idx]) / diagonal;
                }
}

/* SOR::solve, SOR.cpp:22-64. */
int orc_sor_solve(int n, float omega, float floor, const float* b, float* x,
                  const vgx_projection_options* o, vgx_projection_result* result) {
    const size_t cells = (size_t)n * n * n;
    if (o->max_iterations < 0 || o->residual_check_interval <= 0 || !isfinite(o->absolute_tolerance) ||
        o->absolute_tolerance < 0.0 || !isfinite(o->relative_tolerance) || o->relative_tolerance < 0.0)
        return -1;
    if (!all_finite(cells, b) || !all_finite(cells, x)) return -1;
    result->iterations = 0;
    result->initial_residual = result->final_residual = orc_residual_inf(n, b, x);
    const double target = maxd_(o->absolute_tolerance, o->relative_tolerance * result->initial_residual);
    result->converged = result->final_residual <= target;
    if (result->converged && !o->fixed_iterations) return 0;
    for (int iter = 0; iter < o->max_iterations; ++iter) {
        orc_sor_sweep(n, omega, floor, b, x);
        result->iterations = iter + 1;
        if (result->iterations % o->residual_check_interval == 0 || result->iterations == o->max_iterations) {
            result->final_residual = orc_residual_inf(n, b, x);
            result->converged = result->final_residual <= target;
            if (result->converged && !o->fixed_iterations) break;
        }
    }
    return result->iterations;
}

/* ---------------------------------------------- extension: L2 solve policy */

/* FP32 SOR to a compatible relative L2 residual <= tol, in chunks of
 * `check` sweeps rounded up to an even count. The residual is checked in FP64
 * on the iterate each chunk's last sweep starts from; the solve stops after the
 * first chunk whose checked residual meets the tolerance, or once the budget,
 * rounded up to whole chunks, is spent. (The GPU fuses the check into that
 * last sweep; see vgx_api.hip's l2_solve.) Returns the sweeps used; `rel` is
 * the last checked residual. */
static int l2_sor(int n, float omega, float floor, const float* b, float* x, double bnorm, double tol,
                  int max_sweeps, int check, double* rel, int* checks) {
    int sweeps = 0;
    *rel = orc_residual_l2(n, b, x) / bnorm;
    ++*checks;
    if (*rel <= tol || max_sweeps <= 0) return 0;
    const int chunk = check + (check & 1) < 2 ? 2 : check + (check & 1);
    const int chunks = (max_sweeps + chunk - 1) / chunk;
    for (int c = 0; c < chunks; ++c) {
        for (int s = 0; s + 1 < chunk; ++s) orc_sor_sweep(n, omega, floor, b, x);
        *rel = orc_residual_l2(n, b, x) / bnorm;
        ++*checks;
        orc_sor_sweep(n, omega, floor, b, x);
        sweeps += chunk;
        if (*rel <= tol) break;
    }
    return sweeps;
}

/* Inner target of one refinement correction: an equal share of the reduction
 * still missing over the corrections left, but never more than 1e-4 relative
 * from FP32 SOR (the smooth correction equation's FP32 rounding floor is near
 * that at 128^3 and rises with the grid). */
#define ORC_IR_INNER_FLOOR 1e-4

/* FP64 iterative refinement around FP32 SOR: x64 += e with A e = r32, where
 * r32 = float(b - A x64) is formed in FP64. */
static int l2_refine(int n, float omega, float floor, const float* b, float* x, double bnorm, double tol,
                     int max_sweeps, int check, int max_steps, vgx_projection_result* res) {
    const size_t cells = (size_t)n * n * n;
    double* x64 = (double*)malloc(cells * sizeof(double));
    double* r64 = (double*)malloc(cells * sizeof(double));
    float* r32 = (float*)malloc(cells * sizeof(float));
    float* e32 = (float*)malloc(cells * sizeof(float));
    for (size_t i = 0; i < cells; ++i) x64[i] = x[i];
    int sweeps = 0;
    double rel = 0.0;
    for (int step = 0;; ++step) {
        /* r = b - A x64 in FP64 (Neumann matrix, same term order as ax_row). */
        double sum = 0.0;
        const size_t nn = (size_t)n * n;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const size_t idx = idC(n, i, j, k);
                    double ax = (double)diag_at(n, i, j, k) * x64[idx];
                    if (i > 0) ax -= x64[idx - 1];
                    if (i + 1 < n) ax -= x64[idx + 1];
                    if (j > 0) ax -= x64[idx - n];
                    if (j + 1 < n) ax -= x64[idx + n];
                    if (k > 0) ax -= x64[idx - nn];
                    if (k + 1 < n) ax -= x64[idx + nn];
                    r64[idx] = (double)b[idx] - ax;
                    sum += r64[idx] * r64[idx];
                }
        {
            double rsum = 0.0;
            for (size_t i = 0; i < cells; ++i) rsum += r64[i];
            const double v = sum - rsum * rsum / (double)cells;
            rel = sqrt(v > 0.0 ? v : 0.0) / bnorm;
            /* The correction equation's right-hand side: float(r), with the
             * float of its mean removed (as the projection treats b). */
            const float meanf = (float)(rsum / (double)cells);
            for (size_t i = 0; i < cells; ++i) r32[i] = (float)r64[i] - meanf;
        }
        ++res->residual_checks;
        if (step == 0) res->initial_residual_l2 = rel;
        if (rel <= tol || step >= max_steps || sweeps >= max_sweeps) break;
        for (size_t i = 0; i < cells; ++i) e32[i] = 0.f;
        const double rnorm = orc_norm_l2(cells, r32);
        /* An equal share (in decades) of the reduction still missing over the
         * corrections left, with a factor 2 of margin. */
        double inner_tol = 0.5 * pow(tol / rel, 1.0 / (double)(max_steps - step));
        if (inner_tol < ORC_IR_INNER_FLOOR) inner_tol = ORC_IR_INNER_FLOOR;
        double inner_rel;
        if (rnorm > 0.0)
            sweeps += l2_sor(n, omega, floor, r32, e32, rnorm, inner_tol, max_sweeps - sweeps, check,
                             &inner_rel, &res->residual_checks);
        for (size_t i = 0; i < cells; ++i) x64[i] += (double)e32[i];
        ++res->refinement_steps;
    }
    for (size_t i = 0; i < cells; ++i) x[i] = (float)x64[i];
    res->final_residual_l2 = rel;
    free(x64);
    free(r64);
    free(r32);
    free(e32);
    return sweeps;
}

/* -------------------------------------------------------------- projection */

/* CpuPressureSolver3D::project, CpuPressureSolver3D.cpp:25-60. With
 * options->use_l2 the fixed/inf-norm solve is replaced by the L2 tolerance
 * policy (extension); everything around the solve is the reference's. */
int orc_project(int n, float box, const vgx_projection_options* o, float* u, float* v, float* w,
                float* pressure, float* div, float* b, vgx_projection_result* result) {
    memset(result, 0, sizeof(*result));
    if (!isfinite(o->relaxation) || o->relaxation <= 0.f || o->relaxation >= 2.f) return -1;
    if (!isfinite(o->dt) || o->dt <= 0.f) return -1;
    const float h = cell_size(n, box);
    const float scale = h * h / o->dt;
    if (!isfinite(scale) || scale <= 0.f) return -1;
    const size_t cells = (size_t)n * n * n;
    orc_enforce_walls(n, u, v, w);
    orc_divergence(n, box, u, v, w, div);
    for (size_t i = 0; i < cells; ++i) {
        if (!isfinite(div[i])) return -1;
        result->divergence_before = maxf_(result->divergence_before, fabsf(div[i]));
        b[i] = -div[i] * scale;
    }
    double acc = 0.0;
    for (size_t i = 0; i < cells; ++i) acc += b[i];
    result->removed_rhs_mean = acc / (double)cells;
    for (size_t i = 0; i < cells; ++i) b[i] -= (float)result->removed_rhs_mean;
    if (!o->warm_start) memset(pressure, 0, cells * sizeof(float));
    result->warm_started = o->warm_start != 0;
    if (!o->use_l2) {
        if (orc_sor_solve(n, o->relaxation, o->diagonal_floor, b, pressure, o, result) < 0) return -1;
    } else {
        if (o->max_iterations < 0 || o->residual_check_interval <= 0 || !(o->l2_tolerance >= 0.0)) return -1;
        if (!all_finite(cells, pressure)) return -1;
        result->initial_residual = orc_residual_inf(n, b, pressure);
        const double bnorm = orc_norm_l2(cells, b);
        if (bnorm == 0.0) {
            memset(pressure, 0, cells * sizeof(float));
            result->initial_residual_l2 = result->final_residual_l2 = 0.0;
        } else if (o->refinement_steps > 0) {
            result->iterations = l2_refine(n, o->relaxation, o->diagonal_floor, b, pressure, bnorm,
                                           o->l2_tolerance, o->max_iterations, o->residual_check_interval,
                                           o->refinement_steps, result);
        } else {
            double rel0 = orc_residual_l2(n, b, pressure) / bnorm;
            double rel;
            result->iterations = l2_sor(n, o->relaxation, o->diagonal_floor, b, pressure, bnorm,
                                        o->l2_tolerance, o->max_iterations, o->residual_check_interval, &rel,
                                        &result->residual_checks);
            result->initial_residual_l2 = rel0;
            result->final_residual_l2 = rel;
        }
    }
    result->residual_available = 1;
    acc = 0.0;
    for (size_t i = 0; i < cells; ++i) acc += pressure[i];
    const double pressureMean = acc / (double)cells;
    for (size_t i = 0; i < cells; ++i) pressure[i] -= (float)pressureMean;
    result->final_residual = orc_residual_inf(n, b, pressure);
    if (o->use_l2) {
        const double bnorm = orc_norm_l2(cells, b);
        /* A refined solve reports its FP64 iterate's residual: the stored FP32
         * pressure cannot represent it more accurately. */
        if (!(o->refinement_steps > 0 && bnorm > 0.0))
            result->final_residual_l2 = bnorm > 0.0 ? orc_residual_l2(n, b, pressure) / bnorm : 0.0;
        result->converged = result->final_residual_l2 <= o->l2_tolerance;
    } else {
        const double target = maxd_(o->absolute_tolerance, o->relative_tolerance * result->initial_residual);
        result->converged = result->final_residual <= target;
    }
    orc_apply_gradient(n, box, pressure, o->dt, u, v, w);
    orc_enforce_walls(n, u, v, w);
    if (orc_max_divergence(n, box, u, v, w, &result->divergence_after) < 0) return -1;
    return 0;
}

/* ---------------------------------------------------------------- defaults */

/* MacGridParameters3D defaults, MacGridFluidSolver3D.hpp:14-36; extension
 * fields select the reference behaviour. */
void orc_params_default(vgx_params* p) {
    memset(p, 0, sizeof(*p));
    p->emitter_radius = 0.035f;
    p->emitter_center_x = 0.5f;
    p->emitter_center_y = 0.5f;
    p->emitter_center_z = 0.1f;
    p->emitter_kick_multiplier = 1.f;
    p->emitter_kick_base = 0.8f;
    p->emitter_kick_scale = 0.8f;
    p->stir_strength = 0.5f;
    p->random_seed = 0x9E3779B9u;
    p->sor_omega = 1.95f;
    p->buoyancy_split = 0.5f;
    p->cfl_mc = 3.0f;
    p->maccormack_smoke = 1;
    p->maccormack_vel = 1;
    p->smoke_decay = 0.06f;
    p->reflection = 1;
    p->buoyancy = 2.5f;
    p->source_strength = 1.f;
    p->project_iterations = 64;
    p->emitter_profile = VGX_EMITTER_SPHERE;
    p->solve_mode = VGX_SOLVE_FIXED;
    p->warm_start = VGX_WARM_ZERO;
    p->tolerance = 1e-5;
    p->max_sweeps = 20000;
    p->check_interval = 16;
}

/* ProjectionOptions / LinearSolveOptions defaults (ProjectionOptions.hpp:5-9,
 * LinearSolveOptions.hpp:4-10) with the MAC solver's diagonal floor 6. */
void orc_projection_options_default(vgx_projection_options* o) {
    memset(o, 0, sizeof(*o));
    o->dt = 1.f / 60.f;
    o->relaxation = 1.95f;
    o->max_iterations = 200;
    o->residual_check_interval = 8;
    o->absolute_tolerance = 1e-6;
    o->relative_tolerance = 1e-5;
    o->fixed_iterations = 0;
    o->diagonal_floor = 6.f;
    o->l2_tolerance = 1e-5;
}

/* -------------------------------------------------------------- simulation */

struct orc_sim {
    int n;
    float box;
    vgx_params params;
    uint32_t random_state;
    float *rho, *u, *v, *w, *div, *b;
    float* p[2]; /* pressure of projection slot 0 (last of a step) and slot 1 (first of a reflection step) */
    float *u0, *v0, *w0, *ut, *vt, *wt, *uh, *vh, *wh, *ud, *vd, *wd;
    float* scratch;
    vgx_projection_result last[2];
    int projections_this_step;
};

/* MacGridFluidSolver3D::randomUnit, MacGridFluidSolver3D.cpp:33-36. */
static inline float random_unit(uint32_t* state) {
    *state = *state * 1664525u + 1013904223u;
    return (float)(*state >> 8) * (1.0f / 16777216.0f);
}

/* e^v of the capped-Gaussian profile (extension; no reference code). The
 * definition is this fixed sequence of IEEE float operations, the same on the
 * device: Cody-Waite reduction v = k ln2 + r, k = floor(v log2(e) + 1/2),
 * with ln2 = C1 + C2 (C1 with 9 trailing zero bits, so k C1 is exact), the
 * Taylor series of e^r to r^7 on |r| <= ln2/2, and an exact scaling by 2^k.
 * Pinned against libm's double exp (tests/test_oracle_extensions.py). */
float orc_profile_exp(float v) {
    const float k = floorf(v * 1.44269504f + 0.5f);
    const float r = (v - k * 0.693145752f) - k * 1.42860677e-06f;
    float p = 0.000198412701f;
    p = 0.00138888892f + r * p;
    p = 0.00833333377f + r * p;
    p = 0.0416666679f + r * p;
    p = 0.166666672f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    const uint32_t bits = (uint32_t)((int)k + 127) << 23;
    float scale;
    memcpy(&scale, &bits, 4);
    return p * scale;
}

/* MacGridFluidSolver3D::emit, MacGridFluidSolver3D.cpp:140-165, for one
 * sphere. The Gaussian profile (extension) replaces "density = 1" with
 * density = max(density, min(1, 3 s exp(-3.5 (d/r)^2))). */
static void emit_sphere(int n, float box, const vgx_params* P, float cxf, float cyf, float czf, float radius,
                        float strength, float dt, uint32_t* rng, float* rho, float* u, float* v, float* w) {
    const float dx = cell_size(n, box);
    const float dtScale = dt * 60.0f;
    const float ecx = cxf * box, ecy = cyf * box, ecz = czf * box;
    const float erad = radius * box;
    const int emitR = (int)(erad / dx + 1);
    const int cx = (int)(ecx / dx), cy = (int)(ecy / dx), cz = (int)(ecz / dx);
    for (int z = maxi_(0, cz - emitR); z < mini_(n, cz + emitR); ++z)
        for (int y = maxi_(0, cy - emitR); y < mini_(n, cy + emitR); ++y)
            for (int x = maxi_(0, cx - emitR); x < mini_(n, cx + emitR); ++x) {
                const float px = ((float)x + .5f) * dx, py = ((float)y + .5f) * dx, pz = ((float)z + .5f) * dx;
                const float qx = (px - ecx) / erad, qy = (py - ecy) / erad, qz = (pz - ecz) / erad;
                const float q = qx * qx + qy * qy + qz * qz;
                if (q > 1.0f) continue;
                const size_t i = idC(n, x, y, z);
                if (P->emitter_profile == VGX_EMITTER_GAUSSIAN)
                    rho[i] = maxf_(rho[i], minf_(1.f, 3.f * strength * orc_profile_exp(-3.5f * q)));
                else
                    rho[i] = 1.0f;
                const float kick = P->emitter_kick_multiplier * (P->emitter_kick_base + P->emitter_kick_scale * strength);
                if (P->emitter_kick_multiplier > 0.f) {
                    w[idZ(n, x, y, z)] = maxf_(w[idZ(n, x, y, z)], kick);
                    w[idZ(n, x, y, z + 1)] = maxf_(w[idZ(n, x, y, z + 1)], kick);
                }
                const float j = (random_unit(rng) - 0.5f) * P->stir_strength * dtScale;
                u[idX(n, x, y, z)] += j;
                u[idX(n, x + 1, y, z)] += j;
                v[idY(n, x, y, z)] += j;
                v[idY(n, x, y + 1, z)] += j;
            }
}

uint32_t orc_emit(int n, float box, const vgx_params* P, float dt, uint32_t rng, float* rho, float* u,
                  float* v, float* w) {
    const float strength = P->source_strength;
    if (strength <= 0.f) return rng;
    emit_sphere(n, box, P, P->emitter_center_x, P->emitter_center_y, P->emitter_center_z, P->emitter_radius,
                strength, dt, &rng, rho, u, v, w);
    if (P->emitter2_enabled)
        emit_sphere(n, box, P, P->emitter2_center_x, P->emitter2_center_y, P->emitter2_center_z,
                    P->emitter2_radius, strength, dt, &rng, rho, u, v, w);
    return rng;
}

/* MacGridFluidSolver3D::applyBuoyancy, MacGridFluidSolver3D.cpp:167-181. */
static void apply_buoyancy(int n, const float* rho, float* w, float dtStep, float buoyancyVal, float split) {
    for (int z = 1; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float lower = buoyancyVal * rho[idC(n, x, y, z - 1)] * dtStep;
                const float upper = buoyancyVal * rho[idC(n, x, y, z)] * dtStep;
                w[idZ(n, x, y, z)] += split * lower;
                w[idZ(n, x, y, z)] += split * upper;
            }
}

orc_sim* orc_sim_create(int n, float box, const vgx_params* params) {
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->n = n;
    s->box = box;
    s->params = *params;
    const size_t cells = (size_t)n * n * n, faces = (size_t)(n + 1) * n * n;
    s->rho = (float*)calloc(cells, sizeof(float));
    s->p[0] = (float*)calloc(cells, sizeof(float));
    s->p[1] = (float*)calloc(cells, sizeof(float));
    s->div = (float*)calloc(cells, sizeof(float));
    s->b = (float*)calloc(cells, sizeof(float));
    float** faceFields[] = {&s->u, &s->v, &s->w, &s->u0, &s->v0, &s->w0, &s->ut, &s->vt,
                            &s->wt, &s->uh, &s->vh, &s->wh, &s->ud, &s->vd, &s->wd};
    for (size_t k = 0; k < sizeof(faceFields) / sizeof(faceFields[0]); ++k)
        *faceFields[k] = (float*)calloc(faces, sizeof(float));
    s->scratch = (float*)calloc(2 * faces, sizeof(float));
    orc_sim_reset(s);
    return s;
}

void orc_sim_destroy(orc_sim* s) {
    if (!s) return;
    float* all[] = {s->rho, s->p[0], s->p[1], s->div, s->b, s->u, s->v, s->w, s->u0, s->v0, s->w0,
                    s->ut, s->vt, s->wt, s->uh, s->vh, s->wh, s->ud, s->vd, s->wd, s->scratch};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]);
    free(s);
}

void orc_sim_set_params(orc_sim* s, const vgx_params* params) { s->params = *params; }

/* MacGridFluidSolver3D::resetState, MacGridFluidSolver3D.cpp:271-290. */
void orc_sim_reset(orc_sim* s) {
    const size_t cells = (size_t)s->n * s->n * s->n, faces = (size_t)(s->n + 1) * s->n * s->n;
    s->random_state = s->params.random_seed;
    memset(s->rho, 0, cells * sizeof(float));
    memset(s->p[0], 0, cells * sizeof(float));
    memset(s->p[1], 0, cells * sizeof(float));
    memset(s->u, 0, faces * sizeof(float));
    memset(s->v, 0, faces * sizeof(float));
    memset(s->w, 0, faces * sizeof(float));
    memset(s->last, 0, sizeof(s->last));
}

/* MacGridFluidSolver3D::project, MacGridFluidSolver3D.cpp:131-138, plus the
 * extension solve policy selected by the parameters. */
/* slot: 1 for the first projection of a reflection step, 0 otherwise. With
 * VGX_WARM_PREVIOUS each slot starts from its own pressure of the previous
 * step: the two projections of a reflection step solve for pressures of
 * roughly opposite sign (the reflected velocity carries the negated
 * divergence), so each is warm-started from its own history. */
static int sim_project(orc_sim* s, int slot, float dt, int iterations) {
    vgx_projection_options o;
    orc_projection_options_default(&o);
    o.dt = dt;
    o.relaxation = s->params.sor_omega;
    o.max_iterations = iterations;
    o.fixed_iterations = 1;
    o.diagonal_floor = 6.f;
    if (s->params.solve_mode == VGX_SOLVE_TOLERANCE) {
        o.use_l2 = 1;
        o.fixed_iterations = 0;
        o.l2_tolerance = s->params.tolerance;
        o.max_iterations = s->params.max_sweeps;
        o.residual_check_interval = s->params.check_interval;
        o.refinement_steps = s->params.refinement_steps;
    }
    o.warm_start = s->params.warm_start == VGX_WARM_PREVIOUS;
    vgx_projection_result r;
    const int rc = orc_project(s->n, s->box, &o, s->u, s->v, s->w, s->p[slot], s->div, s->b, &r);
    s->last[1] = s->last[0];
    s->last[0] = r;
    ++s->projections_this_step;
    return rc;
}

static void copy_faces(const orc_sim* s, float* du, float* dv, float* dw, const float* su, const float* sv,
                       const float* sw) {
    const size_t faces = (size_t)(s->n + 1) * s->n * s->n;
    memcpy(du, su, faces * sizeof(float));
    memcpy(dv, sv, faces * sizeof(float));
    memcpy(dw, sw, faces * sizeof(float));
}

/* step / stepNoReflect, MacGridFluidSolver3D.cpp:192-245, after the checks
 * of advance (:323-330). */
int orc_sim_advance(orc_sim* s, float dt) {
    const vgx_params* P = &s->params;
    const int n = s->n;
    const float box = s->box;
    if (!isfinite(P->emitter_kick_multiplier) || P->emitter_kick_multiplier < 0.f) return -1;
    const float cfl = P->cfl_mc;
    s->projections_this_step = 0;
    s->random_state = orc_emit(n, box, P, dt, s->random_state, s->rho, s->u, s->v, s->w);
    apply_buoyancy(n, s->rho, s->w, dt, P->buoyancy, P->buoyancy_split);
    orc_enforce_walls(n, s->u, s->v, s->w);
    copy_faces(s, s->u0, s->v0, s->w0, s->u, s->v, s->w);
    const float* flow0[3] = {s->u0, s->v0, s->w0};
    float* state[3] = {s->u, s->v, s->w};
    if (P->reflection) {
        const float halfDt = 0.5f * dt;
        const int halfIters = maxi_(1, P->project_iterations / 2);
        for (int a = 0; a < 3; ++a)
            if (orc_advect_face(n, box, a, flow0[a], s->u0, s->v0, s->w0, halfDt, cfl, P->maccormack_vel, state[a],
                                s->scratch) < 0)
                return -1;
        copy_faces(s, s->ut, s->vt, s->wt, s->u, s->v, s->w);
        if (sim_project(s, 1, halfDt, halfIters) < 0) return -1;
        copy_faces(s, s->ud, s->vd, s->wd, s->u, s->v, s->w);
        const size_t faces = (size_t)(n + 1) * n * n;
        for (size_t i = 0; i < faces; ++i) s->uh[i] = 2.f * s->u[i] - s->ut[i];
        for (size_t i = 0; i < faces; ++i) s->vh[i] = 2.f * s->v[i] - s->vt[i];
        for (size_t i = 0; i < faces; ++i) s->wh[i] = 2.f * s->w[i] - s->wt[i];
        const float* hat[3] = {s->uh, s->vh, s->wh};
        for (int a = 0; a < 3; ++a)
            if (orc_advect_face(n, box, a, hat[a], s->ud, s->vd, s->wd, halfDt, cfl, P->maccormack_vel, state[a],
                                s->scratch) < 0)
                return -1;
        if (sim_project(s, 0, halfDt, halfIters) < 0) return -1;
    } else {
        for (int a = 0; a < 3; ++a)
            if (orc_advect_face(n, box, a, flow0[a], s->u0, s->v0, s->w0, dt, cfl, P->maccormack_vel, state[a],
                                s->scratch) < 0)
                return -1;
        if (sim_project(s, 0, dt, P->project_iterations) < 0) return -1;
    }
    if (orc_advect_scalar(n, box, s->u, s->v, s->w, dt, cfl, P->maccormack_smoke, s->rho, s->scratch) < 0) return -1;
    /* applyDissipation, MacGridFluidSolver3D.cpp:183-190. */
    const float sFactor = clampf_(1.0f - P->smoke_decay * dt, 0.0f, 1.0f);
    const size_t cells = (size_t)n * n * n;
    for (size_t i = 0; i < cells; ++i) s->rho[i] *= sFactor;
    return 0;
}

float* orc_sim_field(orc_sim* s, int which) {
    switch (which) {
    case 0: return s->rho;
    case 1: return s->u;
    case 2: return s->v;
    case 3: return s->w;
    case 4: return s->p[0];
    case 5: return s->p[1];
    default: return NULL;
    }
}

uint32_t orc_sim_random_state(const orc_sim* s) { return s->random_state; }

void orc_sim_last_projection(const orc_sim* s, int index, vgx_projection_result* out) {
    *out = s->last[index == 1 && s->projections_this_step > 1 ? 1 : 0];
}

/* MacGridFluidSolver3D::kineticEnergy, MacGridFluidSolver3D.cpp:247-257. */
float orc_kinetic_energy(int n, float box, const float* u, const float* v, const float* w) {
    const float dx = cell_size(n, box);
    double ke = 0.0;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x <= n; ++x) ke += (double)u[idX(n, x, y, z)] * u[idX(n, x, y, z)];
    for (int z = 0; z < n; ++z)
        for (int y = 0; y <= n; ++y)
            for (int x = 0; x < n; ++x) ke += (double)v[idY(n, x, y, z)] * v[idY(n, x, y, z)];
    for (int z = 0; z <= n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) ke += (double)w[idZ(n, x, y, z)] * w[idZ(n, x, y, z)];
    return (float)(0.5 * ke * dx * dx * dx);
}

/* MacGridFluidSolver3D::maxVelocity, MacGridFluidSolver3D.cpp:259-265. */
float orc_max_velocity(int n, const float* u, const float* v, const float* w) {
    const size_t faces = (size_t)(n + 1) * n * n;
    float m = 0.f;
    for (size_t i = 0; i < faces; ++i) m = maxf_(m, fabsf(u[i]));
    for (size_t i = 0; i < faces; ++i) m = maxf_(m, fabsf(v[i]));
    for (size_t i = 0; i < faces; ++i) m = maxf_(m, fabsf(w[i]));
    return m;
}

void orc_sim_diagnostics(const orc_sim* s, vgx_diagnostics* out) {
    out->kinetic_energy = orc_kinetic_energy(s->n, s->box, s->u, s->v, s->w);
    out->max_velocity = orc_max_velocity(s->n, s->u, s->v, s->w);
    float md = 0.f;
    orc_max_divergence(s->n, s->box, s->u, s->v, s->w, &md);
    out->max_divergence = md;
}


/* ------------------------------------------------------------------
 * Extension: counter-based velocity noise (no reference code).
 * ------------------------------------------------------------------ */

static uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t* hi) {
    const uint64_t product = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(product >> 32);
    return (uint32_t)product;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { /* key schedule: bumped before every round after the first */
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        uint32_t hi0, hi1;
        const uint32_t lo0 = mulhilo32(0xD2511F53u, x[0], &hi0);
        const uint32_t lo1 = mulhilo32(0xCD9E8D57u, x[2], &hi1);
        const uint32_t y0 = hi1 ^ x[1] ^ k[0], y2 = hi0 ^ x[3] ^ k[1];
        x[0] = y0;
        x[1] = lo1;
        x[2] = y2;
        x[3] = lo0;
    }
    memcpy(out, x, sizeof x);
}

float orc_noise_log(float x) {
    uint32_t bits;
    memcpy(&bits, &x, 4);
    int exponent = (int)((bits >> 23) & 0xffu) - 127;
    const uint32_t mbits = (bits & 0x7fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &mbits, 4);
    if (m > 1.41421356f) { /* keep m in (sqrt(1/2), sqrt(2)] */
        m = m * 0.5f;
        exponent += 1;
    }
    /* ln m = 2 atanh(s), s = (m - 1)/(m + 1): 2 (s + s^3/3 + ... + s^11/11) */
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float poly = 0.0909090909f;
    poly = 0.111111111f + s2 * poly;
    poly = 0.142857143f + s2 * poly;
    poly = 0.2f + s2 * poly;
    poly = 0.333333333f + s2 * poly;
    poly = s2 * poly;
    const float twice = 2.0f * s;
    return (float)exponent * 0.693147182f + (twice + twice * poly);
}

float orc_noise_cos2pi(float t) {
    const float quarters = 4.0f * t;
    const int quadrant = (int)quarters;
    const float x = (quarters - (float)quadrant) * 1.57079637f;
    const float x2 = x * x;
    float c = 2.08767570e-09f; /* cos: Taylor to x^12 */
    c = -2.75573192e-07f + x2 * c;
    c = 2.48015873e-05f + x2 * c;
    c = -0.00138888889f + x2 * c;
    c = 0.0416666667f + x2 * c;
    c = -0.5f + x2 * c;
    c = 1.0f + x2 * c;
    float sn = 1.60590438e-10f; /* sin: Taylor to x^13 */
    sn = -2.50521084e-08f + x2 * sn;
    sn = 2.75573192e-06f + x2 * sn;
    sn = -0.000198412698f + x2 * sn;
    sn = 0.00833333333f + x2 * sn;
    sn = -0.166666667f + x2 * sn;
    sn = 1.0f + x2 * sn;
    sn = x * sn;
    switch (quadrant) { /* cos(q pi/2 + x) */
    case 0: return c;
    case 1: return -sn;
    case 2: return -c;
    default: return sn;
    }
}

float orc_noise_normal(uint64_t seed, uint32_t stream, uint32_t i) {
    const uint32_t ctr[4] = {i, stream, 0u, 0u};
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    const float u0 = (float)((w[0] >> 8) + 1u) * 5.96046448e-08f; /* (0, 1] */
    const float u1 = (float)(w[1] >> 8) * 5.96046448e-08f;        /* [0, 1) */
    const float radius = sqrtf(-2.0f * orc_noise_log(u0));
    return radius * orc_noise_cos2pi(u1);
}

void orc_add_velocity_noise(int n, float sigma, uint64_t seed, float* u, float* v, float* w) {
    const uint32_t faces = (uint32_t)(n + 1) * (uint32_t)n * (uint32_t)n;
    float* fields[3] = {u, v, w};
    for (uint32_t a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < faces; ++i) fields[a][i] += sigma * orc_noise_normal(seed, a, i);
}

void orc_sim_add_velocity_noise(orc_sim* sim, float sigma, uint64_t seed) {
    orc_add_velocity_noise(sim->n, sigma, seed, sim->u, sim->v, sim->w);
}
