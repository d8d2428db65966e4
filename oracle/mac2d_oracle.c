/*
 * mac2d_oracle.c - CPU restatement of vapor's two MAC 2D solvers.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the GPU path is compared with. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
 * the product library never does.
 *
 * Restated from (paths relative to the vapor source tree):
 *   MacGridFluidSolver2D ........ src/solvers/mac2d/MacGridFluidSolver2D.cpp:22-167
 *   SemiLagrangian2D ............ src/solvers/mac2d/SemiLagrangian2D.cpp:7-50
 *   CpuPressureSolver2D ......... src/solvers/mac2d/CpuPressureSolver2D.cpp:9-87
 *   MacGridOperators2D .......... src/solvers/mac2d/MacGridOperators2D.cpp:6-41
 *   ReflectionMacFluidSolver2D .. src/solvers/mac2d/ReflectionMacFluidSolver2D.hpp:19-425
 * in the reference's loop and operation order (sequential red-black sweeps,
 * in-place MacCormack limiter in scan order, LCG draws in scan order), with
 * std::lerp / std::clamp / std::min / std::max as libstdc++ 11 defines them.
 *
 * e^v and x^y: with o2_set_libm(1) the C library's expf / powf, exactly as
 * vapor calls them (so the oracle reproduces vapor's own outputs bit for bit,
 * which the golden fixtures check); with o2_set_libm(0) (the default) a fixed
 * sequence of IEEE double operations rounded once to float (o2_det_exp,
 * o2_det_log), the same sequence the GPU evaluates. The two differ by at most
 * one unit in the last place, which tests/test_oracle_mac2d.py checks.
 *
 * Built with -ffp-contract=off on plain x86-64 (no FMA contraction).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/vapor_mac2d.h"

#define O2_EXPORT __attribute__((visibility("default")))

enum { O2_OK = 0, O2_EINVAL = 1, O2_ERANGE = 2, O2_ERUNTIME = 3 };

static int g_libm = 0;
O2_EXPORT void o2_set_libm(int on) { g_libm = on != 0; }

/* ------------------------------------------------------------ arithmetic */

static double bits_d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t d_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static uint32_t f_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* e^t: t = k ln2 + r, k = floor(t/ln2 + 1/2), ln2 = 0.6931471803691238 +
 * 1.9082149292705877e-10, e^r by Taylor to r^13 (Horner), times 2^k as two
 * normal factors. */
O2_EXPORT double o2_det_exp(double t) {
    if (!(t == t)) return t;
    if (t > 709.78) return bits_d(0x7ff0000000000000ull);
    if (t < -745.2) return 0.0;
    const double k = floor(t * 1.4426950408889634 + 0.5);
    const double r = (t - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
    static const double inv_fact[] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                      1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
                                      1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        0.5,
                                      1.0,                1.0};
    double p = inv_fact[0];
    for (int i = 1; i < 14; ++i) p = inv_fact[i] + r * p;
    const int ki = (int)k;
    const int k1 = ki / 2, k2 = ki - k1;
    const double s1 = bits_d((uint64_t)(int64_t)(k1 + 1023) << 52);
    const double s2 = bits_d((uint64_t)(int64_t)(k2 + 1023) << 52);
    return (p * s1) * s2;
}

/* ln x for finite normal x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 s (1 + z/3 + ... + z^12/25), s = (m-1)/(m+1), z = s^2. */
O2_EXPORT double o2_det_log(double x) {
    const uint64_t b = d_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 25.0;
    for (int k = 23; k >= 1; k -= 2) p = 1.0 / (double)k + z * p;
    const double fe = (double)e;
    return fe * 0.6931471803691238 + (fe * 1.9082149292705877e-10 + 2.0 * s * p);
}

static int odd_integer(float y) {
    if (floorf(y) != y || fabsf(y) >= 16777216.f) return 0;
    return (((int64_t)y) & 1) != 0;
}

O2_EXPORT float o2_expf(float v) { return g_libm ? expf(v) : (float)o2_det_exp((double)v); }

/* x^y with C99 Annex F's special values, else det_exp(y det_log|x|). */
O2_EXPORT float o2_powf(float x, float y) {
    if (g_libm) return powf(x, y);
    const float inf = INFINITY;
    if (y == 0.f) return 1.f;
    if (x == 1.f) return 1.f;
    if (!(x == x) || !(y == y)) return x + y;
    if (y == inf || y == -inf) {
        const float ax = fabsf(x);
        if (ax == 1.f) return 1.f;
        return ((ax < 1.f) == (y < 0.f)) ? inf : 0.f;
    }
    const int odd = odd_integer(y);
    if (x == 0.f) {
        const int neg = (int)(f_bits(x) >> 31);
        if (y < 0.f) return (odd && neg) ? -inf : inf;
        return (odd && neg) ? -0.f : 0.f;
    }
    if (x == inf) return y < 0.f ? 0.f : inf;
    if (x == -inf) {
        if (y < 0.f) return odd ? -0.f : 0.f;
        return odd ? -inf : inf;
    }
    if (x < 0.f && floorf(y) != y) return NAN;
    const float r = (float)o2_det_exp((double)y * o2_det_log((double)fabsf(x)));
    return (x < 0.f && odd) ? -r : r;
}

/* libstdc++ 11 std::min / std::max / std::clamp / std::lerp (float). */
static float smin(float a, float b) { return (b < a) ? b : a; }
static float smax(float a, float b) { return (a < b) ? b : a; }
static double smaxd(double a, double b) { return (a < b) ? b : a; }
static float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
static float std_lerp(float a, float b, float t) {
    if ((a <= 0.f && b >= 0.f) || (a >= 0.f && b <= 0.f)) return t * b + (1.f - t) * a;
    if (t == 1.f) return b;
    const float x = a + t * (b - a);
    return ((t > 1.f) == (b > a)) ? ((b < x) ? x : b) : ((b > x) ? x : b);
}
static int finitef_(float v) { return isfinite(v); }

static size_t idC(int n, int x, int y) { return (size_t)x + (size_t)n * (size_t)y; }
static size_t idX(int n, int x, int y) { return (size_t)x + (size_t)(n + 1) * (size_t)y; }
static size_t idY(int n, int x, int y) { return (size_t)x + (size_t)n * (size_t)y; }

/* ============================================== MacGridFluidSolver2D */

/* MacGridDomain2D.cpp:6-19 */
static int check_domain(int n, float box) {
    if (n <= 0) return O2_EINVAL;
    if ((size_t)n > (size_t)INT_MAX / ((size_t)n + 1)) return O2_ERANGE;
    if (!isfinite(box) || box <= 0.f || box / n <= 0.f || !isfinite(1.f / (box / n))) return O2_EINVAL;
    return O2_OK;
}

/* MacGridOperators2D.cpp:14-20 */
static void walls(int n, float* u, float* v) {
    for (int i = 0; i < n; ++i) {
        u[idX(n, 0, i)] = u[idX(n, n, i)] = 0.f;
        v[idY(n, i, 0)] = v[idY(n, i, n)] = 0.f;
    }
}

/* divergenceAt (MacGridOperators2D.cpp:6-11); *ok = 0 where it throws. */
static float divergence_at(int n, float h, const float* u, const float* v, int x, int y, int* ok) {
    const float value = (u[idX(n, x + 1, y)] - u[idX(n, x, y)] + v[idY(n, x, y + 1)] - v[idY(n, x, y)]) / h;
    if (!finitef_(value)) *ok = 0;
    return value;
}

/* maxDivergence (:26-30) */
static int max_divergence(int n, float h, const float* u, const float* v, float* out) {
    float result = 0.f;
    int ok = 1;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float d = divergence_at(n, h, u, v, x, y, &ok);
            if (!ok) return O2_EINVAL;
            result = smax(result, fabsf(d));
        }
    *out = result;
    return O2_OK;
}

/* SemiLagrangian2D.cpp:13-22; returns 0 where the reference throws. */
static int sl_sample(const float* f, int w, int h, float x, float y, float* out) {
    if (!finitef_(x) || !finitef_(y)) return 0;
    x = sclamp(x, 0.f, (float)(w - 1));
    y = sclamp(y, 0.f, (float)(h - 1));
    const int i = (int)floorf(x), j = (int)floorf(y);
    const int ip = (i + 1 < w - 1) ? i + 1 : w - 1, jp = (j + 1 < h - 1) ? j + 1 : h - 1;
    const float tx = x - (float)i, ty = y - (float)j;
    *out = std_lerp(std_lerp(f[i + w * j], f[ip + w * j], tx), std_lerp(f[i + w * jp], f[ip + w * jp], tx), ty);
    return 1;
}

static int overlaps(const float* a, size_t na, const float* b, size_t nb) {
    return na && nb && (uintptr_t)a < (uintptr_t)(b + nb) && (uintptr_t)b < (uintptr_t)(a + na);
}

/* SemiLagrangian2D::advect (SemiLagrangian2D.cpp:24-50). */
O2_EXPORT int o2_advect(int n, float box, int field, float dt, const float* u, const float* v, const float* src,
                        float* out) {
    int rc = check_domain(n, box);
    if (rc) return rc;
    if (field != VGX2_DENSITY && field != VGX2_VELOCITY_X && field != VGX2_VELOCITY_Y) return O2_EINVAL;
    const int w = n + (field == VGX2_VELOCITY_X), h = n + (field == VGX2_VELOCITY_Y);
    const size_t F = (size_t)n * (n + 1), count = (size_t)w * h;
    if (!isfinite(dt) || dt < 0.f || overlaps(src, count, out, count) || overlaps(u, F, out, count) ||
        overlaps(v, F, out, count))
        return O2_EINVAL;
    for (size_t i = 0; i < count; ++i)
        if (!finitef_(src[i])) return O2_EINVAL;
    for (size_t i = 0; i < F; ++i)
        if (!finitef_(u[i]) || !finitef_(v[i])) return O2_EINVAL;
    const float ox = field == VGX2_VELOCITY_X ? 0.f : .5f;
    const float oy = field == VGX2_VELOCITY_Y ? 0.f : .5f;
    const float dt_over_h = dt / (box / (float)n);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float px = (float)x + ox, py = (float)y + oy;
            float uu, vv, value;
            sl_sample(u, n + 1, n, px, py - .5f, &uu);
            sl_sample(v, n, n + 1, px - .5f, py, &vv);
            if (!sl_sample(src, w, h, px - dt_over_h * uu - ox, py - dt_over_h * vv - oy, &value)) return O2_EINVAL;
            out[(size_t)x + (size_t)w * y] = value;
        }
    return O2_OK;
}

static double nbr_sum(int n, const double* p, int x, int y) {
    double sum = 0.0;
    if (x > 0) sum += p[idC(n, x - 1, y)];
    if (x + 1 < n) sum += p[idC(n, x + 1, y)];
    if (y > 0) sum += p[idC(n, x, y - 1)];
    if (y + 1 < n) sum += p[idC(n, x, y + 1)];
    return sum;
}
static int diag_count(int n, int x, int y) { return (x > 0) + (x + 1 < n) + (y > 0) + (y + 1 < n); }
static double residual_inf(int n, const float* rhs, const double* p) {
    double value = 0.0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const size_t i = idC(n, x, y);
            value = smaxd(value, fabs((double)rhs[i] + nbr_sum(n, p, x, y) - (double)diag_count(n, x, y) * p[i]));
        }
    return value;
}

/* CpuPressureSolver2D::project (CpuPressureSolver2D.cpp:9-81) on (u, v) in
 * place with the workspace p (n^2 doubles) and rhs (n^2 floats). */
static int project_impl(int n, float h, const vgx_projection_options* o, float* u, float* v, double* p, float* rhs,
                        vgx_projection_result* result) {
    const float scale = h * h / o->dt;
    if (!isfinite(o->dt) || o->dt <= 0.f || !isfinite(scale) || scale <= 0.f || !isfinite(o->relaxation) ||
        o->relaxation <= 0.f || o->relaxation >= 2.f || o->max_iterations < 0 || o->residual_check_interval <= 0 ||
        !isfinite(o->absolute_tolerance) || o->absolute_tolerance < 0.0 || !isfinite(o->relative_tolerance) ||
        o->relative_tolerance < 0.0)
        return O2_EINVAL;
    const size_t N = (size_t)n * n, F = (size_t)n * (n + 1);
    for (size_t i = 0; i < F; ++i)
        if (!finitef_(u[i]) || !finitef_(v[i])) return O2_EINVAL;
    for (size_t i = 0; i < N; ++i) p[i] = 0.0;
    walls(n, u, v);
    int ok = 1;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            rhs[idC(n, x, y)] = divergence_at(n, h, u, v, x, y, &ok);
            if (!ok) return O2_EINVAL;
        }
    vgx_projection_result r;
    memset(&r, 0, sizeof r);
    for (size_t i = 0; i < N; ++i) {
        r.divergence_before = smax(r.divergence_before, fabsf(rhs[i]));
        rhs[i] *= -scale;
        if (!finitef_(rhs[i])) return O2_EINVAL;
    }
    double acc = 0.0;
    for (size_t i = 0; i < N; ++i) acc += rhs[i];
    r.removed_rhs_mean = acc / (double)N;
    for (size_t i = 0; i < N; ++i) rhs[i] -= (float)r.removed_rhs_mean;
    r.initial_residual = r.final_residual = residual_inf(n, rhs, p);
    const double target = smaxd(o->absolute_tolerance, o->relative_tolerance * r.initial_residual);
    r.converged = r.final_residual <= target;
    for (int iter = 0; iter < o->max_iterations && (o->fixed_iterations || !r.converged); ++iter) {
        for (int parity = 0; parity < 2; ++parity)
            for (int y = 0; y < n; ++y)
                for (int x = (y + parity) & 1; x < n; x += 2) {
                    const int d = diag_count(n, x, y);
                    if (d == 0) continue;
                    const size_t i = idC(n, x, y);
                    p[i] += (double)o->relaxation * (((double)rhs[i] + nbr_sum(n, p, x, y)) / d - p[i]);
                }
        r.iterations = iter + 1;
        if (r.iterations % o->residual_check_interval == 0 || r.iterations == o->max_iterations) {
            r.final_residual = residual_inf(n, rhs, p);
            r.converged = r.final_residual <= target;
        }
    }
    acc = 0.0;
    for (size_t i = 0; i < N; ++i) acc += p[i];
    const double mean = acc / (double)N;
    for (size_t i = 0; i < N; ++i) {
        p[i] -= mean;
        if (!isfinite(p[i])) return O2_ERUNTIME;
    }
    r.final_residual = residual_inf(n, rhs, p);
    r.converged = r.final_residual <= target;
    r.residual_available = 1;
    /* applyPressureGradient (MacGridOperators2D.cpp:31-40) */
    const double gscale = (double)o->dt / (double)h;
    for (int y = 0; y < n; ++y)
        for (int x = 1; x < n; ++x)
            u[idX(n, x, y)] = (float)((double)u[idX(n, x, y)] - gscale * (p[idC(n, x, y)] - p[idC(n, x - 1, y)]));
    for (int y = 1; y < n; ++y)
        for (int x = 0; x < n; ++x)
            v[idY(n, x, y)] = (float)((double)v[idY(n, x, y)] - gscale * (p[idC(n, x, y)] - p[idC(n, x, y - 1)]));
    walls(n, u, v);
    const int rc = max_divergence(n, h, u, v, &r.divergence_after);
    if (rc) return rc;
    *result = r;
    return O2_OK;
}

O2_EXPORT int o2_project(int n, float box, const vgx_projection_options* o, float* u, float* v, double* pressure,
                         vgx_projection_result* result) {
    int rc = check_domain(n, box);
    if (rc) return rc;
    const size_t N = (size_t)n * n;
    double* p = (double*)calloc(N, sizeof(double));
    float* rhs = (float*)calloc(N, sizeof(float));
    vgx_projection_result r;
    memset(&r, 0, sizeof r);
    rc = project_impl(n, box / (float)n, o, u, v, p, rhs, &r);
    if (rc == O2_OK) {
        if (pressure) memcpy(pressure, p, N * sizeof(double));
        if (result) *result = r;
    }
    free(p);
    free(rhs);
    return rc;
}

typedef struct o2_sim {
    int n;
    float box;
    vgx2_params p;
    float *rho, *u, *v, *un, *vn, *rhon, *uh, *vh, *rhs;
    double* pressure;
    vgx_projection_result last, mid;
} o2_sim;

O2_EXPORT void o2_sim_reset(o2_sim* s) {
    const size_t N = (size_t)s->n * s->n, F = (size_t)s->n * (s->n + 1);
    memset(s->rho, 0, N * 4);
    memset(s->rhon, 0, N * 4);
    memset(s->rhs, 0, N * 4);
    memset(s->pressure, 0, N * 8);
    float* faces[] = {s->u, s->v, s->un, s->vn, s->uh, s->vh};
    for (int k = 0; k < 6; ++k) memset(faces[k], 0, F * 4);
    memset(&s->last, 0, sizeof s->last);
    memset(&s->mid, 0, sizeof s->mid);
}

O2_EXPORT o2_sim* o2_sim_create(int n, float box, const vgx2_params* params) {
    if (check_domain(n, box)) return NULL;
    o2_sim* s = (o2_sim*)calloc(1, sizeof(o2_sim));
    s->n = n;
    s->box = box;
    s->p = *params;
    const size_t N = (size_t)n * n, F = (size_t)n * (n + 1);
    s->rho = (float*)calloc(N, 4);
    s->rhon = (float*)calloc(N, 4);
    s->rhs = (float*)calloc(N, 4);
    s->pressure = (double*)calloc(N, 8);
    s->u = (float*)calloc(F, 4);
    s->v = (float*)calloc(F, 4);
    s->un = (float*)calloc(F, 4);
    s->vn = (float*)calloc(F, 4);
    s->uh = (float*)calloc(F, 4);
    s->vh = (float*)calloc(F, 4);
    return s;
}

O2_EXPORT void o2_sim_destroy(o2_sim* s) {
    if (!s) return;
    free(s->rho); free(s->rhon); free(s->rhs); free(s->pressure);
    free(s->u); free(s->v); free(s->un); free(s->vn); free(s->uh); free(s->vh);
    free(s);
}

O2_EXPORT void o2_sim_set_params(o2_sim* s, const vgx2_params* p) { s->p = *p; }
O2_EXPORT float* o2_sim_field(o2_sim* s, int which) { return which == 0 ? s->rho : which == 1 ? s->u : s->v; }
O2_EXPORT double* o2_sim_pressure(o2_sim* s) { return s->pressure; }
O2_EXPORT void o2_sim_projection(o2_sim* s, int midpoint, vgx_projection_result* r) { *r = midpoint ? s->mid : s->last; }

static float cell_size(const o2_sim* s) { return s->box / (float)s->n; }

/* advectVelocity (MacGridFluidSolver2D.cpp:73-79) */
static int advect_velocity(o2_sim* s, const float* su, const float* sv, float dt) {
    int rc = o2_advect(s->n, s->box, VGX2_VELOCITY_X, dt, s->u, s->v, su, s->un);
    if (rc) return rc;
    rc = o2_advect(s->n, s->box, VGX2_VELOCITY_Y, dt, s->u, s->v, sv, s->vn);
    if (rc) return rc;
    const size_t F = (size_t)s->n * (s->n + 1);
    memcpy(s->u, s->un, F * 4);
    memcpy(s->v, s->vn, F * 4);
    return O2_OK;
}

/* addBuoyancy (:81-92) */
static void add_buoyancy(o2_sim* s, float dt) {
    const int n = s->n;
    for (int y = 1; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float smoke = .5f * (o2_powf(s->rho[idC(n, x, y - 1)], .25f) + o2_powf(s->rho[idC(n, x, y)], .25f));
            s->v[idY(n, x, y)] += dt * s->p.buoyancy * smoke;
        }
}

/* emit (:51-71) */
static void emit2(o2_sim* s) {
    const vgx2_params* p = &s->p;
    if (p->source_strength <= 0.f) return;
    const int n = s->n;
    const float r2 = p->emitter_radius * p->emitter_radius;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float dx = ((float)x + .5f) / (float)n - p->emitter_center_x;
            const float dy = ((float)y + .5f) / (float)n - p->emitter_center_y;
            const float q = (dx * dx + dy * dy) / r2;
            if (q > 1.f) continue;
            const float weight = o2_expf(-3.5f * q);
            const size_t i = idC(n, x, y);
            s->rho[i] = smax(s->rho[i], smin(1.f, p->source_strength * 3.f * weight));
            float* vy = &s->v[idY(n, x, y)];
            if (p->emitter_kick_multiplier > 0.f)
                *vy = smax(*vy, p->emitter_kick_multiplier * (.6f + p->source_strength) * weight);
        }
}

/* project(dt) (:94-101) */
static int project2(o2_sim* s, float dt, vgx_projection_result* out) {
    vgx_projection_options o;
    memset(&o, 0, sizeof o);
    o.dt = dt;
    o.relaxation = s->p.sor_omega;
    o.max_iterations = s->p.project_iterations;
    o.residual_check_interval = 8;
    o.absolute_tolerance = 1e-6;
    o.relative_tolerance = 1e-5;
    o.fixed_iterations = 0;
    return project_impl(s->n, cell_size(s), &o, s->u, s->v, s->pressure, s->rhs, out);
}

/* advectDensity (:103-108) */
static int advect_density(o2_sim* s, float dt) {
    const int rc = o2_advect(s->n, s->box, VGX2_DENSITY, dt, s->u, s->v, s->rho, s->rhon);
    if (rc) return rc;
    const float decay = expf(-dt * s->p.smoke_decay);
    const size_t N = (size_t)s->n * s->n;
    for (size_t i = 0; i < N; ++i) s->rho[i] = decay * s->rhon[i];
    return O2_OK;
}

/* advance (:22-45) and advanceReflection (:114-142) */
O2_EXPORT int o2_sim_advance(o2_sim* s, float dt) {
    const vgx2_params* p = &s->p;
    if (!isfinite(p->emitter_kick_multiplier) || p->emitter_kick_multiplier < 0.f) return O2_EINVAL;
    if (!isfinite(dt) || dt <= 0.f || !isfinite(p->source_strength) || p->source_strength < 0.f ||
        !isfinite(p->emitter_radius) || p->emitter_radius <= 0.f || p->emitter_radius > 1.f ||
        !isfinite(p->emitter_center_x) || p->emitter_center_x < 0.f || p->emitter_center_x > 1.f ||
        !isfinite(p->emitter_center_y) || p->emitter_center_y < 0.f || p->emitter_center_y > 1.f ||
        !isfinite(p->buoyancy) || !isfinite(p->smoke_decay) || p->smoke_decay < 0.f || !isfinite(p->sor_omega) ||
        p->sor_omega <= 0.f || p->sor_omega >= 2.f || p->project_iterations < 0)
        return O2_EINVAL;
    const int n = s->n;
    walls(n, s->u, s->v);
    memset(&s->mid, 0, sizeof s->mid);
    int rc;
    if (p->reflection) {
        const float half = .5f * dt;
        if ((rc = advect_density(s, half))) return rc;
        if ((rc = advect_velocity(s, s->u, s->v, half))) return rc;
        add_buoyancy(s, half);
        emit2(s);
        vgx_projection_result r;
        if ((rc = project2(s, half, &r))) return rc;
        s->mid = r;
        const size_t F = (size_t)n * (n + 1);
        for (size_t i = 0; i < F; ++i) s->uh[i] = 2.f * s->u[i] - s->un[i];
        for (size_t i = 0; i < F; ++i) s->vh[i] = 2.f * s->v[i] - s->vn[i];
        for (int i = 0; i < n; ++i) {
            s->uh[idX(n, 0, i)] = s->uh[idX(n, n, i)] = 0.f;
            s->vh[idY(n, i, 0)] = s->vh[idY(n, i, n)] = 0.f;
        }
        if ((rc = advect_density(s, half))) return rc;
        if ((rc = advect_velocity(s, s->uh, s->vh, half))) return rc;
        if ((rc = project2(s, half, &r))) return rc;
        s->last = r;
        return O2_OK;
    }
    if ((rc = advect_velocity(s, s->u, s->v, dt))) return rc;
    add_buoyancy(s, dt);
    emit2(s);
    vgx_projection_result r;
    if ((rc = project2(s, dt, &r))) return rc;
    s->last = r;
    return advect_density(s, dt);
}

/* diagnostics (:157-167) */
O2_EXPORT int o2_sim_diagnostics(o2_sim* s, vgx_diagnostics* out) {
    const size_t F = (size_t)s->n * (s->n + 1);
    double sum = 0.0;
    float vmax = 0.f;
    const float* fields[] = {s->u, s->v};
    for (int f = 0; f < 2; ++f)
        for (size_t i = 0; i < F; ++i) {
            const float value = fields[f][i];
            sum += (double)value * value;
            vmax = smax(vmax, fabsf(value));
        }
    const double h = (double)cell_size(s);
    out->kinetic_energy = (float)(.5 * h * h * sum);
    out->max_velocity = vmax;
    return max_divergence(s->n, cell_size(s), s->u, s->v, &out->max_divergence);
}

/* ========================================= ReflectionMacFluidSolver2D */

typedef struct o2h_sim {
    int n;
    vgx2r_params p;
    uint32_t random_state;
    float *smoke, *vx, *vy, *pressure, *divergence;
    float *fwd, *back, *tmp, *vx0, *vy0, *vxt, *vyt, *vxm, *vym, *vxh, *vyh;
} o2h_sim;

static float hh(const o2h_sim* s) { return 1.0f / (float)s->n; }
static float mix(float a, float b, float t) { return a + (b - a) * t; }
static int mini(int a, int b) { return (b < a) ? b : a; }

/* sampleCell / sampleVxField / sampleVyField (ReflectionMacFluidSolver2D.hpp:113-147) */
static float sample_cell(const o2h_sim* s, const float* f, float px, float py) {
    const int n = s->n;
    const float gx = sclamp(px / hh(s) - 0.5f, 0.f, (float)(n - 1));
    const float gy = sclamp(py / hh(s) - 0.5f, 0.f, (float)(n - 1));
    const int x0 = (int)gx, y0 = (int)gy;
    const int x1 = mini(x0 + 1, n - 1), y1 = mini(y0 + 1, n - 1);
    const float tx = gx - (float)x0, ty = gy - (float)y0;
    return mix(mix(f[idC(n, x0, y0)], f[idC(n, x1, y0)], tx), mix(f[idC(n, x0, y1)], f[idC(n, x1, y1)], tx), ty);
}
static float sample_vx(const o2h_sim* s, const float* f, float px, float py) {
    const int n = s->n;
    const float gx = sclamp(px / hh(s), 0.f, (float)n);
    const float gy = sclamp(py / hh(s) - 0.5f, 0.f, (float)(n - 1));
    const int x0 = mini((int)gx, n - 1), y0 = (int)gy;
    const int x1 = x0 + 1, y1 = mini(y0 + 1, n - 1);
    const float tx = gx - (float)x0, ty = gy - (float)y0;
    return mix(mix(f[idX(n, x0, y0)], f[idX(n, x1, y0)], tx), mix(f[idX(n, x0, y1)], f[idX(n, x1, y1)], tx), ty);
}
static float sample_vy(const o2h_sim* s, const float* f, float px, float py) {
    const int n = s->n;
    const float gx = sclamp(px / hh(s) - 0.5f, 0.f, (float)(n - 1));
    const float gy = sclamp(py / hh(s), 0.f, (float)n);
    const int x0 = (int)gx, y0 = mini((int)gy, n - 1);
    const int x1 = mini(x0 + 1, n - 1), y1 = y0 + 1;
    const float tx = gx - (float)x0, ty = gy - (float)y0;
    return mix(mix(f[idY(n, x0, y0)], f[idY(n, x1, y0)], tx), mix(f[idY(n, x0, y1)], f[idY(n, x1, y1)], tx), ty);
}

static float random_unit(o2h_sim* s) {
    s->random_state = s->random_state * 1664525u + 1013904223u;
    return (float)(s->random_state >> 8) * (1.0f / 16777216.0f);
}

/* advectScalar (:162-205) on the smoke, in place. */
static void advect_scalar(o2h_sim* s, float* field, float dt) {
    const int n = s->n;
    const float dx = hh(s), pad = 0.5f * hh(s);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float px = ((float)x + .5f) * dx, py = ((float)y + .5f) * dx;
            float qx = px - sample_vx(s, s->vx, px, py) * dt;
            float qy = py - sample_vy(s, s->vy, px, py) * dt;
            qx = sclamp(qx, pad, 1.0f - pad);
            qy = sclamp(qy, pad, 1.0f - pad);
            s->fwd[idC(n, x, y)] = sample_cell(s, field, qx, qy);
        }
    const size_t N = (size_t)n * n;
    if (!s->p.maccormack_smoke) {
        memcpy(field, s->fwd, N * 4);
        return;
    }
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float px = ((float)x + .5f) * dx, py = ((float)y + .5f) * dx;
            float qx = px + sample_vx(s, s->vx, px, py) * dt;
            float qy = py + sample_vy(s, s->vy, px, py) * dt;
            qx = sclamp(qx, pad, 1.0f - pad);
            qy = sclamp(qy, pad, 1.0f - pad);
            s->back[idC(n, x, y)] = sample_cell(s, s->fwd, qx, qy);
        }
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const size_t i = idC(n, x, y);
            const float corrected = s->fwd[i] + 0.5f * (field[i] - s->back[i]);
            float fmin = smin(field[i], s->fwd[i]), fmax = smax(field[i], s->fwd[i]);
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx2 = -1; dx2 <= 1; ++dx2) {
                    const int a = x + dx2, b = y + dy;
                    if (a >= 0 && b >= 0 && a < n && b < n) {
                        const float w = s->fwd[idC(n, a, b)], f = field[idC(n, a, b)];
                        float m = fmin;
                        if (w < m) m = w;
                        if (f < m) m = f;
                        fmin = m;
                        float M = fmax;
                        if (M < w) M = w;
                        if (M < f) M = f;
                        fmax = M;
                    }
                }
            field[i] = sclamp(corrected, fmin, fmax);
        }
}

/* advectVx / advectVy (:207-227) into out, from src through (fu, fv). */
static void advect_vx(o2h_sim* s, float dt, const float* src, const float* fu, const float* fv, float* out) {
    const int n = s->n;
    const float dx = hh(s);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x <= n; ++x) {
            const float px = (float)x * dx, py = ((float)y + .5f) * dx;
            float qx = px - sample_vx(s, fu, px, py) * dt;
            float qy = py - sample_vy(s, fv, px, py) * dt;
            qx = sclamp(qx, 0.f, 1.0f);
            qy = sclamp(qy, 0.5f * hh(s), 1.0f - 0.5f * hh(s));
            out[idX(n, x, y)] = sample_vx(s, src, qx, qy);
        }
}
static void advect_vy(o2h_sim* s, float dt, const float* src, const float* fu, const float* fv, float* out) {
    const int n = s->n;
    const float dx = hh(s);
    for (int y = 0; y <= n; ++y)
        for (int x = 0; x < n; ++x) {
            const float px = ((float)x + .5f) * dx, py = (float)y * dx;
            float qx = px - sample_vx(s, fu, px, py) * dt;
            float qy = py - sample_vy(s, fv, px, py) * dt;
            qx = sclamp(qx, 0.5f * hh(s), 1.0f - 0.5f * hh(s));
            qy = sclamp(qy, 0.f, 1.0f);
            out[idY(n, x, y)] = sample_vy(s, src, qx, qy);
        }
}

/* applyBoundary (:229-238), closed walls */
O2_EXPORT void o2h_apply_boundary(o2h_sim* s) { walls(s->n, s->vx, s->vy); }

/* project (:240-289) */
O2_EXPORT void o2h_project(o2h_sim* s, float dt, int iterations) {
    const int n = s->n;
    const float hinv = 1.0f / hh(s), h2 = hh(s) * hh(s);
    o2h_apply_boundary(s);
    const size_t N = (size_t)n * n;
    for (size_t i = 0; i < N; ++i) s->divergence[i] = 0.f;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x)
            s->divergence[idC(n, x, y)] =
                (s->vx[idX(n, x + 1, y)] - s->vx[idX(n, x, y)] + s->vy[idY(n, x, y + 1)] - s->vy[idY(n, x, y)]) * hinv;
    for (size_t i = 0; i < N; ++i) s->pressure[i] = 0.f;
    static const int ddx[4] = {-1, 1, 0, 0}, ddy[4] = {0, 0, -1, 1};
    const float omega = s->p.sor_omega;
    for (int iter = 0; iter < iterations; ++iter)
        for (int parity = 0; parity < 2; ++parity)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    if (((x + y) & 1) != parity) continue;
                    const size_t i = idC(n, x, y);
                    float sum = 0.f;
                    int terms = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int a = x + ddx[k], b = y + ddy[k];
                        if (!(a >= 0 && b >= 0 && a < n && b < n)) continue;
                        ++terms;
                        sum += s->pressure[idC(n, a, b)];
                    }
                    s->pressure[i] = (1.f - omega) * s->pressure[i] +
                                     omega * (sum - s->divergence[i] * h2 / dt) / (float)(terms > 1 ? terms : 1);
                }
    for (int y = 0; y < n; ++y)
        for (int x = 1; x < n; ++x)
            s->vx[idX(n, x, y)] -= dt * (s->pressure[idC(n, x, y)] - s->pressure[idC(n, x - 1, y)]) * hinv;
    for (int y = 1; y < n; ++y)
        for (int x = 0; x < n; ++x)
            s->vy[idY(n, x, y)] -= dt * (s->pressure[idC(n, x, y)] - s->pressure[idC(n, x, y - 1)]) * hinv;
    o2h_apply_boundary(s);
}

/* emit (:291-314) */
O2_EXPORT void o2h_emit(o2h_sim* s, float strength) {
    if (strength <= 0.f) return;
    const vgx2r_params* p = &s->p;
    const int n = s->n;
    const float r2 = p->emitter_radius * p->emitter_radius, dx = hh(s);
    const int emit_h = mini((int)(p->emitter_center_y * (float)n + p->emitter_radius * (float)n + 2.f), n);
    int y0 = (int)(p->emitter_center_y * (float)n - p->emitter_radius * (float)n - 1.f);
    if (y0 < 0) y0 = 0;
    for (int y = y0; y < emit_h; ++y)
        for (int x = 0; x < n; ++x) {
            const float px = ((float)x + .5f) * dx, py = ((float)y + .5f) * dx;
            const float dx2 = (px - p->emitter_center_x) * (px - p->emitter_center_x);
            const float dy2 = (py - p->emitter_center_y) * (py - p->emitter_center_y);
            const float q = (dx2 + dy2) / r2;
            if (q > 1.0f) continue;
            const float w = o2_expf(-p->emitter_falloff * q);
            const size_t i = idC(n, x, y);
            s->smoke[i] = smax(s->smoke[i], smin(1.0f, strength * p->emitter_smoke_scale * w));
            if (p->emitter_kick_multiplier > 0.f)
                s->vy[idY(n, x, y)] = smax(s->vy[idY(n, x, y)],
                                           p->emitter_kick_multiplier * (p->emitter_kick_base +
                                                                         p->emitter_kick_scale * strength) * w);
            if (p->stir_strength > 0.f) {
                s->vx[idX(n, x, y)] += (random_unit(s) - 0.5f) * p->stir_strength * w;
                s->vx[idX(n, x + 1, y)] += (random_unit(s) - 0.5f) * p->stir_strength * w;
            }
        }
}

O2_EXPORT void o2h_advect_smoke(o2h_sim* s, float dt) { advect_scalar(s, s->smoke, dt); }

/* step (:316-408), closed box */
O2_EXPORT void o2h_step(o2h_sim* s, float dt, float buoyancy, float dissipation, float source_strength,
                        int project_iterations) {
    const int n = s->n;
    const size_t N = (size_t)n * n, F = (size_t)n * (n + 1);
    const float half = 0.5f * dt;
    const int half_iters = project_iterations / 2 > 1 ? project_iterations / 2 : 1;
    const float half_decay = powf(dissipation, half * 60.0f);
    advect_scalar(s, s->smoke, half);
    for (size_t i = 0; i < N; ++i) s->smoke[i] *= half_decay;
    memcpy(s->vx0, s->vx, F * 4);
    memcpy(s->vy0, s->vy, F * 4);
    advect_vx(s, half, s->vx0, s->vx0, s->vy0, s->vxt);
    advect_vy(s, half, s->vy0, s->vx0, s->vy0, s->vyt);
    memcpy(s->vx, s->vxt, F * 4);
    memcpy(s->vy, s->vyt, F * 4);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float f = buoyancy * o2_powf(s->smoke[idC(n, x, y)], s->p.buoyancy_exp) * half;
            s->vy[idY(n, x, y)] += s->p.buoyancy_split * f;
            s->vy[idY(n, x, y + 1)] += s->p.buoyancy_split * f;
        }
    o2h_emit(s, source_strength);
    o2h_project(s, half, half_iters);
    memcpy(s->vxm, s->vx, F * 4);
    memcpy(s->vym, s->vy, F * 4);
    for (size_t i = 0; i < F; ++i) s->vxh[i] = 2.f * s->vxm[i] - s->vxt[i];
    for (size_t i = 0; i < F; ++i) s->vyh[i] = 2.f * s->vym[i] - s->vyt[i];
    advect_scalar(s, s->smoke, half);
    for (size_t i = 0; i < N; ++i) s->smoke[i] *= half_decay;
    advect_vx(s, half, s->vxh, s->vxm, s->vym, s->vx);
    advect_vy(s, half, s->vyh, s->vxm, s->vym, s->vy);
    o2h_project(s, half, half_iters);
}

/* advance (:33-43) */
O2_EXPORT int o2h_advance(o2h_sim* s, float dt) {
    const vgx2r_params* p = &s->p;
    if (!isfinite(p->emitter_kick_multiplier) || p->emitter_kick_multiplier < 0.f) return O2_EINVAL;
    if (!isfinite(dt) || dt <= 0.f || !isfinite(p->buoyancy) || !isfinite(p->smoke_decay) || p->smoke_decay < 0.f ||
        !isfinite(p->source_strength) || p->source_strength < 0.f || !isfinite(p->emitter_radius) ||
        p->emitter_radius <= 0.f || p->emitter_radius > 1.f || !isfinite(p->stir_strength) ||
        p->stir_strength < 0.f || !isfinite(p->sor_omega) || p->sor_omega <= 0.f || p->sor_omega >= 2.f ||
        p->project_iterations < 0)
        return O2_EINVAL;
    o2h_apply_boundary(s);
    o2h_step(s, dt, p->buoyancy, expf(-p->smoke_decay / 60.f), p->source_strength, p->project_iterations);
    return O2_OK;
}

O2_EXPORT void o2h_reset(o2h_sim* s) {
    const size_t N = (size_t)s->n * s->n, F = (size_t)s->n * (s->n + 1);
    memset(s->smoke, 0, N * 4);
    memset(s->pressure, 0, N * 4);
    memset(s->divergence, 0, N * 4);
    memset(s->vx, 0, F * 4);
    memset(s->vy, 0, F * 4);
    s->random_state = s->p.random_seed;
}

O2_EXPORT o2h_sim* o2h_create(int n, const vgx2r_params* params) {
    if (n < 2 || n > 512) return NULL;
    o2h_sim* s = (o2h_sim*)calloc(1, sizeof(o2h_sim));
    s->n = n;
    s->p = *params;
    s->random_state = 0x9E3779B9u;
    const size_t N = (size_t)n * n, F = (size_t)n * (n + 1);
    float** cells[] = {&s->smoke, &s->pressure, &s->divergence, &s->fwd, &s->back, &s->tmp};
    for (int k = 0; k < 6; ++k) *cells[k] = (float*)calloc(N, 4);
    float** faces[] = {&s->vx, &s->vy, &s->vx0, &s->vy0, &s->vxt, &s->vyt, &s->vxm, &s->vym, &s->vxh, &s->vyh};
    for (int k = 0; k < 10; ++k) *faces[k] = (float*)calloc(F, 4);
    return s;
}

O2_EXPORT void o2h_destroy(o2h_sim* s) {
    if (!s) return;
    float* all[] = {s->smoke, s->pressure, s->divergence, s->fwd, s->back, s->tmp, s->vx, s->vy,
                    s->vx0, s->vy0, s->vxt, s->vyt, s->vxm, s->vym, s->vxh, s->vyh};
    for (int k = 0; k < 16; ++k) free(all[k]);
    free(s);
}

O2_EXPORT void o2h_set_params(o2h_sim* s, const vgx2r_params* p) { s->p = *p; }
O2_EXPORT float* o2h_field(o2h_sim* s, int which) {
    switch (which) {
        case 0: return s->smoke;
        case 1: return s->vx;
        case 2: return s->vy;
        case 3: return s->pressure;
        default: return s->divergence;
    }
}
O2_EXPORT uint32_t o2h_random_state(o2h_sim* s) { return s->random_state; }
O2_EXPORT void o2h_set_random_state(o2h_sim* s, uint32_t state) { s->random_state = state; }

/* diagnostics (:44-57) */
O2_EXPORT void o2h_diagnostics(o2h_sim* s, vgx_diagnostics* out) {
    const int n = s->n;
    const size_t F = (size_t)n * (n + 1);
    double energy = 0.;
    float vmax = 0.f, dmax = 0.f;
    const float* fields[] = {s->vx, s->vy};
    for (int f = 0; f < 2; ++f)
        for (size_t i = 0; i < F; ++i) {
            const float v = fields[f][i];
            energy += (double)v * v;
            vmax = smax(vmax, fabsf(v));
        }
    const double h = (double)hh(s);
    out->kinetic_energy = (float)(.5 * energy * h * h);
    out->max_velocity = vmax;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x)
            dmax = smax(dmax, fabsf((s->vx[idX(n, x + 1, y)] - s->vx[idX(n, x, y)] + s->vy[idY(n, x, y + 1)] -
                                     s->vy[idY(n, x, y)]) / hh(s)));
    out->max_divergence = dmax;
}
