// vgx2_device.hpp - device-side helpers of the MAC 2D solvers.
//
// Two solvers of vapor's src/solvers/mac2d share these helpers:
//   * MacGridFluidSolver2D (MacGridFluidSolver2D.cpp, SemiLagrangian2D.cpp,
//     CpuPressureSolver2D.cpp, MacGridOperators2D.cpp): bilinear sampling
//     through std::lerp, FP64 pressure;
//   * ReflectionMacFluidSolver2D (ReflectionMacFluidSolver2D.hpp): the
//     historical plume with its own samplers and FP32 pressure.
// Every float and double expression keeps the reference's operation order and
// the library is compiled with -ffp-contract=off, so each value rounds as the
// reference's CPU code rounds it. The only exceptions are e^v and x^y, which
// the reference takes from the C library (expf, powf): here they are a fixed
// sequence of IEEE double operations (det_exp, det_log) rounded once to
// float, the same bits on the host and the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace vgx2 {

// Index maps of MacGridDomain2D.hpp:17-19 and ReflectionMacFluidSolver2D.hpp:102-104.
__host__ __device__ __forceinline__ int idC(int n, int x, int y) { return x + n * y; }
__host__ __device__ __forceinline__ int idX(int n, int x, int y) { return x + (n + 1) * y; }
__host__ __device__ __forceinline__ int idY(int n, int x, int y) { return x + n * y; }

// std::min / std::max / std::clamp as libstdc++ defines them (the comparison
// order decides which of two equal values, +0 or -0, is returned).
template <class T> __host__ __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> __host__ __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> __host__ __device__ __forceinline__ T sclamp(T v, T lo, T hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}
// std::min({a, b, c}) / std::max({a, b, c}): the first smallest / largest.
template <class T> __host__ __device__ __forceinline__ T smin3(T a, T b, T c) {
    T m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}
template <class T> __host__ __device__ __forceinline__ T smax3(T a, T b, T c) {
    T m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}

// std::lerp(a, b, t) for float as libstdc++ 11 computes it (P0811): the
// two-product form when a and b straddle zero, b at t == 1, otherwise
// a + t (b - a) kept monotonic on the side of b.
__host__ __device__ __forceinline__ float std_lerp(float a, float b, float t) {
    if ((a <= 0.f && b >= 0.f) || (a >= 0.f && b <= 0.f)) return t * b + (1.f - t) * a;
    if (t == 1.f) return b;
    const float x = a + t * (b - a);
    return ((t > 1.f) == (b > a)) ? ((b < x) ? x : b) : ((b > x) ? x : b);
}

// int(v) made total (NaN -> 0, magnitude bounded); equal to (int)v wherever
// that is defined.
__host__ __device__ __forceinline__ int trunc_int(float v) {
    return (v == v) ? (int)fmaxf(fminf(v, 1.0e9f), -1.0e9f) : 0;
}

__host__ __device__ __forceinline__ bool finitef(float v) { return v - v == 0.f; }
__host__ __device__ __forceinline__ bool finited(double v) { return v - v == 0.0; }

__host__ __device__ __forceinline__ double bits_double(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
__host__ __device__ __forceinline__ uint64_t double_bits(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
}
__host__ __device__ __forceinline__ uint32_t float_bits(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

// e^t in double for -745 < t < 709.7 from IEEE +, -, * and floor only:
// t = k ln2 + r with k = floor(t / ln2 + 1/2) and ln2 = L1 + L2 (L1 has 21
// trailing zero bits, so k L1 is exact for |k| < 2^21), e^r by its Taylor
// series to r^13 on |r| <= ln2 / 2 (truncation below 5e-18 relative), then
// the exact scaling by 2^k split in two factors so every step stays normal.
__host__ __device__ __forceinline__ double det_exp(double t) {
    if (!(t == t)) return t;
    if (t > 709.78) return bits_double(0x7ff0000000000000ull);
    if (t < -745.2) return 0.0;
    const double k = floor(t * 1.4426950408889634 + 0.5);
    const double r = (t - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    const int ki = (int)k;
    const int k1 = ki / 2, k2 = ki - k1;  // 2^k = 2^k1 * 2^k2, each normal
    const double s1 = bits_double((uint64_t)(int64_t)(k1 + 1023) << 52);
    const double s2 = bits_double((uint64_t)(int64_t)(k2 + 1023) << 52);
    return (p * s1) * s2;
}

// ln x in double for finite x > 0 (normal double) from IEEE +, -, *, /:
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m-1)/(m+1),
// by its series to s^25 (|s| <= 0.1716, truncation below 1e-19 relative), and
// e ln2 with ln2 split as in det_exp.
__host__ __device__ __forceinline__ double det_log(double x) {
    const uint64_t b = double_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = 1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = 1.0 + z * p;
    const double fe = (double)e;
    return fe * 0.6931471803691238 + (fe * 1.9082149292705877e-10 + 2.0 * s * p);
}

// e^v for float v (the reference's std::exp(float)): det_exp rounded once.
__host__ __device__ __forceinline__ float det_expf(float v) { return (float)det_exp((double)v); }

// Whether y (finite) is an integer, and whether it is an odd one.
__host__ __device__ __forceinline__ bool is_integer_f(float y) { return floorf(y) == y; }
__host__ __device__ __forceinline__ bool is_odd_integer_f(float y) {
    if (!is_integer_f(y) || fabsf(y) >= 16777216.f) return false;
    const int64_t i = (int64_t)y;
    return (i & 1) != 0;
}

// x^y for floats (the reference's std::pow(float, float)) with the special
// values of C99 Annex F.9.4.4; otherwise |x|^y = det_exp(y det_log(|x|)) in
// double, rounded once, negated for x < 0 and odd integer y, NaN for x < 0
// and non-integer y.
__host__ __device__ __forceinline__ float det_powf(float x, float y) {
    const float inf = __builtin_huge_valf();
    if (y == 0.f) return 1.f;
    if (x == 1.f) return 1.f;
    if (!(x == x) || !(y == y)) return x + y;
    const bool yinf = (y == inf || y == -inf);
    if (yinf) {
        const float ax = fabsf(x);
        if (ax == 1.f) return 1.f;
        return ((ax < 1.f) == (y < 0.f)) ? inf : 0.f;
    }
    const bool odd = is_odd_integer_f(y);
    if (x == 0.f) {
        const bool neg = float_bits(x) >> 31;
        if (y < 0.f) return (odd && neg) ? -inf : inf;
        return (odd && neg) ? -0.f : 0.f;
    }
    if (x == inf) return y < 0.f ? 0.f : inf;
    if (x == -inf) {
        if (y < 0.f) return odd ? -0.f : 0.f;
        return odd ? -inf : inf;
    }
    if (x < 0.f && !is_integer_f(y)) return __builtin_nanf("");
    const float r = (float)det_exp((double)y * det_log((double)fabsf(x)));
    return (x < 0.f && odd) ? -r : r;
}

// a / 3 correctly rounded without a division: q = a (1/3) rounded, the
// remainder a - 3q is exact in one FMA, and one FMA correction lands on the
// rounded quotient, because a / 3 is never a rounding tie and lies at least
// ulp/12 from one unless it is one. Zeros take q's sign (that of a). Checked
// against IEEE division for every float and for 5e8 doubles, including
// neighbours of exact multiples of 3.
__host__ __device__ __forceinline__ double div3(double a) {
    const double z = 1.0 / 3.0;
    const double q = a * z;
    const double r = __builtin_fma(-3.0, q, a);
    const double q1 = __builtin_fma(r, z, q);
    return q1 == 0.0 ? q : q1;
}
__host__ __device__ __forceinline__ float div3(float a) {
    const float z = 1.0f / 3.0f;
    const float q = a * z;
    const float r = __builtin_fmaf(-3.0f, q, a);
    const float q1 = __builtin_fmaf(r, z, q);
    return q1 == 0.0f ? q : q1;
}
// a / d for a neighbour count d in {2, 3, 4}, branch-free.
template <class V> __host__ __device__ __forceinline__ V div_234(V a, int d) {
    const V q3 = div3(a);
    return d == 4 ? a * (V)0.25 : (d == 3 ? q3 : a * (V)0.5);
}

// ---------------------------------------------------------------------------
// MacGridFluidSolver2D's sampler: SemiLagrangian2D.cpp:13-22. Bilinear in
// index coordinates, clamped to [0, w-1] x [0, h-1]. The caller has checked
// that x and y are finite (the reference throws otherwise). F is a pointer or
// any indexable accessor (the host checks pass bounds-checked ones).
template <class F>
__host__ __device__ __forceinline__ float sl2_sample(const F& f, int w, int h, float x, float y) {
    x = sclamp(x, 0.f, (float)(w - 1));
    y = sclamp(y, 0.f, (float)(h - 1));
    const int i = (int)floorf(x), j = (int)floorf(y);
    const int ip = smin(i + 1, w - 1), jp = smin(j + 1, h - 1);
    const float tx = x - (float)i, ty = y - (float)j;
    return std_lerp(std_lerp(f[i + w * j], f[ip + w * j], tx), std_lerp(f[i + w * jp], f[ip + w * jp], tx), ty);
}

// Field kinds of MacField2D (SemiLagrangian2D.hpp:4).
constexpr int kDensity = 0, kVelocityX = 1, kVelocityY = 2;

// One sample of SemiLagrangian2D::advect (SemiLagrangian2D.cpp:40-49) at
// (x, y) of the field's extent. ok becomes false where the backtrace leaves
// the finite floats (the reference throws there).
template <int FIELD, class F>
__host__ __device__ __forceinline__ float sl2_value(const F& u, const F& v, const F& src, int n, float dt_over_h,
                                                    int x, int y, bool& ok) {
    constexpr float ox = FIELD == kVelocityX ? 0.f : .5f;
    constexpr float oy = FIELD == kVelocityY ? 0.f : .5f;
    const int w = n + (FIELD == kVelocityX), h = n + (FIELD == kVelocityY);
    const float px = (float)x + ox, py = (float)y + oy;
    const float qx = py - .5f, qy = px - .5f;
    const float uu = sl2_sample(u, n + 1, n, px, qx);
    const float vv = sl2_sample(v, n, n + 1, qy, py);
    const float sx = px - dt_over_h * uu - ox;
    const float sy = py - dt_over_h * vv - oy;
    if (!(finitef(sx) && finitef(sy))) {
        ok = false;
        return 0.f;
    }
    return sl2_sample(src, w, h, sx, sy);
}

// ---------------------------------------------------------------------------
// ReflectionMacFluidSolver2D's samplers (ReflectionMacFluidSolver2D.hpp:113-147)
// for the unit box, h = 1/n, mix(a, b, t) = a + (b - a) t.
__host__ __device__ __forceinline__ float hmix(float a, float b, float t) { return a + (b - a) * t; }

struct Hist {
    int n;
    float h;  // 1.0f / n
    template <class F> __host__ __device__ __forceinline__ float cell(const F& f, float px, float py) const {
        const float gx = sclamp(px / h - 0.5f, 0.f, (float)(n - 1));
        const float gy = sclamp(py / h - 0.5f, 0.f, (float)(n - 1));
        const int x0 = trunc_int(gx), y0 = trunc_int(gy);
        const int x1 = smin(x0 + 1, n - 1), y1 = smin(y0 + 1, n - 1);
        const float tx = gx - (float)x0, ty = gy - (float)y0;
        return hmix(hmix(f[idC(n, x0, y0)], f[idC(n, x1, y0)], tx), hmix(f[idC(n, x0, y1)], f[idC(n, x1, y1)], tx),
                    ty);
    }
    template <class F> __host__ __device__ __forceinline__ float vx(const F& f, float px, float py) const {
        const float gx = sclamp(px / h, 0.f, (float)n);
        const float gy = sclamp(py / h - 0.5f, 0.f, (float)(n - 1));
        const int x0 = smin(trunc_int(gx), n - 1), y0 = trunc_int(gy);
        const int x1 = x0 + 1, y1 = smin(y0 + 1, n - 1);
        const float tx = gx - (float)x0, ty = gy - (float)y0;
        return hmix(hmix(f[idX(n, x0, y0)], f[idX(n, x1, y0)], tx), hmix(f[idX(n, x0, y1)], f[idX(n, x1, y1)], tx),
                    ty);
    }
    template <class F> __host__ __device__ __forceinline__ float vy(const F& f, float px, float py) const {
        const float gx = sclamp(px / h - 0.5f, 0.f, (float)(n - 1));
        const float gy = sclamp(py / h, 0.f, (float)n);
        const int x0 = trunc_int(gx), y0 = smin(trunc_int(gy), n - 1);
        const int x1 = smin(x0 + 1, n - 1), y1 = y0 + 1;
        const float tx = gx - (float)x0, ty = gy - (float)y0;
        return hmix(hmix(f[idY(n, x0, y0)], f[idY(n, x1, y0)], tx), hmix(f[idY(n, x0, y1)], f[idY(n, x1, y1)], tx),
                    ty);
    }
    // semiLagrangian (ReflectionMacFluidSolver2D.hpp:154-160) with sign -1;
    // sign +1 is the MacCormack back-trace (:179-183).
    template <int SIGN, class F>
    __host__ __device__ __forceinline__ float trace_cell(const F& src, const F& u, const F& v, int x, int y,
                                                         float dt) const {
        const float px = ((float)x + .5f) * h, py = ((float)y + .5f) * h;
        const float vxp = vx(u, px, py), vyp = vy(v, px, py);
        float qx, qy;
        if constexpr (SIGN < 0) {
            qx = px - vxp * dt;
            qy = py - vyp * dt;
        } else {
            qx = px + vxp * dt;
            qy = py + vyp * dt;
        }
        const float pad = 0.5f * h;
        qx = sclamp(qx, pad, 1.0f - pad);
        qy = sclamp(qy, pad, 1.0f - pad);
        return cell(src, qx, qy);
    }
    // advectVx / advectVy sample (:201-226 and :363-384): face positions,
    // flow (fu, fv), source src.
    template <class F>
    __host__ __device__ __forceinline__ float face_x(const F& src, const F& fu, const F& fv, int x, int y,
                                                     float dt) const {
        const float px = (float)x * h, py = ((float)y + .5f) * h;
        const float a = vx(fu, px, py), b = vy(fv, px, py);
        float qx = px - a * dt, qy = py - b * dt;
        qx = sclamp(qx, 0.f, 1.0f);
        qy = sclamp(qy, 0.5f * h, 1.0f - 0.5f * h);
        return vx(src, qx, qy);
    }
    template <class F>
    __host__ __device__ __forceinline__ float face_y(const F& src, const F& fu, const F& fv, int x, int y,
                                                     float dt) const {
        const float px = ((float)x + .5f) * h, py = (float)y * h;
        const float a = vx(fu, px, py), b = vy(fv, px, py);
        float qx = px - a * dt, qy = py - b * dt;
        qx = sclamp(qx, 0.5f * h, 1.0f - 0.5f * h);
        qy = sclamp(qy, 0.f, 1.0f);
        return vy(src, qx, qy);
    }
};

// The MacCormack limiter of ReflectionMacFluidSolver2D::advectScalar
// (:191-204) for cell (x, y): clamp of the corrected value to the min and max
// of the forward trace and the field over the 3 x 3 neighbourhood, in the
// reference's order (the field as it stands: limited for the neighbours the
// scan has passed, original for the rest).
template <class Fd, class Wd>
__host__ __device__ __forceinline__ float limit_cell(const Fd& field, const Wd& fwd, float corrected, int n, int x,
                                                     int y) {
    const int i = idC(n, x, y);
    float fmn = smin((float)field[i], (float)fwd[i]), fmx = smax((float)field[i], (float)fwd[i]);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int a = x + dx, b = y + dy;
            if (a >= 0 && b >= 0 && a < n && b < n) {
                const int k = idC(n, a, b);
                const float w = fwd[k], f = field[k];
                fmn = smin3(fmn, w, f);
                fmx = smax3(fmx, w, f);
            }
        }
    return sclamp(corrected, fmn, fmx);
}

// Rows [ylo, yhi] of the anti-diagonal x + 2y = t of an n x n grid. The
// limiter's cell (x, y) reads limited values of (x-1, y), (x+1, y-1), (x, y-1)
// and (x-1, y-1), which lie on diagonals t-1, t-1, t-2 and t-3, and original
// values of the rest, which lie on later diagonals: diagonal by diagonal, the
// cells see what the reference's scan order gives them.
__host__ __device__ __forceinline__ void limit_front(int t, int n, int& ylo, int& yhi) {
    ylo = t > n - 1 ? (t - (n - 1) + 1) / 2 : 0;
    yhi = smin(n - 1, t / 2);
}

// Whether cell (x, y) lies in the historical emitter's disc, and its weight
// (ReflectionMacFluidSolver2D.hpp:296-303).
__host__ __device__ __forceinline__ bool hist_in_disc(float cx, float cy, float radius, float falloff, int n,
                                                      float h, int x, int y, float& w) {
    const float px = ((float)x + .5f) * h, py = ((float)y + .5f) * h;
    const float dx2 = (px - cx) * (px - cx);
    const float dy2 = (py - cy) * (py - cy);
    const float q = (dx2 + dy2) / (radius * radius);
    if (q > 1.0f) return false;
    w = det_expf(-falloff * q);
    return true;
}

// The LCG of ReflectionMacFluidSolver2D::randomUnit (:109-112) and its k-step
// jump: s_k = A_k s + C_k with (A_k, C_k) from binary powering of (a, c).
constexpr uint32_t kLcgA = 1664525u, kLcgC = 1013904223u;
__host__ __device__ __forceinline__ void lcg_jump(uint32_t k, uint32_t& A, uint32_t& C) {
    uint32_t a = kLcgA, c = kLcgC;
    A = 1u;
    C = 0u;
    while (k) {
        if (k & 1u) {
            A = A * a;
            C = C * a + c;
        }
        c = c * a + c;
        a = a * a;
        k >>= 1;
    }
}
__host__ __device__ __forceinline__ float lcg_unit(uint32_t s) { return (float)(s >> 8) * (1.0f / 16777216.0f); }

}  // namespace vgx2
