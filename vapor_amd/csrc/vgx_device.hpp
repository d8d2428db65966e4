// vgx_device.hpp - device-side helpers shared by the MAC 3D kernels.
//
// Every float expression here keeps the operation order of vapor's CPU
// solver, and the library is compiled with -ffp-contract=off, so the GPU
// rounds exactly as the CPU reference does (IEEE division and square root are
// the HIP defaults on gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace vgx {

constexpr int kBlock = 256;  // elementwise / gather kernels: 4 waves of 64
constexpr int kWave = 64;

// std::clamp / std::min / std::max with the reference's comparison order.
template <class T> __host__ __device__ __forceinline__ T clampv(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
__host__ __device__ __forceinline__ float clampf(float v, float lo, float hi) { return clampv(v, lo, hi); }
__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
template <class T> __host__ __device__ __forceinline__ T minv(T a, T b) { return (b < a) ? b : a; }
template <class T> __host__ __device__ __forceinline__ T maxv(T a, T b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ float minf(float a, float b) { return minv(a, b); }
__host__ __device__ __forceinline__ float maxf(float a, float b) { return maxv(a, b); }
__host__ __device__ __forceinline__ int mini(int a, int b) { return (b < a) ? b : a; }

// Float -> int truncation as the reference's int(v), made total: NaN maps to 0
// and the magnitude is bounded, so no index is ever derived from a value the
// conversion leaves undefined. Identical to (int)v for finite in-range v.
__host__ __device__ __forceinline__ int trunc_index(float v) {
    return (v == v) ? (int)fmaxf(fminf(v, 1.0e9f), -1.0e9f) : 0;
}
__host__ __device__ __forceinline__ int trunc_index(double v) {
    return (v == v) ? (int)fmax(fmin(v, 1.0e9), -1.0e9) : 0;
}

// Correctly rounded square root in the field precision.
__host__ __device__ __forceinline__ float sqrtv(float v) { return sqrtf(v); }
__host__ __device__ __forceinline__ double sqrtv(double v) { return sqrt(v); }

// e^v for the capped-Gaussian emitter profile, whose argument -3.5 (d/r)^2
// lies in [-3.5, 0]: bit-identical on the CPU and the GPU because it uses
// only IEEE +, -, * and floor. Cody-Waite reduction v = k ln2 + r with
// k = floor(v log2(e) + 1/2) and ln2 split as C1 + C2 (C1 has 9 trailing zero
// bits, so k C1 is exact for |k| < 2^9), e^r by its Taylor series to r^7 on
// |r| <= ln2 / 2 (truncation below 6e-9 relative), then the exact scaling by
// 2^k. Valid for -87 <= v <= 88.
__host__ __device__ __forceinline__ float profile_exp(float v) {
    const float k = floorf(v * 1.44269504f + 0.5f);
    const float r = (v - k * 0.693145752f) - k * 1.42860677e-06f;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f +
                    r * (0.00833333377f + r * (0.00138888892f + r * 0.000198412701f))))));
    const uint32_t scale = (uint32_t)((int)k + 127) << 23;
    float s;
    memcpy(&s, &scale, 4);
    return p * s;
}
__host__ __device__ __forceinline__ float expv(float v) { return profile_exp(v); }
__host__ __device__ __forceinline__ double expv(double v) { return exp(v); }

// Index maps, MacGridDomain3D.hpp:22-31 (n^3 < 2^31 is enforced at creation).
__host__ __device__ __forceinline__ int idC(int n, int x, int y, int z) { return x + n * (y + n * z); }
__host__ __device__ __forceinline__ int idX(int n, int x, int y, int z) { return x + (n + 1) * (y + n * z); }
__host__ __device__ __forceinline__ int idY(int n, int x, int y, int z) { return x + n * (y + (n + 1) * z); }
__host__ __device__ __forceinline__ int idZ(int n, int x, int y, int z) { return x + n * (y + n * z); }

// AXIS: -1 cell centred, 0/1/2 staggered x/y/z faces.
template <int AXIS> __host__ __device__ __forceinline__ int idA(int n, int x, int y, int z) {
    if constexpr (AXIS == 0) return idX(n, x, y, z);
    else if constexpr (AXIS == 1) return idY(n, x, y, z);
    else return idC(n, x, y, z);
}

// a * b for 0 <= a, b < 2^24 with a product below 2^31: on the device one
// full-rate 24-bit multiply instead of a quarter-rate 32-bit one.
__host__ __device__ __forceinline__ int mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (int)__umul24((unsigned)a, (unsigned)b);
#else
    return a * b;
#endif
}
// idA for non-negative coordinates (a clamped sample corner or limiter
// neighbour): x + ex (y + ey z) with the field's x and y extents, the same
// index with 24-bit multiplies (n < 4096).
template <int AXIS> __host__ __device__ __forceinline__ int idA24(int n, int x, int y, int z) {
    const int ex = n + (AXIS == 0), ey = n + (AXIS == 1);
    return x + mul24(ex, y + mul24(ey, z));
}

// A field read through an optional reflection: value = 2 S[i] - T[i]
// (MacGridFluidSolver3D.cpp:211-218 materialises the same expression), in the
// precision of the field's elements.
template <bool HAT, class P = const float*> struct Src {
    P s;
    P t;
    __host__ __device__ __forceinline__ auto operator[](int i) const {
        using V = decltype(+s[i]);
        if constexpr (HAT) return V(2) * s[i] - t[i];
        else return V(s[i]);
    }
};

// p / h for the grid spacing h. With POW2 (h an exact power of two) it is
// p * (1 / h): both round the same real number p / h, so the value is the
// same, with one multiply in place of a division.
template <class T, bool POW2 = false> struct Spacing {
    T h, inv;
    __host__ __device__ __forceinline__ T over(T p) const {
        if constexpr (POW2) return p * inv;
        else return p / h;
    }
};

// Whether h is an exact power of two 2^k, -125 <= k <= 124, so that 1 / h is
// exact and normal (Spacing's POW2 path).
__host__ __device__ __forceinline__ bool pow2_spacing(float h) {
    unsigned u;
    memcpy(&u, &h, sizeof u);
    const unsigned e = (u >> 23) & 0xffu;
    return (u & 0x807fffffu) == 0u && e >= 2u && e <= 251u;
}

// int(v) for v in [0, 2^30] or NaN (a clamped sample coordinate): NaN maps
// to 0, as in trunc_index.
template <class T> __host__ __device__ __forceinline__ int trunc_clamped(T v) { return (v == v) ? (int)v : 0; }

// Clamped trilinear sample, MacGridSampler3D.cpp:5-71, in the precision T of
// the coordinates. On the staggered axis g = p/h in [0, n] with the lower
// index capped at n-1; elsewhere g = p/h - 0.5 in [0, n-1] with the upper
// index capped at n-1. The eight corners are the first corner's index plus
// the x, row and plane steps (idA is linear in x, y, z).
template <int AXIS, class F, class T, bool POW2>
__host__ __device__ __forceinline__ T sample(const F& f, int n, const Spacing<T, POW2>& sp, T px, T py, T pz) {
    const T fn = (T)n, fn1 = (T)(n - 1), half = (T)0.5, zero = (T)0;
    T gx, gy, gz;
    int x0, y0, z0, x1, y1, z1;
    if constexpr (AXIS == 0) {
        gx = clampv(sp.over(px), zero, fn);
        x0 = mini(trunc_clamped(gx), n - 1);
        x1 = x0 + 1;
    } else {
        gx = clampv(sp.over(px) - half, zero, fn1);
        x0 = trunc_clamped(gx);
        x1 = mini(x0 + 1, n - 1);
    }
    if constexpr (AXIS == 1) {
        gy = clampv(sp.over(py), zero, fn);
        y0 = mini(trunc_clamped(gy), n - 1);
        y1 = y0 + 1;
    } else {
        gy = clampv(sp.over(py) - half, zero, fn1);
        y0 = trunc_clamped(gy);
        y1 = mini(y0 + 1, n - 1);
    }
    if constexpr (AXIS == 2) {
        gz = clampv(sp.over(pz), zero, fn);
        z0 = mini(trunc_clamped(gz), n - 1);
        z1 = z0 + 1;
    } else {
        gz = clampv(sp.over(pz) - half, zero, fn1);
        z0 = trunc_clamped(gz);
        z1 = mini(z0 + 1, n - 1);
    }
    const T tx = gx - (T)x0, ty = gy - (T)y0, tz = gz - (T)z0;
    const int row = AXIS == 0 ? n + 1 : n, plane = AXIS == 0 || AXIS == 1 ? n * (n + 1) : n * n;
    const int i0 = idA24<AXIS>(n, x0, y0, z0);
    const int ox = x1 - x0, oy = y1 != y0 ? row : 0, oz = z1 != z0 ? plane : 0;
    const T f000 = f[i0], f100 = f[i0 + ox];
    const T f010 = f[i0 + oy], f110 = f[i0 + ox + oy];
    const T f001 = f[i0 + oz], f101 = f[i0 + ox + oz];
    const T f011 = f[i0 + oy + oz], f111 = f[i0 + ox + oy + oz];
    return f000 * (1 - tx) * (1 - ty) * (1 - tz) + f100 * tx * (1 - ty) * (1 - tz) +
           f010 * (1 - tx) * ty * (1 - tz) + f110 * tx * ty * (1 - tz) +
           f001 * (1 - tx) * (1 - ty) * tz + f101 * tx * (1 - ty) * tz +
           f011 * (1 - tx) * ty * tz + f111 * tx * ty * tz;
}
template <int AXIS, class F, class T>
__host__ __device__ __forceinline__ T sample(const F& f, int n, T h, T px, T py, T pz) {
    return sample<AXIS>(f, n, Spacing<T, false>{h, (T)0}, px, py, pz);
}

// Sample position of entry (x, y, z): cell centre or face centre
// (CpuAdvectionKernels.cpp:25, :91, :154, :217).
template <int AXIS, class T>
__host__ __device__ __forceinline__ void sample_point(T dx, int x, int y, int z, T& px, T& py, T& pz) {
    const T half = (T)0.5;
    px = AXIS == 0 ? (T)x * dx : ((T)x + half) * dx;
    py = AXIS == 1 ? (T)y * dx : ((T)y + half) * dx;
    pz = AXIS == 2 ? (T)z * dx : ((T)z + half) * dx;
}

// Decomposition of a linear index of field AXIS into (x, y, z).
template <int AXIS> __host__ __device__ __forceinline__ void decompose(int i, int n, int& x, int& y, int& z) {
    const int ex = n + (AXIS == 0), ey = n + (AXIS == 1);
    x = i % ex;
    const int t = i / ex;
    y = t % ey;
    z = t / ey;
}

// Forward (semi-Lagrangian) trace of entry i of field AXIS,
// CpuAdvectionKernels.cpp:22-28 and :88-95 (same for y, z).
template <int AXIS, class U, class S, class T, bool POW2>
__host__ __device__ __forceinline__ T sl_value(const U& u, const U& v, const U& w, const S& src, int n,
                                               const Spacing<T, POW2>& sp, T dt, int i) {
    int x, y, z;
    decompose<AXIS>(i, n, x, y, z);
    T px, py, pz;
    sample_point<AXIS>(sp.h, x, y, z, px, py, pz);
    const T vx = sample<0>(u, n, sp, px, py, pz);
    const T vy = sample<1>(v, n, sp, px, py, pz);
    const T vz = sample<2>(w, n, sp, px, py, pz);
    return sample<AXIS>(src, n, sp, px - vx * dt, py - vy * dt, pz - vz * dt);
}
template <int AXIS, class U, class S, class T>
__host__ __device__ __forceinline__ T sl_value(const U& u, const U& v, const U& w, const S& src, int n, T h, T dt,
                                               int i) {
    return sl_value<AXIS>(u, v, w, src, n, Spacing<T, false>{h, (T)0}, dt, i);
}

// MacCormack correction with its fallback and limiter for entry i:
// CpuAdvectionKernels.cpp:45-78 (scalar), :111-142 (u), :174-205 (v),
// :237-268 (w). The backward trace of the forward field is evaluated only
// where the correction is used.
template <int AXIS, class U, class S, class W, class T, bool POW2>
__host__ __device__ __forceinline__ T mc_value(const U& u, const U& v, const U& w, const S& src, const W& fwd, int n,
                                               const Spacing<T, POW2>& sp, T box, T dt, T cflMc, int i) {
    const T half = (T)0.5;
    int x, y, z;
    decompose<AXIS>(i, n, x, y, z);
    T px, py, pz;
    sample_point<AXIS>(sp.h, x, y, z, px, py, pz);
    const T vx = sample<0>(u, n, sp, px, py, pz);
    const T vy = sample<1>(v, n, sp, px, py, pz);
    const T vz = sample<2>(w, n, sp, px, py, pz);
    const T fx = px - vx * dt, fy = py - vy * dt, fz = pz - vz * dt;
    const T bx = px + vx * dt, by = py + vy * dt, bz = pz + vz * dt;
    const T cfl = sp.over(sqrtv(vx * vx + vy * vy + vz * vz) * dt);
    if (fx < 0 || fx > box || fy < 0 || fy > box || fz < 0 || fz > box || bx < 0 || bx > box || by < 0 ||
        by > box || bz < 0 || bz > box || cfl > cflMc)
        return fwd[i];
    const T back = sample<AXIS>(fwd, n, sp, bx, by, bz);
    const T corrected = fwd[i] + half * (src[i] - back);
    const int h0 = AXIS == 0 ? n : n - 1, h1 = AXIS == 1 ? n : n - 1, h2 = AXIS == 2 ? n : n - 1;
    const int g0 = clampi(AXIS == 0 ? trunc_index(sp.over(fx)) : trunc_index(sp.over(fx) - half), 0, h0);
    const int g1 = clampi(AXIS == 1 ? trunc_index(sp.over(fy)) : trunc_index(sp.over(fy) - half), 0, h1);
    const int g2 = clampi(AXIS == 2 ? trunc_index(sp.over(fz)) : trunc_index(sp.over(fz) - half), 0, h2);
    T fMin = src[idA24<AXIS>(n, g0, g1, g2)], fMax = fMin;
#pragma unroll
    for (int dz = 0; dz <= 1; ++dz)
#pragma unroll
        for (int dy = 0; dy <= 1; ++dy)
#pragma unroll
            for (int d0 = 0; d0 <= 1; ++d0) {
                const T val =
                    src[idA24<AXIS>(n, clampi(g0 + d0, 0, h0), clampi(g1 + dy, 0, h1), clampi(g2 + dz, 0, h2))];
                fMin = minv(fMin, val);
                fMax = maxv(fMax, val);
            }
    return clampv(corrected, fMin, fMax);
}
template <int AXIS, class U, class S, class W, class T>
__host__ __device__ __forceinline__ T mc_value(const U& u, const U& v, const U& w, const S& src, const W& fwd, int n,
                                               T dx, T box, T dt, T cflMc, int i) {
    return mc_value<AXIS>(u, v, w, src, fwd, n, Spacing<T, false>{dx, (T)0}, box, dt, cflMc, i);
}

// ---- initial velocity noise: Philox4x32-10 + Box-Muller ----
// Counter-based normal variates, bit-identical on the CPU and the GPU: the
// generator is Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), and the
// Box-Muller transform uses only IEEE +, -, *, / and sqrt (correctly rounded
// on both), with log and cos evaluated by fixed polynomials below.
struct Philox4 {
    uint32_t v[4];
};
__host__ __device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                         uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    return Philox4{{c0, c1, c2, c3}};
}

__host__ __device__ __forceinline__ float bits_to_float(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}
__host__ __device__ __forceinline__ uint32_t float_to_bits(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

// ln(x) for a normal positive float: x = m 2^e with m in (sqrt(1/2), sqrt(2)],
// ln m = 2 atanh(s) with s = (m - 1)/(m + 1), |s| <= 0.172, by its series to s^11.
__host__ __device__ __forceinline__ float noise_log(float x) {
    const uint32_t b = float_to_bits(x);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = bits_to_float((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f), s2 = s * s;
    const float p = s2 * (0.333333333f + s2 * (0.2f + s2 * (0.142857143f + s2 * (0.111111111f + s2 * 0.0909090909f))));
    const float lnm = 2.0f * s + 2.0f * s * p;
    return (float)e * 0.693147182f + lnm;
}

// cos(2 pi t) for t in [0, 1): quadrant q = floor(4t), x = (4t - q) pi/2 in
// [0, pi/2), and the Taylor polynomials of cos and sin to x^12 and x^13.
__host__ __device__ __forceinline__ float noise_cos2pi(float t) {
    const float r = 4.0f * t;
    const int q = (int)r;
    const float x = (r - (float)q) * 1.57079637f, x2 = x * x;
    const float c = 1.0f + x2 * (-0.5f + x2 * (0.0416666667f + x2 * (-0.00138888889f + x2 * (2.48015873e-05f +
                    x2 * (-2.75573192e-07f + x2 * 2.08767570e-09f)))));
    const float sn = x * (1.0f + x2 * (-0.166666667f + x2 * (0.00833333333f + x2 * (-0.000198412698f +
                     x2 * (2.75573192e-06f + x2 * (-2.50521084e-08f + x2 * 1.60590438e-10f))))));
    return q == 0 ? c : q == 1 ? -sn : q == 2 ? -c : sn;
}

// The standard normal variate of entry i of field `stream` under `seed`:
// Philox counter (i, stream, 0, 0), key (seed low, seed high); u0 in (0, 1]
// and u1 in [0, 1) from the top 24 bits of the first two words.
__host__ __device__ __forceinline__ float noise_normal(uint64_t seed, uint32_t stream, uint32_t i) {
    const Philox4 r = philox4x32_10(i, stream, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u0 = (float)((r.v[0] >> 8) + 1u) * 5.96046448e-08f;
    const float u1 = (float)(r.v[1] >> 8) * 5.96046448e-08f;
    return sqrtf(-2.0f * noise_log(u0)) * noise_cos2pi(u1);
}

// ---- fused red-black sweep: tile geometry and phases ----
// A block owns a kRbTx x kRbTy x kRbTz tile (kRbTx = 64 = one wave of lanes)
// and stages in LDS the input iterate over the tile and a halo (4 cells in x,
// 2 in y and z) and the right-hand side at the cells the update phases
// update; the update phases then read LDS only.
//
// LDS layout. A staged row (fixed y, z) holds the 72 x-positions X0-4 ..
// X0+67 split by x-parity: words [0, 36) hold the even x (E[j] = X0-4+2j),
// words [36, 72) the odd x (O[j] = X0-3+2j); the tile's cells are E[2..33]
// and O[2..33]. The cells one colour updates in a row are 32 consecutive
// words, as are each of their six neighbours, and all seven sit at fixed
// offsets from the cell's word:
//   E[j]: x-1 = O[j-1] (+35), x+1 = O[j] (+36);  O[j]: x-1 = E[j] (-36), x+1 = E[j+1] (-35);
//   y+-1: +-kRbRow;  z+-1: +-kRbPlane.
// The float4 of x-positions X0-4+4q .. X0-1+4q is (E[2q], O[2q], E[2q+1],
// O[2q+1]): staging and writing move whole float4s, two 8-byte LDS words each.
// Rows are numbered ly + kRbHy * lz over the staged y and z (offsets -2 ..).
// The right-hand side uses the same row pitch over planes lz = 1 .. kRbTz + 2,
// so its word for a cell is the iterate's word minus kRbPlane, in its own array.
//
// Cells outside the domain are staged as +0.0 and never written, so an update
// adds every neighbour: the reference skips an absent neighbour's term, and
// adding -1 * (+0.0) = -0.0 leaves any partial sum, signed zeros included,
// bit-identical to skipping it.
//
// The phases are __host__ __device__ over accessor types so that a CPU build
// can run them, lane by lane, with bounds-checked accessors.
constexpr int kRbTx = 64, kRbTy = 8, kRbTz = 8, kRbThreads = 1024, kRbWaves = kRbThreads / 64;
constexpr int kRbHx = kRbTx + 8, kRbHy = kRbTy + 4, kRbHz = kRbTz + 4;
constexpr int kRbHalf = kRbHx / 2, kRbRow = kRbHx, kRbPlane = kRbRow * kRbHy;
constexpr int kRbQ = kRbHx / 4;                    // float4 per staged row
constexpr int kRbXWords = kRbPlane * kRbHz;        // staged iterate
constexpr int kRbBWords = kRbPlane * (kRbTz + 2);  // staged right-hand side
constexpr int kRbBRows = kRbTy * kRbTz + 2 * kRbTz + 2 * kRbTy;  // rows holding updated cells

struct RbTile {
    int X0, Y0, Z0;
    bool valid;
};

__host__ __device__ __forceinline__ int rb_tiles(int n) {
    return ((n + kRbTx - 1) / kRbTx) * ((n + kRbTy - 1) / kRbTy) * ((n + kRbTz - 1) / kRbTz);
}
// Grid size: a multiple of 8 (one share per XCD).
__host__ __device__ __forceinline__ int rb_blocks(int n) { return 8 * ((rb_tiles(n) + 7) / 8); }

// Tile counts of an n^3 sweep and, for the two divisions of rb_tile, the
// multipliers m = ceil(2^40 / d): floor(t / d) = (t m) >> 40 exactly for
// t, d < 2^20 (the error t (m d - 2^40) / (d 2^40) stays below 1 / d), which
// holds for every n with n^3 < 2^31. Computed once per launch on the host, so
// a block finds its tile with a few scalar instructions instead of three
// runtime integer divisions.
struct RbGeom {
    int nbx, nby, per, tiles;
    unsigned long long mx, mxy;  // multipliers for d = nbx and d = nbx * nby
};
__host__ __forceinline__ RbGeom rb_geom(int n) {
    RbGeom g;
    g.nbx = (n + kRbTx - 1) / kRbTx;
    g.nby = (n + kRbTy - 1) / kRbTy;
    g.tiles = rb_tiles(n);
    g.per = (g.tiles + 7) / 8;
    const unsigned long long one = 1ull << 40;
    g.mx = (one + (unsigned long long)g.nbx - 1) / (unsigned long long)g.nbx;
    const unsigned long long dxy = (unsigned long long)g.nbx * (unsigned long long)g.nby;
    g.mxy = (one + dxy - 1) / dxy;
    return g;
}

// Block b takes tile (b % 8) * ceil(T / 8) + b / 8; tiles are numbered
// x-fastest, then y, then z, so the blocks the hardware deals to one XCD
// (b % 8 equal) cover one contiguous z-slab.
// (Without a branch, so that the kernel loads all its arguments at once; an
// invalid tile's origin is never used.)
__host__ __device__ __forceinline__ RbTile rb_tile(const RbGeom& g, int block) {
    const int t = (block & 7) * g.per + (block >> 3);
    const int qx = (int)(((unsigned long long)t * g.mx) >> 40);    // t / nbx
    const int qxy = (int)(((unsigned long long)t * g.mxy) >> 40);  // t / (nbx nby)
    return RbTile{(t - qx * g.nbx) * kRbTx, (qx - qxy * g.nby) * kRbTy, qxy * kRbTz, t < g.tiles};
}
// The same with the divisions done in place (the batch's scenes each have
// their own n).
__host__ __device__ __forceinline__ RbTile rb_tile(int n, int block) {
    const int nbx = (n + kRbTx - 1) / kRbTx, nby = (n + kRbTy - 1) / kRbTy;
    const int tiles = rb_tiles(n), per = (tiles + 7) / 8;
    const int t = (block % 8) * per + block / 8;
    RbTile r{0, 0, 0, t < tiles};
    if (r.valid) {
        r.X0 = (t % nbx) * kRbTx;
        r.Y0 = ((t / nbx) % nby) * kRbTy;
        r.Z0 = (t / (nbx * nby)) * kRbTz;
    }
    return r;
}

// A value the whole wave holds (the wave index): kept in a scalar register.
__host__ __device__ __forceinline__ int rb_uniform(int v) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// Word of the lane's x (X0 + lane) in staged row `row`.
__host__ __device__ __forceinline__ int rb_lane_word(int row, int lane) {
    return row * kRbRow + (lane & 1) * kRbHalf + (lane >> 1) + 2;
}

// Number of neighbours a cell has in the domain (the matrix diagonal).
__host__ __device__ __forceinline__ int rb_degree(int gx, int gy, int gz, int n) {
    return (gx > 0) + (gx + 1 < n) + (gy > 0) + (gy + 1 < n) + (gz > 0) + (gz + 1 < n);
}

// Four consecutive floats from p[i]: one 16-byte load for a raw pointer
// (callers guarantee i % 4 == 0 with 16-byte aligned rows), element-wise for
// the host check's accessors.
__host__ __device__ __forceinline__ float4 rb_load4(const float* p, int i) {
    return *reinterpret_cast<const float4*>(p + i);
}
template <class A> __host__ __device__ __forceinline__ float4 rb_load4(const A& a, int i) {
    return make_float4(a[i], a[i + 1], a[i + 2], a[i + 3]);
}
// The tile write is non-temporal: the next sweep, a separate launch that may
// run on another XCD, reads it from beyond this XCD's L2 anyway, and the
// stores then skip the L2 write-back at the end of the launch (about 5% of a
// 128^3 sweep, measured).
__host__ __device__ __forceinline__ void rb_store4(float* p, int i, float4 v) {
#ifdef __HIP_DEVICE_COMPILE__
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(p + i));
#else
    *reinterpret_cast<float4*>(p + i) = v;
#endif
}
template <class A> __host__ __device__ __forceinline__ void rb_store4(A& a, int i, float4 v) {
    a[i] = v.x;
    a[i + 1] = v.y;
    a[i + 2] = v.z;
    a[i + 3] = v.w;
}

// Correctly rounded a / D for a diagonal D in {1, ..., 6}, given r = 1.0f / D
// rounded to nearest: q = a r, e = a - q D (exact, one FMA), q + e r (one
// FMA). This estimate equals the IEEE quotient for 2^-125 <= |a| < inf
// (checked over all 2^32 floats for each such D). Where the estimate is a
// normal number, |a| >= D 2^-126 >= 2^-125 for D >= 2, and D = 1 is exact
// everywhere, so the estimate is kept exactly when it is normal (one class
// test); zeros, subnormals, infinities and NaN take the IEEE division.
__host__ __device__ __forceinline__ float rb_div_estimate(float a, float D, float r) {
    const float q = a * r;
    const float e = fmaf(-q, D, a);
    return fmaf(e, r, q);
}
// The IEEE quotient where rb_div_estimate's value v is not normal.
__host__ __device__ __forceinline__ float rb_div_fix(float a, float D, float v) {
    if (!__builtin_isnormal(v)) {
#ifdef __HIP_DEVICE_COMPILE__
        // keeps the rare IEEE division behind its branch instead of being
        // computed for every lane and selected (an empty asm, no instruction)
        asm volatile("" : "+v"(a));
#endif
        v = a / D;
    }
    return v;
}
__host__ __device__ __forceinline__ float rb_div(float a, float D, float r) {
    return rb_div_fix(a, D, rb_div_estimate(a, D, r));
}

// Sweep constants. With FAST kernels the diagonal max(d, floor) is the same
// D for every cell (floor = 6 >= every degree) and r = 1.0f / D.
struct RbArgs {
    float omega, floor, D, r;
};

// x-neighbour offset of a cell in the even (S = 0) or odd (S = 1) half-row.
template <int S> constexpr int rb_dm() { return S ? -kRbHalf : kRbHalf - 1; }

// Right-hand side of an update: the staged b' word of the cell (LDS), or,
// for the register path, the b value the lane loaded for pass p at the start
// of the launch, b' = b - meanf formed here (the same FP32 subtraction the
// staging does).
template <int NP> struct RbBReg {
    float v[NP];
    float meanf;
};
template <class BS> __host__ __device__ __forceinline__ float rb_bval(const BS& bs, int, int w) {
    return bs[w - kRbPlane];
}
template <int NP> __host__ __device__ __forceinline__ float rb_bval(const RbBReg<NP>& r, int p, int) {
    return r.v[p] - r.meanf;
}

// Tiles whose update rows (the tile and its 1-cell ring in y and z) lie at
// least one cell inside the domain in y and z: every updated cell there has
// both y- and both z-neighbours.
__host__ __device__ __forceinline__ bool rb_inner(int n, const RbTile& t) {
    return t.Y0 >= 2 && t.Z0 >= 2 && t.Y0 + kRbTy + 2 <= n && t.Z0 + kRbTz + 2 <= n;
}

// NP update passes of x-parity S, run together: pass p gives lanes 0-31 the
// 32 cells of its colour in row (ly[p], lz[p]) and lanes 32-63 those of row
// (ly[p] + dly, lz[p] + dlz). Each cell gets the reference's
//   x += omega * (b - sum(offdiag x) - d x) / max(d, floor)
// with the neighbour terms in its order (SOR.cpp:36-49). All LDS reads come
// before any write (a pass reads only the other colour and its own cells), so
// the passes' dependency chains interleave. A lane whose cell lies outside
// the domain writes back the +0.0 it read.
template <int S, bool FAST, bool INNER, int NP, class XS, class BS>
__host__ __device__ __forceinline__ void rb_passes(XS& xs, const BS& bs, const RbArgs& a, int n, const RbTile& t,
                                                   const int (&ly)[NP], const int (&lz)[NP], int dly, int dlz,
                                                   int lane) {
    constexpr int dm = rb_dm<S>();
    const int half = lane >> 5, j = (lane & 31) + 2;
    const int gx = t.X0 - 4 + 2 * j + S;  // in [X0, X0 + 63]
    const int lane_w = half * (dly * kRbRow + dlz * kRbPlane) + S * kRbHalf + j;
    int w[NP];
    bool ok[NP];
    float d[NP], c[NP], off[NP], bi[NP], v[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        w[p] = (ly[p] + kRbHy * lz[p]) * kRbRow + lane_w;
        const int gy = t.Y0 - 2 + ly[p] + half * dly, gz = t.Z0 - 2 + lz[p] + half * dlz;
        if (INNER) {
            ok[p] = gx < n;
            d[p] = (float)((gx > 0) + (gx + 1 < n) + 4);
        } else {
            ok[p] = gx < n && gy >= 0 && gz >= 0 && gy < n && gz < n;
            d[p] = (float)rb_degree(gx, gy, gz, n);
        }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        c[p] = xs[w[p]];
        float o = 0.f;
        o += -1.f * xs[w[p] + dm];
        o += -1.f * xs[w[p] + dm + 1];
        o += -1.f * xs[w[p] - kRbRow];
        o += -1.f * xs[w[p] + kRbRow];
        o += -1.f * xs[w[p] - kRbPlane];
        o += -1.f * xs[w[p] + kRbPlane];
        off[p] = o;
        bi[p] = rb_bval(bs, p, w[p]);
    }
    float tp[NP], q[NP];
    bool slow = false;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        tp[p] = a.omega * (bi[p] - off[p] - d[p] * c[p]);
        if (FAST) {
            q[p] = rb_div_estimate(tp[p], a.D, a.r);
            slow = slow || !__builtin_isnormal(q[p]);
        } else {
            q[p] = tp[p] / maxf(d[p], a.floor);
        }
    }
    // one branch for the passes' rare non-normal quotients (rb_div)
    if (FAST && slow) {
#pragma unroll
        for (int p = 0; p < NP; ++p) q[p] = rb_div_fix(tp[p], a.D, q[p]);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) v[p] = c[p] + q[p];
#pragma unroll
    for (int p = 0; p < NP; ++p) xs[w[p]] = ok[p] ? v[p] : c[p];
}

// The NP passes of colour COLOR at rows (ly[p], lz[p]) of one parity.
template <int COLOR, bool FAST, bool INNER, int NP, class XS, class BS>
__host__ __device__ __forceinline__ void rb_passes_at(XS& xs, const BS& bs, const RbArgs& a, int n, const RbTile& t,
                                                      const int (&ly)[NP], const int (&lz)[NP], int dly, int dlz,
                                                      int lane) {
    const int S = ((t.Y0 + ly[0] + t.Z0 + lz[0]) & 1) ^ COLOR;  // red: even x in an even row
    if (S) rb_passes<1, FAST, INNER, NP>(xs, bs, a, n, t, ly, lz, dly, dlz, lane);
    else rb_passes<0, FAST, INNER, NP>(xs, bs, a, n, t, ly, lz, dly, dlz, lane);
}

// One update of the cell at word w with x-neighbours at w + dm, w + dm + 1
// (the x-ring cells, whose parity differs from lane to lane).
template <bool FAST, class XS, class BS>
__host__ __device__ __forceinline__ void rb_cell(XS& xs, const BS& bs, int w, int dm, float d, const RbArgs& a) {
    const float c = xs[w];
    float off = 0.f;
    off += -1.f * xs[w + dm];
    off += -1.f * xs[w + dm + 1];
    off += -1.f * xs[w - kRbRow];
    off += -1.f * xs[w + kRbRow];
    off += -1.f * xs[w - kRbPlane];
    off += -1.f * xs[w + kRbPlane];
    const float bi = rb_bval(bs, 0, w);
    const float tp = a.omega * (bi - off - d * c);
    xs[w] = c + (FAST ? rb_div(tp, a.D, a.r) : tp / maxf(d, a.floor));
}

// Staged row (ly, lz) of the right-hand side's r-th row: the 64 tile rows,
// then the y-ring rows (ly = 1, 10; lz = 2..9), then the z-ring rows
// (lz = 1, 10; ly = 2..9).
static_assert(kRbTy == 8 && kRbTz == 8 && kRbWaves == 16, "row schedules below are written for 8 x 8 rows, 16 waves");
__host__ __device__ __forceinline__ void rb_brow(int r, int& ly, int& lz) {
    if (r < 64) {
        ly = 2 + (r & 7);
        lz = 2 + (r >> 3);
    } else if (r < 80) {
        ly = r < 72 ? 1 : 10;
        lz = 2 + (r & 7);
    } else {
        lz = r < 88 ? 1 : 10;
        ly = 2 + (r & 7);
    }
}

// Phase 1: stage xin over the tile and its halo and b' = b - meanf on the
// rows the update phases update (+0.0 outside the domain), one float4 of
// each row per lane and step. VEC (n % 4 == 0, so every float4 lies wholly in
// or out of the domain and rows are 16-byte aligned) loads whole float4s;
// otherwise each element is loaded and bounds-checked alone. All loads are
// issued before the first LDS store.
// The float4 of one staged row segment: element i of the field and, per
// component, whether it lies in the domain (and the lane has work: live).
struct RbSpan {
    int i;
    bool o[4];
};
__host__ __device__ __forceinline__ RbSpan rb_span(int n, int gx, int gy, int gz, bool live) {
    const bool row_in = live && gy >= 0 && gz >= 0 && gy < n && gz < n;
    RbSpan s;
    s.i = idC(n, gx, gy, gz);
#pragma unroll
    for (int k = 0; k < 4; ++k) s.o[k] = row_in && gx + k >= 0 && gx + k < n;
    return s;
}
// Staged iterate row `row` (ly + kRbHy * lz), float4 q; right-hand-side row
// (ly, lz), float4 q.
__host__ __device__ __forceinline__ RbSpan rb_x_span(int n, const RbTile& t, int row, int q, bool live) {
    return rb_span(n, t.X0 - 4 + 4 * q, t.Y0 - 2 + row % kRbHy, t.Z0 - 2 + row / kRbHy, live);
}
__host__ __device__ __forceinline__ RbSpan rb_b_span(int n, const RbTile& t, int ly, int lz, int q, bool live) {
    return rb_span(n, t.X0 - 4 + 4 * q, t.Y0 - 2 + ly, t.Z0 - 2 + lz, live);
}
// Every lane loads (from element 0 where a value lies outside the domain) and
// selects afterwards (rb_keep): a load under a branch makes the compiler wait
// for it before the branch joins, one memory round trip per load. VEC
// (n % 4 == 0): the four components are in or out together.
template <bool VEC, class P> __host__ __device__ __forceinline__ float4 rb_raw(const P& p, const RbSpan& s) {
    if (VEC) return rb_load4(p, s.o[0] ? s.i : 0);
    float4 u;
    u.x = p[s.o[0] ? s.i : 0];
    u.y = p[s.o[1] ? s.i + 1 : 0];
    u.z = p[s.o[2] ? s.i + 2 : 0];
    u.w = p[s.o[3] ? s.i + 3 : 0];
    return u;
}
// The staged iterate (+0.0 outside the domain) and b' = b - meanf (+0.0 outside).
template <bool VEC> __host__ __device__ __forceinline__ float4 rb_keep(float4 u, const RbSpan& s) {
    float4 v;
    v.x = (VEC ? s.o[0] : s.o[0]) ? u.x : 0.f;
    v.y = (VEC ? s.o[0] : s.o[1]) ? u.y : 0.f;
    v.z = (VEC ? s.o[0] : s.o[2]) ? u.z : 0.f;
    v.w = (VEC ? s.o[0] : s.o[3]) ? u.w : 0.f;
    return v;
}
template <bool VEC> __host__ __device__ __forceinline__ float4 rb_keep_b(float4 u, const RbSpan& s, float meanf) {
    float4 w;
    w.x = (VEC ? s.o[0] : s.o[0]) ? u.x - meanf : 0.f;
    w.y = (VEC ? s.o[0] : s.o[1]) ? u.y - meanf : 0.f;
    w.z = (VEC ? s.o[0] : s.o[2]) ? u.z - meanf : 0.f;
    w.w = (VEC ? s.o[0] : s.o[3]) ? u.w - meanf : 0.f;
    return w;
}
template <bool VEC, class XIN>
__host__ __device__ __forceinline__ float4 rb_fetch_x(const XIN& xin, int n, const RbTile& t, int row, int q,
                                                      bool live) {
    const RbSpan s = rb_x_span(n, t, row, q, live);
    return rb_keep<VEC>(rb_raw<VEC>(xin, s), s);
}
template <bool VEC, class B>
__host__ __device__ __forceinline__ float4 rb_fetch_b(const B& b, float meanf, int n, const RbTile& t, int ly, int lz,
                                                      int q, bool live) {
    const RbSpan s = rb_b_span(n, t, ly, lz, q, live);
    return rb_keep_b<VEC>(rb_raw<VEC>(b, s), s, meanf);
}
// LDS stores of a staged float4: x-positions 4q .. 4q+3 of the row go to
// E[2q], O[2q], E[2q+1], O[2q+1].
template <class XS> __host__ __device__ __forceinline__ void rb_put(XS& xs, int wd, float4 v) {
    xs[wd] = v.x;
    xs[wd + 1] = v.z;
    xs[wd + kRbHalf] = v.y;
    xs[wd + kRbHalf + 1] = v.w;
}
__host__ __device__ __forceinline__ int rb_x_word(int row, int q) { return row * kRbRow + 2 * q; }
__host__ __device__ __forceinline__ int rb_b_word(int ly, int lz, int q) {
    return (ly + kRbHy * lz) * kRbRow + 2 * q - kRbPlane;
}

template <bool VEC, class XS, class BS, class XIN, class B>
__host__ __device__ __forceinline__ void rb_stage(XS& xs, BS& bs, const XIN& xin, const B& b, float meanf, int n,
                                                  const RbTile& t, int tid) {
    constexpr int kXVec = kRbHy * kRbHz * kRbQ, kXIter = (kXVec + kRbThreads - 1) / kRbThreads;
    constexpr int kBVec = kRbBRows * kRbQ, kBIter = (kBVec + kRbThreads - 1) / kRbThreads;
    float4 v[kXIter], w[kBIter];
#pragma unroll
    for (int k = 0; k < kXIter; ++k) {
        const int e = tid + kRbThreads * k, row = e / kRbQ, q = e - row * kRbQ;
        v[k] = rb_fetch_x<VEC>(xin, n, t, row, q, e < kXVec);
    }
#pragma unroll
    for (int k = 0; k < kBIter; ++k) {
        const int e = tid + kRbThreads * k, r = e / kRbQ, q = e - r * kRbQ;
        int ly, lz;
        rb_brow(r, ly, lz);
        w[k] = rb_fetch_b<VEC>(b, meanf, n, t, ly, lz, q, e < kBVec);
    }
#pragma unroll
    for (int k = 0; k < kXIter; ++k) {
        const int e = tid + kRbThreads * k, row = e / kRbQ, q = e - row * kRbQ;
        if (e < kXVec) rb_put(xs, rb_x_word(row, q), v[k]);
    }
#pragma unroll
    for (int k = 0; k < kBIter; ++k) {
        const int e = tid + kRbThreads * k, r = e / kRbQ, q = e - r * kRbQ;
        if (e < kBVec) {
            int ly, lz;
            rb_brow(r, ly, lz);
            rb_put(bs, rb_b_word(ly, lz, q), w[k]);
        }
    }
}

// Phase 1b (optional): FP64 residual of the staged iterate on the tile, the
// seven partial fields of the residual check (StencilLinearSystem.cpp:60-67
// term order; absent neighbours add -0.0, which changes no sum). Without
// RHS_STATS, fields 2, 3 and 6 (sums of b'^2, x and b') are left alone: a
// tolerance solve's chunk checks take b''s from the solve's start residual,
// since b does not change during the solve.
template <bool RHS_STATS = true, class XS, class BS>
__host__ __device__ __forceinline__ void rb_residual(const XS& xs, const BS& bs, int n, const RbTile& t, int tid,
                                                     double* acc) {
    const int wave = rb_uniform(tid >> 6), lane = tid & 63;
    const int gx = t.X0 + lane;
    const int dm = (lane & 1) ? -kRbHalf : kRbHalf - 1;
    for (int row = wave; row < kRbTy * kRbTz; row += kRbWaves) {
        const int ly = row % kRbTy, lz = row / kRbTy, gy = t.Y0 + ly, gz = t.Z0 + lz;
        if (gy >= n || gz >= n || gx >= n) continue;
        const int w = rb_lane_word(ly + 2 + kRbHy * (lz + 2), lane);
        const float c = xs[w];
        double ax = (double)rb_degree(gx, gy, gz, n) * (double)c;
        ax += -1.0 * (double)xs[w + dm];
        ax += -1.0 * (double)xs[w + dm + 1];
        ax += -1.0 * (double)xs[w - kRbRow];
        ax += -1.0 * (double)xs[w + kRbRow];
        ax += -1.0 * (double)xs[w - kRbPlane];
        ax += -1.0 * (double)xs[w + kRbPlane];
        const float bi = bs[w - kRbPlane];
        const double res = (double)bi - ax;
        acc[0] += res * res;
        const double ar = res < 0 ? -res : res;
        acc[1] = acc[1] < ar ? ar : acc[1];
        if (RHS_STATS) {
            acc[2] += (double)bi * (double)bi;
            acc[3] += (double)c;
        }
        if (!(c == c) || c - c != 0.f) acc[4] += 1.0;
        acc[5] += res;
        if (RHS_STATS) acc[6] += (double)bi;
    }
}

// The red x-ring cell of tile row (2 + lane % 8, 2 + lane / 8): x = X0 - 1 or
// X0 + 64, whichever is red, read by the black cell at the tile's x-edge.
template <bool FAST, class XS, class BS>
__host__ __device__ __forceinline__ void rb_xring(XS& xs, const BS& bs, const RbArgs& a, int n, const RbTile& t,
                                                  int lane) {
    const int ly = 2 + (lane & 7), lz = 2 + (lane >> 3);
    const int gy = t.Y0 - 2 + ly, gz = t.Z0 - 2 + lz;
    const int s = (gy + gz) & 1, jr = s ? 1 : kRbHalf - 2;
    const int gx = t.X0 - 4 + 2 * jr + s;  // X0 - 1 or X0 + 64
    if (gx >= 0 && gx < n && gy < n && gz < n) {
        const int w = (ly + kRbHy * lz) * kRbRow + s * kRbHalf + jr;
        rb_cell<FAST>(xs, bs, w, s ? -kRbHalf : kRbHalf - 1, (float)rb_degree(gx, gy, gz, n), a);
    }
}

// Phase 2: the red cells (x + y + z even) the tile's black cells read: those
// of the tile, of its y- and z-ring rows (x in the tile) and of its x-ring
// (tile rows). Ring reds are also updated by the neighbouring block from the
// same old values, identically. Each wave: two tile row pairs ((ly, lz) with
// (ly + 4, lz)) and one ring row pair (y-ring rows pair with lz + 2, z-ring
// rows with ly + 2); the last wave also takes the 64 x-ring cells.
// The b sources of the three pass groups (tile rows, ring rows, x-ring
// cell): the staged LDS array for all three, or the lane's registers.
template <bool FAST, bool INNER, class XS, class BT, class BR, class BX>
__host__ __device__ __forceinline__ void rb_red(XS& xs, const BT& btile, const BR& bring, const BX& bx,
                                                const RbArgs& a, int n, const RbTile& t, int tid) {
    const int wave = rb_uniform(tid >> 6), lane = tid & 63;
    {
        // rows (2 + w % 4, 2 + w / 4) and 4 planes up: the same parity
        const int ly[2] = {2 + wave % 4, 2 + wave % 4}, lz[2] = {2 + wave / 4, 6 + wave / 4};
        rb_passes_at<0, FAST, INNER, 2>(xs, btile, a, n, t, ly, lz, 4, 0, lane);
    }
    const int i = wave & 7, first = 2 + (i & 1) + 4 * ((i >> 1) & 1);  // 2, 3, 6, 7
    if (wave < 8) {
        const int ly[1] = {i < 4 ? 1 : 10}, lz[1] = {first};
        rb_passes_at<0, FAST, INNER, 1>(xs, bring, a, n, t, ly, lz, 0, 2, lane);
    } else {
        const int ly[1] = {first}, lz[1] = {i < 4 ? 1 : 10};
        rb_passes_at<0, FAST, INNER, 1>(xs, bring, a, n, t, ly, lz, 2, 0, lane);
    }
    if (wave == kRbWaves - 1) rb_xring<FAST>(xs, bx, a, n, t, lane);
}
template <bool FAST, bool INNER, class XS, class BS>
__host__ __device__ __forceinline__ void rb_red(XS& xs, const BS& bs, const RbArgs& a, int n, const RbTile& t,
                                                int tid) {
    rb_red<FAST, INNER>(xs, bs, bs, bs, a, n, t, tid);
}

// Phase 3: black cells (x + y + z odd) of the tile, two row pairs per wave.
template <bool FAST, bool INNER, class XS, class BS>
__host__ __device__ __forceinline__ void rb_black(XS& xs, const BS& bs, const RbArgs& a, int n, const RbTile& t,
                                                  int tid) {
    const int wave = rb_uniform(tid >> 6), lane = tid & 63;
    const int ly[2] = {2 + wave % 4, 2 + wave % 4}, lz[2] = {2 + wave / 4, 6 + wave / 4};
    rb_passes_at<1, FAST, INNER, 2>(xs, bs, a, n, t, ly, lz, 4, 0, lane);
}

// Phase 4: write the tile, one float4 per lane; returns the lane's sum of
// the values it wrote (FP64, in x order).
template <bool VEC, class XS, class XOUT>
__host__ __device__ __forceinline__ double rb_write(const XS& xs, XOUT& xout, int n, const RbTile& t, int tid) {
    static_assert(kRbTy * kRbTz * (kRbTx / 4) == kRbThreads, "one float4 of the tile per lane");
    const int row = tid / (kRbTx / 4), q = tid % (kRbTx / 4);
    const int ly = row % kRbTy, lz = row / kRbTy;
    const int gx = t.X0 + 4 * q, gy = t.Y0 + ly, gz = t.Z0 + lz;
    if (gy >= n || gz >= n || gx >= n) return 0.0;
    const int wd = (ly + 2 + kRbHy * (lz + 2)) * kRbRow + 2 * q + 2;
    const float4 v = make_float4(xs[wd], xs[wd + kRbHalf], xs[wd + 1], xs[wd + kRbHalf + 1]);
    const int i = idC(n, gx, gy, gz);
    double sum = 0.0;
    if (VEC) {
        rb_store4(xout, i, v);
        sum = (((sum + (double)v.x) + (double)v.y) + (double)v.z) + (double)v.w;
    } else {
        xout[i] = v.x;
        sum += (double)v.x;
        if (gx + 1 < n) { xout[i + 1] = v.y; sum += (double)v.y; }
        if (gx + 2 < n) { xout[i + 2] = v.z; sum += (double)v.z; }
        if (gx + 3 < n) { xout[i + 3] = v.w; sum += (double)v.w; }
    }
    return sum;
}

// ---- the register path of the sweep (k_sor_rb_reg) ----
//
// The same tile, phases and LDS layout for the iterate, but the staging
// loads are buffer loads at 32-bit byte offsets that each lane advances by a
// fixed stride, and the right-hand side never enters LDS: each lane loads, at
// the start of the launch, b at the (at most six) cells its update passes
// will update, into registers. A load whose cell lies outside the domain
// takes the offset kRbOOB, past the field's records, and the buffer unit
// returns 0.0 for it (the staged +0.0 of the LDS path). Requires n % 4 == 0
// (whole float4 rows) and n^3 * 4 < 2^31 (rb_reg_fits).
constexpr unsigned kRbOOB = 0x80000000u;
__host__ __device__ __forceinline__ bool rb_reg_fits(int n) {
    return n % 4 == 0 && (long long)n * n * n * 4 < (1ll << 31);
}

// Host accessors: element off / 4 of the field, 0.0 at kRbOOB.
template <class A> __host__ __device__ __forceinline__ float rb_bload1(const A& a, unsigned off) {
    return off == kRbOOB ? 0.f : a[(long)(off >> 2)];
}
template <class A> __host__ __device__ __forceinline__ float4 rb_bload4(const A& a, unsigned off) {
    if (off == kRbOOB) return make_float4(0.f, 0.f, 0.f, 0.f);
    const long i = (long)(off >> 2);
    return make_float4(a[i], a[i + 1], a[i + 2], a[i + 3]);
}
#if defined(__HIP__)
// A device field as a raw buffer of n^3 floats (num_records in bytes;
// 0x00020000 is the gfx9 family's 32-bit data format word).
struct RbField {
    __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ RbField rb_field(const float* p, int n) {
    return RbField{__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, n * n * n * 4, 0x00020000)};
}
__device__ __forceinline__ float rb_bload1(const RbField& f, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)off, 0, 0));
}
__device__ __forceinline__ float4 rb_bload4(const RbField& f, unsigned off) {
    // (an int vector: with an unsigned one this compiler emits a one-dword load)
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f u = __builtin_bit_cast(v4f, (v4i)__builtin_amdgcn_raw_buffer_load_b128(f.r, (int)off, 0, 0));
    return make_float4(u.x, u.y, u.z, u.w);
}
#endif

// Byte offset of cell (gx, gy, gz), or kRbOOB outside the domain.
__host__ __device__ __forceinline__ unsigned rb_cell_off(int n, int gx, int gy, int gz) {
    const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n;
    return in ? (unsigned)(((gz * n + gy) * n + gx) * 4) : kRbOOB;
}

// Staging: lanes 0 .. kRbSLanes - 1 each own float4 q of staged row ly in
// planes lz0, lz0 + kRbSPlanes, ... (kRbSIter planes); the offset advances
// by kRbSPlanes planes per step.
constexpr int kRbSPlanes = kRbThreads / (kRbHy * kRbQ);  // 4 planes per step
constexpr int kRbSLanes = kRbSPlanes * kRbHy * kRbQ;     // 864 staging lanes
constexpr int kRbSIter = kRbHz / kRbSPlanes;             // 3 steps
static_assert(kRbHz % kRbSPlanes == 0, "the staging steps cover the staged planes");

template <class XIN>
__host__ __device__ __forceinline__ void rb_fetch_x_reg(const XIN& xin, int n, const RbTile& t, int tid,
                                                        float4 (&v)[kRbSIter]) {
    const int q = tid % kRbQ, r = tid / kRbQ, ly = r % kRbHy, lz0 = r / kRbHy;
    const int gx = t.X0 - 4 + 4 * q, gy = t.Y0 - 2 + ly, gz0 = t.Z0 - 2 + lz0;
    const bool row = tid < kRbSLanes && gx >= 0 && gx < n && gy >= 0 && gy < n;
    const int step = kRbSPlanes * n * n;
    int idx = (gz0 * n + gy) * n + gx;
#pragma unroll
    for (int k = 0; k < kRbSIter; ++k, idx += step) {
        const int gz = gz0 + kRbSPlanes * k;
        v[k] = rb_bload4(xin, row && gz >= 0 && gz < n ? (unsigned)(idx * 4) : kRbOOB);
    }
}
template <class XS>
__host__ __device__ __forceinline__ void rb_put_x_reg(XS& xs, int tid, const float4 (&v)[kRbSIter]) {
    if (tid >= kRbSLanes) return;
    const int q = tid % kRbQ, r = tid / kRbQ;  // r = ly + kRbHy * lz0
#pragma unroll
    for (int k = 0; k < kRbSIter; ++k) rb_put(xs, rb_x_word(r + kRbSPlanes * kRbHy * k, q), v[k]);
}

// The lane's b values: red tile passes 0 and 1, red ring pass, x-ring cell
// (last wave), black passes 0 and 1 -- the cells rb_red / rb_black give it.
struct RbRegs {
    float b[6];
};
__host__ __device__ __forceinline__ unsigned rb_pass_off(int n, const RbTile& t, int color, int ly, int lz, int ly0,
                                                         int lz0, int half, int dly, int dlz, int j) {
    const int S = ((t.Y0 + ly0 + t.Z0 + lz0) & 1) ^ color;
    return rb_cell_off(n, t.X0 - 4 + 2 * j + S, t.Y0 - 2 + ly + half * dly, t.Z0 - 2 + lz + half * dlz);
}
template <class B>
__host__ __device__ __forceinline__ void rb_fetch_b_reg(RbRegs& r, const B& b, int n, const RbTile& t, int tid) {
    const int wave = rb_uniform(tid >> 6), lane = tid & 63, half = lane >> 5, j = (lane & 31) + 2;
    const int ly = 2 + wave % 4, lz0 = 2 + wave / 4, lz1 = 6 + wave / 4;
    r.b[0] = rb_bload1(b, rb_pass_off(n, t, 0, ly, lz0, ly, lz0, half, 4, 0, j));
    r.b[1] = rb_bload1(b, rb_pass_off(n, t, 0, ly, lz1, ly, lz0, half, 4, 0, j));
    r.b[4] = rb_bload1(b, rb_pass_off(n, t, 1, ly, lz0, ly, lz0, half, 4, 0, j));
    r.b[5] = rb_bload1(b, rb_pass_off(n, t, 1, ly, lz1, ly, lz0, half, 4, 0, j));
    const int i = wave & 7, first = 2 + (i & 1) + 4 * ((i >> 1) & 1), edge = i < 4 ? 1 : 10;
    if (wave < 8) r.b[2] = rb_bload1(b, rb_pass_off(n, t, 0, edge, first, edge, first, half, 0, 2, j));
    else r.b[2] = rb_bload1(b, rb_pass_off(n, t, 0, first, edge, first, edge, half, 2, 0, j));
    r.b[3] = 0.f;
    if (wave == kRbWaves - 1) {  // rb_xring's cell
        const int gy = t.Y0 + (lane & 7), gz = t.Z0 + (lane >> 3);
        const int s = (gy + gz) & 1, jr = s ? 1 : kRbHalf - 2;
        r.b[3] = rb_bload1(b, rb_cell_off(n, t.X0 - 4 + 2 * jr + s, gy, gz));
    }
}

template <bool FAST, bool INNER, class XS>
__host__ __device__ __forceinline__ void rb_red_reg(XS& xs, const RbRegs& r, float meanf, const RbArgs& a, int n,
                                                    const RbTile& t, int tid) {
    const RbBReg<2> tile{{r.b[0], r.b[1]}, meanf};
    const RbBReg<1> ring{{r.b[2]}, meanf}, xr{{r.b[3]}, meanf};
    rb_red<FAST, INNER>(xs, tile, ring, xr, a, n, t, tid);
}
template <bool FAST, bool INNER, class XS>
__host__ __device__ __forceinline__ void rb_black_reg(XS& xs, const RbRegs& r, float meanf, const RbArgs& a, int n,
                                                      const RbTile& t, int tid) {
    const RbBReg<2> tile{{r.b[4], r.b[5]}, meanf};
    rb_black<FAST, INNER>(xs, tile, a, n, t, tid);
}

// Block reductions (deterministic: fixed tree, fixed order).
__device__ __forceinline__ int block_linear_thread() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ int block_threads() { return blockDim.x * blockDim.y * blockDim.z; }
__device__ __forceinline__ int block_linear_id() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

enum class Red { Sum, Max, Min };

template <Red OP, class T> __device__ __forceinline__ T red_op(T a, T b) {
    if constexpr (OP == Red::Sum) return a + b;
    else if constexpr (OP == Red::Max) return (a < b) ? b : a;
    else return (b < a) ? b : a;
}

// DPP row shift of a double: lane i of each row of 16 lanes reads lane i - N
// of its row; lanes i < N read `edge`.
template <int N> __device__ __forceinline__ double row_shr(double v, double edge) {
#ifdef __HIP_DEVICE_COMPILE__
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long e = __builtin_bit_cast(unsigned long long, edge);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)u, 0x110 + N, 0xf, 0xf, false);
    const int hi =
        __builtin_amdgcn_update_dpp((int)(unsigned)(e >> 32), (int)(unsigned)(u >> 32), 0x110 + N, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
#else
    return v;
#endif
}

__device__ __forceinline__ double read_lane(double v, int lane) {
#ifdef __HIP_DEVICE_COMPILE__
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
#else
    return v;
#endif
}

// Reduction over the 64 lanes of a full wave with a fixed tree (so the result
// is bit-reproducible): a shift-and-combine scan inside each row of 16 (DPP),
// then the four row results, in row order. The result is wave-uniform. Sums
// take 0 at the row edges, max/min the lane's own value.
template <Red OP> __device__ __forceinline__ double wave_reduce(double v) {
    v = red_op<OP>(v, row_shr<1>(v, OP == Red::Sum ? 0.0 : v));
    v = red_op<OP>(v, row_shr<2>(v, OP == Red::Sum ? 0.0 : v));
    v = red_op<OP>(v, row_shr<4>(v, OP == Red::Sum ? 0.0 : v));
    v = red_op<OP>(v, row_shr<8>(v, OP == Red::Sum ? 0.0 : v));
    double r = read_lane(v, 15);
    r = red_op<OP>(r, read_lane(v, 31));
    r = red_op<OP>(r, read_lane(v, 47));
    return red_op<OP>(r, read_lane(v, 63));
}
__device__ __forceinline__ double wave_reduce(double v, Red op) {
    return op == Red::Sum ? wave_reduce<Red::Sum>(v) : op == Red::Max ? wave_reduce<Red::Max>(v)
                                                                       : wave_reduce<Red::Min>(v);
}

// Reduce kSlots doubles across a block of full waves (at most 16), called
// once per kernel by every thread; thread 0 receives the results. Each wave
// reduces its lanes, lane 0 of each wave stores the wave's values in lds
// (kSlots * 16 entries), and wave 0 reduces those.
template <int kSlots>
__device__ __forceinline__ void block_reduce(double (&v)[kSlots], const Red (&op)[kSlots], double* lds) {
    const int tid = block_linear_thread();
    const int lane = tid % kWave, wid = tid / kWave, waves = block_threads() / kWave;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        v[s] = wave_reduce(v[s], op[s]);
        if (lane == 0) lds[s * 16 + wid] = v[s];
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double first = lds[s * 16];
            const double x = lane < waves ? lds[s * 16 + lane] : (op[s] == Red::Sum ? 0.0 : first);
            v[s] = wave_reduce(x, op[s]);
        }
    }
}

}  // namespace vgx
