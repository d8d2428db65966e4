// vgx_device.hpp - device-side helpers shared by the MAC 3D kernels.
//
// Every float expression here keeps the operation order of vapor's CPU
// solver, and the library is compiled with -ffp-contract=off, so the GPU
// rounds exactly as the CPU reference does (IEEE division and square root are
// the HIP defaults on gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgx {

constexpr int kBlock = 256;  // elementwise / gather kernels: 4 waves of 64
constexpr int kWave = 64;

// std::clamp / std::min / std::max with the reference's comparison order.
__host__ __device__ __forceinline__ float clampf(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
__host__ __device__ __forceinline__ float minf(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ float maxf(float a, float b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ int mini(int a, int b) { return (b < a) ? b : a; }

// Float -> int truncation as the reference's int(v), made total: NaN maps to 0
// and the magnitude is bounded, so no index is ever derived from a value the
// conversion leaves undefined. Identical to (int)v for finite in-range v.
__host__ __device__ __forceinline__ int trunc_index(float v) {
    return (v == v) ? (int)fmaxf(fminf(v, 1.0e9f), -1.0e9f) : 0;
}

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

// A field read through an optional reflection: value = 2 S[i] - T[i]
// (MacGridFluidSolver3D.cpp:211-218 materialises the same expression).
template <bool HAT, class P = const float*> struct Src {
    P s;
    P t;
    __host__ __device__ __forceinline__ float operator[](int i) const {
        if constexpr (HAT) return 2.f * s[i] - t[i];
        else return s[i];
    }
};

// Clamped trilinear sample, MacGridSampler3D.cpp:5-71. On the staggered axis
// g = p/h in [0, n] with the lower index capped at n-1; elsewhere
// g = p/h - 0.5 in [0, n-1] with the upper index capped at n-1.
template <int AXIS, class F>
__host__ __device__ __forceinline__ float sample(const F& f, int n, float h, float px, float py, float pz) {
    const float fn = (float)n, fn1 = (float)(n - 1);
    float gx, gy, gz;
    int x0, y0, z0, x1, y1, z1;
    if constexpr (AXIS == 0) {
        gx = clampf(px / h, 0.f, fn);
        x0 = mini(trunc_index(gx), n - 1);
        x1 = x0 + 1;
    } else {
        gx = clampf(px / h - 0.5f, 0.f, fn1);
        x0 = trunc_index(gx);
        x1 = mini(x0 + 1, n - 1);
    }
    if constexpr (AXIS == 1) {
        gy = clampf(py / h, 0.f, fn);
        y0 = mini(trunc_index(gy), n - 1);
        y1 = y0 + 1;
    } else {
        gy = clampf(py / h - 0.5f, 0.f, fn1);
        y0 = trunc_index(gy);
        y1 = mini(y0 + 1, n - 1);
    }
    if constexpr (AXIS == 2) {
        gz = clampf(pz / h, 0.f, fn);
        z0 = mini(trunc_index(gz), n - 1);
        z1 = z0 + 1;
    } else {
        gz = clampf(pz / h - 0.5f, 0.f, fn1);
        z0 = trunc_index(gz);
        z1 = mini(z0 + 1, n - 1);
    }
    const float tx = gx - (float)x0, ty = gy - (float)y0, tz = gz - (float)z0;
    const float f000 = f[idA<AXIS>(n, x0, y0, z0)], f100 = f[idA<AXIS>(n, x1, y0, z0)];
    const float f010 = f[idA<AXIS>(n, x0, y1, z0)], f110 = f[idA<AXIS>(n, x1, y1, z0)];
    const float f001 = f[idA<AXIS>(n, x0, y0, z1)], f101 = f[idA<AXIS>(n, x1, y0, z1)];
    const float f011 = f[idA<AXIS>(n, x0, y1, z1)], f111 = f[idA<AXIS>(n, x1, y1, z1)];
    return f000 * (1 - tx) * (1 - ty) * (1 - tz) + f100 * tx * (1 - ty) * (1 - tz) +
           f010 * (1 - tx) * ty * (1 - tz) + f110 * tx * ty * (1 - tz) +
           f001 * (1 - tx) * (1 - ty) * tz + f101 * tx * (1 - ty) * tz +
           f011 * (1 - tx) * ty * tz + f111 * tx * ty * tz;
}

// Sample position of entry (x, y, z): cell centre or face centre
// (CpuAdvectionKernels.cpp:25, :91, :154, :217).
template <int AXIS>
__host__ __device__ __forceinline__ void sample_point(float dx, int x, int y, int z, float& px, float& py, float& pz) {
    px = AXIS == 0 ? (float)x * dx : ((float)x + .5f) * dx;
    py = AXIS == 1 ? (float)y * dx : ((float)y + .5f) * dx;
    pz = AXIS == 2 ? (float)z * dx : ((float)z + .5f) * dx;
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
template <int AXIS, class U, class S>
__host__ __device__ __forceinline__ float sl_value(const U& u, const U& v, const U& w, const S& src, int n, float h,
                                                   float dt, int i) {
    int x, y, z;
    decompose<AXIS>(i, n, x, y, z);
    float px, py, pz;
    sample_point<AXIS>(h, x, y, z, px, py, pz);
    const float vx = sample<0>(u, n, h, px, py, pz);
    const float vy = sample<1>(v, n, h, px, py, pz);
    const float vz = sample<2>(w, n, h, px, py, pz);
    return sample<AXIS>(src, n, h, px - vx * dt, py - vy * dt, pz - vz * dt);
}

// MacCormack correction with its fallback and limiter for entry i:
// CpuAdvectionKernels.cpp:45-78 (scalar), :111-142 (u), :174-205 (v),
// :237-268 (w). The backward trace of the forward field is evaluated only
// where the correction is used.
template <int AXIS, class U, class S, class W>
__host__ __device__ __forceinline__ float mc_value(const U& u, const U& v, const U& w, const S& src, const W& fwd,
                                                   int n, float dx, float box, float dt, float cflMc, int i) {
    int x, y, z;
    decompose<AXIS>(i, n, x, y, z);
    float px, py, pz;
    sample_point<AXIS>(dx, x, y, z, px, py, pz);
    const float vx = sample<0>(u, n, dx, px, py, pz);
    const float vy = sample<1>(v, n, dx, px, py, pz);
    const float vz = sample<2>(w, n, dx, px, py, pz);
    const float fx = px - vx * dt, fy = py - vy * dt, fz = pz - vz * dt;
    const float bx = px + vx * dt, by = py + vy * dt, bz = pz + vz * dt;
    const float cfl = sqrtf(vx * vx + vy * vy + vz * vz) * dt / dx;
    if (fx < 0 || fx > box || fy < 0 || fy > box || fz < 0 || fz > box || bx < 0 || bx > box || by < 0 ||
        by > box || bz < 0 || bz > box || cfl > cflMc)
        return fwd[i];
    const float back = sample<AXIS>(fwd, n, dx, bx, by, bz);
    const float corrected = fwd[i] + 0.5f * (src[i] - back);
    const int h0 = AXIS == 0 ? n : n - 1, h1 = AXIS == 1 ? n : n - 1, h2 = AXIS == 2 ? n : n - 1;
    const int g0 = clampi(AXIS == 0 ? trunc_index(fx / dx) : trunc_index(fx / dx - 0.5f), 0, h0);
    const int g1 = clampi(AXIS == 1 ? trunc_index(fy / dx) : trunc_index(fy / dx - 0.5f), 0, h1);
    const int g2 = clampi(AXIS == 2 ? trunc_index(fz / dx) : trunc_index(fz / dx - 0.5f), 0, h2);
    float fMin = src[idA<AXIS>(n, g0, g1, g2)], fMax = fMin;
#pragma unroll
    for (int dz = 0; dz <= 1; ++dz)
#pragma unroll
        for (int dy = 0; dy <= 1; ++dy)
#pragma unroll
            for (int d0 = 0; d0 <= 1; ++d0) {
                const float val =
                    src[idA<AXIS>(n, clampi(g0 + d0, 0, h0), clampi(g1 + dy, 0, h1), clampi(g2 + dz, 0, h2))];
                fMin = minf(fMin, val);
                fMax = maxf(fMax, val);
            }
    return clampf(corrected, fMin, fMax);
}

// ---- fused red-black sweep: tile geometry and phases ----
// A block owns a kRbTx x kRbTy x kRbTz tile and stages the input iterate over
// the tile plus a 2-cell halo (kRbHx x kRbHy x kRbHz) and the right-hand side
// over the tile plus a 1-cell ring (kRbBx x kRbBy x kRbBz) in LDS, so the update
// phases read LDS only. The phases are __host__ __device__ over accessor
// types so that a CPU build can run them with bounds-checked accessors.
constexpr int kRbTx = 64, kRbTy = 8, kRbTz = 8, kRbThreads = 1024;
constexpr int kRbHx = kRbTx + 4, kRbHy = kRbTy + 4, kRbHz = kRbTz + 4;
constexpr int kRbBx = kRbTx + 2, kRbBy = kRbTy + 2, kRbBz = kRbTz + 2;
constexpr int kRbStage = kRbHx * kRbHy * kRbHz;
constexpr int kRbStageB = kRbBx * kRbBy * kRbBz;

struct RbTile {
    int X0, Y0, Z0;
    bool valid;
};

__host__ __device__ __forceinline__ int rb_tiles(int n) {
    return ((n + kRbTx - 1) / kRbTx) * ((n + kRbTy - 1) / kRbTy) * ((n + kRbTz - 1) / kRbTz);
}
// Grid size: a multiple of 8 (one share per XCD).
__host__ __device__ __forceinline__ int rb_blocks(int n) { return 8 * ((rb_tiles(n) + 7) / 8); }

// Block b takes tile (b % 8) * ceil(T / 8) + b / 8; tiles are numbered
// x-fastest, then y, then z, so the blocks the hardware deals to one XCD
// (b % 8 equal) cover one contiguous z-slab.
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

// One red-black update of the cell at staged index li: the reference's
// x += omega (b - sum(offdiag x) - d x) / max(d, floor) with the neighbour
// terms in its order (SOR.cpp:36-49).
template <class XS>
__host__ __device__ __forceinline__ float rb_update(const XS& xs, int li, int cx, int cy, int cz, int n, float bi,
                                                    float omega, float floor) {
    const float c = xs[li];
    float off = 0.f;
    if (cx > 0) off += -1.f * xs[li - 1];
    if (cx + 1 < n) off += -1.f * xs[li + 1];
    if (cy > 0) off += -1.f * xs[li - kRbHx];
    if (cy + 1 < n) off += -1.f * xs[li + kRbHx];
    if (cz > 0) off += -1.f * xs[li - kRbHx * kRbHy];
    if (cz + 1 < n) off += -1.f * xs[li + kRbHx * kRbHy];
    const float d = (float)((cx > 0) + (cx + 1 < n) + (cy > 0) + (cy + 1 < n) + (cz > 0) + (cz + 1 < n));
    if (d == 0.f) return c;
    const float diagonal = maxf(d, floor);
    return c + omega * (bi - off - d * c) / diagonal;
}

// Phase 1: stage xin over the tile and its 2-cell halo and b' = b - meanf over
// the tile and its 1-cell ring (zero outside the domain; those entries are
// never read). The loops have fixed trip counts so that every lane issues all
// its loads before the first one is consumed.
template <class XS, class BS, class XIN, class B>
__host__ __device__ __forceinline__ void rb_stage(XS& xs, BS& bs, const XIN& xin, const B& b, float meanf, int n,
                                                  const RbTile& t, int tid) {
    constexpr int kIter = (kRbStage + kRbThreads - 1) / kRbThreads;
#pragma unroll
    for (int k = 0; k < kIter; ++k) {
        const int l = tid + k * kRbThreads;
        if (l < kRbStage) {
            const int lx = l % kRbHx, r = l / kRbHx, ly = r % kRbHy, lz = r / kRbHy;
            const int gx = t.X0 - 2 + lx, gy = t.Y0 - 2 + ly, gz = t.Z0 - 2 + lz;
            xs[l] = (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) ? xin[idC(n, gx, gy, gz)] : 0.f;
        }
    }
    constexpr int kIterB = (kRbStageB + kRbThreads - 1) / kRbThreads;
#pragma unroll
    for (int k = 0; k < kIterB; ++k) {
        const int l = tid + k * kRbThreads;
        if (l < kRbStageB) {
            const int lx = l % kRbBx, r = l / kRbBx, ly = r % kRbBy, lz = r / kRbBy;
            const int gx = t.X0 - 1 + lx, gy = t.Y0 - 1 + ly, gz = t.Z0 - 1 + lz;
            bs[l] = (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) ? b[idC(n, gx, gy, gz)] - meanf
                                                                                  : 0.f;
        }
    }
}

// Staged-b index of cell (cx, cy, cz).
__host__ __device__ __forceinline__ int rb_bindex(const RbTile& t, int cx, int cy, int cz) {
    return (cx - t.X0 + 1) + kRbBx * ((cy - t.Y0 + 1) + kRbBy * (cz - t.Z0 + 1));
}

// Phase 1b (optional): FP64 residual of the staged iterate on the tile, the
// seven partial fields of the residual check.
template <class XS, class BS>
__host__ __device__ __forceinline__ void rb_residual(const XS& xs, const BS& bs, int n, const RbTile& t, int tid,
                                                     double* acc) {
    for (int l = tid; l < kRbTx * kRbTy * kRbTz; l += kRbThreads) {
        const int lx = l % kRbTx, r = l / kRbTx, ly = r % kRbTy, lz = r / kRbTy;
        const int cx = t.X0 + lx, cy = t.Y0 + ly, cz = t.Z0 + lz;
        if (cx >= n || cy >= n || cz >= n) continue;
        const int li = (lx + 2) + kRbHx * ((ly + 2) + kRbHy * (lz + 2));
        const float c = xs[li];
        const double d = (double)((cx > 0) + (cx + 1 < n) + (cy > 0) + (cy + 1 < n) + (cz > 0) + (cz + 1 < n));
        double ax = d * (double)c;
        if (cx > 0) ax += -1.0 * (double)xs[li - 1];
        if (cx + 1 < n) ax += -1.0 * (double)xs[li + 1];
        if (cy > 0) ax += -1.0 * (double)xs[li - kRbHx];
        if (cy + 1 < n) ax += -1.0 * (double)xs[li + kRbHx];
        if (cz > 0) ax += -1.0 * (double)xs[li - kRbHx * kRbHy];
        if (cz + 1 < n) ax += -1.0 * (double)xs[li + kRbHx * kRbHy];
        const float bi = bs[rb_bindex(t, cx, cy, cz)];
        const double res = (double)bi - ax;
        acc[0] += res * res;
        const double ar = res < 0 ? -res : res;
        acc[1] = acc[1] < ar ? ar : acc[1];
        acc[2] += (double)bi * (double)bi;
        acc[3] += (double)c;
        if (!(c == c) || c - c != 0.f) acc[4] += 1.0;
        acc[5] += res;
        acc[6] += (double)bi;
    }
}

// Phase 2: red cells (parity 0) of the tile and its 1-cell ring, one pair of
// x-neighbours per lane with the pair's red member updated. Ring reds are also
// updated by the neighbouring block from the same old values, identically.
template <class XS, class BS>
__host__ __device__ __forceinline__ void rb_red(XS& xs, const BS& bs, float omega, float floor, int n,
                                                const RbTile& t, int tid) {
    constexpr int kRx = (kRbTx + 2) / 2, kRy = kRbTy + 2, kRz = kRbTz + 2;
    for (int l = tid; l < kRx * kRy * kRz; l += kRbThreads) {
        const int px = l % kRx, r = l / kRx, ly = r % kRy, lz = r / kRy;
        const int cy = t.Y0 - 1 + ly, cz = t.Z0 - 1 + lz;
        const int cx0 = t.X0 - 1 + 2 * px;
        const int cx = cx0 + ((cx0 + cy + cz) & 1);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= n || cy >= n || cz >= n) continue;
        const int li = (cx - t.X0 + 2) + kRbHx * ((ly + 1) + kRbHy * (lz + 1));
        xs[li] = rb_update(xs, li, cx, cy, cz, n, bs[rb_bindex(t, cx, cy, cz)], omega, floor);
    }
}

// Phase 3: black cells (parity 1) of the tile.
template <class XS, class BS>
__host__ __device__ __forceinline__ void rb_black(XS& xs, const BS& bs, float omega, float floor, int n,
                                                  const RbTile& t, int tid) {
    constexpr int kBx = kRbTx / 2;
    for (int l = tid; l < kBx * kRbTy * kRbTz; l += kRbThreads) {
        const int px = l % kBx, r = l / kBx, ly = r % kRbTy, lz = r / kRbTy;
        const int cy = t.Y0 + ly, cz = t.Z0 + lz;
        const int cx0 = t.X0 + 2 * px;
        const int cx = cx0 + (1 - ((cx0 + cy + cz) & 1));
        if (cx >= n || cy >= n || cz >= n) continue;
        const int li = (cx - t.X0 + 2) + kRbHx * ((ly + 2) + kRbHy * (lz + 2));
        xs[li] = rb_update(xs, li, cx, cy, cz, n, bs[rb_bindex(t, cx, cy, cz)], omega, floor);
    }
}

// Phase 4: write the tile; returns the sum of the written values.
template <class XS, class XOUT>
__host__ __device__ __forceinline__ double rb_write(const XS& xs, XOUT& xout, int n, const RbTile& t, int tid) {
    double sum = 0.0;
    for (int l = tid; l < kRbTx * kRbTy * kRbTz; l += kRbThreads) {
        const int lx = l % kRbTx, r = l / kRbTx, ly = r % kRbTy, lz = r / kRbTy;
        const int cx = t.X0 + lx, cy = t.Y0 + ly, cz = t.Z0 + lz;
        if (cx >= n || cy >= n || cz >= n) continue;
        const float v = xs[(lx + 2) + kRbHx * ((ly + 2) + kRbHy * (lz + 2))];
        xout[idC(n, cx, cy, cz)] = v;
        sum += (double)v;
    }
    return sum;
}

// Wave / block reductions (deterministic: fixed shuffle tree, fixed order).
template <class T> __device__ __forceinline__ T wave_sum(T v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
template <class T> __device__ __forceinline__ T wave_max(T v) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const T other = __shfl_xor(v, o);
        v = (v < other) ? other : v;
    }
    return v;
}
template <class T> __device__ __forceinline__ T wave_min(T v) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const T other = __shfl_xor(v, o);
        v = (other < v) ? other : v;
    }
    return v;
}

__device__ __forceinline__ int block_linear_thread() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ int block_threads() { return blockDim.x * blockDim.y * blockDim.z; }
__device__ __forceinline__ int block_linear_id() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// Reduce up to kSlots values across a block; every thread receives the result.
// lds holds kSlots * (threads / 64) entries.
enum class Red { Sum, Max, Min };
template <int kSlots, class T>
__device__ __forceinline__ void block_reduce(T (&v)[kSlots], const Red (&op)[kSlots], T* lds) {
    const int tid = block_linear_thread();
    const int lane = tid % kWave, wid = tid / kWave, waves = (block_threads() + kWave - 1) / kWave;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        v[s] = op[s] == Red::Sum ? wave_sum(v[s]) : op[s] == Red::Max ? wave_max(v[s]) : wave_min(v[s]);
        if (lane == 0) lds[s * 16 + wid] = v[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        T acc = lds[s * 16];
        for (int k = 1; k < waves; ++k) {
            const T o = lds[s * 16 + k];
            acc = op[s] == Red::Sum ? acc + o : op[s] == Red::Max ? ((acc < o) ? o : acc) : ((o < acc) ? o : acc);
        }
        v[s] = acc;
    }
    __syncthreads();
}

}  // namespace vgx
