// vgx_kernels.hip - gfx950 kernels of the MAC 3D smoke step.
//
// All fields are FP32 SoA in the reference layouts (MacGridDomain3D.hpp:22-31).
// Every kernel is HBM/MALL-bandwidth bound: the advection kernels gather
// trilinear stencils that neighbouring lanes share through L2, the stencil
// kernels (divergence, red-black SOR, residual, gradient) read x-contiguous
// rows with one lane per cell, 64 lanes per row segment.
//
// Reductions are deterministic: per-block partials in a fixed shuffle tree,
// then one block folds them in a fixed order, so repeated runs agree bitwise.
#include "vgx_device.hpp"
#include "vgx_internal.hpp"

#include "../../include/vapor_mac3d.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>

namespace vgx {

const char* const kKindNames[K_COUNT] = {
    "emit",         "forces",    "advect_sl_faces", "advect_mc_faces", "advect_sl_scalar",
    "advect_mc_scalar", "divergence_rhs", "sor_sweep", "residual_fp64", "gradient_project",
    "reduce",       "diagnostics", "fill",            "sor_sweep_check", "learned_start"};

ProbeSource*& current_probe_source() {
    thread_local ProbeSource* source = nullptr;
    return source;
}

namespace {

// Every kernel goes through here: timed launches attach their events to the
// dispatch itself.
template <class... KArgs, class... Args>
void vlaunch(void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t s, Args... args) {
    const Probe p = current_probe_source() ? current_probe_source()->take() : Probe{};
    if (p.start) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, p.start, p.stop, 0, static_cast<KArgs>(args)...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, s, static_cast<KArgs>(args)...);
}

constexpr int kMaxReduceBlocks = 16384;

struct Launch {
    dim3 grid, block;
};

// Cell-stencil launches: 64 x 4 lanes per block over (x, y), each lane
// walking a contiguous run of `chunk` z planes, so that its neighbours along
// z stay in registers from one plane to the next; about kStencilBlocks
// blocks (4 waves each), so that the per-block FP64 reductions are few and
// every CU still holds several blocks.
constexpr int kStencilBlocks = 2048;
struct CellGeom {
    int gx, gy, gz, chunk;
};
__host__ __device__ inline CellGeom cell_geom(int n) {
    const int gx = (n + 63) / 64, gy = (n + 3) / 4;
    int want = kStencilBlocks / (gx * gy);
    want = want < 1 ? 1 : (want > n ? n : want);
    const int chunk = (n + want - 1) / want;
    return {gx, gy, (n + chunk - 1) / chunk, chunk};
}
Launch cell_launch(int n) {
    const CellGeom g = cell_geom(n);
    return {dim3(g.gx, g.gy, g.gz), dim3(64, 4, 1)};
}

int linear_blocks(size_t count) {
    const size_t b = (count + kBlock - 1) / kBlock;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), kMaxReduceBlocks);
}

constexpr double kInf = __builtin_huge_val();

// ------------------------------------------------------------------- fills

__global__ void __launch_bounds__(kBlock) k_fill_f32(float* p, size_t count, float value) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        p[i] = value;
}
__global__ void __launch_bounds__(kBlock) k_fill_f64(double* p, size_t count, double value) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        p[i] = value;
}
__global__ void k_set_u32(uint32_t* p, uint32_t value) { *p = value; }

// out = a + b (element-wise).
__global__ void __launch_bounds__(kBlock) k_add_f32(const float* __restrict__ a, const float* __restrict__ b,
                                                    float* __restrict__ out, size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = a[i] + b[i];
}
// out = a + s b (element-wise).
__global__ void __launch_bounds__(kBlock) k_add_scaled_f32(const float* __restrict__ a, float s,
                                                           const float* __restrict__ b, float* __restrict__ out,
                                                           size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = a[i] + s * b[i];
}
// out = a * s (element-wise).
__global__ void __launch_bounds__(kBlock) k_scale_f32(const float* __restrict__ a, float s, float* __restrict__ out,
                                                      size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = a[i] * s;
}
// Training target of the learned warm start: the gauged solution minus the
// starting pressure, out = ((x - ps->x_meanf) - prev) * s.
__global__ void __launch_bounds__(kBlock) k_correction(const float* __restrict__ x, const float* __restrict__ prev,
                                                       const ProjState* __restrict__ ps, float s,
                                                       float* __restrict__ out, size_t count) {
    const float meanf = ps->x_meanf;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = ((x[i] - meanf) - prev[i]) * s;
}

// field[i] += sigma * N(0, 1) with the counter-based variate of entry i of
// noise stream `stream` (noise_normal): the same value the host computes.
__global__ void __launch_bounds__(kBlock) k_add_noise(float* field, uint32_t count, float sigma, uint64_t seed,
                                                      uint32_t stream) {
    for (uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x; i < count; i += gridDim.x * (uint32_t)kBlock)
        field[i] += sigma * noise_normal(seed, stream, i);
}

// ------------------------------------------------------------------ emitter

// In-sphere test of MacGridFluidSolver3D.cpp:150-153, same operations in the
// field precision T.
template <class T> __device__ __forceinline__ bool emit_inside(const EmitArgsT<T>& a, int gx, int gy, int gz, T& q) {
    const T half = (T)0.5;
    const T px = ((T)gx + half) * a.dx, py = ((T)gy + half) * a.dx, pz = ((T)gz + half) * a.dx;
    const T qx = (px - a.ecx) / a.erad, qy = (py - a.ecy) / a.erad, qz = (pz - a.ecz) / a.erad;
    q = qx * qx + qy * qy + qz * qz;
    return !(q > (T)1);
}

// The reference LCG s -> 1664525 s + 1013904223 (MacGridFluidSolver3D.cpp:33-36)
// advanced k steps at once by binary powering of the affine map.
__device__ __forceinline__ uint32_t lcg_jump(uint32_t s, uint32_t k) {
    uint32_t A = 1664525u, C = 1013904223u, Ar = 1u, Cr = 0u;
    while (k) {
        if (k & 1u) {
            Ar = A * Ar;
            Cr = A * Cr + C;
        }
        C = A * C + C;
        A = A * A;
        k >>= 1;
    }
    return Ar * s + Cr;
}

// Stir noise of every in-sphere cell of the patch. The reference draws one LCG
// value per in-sphere cell in z, y, x order; here each cell's rank among the
// in-sphere cells comes from a block scan, and its value from an LCG jump of
// rank + 1 steps. One block walks the patch in chunks of 1024 cells.
template <class T> __device__ __forceinline__ void emit_noise_body(const EmitArgsT<T>& a, uint32_t* rng, T* noise) {
    __shared__ uint32_t wave_total[16], wave_offset[16], chunk_total, carry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t s0 = *rng;
    const int count = a.ext[0] * a.ext[1] * a.ext[2];
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < count; base += 1024) {
        const int k = base + tid;
        bool in = false;
        T q;
        if (k < count) {
            const int lx = k % a.ext[0], t = k / a.ext[0], ly = t % a.ext[1], lz = t / a.ext[1];
            in = emit_inside(a, a.lo[0] + lx, a.lo[1] + ly, a.lo[2] + lz, q);
        }
        const uint64_t mask = __ballot(in);
        const uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wave_total[wid] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int w = 0; w < 16; ++w) {
                wave_offset[w] = acc;
                acc += wave_total[w];
            }
            chunk_total = acc;
        }
        __syncthreads();
        if (in) {
            const uint32_t s = lcg_jump(s0, carry + wave_offset[wid] + before + 1u);
            const T r = (T)(s >> 8) * (T)(1.0f / 16777216.0f);
            noise[k] = (r - (T)0.5) * a.stir * a.dt_scale;
        }
        __syncthreads();
        if (tid == 0) carry += chunk_total;
        __syncthreads();
    }
    if (tid == 0) *rng = lcg_jump(s0, carry);
}
__global__ void __launch_bounds__(1024) k_emit_noise(EmitArgs a, uint32_t* rng, float* noise) {
    emit_noise_body(a, rng, noise);
}

// Gather form of MacGridFluidSolver3D::emit: one thread per patch point
// updates the cell and the low x/y/z faces there. A face shared by two cells
// receives the lower cell's stir value first, as in the reference's traversal.
template <class T>
__device__ __forceinline__ void emit_apply_body(const EmitArgsT<T>& a, const T* __restrict__ noise, T* rho, T* u, T* v,
                                                T* w, int block) {
    const int E0 = a.ext[0] + 1, E1 = a.ext[1] + 1, E2 = a.ext[2] + 1;
    const int k = block * kBlock + threadIdx.x;
    if (k >= E0 * E1 * E2) return;
    const int lx = k % E0, t = k / E0, ly = t % E1, lz = t / E1;
    const int gx = a.lo[0] + lx, gy = a.lo[1] + ly, gz = a.lo[2] + lz;
    const int n = a.n;
    auto cell = [&](int cx, int cy, int cz, T& j, T& q) -> bool {
        if (cx < 0 || cy < 0 || cz < 0 || cx >= a.ext[0] || cy >= a.ext[1] || cz >= a.ext[2]) return false;
        if (!emit_inside(a, a.lo[0] + cx, a.lo[1] + cy, a.lo[2] + cz, q)) return false;
        j = noise[cx + a.ext[0] * (cy + a.ext[1] * cz)];
        return true;
    };
    T js = 0, qs = 0, jo = 0, qo = 0;
    const bool self = cell(lx, ly, lz, js, qs);
    if (self) {
        const int i = idC(n, gx, gy, gz);
        rho[i] = a.profile == VGX_EMITTER_GAUSSIAN ? maxv(rho[i], minv((T)1, (T)3 * a.strength * expv((T)-3.5 * qs)))
                                                   : (T)1;
    }
    if (ly < a.ext[1] && lz < a.ext[2]) {
        const bool lower = cell(lx - 1, ly, lz, jo, qo);
        if (lower || self) {
            const int i = idX(n, gx, gy, gz);
            T val = u[i];
            if (lower) val += jo;
            if (self) val += js;
            u[i] = val;
        }
    }
    if (lx < a.ext[0] && lz < a.ext[2]) {
        const bool lower = cell(lx, ly - 1, lz, jo, qo);
        if (lower || self) {
            const int i = idY(n, gx, gy, gz);
            T val = v[i];
            if (lower) val += jo;
            if (self) val += js;
            v[i] = val;
        }
    }
    if (a.kick_on && lx < a.ext[0] && ly < a.ext[1]) {
        const bool lower = cell(lx, ly, lz - 1, jo, qo);
        if (lower || self) {
            const int i = idZ(n, gx, gy, gz);
            w[i] = maxv(w[i], a.kick);
        }
    }
}
__global__ void __launch_bounds__(kBlock) k_emit_apply(EmitArgs a, const float* __restrict__ noise,
                                                       float* rho, float* u, float* v, float* w) {
    emit_apply_body(a, noise, rho, u, v, w, blockIdx.x);
}

// ------------------------------------------------------------------- forces

// applyBuoyancy + enforceClosedBoundaries (MacGridFluidSolver3D.cpp:167-181,
// MacGridOperators3D.cpp:16-28). The three face arrays have equal length, so
// one lane serves index i of each.
template <class T>
__device__ __forceinline__ void forces_body(int n, const T* __restrict__ rho, T* u, T* v, T* w, T buoyancy, T dt,
                                            T split, int block, int nblocks) {
    const int count = (n + 1) * n * n;
    for (int i = block * kBlock + threadIdx.x; i < count; i += nblocks * kBlock) {
        {
            const int x = i % (n + 1);
            if (x == 0 || x == n) u[i] = 0;
        }
        {
            const int y = (i / n) % (n + 1);
            if (y == 0 || y == n) v[i] = 0;
        }
        {
            const int x = i % n, t = i / n, y = t % n, z = t / n;
            if (z == 0 || z == n) {
                w[i] = 0;
            } else {
                const T lower = buoyancy * rho[idC(n, x, y, z - 1)] * dt;
                const T upper = buoyancy * rho[idC(n, x, y, z)] * dt;
                T val = w[i];
                val += split * lower;
                val += split * upper;
                w[i] = val;
            }
        }
    }
}
__global__ void __launch_bounds__(kBlock) k_forces(int n, const float* __restrict__ rho, float* u, float* v,
                                                   float* w, float buoyancy, float dt, float split) {
    forces_body(n, rho, u, v, w, buoyancy, dt, split, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- advection

template <int AXIS, bool HAT, bool POW2, class T>
__device__ __forceinline__ void sl_entry(const AdvArgsT<T>& a, const AdvFieldT<T>& f, const Spacing<T, POW2>& sp,
                                         int i) {
    const Src<HAT, const T*> src{f.src, f.tsrc};
    T r = sl_value<AXIS>(a.u, a.v, a.w, src, a.n, sp, a.dt, i);
    if (AXIS < 0 && a.scale_out) r = r * a.out_scale;
    f.out[i] = r;
}

template <int AXIS, bool HAT, bool POW2, class T>
__device__ __forceinline__ void mc_entry(const AdvArgsT<T>& a, const AdvFieldT<T>& f, const Spacing<T, POW2>& sp,
                                         int i) {
    const Src<HAT, const T*> src{f.src, f.tsrc};
    T r = mc_value<AXIS>(a.u, a.v, a.w, src, f.fwd, a.n, sp, a.box, a.dt, a.cfl, i);
    if (AXIS < 0 && a.scale_out) r = r * a.out_scale;
    f.out[i] = r;
}

// A float field read by raw buffer loads: a 32-bit byte offset per gathered
// element instead of 64-bit address arithmetic (the same loads, the same
// values). Used while every field's bytes fit in 2^31 (advect_buffers).
struct BufField {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float operator[](int i) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0));
    }
};
__device__ __forceinline__ BufField buf_field(const float* p, int bytes) {
    return BufField{__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, bytes, 0x00020000)};
}

// One kernel per field (AXIS -1 cell, 0/1/2 faces): every instantiation
// stores through the output pointer of its own field, loaded straight from the
// kernel arguments. (A runtime switch over the field inside the grid-stride
// loop compiled, with this toolchain, to a store through an output-pointer
// register that the cell-field path never assigned.)
// POW2: the spacing is a power of two (Spacing).
template <bool MC, int AXIS, bool HAT, bool POW2 = false, class T>
__device__ __forceinline__ void advect_body(const AdvArgsT<T>& a, int block, int nblocks) {
    const int n = a.n;
    const int count = AXIS < 0 ? n * n * n : (n + 1) * n * n;
    const AdvFieldT<T>& f = a.f[AXIS < 0 ? 3 : AXIS];
    const Spacing<T, POW2> sp{a.h, POW2 ? (T)1 / a.h : (T)0};
    for (int i = block * kBlock + threadIdx.x; i < count; i += nblocks * kBlock) {
        if (MC) mc_entry<AXIS, HAT>(a, f, sp, i);
        else sl_entry<AXIS, HAT>(a, f, sp, i);
    }
}
// advect_body with every gathered field read through buffer loads (float).
template <bool MC, int AXIS, bool HAT, bool POW2>
__device__ __forceinline__ void advect_body_buf(const AdvArgs& a, int block, int nblocks) {
    const int n = a.n;
    const int count = AXIS < 0 ? n * n * n : (n + 1) * n * n;
    const int bytes = (n + 1) * n * n * 4;  // the largest field; a bound only
    const AdvField& f = a.f[AXIS < 0 ? 3 : AXIS];
    const Spacing<float, POW2> sp{a.h, POW2 ? 1.f / a.h : 0.f};
    const BufField u = buf_field(a.u, bytes), v = buf_field(a.v, bytes), w = buf_field(a.w, bytes);
    const Src<HAT, BufField> src{buf_field(f.src, bytes), buf_field(HAT ? f.tsrc : f.src, bytes)};
    for (int i = block * kBlock + threadIdx.x; i < count; i += nblocks * kBlock) {
        float r;
        if (MC) r = mc_value<AXIS>(u, v, w, src, buf_field(f.fwd, bytes), n, sp, a.box, a.dt, a.cfl, i);
        else r = sl_value<AXIS>(u, v, w, src, n, sp, a.dt, i);
        if (AXIS < 0 && a.scale_out) r = r * a.out_scale;
        f.out[i] = r;
    }
}
template <int AXIS, bool HAT, bool POW2, bool BUF = false>
__global__ void __launch_bounds__(kBlock) k_advect_sl(AdvArgs a) {
    if (BUF) advect_body_buf<false, AXIS, HAT, POW2>(a, blockIdx.x, gridDim.x);
    else advect_body<false, AXIS, HAT, POW2>(a, blockIdx.x, gridDim.x);
}
template <int AXIS, bool HAT, bool POW2, bool BUF = false>
__global__ void __launch_bounds__(kBlock) k_advect_mc(AdvArgs a) {
    if (BUF) advect_body_buf<true, AXIS, HAT, POW2>(a, blockIdx.x, gridDim.x);
    else advect_body<true, AXIS, HAT, POW2>(a, blockIdx.x, gridDim.x);
}

// --------------------------------------------------------------- projection

__device__ __forceinline__ void write_partials(double* partials, const double* v, int count, int id) {
    if (block_linear_thread() == 0) {
        double* dst = partials + (size_t)id * kPartialStride;
        for (int k = 0; k < count; ++k) dst[k] = v[k];
    }
}
__device__ __forceinline__ void write_partials(double* partials, const double* v, int count) {
    write_partials(partials, v, count, block_linear_id());
}

// Block coordinates of a cell-stencil launch (64 x 4 lanes over x, y; the
// z planes [z0, z1)) and the block's partial-record index.
struct Blk {
    int x, y, z0, z1, id;
};
__device__ __forceinline__ Blk blk_of(int n, int bx, int by, int bz, int id) {
    const int chunk = cell_geom(n).chunk, z0 = bz * chunk;
    return {bx, by, z0, z0 + chunk < n ? z0 + chunk : n, id};
}
// The tile of a cell-stencil block. The hardware deals blocks to the 8 XCDs
// round-robin in launch order; the tiles are renumbered so that each XCD
// takes one contiguous run of them (a slab of z planes), and the y and z halo
// rows a block reads were mostly staged by blocks on the same XCD, in its L2.
// A tile keeps its own index for its partial record, so reductions run in the
// same order as with the plain mapping.
__device__ __forceinline__ Blk hw_blk(int n) {
    const CellGeom g = cell_geom(n);
    const int total = g.gx * g.gy * g.gz, full = total & ~7;
    const int l = block_linear_id();
    const int m = l < full ? (l & 7) * (full >> 3) + (l >> 3) : l;
    const int bx = m % g.gx, by = (m / g.gx) % g.gy, bz = m / (g.gx * g.gy);
    return blk_of(n, bx, by, bz, m);
}

// Per-block FP64 partials of r . e and e . (A e) (A the Neumann matrix of
// the projection): the energy-optimal step along a correction e.
__global__ void __launch_bounds__(kBlock) k_energy_dots(int n, const float* __restrict__ r, const float* __restrict__ e,
                                                        double* partials) {
    __shared__ double lds[2 * 16];
    const int nn = n * n;
    const Blk k = hw_blk(n);
    double acc[2] = {0.0, 0.0};
    const int xi = k.x * 64 + threadIdx.x, y = k.y * 4 + threadIdx.y;
    if (xi < n && y < n) {
        int i = idC(n, xi, y, k.z0);
        float em = k.z0 > 0 ? e[i - nn] : 0.f, ec = e[i];
        for (int z = k.z0; z < k.z1; ++z, i += nn) {
            const float ep = z + 1 < n ? e[i + nn] : 0.f;
            const double c = ec;
            const double d = (double)((xi > 0) + (xi + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
            double ae = d * c;
            if (xi > 0) ae -= e[i - 1];
            if (xi + 1 < n) ae -= e[i + 1];
            if (y > 0) ae -= e[i - n];
            if (y + 1 < n) ae -= e[i + n];
            if (z > 0) ae -= em;
            if (z + 1 < n) ae -= ep;
            acc[0] += (double)r[i] * c;
            acc[1] += c * ae;
            em = ec;
            ec = ep;
        }
    }
    const Red ops[2] = {Red::Sum, Red::Sum};
    block_reduce<2>(acc, ops, lds);
    write_partials(partials, acc, 2);
}

// ---- the learned start's decisions on the device ----
__device__ __forceinline__ double compat_rel_dev(double r2, double rsum, double b2, double bsum, double count) {
    const double num = r2 - rsum * rsum / count, den = b2 - bsum * bsum / count;
    if (!(den > 0.0)) return 0.0;
    return sqrt((num > 0.0 ? num : 0.0) / den);
}
// One block: each thread sums a strided share of the partial records in
// order, then the fixed block tree (deterministic run to run).
__global__ void __launch_bounds__(kBlock) k_line_search(const double* __restrict__ partials, int nblocks,
                                                        ProjState* ps) {
    __shared__ double lds[2 * 16];
    double acc[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < nblocks; k += kBlock) {
        acc[0] += partials[(size_t)k * kPartialStride];
        acc[1] += partials[(size_t)k * kPartialStride + 1];
    }
    const Red ops[2] = {Red::Sum, Red::Sum};
    block_reduce<2>(acc, ops, lds);
    if (threadIdx.x == 0) {
        const double ratio = acc[0] / acc[1];
        ps->learned_alpha = (acc[1] > 0.0 && isfinite(ratio)) ? ratio : 0.0;
    }
}
// out = x + alpha e; out may be x (each element is read and written by one lane).
__global__ void __launch_bounds__(kBlock) k_axpy_alpha(const float* x, const ProjState* __restrict__ ps,
                                                       const float* __restrict__ e, float* out, size_t count) {
    const float a = (float)ps->learned_alpha;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = x[i] + a * e[i];
}
__device__ __forceinline__ void judge_body(ProjState* plain, const ProjState* learned, double count) {
    const double rel_plain = compat_rel_dev(plain->res2, plain->res_sum, plain->bnorm2, plain->b_sum, count);
    const double rel_learned =
        compat_rel_dev(learned->res2, learned->res_sum, learned->bnorm2, learned->b_sum, count);
    const double alpha = plain->learned_alpha;
    if (!learned->nonfinite_x && rel_learned < rel_plain) {
        *plain = *learned;
        plain->learned = 1;
    } else {
        plain->learned = 2;
    }
    plain->learned_alpha = alpha;
}
__global__ void k_judge(ProjState* plain, const ProjState* learned, double count) { judge_body(plain, learned, count); }
__global__ void __launch_bounds__(kBlock) k_copy_if_learned(const float* __restrict__ src, float* __restrict__ dst,
                                                            const ProjState* __restrict__ ps, size_t count) {
    if (ps->learned != 1) return;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

// computeDivergence on the state with closed walls, b = -div h^2/dt
// (CpuPressureSolver3D.cpp:35-42, MacGridOperators3D.cpp:8-13). Wall faces are
// read as zero without being written, so a reflected pass keeps its advected
// wall values (MacGridFluidSolver3D.cpp:203-207 copies them before projecting).
__device__ __forceinline__ void divergence_body(int n, float h, float scale, const float* __restrict__ u,
                                                const float* __restrict__ v, const float* __restrict__ w,
                                                float* __restrict__ b, double* partials, Blk k) {
    __shared__ double lds[5 * 16];
    double acc[5] = {0.0, 0.0, -kInf, kInf, 0.0};
    const int x = k.x * 64 + threadIdx.x, y = k.y * 4 + threadIdx.y;
    if (x < n && y < n) {
        float wlo = k.z0 == 0 ? 0.f : w[idZ(n, x, y, k.z0)];
        for (int z = k.z0; z < k.z1; ++z) {
            const float ulo = x == 0 ? 0.f : u[idX(n, x, y, z)], uhi = x + 1 == n ? 0.f : u[idX(n, x + 1, y, z)];
            const float vlo = y == 0 ? 0.f : v[idY(n, x, y, z)], vhi = y + 1 == n ? 0.f : v[idY(n, x, y + 1, z)];
            const float whi = z + 1 == n ? 0.f : w[idZ(n, x, y, z + 1)];
            const float div = (uhi - ulo + vhi - vlo + whi - wlo) / h;
            if (!isfinite(div)) acc[4] += 1.0;
            const double ad = (double)fabsf(div);
            acc[1] = acc[1] < ad ? ad : acc[1];
            const float bv = -div * scale;
            b[idC(n, x, y, z)] = bv;
            acc[0] += (double)bv;
            acc[2] = acc[2] < (double)bv ? (double)bv : acc[2];
            acc[3] = (double)bv < acc[3] ? (double)bv : acc[3];
            wlo = whi;
        }
    }
    const Red ops[5] = {Red::Sum, Red::Max, Red::Max, Red::Min, Red::Sum};
    block_reduce<5>(acc, ops, lds);
    write_partials(partials, acc, 5, k.id);
}
__global__ void __launch_bounds__(kBlock) k_divergence_rhs(int n, float h, float scale, const float* __restrict__ u,
                                                           const float* __restrict__ v, const float* __restrict__ w,
                                                           float* __restrict__ b, double* partials) {
    divergence_body(n, h, scale, u, v, w, b, partials, hw_blk(n));
}

// ---------------------------------------------------------- fused RB sweep

// One full red-black sweep per launch (both colours), out of place: reads
// xin, writes xout, so the halos a block reads are never written in the same
// launch. Each block owns a 64 x 8 x 8 tile and stages xin over the tile plus a
// 2-cell halo in LDS; it updates the red cells (parity 0) the tile's black
// cells read (the tile's own and those of its 1-cell ring, which the
// neighbouring block recomputes from the same old values, identically), then
// the black cells of the tile, and writes the tile. Both colours see exactly
// the values the reference's sequential red-then-black sweep gives them
// (SOR.cpp:36-49), with the same neighbour-sum order, so results are
// bit-identical to it. Global traffic is the compulsory read x, read b,
// write x: 12 B per cell.
//
// Tiles are numbered x-fastest, then y, then z, and block b takes tile
// (b % 8) * ceil(T / 8) + b / 8: the hardware deals blocks round-robin to the
// 8 XCDs, so each XCD gets a contiguous z-slab of tiles and the halo rows its
// blocks share stay in that XCD's L2.
//
// MODE 0: sweep only. MODE 1: also the FP64 residual of xin on the tile
// (the partials of k_residual). MODE 2: also sum(xout) (Finalize::XSum).
// MODE 3: the partials of a chunk check (check_body): the residual of xin
// without the right-hand side's sums, and sum(xout). FAST: every cell's
// diagonal is a.D (see rb_div). With `done` set (a converged tolerance solve),
// the launch does nothing.
struct RbShared {
    float xs[kRbXWords];
    float bs[kRbBWords];
    double lds[8 * 16];
};

template <int MODE, bool FAST, bool VEC>
__device__ __forceinline__ void sor_body(int n, const RbArgs& a, const float* __restrict__ b,
                                         const float* __restrict__ xin, float* __restrict__ xout,
                                         const ProjState* __restrict__ ps, const int* __restrict__ done,
                                         double* partials, int block, const RbTile& t, RbShared& sh) {
    float* xs = sh.xs;
    float* bs = sh.bs;
    double* lds = sh.lds;
    // issued before the staging loads (a vector load, so waiting for it does
    // not wait for the scalar loads), tested after them: a converged solve's
    // launches return without writing anything
    const int stop_v = done ? __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    const int tid = threadIdx.x;
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t.valid) rb_stage<VEC>(xs, bs, xin, b, ps->rhs_meanf, n, t, tid);
    if (__builtin_amdgcn_readfirstlane(stop_v)) return;
    if (t.valid) {
        __syncthreads();
        if (MODE == 1 || MODE == 3) {
            rb_residual<MODE == 1>(xs, bs, n, t, tid, acc);
            // every wave's residual reads the staged input before any wave's
            // red pass overwrites red cells of it
            __syncthreads();
        }
        if (n > 1) {  // a single cell has no neighbours and is never updated
            if (rb_inner(n, t)) {
                rb_red<FAST, true>(xs, bs, a, n, t, tid);
                __syncthreads();
                rb_black<FAST, true>(xs, bs, a, n, t, tid);
            } else {
                rb_red<FAST, false>(xs, bs, a, n, t, tid);
                __syncthreads();
                rb_black<FAST, false>(xs, bs, a, n, t, tid);
            }
        }
        __syncthreads();
        const double sum = rb_write<VEC>(xs, xout, n, t, tid);
        if (MODE == 2) acc[0] = sum;
        if (MODE == 3) acc[7] = sum;
    }
    if (MODE == 1) {
        const Red ops[8] = {Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum};
        block_reduce<8>(acc, ops, lds);
        write_partials(partials, acc, 8, block);
    } else if (MODE == 3) {
        // sum r^2, max |r|, non-finite x, sum r, sum xout (check_body)
        double c[5] = {acc[0], acc[1], acc[4], acc[5], acc[7]};
        const Red ops[5] = {Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum};
        block_reduce<5>(c, ops, lds);
        write_partials(partials, c, 5, block);
    } else if (MODE == 2) {
        double one[1] = {acc[0]};
        const Red ops[1] = {Red::Sum};
        block_reduce<1>(one, ops, lds);
        write_partials(partials, one, 1, block);
    }
}
template <int MODE, bool FAST, bool VEC>
__global__ void __launch_bounds__(kRbThreads) k_sor_rb(int n, RbGeom geo, RbArgs a, const float* __restrict__ b,
                                                       const float* __restrict__ xin, float* __restrict__ xout,
                                                       const ProjState* __restrict__ ps, const int* __restrict__ done,
                                                       double* partials) {
    __shared__ RbShared sh;
    sor_body<MODE, FAST, VEC>(n, a, b, xin, xout, ps, done, partials, blockIdx.x, rb_tile(geo, blockIdx.x), sh);
}

// The register path of the sweep (rb_fetch_x_reg / rb_fetch_b_reg in
// vgx_device.hpp): the same phases with the right-hand side in registers and
// the staging loads at strided buffer offsets, for MODE 0 and MODE 2 when
// rb_reg_fits(n). Results are bit-identical to sor_body's.
struct RbSharedReg {
    float xs[kRbXWords];
    double lds[16];
};
template <int MODE, bool FAST>
__device__ __forceinline__ void sor_body_reg(int n, const RbArgs& a, const float* __restrict__ b,
                                             const float* __restrict__ xin, float* __restrict__ xout,
                                             const ProjState* __restrict__ ps, const int* __restrict__ done,
                                             double* partials, int block, const RbTile& t, RbSharedReg& sh) {
    static_assert(MODE == 0 || MODE == 2, "the register path has no residual modes");
    const int stop_v = done ? __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    const int tid = threadIdx.x;
    float4 v[kRbSIter];
    RbRegs r;
    if (t.valid) {
        rb_fetch_x_reg(rb_field(xin, n), n, t, tid, v);
        rb_fetch_b_reg(r, rb_field(b, n), n, t, tid);
    }
    if (__builtin_amdgcn_readfirstlane(stop_v)) return;
    double sum = 0.0;
    if (t.valid) {
        const float meanf = ps->rhs_meanf;
        rb_put_x_reg(sh.xs, tid, v);
        __syncthreads();
        if (n > 1) {
            if (rb_inner(n, t)) {
                rb_red_reg<FAST, true>(sh.xs, r, meanf, a, n, t, tid);
                __syncthreads();
                rb_black_reg<FAST, true>(sh.xs, r, meanf, a, n, t, tid);
            } else {
                rb_red_reg<FAST, false>(sh.xs, r, meanf, a, n, t, tid);
                __syncthreads();
                rb_black_reg<FAST, false>(sh.xs, r, meanf, a, n, t, tid);
            }
        }
        __syncthreads();
        sum = rb_write<true>(sh.xs, xout, n, t, tid);
    }
    if (MODE == 2) {
        double one[1] = {sum};
        const Red ops[1] = {Red::Sum};
        block_reduce<1>(one, ops, sh.lds);
        write_partials(partials, one, 1, block);
    }
}
template <int MODE, bool FAST>
__global__ void __launch_bounds__(kRbThreads) k_sor_rb_reg(int n, RbGeom geo, RbArgs a, const float* __restrict__ b,
                                                           const float* __restrict__ xin, float* __restrict__ xout,
                                                           const ProjState* __restrict__ ps,
                                                           const int* __restrict__ done, double* partials) {
    __shared__ RbSharedReg sh;
    sor_body_reg<MODE, FAST>(n, a, b, xin, xout, ps, done, partials, blockIdx.x, rb_tile(geo, blockIdx.x), sh);
}

// End of a tolerance chunk: folds the residual partials of the chunk's last
// sweep (the iterate that sweep started from) into the record, counts the
// chunk's sweeps and raises `done` when the compatible relative L2 residual
// meets the tolerance. Does nothing once `done` is set, like the sweeps.
// The record goes to ring[(chunks % ring_size) * ring_stride + ring_off]: a
// batch of scenes shares one ring of per-scene rows.
__device__ __forceinline__ void check_body(const double* __restrict__ partials, int nblocks, double count,
                                           ProjState* ps, double tol, int chunk, ProjState* ring, int ring_size,
                                           int ring_stride, int ring_off) {
    __shared__ double lds[8 * 16];
    if (ps->done) {
        if (threadIdx.x == 0) {
            ring[(ps->chunks % ring_size) * ring_stride + ring_off] = *ps;
            ps->chunks += 1;
        }
        return;
    }
    // the MODE 3 partials: sum r^2, max |r|, non-finite x, sum r, sum xout
    const Red ops[5] = {Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum};
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int bidx = threadIdx.x; bidx < nblocks; bidx += blockDim.x) {
        const double* src = partials + (size_t)bidx * kPartialStride;
        for (int k = 0; k < 5; ++k) acc[k] = ops[k] == Red::Max ? (acc[k] < src[k] ? src[k] : acc[k]) : acc[k] + src[k];
    }
    block_reduce<5>(acc, ops, lds);
    if (threadIdx.x != 0) return;
    ps->res2 = acc[0];
    ps->res_inf = acc[1];
    ps->nonfinite_x = acc[2] > 0.0;
    ps->res_sum = acc[3];
    // sum of the chunk's output iterate: the zero-mean gauge of the result
    ps->x_sum = acc[4];
    ps->x_meanf = (float)(acc[4] / count);
    // bnorm2 and b_sum are the solve's start residual's: b does not change
    const double num = acc[0] - acc[3] * acc[3] / count, den = ps->bnorm2 - ps->b_sum * ps->b_sum / count;
    const double rel = den > 0.0 ? sqrt((num > 0.0 ? num : 0.0) / den) : 0.0;
    ps->checked_rel = rel;
    ps->sweeps_run += chunk;
    if (rel <= tol || acc[2] > 0.0) {
        ps->done = 1;
        ps->done_clock = (long long)wall_clock64();
    }
    ring[(ps->chunks % ring_size) * ring_stride + ring_off] = *ps;
    ps->chunks += 1;
}
__global__ void __launch_bounds__(1024) k_check(const double* __restrict__ partials, int nblocks, double count,
                                                ProjState* ps, double tol, int chunk, ProjState* ring,
                                                int ring_size) {
    check_body(partials, nblocks, count, ps, tol, chunk, ring, ring_size, 1, 0);
}

__global__ void k_reset_status(ProjState* ps) {
    ps->done = 0;
    ps->chunks = 0;
    ps->sweeps_run = 0;
    ps->checked_rel = -1.0;
}

// A x at one row in FP64, StencilLinearSystem.cpp:60-67 (same term order);
// the z neighbours zm, zp come from the caller's register window.
__device__ __forceinline__ double ax_row(int n, const float* __restrict__ x, int i, int xi, int y, int z, float c,
                                         float zm, float zp) {
    const double d = (double)((xi > 0) + (xi + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
    double ax = d * (double)c;
    if (xi > 0) ax += -1.0 * (double)x[i - 1];
    if (xi + 1 < n) ax += -1.0 * (double)x[i + 1];
    if (y > 0) ax += -1.0 * (double)x[i - n];
    if (y + 1 < n) ax += -1.0 * (double)x[i + n];
    if (z > 0) ax += -1.0 * (double)zm;
    if (z + 1 < n) ax += -1.0 * (double)zp;
    return ax;
}

// FP64 residual check: ||b' - A x||_2^2, its inf-norm, ||b'||_2^2 and sum(x).
__device__ __forceinline__ void residual_body(int n, const float* __restrict__ b, const float* __restrict__ x,
                                              const ProjState* __restrict__ ps, double* partials, Blk k) {
    __shared__ double lds[7 * 16];
    const float meanf = ps->rhs_meanf;
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int xi = k.x * 64 + threadIdx.x, y = k.y * 4 + threadIdx.y;
    if (xi < n && y < n) {
        const int nn = n * n;
        int i = idC(n, xi, y, k.z0);
        float xm = k.z0 > 0 ? x[i - nn] : 0.f, c = x[i];
        for (int z = k.z0; z < k.z1; ++z, i += nn) {
            const float xp = z + 1 < n ? x[i + nn] : 0.f;
            const float bi = b[i] - meanf;
            const double r = (double)bi - ax_row(n, x, i, xi, y, z, c, xm, xp);
            acc[0] += r * r;
            const double ar = fabs(r);
            acc[1] = acc[1] < ar ? ar : acc[1];
            acc[2] += (double)bi * (double)bi;
            acc[3] += (double)c;
            if (!isfinite(c)) acc[4] += 1.0;
            acc[5] += r;
            acc[6] += (double)bi;
            xm = c;
            c = xp;
        }
    }
    const Red ops[7] = {Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum};
    block_reduce<7>(acc, ops, lds);
    write_partials(partials, acc, 7, k.id);
}
__global__ void __launch_bounds__(kBlock) k_residual(int n, const float* __restrict__ b, const float* __restrict__ x,
                                                     const ProjState* __restrict__ ps, double* partials) {
    residual_body(n, b, x, ps, partials, hw_blk(n));
}
// Two residual checks with one right-hand side in one launch (the learned
// start's judge): each block computes k_residual's block for x1, then for x2,
// so both partial sets are k_residual's bit for bit.
__global__ void __launch_bounds__(kBlock) k_residual_pair(int n, const float* __restrict__ b,
                                                          const float* __restrict__ x1, const ProjState* __restrict__ ps1,
                                                          double* partials1, const float* __restrict__ x2,
                                                          const ProjState* __restrict__ ps2, double* partials2) {
    const Blk k = hw_blk(n);
    residual_body(n, b, x1, ps1, partials1, k);
    __syncthreads();
    residual_body(n, b, x2, ps2, partials2, k);
}

// FP64 residual field: r = b' - A x in FP64, stored as FP32, for an FP64
// iterate (the iterative-refinement loop's next right-hand side) or an FP32
// one (the residual the learned start reads). The statistics are those of
// k_residual: the same blocks, cells and order, and for an FP32 iterate the
// same values (x widens exactly, a - y equals a + (-1) y), so its partials
// are k_residual's bit for bit.
// With xcopy (FP32 iterates), every cell's x is also written there.
template <class T>
__global__ void __launch_bounds__(kBlock) k_residual_field(int n, const float* __restrict__ b, const T* __restrict__ x,
                                                           const ProjState* __restrict__ ps, float* __restrict__ r32,
                                                           double* partials, float* __restrict__ xcopy) {
    __shared__ double lds[7 * 16];
    const float meanf = ps->rhs_meanf;
    const int nn = n * n;
    const Blk k = hw_blk(n);
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int xi = k.x * 64 + threadIdx.x, y = k.y * 4 + threadIdx.y;
    if (xi < n && y < n) {
        int i = idC(n, xi, y, k.z0);
        double xm = k.z0 > 0 ? (double)x[i - nn] : 0.0, c = (double)x[i];
        for (int z = k.z0; z < k.z1; ++z, i += nn) {
            const double xp = z + 1 < n ? (double)x[i + nn] : 0.0;
            const double d = (double)((xi > 0) + (xi + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
            double ax = d * c;
            if (xi > 0) ax -= (double)x[i - 1];
            if (xi + 1 < n) ax -= (double)x[i + 1];
            if (y > 0) ax -= (double)x[i - n];
            if (y + 1 < n) ax -= (double)x[i + n];
            if (z > 0) ax -= xm;
            if (z + 1 < n) ax -= xp;
            const float bi = b[i] - meanf;
            const double r = (double)bi - ax;
            r32[i] = (float)r;
            if (xcopy) xcopy[i] = (float)c;
            acc[0] += r * r;
            const double ar = fabs(r);
            acc[1] = acc[1] < ar ? ar : acc[1];
            acc[2] += (double)bi * (double)bi;
            acc[3] += c;
            if (!isfinite(c)) acc[4] += 1.0;
            acc[5] += r;
            acc[6] += (double)bi;
            xm = c;
            c = xp;
        }
    }
    const Red ops[7] = {Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum};
    block_reduce<7>(acc, ops, lds);
    write_partials(partials, acc, 7, k.id);
}

// x64 += e (refinement correction), x64 <- x32, x32 <- x64.
__global__ void __launch_bounds__(kBlock) k_axpy_f64(double* x, const float* __restrict__ e, size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        x[i] += (double)e[i];
}
__global__ void __launch_bounds__(kBlock) k_widen(double* x, const float* __restrict__ e, size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        x[i] = (double)e[i];
}
__global__ void __launch_bounds__(kBlock) k_narrow(float* x, const double* __restrict__ e, size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        x[i] = (float)e[i];
}

// Zero-mean gauge, applyPressureGradient and closed walls
// (CpuPressureSolver3D.cpp:50-58, MacGridOperators3D.cpp:51-67), out of place:
// reads the unprojected faces `in`, writes the projected faces `out` and the
// gauged pressure p_out. Each cell writes its low faces (and the high wall
// face on the last layer) and recomputes its high faces to obtain the max
// divergence afterwards (MacGridOperators3D.cpp:40-49); it also evaluates the
// final residual b' - A p' in FP64.
__device__ __forceinline__ void gradient_body(int n, float h, float dt, const float* __restrict__ uin,
                                              const float* __restrict__ vin, const float* __restrict__ win,
                                              const float* __restrict__ p, float* __restrict__ p_out,
                                              const float* __restrict__ b, float* __restrict__ uout,
                                              float* __restrict__ vout, float* __restrict__ wout,
                                              const ProjState* __restrict__ ps, double* partials, Blk k) {
    __shared__ double lds[7 * 16];
    const float xm = ps->x_meanf, bm = ps->rhs_meanf;
    const float hInv = 1.f / h;
    const int nn = n * n;
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int x = k.x * 64 + threadIdx.x, y = k.y * 4 + threadIdx.y;
    if (x < n && y < n) {
        // register windows along z: the pressure below / at / above the
        // cell, and the w face between this cell and the next
        int i = idC(n, x, y, k.z0);
        float rzm = k.z0 > 0 ? p[i - nn] : 0.f, rc = p[i];
        float wraw = k.z0 > 0 ? win[idZ(n, x, y, k.z0)] : 0.f;
        for (int z = k.z0; z < k.z1; ++z, i += nn) {
            const float rzp = z + 1 < n ? p[i + nn] : 0.f;
            const float wnext = z + 1 < n ? win[idZ(n, x, y, z + 1)] : 0.f;
            const float pc = rc - xm;
            const float pxm = x > 0 ? p[i - 1] - xm : 0.f, pxp = x + 1 < n ? p[i + 1] - xm : 0.f;
            const float pym = y > 0 ? p[i - n] - xm : 0.f, pyp = y + 1 < n ? p[i + n] - xm : 0.f;
            const float pzm = z > 0 ? rzm - xm : 0.f, pzp = z + 1 < n ? rzp - xm : 0.f;
            p_out[i] = pc;
            const float ulo = x == 0 ? 0.f : uin[idX(n, x, y, z)] - dt * (pc - pxm) * hInv;
            const float uhi = x + 1 == n ? 0.f : uin[idX(n, x + 1, y, z)] - dt * (pxp - pc) * hInv;
            const float vlo = y == 0 ? 0.f : vin[idY(n, x, y, z)] - dt * (pc - pym) * hInv;
            const float vhi = y + 1 == n ? 0.f : vin[idY(n, x, y + 1, z)] - dt * (pyp - pc) * hInv;
            const float wlo = z == 0 ? 0.f : wraw - dt * (pc - pzm) * hInv;
            const float whi = z + 1 == n ? 0.f : wnext - dt * (pzp - pc) * hInv;
            rzm = rc;
            rc = rzp;
            wraw = wnext;
            uout[idX(n, x, y, z)] = ulo;
            if (x + 1 == n) uout[idX(n, n, y, z)] = 0.f;
            vout[idY(n, x, y, z)] = vlo;
            if (y + 1 == n) vout[idY(n, x, n, z)] = 0.f;
            wout[idZ(n, x, y, z)] = wlo;
            if (z + 1 == n) wout[idZ(n, x, y, n)] = 0.f;
            const float div = (uhi - ulo + vhi - vlo + whi - wlo) / h;
            if (!isfinite(div)) acc[3] += 1.0;
            const double ad = (double)fabsf(div);
            acc[2] = acc[2] < ad ? ad : acc[2];
            // Final residual on the gauged pressure, same term order as ax_row.
            const double d = (double)((x > 0) + (x + 1 < n) + (y > 0) + (y + 1 < n) + (z > 0) + (z + 1 < n));
            double ax = d * (double)pc;
            if (x > 0) ax += -1.0 * (double)pxm;
            if (x + 1 < n) ax += -1.0 * (double)pxp;
            if (y > 0) ax += -1.0 * (double)pym;
            if (y + 1 < n) ax += -1.0 * (double)pyp;
            if (z > 0) ax += -1.0 * (double)pzm;
            if (z + 1 < n) ax += -1.0 * (double)pzp;
            const float bi = b[i] - bm;
            const double r = (double)bi - ax;
            acc[0] += r * r;
            const double ar = fabs(r);
            acc[1] = acc[1] < ar ? ar : acc[1];
            acc[4] += (double)bi * (double)bi;
            acc[5] += r;
            acc[6] += (double)bi;
        }
    }
    const Red ops[7] = {Red::Sum, Red::Max, Red::Max, Red::Sum, Red::Sum, Red::Sum, Red::Sum};
    block_reduce<7>(acc, ops, lds);
    write_partials(partials, acc, 7, k.id);
}
__global__ void __launch_bounds__(kBlock) k_gradient(int n, float h, float dt, const float* __restrict__ uin,
                                                     const float* __restrict__ vin, const float* __restrict__ win,
                                                     const float* __restrict__ p, float* __restrict__ p_out,
                                                     const float* __restrict__ b, float* __restrict__ uout,
                                                     float* __restrict__ vout, float* __restrict__ wout,
                                                     const ProjState* __restrict__ ps, double* partials) {
    gradient_body(n, h, dt, uin, vin, win, p, p_out, b, uout, vout, wout, ps, partials, hw_blk(n));
}

// ------------------------------------------------------------------ finalize

enum Mode { MODE_RHS = 0, MODE_XSUM, MODE_RESIDUAL, MODE_FINAL, MODE_DIAG };

// One block folds the per-block partials in a fixed order and derives the
// projection record.
template <int MODE>
__device__ __forceinline__ void finalize_body(const double* __restrict__ partials, int nblocks, double count,
                                              ProjState* ps, DiagState* ds) {
    __shared__ double lds[7 * 16];
    constexpr int kFields = MODE == MODE_XSUM ? 1 : MODE == MODE_DIAG ? 4 : MODE == MODE_RHS ? 5 : 7;
    Red ops[7] = {Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum, Red::Sum};
    if (MODE == MODE_RHS) {
        ops[1] = Red::Max;
        ops[2] = Red::Max;
        ops[3] = Red::Min;
    } else if (MODE == MODE_RESIDUAL) {
        ops[1] = Red::Max;
    } else if (MODE == MODE_FINAL) {
        ops[1] = Red::Max;
        ops[2] = Red::Max;
    } else if (MODE == MODE_DIAG) {
        ops[1] = Red::Max;
        ops[2] = Red::Max;
    }
    double acc[7];
    for (int k = 0; k < 7; ++k) acc[k] = ops[k] == Red::Sum ? 0.0 : ops[k] == Red::Max ? -kInf : kInf;
    for (int bidx = threadIdx.x; bidx < nblocks; bidx += blockDim.x) {
        const double* src = partials + (size_t)bidx * kPartialStride;
        for (int k = 0; k < kFields; ++k) {
            const double o = src[k];
            acc[k] = ops[k] == Red::Sum ? acc[k] + o : ops[k] == Red::Max ? (acc[k] < o ? o : acc[k])
                                                                           : (o < acc[k] ? o : acc[k]);
        }
    }
    block_reduce<7>(acc, ops, lds);
    if (threadIdx.x != 0) return;
    if (MODE == MODE_RHS) {
        ps->rhs_sum = acc[0];
        ps->rhs_mean = acc[0] / count;
        const float m = (float)ps->rhs_mean;
        ps->rhs_meanf = m;
        ps->div_before = (float)acc[1];
        // inf-norm of b' = b - m at x = 0: float subtraction is monotonic, so
        // the extremes of b map to the extremes of b'.
        const double hi = fabs((double)((float)acc[2] - m)), lo = fabs((double)((float)acc[3] - m));
        ps->initial_inf = count > 0 ? (hi < lo ? lo : hi) : 0.0;
        ps->nonfinite_in = acc[4] > 0.0;
    } else if (MODE == MODE_XSUM) {
        ps->x_sum = acc[0];
        ps->x_meanf = (float)(acc[0] / count);
    } else if (MODE == MODE_RESIDUAL) {
        ps->res2 = acc[0];
        ps->res_inf = acc[1];
        ps->bnorm2 = acc[2];
        ps->x_sum = acc[3];
        ps->x_meanf = (float)(acc[3] / count);
        ps->nonfinite_x = acc[4] > 0.0;
        ps->res_sum = acc[5];
        ps->b_sum = acc[6];
    } else if (MODE == MODE_FINAL) {
        ps->final_res2 = acc[0];
        ps->final_inf = acc[1];
        ps->div_after = (float)acc[2];
        ps->nonfinite_out = acc[3] > 0.0;
        ps->final_bnorm2 = acc[4];
        ps->final_res_sum = acc[5];
        ps->final_b_sum = acc[6];
    } else {
        ds->ke_sum = acc[0];
        ds->max_vel = acc[1];
        ds->max_div = acc[2];
        ds->nonfinite = acc[3];
    }
}
template <int MODE>
__global__ void __launch_bounds__(1024) k_finalize(const double* __restrict__ partials, int nblocks, double count,
                                                   ProjState* ps, DiagState* ds) {
    finalize_body<MODE>(partials, nblocks, count, ps, ds);
}

// --------------------------------------------------------------- diagnostics

// kineticEnergy / maxVelocity / maxDivergence (MacGridFluidSolver3D.cpp:247-265,
// MacGridOperators3D.cpp:40-49) in one pass: faces by index, cells by index.
__global__ void __launch_bounds__(kBlock) k_diagnostics(int n, float h, const float* __restrict__ u,
                                                        const float* __restrict__ v, const float* __restrict__ w,
                                                        double* partials) {
    __shared__ double lds[4 * 16];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int faces = (n + 1) * n * n, cells = n * n * n;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < faces; i += gridDim.x * kBlock) {
        const float a = u[i], bb = v[i], c = w[i];
        acc[0] += (double)a * a + (double)bb * bb + (double)c * c;
        const double m = (double)maxf(maxf(fabsf(a), fabsf(bb)), fabsf(c));
        acc[1] = acc[1] < m ? m : acc[1];
        if (i < cells) {
            const int x = i % n, t = i / n, y = t % n, z = t / n;
            const float div = (u[idX(n, x + 1, y, z)] - u[idX(n, x, y, z)] + v[idY(n, x, y + 1, z)] -
                               v[idY(n, x, y, z)] + w[idZ(n, x, y, z + 1)] - w[idZ(n, x, y, z)]) / h;
            if (!isfinite(div)) acc[3] += 1.0;
            const double ad = (double)fabsf(div);
            acc[2] = acc[2] < ad ? ad : acc[2];
        }
    }
    const Red ops[4] = {Red::Sum, Red::Max, Red::Max, Red::Sum};
    block_reduce<4>(acc, ops, lds);
    write_partials(partials, acc, 4);
}

// The judge's two residual records folded (Finalize::Residual each, in the
// same order as k_finalize) and the judge applied, in one block.
__global__ void __launch_bounds__(1024) k_judge_pair(const double* __restrict__ partials1,
                                                     const double* __restrict__ partials2, int nblocks, double count,
                                                     ProjState* plain, ProjState* learned) {
    finalize_body<MODE_RESIDUAL>(partials1, nblocks, count, plain, nullptr);
    __syncthreads();
    finalize_body<MODE_RESIDUAL>(partials2, nblocks, count, learned, nullptr);
    __syncthreads();
    if (threadIdx.x == 0) judge_body(plain, learned, count);
}

template <int MODE> hipError_t finalize(hipStream_t s, const double* partials, int nblocks, double count,
                                        ProjState* ps, DiagState* ds) {
    vlaunch(k_finalize<MODE>, dim3(1), dim3(1024), s, partials, nblocks, count, ps, ds);
    return hipGetLastError();
}

int blocks_of(const Launch& l) { return (int)(l.grid.x * l.grid.y * l.grid.z); }

// ------------------------------------------------------------------ batches
//
// A batched launch runs one kernel over every scene of a batch. Scene s owns
// items [prefix[s], prefix[s+1]) of the launch's work list; item w is block
// w - prefix[s] of the single-scene launch with scene s's arguments, so each
// scene computes exactly what it computes alone (the same decomposition, the
// same reduction partials); a table gives each item's scene. Blocks take
// items w = blockIdx.x + k gridDim.x: the grid may be smaller than the list
// (a persistent grid), which the autotuner chooses per kernel family.

// Block coordinates of item `local` of an n^3 cell-stencil launch (the
// decomposition of cell_launch).
__device__ __forceinline__ Blk cell_blk(int n, int local) {
    const CellGeom g = cell_geom(n);
    return blk_of(n, local % g.gx, (local / g.gx) % g.gy, local / (g.gx * g.gy), local);
}

__global__ void __launch_bounds__(1024) kb_emit_noise(const EmitSlot* __restrict__ t) {
    const EmitSlot& e = t[blockIdx.x];
    if (e.active) emit_noise_body(e.a, e.rng, e.noise);
}

__global__ void __launch_bounds__(kBlock) kb_emit_apply(const EmitSlot* __restrict__ t, BatchList L) {
    for (int w = blockIdx.x; w < L.items; w += gridDim.x) {
        const int s = L.owner[w];
        const EmitSlot& e = t[s];
        emit_apply_body(e.a, e.noise, e.rho, e.u, e.v, e.w, w - L.prefix[s]);
    }
}

__global__ void __launch_bounds__(kBlock) kb_forces(const ForceArgs* __restrict__ t, BatchList L) {
    for (int w = blockIdx.x; w < L.items; w += gridDim.x) {
        const int s = L.owner[w];
        const ForceArgs& f = t[s];
        forces_body(f.n, f.rho, f.u, f.v, f.w, f.buoyancy, f.dt, f.split, w - L.prefix[s],
                    L.prefix[s + 1] - L.prefix[s]);
    }
}

template <bool MC, int AXIS, bool HAT>
__global__ void __launch_bounds__(kBlock) kb_advect(const AdvArgs* __restrict__ t, BatchList L) {
    for (int w = blockIdx.x; w < L.items; w += gridDim.x) {
        const int s = L.owner[w];
        const AdvArgs& a = t[s];
        // scenes differ in spacing: the multiply-by-1/h path where it is exact
        if (pow2_spacing(a.h)) advect_body<MC, AXIS, HAT, true>(a, w - L.prefix[s], L.prefix[s + 1] - L.prefix[s]);
        else advect_body<MC, AXIS, HAT, false>(a, w - L.prefix[s], L.prefix[s + 1] - L.prefix[s]);
    }
}

template <int OP>
__global__ void __launch_bounds__(kBlock) kb_cell(const CellArgs* __restrict__ t, BatchList L) {
    for (int w = blockIdx.x; w < L.items; w += gridDim.x) {
        const int s = L.owner[w];
        const CellArgs& c = t[s];
        const Blk k = cell_blk(c.n, w - L.prefix[s]);
        if (w != (int)blockIdx.x) __syncthreads();  // the reduction's LDS is reused
        if (OP == 0) divergence_body(c.n, c.h, c.scale, c.u, c.v, c.w, c.b, c.partials, k);
        else if (OP == 1) residual_body(c.n, c.b, c.p, c.ps, c.partials, k);
        else gradient_body(c.n, c.h, c.dt, c.u, c.v, c.w, c.p, c.p_out, c.b, c.uo, c.vo, c.wo, c.ps, c.partials, k);
    }
}

// Sweeps over the whole tile list L, or over the active list (compacted:
// only scenes still solving) when act.count is set.
template <int MODE, bool FAST, bool VEC>
__global__ void __launch_bounds__(kRbThreads) kb_sor(RbArgs a, const RbSlot* __restrict__ t, BatchList L,
                                                     ActiveList act) {
    __shared__ RbShared sh;
    const int items = act.count ? *act.count : L.items;
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        int s, local;
        if (act.count) {
            s = act.owner[w];
            local = act.local[w];
        } else {
            s = L.owner[w];
            local = w - L.prefix[s];
        }
        const RbSlot& r = t[s];
        if (w != (int)blockIdx.x) __syncthreads();  // the previous item's reads of the tile are done
        sor_body<MODE, FAST, VEC>(r.n, a, r.b, r.xin, r.xout, r.ps, nullptr, r.partials, local, rb_tile(r.geo, local),
                                  sh);
    }
}

// kb_sor's plain and sum-partial sweeps on the register path, for batches
// whose every scene allows it (rb_reg_fits): bit-identical, fewer
// instructions and a smaller LDS footprint than sor_body.
template <int MODE, bool FAST>
__global__ void __launch_bounds__(kRbThreads) kb_sor_reg(RbArgs a, const RbSlot* __restrict__ t, BatchList L,
                                                         ActiveList act) {
    __shared__ RbSharedReg sh;
    const int items = act.count ? *act.count : L.items;
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        int s, local;
        if (act.count) {
            s = act.owner[w];
            local = act.local[w];
        } else {
            s = L.owner[w];
            local = w - L.prefix[s];
        }
        const RbSlot& r = t[s];
        if (w != (int)blockIdx.x) __syncthreads();  // the previous item's reads of the tile are done
        sor_body_reg<MODE, FAST>(r.n, a, r.b, r.xin, r.xout, r.ps, nullptr, r.partials, local, rb_tile(r.geo, local),
                                 sh);
    }
}

// Compacts the tile list L to the scenes whose records are not done, in
// scene order: chunks of 1024 scenes get an exclusive scan of their active
// tile counts in LDS, then every item of an active scene is written to its
// place. Each scene's items stay a multiple of 8 long and start at a
// multiple of 8, so item w keeps running on XCD w % 8 with the tile order of
// the single-scene launch.
__global__ void __launch_bounds__(1024) kb_active(const FinArgs* __restrict__ t, BatchList L, ActiveList act) {
    __shared__ int sbase[1024];
    __shared__ int wsum[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int s0 = 0; s0 < L.scenes; s0 += 1024) {
        const int s = s0 + tid;
        const int c = (s < L.scenes && !t[s].ps->done) ? L.prefix[s + 1] - L.prefix[s] : 0;
        int v = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(v, o, 64);
            if (lane >= o) v += y;
        }
        if (lane == 63) wsum[wid] = v;
        __syncthreads();
        if (wid == 0) {
            int x = lane < 16 ? wsum[lane] : 0;
            for (int o = 1; o < 16; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane < 16) wsum[lane] = x;
        }
        __syncthreads();
        const int incl = v + (wid > 0 ? wsum[wid - 1] : 0);
        sbase[tid] = carry + incl - c;
        __syncthreads();
        if (tid == 1023) carry += incl;
        const int s1 = min(s0 + 1024, L.scenes);
        for (int w = L.prefix[s0] + tid; w < L.prefix[s1]; w += 1024) {
            const int o = L.owner[w];
            if (t[o].ps->done) continue;
            const int local = w - L.prefix[o];
            const int idx = sbase[o - s0] + local;
            act.owner[idx] = (uint16_t)o;
            act.local[idx] = local;
        }
        __syncthreads();
    }
    if (tid == 0) *act.count = carry;
}

template <int MODE> __global__ void __launch_bounds__(1024) kb_finalize(const FinArgs* __restrict__ t) {
    const FinArgs& f = t[blockIdx.x];
    finalize_body<MODE>(f.partials, f.nblocks, f.count, f.ps, nullptr);
}

__global__ void __launch_bounds__(1024) kb_check(const FinArgs* __restrict__ t, double tol, int chunk,
                                                 ProjState* ring, int ring_size) {
    const FinArgs& f = t[blockIdx.x];
    check_body(f.partials, f.nblocks, f.count, f.ps, tol, chunk, ring, ring_size, (int)gridDim.x, (int)blockIdx.x);
}

// Start of a batched tolerance solve, after the residual check of the
// starting iterate (Finalize::Residual): the host decisions of the
// single-scene solve, taken per scene on the device. A warm start whose
// right-hand side has no compatible part restarts from zero (zero_start; the
// record then holds the residual check of x = 0, which follows from the
// right-hand side's own statistics: r = b' exactly). A start that meets the
// tolerance, a non-finite iterate or a zero budget is done at once. The
// record is then written to the ring as check 0.
__global__ void kb_start(const FinArgs* __restrict__ t, double tol, int budget, int warm, ProjState* ring,
                         int ring_size) {
    if (threadIdx.x != 0) return;
    ProjState* ps = t[blockIdx.x].ps;
    const double count = t[blockIdx.x].count;
    ps->zero_start = 0;
    if (warm && !(ps->bnorm2 - ps->b_sum * ps->b_sum / count > 0.0) && !ps->nonfinite_x) {
        ps->zero_start = 1;
        ps->res2 = ps->bnorm2;
        ps->res_inf = ps->initial_inf;
        ps->x_sum = 0.0;
        ps->x_meanf = 0.0f;
        ps->res_sum = ps->b_sum;
    }
    const double num = ps->res2 - ps->res_sum * ps->res_sum / count, den = ps->bnorm2 - ps->b_sum * ps->b_sum / count;
    const double rel = den > 0.0 ? sqrt((num > 0.0 ? num : 0.0) / den) : 0.0;
    ps->start_rel = rel;
    ps->start_inf = ps->res_inf;
    ps->checked_rel = rel;
    ps->sweeps_run = 0;
    ps->chunks = 0;
    ps->done = (rel <= tol || budget <= 0 || ps->nonfinite_x) ? 1 : 0;
    if (ps->done) ps->done_clock = (long long)wall_clock64();
    ring[(ps->chunks % ring_size) * (int)gridDim.x + (int)blockIdx.x] = *ps;
    ps->chunks = 1;
}

// p = 0 for every scene (cold start), or only where the start kernel chose a
// zero start.
__global__ void __launch_bounds__(kBlock) kb_zero(const FillSlot* __restrict__ t, BatchList L, int only_zero_start) {
    for (int w = blockIdx.x; w < L.items; w += gridDim.x) {
        const int s = L.owner[w];
        const FillSlot& f = t[s];
        if (only_zero_start && !f.ps->zero_start) continue;
        const int nb = L.prefix[s + 1] - L.prefix[s];
        for (size_t i = (size_t)(w - L.prefix[s]) * kBlock + threadIdx.x; i < f.count; i += (size_t)nb * kBlock)
            f.p[i] = 0.0f;
    }
}

__global__ void kb_stamp(long long* dst) {
    if (threadIdx.x == 0) *dst = (long long)wall_clock64();
}

// out[s * per + k] = records[s][k] (out: device-visible pinned host memory).
__global__ void kb_gather(ProjState* const* __restrict__ records, int per, ProjState* out) {
    if (threadIdx.x < per) out[blockIdx.x * per + threadIdx.x] = records[blockIdx.x][threadIdx.x];
}

// ------------------------------------------ precision-generic reference step
// vapor's CPU step as one plain kernel per operation in the field precision T
// (CpuPressureSolver3D.cpp:25-60, SOR.cpp:36-49, MacGridOperators3D.cpp:8-67):
// with T = double it is the all-FP64 run the FP32 path is measured against;
// with T = float it is the reference's FP32 step again, by a second route.
// The emitter, forces and advection share the FP32 path's bodies; the
// projection keeps the reference's own order: walls, divergence, b = -div
// h^2/dt, minus its FP64 mean, red-black SOR in place, minus the pressure's
// FP64 mean, gradient, walls. Lanes walk the arrays linearly (grid-stride).

template <class T> __global__ void __launch_bounds__(1024) kw_emit_noise(EmitArgsT<T> a, uint32_t* rng, T* noise) {
    emit_noise_body(a, rng, noise);
}
template <class T>
__global__ void __launch_bounds__(kBlock) kw_emit_apply(EmitArgsT<T> a, const T* __restrict__ noise, T* rho, T* u,
                                                        T* v, T* w) {
    emit_apply_body(a, noise, rho, u, v, w, blockIdx.x);
}
template <class T>
__global__ void __launch_bounds__(kBlock) kw_forces(int n, const T* __restrict__ rho, T* u, T* v, T* w, T buoyancy,
                                                    T dt, T split) {
    forces_body(n, rho, u, v, w, buoyancy, dt, split, blockIdx.x, gridDim.x);
}
template <bool MC, int AXIS, bool HAT, class T> __global__ void __launch_bounds__(kBlock) kw_advect(AdvArgsT<T> a) {
    advect_body<MC, AXIS, HAT>(a, blockIdx.x, gridDim.x);
}

// enforceClosedBoundaries (MacGridOperators3D.cpp:16-28): the three face
// arrays have (n+1) n^2 entries each; lane i zeroes whichever of u[i], v[i],
// w[i] lies on a wall.
template <class T> __global__ void __launch_bounds__(kBlock) kw_walls(int n, T* u, T* v, T* w) {
    const size_t count = (size_t)(n + 1) * n * n, nn = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) {
        const int x = (int)(i % (size_t)(n + 1));
        if (x == 0 || x == n) u[i] = 0;
        const int y = (int)((i / n) % (size_t)(n + 1));
        if (y == 0 || y == n) v[i] = 0;
        const int z = (int)(i / nn);
        if (z == 0 || z == n) w[i] = 0;
    }
}

// divergenceAt / computeDivergence (MacGridOperators3D.cpp:8-13,30-38) on
// faces whose walls are already zero; with b, also b = -div scale
// (CpuPressureSolver3D.cpp:35-42). Partials: sum of b, max |div|, count of
// non-finite divergences.
template <class T>
__global__ void __launch_bounds__(kBlock) kw_divergence(int n, T h, T scale, const T* __restrict__ u,
                                                        const T* __restrict__ v, const T* __restrict__ w, T* b,
                                                        double* partials) {
    __shared__ double lds[3 * 16];
    double acc[3] = {0.0, 0.0, 0.0};
    const size_t cells = (size_t)n * n * n;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < cells; i += (size_t)gridDim.x * kBlock) {
        const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / ((size_t)n * n));
        const T div = (u[idX(n, x + 1, y, z)] - u[idX(n, x, y, z)] + v[idY(n, x, y + 1, z)] - v[idY(n, x, y, z)] +
                       w[idZ(n, x, y, z + 1)] - w[idZ(n, x, y, z)]) /
                      h;
        if (!isfinite(div)) acc[2] += 1.0;
        const double ad = (double)fabs((double)div);
        acc[1] = acc[1] < ad ? ad : acc[1];
        if (b) {
            const T bv = -div * scale;
            b[i] = bv;
            acc[0] += (double)bv;
        }
    }
    const Red ops[3] = {Red::Sum, Red::Max, Red::Sum};
    block_reduce<3>(acc, ops, lds);
    write_partials(partials, acc, 3);
}

// x[i] -= m: the removed right-hand-side mean (CpuPressureSolver3D.cpp:44-45)
// and the pressure's zero-mean gauge (:50-51).
template <class T> __global__ void __launch_bounds__(kBlock) kw_shift(T* x, T m, size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) x[i] -= m;
}

// One colour of a red-black sweep in place (SOR.cpp:36-49): cells with
// (x + y + z) % 2 == parity. Same-colour cells never read each other, so the
// order in which lanes update them does not matter.
//   x += omega (b - sum(offdiag x) - d x) / max(d, floor)
// with the neighbour terms in the reference's order; d = 0 rows are skipped.
template <class T>
__global__ void __launch_bounds__(kBlock) kw_sor_color(int n, T omega, T floor, const T* __restrict__ b, T* x,
                                                       int parity) {
    const int half = (n + 1) / 2;
    const size_t count = (size_t)n * n * half, nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < count; t += (size_t)gridDim.x * kBlock) {
        const int j = (int)(t % half);
        const size_t row = t / half;
        const int y = (int)(row % n), z = (int)(row / n);
        const int xi = 2 * j + ((y + z + parity) & 1);
        if (xi >= n) continue;
        const size_t idx = (size_t)xi + (size_t)n * row;
        const T d = (T)rb_degree(xi, y, z, n);
        if (d == 0) continue;
        T off = 0;
        if (xi > 0) off += (T)-1 * x[idx - 1];
        if (xi + 1 < n) off += (T)-1 * x[idx + 1];
        if (y > 0) off += (T)-1 * x[idx - n];
        if (y + 1 < n) off += (T)-1 * x[idx + n];
        if (z > 0) off += (T)-1 * x[idx - nn];
        if (z + 1 < n) off += (T)-1 * x[idx + nn];
        const T diagonal = maxv(d, floor);
        x[idx] += omega * (b[idx] - off - d * x[idx]) / diagonal;
    }
}

// FP64 residual r = b - A x (StencilLinearSystem.cpp:53-70 term order).
// Partials: sum r, sum r^2, max |r|, sum b, sum b^2, non-finite x count.
template <class T>
__global__ void __launch_bounds__(kBlock) kw_residual(int n, const T* __restrict__ b, const T* __restrict__ x,
                                                      double* partials) {
    __shared__ double lds[6 * 16];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t cells = (size_t)n * n * n, nn = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < cells; i += (size_t)gridDim.x * kBlock) {
        const int xi = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / nn);
        if (!isfinite(x[i])) acc[5] += 1.0;
        double ax = (double)rb_degree(xi, y, z, n) * (double)x[i];
        if (xi > 0) ax += -1.0 * (double)x[i - 1];
        if (xi + 1 < n) ax += -1.0 * (double)x[i + 1];
        if (y > 0) ax += -1.0 * (double)x[i - n];
        if (y + 1 < n) ax += -1.0 * (double)x[i + n];
        if (z > 0) ax += -1.0 * (double)x[i - nn];
        if (z + 1 < n) ax += -1.0 * (double)x[i + nn];
        const double bi = (double)b[i], r = bi - ax, ar = fabs(r);
        acc[0] += r;
        acc[1] += r * r;
        acc[2] = acc[2] < ar ? ar : acc[2];
        acc[3] += bi;
        acc[4] += bi * bi;
    }
    const Red ops[6] = {Red::Sum, Red::Sum, Red::Max, Red::Sum, Red::Sum, Red::Sum};
    block_reduce<6>(acc, ops, lds);
    write_partials(partials, acc, 6);
}

// FP64 sum of an array (the pressure's mean).
template <class T> __global__ void __launch_bounds__(kBlock) kw_sum(const T* __restrict__ x, size_t count,
                                                                   double* partials) {
    __shared__ double lds[16];
    double acc[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        acc[0] += (double)x[i];
    const Red ops[1] = {Red::Sum};
    block_reduce<1>(acc, ops, lds);
    write_partials(partials, acc, 1);
}

// applyPressureGradient then enforceClosedBoundaries (MacGridOperators3D.cpp:
// 51-67, 16-28): interior faces u -= dt (p_i - p_{i-1}) / h, wall faces 0.
template <class T>
__global__ void __launch_bounds__(kBlock) kw_gradient(int n, T hInv, T dt, const T* __restrict__ p, T* u, T* v,
                                                      T* w) {
    const size_t count = (size_t)(n + 1) * n * n, nn = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) {
        {
            const int x = (int)(i % (size_t)(n + 1));
            const size_t r = i / (size_t)(n + 1);  // y + n z
            if (x == 0 || x == n) u[i] = 0;
            else u[i] -= dt * (p[(size_t)x + n * r] - p[(size_t)x - 1 + n * r]) * hInv;
        }
        {
            const int x = (int)(i % n), y = (int)((i / n) % (size_t)(n + 1)), z = (int)(i / ((size_t)n * (n + 1)));
            if (y == 0 || y == n) v[i] = 0;
            else v[i] -= dt * (p[idC(n, x, y, z)] - p[idC(n, x, y - 1, z)]) * hInv;
        }
        {
            const int z = (int)(i / nn);
            if (z == 0 || z == n) w[i] = 0;
            else w[i] -= dt * (p[i] - p[i - nn]) * hInv;
        }
    }
}

// Error of a field against a reference field. Partials: max |a - r|,
// sum (a - r)^2, max |r|, sum r^2.
template <class A, class B>
__global__ void __launch_bounds__(kBlock) kw_compare(const A* __restrict__ a, const B* __restrict__ ref, size_t count,
                                                     double* partials) {
    __shared__ double lds[4 * 16];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) {
        const double r = (double)ref[i], e = (double)a[i] - r, ae = fabs(e), ar = fabs(r);
        acc[0] = acc[0] < ae ? ae : acc[0];
        acc[1] += e * e;
        acc[2] = acc[2] < ar ? ar : acc[2];
        acc[3] += r * r;
    }
    const Red ops[4] = {Red::Max, Red::Sum, Red::Max, Red::Sum};
    block_reduce<4>(acc, ops, lds);
    write_partials(partials, acc, 4);
}

template <class T> __global__ void __launch_bounds__(kBlock) kw_convert(T* out, const float* __restrict__ in,
                                                                       size_t count) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        out[i] = (T)in[i];
}

}  // namespace

// ---------------------------------------------------------------- launchers

size_t partials_capacity(int n) {
    return (size_t)std::max(kMaxReduceBlocks, rb_blocks(n)) * kPartialStride;
}

hipError_t launch_add_f32(hipStream_t s, const float* a, const float* b, float* out, size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_add_f32, dim3(linear_blocks(count)), dim3(kBlock), s, a, b, out, count);
    return hipGetLastError();
}
hipError_t launch_energy_dots(hipStream_t s, int n, const float* r, const float* e, double* partials, int* nblocks) {
    const Launch l = cell_launch(n);
    vlaunch(k_energy_dots, l.grid, l.block, s, n, r, e, partials);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}
hipError_t launch_line_search(hipStream_t s, const double* partials, int nblocks, ProjState* ps) {
    vlaunch(k_line_search, dim3(1), dim3(kBlock), s, partials, nblocks, ps);
    return hipGetLastError();
}
hipError_t launch_axpy_alpha(hipStream_t s, const float* x, const ProjState* ps, const float* e, float* out,
                             size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_axpy_alpha, dim3(linear_blocks(count)), dim3(kBlock), s, x, ps, e, out, count);
    return hipGetLastError();
}
hipError_t launch_judge(hipStream_t s, ProjState* plain, const ProjState* learned, double count) {
    vlaunch(k_judge, dim3(1), dim3(1), s, plain, learned, count);
    return hipGetLastError();
}
hipError_t launch_copy_if_learned(hipStream_t s, const float* src, float* dst, const ProjState* ps, size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_copy_if_learned, dim3(linear_blocks(count)), dim3(kBlock), s, src, dst, ps, count);
    return hipGetLastError();
}
hipError_t launch_add_scaled_f32(hipStream_t s, const float* a, float scale, const float* b, float* out, size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_add_scaled_f32, dim3(linear_blocks(count)), dim3(kBlock), s, a, scale, b, out, count);
    return hipGetLastError();
}
hipError_t launch_scale_f32(hipStream_t s, const float* a, float scale, float* out, size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_scale_f32, dim3(linear_blocks(count)), dim3(kBlock), s, a, scale, out, count);
    return hipGetLastError();
}
hipError_t launch_correction(hipStream_t s, const float* x, const float* prev, const ProjState* ps, float scale,
                             float* out, size_t count) {
    if (!count) return hipSuccess;
    vlaunch(k_correction, dim3(linear_blocks(count)), dim3(kBlock), s, x, prev, ps, scale, out, count);
    return hipGetLastError();
}
hipError_t launch_add_noise(hipStream_t s, float* field, size_t count, float sigma, uint64_t seed, uint32_t stream) {
    if (!count) return hipSuccess;
    vlaunch(k_add_noise, dim3(linear_blocks(count)), dim3(kBlock), s, field, (uint32_t)count, sigma, seed, stream);
    return hipGetLastError();
}
hipError_t launch_fill_f32(hipStream_t s, float* p, size_t count, float value) {
    if (!count) return hipSuccess;
    vlaunch(k_fill_f32, dim3(linear_blocks(count)), dim3(kBlock), s, p, count, value);
    return hipGetLastError();
}
hipError_t launch_fill_f64(hipStream_t s, double* p, size_t count, double value) {
    if (!count) return hipSuccess;
    vlaunch(k_fill_f64, dim3(linear_blocks(count)), dim3(kBlock), s, p, count, value);
    return hipGetLastError();
}
hipError_t launch_set_u32(hipStream_t s, uint32_t* p, uint32_t value) {
    vlaunch(k_set_u32, dim3(1), dim3(1), s, p, value);
    return hipGetLastError();
}

hipError_t launch_emit(hipStream_t s, const EmitArgs& a, uint32_t* rng, float* noise, float* rho, float* u,
                       float* v, float* w) {
    const int count = a.ext[0] * a.ext[1] * a.ext[2];
    if (count <= 0) return hipSuccess;
    vlaunch(k_emit_noise, dim3(1), dim3(1024), s, a, rng, noise);
    const int points = (a.ext[0] + 1) * (a.ext[1] + 1) * (a.ext[2] + 1);
    vlaunch(k_emit_apply, dim3((points + kBlock - 1) / kBlock), dim3(kBlock), s, a, noise, rho, u, v,
                       w);
    return hipGetLastError();
}

hipError_t launch_forces(hipStream_t s, int n, const float* rho, float* u, float* v, float* w, float buoyancy,
                         float dt, float split) {
    const size_t faces = (size_t)(n + 1) * n * n;
    vlaunch(k_forces, dim3(linear_blocks(faces)), dim3(kBlock), s, n, rho, u, v, w, buoyancy, dt,
                       split);
    return hipGetLastError();
}

static dim3 advect_grid(const AdvArgs& a) {
    const size_t count = (size_t)(a.n + 1) * a.n * a.n;
    const size_t blocks = std::min<size_t>((count + kBlock - 1) / kBlock, 65535);
    return dim3((unsigned)blocks, 1, 1);
}

// Buffer loads take 32-bit byte offsets: every field must fit in 2^31 bytes.
static bool advect_buffers(int n) { return (size_t)(n + 1) * n * n * 4 < ((size_t)1 << 31); }

template <bool MC, bool HAT, bool P, bool B> static void launch_field_pb(hipStream_t s, const AdvArgs& a, int field) {
    const dim3 grid = advect_grid(a);
    switch (field) {
    case 0:
        if (MC) vlaunch((k_advect_mc<0, HAT, P, B>), grid, dim3(kBlock), s, a);
        else vlaunch((k_advect_sl<0, HAT, P, B>), grid, dim3(kBlock), s, a);
        break;
    case 1:
        if (MC) vlaunch((k_advect_mc<1, HAT, P, B>), grid, dim3(kBlock), s, a);
        else vlaunch((k_advect_sl<1, HAT, P, B>), grid, dim3(kBlock), s, a);
        break;
    case 2:
        if (MC) vlaunch((k_advect_mc<2, HAT, P, B>), grid, dim3(kBlock), s, a);
        else vlaunch((k_advect_sl<2, HAT, P, B>), grid, dim3(kBlock), s, a);
        break;
    default:
        if (MC) vlaunch((k_advect_mc<-1, false, P, B>), grid, dim3(kBlock), s, a);
        else vlaunch((k_advect_sl<-1, false, P, B>), grid, dim3(kBlock), s, a);
        break;
    }
}
template <bool MC, bool HAT> static void launch_field(hipStream_t s, const AdvArgs& a, int field) {
    const bool p = pow2_spacing(a.h), b = advect_buffers(a.n);
    if (p && b) launch_field_pb<MC, HAT, true, true>(s, a, field);
    else if (p) launch_field_pb<MC, HAT, true, false>(s, a, field);
    else if (b) launch_field_pb<MC, HAT, false, true>(s, a, field);
    else launch_field_pb<MC, HAT, false, false>(s, a, field);
}

hipError_t launch_advect_sl(hipStream_t s, const AdvArgs& a, int fields, bool hat) {
    for (int k = 0; k < fields; ++k) {
        if (hat) launch_field<false, true>(s, a, a.first_field + k);
        else launch_field<false, false>(s, a, a.first_field + k);
    }
    return hipGetLastError();
}
hipError_t launch_advect_mc(hipStream_t s, const AdvArgs& a, int fields, bool hat) {
    for (int k = 0; k < fields; ++k) {
        if (hat) launch_field<true, true>(s, a, a.first_field + k);
        else launch_field<true, false>(s, a, a.first_field + k);
    }
    return hipGetLastError();
}

hipError_t launch_divergence_rhs(hipStream_t s, int n, float h, float scale, const float* u, const float* v,
                                 const float* w, float* b, double* partials, int* nblocks) {
    const Launch l = cell_launch(n);
    vlaunch(k_divergence_rhs, l.grid, l.block, s, n, h, scale, u, v, w, b, partials);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}

hipError_t launch_sor_rb(hipStream_t s, int n, float omega, float floor, const float* b, const float* xin,
                         float* xout, const ProjState* ps, const int* done, int mode, double* partials, int* nblocks) {
    const int blocks = rb_blocks(n);
    // max(d, floor) = 6 for every degree d <= 6: one divisor, one reciprocal
    // float4 staging needs 16-byte aligned rows: n % 4 == 0 and aligned bases
    const bool aligned = ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(xin) |
                           reinterpret_cast<uintptr_t>(xout)) & 15) == 0;
    const bool fast = floor == 6.0f, vec = n % 4 == 0 && aligned;
    const RbArgs a{omega, floor, 6.0f, 1.0f / 6.0f};
    const RbGeom geo = rb_geom(n);
    const dim3 g(blocks), blk(kRbThreads);
    if (vec && rb_reg_fits(n) && (mode == 0 || mode == 2)) {
        if (mode == 2)
            fast ? vlaunch(k_sor_rb_reg<2, true>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials)
                 : vlaunch(k_sor_rb_reg<2, false>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials);
        else
            fast ? vlaunch(k_sor_rb_reg<0, true>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials)
                 : vlaunch(k_sor_rb_reg<0, false>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials);
        if (nblocks) *nblocks = blocks;
        return hipGetLastError();
    }
#define VGX_RB(M)                                                                                              \
    (fast ? (vec ? vlaunch(k_sor_rb<M, true, true>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials)      \
                 : vlaunch(k_sor_rb<M, true, false>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials))    \
          : (vec ? vlaunch(k_sor_rb<M, false, true>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials)     \
                 : vlaunch(k_sor_rb<M, false, false>, g, blk, s, n, geo, a, b, xin, xout, ps, done, partials)))
    switch (mode) {
    case 1: VGX_RB(1); break;
    case 2: VGX_RB(2); break;
    case 3: VGX_RB(3); break;
    default: VGX_RB(0); break;
    }
#undef VGX_RB
    if (nblocks) *nblocks = blocks;
    return hipGetLastError();
}

hipError_t launch_check(hipStream_t s, const double* partials, int nblocks, int n, ProjState* ps, double tol,
                        int chunk, ProjState* ring, int ring_size) {
    vlaunch(k_check, dim3(1), dim3(1024), s, partials, nblocks, (double)n * n * n, ps, tol, chunk, ring, ring_size);
    return hipGetLastError();
}

hipError_t launch_reset_status(hipStream_t s, ProjState* ps) {
    vlaunch(k_reset_status, dim3(1), dim3(1), s, ps);
    return hipGetLastError();
}

hipError_t launch_residual(hipStream_t s, int n, const float* b, const float* x, const ProjState* ps,
                           double* partials, int* nblocks) {
    const Launch l = cell_launch(n);
    vlaunch(k_residual, l.grid, l.block, s, n, b, x, ps, partials);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}
hipError_t launch_judge_pair(hipStream_t s, int n, const float* b, const float* x_plain, ProjState* plain,
                             double* partials_plain, const float* x_learned, ProjState* learned,
                             double* partials_learned) {
    if (partials_plain == partials_learned) return hipErrorInvalidValue;
    const Launch l = cell_launch(n);
    vlaunch(k_residual_pair, l.grid, l.block, s, n, b, x_plain, plain, partials_plain, x_learned, learned,
            partials_learned);
    vlaunch(k_judge_pair, dim3(1), dim3(1024), s, partials_plain, partials_learned, blocks_of(l),
            (double)n * n * n, plain, learned);
    return hipGetLastError();
}

hipError_t launch_residual64(hipStream_t s, int n, const float* b, const double* x, const ProjState* ps, float* r32,
                             double* partials, int* nblocks) {
    const Launch l = cell_launch(n);
    vlaunch(k_residual_field<double>, l.grid, l.block, s, n, b, x, ps, r32, partials, (float*)nullptr);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}
hipError_t launch_residual_field(hipStream_t s, int n, const float* b, const float* x, const ProjState* ps, float* r32,
                                 double* partials, int* nblocks, float* xcopy) {
    if (xcopy == x) return hipErrorInvalidValue;
    const Launch l = cell_launch(n);
    vlaunch(k_residual_field<float>, l.grid, l.block, s, n, b, x, ps, r32, partials, xcopy);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}

hipError_t launch_axpy_f64(hipStream_t s, double* x, const float* e, size_t count) {
    vlaunch(k_axpy_f64, dim3(linear_blocks(count)), dim3(kBlock), s, x, e, count);
    return hipGetLastError();
}
hipError_t launch_widen(hipStream_t s, double* x, const float* e, size_t count) {
    vlaunch(k_widen, dim3(linear_blocks(count)), dim3(kBlock), s, x, e, count);
    return hipGetLastError();
}
hipError_t launch_narrow(hipStream_t s, float* x, const double* e, size_t count) {
    vlaunch(k_narrow, dim3(linear_blocks(count)), dim3(kBlock), s, x, e, count);
    return hipGetLastError();
}

hipError_t launch_gradient(hipStream_t s, int n, float h, float dt, const float* uin, const float* vin,
                           const float* win, const float* p, float* p_out, const float* b, float* uout, float* vout,
                           float* wout, const ProjState* ps, double* partials, int* nblocks) {
    const Launch l = cell_launch(n);
    vlaunch(k_gradient, l.grid, l.block, s, n, h, dt, uin, vin, win, p, p_out, b, uout, vout, wout, ps,
                       partials);
    *nblocks = blocks_of(l);
    return hipGetLastError();
}

hipError_t launch_diagnostics(hipStream_t s, int n, float h, const float* u, const float* v, const float* w,
                              double* partials, int* nblocks) {
    const size_t faces = (size_t)(n + 1) * n * n;
    const int blocks = linear_blocks(faces);
    vlaunch(k_diagnostics, dim3(blocks), dim3(kBlock), s, n, h, u, v, w, partials);
    *nblocks = blocks;
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t s, Finalize mode, const double* partials, int nblocks, int n, ProjState* ps,
                           DiagState* ds) {
    const double count = (double)n * n * n;
    switch (mode) {
    case Finalize::Rhs: return finalize<MODE_RHS>(s, partials, nblocks, count, ps, ds);
    case Finalize::XSum: return finalize<MODE_XSUM>(s, partials, nblocks, count, ps, ds);
    case Finalize::Residual: return finalize<MODE_RESIDUAL>(s, partials, nblocks, count, ps, ds);
    case Finalize::Final: return finalize<MODE_FINAL>(s, partials, nblocks, count, ps, ds);
    default: return finalize<MODE_DIAG>(s, partials, nblocks, count, ps, ds);
    }
}

// ---------------------------------------------------------- batch launchers

static int grid_of(const BatchList& L, int grid) { return std::max(1, grid > 0 ? std::min(grid, L.items) : L.items); }

hipError_t batch_emit(hipStream_t s, const EmitSlot* t, int scenes, BatchList apply, int grid) {
    vlaunch(kb_emit_noise, dim3(scenes), dim3(1024), s, t);
    if (apply.items > 0) vlaunch(kb_emit_apply, dim3(grid_of(apply, grid)), dim3(kBlock), s, t, apply);
    return hipGetLastError();
}

hipError_t batch_forces(hipStream_t s, const ForceArgs* t, BatchList L, int grid) {
    vlaunch(kb_forces, dim3(grid_of(L, grid)), dim3(kBlock), s, t, L);
    return hipGetLastError();
}

template <bool MC, bool HAT> static void batch_field(hipStream_t s, int field, const AdvArgs* t, BatchList L, dim3 g) {
    switch (field) {
    case 0: vlaunch((kb_advect<MC, 0, HAT>), g, dim3(kBlock), s, t, L); break;
    case 1: vlaunch((kb_advect<MC, 1, HAT>), g, dim3(kBlock), s, t, L); break;
    case 2: vlaunch((kb_advect<MC, 2, HAT>), g, dim3(kBlock), s, t, L); break;
    default: vlaunch((kb_advect<MC, -1, false>), g, dim3(kBlock), s, t, L); break;
    }
}

hipError_t batch_advect(hipStream_t s, bool mc, int field, bool hat, const AdvArgs* t, BatchList L, int grid) {
    const dim3 g(grid_of(L, grid));
    if (mc) {
        if (hat) batch_field<true, true>(s, field, t, L, g);
        else batch_field<true, false>(s, field, t, L, g);
    } else {
        if (hat) batch_field<false, true>(s, field, t, L, g);
        else batch_field<false, false>(s, field, t, L, g);
    }
    return hipGetLastError();
}

hipError_t batch_cell(hipStream_t s, CellOp op, const CellArgs* t, BatchList L, int grid) {
    const dim3 g(grid_of(L, grid)), blk(64, 4, 1);
    switch (op) {
    case CellOp::Divergence: vlaunch(kb_cell<0>, g, blk, s, t, L); break;
    case CellOp::Residual: vlaunch(kb_cell<1>, g, blk, s, t, L); break;
    default: vlaunch(kb_cell<2>, g, blk, s, t, L); break;
    }
    return hipGetLastError();
}

hipError_t batch_sor(hipStream_t s, float omega, float floor, bool vec, bool reg, int mode, const ActiveList* act,
                     const RbSlot* t, BatchList L, int grid) {
    const bool fast = floor == 6.0f;
    const RbArgs a{omega, floor, 6.0f, 1.0f / 6.0f};
    const dim3 g(grid_of(L, grid)), blk(kRbThreads);
    const ActiveList ud = act ? *act : ActiveList{nullptr, nullptr, nullptr};
    if (vec && reg && (mode == 0 || mode == 2)) {
        if (mode == 2)
            fast ? vlaunch(kb_sor_reg<2, true>, g, blk, s, a, t, L, ud) : vlaunch(kb_sor_reg<2, false>, g, blk, s, a, t, L, ud);
        else
            fast ? vlaunch(kb_sor_reg<0, true>, g, blk, s, a, t, L, ud) : vlaunch(kb_sor_reg<0, false>, g, blk, s, a, t, L, ud);
        return hipGetLastError();
    }
#define VGX_RBB(M)                                                                                       \
    (fast ? (vec ? vlaunch(kb_sor<M, true, true>, g, blk, s, a, t, L, ud)                                 \
                 : vlaunch(kb_sor<M, true, false>, g, blk, s, a, t, L, ud))                               \
          : (vec ? vlaunch(kb_sor<M, false, true>, g, blk, s, a, t, L, ud)                                \
                 : vlaunch(kb_sor<M, false, false>, g, blk, s, a, t, L, ud)))
    switch (mode) {
    case 1: VGX_RBB(1); break;
    case 2: VGX_RBB(2); break;
    case 3: VGX_RBB(3); break;
    default: VGX_RBB(0); break;
    }
#undef VGX_RBB
    return hipGetLastError();
}

hipError_t batch_finalize(hipStream_t s, Finalize mode, const FinArgs* t, int scenes) {
    const dim3 g(scenes), blk(1024);
    switch (mode) {
    case Finalize::Rhs: vlaunch(kb_finalize<MODE_RHS>, g, blk, s, t); break;
    case Finalize::XSum: vlaunch(kb_finalize<MODE_XSUM>, g, blk, s, t); break;
    case Finalize::Residual: vlaunch(kb_finalize<MODE_RESIDUAL>, g, blk, s, t); break;
    case Finalize::Final: vlaunch(kb_finalize<MODE_FINAL>, g, blk, s, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t batch_check(hipStream_t s, const FinArgs* t, int scenes, double tol, int chunk, ProjState* ring,
                       int ring_size) {
    vlaunch(kb_check, dim3(scenes), dim3(1024), s, t, tol, chunk, ring, ring_size);
    return hipGetLastError();
}

hipError_t batch_active(hipStream_t s, const FinArgs* t, BatchList L, ActiveList act) {
    vlaunch(kb_active, dim3(1), dim3(1024), s, t, L, act);
    return hipGetLastError();
}

hipError_t batch_start(hipStream_t s, const FinArgs* t, int scenes, double tol, int budget, bool warm, ProjState* ring,
                       int ring_size) {
    vlaunch(kb_start, dim3(scenes), dim3(64), s, t, tol, budget, warm ? 1 : 0, ring, ring_size);
    return hipGetLastError();
}

hipError_t batch_zero(hipStream_t s, const FillSlot* t, BatchList L, int grid, bool only_zero_start) {
    vlaunch(kb_zero, dim3(grid_of(L, grid)), dim3(kBlock), s, t, L, only_zero_start ? 1 : 0);
    return hipGetLastError();
}

hipError_t batch_stamp(hipStream_t s, long long* dst) {
    vlaunch(kb_stamp, dim3(1), dim3(64), s, dst);
    return hipGetLastError();
}

hipError_t batch_gather(hipStream_t s, ProjState* const* records, int scenes, int per, ProjState* out) {
    vlaunch(kb_gather, dim3(scenes), dim3(64), s, records, per, out);
    return hipGetLastError();
}

// ---- precision-generic reference step ----
namespace wide {

// Grid of a grid-stride element-wise launch over `count` entries.
static int wide_blocks(size_t count) { return linear_blocks(count); }

template <class T>
hipError_t emit(hipStream_t s, const EmitArgsT<T>& a, uint32_t* rng, T* noise, T* rho, T* u, T* v, T* w) {
    vlaunch(kw_emit_noise<T>, dim3(1), dim3(1024), s, a, rng, noise);
    const int count = (a.ext[0] + 1) * (a.ext[1] + 1) * (a.ext[2] + 1);
    vlaunch(kw_emit_apply<T>, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), s, a, (const T*)noise, rho, u, v, w);
    return hipGetLastError();
}
template <class T>
hipError_t forces(hipStream_t s, int n, const T* rho, T* u, T* v, T* w, T buoyancy, T dt, T split) {
    vlaunch(kw_forces<T>, dim3(wide_blocks((size_t)(n + 1) * n * n)), dim3(kBlock), s, n, rho, u, v, w, buoyancy,
            dt, split);
    return hipGetLastError();
}
template <bool MC, bool HAT, class T> static void advect_field(hipStream_t s, const AdvArgsT<T>& a, int field) {
    const size_t count = field == 3 ? (size_t)a.n * a.n * a.n : (size_t)(a.n + 1) * a.n * a.n;
    const dim3 g(wide_blocks(count));
    switch (field) {
    case 0: vlaunch(kw_advect<MC, 0, HAT, T>, g, dim3(kBlock), s, a); break;
    case 1: vlaunch(kw_advect<MC, 1, HAT, T>, g, dim3(kBlock), s, a); break;
    case 2: vlaunch(kw_advect<MC, 2, HAT, T>, g, dim3(kBlock), s, a); break;
    default: vlaunch(kw_advect<MC, -1, false, T>, g, dim3(kBlock), s, a); break;
    }
}
template <class T> hipError_t advect(hipStream_t s, const AdvArgsT<T>& a, int field, bool mc, bool hat) {
    if (mc) hat ? advect_field<true, true>(s, a, field) : advect_field<true, false>(s, a, field);
    else hat ? advect_field<false, true>(s, a, field) : advect_field<false, false>(s, a, field);
    return hipGetLastError();
}
template <class T> hipError_t walls(hipStream_t s, int n, T* u, T* v, T* w) {
    vlaunch(kw_walls<T>, dim3(wide_blocks((size_t)(n + 1) * n * n)), dim3(kBlock), s, n, u, v, w);
    return hipGetLastError();
}
template <class T>
hipError_t divergence(hipStream_t s, int n, T h, T scale, const T* u, const T* v, const T* w, T* b, double* partials,
                      int* nblocks) {
    *nblocks = wide_blocks((size_t)n * n * n);
    vlaunch(kw_divergence<T>, dim3(*nblocks), dim3(kBlock), s, n, h, scale, u, v, w, b, partials);
    return hipGetLastError();
}
template <class T> hipError_t shift(hipStream_t s, T* x, T m, size_t count) {
    vlaunch(kw_shift<T>, dim3(wide_blocks(count)), dim3(kBlock), s, x, m, count);
    return hipGetLastError();
}
template <class T> hipError_t sor_color(hipStream_t s, int n, T omega, T floor, const T* b, T* x, int parity) {
    vlaunch(kw_sor_color<T>, dim3(wide_blocks((size_t)n * n * ((n + 1) / 2))), dim3(kBlock), s, n, omega, floor, b,
            x, parity);
    return hipGetLastError();
}
template <class T> hipError_t residual(hipStream_t s, int n, const T* b, const T* x, double* partials, int* nblocks) {
    *nblocks = wide_blocks((size_t)n * n * n);
    vlaunch(kw_residual<T>, dim3(*nblocks), dim3(kBlock), s, n, b, x, partials);
    return hipGetLastError();
}
template <class T> hipError_t sum(hipStream_t s, const T* x, size_t count, double* partials, int* nblocks) {
    *nblocks = wide_blocks(count);
    vlaunch(kw_sum<T>, dim3(*nblocks), dim3(kBlock), s, x, count, partials);
    return hipGetLastError();
}
template <class T> hipError_t gradient(hipStream_t s, int n, T hInv, T dt, const T* p, T* u, T* v, T* w) {
    vlaunch(kw_gradient<T>, dim3(wide_blocks((size_t)(n + 1) * n * n)), dim3(kBlock), s, n, hInv, dt, p, u, v, w);
    return hipGetLastError();
}
template <class A, class B>
hipError_t compare(hipStream_t s, const A* a, const B* ref, size_t count, double* partials, int* nblocks) {
    *nblocks = wide_blocks(count);
    vlaunch(kw_compare<A, B>, dim3(*nblocks), dim3(kBlock), s, a, ref, count, partials);
    return hipGetLastError();
}
template <class T> hipError_t convert(hipStream_t s, T* out, const float* in, size_t count) {
    vlaunch(kw_convert<T>, dim3(wide_blocks(count)), dim3(kBlock), s, out, in, count);
    return hipGetLastError();
}

#define VGX_WIDE_INSTANTIATE(T)                                                                                    \
    template hipError_t emit<T>(hipStream_t, const EmitArgsT<T>&, uint32_t*, T*, T*, T*, T*, T*);                \
    template hipError_t forces<T>(hipStream_t, int, const T*, T*, T*, T*, T, T, T);                              \
    template hipError_t advect<T>(hipStream_t, const AdvArgsT<T>&, int, bool, bool);                             \
    template hipError_t walls<T>(hipStream_t, int, T*, T*, T*);                                                  \
    template hipError_t divergence<T>(hipStream_t, int, T, T, const T*, const T*, const T*, T*, double*, int*);  \
    template hipError_t shift<T>(hipStream_t, T*, T, size_t);                                                    \
    template hipError_t sor_color<T>(hipStream_t, int, T, T, const T*, T*, int);                                 \
    template hipError_t residual<T>(hipStream_t, int, const T*, const T*, double*, int*);                        \
    template hipError_t sum<T>(hipStream_t, const T*, size_t, double*, int*);                                    \
    template hipError_t gradient<T>(hipStream_t, int, T, T, const T*, T*, T*, T*);                               \
    template hipError_t compare<float, T>(hipStream_t, const float*, const T*, size_t, double*, int*);           \
    template hipError_t convert<T>(hipStream_t, T*, const float*, size_t);
VGX_WIDE_INSTANTIATE(float)
VGX_WIDE_INSTANTIATE(double)
#undef VGX_WIDE_INSTANTIATE

}  // namespace wide

// Per-scene work-list sizes of the batched families.
int batch_items(BatchFamily f, int n, int per_thread) {
    const size_t cells = (size_t)n * n * n, faces = (size_t)(n + 1) * n * n;
    const size_t pt = (size_t)std::max(1, per_thread);
    switch (f) {
    case BatchFamily::Elementwise: return (int)std::max<size_t>(1, (faces + kBlock * pt - 1) / (kBlock * pt));
    case BatchFamily::Cells: return (int)std::max<size_t>(1, (cells + kBlock * pt - 1) / (kBlock * pt));
    case BatchFamily::Stencil: return blocks_of(cell_launch(n));
    default: return rb_blocks(n);
    }
}

}  // namespace vgx
