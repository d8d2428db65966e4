// vgx2_kernels.hip - gfx950 kernels of vapor's MAC 2D solvers.
//
// MacGridFluidSolver2D (src/solvers/mac2d/MacGridFluidSolver2D.cpp) and the
// historical ReflectionMacFluidSolver2D (ReflectionMacFluidSolver2D.hpp),
// device-resident. The grids are small (vapor's defaults are 128^2), so the
// work splits two ways:
//   * gathers (advection, forces, emitter) run one lane per sample over the
//     whole GPU;
//   * each pressure projection is ONE workgroup of 1024 lanes that runs the
//     whole solve: divergence, the FP64 mean sums in the reference's order,
//     every red-black sweep with its residual checks, the gradient and the
//     final divergence. Each lane keeps a T x T tile of cells in registers
//     and the tile's edge cells in LDS, where the neighbouring lanes read
//     them; the iteration count is data-dependent and no host round trip or
//     relaunch happens between sweeps. Grids larger than one workgroup's LDS
//     run the same schedule with the pressure in global memory (L2-resident).
//   * the historical MacCormack limiter updates the smoke in place in scan
//     order (each cell reads the already-limited values of its earlier
//     neighbours); it runs as a wavefront over the anti-diagonals x + 2y = t,
//     on which every cell's earlier neighbours are complete.
#include "vgx2_device.hpp"
#include "vgx2_kernels.hpp"

#include <type_traits>

namespace vgx2 {

namespace {

constexpr int kThreads = 256;
constexpr int kProjThreads = 1024;
constexpr int kProjLdsBytes = 160 * 1024 - 4096;  // dynamic LDS of the single-workgroup solves

inline unsigned blocks_for(size_t count, int threads = kThreads) {
    return (unsigned)((count + (size_t)threads - 1) / (size_t)threads);
}

template <class T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = smax(v, (T)__shfl_xor(v, o));
    return v;
}

// Maximum over the block of non-NaN values, returned to every lane.
template <class T> __device__ T block_max(T v, T* red) {
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    T m = red[0];
    const int waves = (int)(blockDim.x >> 6);
    for (int w = 1; w < waves; ++w) m = smax(m, red[w]);
    return m;
}

// Lane 0's part of the reference's std::accumulate(begin, end, 0.0): adds
// stage[0, len16) to s one by one in index order. The next 16 values are read
// from LDS before the current 16 are added, so the read latency hides behind
// the chain of dependent adds. stage holds zeros past the data up to len16, a
// multiple of 16: the running sum starts at +0 and is never -0, so adding a
// zero of either sign leaves it unchanged.
template <class V> __device__ double lane_sum16(const V* stage, int len16, double s) {
    V cur[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cur[k] = stage[k];
    for (int base = 0; base < len16; base += 16) {
        V next[16];
        const int ahead = base + 16 < len16 ? base + 16 : base;
#pragma unroll
        for (int k = 0; k < 16; ++k) next[k] = stage[ahead + k];
#pragma unroll
        for (int k = 0; k < 16; ++k) s += (double)cur[k];
#pragma unroll
        for (int k = 0; k < 16; ++k) cur[k] = next[k];
    }
    return s;
}

// std::accumulate over values[0, count) in index order, returned to every
// lane. With `staged` the values are already in `stage` followed by zeros up
// to a multiple of 16; otherwise they are staged through `stage` (room for
// `chunk` values, a multiple of 16) one chunk at a time.
template <class V>
__device__ double sequential_sum(const V* values, int count, V* stage, int chunk, bool staged, double* red) {
    double s = 0.0;
    if (staged) {
        __syncthreads();
        if (threadIdx.x == 0) s = lane_sum16(stage, (count + 15) & ~15, s);
    } else {
        for (int base = 0; base < count; base += chunk) {
            const int len = min(chunk, count - base), len16 = (len + 15) & ~15;
            __syncthreads();
            for (int k = threadIdx.x; k < len16; k += blockDim.x) stage[k] = k < len ? values[base + k] : (V)0;
            __syncthreads();
            if (threadIdx.x == 0) s = lane_sum16(stage, len16, s);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) red[0] = s;
    __syncthreads();
    s = red[0];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double div_by_count(double a, int d) {
    // a / d for the neighbour count d: a power of two divides exactly as the
    // multiply by its reciprocal does.
    if (d == 4) return a * 0.25;
    if (d == 2) return a * 0.5;
    if (d == 1) return a;
    return a / (double)d;
}
__device__ __forceinline__ float div_by_count(float a, int d) {
    if (d == 4) return a * 0.25f;
    if (d == 2) return a * 0.5f;
    if (d == 1) return a;
    return a / (float)d;
}

// ---------------------------------------------------------------------------
// MacGridFluidSolver2D: advection (SemiLagrangian2D::advect).

template <int FIELD>
__global__ void k2_advect(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ src,
                          float* __restrict__ out, int n, float dt_over_h, Status2* st) {
    if (st->error) return;
    const int F = n * (n + 1);
    const int w = n + (FIELD == kVelocityX), h = n + (FIELD == kVelocityY);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= F) return;
    uint32_t bits = 0;
    if (!(finitef(u[i]) && finitef(v[i]))) bits |= kErrAdvectionInput;
    if (i < w * h) {
        if (!finitef(src[i])) bits |= kErrAdvectionInput;
        bool ok = true;
        const float value = sl2_value<FIELD>(u, v, src, n, dt_over_h, i % w, i / w, ok);
        if (!ok) bits |= kErrBacktrace;
        out[i] = value;
    }
    if (bits) atomicOr(&st->error, bits);
}

// Both components of advectVelocity (MacGridFluidSolver2D.cpp:73-79) in one
// launch: lanes [0, F) the x faces from su, lanes [F, 2F) the y faces from sv,
// both traced through the same flow (u, v). The y pass's error bits are
// shifted by 8, so the host can report the x pass's failure first, as the
// reference does.
__global__ void k2_advect_velocity(const float* __restrict__ u, const float* __restrict__ v,
                                   const float* __restrict__ su, const float* __restrict__ sv,
                                   float* __restrict__ out_u, float* __restrict__ out_v, int n, float dt_over_h,
                                   Status2* st) {
    if (st->error) return;
    const int F = n * (n + 1);
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= 2 * F) return;
    uint32_t bits = 0;
    bool ok = true;
    if (g < F) {
        const int i = g;
        if (!(finitef(u[i]) && finitef(v[i]) && finitef(su[i]))) bits |= kErrAdvectionInput;
        out_u[i] = sl2_value<kVelocityX>(u, v, su, n, dt_over_h, i % (n + 1), i / (n + 1), ok);
        if (!ok) bits |= kErrBacktrace;
    } else {
        const int i = g - F;
        if (!(finitef(u[i]) && finitef(v[i]) && finitef(sv[i]))) bits |= kErrAdvectionInput;
        out_v[i] = sl2_value<kVelocityY>(u, v, sv, n, dt_over_h, i % n, i / n, ok);
        if (!ok) bits |= kErrBacktrace;
        bits <<= 8;
    }
    if (bits) atomicOr(&st->error, bits);
}

// Commit of the advected velocity (MacGridFluidSolver2D.cpp:77-78) fused with
// addBuoyancy (:81-92): interior y faces get dt b (rho_below^1/4 + rho^1/4)/2.
__global__ void k2_commit_velocity(const float* __restrict__ un, const float* __restrict__ vn, float* __restrict__ u,
                                   float* __restrict__ v, const float* __restrict__ rho, int n, float dt_buoyancy,
                                   int buoyancy, Status2* st) {
    if (st->error) return;
    const int F = n * (n + 1);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= F) return;
    u[i] = un[i];
    float value = vn[i];
    const int x = i % n, y = i / n;
    if (buoyancy && y >= 1 && y <= n - 1) {
        const float smoke = .5f * (det_powf(rho[idC(n, x, y - 1)], .25f) + det_powf(rho[idC(n, x, y)], .25f));
        value += dt_buoyancy * smoke;
    }
    v[i] = value;
}

// emit (MacGridFluidSolver2D.cpp:51-71): the replenished Gaussian source and
// its inlet kick, one lane per cell.
__global__ void k2_emit(float* __restrict__ rho, float* __restrict__ v, int n, float cx, float cy, float r2,
                        float strength, float kick, Status2* st) {
    if (st->error) return;
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n * n) return;
    const int x = c % n, y = c / n;
    const float fn = (float)n;
    const float dx = ((float)x + .5f) / fn - cx;
    const float dy = ((float)y + .5f) / fn - cy;
    const float q = (dx * dx + dy * dy) / r2;
    if (q > 1.f) return;
    const float weight = det_expf(-3.5f * q);
    rho[c] = smax(rho[c], smin(1.f, strength * 3.f * weight));
    if (kick > 0.f) {
        float& vy = v[idY(n, x, y)];
        vy = smax(vy, kick * (.6f + strength) * weight);
    }
}

__global__ void k2_scale(const float* __restrict__ src, float* __restrict__ dst, size_t count, float factor,
                         Status2* st) {
    if (st->error) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = factor * src[i];
}

// Reflected source (MacGridFluidSolver2D.cpp:127-137): 2u - u~ with the wall
// faces zeroed.
__global__ void k2_hat(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ un,
                       const float* __restrict__ vn, float* __restrict__ uh, float* __restrict__ vh, int n,
                       Status2* st) {
    if (st->error) return;
    const int F = n * (n + 1);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= F) return;
    const int xu = i % (n + 1);
    uh[i] = (xu == 0 || xu == n) ? 0.f : 2.f * u[i] - un[i];
    const int yv = i / n;
    vh[i] = (yv == 0 || yv == n) ? 0.f : 2.f * v[i] - vn[i];
}

// ---------------------------------------------------------------------------
// CpuPressureSolver2D::project (CpuPressureSolver2D.cpp:9-81) as one workgroup.

// Steps before the sweeps: velocity check, walls, divergence, right-hand side
// and its removed mean. Returns false (error raised) where the reference throws.
// With `stage_all`, the right-hand side is left staged in `stage` as floats.
__device__ bool project_prologue(float* __restrict__ u, float* __restrict__ v, float* __restrict__ rhs_g, int n,
                                 const ProjOpts2& o, Status2* st, float* stage, int stage_floats, bool stage_all,
                                 float& div_before, double& mean, double* redd, float* redf) {
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int N = n * n, F = n * (n + 1);
    int bad = 0;
    for (int i = tid; i < F; i += nt) bad |= !(finitef(u[i]) && finitef(v[i]));
    if (__syncthreads_or(bad)) {
        if (tid == 0) atomicOr(&st->error, kErrVelocity);
        return false;
    }
    for (int i = tid; i < n; i += nt) {
        u[idX(n, 0, i)] = 0.f;
        u[idX(n, n, i)] = 0.f;
        v[idY(n, i, 0)] = 0.f;
        v[idY(n, i, n)] = 0.f;
    }
    __syncthreads();
    float dmax = 0.f;
    int nonfinite = 0, overflow = 0;
    for (int c = tid; c < N; c += nt) {
        const int x = c % n, y = c / n;
        const float d = (u[idX(n, x + 1, y)] - u[idX(n, x, y)] + v[idY(n, x, y + 1)] - v[idY(n, x, y)]) / o.h;
        nonfinite |= !finitef(d);
        dmax = smax(dmax, fabsf(d));
        const float r = d * -o.scale;
        overflow |= !finitef(r);
        rhs_g[c] = r;
        if (stage_all) stage[c] = r;
    }
    if (stage_all)
        for (int c = N + tid; c < ((N + 15) & ~15); c += nt) stage[c] = 0.f;
    if (__syncthreads_or(nonfinite)) {
        if (tid == 0) atomicOr(&st->error, kErrDivergence);
        return false;
    }
    div_before = block_max(dmax, redf);
    if (__syncthreads_or(overflow)) {
        if (tid == 0) atomicOr(&st->error, kErrRhs);
        return false;
    }
    __syncthreads();
    const double sum = sequential_sum(rhs_g, N, stage, stage_all ? N : stage_floats, stage_all, redd);
    mean = sum / (double)N;
    return true;
}

// Steps after the mean-free pressure is known (P(x, y) = P[x + stride y], in
// LDS or global memory): gradient, walls, divergence after, record.
__device__ void project_epilogue(float* __restrict__ u, float* __restrict__ v, const double* P, int stride, int n,
                                 const ProjOpts2& o, Status2* st, int slot, int iterations, int converged,
                                 double initial, double final_residual, float div_before, double mean, float* redf) {
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int N = n * n;
    const double scale = (double)o.dt / (double)o.h;
    const int inner = n * (n - 1);
    for (int k = tid; k < inner; k += nt) {
        const int x = 1 + k % (n - 1), y = k / (n - 1);
        float& f = u[idX(n, x, y)];
        f = (float)((double)f - scale * (P[x + stride * y] - P[x - 1 + stride * y]));
    }
    for (int k = tid; k < inner; k += nt) {
        const int x = k % n, y = 1 + k / n;
        float& f = v[idY(n, x, y)];
        f = (float)((double)f - scale * (P[x + stride * y] - P[x + stride * (y - 1)]));
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        u[idX(n, 0, i)] = 0.f;
        u[idX(n, n, i)] = 0.f;
        v[idY(n, i, 0)] = 0.f;
        v[idY(n, i, n)] = 0.f;
    }
    __syncthreads();
    float dmax = 0.f;
    int nonfinite = 0;
    for (int c = tid; c < N; c += nt) {
        const int x = c % n, y = c / n;
        const float d = (u[idX(n, x + 1, y)] - u[idX(n, x, y)] + v[idY(n, x, y + 1)] - v[idY(n, x, y)]) / o.h;
        nonfinite |= !finitef(d);
        dmax = smax(dmax, fabsf(d));
    }
    if (__syncthreads_or(nonfinite)) {
        if (tid == 0) atomicOr(&st->error, kErrDivergence);
        return;
    }
    const float div_after = block_max(dmax, redf);
    if (tid == 0) {
        ProjRecord2& r = st->proj[slot];
        r.iterations = iterations;
        r.converged = converged;
        r.initial_residual = initial;
        r.final_residual = final_residual;
        r.residual_available = 1;
        r.divergence_before = div_before;
        r.divergence_after = div_after;
        r.removed_rhs_mean = mean;
    }
}

// Register-tile schedule: lane (tx, ty) of a 32 x 32 grid owns cells
// [tx T, tx T + T) x [ty T, ty T + T), T even, so a cell's colour is fixed by
// its place in the tile and each pass's cells are known at compile time. P
// (LDS) is the pressure on an (n+2)^2 grid whose outer ring stays zero; it
// holds every cell on a tile edge, the only values other lanes read. Cells
// outside the domain (ring or partial tiles) read as +0: the reference skips
// those neighbours, and adding +0 to its neighbour sum, which is never -0
// (the pressure starts at +0 and an update p + w (q - p) is -0 only from
// -0), leaves the sum bit-identical. The neighbour count d divides the sum:
// by 4 (a multiply by 0.25, exact) inside, true division on the walls.
// FULL: n is a multiple of T, so every tile a lane owns lies inside the
// domain and a cell can meet a wall only in the tile's edge slots; the
// neighbour count is then 4 at compile time everywhere else.
template <int T, bool FULL>
__global__ __launch_bounds__(kProjThreads) void k2_project_tiles(float* __restrict__ u, float* __restrict__ v,
                                                                 float* __restrict__ rhs_g, double* __restrict__ p_g,
                                                                 int n, ProjOpts2 o, Status2* st, int slot) {
    static_assert(T % 2 == 0, "even tiles fix each cell's colour");
    extern __shared__ __align__(16) unsigned char dyn[];
    double* P = reinterpret_cast<double*>(dyn);
    float* S = reinterpret_cast<float*>(dyn);
    __shared__ double redd[16];
    __shared__ float redf[16];
    if (st->error) return;
    const int tid = (int)threadIdx.x;
    const int N = n * n, W = n + 2, PN = W * W, PN16 = (PN + 15) & ~15;
    float div_before;
    double mean;
    if (!project_prologue(u, v, rhs_g, n, o, st, S, N, true, div_before, mean, redd, redf)) return;
    const float meanf = (float)mean;
    const int x0 = (tid & 31) * T, y0 = (tid >> 5) * T;
    const bool full = x0 + T <= n && y0 + T <= n;
    const bool active = x0 < n && y0 < n;
    const int xlo = x0 == 0, xhi = x0 + T == n, ylo = y0 == 0, yhi = y0 + T == n;
    double pr[T][T];
    float br[T][T];
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const int x = x0 + i, y = y0 + j;
            pr[j][i] = 0.0;
            br[j][i] = (x < n && y < n) ? S[idC(n, x, y)] - meanf : 0.f;
        }
    __syncthreads();
    for (int c = tid; c < PN16; c += kProjThreads) P[c] = 0.0;
    __syncthreads();
    auto pid = [W](int x, int y) { return (y + 1) * W + (x + 1); };
    // Neighbour count of slot (i, j) at (x, y): 4 - the walls it touches.
    auto count = [&](int i, int j, int x, int y) {
        if constexpr (FULL)
            return 4 - (i == 0 ? xlo : 0) - (i == T - 1 ? xhi : 0) - (j == 0 ? ylo : 0) - (j == T - 1 ? yhi : 0);
        else
            return 4 - (x == 0) - (x == n - 1) - (y == 0) - (y == n - 1);
    };

    // Neighbour sum of CpuPressureSolver2D.cpp:34-41 in its order
    // (x-1, x+1, y-1, y+1): registers inside the tile, P across its edge.
    auto neighbours = [&](int i, int j, int x, int y) -> double {
        const double l = (i > 0) ? pr[j][i - 1] : P[pid(x - 1, y)];
        const double r = (i < T - 1) ? pr[j][i + 1] : P[pid(x + 1, y)];
        const double d = (j > 0) ? pr[j - 1][i] : P[pid(x, y - 1)];
        const double t = (j < T - 1) ? pr[j + 1][i] : P[pid(x, y + 1)];
        return ((l + r) + d) + t;
    };
    auto residual = [&]() -> double {
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < T; ++j)
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int x = x0 + i, y = y0 + j;
                if (FULL ? active : (x < n && y < n)) {
                    float b = br[j][i];
                    asm volatile("" : "+v"(b));
                    m = smax(m, fabs((double)b + neighbours(i, j, x, y) - (double)count(i, j, x, y) * pr[j][i]));
                }
            }
        return block_max(m, redd);
    };
    const double omega = (double)o.relaxation;
    auto pass = [&](auto parity) {
        if (FULL ? active : true) {
#pragma unroll
            for (int j = 0; j < T; ++j)
#pragma unroll
                for (int i = 0; i < T; ++i) {
                    if (((i + j) & 1) != decltype(parity)::value) continue;
                    const int x = x0 + i, y = y0 + j;
                    if (!FULL && !full && !(x < n && y < n)) continue;
                    // The right-hand side stays FP32 in registers (16 VGPRs at
                    // T = 4): the empty asm keeps its widening at the point of use.
                    float b = br[j][i];
                    asm volatile("" : "+v"(b));
                    const double a = (double)b + neighbours(i, j, x, y);
                    const bool edge = i == 0 || i == T - 1 || j == 0 || j == T - 1;
                    double q;
                    if (FULL && !edge) {
                        q = a * 0.25;
                    } else {
                        const int d = count(i, j, x, y);
                        if (d == 0) continue;  // the single cell of n = 1
                        q = div_234(a, d);
                    }
                    pr[j][i] += omega * (q - pr[j][i]);
                    if (edge) P[pid(x, y)] = pr[j][i];
                }
        }
        __syncthreads();
    };

    const double initial = residual();
    double final_residual = initial;
    const double target = smax(o.absolute_tolerance, o.relative_tolerance * initial);
    bool converged = final_residual <= target;
    int iterations = 0;
    for (int iter = 0; iter < o.max_iterations && (o.fixed_iterations || !converged); ++iter) {
        pass(std::integral_constant<int, 0>{});
        pass(std::integral_constant<int, 1>{});
        iterations = iter + 1;
        if (iterations % o.check_interval == 0 || iterations == o.max_iterations) {
            final_residual = residual();
            converged = final_residual <= target;
        }
    }
    // Pressure mean (CpuPressureSolver2D.cpp:70-74): every cell into P, then
    // the padded grid summed in linear order (its ring adds zeros).
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
        for (int i = 0; i < T; ++i)
            if (x0 + i < n && y0 + j < n) P[pid(x0 + i, y0 + j)] = pr[j][i];
    const double pmean = sequential_sum(P, PN, P, PN16, true, redd) / (double)N;
    int nonfinite = 0;
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
        for (int i = 0; i < T; ++i) {
            // reloaded from P, so the tile is not held in registers across the sum
            pr[j][i] = (x0 + i < n && y0 + j < n) ? P[pid(x0 + i, y0 + j)] : 0.0;
            pr[j][i] -= pmean;
            nonfinite |= (x0 + i < n && y0 + j < n) && !finited(pr[j][i]);
        }
    if (__syncthreads_or(nonfinite)) {
        if (tid == 0) atomicOr(&st->error, kErrPressure);
        return;
    }
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const int x = x0 + i, y = y0 + j;
            if (x < n && y < n) {
                P[pid(x, y)] = pr[j][i];
                p_g[idC(n, x, y)] = pr[j][i];
            } else {
                pr[j][i] = 0.0;  // cells outside keep reading as +0
            }
        }
    __syncthreads();
    final_residual = residual();
    converged = final_residual <= target;
    project_epilogue(u, v, P + W + 1, W, n, o, st, slot, iterations, converged, initial, final_residual, div_before,
                     mean, redf);
}

// Grids beyond one workgroup's LDS: the same steps with the right-hand side
// and the pressure in global memory, lanes striding over each colour.
__global__ __launch_bounds__(kProjThreads) void k2_project_global(float* __restrict__ u, float* __restrict__ v,
                                                                  float* __restrict__ rhs_g, double* __restrict__ p_g,
                                                                  int n, ProjOpts2 o, Status2* st, int slot) {
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ double redd[16];
    __shared__ float redf[16];
    if (st->error) return;
    const int tid = (int)threadIdx.x;
    const int N = n * n;
    float div_before;
    double mean;
    if (!project_prologue(u, v, rhs_g, n, o, st, reinterpret_cast<float*>(dyn), kProjLdsBytes / 4, false, div_before,
                          mean, redd, redf))
        return;
    const float meanf = (float)mean;
    for (int c = tid; c < N; c += kProjThreads) {
        rhs_g[c] -= meanf;
        p_g[c] = 0.0;
    }
    __syncthreads();
    auto neighbours = [&](int x, int y, int& d) -> double {
        double sum = 0.0;
        d = 0;
        if (x > 0) { sum += p_g[idC(n, x - 1, y)]; ++d; }
        if (x + 1 < n) { sum += p_g[idC(n, x + 1, y)]; ++d; }
        if (y > 0) { sum += p_g[idC(n, x, y - 1)]; ++d; }
        if (y + 1 < n) { sum += p_g[idC(n, x, y + 1)]; ++d; }
        return sum;
    };
    auto residual = [&]() -> double {
        double m = 0.0;
        for (int c = tid; c < N; c += kProjThreads) {
            int d;
            const double s = neighbours(c % n, c / n, d);
            m = smax(m, fabs((double)rhs_g[c] + s - (double)d * p_g[c]));
        }
        return block_max(m, redd);
    };
    const double omega = (double)o.relaxation;
    const int half = (n + 1) / 2;  // cells of one colour per row, at most
    auto pass = [&](int parity) {
        for (int k = tid; k < n * half; k += kProjThreads) {
            const int y = k / half;
            const int x = ((y + parity) & 1) + 2 * (k % half);
            if (x < n) {
                int d;
                const double s = neighbours(x, y, d);
                if (d != 0) {
                    double& p = p_g[idC(n, x, y)];
                    p += omega * (div_by_count((double)rhs_g[idC(n, x, y)] + s, d) - p);
                }
            }
        }
        __syncthreads();
    };
    const double initial = residual();
    double final_residual = initial;
    const double target = smax(o.absolute_tolerance, o.relative_tolerance * initial);
    bool converged = final_residual <= target;
    int iterations = 0;
    for (int iter = 0; iter < o.max_iterations && (o.fixed_iterations || !converged); ++iter) {
        pass(0);
        pass(1);
        iterations = iter + 1;
        if (iterations % o.check_interval == 0 || iterations == o.max_iterations) {
            final_residual = residual();
            converged = final_residual <= target;
        }
    }
    const double pmean = sequential_sum(p_g, N, reinterpret_cast<double*>(dyn), kProjLdsBytes / 8, false, redd) /
                         (double)N;
    int nonfinite = 0;
    for (int c = tid; c < N; c += kProjThreads) {
        p_g[c] -= pmean;
        nonfinite |= !finited(p_g[c]);
    }
    if (__syncthreads_or(nonfinite)) {
        if (tid == 0) atomicOr(&st->error, kErrPressure);
        return;
    }
    final_residual = residual();
    converged = final_residual <= target;
    project_epilogue(u, v, p_g, n, n, o, st, slot, iterations, converged, initial, final_residual, div_before, mean,
                     redf);
}

// ---------------------------------------------------------------------------
// FluidDiagnostics (MacGridFluidSolver2D.cpp:144-153; ReflectionMacFluidSolver2D.hpp:44-57):
// kinetic energy from the reference's sequential FP64 sum over u then v,
// max |velocity|, max |div u| (MacGridFluidSolver2D throws on a non-finite one).
__global__ __launch_bounds__(kProjThreads) void k2_diagnostics(const float* __restrict__ u,
                                                               const float* __restrict__ v, int n, float h,
                                                               int historical, Status2* st) {
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ float redf[16];
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int F = n * (n + 1), N = n * n;
    float* stage = reinterpret_cast<float*>(dyn);
    const int chunk = kProjLdsBytes / 4;
    double sum = 0.0;
    float vmax = 0.f;
    for (int field = 0; field < 2; ++field) {
        const float* f = field == 0 ? u : v;
        for (int base = 0; base < F; base += chunk) {
            const int len = min(chunk, F - base);
            __syncthreads();
            for (int k = tid; k < len; k += nt) {
                const float x = f[base + k];
                stage[k] = x;
                vmax = smax(vmax, fabsf(x));
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll 8
                for (int k = 0; k < len; ++k) sum += (double)stage[k] * (double)stage[k];
            }
        }
    }
    vmax = block_max(vmax, redf);
    float dmax = 0.f;
    int nonfinite = 0;
    for (int c = tid; c < N; c += nt) {
        const int x = c % n, y = c / n;
        const float d = (u[idX(n, x + 1, y)] - u[idX(n, x, y)] + v[idY(n, x, y + 1)] - v[idY(n, x, y)]) / h;
        nonfinite |= !finitef(d);
        dmax = smax(dmax, fabsf(d));
    }
    const int bad = __syncthreads_or(nonfinite);
    dmax = block_max(dmax, redf);
    if (tid == 0) {
        const double hd = (double)h;
        st->kinetic_energy = historical ? (float)(.5 * sum * hd * hd) : (float)(.5 * hd * hd * sum);
        st->max_velocity = vmax;
        st->max_divergence = dmax;
        if (bad && !historical) atomicOr(&st->error, kErrDivergence);
    }
}

// ---------------------------------------------------------------------------
// ReflectionMacFluidSolver2D.

__global__ void kh_walls(float* __restrict__ vx, float* __restrict__ vy, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    vx[idX(n, 0, i)] = 0.f;
    vx[idX(n, n, i)] = 0.f;
    vy[idY(n, i, 0)] = 0.f;
    vy[idY(n, i, n)] = 0.f;
}

// advectScalar's forward trace (ReflectionMacFluidSolver2D.hpp:162-167). For
// semi-Lagrangian transport the result is final and carries the decay factor.
__global__ void kh_trace_forward(const float* __restrict__ smoke, const float* __restrict__ vx,
                                 const float* __restrict__ vy, float* __restrict__ out, int n, float dt,
                                 float factor, int apply_factor) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n * n) return;
    const Hist H{n, 1.0f / (float)n};
    const float f = H.trace_cell<-1>(smoke, vx, vy, c % n, c / n, dt);
    out[c] = apply_factor ? f * factor : f;
}

// MacCormack back-trace and correction (:174-189), before the limiter.
__global__ void kh_trace_back(const float* __restrict__ smoke, const float* __restrict__ fwd,
                              const float* __restrict__ vx, const float* __restrict__ vy, float* __restrict__ corr,
                              int n, float dt) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n * n) return;
    const Hist H{n, 1.0f / (float)n};
    const float back = H.trace_cell<1>(fwd, vx, vy, c % n, c / n, dt);
    corr[c] = fwd[c] + 0.5f * (smoke[c] - back);
}

// The limiter (:191-204) in place, wavefront over x + 2y = t, then the
// factor. With LDS the field and the forward trace are staged (2 n^2 floats).
template <bool LDS>
__global__ __launch_bounds__(128) void kh_limit(float* __restrict__ smoke, const float* __restrict__ fwd,
                                                const float* __restrict__ corr, int n, float factor) {
    extern __shared__ __align__(16) unsigned char dyn[];
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int N = n * n;
    float* Fd = smoke;
    const float* Wd = fwd;
    if constexpr (LDS) {
        float* a = reinterpret_cast<float*>(dyn);
        float* b = a + N;
        for (int c = tid; c < N; c += nt) {
            a[c] = smoke[c];
            b[c] = fwd[c];
        }
        Fd = a;
        Wd = b;
        __syncthreads();
    }
    for (int t = 0; t <= 3 * (n - 1); ++t) {
        int ylo, yhi;
        limit_front(t, n, ylo, yhi);
        for (int y = ylo + tid; y <= yhi; y += nt) {
            const int x = t - 2 * y;
            const int i = idC(n, x, y);
            Fd[i] = limit_cell(Fd, Wd, corr[i], n, x, y);
        }
        __syncthreads();
    }
    for (int c = tid; c < N; c += nt) smoke[c] = Fd[c] * factor;
}

// advectVx / advectVy (:196-226), or the reflected pass (:363-384): lanes
// [0, F) the x faces, [F, 2F) the y faces.
__global__ void kh_advect_velocity(const float* __restrict__ srcx, const float* __restrict__ srcy,
                                   const float* __restrict__ fu, const float* __restrict__ fv,
                                   float* __restrict__ outx, float* __restrict__ outy, int n, float dt) {
    const int F = n * (n + 1);
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= 2 * F) return;
    const Hist H{n, 1.0f / (float)n};
    if (g < F) {
        outx[g] = H.face_x(srcx, fu, fv, g % (n + 1), g / (n + 1), dt);
    } else {
        const int i = g - F;
        outy[i] = H.face_y(srcy, fu, fv, i % n, i / n, dt);
    }
}

// vx = vx~; vy = vy~ plus the buoyancy of the cells below and above the face
// (:327-332), in the reference's order (the lower cell first).
__global__ void kh_forces(const float* __restrict__ vxt, const float* __restrict__ vyt, float* __restrict__ vx,
                          float* __restrict__ vy, const float* __restrict__ smoke, int n, float buoyancy,
                          float exponent, float half_dt, float split) {
    const int F = n * (n + 1);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= F) return;
    vx[i] = vxt[i];
    const int x = i % n, y = i / n;
    float value = vyt[i];
    if (y >= 1) {
        const float f = buoyancy * det_powf(smoke[idC(n, x, y - 1)], exponent) * half_dt;
        value += split * f;
    }
    if (y <= n - 1) {
        const float f = buoyancy * det_powf(smoke[idC(n, x, y)], exponent) * half_dt;
        value += split * f;
    }
    vy[i] = value;
}

__global__ void kh_hat(const float* __restrict__ mx, const float* __restrict__ my, const float* __restrict__ tx,
                       const float* __restrict__ ty, float* __restrict__ hx, float* __restrict__ hy, int n) {
    const int F = n * (n + 1);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= F) return;
    hx[i] = 2.f * mx[i] - tx[i];
    hy[i] = 2.f * my[i] - ty[i];
}

// emit (:291-314) as one workgroup: in-circle cells of the scanned rows in
// scan order get their ranks by a block prefix count; rank k draws the LCG
// outputs 2k+1 and 2k+2 past the state at entry (jump-ahead), so the stirring
// of face x of a row adds the second draw of cell x-1, then the first of cell x.
__device__ __forceinline__ bool hist_in_circle(const HistEmit& e, int n, float h, int x, int y, float& w) {
    return hist_in_disc(e.cx, e.cy, e.radius, e.falloff, n, h, x, y, w);
}

__global__ __launch_bounds__(1024) void kh_emit(float* __restrict__ smoke, float* __restrict__ vx,
                                                float* __restrict__ vy, int* __restrict__ ranks, int n, HistEmit e,
                                                Status2* st) {
    __shared__ int wave_counts[16];
    __shared__ int chunk_total;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float h = 1.0f / (float)n;
    const int rows = e.y_end - e.y_begin;
    if (rows <= 0) return;
    const int L = rows * n;
    int base = 0;
    for (int start = 0; start < L; start += 1024) {
        const int b = start + tid;
        float w = 0.f;
        bool in = false;
        int x = 0, y = 0;
        if (b < L) {
            x = b % n;
            y = e.y_begin + b / n;
            in = hist_in_circle(e, n, h, x, y, w);
        }
        const unsigned long long mask = __ballot(in);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wave_counts[wave] = __popcll(mask);
        __syncthreads();
        int offset = 0;
        for (int k = 0; k < wave; ++k) offset += wave_counts[k];
        if (tid == 0) {
            int t = 0;
            for (int k = 0; k < 16; ++k) t += wave_counts[k];
            chunk_total = t;
        }
        if (in) {
            const int i = idC(n, x, y);
            smoke[i] = smax(smoke[i], smin(1.0f, e.strength * e.smoke_scale * w));
            if (e.kick > 0.f) {
                float& f = vy[idY(n, x, y)];
                f = smax(f, e.kick * (e.kick_base + e.kick_scale * e.strength) * w);
            }
            ranks[b] = base + offset + before;
        }
        __syncthreads();
        base += chunk_total;
        __syncthreads();
    }
    if (!(e.stir > 0.f)) return;
    const uint32_t s0 = st->random_state;
    __syncthreads();
    const int faces = rows * (n + 1);
    for (int k = tid; k < faces; k += 1024) {
        const int x = k % (n + 1), y = e.y_begin + k / (n + 1);
        float& f = vx[idX(n, x, y)];
        float value = f;
        float w;
        if (x >= 1 && hist_in_circle(e, n, h, x - 1, y, w)) {
            const int r = ranks[(y - e.y_begin) * n + (x - 1)];
            uint32_t A, C;
            lcg_jump(2u * (uint32_t)r + 2u, A, C);
            value += (lcg_unit(A * s0 + C) - 0.5f) * e.stir * w;
        }
        if (x <= n - 1 && hist_in_circle(e, n, h, x, y, w)) {
            const int r = ranks[(y - e.y_begin) * n + x];
            uint32_t A, C;
            lcg_jump(2u * (uint32_t)r + 1u, A, C);
            value += (lcg_unit(A * s0 + C) - 0.5f) * e.stir * w;
        }
        f = value;
    }
    if (tid == 0) {
        uint32_t A, C;
        lcg_jump(2u * (uint32_t)base, A, C);
        st->random_state = A * s0 + C;
    }
}

// project (:239-289): walls, divergence (kept in div), fixed-count FP32
// red-black SOR from zero, gradient, walls. With T > 0 (even) the register
// tiles of k2_project_tiles over an (n+2)^2 LDS grid with a zero ring; the
// reference's neighbour sum starts at +0 and the pressure is never -0 (the
// relaxed update (1 - w) p + c gives -0 only from c = -0, which needs a -0
// sum), so the ring's +0 terms leave it bit-identical. T == 0: global memory.
template <int T, bool FULL>
__global__ __launch_bounds__(kProjThreads) void kh_project(float* __restrict__ vx, float* __restrict__ vy,
                                                           float* __restrict__ p_g, float* __restrict__ div_g, int n,
                                                           float dt, int iterations, float omega) {
    static_assert(T % 2 == 0, "even tiles fix each cell's colour");
    extern __shared__ __align__(16) unsigned char dyn[];
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int N = n * n;
    const float h = 1.0f / (float)n;
    const float hinv = 1.0f / h, h2 = h * h;
    for (int i = tid; i < n; i += nt) {
        vx[idX(n, 0, i)] = 0.f;
        vx[idX(n, n, i)] = 0.f;
        vy[idY(n, i, 0)] = 0.f;
        vy[idY(n, i, n)] = 0.f;
    }
    __syncthreads();
    for (int c = tid; c < N; c += nt) {
        const int x = c % n, y = c / n;
        div_g[c] = (vx[idX(n, x + 1, y)] - vx[idX(n, x, y)] + vy[idY(n, x, y + 1)] - vy[idY(n, x, y)]) * hinv;
    }
    __syncthreads();
    const float keep = 1.f - omega;
    const int W = T > 0 ? n + 2 : n;
    // P(x, y) = Pb[x + W y]; with tiles, Pb sits inside the zero ring.
    float* P = T > 0 ? reinterpret_cast<float*>(dyn) : p_g;
    float* Pb = T > 0 ? P + W + 1 : p_g;
    const int PN = T > 0 ? W * W : N;
    for (int c = tid; c < PN; c += nt) P[c] = 0.f;
    if constexpr (T > 0) {
        const int x0 = (tid & 31) * T, y0 = (tid >> 5) * T;
        const bool full = x0 + T <= n && y0 + T <= n;
        const bool active = x0 < n && y0 < n;
        const int xlo = x0 == 0, xhi = x0 + T == n, ylo = y0 == 0, yhi = y0 + T == n;
        float pr[T][T], br[T][T];
#pragma unroll
        for (int j = 0; j < T; ++j)
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int x = x0 + i, y = y0 + j;
                pr[j][i] = 0.f;
                br[j][i] = (x < n && y < n) ? div_g[idC(n, x, y)] * h2 / dt : 0.f;
            }
        __syncthreads();
        auto pass = [&](auto parity) {
            if (FULL ? active : true) {
#pragma unroll
                for (int j = 0; j < T; ++j)
#pragma unroll
                    for (int i = 0; i < T; ++i) {
                        if (((i + j) & 1) != decltype(parity)::value) continue;
                        const int x = x0 + i, y = y0 + j;
                        if (!FULL && !full && !(x < n && y < n)) continue;
                        const float l = (i > 0) ? pr[j][i - 1] : Pb[x - 1 + W * y];
                        const float r = (i < T - 1) ? pr[j][i + 1] : Pb[x + 1 + W * y];
                        const float d = (j > 0) ? pr[j - 1][i] : Pb[x + W * (y - 1)];
                        const float t = (j < T - 1) ? pr[j + 1][i] : Pb[x + W * (y + 1)];
                        const float sum = ((l + r) + d) + t;
                        const float c = omega * (sum - br[j][i]);
                        const bool edge = i == 0 || i == T - 1 || j == 0 || j == T - 1;
                        float q;
                        if (FULL && !edge) {
                            q = c * 0.25f;
                        } else {
                            const int terms = FULL ? 4 - (i == 0 ? xlo : 0) - (i == T - 1 ? xhi : 0) -
                                                         (j == 0 ? ylo : 0) - (j == T - 1 ? yhi : 0)
                                                   : 4 - (x == 0) - (x == n - 1) - (y == 0) - (y == n - 1);
                            q = FULL ? div_234(c, terms) : div_by_count(c, smax(terms, 1));
                        }
                        pr[j][i] = keep * pr[j][i] + q;
                        if (edge) Pb[x + W * y] = pr[j][i];
                    }
            }
            __syncthreads();
        };
        for (int it = 0; it < iterations; ++it) {
            pass(std::integral_constant<int, 0>{});
            pass(std::integral_constant<int, 1>{});
        }
#pragma unroll
        for (int j = 0; j < T; ++j)
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int x = x0 + i, y = y0 + j;
                if (x < n && y < n) {
                    Pb[x + W * y] = pr[j][i];
                    p_g[idC(n, x, y)] = pr[j][i];
                }
            }
    } else {
        __syncthreads();
        const int half = (n + 1) / 2;
        for (int it = 0; it < iterations; ++it) {
            for (int parity = 0; parity < 2; ++parity) {
                for (int k = tid; k < n * half; k += nt) {
                    const int y = k / half;
                    const int x = ((y + parity) & 1) + 2 * (k % half);
                    if (x < n) {
                        float sum = 0.f;
                        int terms = 0;
                        if (x > 0) { sum += P[idC(n, x - 1, y)]; ++terms; }
                        if (x + 1 < n) { sum += P[idC(n, x + 1, y)]; ++terms; }
                        if (y > 0) { sum += P[idC(n, x, y - 1)]; ++terms; }
                        if (y + 1 < n) { sum += P[idC(n, x, y + 1)]; ++terms; }
                        const int c = idC(n, x, y);
                        P[c] = keep * P[c] + div_by_count(omega * (sum - div_g[c] * h2 / dt), smax(terms, 1));
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    const int inner = n * (n - 1);
    for (int k = tid; k < inner; k += nt) {
        const int x = 1 + k % (n - 1), y = k / (n - 1);
        vx[idX(n, x, y)] -= dt * (Pb[x + W * y] - Pb[x - 1 + W * y]) * hinv;
    }
    for (int k = tid; k < inner; k += nt) {
        const int x = k % n, y = 1 + k / n;
        vy[idY(n, x, y)] -= dt * (Pb[x + W * y] - Pb[x + W * (y - 1)]) * hinv;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        vx[idX(n, 0, i)] = 0.f;
        vx[idX(n, n, i)] = 0.f;
        vy[idY(n, i, 0)] = 0.f;
        vy[idY(n, i, n)] = 0.f;
    }
}

// Raises a kernel's dynamic LDS limit above the 64 KiB default (a host-side
// attribute; set before each launch that needs it).
template <class K> void set_lds(K kernel, int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes);
}

}  // namespace

// Largest n whose n^2 pressure values (FP64 for MacGridFluidSolver2D, FP32
// for the historical solver) the register-tile schedule holds in LDS.
constexpr int kTileMaxDouble = 128;  // 32 lanes x T = 4
constexpr int kTileMaxFloat = 192;   // 32 lanes x T = 6

void launch_advect(hipStream_t s, int field, const float* u, const float* v, const float* src, float* out, int n,
                   float dt_over_h, Status2* st) {
    const unsigned g = blocks_for((size_t)n * (n + 1));
    if (field == kDensity) k2_advect<kDensity><<<g, kThreads, 0, s>>>(u, v, src, out, n, dt_over_h, st);
    else if (field == kVelocityX) k2_advect<kVelocityX><<<g, kThreads, 0, s>>>(u, v, src, out, n, dt_over_h, st);
    else k2_advect<kVelocityY><<<g, kThreads, 0, s>>>(u, v, src, out, n, dt_over_h, st);
}

void launch_advect_velocity(hipStream_t s, const float* u, const float* v, const float* su, const float* sv,
                            float* out_u, float* out_v, int n, float dt_over_h, Status2* st) {
    k2_advect_velocity<<<blocks_for(2 * (size_t)n * (n + 1)), kThreads, 0, s>>>(u, v, su, sv, out_u, out_v, n,
                                                                                dt_over_h, st);
}

void launch_commit_velocity(hipStream_t s, const float* un, const float* vn, float* u, float* v, const float* rho,
                            int n, float dt_buoyancy, bool buoyancy, Status2* st) {
    k2_commit_velocity<<<blocks_for((size_t)n * (n + 1)), kThreads, 0, s>>>(un, vn, u, v, rho, n, dt_buoyancy,
                                                                          buoyancy ? 1 : 0, st);
}

void launch_emit(hipStream_t s, float* rho, float* v, int n, float cx, float cy, float r2, float strength,
                 float kick, Status2* st) {
    k2_emit<<<blocks_for((size_t)n * n), kThreads, 0, s>>>(rho, v, n, cx, cy, r2, strength, kick, st);
}

void launch_project(hipStream_t s, float* u, float* v, float* rhs, double* p, int n, const ProjOpts2& o,
                    Status2* st, int slot) {
    if (n <= kTileMaxDouble) {
        const int bytes = (((n + 2) * (n + 2) + 15) & ~15) * (int)sizeof(double);
        if (n <= 64) {
            if (n % 2 == 0) {
                set_lds(k2_project_tiles<2, true>, kProjLdsBytes);
                k2_project_tiles<2, true><<<1, kProjThreads, bytes, s>>>(u, v, rhs, p, n, o, st, slot);
            } else {
                set_lds(k2_project_tiles<2, false>, kProjLdsBytes);
                k2_project_tiles<2, false><<<1, kProjThreads, bytes, s>>>(u, v, rhs, p, n, o, st, slot);
            }
        } else if (n % 4 == 0) {
            set_lds(k2_project_tiles<4, true>, kProjLdsBytes);
            k2_project_tiles<4, true><<<1, kProjThreads, bytes, s>>>(u, v, rhs, p, n, o, st, slot);
        } else {
            set_lds(k2_project_tiles<4, false>, kProjLdsBytes);
            k2_project_tiles<4, false><<<1, kProjThreads, bytes, s>>>(u, v, rhs, p, n, o, st, slot);
        }
    } else {
        set_lds(k2_project_global, kProjLdsBytes);
        k2_project_global<<<1, kProjThreads, kProjLdsBytes, s>>>(u, v, rhs, p, n, o, st, slot);
    }
}

void launch_scale(hipStream_t s, const float* src, float* dst, size_t count, float factor, Status2* st) {
    k2_scale<<<blocks_for(count), kThreads, 0, s>>>(src, dst, count, factor, st);
}

void launch_hat(hipStream_t s, const float* u, const float* v, const float* un, const float* vn, float* uh,
                float* vh, int n, Status2* st) {
    k2_hat<<<blocks_for((size_t)n * (n + 1)), kThreads, 0, s>>>(u, v, un, vn, uh, vh, n, st);
}

void launch_diagnostics(hipStream_t s, const float* u, const float* v, int n, float h, bool historical,
                        Status2* st) {
    set_lds(k2_diagnostics, kProjLdsBytes);
    k2_diagnostics<<<1, kProjThreads, kProjLdsBytes, s>>>(u, v, n, h, historical ? 1 : 0, st);
}

void launch_hist_walls(hipStream_t s, float* vx, float* vy, int n) {
    kh_walls<<<blocks_for((size_t)n), kThreads, 0, s>>>(vx, vy, n);
}

void launch_hist_advect_scalar(hipStream_t s, float* smoke, float* fwd, float* corr, const float* vx,
                               const float* vy, int n, float dt, bool maccormack, float decay, float* out) {
    const unsigned g = blocks_for((size_t)n * n);
    if (!maccormack) {
        kh_trace_forward<<<g, kThreads, 0, s>>>(smoke, vx, vy, out, n, dt, decay, 1);
        return;
    }
    kh_trace_forward<<<g, kThreads, 0, s>>>(smoke, vx, vy, fwd, n, dt, 1.f, 0);
    kh_trace_back<<<g, kThreads, 0, s>>>(smoke, fwd, vx, vy, corr, n, dt);
    const int bytes = 2 * n * n * (int)sizeof(float);
    if (bytes <= kProjLdsBytes) {
        set_lds(kh_limit<true>, kProjLdsBytes);
        kh_limit<true><<<1, 128, bytes, s>>>(smoke, fwd, corr, n, decay);
    } else {
        kh_limit<false><<<1, 128, 0, s>>>(smoke, fwd, corr, n, decay);
    }
}

void launch_hist_advect_velocity(hipStream_t s, const float* srcx, const float* srcy, const float* fu,
                                 const float* fv, float* outx, float* outy, int n, float dt) {
    kh_advect_velocity<<<blocks_for(2 * (size_t)n * (n + 1)), kThreads, 0, s>>>(srcx, srcy, fu, fv, outx, outy, n,
                                                                                dt);
}

void launch_hist_forces(hipStream_t s, const float* vxt, const float* vyt, float* vx, float* vy,
                        const float* smoke, int n, float buoyancy, float exponent, float half_dt, float split) {
    kh_forces<<<blocks_for((size_t)n * (n + 1)), kThreads, 0, s>>>(vxt, vyt, vx, vy, smoke, n, buoyancy, exponent,
                                                                  half_dt, split);
}

void launch_hist_emit(hipStream_t s, float* smoke, float* vx, float* vy, int* ranks, int n, const HistEmit& e,
                      Status2* st) {
    kh_emit<<<1, 1024, 0, s>>>(smoke, vx, vy, ranks, n, e, st);
}

void launch_hist_project(hipStream_t s, float* vx, float* vy, float* p, float* div, int n, float dt,
                         int iterations, float omega) {
    if (n <= kTileMaxFloat) {
        const int bytes = (n + 2) * (n + 2) * (int)sizeof(float);
        const int T = n <= 64 ? 2 : (n <= 128 ? 4 : 6);
        const bool full = n % T == 0;
#define VGX2_HIST_LAUNCH(TT, FF)                                                                \
    do {                                                                                        \
        set_lds(kh_project<TT, FF>, kProjLdsBytes);                                             \
        kh_project<TT, FF><<<1, kProjThreads, bytes, s>>>(vx, vy, p, div, n, dt, iterations, omega); \
    } while (0)
        if (T == 2) {
            if (full) VGX2_HIST_LAUNCH(2, true);
            else VGX2_HIST_LAUNCH(2, false);
        } else if (T == 4) {
            if (full) VGX2_HIST_LAUNCH(4, true);
            else VGX2_HIST_LAUNCH(4, false);
        } else {
            if (full) VGX2_HIST_LAUNCH(6, true);
            else VGX2_HIST_LAUNCH(6, false);
        }
#undef VGX2_HIST_LAUNCH
    } else {
        kh_project<0, false><<<1, kProjThreads, 0, s>>>(vx, vy, p, div, n, dt, iterations, omega);
    }
}

void launch_hist_hat(hipStream_t s, const float* mid_x, const float* mid_y, const float* tx, const float* ty,
                     float* hx, float* hy, int n) {
    kh_hat<<<blocks_for((size_t)n * (n + 1)), kThreads, 0, s>>>(mid_x, mid_y, tx, ty, hx, hy, n);
}

}  // namespace vgx2
