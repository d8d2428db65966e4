// vgx_unet_kernels.hip - kernels of the learned pressure warm start: a 3D
// U-Net in BF16 with FP32 accumulation on the gfx950 matrix cores.
//
// Tensors are channels-last: element (voxel v, channel c) of a view lives at
// p[v * stride + off + c], voxels numbered x + X (y + Y z). Strides, offsets
// and channel counts are multiples of 8, so eight channels are one 16-byte
// load. A 3^3 convolution is an implicit GEMM: rows are voxels, columns are
// output channels, and k = tap * cin + ci runs over the 27 taps (tap =
// (dz + 1) 9 + (dy + 1) 3 + (dx + 1)) and the input channels; taps outside
// the grid read zero. Weights are [cout][kp] with kp = 27 cin rounded up to
// 32 and zero padding.
//
// MFMA v_mfma_f32_16x16x32_bf16 operand layout (checked on the device): A
// (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. + 7; B (32 x 16):
// lane l holds column l % 16, the same k; C (16 x 16): lane l holds column
// l % 16, rows 4 (l / 16) .. + 3.
#include "vgx_unet.hpp"

namespace vgx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2bf(float v) { return (bf16)v; }

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (bf16)0.f;
    return z;
}

// ------------------------------------------------------------ convolution
// One wave computes 32 voxels (two row tiles) x NT column tiles of 16 output
// channels over the k-steps k0 = 32 (first + j * step); k_conv3 runs four
// independent waves per block over all k-steps, k_conv3_split gives one
// 32-voxel tile to a block of S waves that split the k-steps and sum their
// partial tiles in LDS (a fixed order) -- the coarse levels have too few
// voxels to fill the GPU one wave per tile.
constexpr int kConvWaves = 4, kConvRows = 32;

// k-steps in flight per wave: each step's operands are loaded D - 1 steps
// ahead of its MFMAs, so a wave waits on one memory latency per D steps
// instead of one per step (2 + NT 16-byte operands a step).
template <int NT> constexpr int conv_depth() { return NT == 1 ? 4 : (NT == 2 ? 3 : 2); }

template <int NT>
__device__ __forceinline__ void conv_tile(const ConvArgs& a, int base, int lane, int first, int step,
                                          f32x4 (&acc)[2][NT]) {
    constexpr int D = conv_depth<NT>();
    const int V = a.X * a.Y * a.Z;
    const int K = 27 * a.cin;
    // the two voxels this lane loads A rows for
    int vx[2], vy[2], vz[2];
    bool vin[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int v = base + r * 16 + (lane & 15);
        vin[r] = v < V;
        const int vv = vin[r] ? v : 0;
        vx[r] = vv % a.X;
        vy[r] = (vv / a.X) % a.Y;
        vz[r] = vv / (a.X * a.Y);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < NT; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this wave's k-steps: k0 = 32 (first + j step), j = 0 .. nsteps - 1, in
    // increasing k (the accumulation order); the load cursor k = k0 + ksub
    // walks them with its tap t and channel ci (k = t cin + ci) advanced
    // without divisions
    const int total = a.kp / 32;
    const int nsteps = first < total ? (total - first + step - 1) / step : 0;
    const int kstride = 32 * step, dtap = kstride / a.cin, dci = kstride - dtap * a.cin;
    int k = 32 * first + 8 * (lane >> 4);
    int t = k / a.cin, ci = k - t * a.cin;
    bf16x8 A[D][2], B[D][NT];
    bool ok[D][2];
    // Operands of the cursor's step into slot q. A tap outside the grid (or
    // k in the zero padding beyond 27 cin) loads voxel 0 and is replaced by
    // zeros at use, so no load waits under a branch.
    auto load = [&](int q) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int x = vx[r] + dx, y = vy[r] + dy, z = vz[r] + dz;
            const bool in = vin[r] && k < K && x >= 0 && y >= 0 && z >= 0 && x < a.X && y < a.Y && z < a.Z;
            const size_t e = in ? (size_t)(x + a.X * (y + a.Y * z)) * a.in_stride + a.in_off + ci : 0;
            ok[q][r] = in;
            A[q][r] = *reinterpret_cast<const bf16x8*>(a.in + e);
        }
#pragma unroll
        for (int c = 0; c < NT; ++c)
            B[q][c] = *reinterpret_cast<const bf16x8*>(a.w + (size_t)(c * 16 + (lane & 15)) * a.kp + k);
        k += kstride;
        t += dtap;
        ci += dci;
        if (ci >= a.cin) {
            ci -= a.cin;
            ++t;
        }
    };
#pragma unroll
    for (int q = 0; q < D - 1; ++q)
        if (q < nsteps) load(q);
    for (int j0 = 0; j0 < nsteps; j0 += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int j = j0 + s;
            if (j >= nsteps) break;
            if (j + D - 1 < nsteps) load((s + D - 1) % D);
            const bf16x8 a0 = ok[s][0] ? A[s][0] : zero8(), a1 = ok[s][1] ? A[s][1] : zero8();
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, B[s][c], acc[0][c], 0, 0, 0);
                acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, B[s][c], acc[1][c], 0, 0, 0);
            }
        }
    }
}
template <int NT>
__device__ __forceinline__ void conv_store(const ConvArgs& a, int base, int lane, const f32x4 (&acc)[2][NT]) {
    const int V = a.X * a.Y * a.Z;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < NT; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int v = base + r * 16 + 4 * (lane >> 4) + i;
                if (v >= V) continue;
                float y = acc[r][c][i];
                if (a.relu) y = y > 0.f ? y : 0.f;
                a.out[(size_t)v * a.out_stride + a.out_off + c * 16 + (lane & 15)] = f2bf(y);
            }
}

template <int NT>
__global__ void __launch_bounds__(64 * kConvWaves) k_conv3(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = (blockIdx.x * kConvWaves + wave) * kConvRows;
    if (base >= a.X * a.Y * a.Z) return;
    f32x4 acc[2][NT];
    conv_tile<NT>(a, base, lane, 0, 1, acc);
    conv_store<NT>(a, base, lane, acc);
}

template <int NT, int S>
__global__ void __launch_bounds__(64 * S) k_conv3_split(ConvArgs a) {
    __shared__ float part[(S - 1) * 2 * NT * 4 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * kConvRows;
    f32x4 acc[2][NT];
    conv_tile<NT>(a, base, lane, wave, S, acc);
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < NT; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) part[(((wave - 1) * 2 + r) * NT * 4 + c * 4 + i) * 64 + lane] = acc[r][c][i];
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll 1
    for (int w = 1; w < S; ++w)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < NT; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[r][c][i] += part[(((w - 1) * 2 + r) * NT * 4 + c * 4 + i) * 64 + lane];
    conv_store<NT>(a, base, lane, acc);
}

// --------------------------------------------------------- weight gradients
// dW[co][t cin + ci] = sum_v dz[v][co] x[v + off(t)][ci]: block (chunk, tap)
// sums a voxel chunk; per 32 voxels the block stages dz and the tap-shifted x
// transposed in LDS ([channel][voxel]) so both MFMA operands are contiguous.
// Partials go to part[chunk][co][t][ci] and k_wgrad_reduce sums them in
// chunk order (deterministic).
constexpr int kWgradVoxels = 1024, kWgradMaxCout = 64, kWgradMaxCin = 112;

__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
    __shared__ bf16 dzT[kWgradMaxCout][32 + 8];
    __shared__ bf16 xT[kWgradMaxCin][32 + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int chunk = blockIdx.x, t = blockIdx.y;
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const int V = a.X * a.Y * a.Z;
    const int ct_n = a.cout / 16, it_n = a.cin / 16, tiles = ct_n * it_n;
    f32x4 acc[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int v_begin = chunk * kWgradVoxels, v_end = min(V, v_begin + kWgradVoxels);
    for (int vs = v_begin; vs < v_end; vs += 32) {
        // stage: dz rows (cout / 8 loads of 8 channels per voxel), x rows (cin / 8)
        const int dz_loads = 32 * (a.cout / 8), x_loads = 32 * (a.cin / 8);
        for (int e = threadIdx.x; e < dz_loads + x_loads; e += 256) {
            if (e < dz_loads) {
                const int vv = e / (a.cout / 8), c8 = (e % (a.cout / 8)) * 8, v = vs + vv;
                bf16x8 val = zero8();
                if (v < v_end) val = *reinterpret_cast<const bf16x8*>(a.dz + (size_t)v * a.dz_stride + a.dz_off + c8);
#pragma unroll
                for (int j = 0; j < 8; ++j) dzT[c8 + j][vv] = val[j];
            } else {
                const int ee = e - dz_loads;
                const int vv = ee / (a.cin / 8), c8 = (ee % (a.cin / 8)) * 8, v = vs + vv;
                bf16x8 val = zero8();
                if (v < v_end) {
                    const int x = v % a.X + dx, y = (v / a.X) % a.Y + dy, z = v / (a.X * a.Y) + dz;
                    if (x >= 0 && y >= 0 && z >= 0 && x < a.X && y < a.Y && z < a.Z)
                        val = *reinterpret_cast<const bf16x8*>(a.x + (size_t)(x + a.X * (y + a.Y * z)) * a.x_stride +
                                                               a.x_off + c8);
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) xT[c8 + j][vv] = val[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int q = wave + 4 * j;
            if (q < tiles) {
                const int ct = q / it_n, it = q % it_n;
                bf16x8 av, bv;
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    av[s] = dzT[ct * 16 + (lane & 15)][8 * (lane >> 4) + s];
                    bv[s] = xT[it * 16 + (lane & 15)][8 * (lane >> 4) + s];
                }
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int q = wave + 4 * j;
        if (q >= tiles) continue;
        const int ct = q / it_n, it = q % it_n;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = ct * 16 + 4 * (lane >> 4) + i, ci = it * 16 + (lane & 15);
            a.part[(((size_t)chunk * a.cout + co) * 27 + t) * a.cin + ci] = acc[j][i];
        }
    }
}

// grad[co][k] (+)= sum over chunks of part[chunk][co][k], k < 27 cin.
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float* __restrict__ part, int chunks, int count,
                                                      float* grad, int accumulate) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * count + i];
    grad[i] = accumulate ? grad[i] + s : s;
}

// ------------------------------------------------------------- elementwise
// out[v][c] = g[v][c] * (y[v][c] > 0) (ReLU backward), out contiguous [V][C].
__global__ void __launch_bounds__(256) k_relu_mask(View g, View y, int C, int V, bf16* out) {
    const size_t n = (size_t)V * C;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t v = i / C, c = i % C;
        const float yy = bf2f(y.p[v * y.stride + y.off + c]);
        out[i] = yy > 0.f ? g.p[v * g.stride + g.off + c] : f2bf(0.f);
    }
}

// 2x2x2 average pool of a view of a (2X, 2Y, 2Z) grid into out [X Y Z][C].
__global__ void __launch_bounds__(256) k_pool(View in, int C, int X, int Y, int Z, bf16* out) {
    const size_t n = (size_t)X * Y * Z * C;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i % C);
        const size_t v = i / C;
        const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((size_t)X * Y));
        float s = 0.f;
        for (int d = 0; d < 8; ++d) {
            const size_t u = (size_t)(2 * x + (d & 1)) + 2 * X * ((size_t)(2 * y + ((d >> 1) & 1)) +
                                                                  2 * (size_t)Y * (2 * z + (d >> 2)));
            s += bf2f(in.p[u * in.stride + in.off + c]);
        }
        out[i] = f2bf(s * 0.125f);
    }
}

// g[u][c] += gl[parent(u)][c] / 8 on the fine grid (2X, 2Y, 2Z) (pool backward).
__global__ void __launch_bounds__(256) k_pool_back_add(const bf16* __restrict__ gl, int C, int X, int Y, int Z,
                                                       View g) {
    const size_t n = (size_t)8 * X * Y * Z * C;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i % C);
        const size_t u = i / C;
        const int x = (int)(u % (2 * X)), y = (int)((u / (2 * X)) % (2 * Y)), z = (int)(u / ((size_t)4 * X * Y));
        const size_t p = (size_t)(x / 2) + X * ((size_t)(y / 2) + (size_t)Y * (z / 2));
        bf16* dst = g.p + u * g.stride + g.off + c;
        *dst = f2bf(bf2f(*dst) + bf2f(gl[p * C + c]) * 0.125f);
    }
}

// Nearest upsample of in [X Y Z][C] onto the fine grid, into a view.
__global__ void __launch_bounds__(256) k_upsample(const bf16* __restrict__ in, int C, int X, int Y, int Z, View out) {
    const size_t n = (size_t)8 * X * Y * Z * C;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i % C);
        const size_t u = i / C;
        const int x = (int)(u % (2 * X)), y = (int)((u / (2 * X)) % (2 * Y)), z = (int)(u / ((size_t)4 * X * Y));
        const size_t p = (size_t)(x / 2) + X * ((size_t)(y / 2) + (size_t)Y * (z / 2));
        out.p[u * out.stride + out.off + c] = in[p * C + c];
    }
}

// gl[p][c] = sum of the 8 children of p in a view of the fine grid (upsample backward).
__global__ void __launch_bounds__(256) k_upsample_back(View g, int C, int X, int Y, int Z, bf16* gl) {
    const size_t n = (size_t)X * Y * Z * C;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i % C);
        const size_t v = i / C;
        const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((size_t)X * Y));
        float s = 0.f;
        for (int d = 0; d < 8; ++d) {
            const size_t u = (size_t)(2 * x + (d & 1)) + 2 * X * ((size_t)(2 * y + ((d >> 1) & 1)) +
                                                                  2 * (size_t)Y * (2 * z + (d >> 2)));
            s += bf2f(g.p[u * g.stride + g.off + c]);
        }
        gl[i] = f2bf(s);
    }
}

// Network input: channel 0 = field * scale, channel 1 = boundary mask (1 at
// cells with fewer than 6 in-domain neighbours), channels 2..15 = 0.
__global__ void __launch_bounds__(256) k_unet_input(const float* __restrict__ field, float scale, int n,
                                                    bf16* __restrict__ x0) {
    const size_t V = (size_t)n * n * n;
    for (size_t v = blockIdx.x * 256ull + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((size_t)n * n));
        const bool wall = x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
        bf16* row = x0 + v * kUnetIn;
        row[0] = f2bf(field[v] * scale);
        row[1] = f2bf(wall ? 1.f : 0.f);
        for (int c = 2; c < kUnetIn; ++c) row[c] = f2bf(0.f);
    }
}

// out[v] = scale * sum_c d0[v][c] wh[c] (1x1x1 head, FP32 weights).
__global__ void __launch_bounds__(256) k_head(const bf16* __restrict__ d0, const float* __restrict__ wh, int C,
                                              size_t V, float scale, float* __restrict__ out) {
    for (size_t v = blockIdx.x * 256ull + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        float s = 0.f;
        for (int c = 0; c < C; ++c) s += bf2f(d0[v * C + c]) * wh[c];
        out[v] = s * scale;
    }
}

// Loss pieces: dout[v] = 2 (out - target) / V and per-block sums of
// (out - target)^2 in FP64 (deterministic second pass on the host).
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ out, const float* __restrict__ target, size_t V,
                                             float* __restrict__ dout, double* __restrict__ part) {
    __shared__ double lds[16 * 16];
    double acc[1] = {0.0};
    for (size_t v = blockIdx.x * 256ull + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const float d = out[v] - target[v];
        dout[v] = 2.f * d / (float)V;
        acc[0] += (double)d * (double)d;
    }
    const Red ops[1] = {Red::Sum};
    block_reduce<1>(acc, ops, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Head backward: g_d0[v][c] = dout[v] wh[c] (bf16), and per-block partial
// sums of dout[v] d0[v][c] for dwh (deterministic).
__global__ void __launch_bounds__(256) k_head_back(const bf16* __restrict__ d0, const float* __restrict__ wh,
                                                   const float* __restrict__ dout, int C, size_t V,
                                                   bf16* __restrict__ gd0, double* __restrict__ part) {
    __shared__ double lds[16 * 16];
    double acc[kUnetHead];
    for (int c = 0; c < kUnetHead; ++c) acc[c] = 0.0;
    for (size_t v = blockIdx.x * 256ull + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const float g = dout[v];
        for (int c = 0; c < C; ++c) {
            gd0[v * C + c] = f2bf(g * wh[c]);
            acc[c] += (double)g * (double)bf2f(d0[v * C + c]);
        }
    }
    Red ops[kUnetHead];
    for (int c = 0; c < kUnetHead; ++c) ops[c] = Red::Sum;
    block_reduce<kUnetHead>(acc, ops, lds);
    if (threadIdx.x == 0)
        for (int c = 0; c < kUnetHead; ++c) part[(size_t)blockIdx.x * kUnetHead + c] = acc[c];
}

// Adam (Kingma & Ba) on FP32 masters: m, v moments; bias-corrected step.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ vv, size_t count, float gscale, float lr, float beta1,
                                              float beta2, float eps, float c1, float c2) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) {
        const float gi = g[i] * gscale;
        const float mi = beta1 * m[i] + (1.f - beta1) * gi;
        const float vi = beta2 * vv[i] + (1.f - beta2) * gi * gi;
        m[i] = mi;
        vv[i] = vi;
        p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 copies of one layer's FP32 weights: w [cout][kp] (zero-padded k) for
// the forward pass and wt [cin][kpt] with wt[ci][t' cout + co] =
// w[co][(26 - t') cin + ci] (mirrored taps) for the backward-data pass.
__global__ void __launch_bounds__(256) k_pack_weights(const float* __restrict__ w32, int cout, int cin, int kp, int kpt,
                                                      bf16* __restrict__ w, bf16* __restrict__ wt) {
    const int K = 27 * cin, KT = 27 * cout;
    const size_t n1 = (size_t)cout * kp, n2 = (size_t)cin * kpt;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n1 + n2; i += (size_t)gridDim.x * 256) {
        if (i < n1) {
            const int co = (int)(i / kp), k = (int)(i % kp);
            w[i] = f2bf(k < K ? w32[(size_t)co * K + k] : 0.f);
        } else {
            const size_t j = i - n1;
            const int ci = (int)(j / kpt), k = (int)(j % kpt);
            float val = 0.f;
            if (k < KT) {
                const int tp = k / cout, co = k % cout;
                val = w32[(size_t)co * K + (26 - tp) * cin + ci];
            }
            wt[j] = f2bf(val);
        }
    }
}

// He-normal initial weights, std = sqrt(2 / fan_in), from the counter-based
// generator (stream = layer index).
__global__ void __launch_bounds__(256) k_init_weights(float* __restrict__ w, size_t count, float stdev, uint64_t seed,
                                                      uint32_t stream) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (size_t)gridDim.x * 256)
        w[i] = stdev * noise_normal(seed, stream, (uint32_t)i);
}

int grid_for(size_t n) { return (int)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 8192); }

template <int NT>
hipError_t conv_split(hipStream_t s, const ConvArgs& a, int tiles, int S) {
    const dim3 g(tiles);
    switch (S) {
    case 2: hipLaunchKernelGGL((k_conv3_split<NT, 2>), g, dim3(128), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_conv3_split<NT, 4>), g, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_conv3_split<NT, 8>), g, dim3(512), 0, s, a); break;
    default:
        // (S - 1) NT 2 KiB of partials: 16 waves fit LDS up to NT = 5
        if constexpr (NT <= 5) hipLaunchKernelGGL((k_conv3_split<NT, 16>), g, dim3(1024), 0, s, a);
        else hipLaunchKernelGGL((k_conv3_split<NT, 8>), g, dim3(512), 0, s, a);
        break;
    }
    return hipGetLastError();
}

}  // namespace

// Waves a convolution should occupy: with fewer 32-voxel tiles than this,
// the k-steps are split over up to 16 waves per tile.
constexpr int kConvTargetWaves = 2048;

hipError_t unet_conv(hipStream_t s, const ConvArgs& a) {
    const int V = a.X * a.Y * a.Z;
    const int tiles = (V + kConvRows - 1) / kConvRows, steps = a.kp / 32;
    if (tiles < kConvTargetWaves && steps >= 4) {
        int S = 2;
        while (S < 16 && tiles * S < kConvTargetWaves && 2 * S <= steps) S *= 2;
        switch (a.cout / 16) {
        case 1: return conv_split<1>(s, a, tiles, S);
        case 2: return conv_split<2>(s, a, tiles, S);
        case 3: return conv_split<3>(s, a, tiles, S);
        case 4: return conv_split<4>(s, a, tiles, S);
        case 5: return conv_split<5>(s, a, tiles, S);
        case 6: return conv_split<6>(s, a, tiles, S);
        case 7: return conv_split<7>(s, a, tiles, S);
        default: return hipErrorInvalidValue;
        }
    }
    const dim3 g((V + kConvWaves * kConvRows - 1) / (kConvWaves * kConvRows)), b(64 * kConvWaves);
    switch (a.cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv3<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_conv3<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_conv3<3>, g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_conv3<4>, g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_conv3<5>, g, b, 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_conv3<6>, g, b, 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_conv3<7>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int unet_wgrad_chunks(int V) { return (V + kWgradVoxels - 1) / kWgradVoxels; }

hipError_t unet_wgrad(hipStream_t s, const WgradArgs& a, float* grad, int accumulate) {
    if (a.cout > kWgradMaxCout || a.cin > kWgradMaxCin || a.cout % 16 || a.cin % 16 ||
        (a.cout / 16) * (a.cin / 16) > 28)
        return hipErrorInvalidValue;
    const int V = a.X * a.Y * a.Z, chunks = unet_wgrad_chunks(V);
    hipLaunchKernelGGL(k_wgrad, dim3(chunks, 27), dim3(256), 0, s, a);
    const int count = a.cout * 27 * a.cin;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((count + 255) / 256), dim3(256), 0, s, a.part, chunks, count, grad,
                       accumulate);
    return hipGetLastError();
}

hipError_t unet_relu_mask(hipStream_t s, View g, View y, int C, int V, bf16* out) {
    hipLaunchKernelGGL(k_relu_mask, dim3(grid_for((size_t)V * C)), dim3(256), 0, s, g, y, C, V, out);
    return hipGetLastError();
}
hipError_t unet_pool(hipStream_t s, View in, int C, int X, int Y, int Z, bf16* out) {
    hipLaunchKernelGGL(k_pool, dim3(grid_for((size_t)X * Y * Z * C)), dim3(256), 0, s, in, C, X, Y, Z, out);
    return hipGetLastError();
}
hipError_t unet_pool_back_add(hipStream_t s, const bf16* gl, int C, int X, int Y, int Z, View g) {
    hipLaunchKernelGGL(k_pool_back_add, dim3(grid_for((size_t)8 * X * Y * Z * C)), dim3(256), 0, s, gl, C, X, Y, Z, g);
    return hipGetLastError();
}
hipError_t unet_upsample(hipStream_t s, const bf16* in, int C, int X, int Y, int Z, View out) {
    hipLaunchKernelGGL(k_upsample, dim3(grid_for((size_t)8 * X * Y * Z * C)), dim3(256), 0, s, in, C, X, Y, Z, out);
    return hipGetLastError();
}
hipError_t unet_upsample_back(hipStream_t s, View g, int C, int X, int Y, int Z, bf16* gl) {
    hipLaunchKernelGGL(k_upsample_back, dim3(grid_for((size_t)X * Y * Z * C)), dim3(256), 0, s, g, C, X, Y, Z, gl);
    return hipGetLastError();
}
// ---- the coarse grid the network can run on ----
// Cell-centred restriction: the mean of each f^3 block of an n^3 field
// (n = m f), summed in x-fastest order.
// Sum of the f^3 fine cells of coarse cell (X, Y, Z), in z, y, x order. F is
// f when it is 2, 4 or 8 (0 otherwise): the loops unroll and every load issues
// before the sum's first add.
template <int F>
__device__ __forceinline__ float restrict_cell(int m, int f, const float* __restrict__ fine, int X, int Y, int Z) {
    float sum = 0.f;
    if constexpr (F > 0) {
        const int n = m * F;
#pragma unroll
        for (int dz = 0; dz < F; ++dz)
#pragma unroll
            for (int dy = 0; dy < F; ++dy) {
                const float* row = fine + idC(n, X * F, Y * F + dy, Z * F + dz);
#pragma unroll
                for (int dx = 0; dx < F; ++dx) sum += row[dx];
            }
    } else {
        const int n = m * f;
        for (int dz = 0; dz < f; ++dz)
            for (int dy = 0; dy < f; ++dy) {
                const float* row = fine + idC(n, X * f, Y * f + dy, Z * f + dz);
                for (int dx = 0; dx < f; ++dx) sum += row[dx];
            }
    }
    return sum;
}
// One lane per coarse cell: x along the block's 64 lanes, y and z from the
// grid (grid (ceil(m / 64), m, m)), so no index divides and every coarse row
// is its own block.
template <int F>
__global__ void __launch_bounds__(64) k_restrict_mean(int m, int f, float scale, const float* __restrict__ fine,
                                                      float* __restrict__ coarse) {
    const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y, Z = blockIdx.z;
    if (X < m) coarse[idC(m, X, Y, Z)] = restrict_cell<F>(m, f, fine, X, Y, Z) * scale;
}
// Cell-centred trilinear prolongation of an m^3 field to n^3 = (m f)^3:
// fine cell x sits at coarse coordinate (x + 1/2) / f - 1/2, clamped to
// [0, m - 1] (constant extension beyond the outermost coarse centres).
__device__ __forceinline__ void prolong_axis(int x, int f, int m, int& i0, int& i1, float& t) {
    float p = ((float)x + 0.5f) / (float)f - 0.5f;
    p = fminf(fmaxf(p, 0.f), (float)(m - 1));
    i0 = (int)p;
    i1 = min(i0 + 1, m - 1);
    t = p - (float)i0;
}
// The prolonged value at fine cell (x, y, z).
__device__ __forceinline__ float prolong_cell(int m, int f, const float* __restrict__ coarse, int x, int y, int z) {
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    prolong_axis(x, f, m, x0, x1, tx);
    prolong_axis(y, f, m, y0, y1, ty);
    prolong_axis(z, f, m, z0, z1, tz);
    auto at = [&](int a, int b, int c) { return coarse[idC(m, a, b, c)]; };
    const float c00 = at(x0, y0, z0) * (1.f - tx) + at(x1, y0, z0) * tx;
    const float c10 = at(x0, y1, z0) * (1.f - tx) + at(x1, y1, z0) * tx;
    const float c01 = at(x0, y0, z1) * (1.f - tx) + at(x1, y0, z1) * tx;
    const float c11 = at(x0, y1, z1) * (1.f - tx) + at(x1, y1, z1) * tx;
    const float c0 = c00 * (1.f - ty) + c10 * ty, c1 = c01 * (1.f - ty) + c11 * ty;
    return c0 * (1.f - tz) + c1 * tz;
}
// Fine cells on blocks of 64 x 4 lanes over grid (ceil(n / 64), ceil(n / 4),
// ceil(n / kProlongZ)); each lane walks kProlongZ cells along z.
constexpr int kProlongZ = 8;
__global__ void __launch_bounds__(256) k_prolong_linear(int m, int f, const float* __restrict__ coarse,
                                                        float* __restrict__ fine) {
    const int n = m * f;
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= n || y >= n) return;
    const int z1 = min(n, (int)(blockIdx.z + 1) * kProlongZ);
    for (int z = blockIdx.z * kProlongZ; z < z1; ++z) fine[idC(n, x, y, z)] = prolong_cell(m, f, coarse, x, y, z);
}

// RMS of the compatible residual of a record (its mean is out of reach).
__device__ __forceinline__ double record_rms(const ProjState* ps, double count) {
    const double v = ps->res2 - ps->res_sum * ps->res_sum / count;
    return sqrt((v > 0.0 ? v : 0.0) / count);
}
template <int F>
__global__ void __launch_bounds__(64) k_restrict_rms(int m, int f, const ProjState* __restrict__ ps, double count,
                                                     const float* __restrict__ fine, float* __restrict__ coarse) {
    const double rms = record_rms(ps, count);
    const float scale = rms > 0.0 ? (float)(1.0 / rms) / (float)(f * f * f) : 0.f;
    const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y, Z = blockIdx.z;
    if (X < m) coarse[idC(m, X, Y, Z)] = restrict_cell<F>(m, f, fine, X, Y, Z) * scale;
}
// The record's scale (an FP64 divide and square root) is evaluated once per
// block and shared through LDS, not once per fine cell.
__global__ void __launch_bounds__(256) k_prolong_rms(int m, int f, const ProjState* __restrict__ ps, double count,
                                                     const float* __restrict__ coarse, float* __restrict__ fine) {
    __shared__ float scale_s;
    if (threadIdx.x == 0 && threadIdx.y == 0) scale_s = (float)(record_rms(ps, count) * (double)(f * f));
    __syncthreads();
    const float scale = scale_s;
    const int n = m * f;
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= n || y >= n) return;
    const int z1 = min(n, (int)(blockIdx.z + 1) * kProlongZ);
    for (int z = blockIdx.z * kProlongZ; z < z1; ++z)
        fine[idC(n, x, y, z)] = prolong_cell(m, f, coarse, x, y, z) * scale;
}
namespace {
dim3 coarse_grid(int m) { return dim3((unsigned)(m + 63) / 64, (unsigned)m, (unsigned)m); }
dim3 fine_grid(int n) {
    return dim3((unsigned)(n + 63) / 64, (unsigned)(n + 3) / 4, (unsigned)(n + kProlongZ - 1) / kProlongZ);
}
}  // namespace
hipError_t restrict_rms(hipStream_t s, int m, int f, const ProjState* ps, const float* fine, float* coarse) {
    const double count = (double)m * f * m * f * m * f;
    const dim3 g = coarse_grid(m), b(64);
    switch (f) {
    case 2: hipLaunchKernelGGL(k_restrict_rms<2>, g, b, 0, s, m, f, ps, count, fine, coarse); break;
    case 4: hipLaunchKernelGGL(k_restrict_rms<4>, g, b, 0, s, m, f, ps, count, fine, coarse); break;
    case 8: hipLaunchKernelGGL(k_restrict_rms<8>, g, b, 0, s, m, f, ps, count, fine, coarse); break;
    default: hipLaunchKernelGGL(k_restrict_rms<0>, g, b, 0, s, m, f, ps, count, fine, coarse); break;
    }
    return hipGetLastError();
}
hipError_t prolong_rms(hipStream_t s, int m, int f, const ProjState* ps, const float* coarse, float* fine) {
    const int n = m * f;
    hipLaunchKernelGGL(k_prolong_rms, fine_grid(n), dim3(64, 4), 0, s, m, f, ps, (double)n * n * n, coarse, fine);
    return hipGetLastError();
}

hipError_t restrict_mean(hipStream_t s, int m, int f, float scale, const float* fine, float* coarse) {
    const dim3 g = coarse_grid(m), b(64);
    const float sc = scale / (float)(f * f * f);
    switch (f) {
    case 2: hipLaunchKernelGGL(k_restrict_mean<2>, g, b, 0, s, m, f, sc, fine, coarse); break;
    case 4: hipLaunchKernelGGL(k_restrict_mean<4>, g, b, 0, s, m, f, sc, fine, coarse); break;
    case 8: hipLaunchKernelGGL(k_restrict_mean<8>, g, b, 0, s, m, f, sc, fine, coarse); break;
    default: hipLaunchKernelGGL(k_restrict_mean<0>, g, b, 0, s, m, f, sc, fine, coarse); break;
    }
    return hipGetLastError();
}
hipError_t prolong_linear(hipStream_t s, int m, int f, const float* coarse, float* fine) {
    hipLaunchKernelGGL(k_prolong_linear, fine_grid(m * f), dim3(64, 4), 0, s, m, f, coarse, fine);
    return hipGetLastError();
}

hipError_t unet_input(hipStream_t s, const float* field, float scale, int n, bf16* x0) {
    hipLaunchKernelGGL(k_unet_input, dim3(grid_for((size_t)n * n * n)), dim3(256), 0, s, field, scale, n, x0);
    return hipGetLastError();
}
hipError_t unet_head(hipStream_t s, const bf16* d0, const float* wh, int C, size_t V, float scale, float* out) {
    hipLaunchKernelGGL(k_head, dim3(grid_for(V)), dim3(256), 0, s, d0, wh, C, V, scale, out);
    return hipGetLastError();
}
int unet_reduce_blocks(size_t V) { return (int)std::min<size_t>(std::max<size_t>((V + 255) / 256, 1), 1024); }
hipError_t unet_mse(hipStream_t s, const float* out, const float* target, size_t V, float* dout, double* part) {
    hipLaunchKernelGGL(k_mse, dim3(unet_reduce_blocks(V)), dim3(256), 0, s, out, target, V, dout, part);
    return hipGetLastError();
}
hipError_t unet_head_back(hipStream_t s, const bf16* d0, const float* wh, const float* dout, int C, size_t V,
                          bf16* gd0, double* part) {
    hipLaunchKernelGGL(k_head_back, dim3(unet_reduce_blocks(V)), dim3(256), 0, s, d0, wh, dout, C, V, gd0, part);
    return hipGetLastError();
}
hipError_t unet_adam(hipStream_t s, float* p, const float* g, float* m, float* v, size_t count, float gscale, float lr,
                     float beta1, float beta2, float eps, float c1, float c2) {
    hipLaunchKernelGGL(k_adam, dim3(grid_for(count)), dim3(256), 0, s, p, g, m, v, count, gscale, lr, beta1, beta2, eps,
                       c1, c2);
    return hipGetLastError();
}
hipError_t unet_pack_weights(hipStream_t s, const float* w32, int cout, int cin, int kp, int kpt, bf16* w, bf16* wt) {
    hipLaunchKernelGGL(k_pack_weights, dim3(grid_for((size_t)cout * kp + (size_t)cin * kpt)), dim3(256), 0, s, w32,
                       cout, cin, kp, kpt, w, wt);
    return hipGetLastError();
}
hipError_t unet_init_weights(hipStream_t s, float* w, size_t count, float stdev, uint64_t seed, uint32_t stream) {
    hipLaunchKernelGGL(k_init_weights, dim3(grid_for(count)), dim3(256), 0, s, w, count, stdev, seed, stream);
    return hipGetLastError();
}

}  // namespace vgx
