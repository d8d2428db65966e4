// advect_hostcheck.cpp - CPU build of the device advection code, with every
// field read bounds-checked (test infrastructure).
//
// The advection kernels call vgx::sl_value / vgx::mc_value, which are
// __host__ __device__ templates over the field accessor. Here they run on the
// host over accessors that count out-of-range indices instead of reading
// them, so the kernels' indexing is checked on the CPU for any input
// (including NaN, infinite and huge velocities) before it runs on a GPU.
#include "../../vapor_amd/csrc/vgx_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace {
struct Checked {
    const float* p;
    long size;
    long* oob;
    float operator[](long i) const {
        if (i < 0 || i >= size) {
            ++*oob;
            return 0.f;
        }
        return p[i];
    }
};

template <int AXIS, bool POW2 = false>
long run(int n, float box, int method, float dt, float cfl, const float* u, const float* v, const float* w,
         const float* src, const float* tsrc, float* out) {
    long oob = 0;
    const long faces = (long)(n + 1) * n * n, cells = (long)n * n * n;
    const long count = AXIS < 0 ? cells : faces;
    const float h0 = box / (float)n;
    const vgx::Spacing<float, POW2> h{h0, POW2 ? 1.0f / h0 : 0.0f};
    const Checked cu{u, faces, &oob}, cv{v, faces, &oob}, cw{w, faces, &oob};
    std::vector<float> fwd(count);
    auto advect = [&](auto src_acc) {
        for (long i = 0; i < count; ++i) fwd[i] = vgx::sl_value<AXIS>(cu, cv, cw, src_acc, n, h, dt, (int)i);
        if (!method) {
            for (long i = 0; i < count; ++i) out[i] = fwd[i];
            return;
        }
        const Checked cf{fwd.data(), count, &oob};
        for (long i = 0; i < count; ++i)
            out[i] = vgx::mc_value<AXIS>(cu, cv, cw, src_acc, cf, n, h, box, dt, cfl, (int)i);
    };
    if (tsrc) advect(vgx::Src<true, Checked>{Checked{src, count, &oob}, Checked{tsrc, count, &oob}});
    else advect(vgx::Src<false, Checked>{Checked{src, count, &oob}, Checked{src, count, &oob}});
    return oob;
}
}  // namespace

extern "C" {
// field: -1 cell, 0..2 faces. tsrc may be null (reflected source 2 src - tsrc
// otherwise). Returns the number of out-of-range reads.
long hc_advect(int n, float box, int field, int method, float dt, float cfl, const float* u, const float* v,
               const float* w, const float* src, const float* tsrc, float* out) {
    switch (field) {
    case 0: return run<0>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    case 1: return run<1>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    case 2: return run<2>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    default: return run<-1>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    }
}
int hc_pow2_spacing(float h) { return vgx::pow2_spacing(h) ? 1 : 0; }

// The same with the power-of-two spacing path (box / n must be a power of two).
long hc_advect_pow2(int n, float box, int field, int method, float dt, float cfl, const float* u, const float* v,
                    const float* w, const float* src, const float* tsrc, float* out) {
    switch (field) {
    case 0: return run<0, true>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    case 1: return run<1, true>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    case 2: return run<2, true>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    default: return run<-1, true>(n, box, method, dt, cfl, u, v, w, src, tsrc, out);
    }
}
}

// ---- fused red-black sweep (k_sor_rb), run block by block, phase by phase ----
namespace {
struct CheckedRW {
    float* p;
    long size;
    long* oob;
    float dummy = 0.f;
    float& operator[](long i) {
        if (i < 0 || i >= size) {
            ++*oob;
            dummy = 0.f;
            return dummy;
        }
        return p[i];
    }
    float operator[](long i) const {
        if (i < 0 || i >= size) {
            ++*oob;
            return 0.f;
        }
        return p[i];
    }
};
}  // namespace

extern "C" {
// One k_sor_rb launch emulated on the host: every block's phases in order, all
// lanes of a phase before the next (the kernel's __syncthreads). Returns the
// out-of-range reads and writes; residual[7] receives the mode-1 partial sums
// over all blocks (max for field 1).
long hc_sor_rb(int n, float omega, float floor, float meanf, const float* b, const float* xin, float* xout,
               int mode, double* residual, double* out_sum) {
    long oob = 0;
    const long cells = (long)n * n * n;
    const Checked cb{b, cells, &oob}, cxin{xin, cells, &oob};
    CheckedRW cxout{xout, cells, &oob};
    std::vector<float> stage(vgx::kRbXWords), stage_b(vgx::kRbBWords);
    CheckedRW xs{stage.data(), vgx::kRbXWords, &oob};
    CheckedRW bs{stage_b.data(), vgx::kRbBWords, &oob};
    double total[7] = {0, 0, 0, 0, 0, 0, 0};
    double sum = 0.0;
    for (int block = 0; block < vgx::rb_blocks(n); ++block) {
        const vgx::RbTile t = vgx::rb_tile(n, block);
        if (!t.valid) continue;
        // unstaged LDS words hold NaN, so reading one shows in the result
        std::fill(stage.begin(), stage.end(), std::nanf(""));
        std::fill(stage_b.begin(), stage_b.end(), std::nanf(""));
        const bool vec = n % 4 == 0;
        const vgx::RbArgs a{omega, floor, 6.0f, 1.0f / 6.0f};
        for (int tid = 0; tid < vgx::kRbThreads; ++tid) {
            if (vec) vgx::rb_stage<true>(xs, bs, cxin, cb, meanf, n, t, tid);
            else vgx::rb_stage<false>(xs, bs, cxin, cb, meanf, n, t, tid);
        }
        if (mode == 1)
            for (int tid = 0; tid < vgx::kRbThreads; ++tid) {
                double acc[7] = {0, 0, 0, 0, 0, 0, 0};
                vgx::rb_residual(xs, bs, n, t, tid, acc);
                for (int k = 0; k < 7; ++k) total[k] = k == 1 ? (total[k] < acc[k] ? acc[k] : total[k]) : total[k] + acc[k];
            }
        // the kernel's schedule: fast divisor iff floor == 6, inner tiles, no updates for n == 1
        if (n > 1) {
            const bool inner = vgx::rb_inner(n, t), fast = floor == 6.0f;
            auto phases = [&](auto fast_c, auto inner_c) {
                constexpr bool F = decltype(fast_c)::value, I = decltype(inner_c)::value;
                for (int tid = 0; tid < vgx::kRbThreads; ++tid) vgx::rb_red<F, I>(xs, bs, a, n, t, tid);
                for (int tid = 0; tid < vgx::kRbThreads; ++tid) vgx::rb_black<F, I>(xs, bs, a, n, t, tid);
            };
            using T = std::true_type;
            using F = std::false_type;
            if (fast && inner) phases(T{}, T{});
            else if (fast) phases(T{}, F{});
            else if (inner) phases(F{}, T{});
            else phases(F{}, F{});
        }
        for (int tid = 0; tid < vgx::kRbThreads; ++tid)
            sum += vec ? vgx::rb_write<true>(xs, cxout, n, t, tid) : vgx::rb_write<false>(xs, cxout, n, t, tid);
    }
    if (residual)
        for (int k = 0; k < 7; ++k) residual[k] = total[k];
    if (out_sum) *out_sum = sum;
    return oob;
}

// The register path (k_sor_rb_reg) emulated the same way: per block, every
// lane's staging and right-hand-side loads (buffer loads at byte offsets,
// kRbOOB reading 0.0), then the phases. Only for rb_reg_fits(n).
long hc_sor_rb_reg(int n, float omega, float floor, float meanf, const float* b, const float* xin, float* xout,
                   double* out_sum) {
    long oob = 0;
    if (!vgx::rb_reg_fits(n)) return -1;
    const long cells = (long)n * n * n;
    const Checked cb{b, cells, &oob}, cxin{xin, cells, &oob};
    CheckedRW cxout{xout, cells, &oob};
    std::vector<float> stage(vgx::kRbXWords);
    CheckedRW xs{stage.data(), vgx::kRbXWords, &oob};
    std::vector<vgx::RbRegs> regs(vgx::kRbThreads);
    const vgx::RbArgs a{omega, floor, 6.0f, 1.0f / 6.0f};
    double sum = 0.0;
    for (int block = 0; block < vgx::rb_blocks(n); ++block) {
        const vgx::RbTile t = vgx::rb_tile(n, block);
        if (!t.valid) continue;
        std::fill(stage.begin(), stage.end(), std::nanf(""));
        for (int tid = 0; tid < vgx::kRbThreads; ++tid) {
            float4 v[vgx::kRbSIter];
            vgx::rb_fetch_x_reg(cxin, n, t, tid, v);
            vgx::rb_fetch_b_reg(regs[tid], cb, n, t, tid);
            vgx::rb_put_x_reg(xs, tid, v);
        }
        if (n > 1) {
            const bool inner = vgx::rb_inner(n, t), fast = floor == 6.0f;
            auto phases = [&](auto fast_c, auto inner_c) {
                constexpr bool F = decltype(fast_c)::value, I = decltype(inner_c)::value;
                for (int tid = 0; tid < vgx::kRbThreads; ++tid) vgx::rb_red_reg<F, I>(xs, regs[tid], meanf, a, n, t, tid);
                for (int tid = 0; tid < vgx::kRbThreads; ++tid) vgx::rb_black_reg<F, I>(xs, regs[tid], meanf, a, n, t, tid);
            };
            using T = std::true_type;
            using F = std::false_type;
            if (fast && inner) phases(T{}, T{});
            else if (fast) phases(T{}, F{});
            else if (inner) phases(F{}, T{});
            else phases(F{}, F{});
        }
        for (int tid = 0; tid < vgx::kRbThreads; ++tid) sum += vgx::rb_write<true>(xs, cxout, n, t, tid);
    }
    if (out_sum) *out_sum = sum;
    return oob;
}

// rb_tile through rb_geom's multipliers against the in-place divisions, for
// every block of every n from 1 to nmax; returns the number of mismatches.
long hc_rb_geom_check(int nmax) {
    long bad = 0;
    for (int n = 1; n <= nmax; ++n) {
        const vgx::RbGeom g = vgx::rb_geom(n);
        for (int block = 0; block < vgx::rb_blocks(n); ++block) {
            const vgx::RbTile a = vgx::rb_tile(g, block), b = vgx::rb_tile(n, block);
            if (a.valid != b.valid || (a.valid && (a.X0 != b.X0 || a.Y0 != b.Y0 || a.Z0 != b.Z0))) ++bad;
        }
    }
    return bad;
}

// Tiles covered by the fused sweep's grid: every cell exactly once.
long hc_rb_coverage(int n, int* counts) {
    long bad = 0;
    for (int block = 0; block < vgx::rb_blocks(n); ++block) {
        const vgx::RbTile t = vgx::rb_tile(n, block);
        if (!t.valid) continue;
        for (int z = t.Z0; z < t.Z0 + vgx::kRbTz; ++z)
            for (int y = t.Y0; y < t.Y0 + vgx::kRbTy; ++y)
                for (int x = t.X0; x < t.X0 + vgx::kRbTx; ++x) {
                    if (x >= n || y >= n || z >= n) continue;
                    if (x < 0 || y < 0 || z < 0) ++bad;
                    else ++counts[vgx::idC(n, x, y, z)];
                }
    }
    return bad;
}

// rb_div against IEEE division for every float a whose bit pattern is a
// multiple of `stride` (1 = all 2^32), divisor D, reciprocal 1.0f / D.
// Returns the number of mismatching finite quotients (NaN matches NaN).
long hc_rb_div_check(float D, unsigned stride, int threads) {
    const float r = 1.0f / D;
    std::vector<long> bad(threads, 0);
    std::vector<std::thread> pool;
    const uint64_t total = 1ull << 32;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            const uint64_t lo = total * k / threads, hi = total * (k + 1) / threads;
            volatile float dv = D;
            for (uint64_t u = (lo + stride - 1) / stride * stride; u < hi; u += stride) {
                const uint32_t bits = (uint32_t)u;
                float a;
                std::memcpy(&a, &bits, 4);
                const float want = a / dv, got = vgx::rb_div(a, D, r);
                uint32_t wb, gb;
                std::memcpy(&wb, &want, 4);
                std::memcpy(&gb, &got, 4);
                if (wb != gb && !(want != want && got != got)) ++bad[k];
            }
        });
    for (auto& th : pool) th.join();
    long sum = 0;
    for (long v : bad) sum += v;
    return sum;
}

// The device's counter-based normal variates for entries [0, count) of one
// stream, computed by the same __host__ __device__ code on the host.
void hc_noise_normal(uint64_t seed, uint32_t stream, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = vgx::noise_normal(seed, stream, i);
}
}
