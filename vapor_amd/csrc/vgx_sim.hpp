// vgx_sim.hpp - host-side state of one MAC 3D simulation: the projection
// workspace (RHS, pressures, records, tolerance chunks, refinement, learned
// start), kernel timing, and the resident step (MacGridFluidSolver3D::step).
// Shared by the single-scene C ABI (vgx_api.hip) and the batch (vgx_batch.hip).
#pragma once

#include "vgx_internal.hpp"
#include "vgx_host.hpp"
#include "vgx_unet.hpp"

#include "../../include/vapor_mac3d.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

namespace vgx_impl {

using namespace vgx_host;

// MacGridDomain3D's checks (MacGridDomain3D.cpp:7-23): n > 0, n^2 (n+1) within
// int, finite positive box and spacing.
inline void validate_grid(int n, float box) {
    if (n <= 0) invalid("MAC grid resolution must be positive");
    const size_t nn = (size_t)n, limit = (size_t)std::numeric_limits<int>::max();
    if (nn > limit / nn || nn * nn > limit / (nn + 1))
        throw Failure{VGX_ERANGE, "MAC grid exceeds the supported index range"};
    if (!std::isfinite(box) || box <= 0.f || box / (float)n <= 0.f)
        invalid("MAC grid requires a finite positive box size and spacing");
}

// Projection policy, the mirror of ProjectionOptions / LinearSolveOptions.
struct SolvePolicy {
    float dt = 1.f / 60.f;
    float omega = 1.95f;
    float floor = 6.f;
    int max_iterations = 0;
    int check_interval = 8;
    int final_interval = 0;  // tolerance solves: checks this often once the trend says the next chunk converges
    double abs_tol = 1e-6, rel_tol = 1e-5;
    bool fixed = true;
    bool use_l2 = false;
    double l2_tol = 1e-5;
    bool warm = false;
    int refinement = 0;
    bool graphs = false;  // replay tolerance chunks from captured HIP graphs
};

// ||r - mean(r)||_2 / ||b - mean(b)||_2 from the sums a residual check
// reduces: the Neumann matrix's null space (constants) makes the residual's
// mean unreachable, so the solve is measured on the compatible part.
inline double compat_rel(double r2, double rsum, double b2, double bsum, double count) {
    const double num = r2 - rsum * rsum / count, den = b2 - bsum * bsum / count;
    if (!(den > 0.0)) return 0.0;
    return std::sqrt(std::max(num, 0.0) / den);
}
inline double compat_norm2(double b2, double bsum, double count) { return b2 - bsum * bsum / count; }

// Timing with HIP events attached to each dispatch (hipExtLaunchKernelGGL):
// while a Timed scope is open, every kernel launched on this thread records
// its own start/stop, so a kind's total is kernel execution time. The sweeps,
// thousands per step, are timed instead in runs: two events recorded on the
// stream around a run of back-to-back sweeps give their stream time, the gaps
// between the launches included.
struct Timer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    struct Range {
        int kind;
        uint64_t count;
        hipEvent_t start, stop;
    };
    std::vector<Range> ranges;
    double ms[vgx::K_COUNT] = {};          // summed over the sampled launches
    uint64_t launches[vgx::K_COUNT] = {};  // all launches while enabled
    uint64_t sampled[vgx::K_COUNT] = {};
    // Sweeps timed one dispatch at a time (the batch's): one launch in
    // kSweepStride (the mean of the sample estimates the mean launch), so the
    // event bookkeeping does not slow the host loop that feeds the GPU.
    static constexpr uint64_t kSweepStride = 16;

    hipEvent_t take_event() {
        if (pool.empty()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    vgx::Probe probe(int kind) {
        const uint64_t stride = kind == vgx::K_SOR ? kSweepStride : 1;
        if (launches[kind]++ % stride != 0) return {};
        vgx::Probe p{take_event(), take_event()};
        pending.push_back({kind, {p.start, p.stop}});
        return p;
    }
    // A run of `count` launches of `kind` on stream s, between range_begin and
    // range_end (both return or take null when timing is off).
    hipEvent_t range_begin(hipStream_t s) {
        if (!enabled) return nullptr;
        hipEvent_t e = take_event();
        check(hipEventRecord(e, s), "hipEventRecord");
        return e;
    }
    void range_end(hipStream_t s, hipEvent_t start, int kind, uint64_t count) {
        if (!enabled || !start) return;
        hipEvent_t e = take_event();
        check(hipEventRecord(e, s), "hipEventRecord");
        ranges.push_back({kind, count, start, e});
        launches[kind] += count;
    }
    // Call after the stream is synchronised.
    void flush() {
        for (auto& r : ranges) {
            float elapsed = 0.f;
            check(hipEventElapsedTime(&elapsed, r.start, r.stop), "hipEventElapsedTime");
            ms[r.kind] += elapsed;
            sampled[r.kind] += r.count;
            pool.push_back(r.start);
            pool.push_back(r.stop);
        }
        ranges.clear();
        for (auto& p : pending) {
            float elapsed = 0.f;
            check(hipEventElapsedTime(&elapsed, p.second.first, p.second.second), "hipEventElapsedTime");
            ms[p.first] += elapsed;
            sampled[p.first] += 1;
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        pending.clear();
    }
    // Launches that ran without a probe (inside a replayed graph).
    void count(int kind, uint64_t launches_run) {
        if (enabled) launches[kind] += launches_run;
    }
    // Estimated total device milliseconds of a kind.
    double total_ms(int kind) const {
        return sampled[kind] ? ms[kind] * (double)launches[kind] / (double)sampled[kind] : 0.0;
    }
    void reset() {
        for (auto& v : ms) v = 0.0;
        for (auto& v : launches) v = 0;
        for (auto& v : sampled) v = 0;
    }
    ~Timer() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        for (auto& r : ranges) {
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Scope whose kernel launches are timed under `kind` (when timing is on).
struct Timed final : vgx::ProbeSource {
    Timer& t;
    int kind;
    vgx::ProbeSource* previous;
    Timed(Timer& timer, hipStream_t, int k) : t(timer), kind(k), previous(vgx::current_probe_source()) {
        if (t.enabled) vgx::current_probe_source() = this;
    }
    vgx::Probe take() override { return t.probe(kind); }
    ~Timed() override { vgx::current_probe_source() = previous; }
};

// Device workspace of the projections: RHS, reduction partials, and per
// projection slot (0: the step's last projection, 1: the first projection of
// a reflection step) a pressure ping-pong pair and a device record. Each slot
// keeps its own pressure history, which is what a warm start reads: the two
// projections of a reflection step solve for pressures of roughly opposite
// sign (the reflected velocity carries the negated divergence).
struct ProjectionWorkspace {
    static constexpr int kInner = 2;  // record of the refinement's correction solves
    int n = 0;
    float box = 0.f;
    hipStream_t stream = nullptr;
    float* b = nullptr;
    float* p[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int p_cur[2] = {0, 0};
    double* partials = nullptr;
    vgx::ProjState* ps_dev = nullptr;   // [3]: slot 0, slot 1, refinement corrections
    vgx::ProjState* ps_host = nullptr;  // [3], pinned
    // FP64 iterative refinement (allocated on first use).
    double* x64 = nullptr;
    float* r32 = nullptr;
    float* e32[2] = {nullptr, nullptr};
    // the coarse grid of a learned start or of recorded pairs: restricted
    // residual and prediction (m^3), and the fine starting pressure of a
    // recorded solve
    int coarse_m = 0;
    float* rc = nullptr;
    float* ec = nullptr;
    float* rec_prev = nullptr;
    uint64_t sweeps = 0;
    // Tolerance chunks run ahead of the host: each chunk's check kernel
    // writes its record into a pinned ring slot, marked by an event the host
    // waits on. Two chunks in flight keep the device busy (the host turns a
    // record around in tens of microseconds, a chunk runs for hundreds); each
    // chunk queued beyond the converged one still costs its launches.
    static constexpr int kRing = 8, kLookahead = 2;
    vgx::ProjState* ring = nullptr;      // pinned [kRing]
    vgx::ProjState* ring_dev = nullptr;  // the same memory as the device sees it
    hipEvent_t ring_ev[kRing] = {};
    // The learned start's second stream (default priority, created on first
    // use): the judge's plain candidate sweeps there while the network runs
    // on `stream`; fork and join are events.
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    double* judge_partials = nullptr;  // the learned candidate's residual partials (the judge)

    void allocate(int n_, float box_, hipStream_t s) {
        n = n_;
        box = box_;
        stream = s;
        const size_t cells = (size_t)n * n * n;
        b = device_alloc<float>(cells, "hipMalloc(b)");
        for (auto& slot : p)
            for (auto& q : slot) q = device_alloc<float>(cells, "hipMalloc(p)");
        partials = device_alloc<double>(vgx::partials_capacity(n), "hipMalloc(partials)");
        ps_dev = device_alloc<vgx::ProjState>(3, "hipMalloc(projection record)");
        // The correction solves' record keeps rhs_meanf = 0 (their right-hand
        // side is already a residual); zero all records once here.
        check(hipMemset(ps_dev, 0, 3 * sizeof(vgx::ProjState)), "hipMemset(records)");
        void* host = nullptr;
        check(hipHostMalloc(&host, 3 * sizeof(vgx::ProjState), hipHostMallocDefault), "hipHostMalloc");
        ps_host = static_cast<vgx::ProjState*>(host);
        std::memset(ps_host, 0, 3 * sizeof(vgx::ProjState));
        check(hipHostMalloc(&host, kRing * sizeof(vgx::ProjState), hipHostMallocDefault), "hipHostMalloc");
        ring = static_cast<vgx::ProjState*>(host);
        void* dev = nullptr;
        check(hipHostGetDevicePointer(&dev, host, 0), "hipHostGetDevicePointer");
        ring_dev = static_cast<vgx::ProjState*>(dev);
        for (auto& e : ring_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    void ensure_refinement_buffers() {
        if (x64) return;
        const size_t cells = (size_t)n * n * n;
        x64 = device_alloc<double>(cells, "hipMalloc(x64)");
        r32 = device_alloc<float>(cells, "hipMalloc(r32)");
        e32[0] = device_alloc<float>(cells, "hipMalloc(e32)");
        e32[1] = device_alloc<float>(cells, "hipMalloc(e32)");
    }
    // A network on grid m serves an n^3 solve when it runs on n^3 itself
    // (m = 0, n % 8 == 0) or on a coarser m^3 that divides it; a dataset of
    // grid m records from an n^3 solve on the same terms.
    bool learnable(int m) const { return m == 0 ? n % 8 == 0 : recordable(m); }
    bool recordable(int m) const { return m > 0 && m % 8 == 0 && m <= n && n % m == 0; }
    void ensure_coarse_buffers(int m) {
        if (coarse_m == m) return;
        for (float* q : {rc, ec})
            if (q) check(hipFree(q), "hipFree(coarse)");
        const size_t cells = (size_t)m * m * m;
        rc = device_alloc<float>(cells, "hipMalloc(coarse)");
        ec = device_alloc<float>(cells, "hipMalloc(coarse)");
        coarse_m = m;
    }
    void release() {
        release_graphs();
        for (void* q : {(void*)b, (void*)partials, (void*)ps_dev, (void*)x64, (void*)r32, (void*)e32[0],
                        (void*)e32[1], (void*)rc, (void*)ec, (void*)rec_prev, (void*)judge_partials})
            if (q) (void)hipFree(q);
        rc = ec = rec_prev = nullptr;
        judge_partials = nullptr;
        coarse_m = 0;
        if (ring) (void)hipHostFree(ring);
        ring = nullptr;
        for (hipEvent_t* e : {&fork_ev, &join_ev})
            if (*e) {
                (void)hipEventDestroy(*e);
                *e = nullptr;
            }
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
        ring_dev = nullptr;
        for (auto& e : ring_ev)
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        for (auto& slot : p)
            for (auto& q : slot) {
                if (q) (void)hipFree(q);
                q = nullptr;
            }
        if (ps_host) (void)hipHostFree(ps_host);
        b = nullptr;
        partials = nullptr;
        ps_dev = nullptr;
        ps_host = nullptr;
        x64 = nullptr;
        r32 = e32[0] = e32[1] = nullptr;
    }
    void clear(Timer& timer) {
        const size_t cells = (size_t)n * n * n;
        Timed t(timer, stream, vgx::K_FILL);
        for (auto& slot : p)
            for (auto& q : slot) check(vgx::launch_fill_f32(stream, q, cells, 0.f), "fill");
        check(vgx::launch_fill_f32(stream, b, cells, 0.f), "fill");
        p_cur[0] = p_cur[1] = 0;
    }
    float h() const { return box / (float)n; }
    float* current(int slot) const { return p[slot][p_cur[slot]]; }

    void read_record(int rec) {
        check(hipMemcpyAsync(ps_host + rec, ps_dev + rec, sizeof(vgx::ProjState), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(record)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // FP64 residual check of rhs' - A x into record `rec`.
    void residual(Timer& timer, int rec, const float* rhs, const float* x) {
        int blocks = 0;
        {
            Timed t(timer, stream, vgx::K_RESIDUAL);
            check(vgx::launch_residual(stream, n, rhs, x, ps_dev + rec, partials, &blocks), "residual");
        }
        Timed t(timer, stream, vgx::K_REDUCE);
        check(vgx::launch_finalize(stream, vgx::Finalize::Residual, partials, blocks, n, ps_dev + rec, nullptr),
              "finalize");
    }
    void residual(Timer& timer, int slot) { residual(timer, slot, b, current(slot)); }

    // `count` full red-black sweeps on A x = rhs' (rhs' = rhs - record's
    // rhs_meanf), ping-ponging between buf[cur] and buf[1 - cur]; cur follows
    // the iterate. With sum_last the last sweep also reduces sum(x) into the
    // record (Finalize::XSum).
    void sweeps_on(Timer& timer, int rec, const SolvePolicy& pol, int count, bool sum_last, const float* rhs,
                   float* const buf[2], int& cur, hipStream_t on = nullptr) {
        hipStream_t st = on ? on : stream;
        // the sweeps are timed as one run (Timer), the summing last one as well
        hipEvent_t run = count > 0 ? timer.range_begin(st) : nullptr;
        for (int it = 0; it < count; ++it) {
            const bool last = sum_last && it + 1 == count;
            int blocks = 0;
            check(vgx::launch_sor_rb(st, n, pol.omega, pol.floor, rhs, buf[cur], buf[1 - cur], ps_dev + rec, nullptr,
                                     last ? 2 : 0, partials, &blocks),
                  "sor");
            if (it + 1 == count) timer.range_end(st, run, vgx::K_SOR, (uint64_t)count);
            cur = 1 - cur;
            if (last) {
                Timed t(timer, st, vgx::K_REDUCE);
                check(vgx::launch_finalize(st, vgx::Finalize::XSum, partials, blocks, n, ps_dev + rec, nullptr),
                      "finalize");
            }
        }
        sweeps += (uint64_t)count;
    }
    void sweeps_fixed(Timer& timer, int slot, const SolvePolicy& pol, int count, bool sum_last) {
        sweeps_on(timer, slot, pol, count, sum_last, b, p[slot], p_cur[slot]);
    }

    // Sweeps per tolerance chunk: the check interval rounded up to an even
    // count, so a chunk leaves the iterate in the buffer it started in.
    static int chunk_of(const SolvePolicy& pol) { return std::max(2, pol.check_interval + (pol.check_interval & 1)); }

    // CpuPressureSolver3D::project (CpuPressureSolver3D.cpp:25-60) out of place:
    // faces `in` are projected into `out`. Enqueues everything; returns after
    // the solve's own residual checks (tolerance modes synchronise there).
    // The caller reads the record with finish().
    struct Pending {
        int slot;
        SolvePolicy pol;
        int iterations;
        double initial_inf_warm;
        bool initial_from_residual;
        double initial_l2;
        int checks;
        int refinements;
        double refined_l2;  // FP64 relative residual of the refined iterate (< 0: unused)
        bool warm_started;
        int learned_start = 0;  // 0 not tried, 1 learned start used, 2 rejected
        double learned_alpha = 0.0;
        bool learned_on_device = false;  // decided by k_judge: read from the record at finish
    };

    // Learned warm start (borrowed network) and training-pair recorder.
    vgx::UnetNet* net = nullptr;
    vgx::Dataset* recorder = nullptr;
    static constexpr int kJudgeSweeps = 10;
    // Sweeps from the previous pressure before a coarse-grid learned start
    // (or a coarse recorded pair) reads the residual: they remove the
    // high-frequency residual, which restriction would otherwise alias into
    // the coarse grid. They are the plain path's first sweeps and count as
    // such.
    static constexpr int kPreSweeps = 32;

    // The residual field r32 = b' - A x of the slot's current pressure (FP64
    // evaluation, stored in FP32). With `record` the same pass is the slot's
    // residual check (its partials are residual()'s bit for bit) and leaves a
    // copy of the pressure in e32[0] (the learned start's candidate).
    void residual_field(Timer& timer, int slot, bool record = false) {
        ensure_refinement_buffers();
        int blocks = 0;
        {
            Timed t(timer, stream, vgx::K_RESIDUAL);
            check(vgx::launch_residual_field(stream, n, b, current(slot), ps_dev + slot, r32, partials, &blocks,
                                             record ? e32[0] : nullptr),
                  "residual");
        }
        if (record) {
            Timed t(timer, stream, vgx::K_REDUCE);
            check(vgx::launch_finalize(stream, vgx::Finalize::Residual, partials, blocks, n, ps_dev + slot, nullptr),
                  "finalize");
        }
    }

    // Learned start of a warm tolerance solve (MacGridFluidSolver3D has no
    // counterpart): e = rms * f(r / rms) for the previous pressure's
    // residual r, taken with the energy-optimal step alpha = (r . e) /
    // (e . A e), which minimises the A-norm error along e (so the candidate
    // is never worse than the previous pressure in that norm; alpha = 0 if
    // e is useless). Both starts then get kJudgeSweeps sweeps; the one with
    // the smaller compatible L2 residual continues, so the plain
    // previous-pressure start is both the fallback and the judge. Returns the
    // sweeps spent on the kept start (the other start's are counted in
    // `sweeps` only); the slot's host record holds the kept start's check.
    int learned_start(Timer& timer, int slot, const SolvePolicy& pol, double rms, Pending& pend) {
        const size_t cells = (size_t)n * n * n;
        const float inv = (float)(1.0 / rms);
        {
            Timed t(timer, stream, vgx::K_LEARNED);
            const int m = net->grid();
            if (m > 0 && m < n) {
                // the network on the residual restricted to m^3, its
                // prediction prolonged back: the smooth modes SOR is slow on
                const int f = n / m;
                ensure_coarse_buffers(m);
                check(vgx::restrict_mean(stream, m, f, 1.f, r32, rc), "restrict");
                // the network predicts the correction in coarse-grid units:
                // the fine operator on smooth fields is the coarse one over f^2
                net->predict(stream, m, rc, inv, (float)(rms * f * f), ec);
                check(vgx::prolong_linear(stream, m, f, ec, e32[1]), "prolong");
            } else {
                net->predict(stream, n, r32, inv, (float)rms, e32[1]);
            }
            int blocks = 0;
            check(vgx::launch_energy_dots(stream, n, r32, e32[1], partials, &blocks), "dots");
            std::vector<double> part((size_t)blocks * vgx::kPartialStride);
            check(hipMemcpyAsync(part.data(), partials, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync(dots)");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            double re = 0.0, eae = 0.0;
            for (int k = 0; k < blocks; ++k) {
                re += part[(size_t)k * vgx::kPartialStride];
                eae += part[(size_t)k * vgx::kPartialStride + 1];
            }
            const double alpha = (eae > 0.0 && std::isfinite(re / eae)) ? re / eae : 0.0;
            pend.learned_alpha = alpha;
            check(vgx::launch_add_scaled_f32(stream, current(slot), (float)alpha, e32[1], e32[0], cells), "add");
        }
        // the learned path uses a copy of the slot's record (same b' mean)
        check(hipMemcpyAsync(ps_dev + kInner, ps_dev + slot, sizeof(vgx::ProjState), hipMemcpyDeviceToDevice, stream),
              "hipMemcpyAsync(record)");
        sweeps_fixed(timer, slot, pol, kJudgeSweeps, false);
        int lcur = 0;
        sweeps_on(timer, kInner, pol, kJudgeSweeps, false, b, e32, lcur);
        residual(timer, slot);
        residual(timer, kInner, b, e32[lcur]);
        read_record(slot);
        read_record(kInner);
        pend.checks += 2;
        const double count = (double)cells;
        const vgx::ProjState &rp = ps_host[slot], &rl = ps_host[kInner];
        const double rel_plain = compat_rel(rp.res2, rp.res_sum, rp.bnorm2, rp.b_sum, count);
        const double rel_learned = compat_rel(rl.res2, rl.res_sum, rl.bnorm2, rl.b_sum, count);
        if (!rl.nonfinite_x && rel_learned < rel_plain) {
            check(hipMemcpyAsync(current(slot), e32[lcur], cells * sizeof(float), hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync(start)");
            check(hipMemcpyAsync(ps_dev + slot, ps_dev + kInner, sizeof(vgx::ProjState), hipMemcpyDeviceToDevice,
                                 stream),
                  "hipMemcpyAsync(record)");
            ps_host[slot] = ps_host[kInner];
            pend.learned_start = 1;
        } else {
            pend.learned_start = 2;
        }
        return kJudgeSweeps;
    }

    // The coarse-grid learned start with every decision on the device, so the
    // host enqueues it and the solve after it without waiting. It reads the
    // residual field r32 of the pre-swept start and that start's residual
    // record (the caller's residual_field with the record); the line search,
    // the candidate and the 10-sweep judge run as kernels, and k_judge leaves
    // its decision (and, if the learned start wins, its record) in the slot's
    // record, where finish() reads it.
    int learned_start_device(Timer& timer, int slot, const SolvePolicy& pol, Pending& pend) {
        const size_t cells = (size_t)n * n * n;
        const int m = net->grid(), f = n / m;
        ensure_coarse_buffers(m);
        if (!side) {
            check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
            check(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "hipEventCreate");
            check(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming), "hipEventCreate");
        }
        if (!judge_partials)
            judge_partials = device_alloc<double>(vgx::partials_capacity(n), "hipMalloc(judge partials)");
        // The learned candidate starts from the copy of the pre-swept
        // pressure in e32[0] (residual_field with the record), so the judge's
        // plain candidate can sweep the slot's own buffers on the side stream
        // meanwhile. Neither stream writes what the other reads: the side
        // stream's sweeps read b and the record's rhs_meanf and write the
        // slot's pressure pair; this stream writes e32, the coarse buffers,
        // the partials and the records' learned_alpha and kInner.
        check(hipEventRecord(fork_ev, stream), "hipEventRecord");
        check(hipStreamWaitEvent(side, fork_ev, 0), "hipStreamWaitEvent");
        sweeps_on(timer, slot, pol, kJudgeSweeps, false, b, p[slot], p_cur[slot], side);
        check(hipEventRecord(join_ev, side), "hipEventRecord");
        {
            Timed t(timer, stream, vgx::K_LEARNED);
            check(vgx::restrict_rms(stream, m, f, ps_dev + slot, r32, rc), "restrict");
            net->predict(stream, m, rc, 1.f, 1.f, ec);
            check(vgx::prolong_rms(stream, m, f, ps_dev + slot, ec, e32[1]), "prolong");
            int blocks = 0;
            check(vgx::launch_energy_dots(stream, n, r32, e32[1], partials, &blocks), "dots");
            check(vgx::launch_line_search(stream, partials, blocks, ps_dev + slot), "line search");
            check(vgx::launch_axpy_alpha(stream, e32[0], ps_dev + slot, e32[1], e32[0], cells), "add");
        }
        check(hipMemcpyAsync(ps_dev + kInner, ps_dev + slot, sizeof(vgx::ProjState), hipMemcpyDeviceToDevice, stream),
              "hipMemcpyAsync(record)");
        int lcur = 0;
        sweeps_on(timer, kInner, pol, kJudgeSweeps, false, b, e32, lcur);
        check(hipStreamWaitEvent(stream, join_ev, 0), "hipStreamWaitEvent");
        {
            // both candidates' residual checks and the judge in two launches
            Timed t(timer, stream, vgx::K_RESIDUAL);
            check(vgx::launch_judge_pair(stream, n, b, current(slot), ps_dev + slot, partials, e32[lcur],
                                         ps_dev + kInner, judge_partials),
                  "judge");
        }
        check(vgx::launch_copy_if_learned(stream, e32[lcur], current(slot), ps_dev + slot, cells), "start");
        pend.checks += 2;
        pend.learned_on_device = true;
        return kJudgeSweeps;
    }

    // Relative L2 tolerance solve in FP32 SOR with FP64 residual checks, in
    // chunks of chunk_of(pol) sweeps. Each chunk's last sweep also evaluates
    // the residual of the iterate it starts from, and k_check raises the
    // record's `done` flag once that meets `tol`; every later sweep and check
    // then returns at once. The host keeps kLookahead chunks queued and reads
    // the chunks' records from a pinned ring as they complete, so the GPU does
    // not idle on a round trip per check. The iterate stays in buf[cur].
    // Returns the sweeps that ran; `rel` is the last checked residual.
    // With start_on_device the start's check is on the device only (the
    // learned start decided there): the chunks start without a host round
    // trip, and a start that already meets the tolerance costs one chunk.
    int l2_solve(Timer& timer, int rec, const SolvePolicy& pol, const float* rhs, float* const buf[2], int cur,
                 double tol, int budget, double& rel, int& checks, bool start_checked = false,
                 bool start_on_device = false) {
        if (!start_checked && !start_on_device) {
            residual(timer, rec, rhs, buf[cur]);
            read_record(rec);
            ++checks;
        }
        if (!start_on_device) {
            const vgx::ProjState& r = ps_host[rec];
            if (r.nonfinite_x) invalid("Invalid stencil coefficient, boundary entry, or vector value");
            const double count = (double)n * n * n;
            rel = compat_rel(r.res2, r.res_sum, r.bnorm2, r.b_sum, count);
            if (rel <= tol) return 0;
        }
        if (budget <= 0) return 0;
        const int chunk = chunk_of(pol);
        const int limit = (budget + chunk - 1) / chunk * chunk;  // sweeps the solve may enqueue
        // the final checks' spacing (even, shorter than a chunk), 0 if off
        const int fine_len = pol.final_interval > 0 && pol.final_interval + (pol.final_interval & 1) < chunk
                                 ? pol.final_interval + (pol.final_interval & 1)
                                 : 0;
        check(vgx::launch_reset_status(stream, ps_dev + rec), "reset");
        int enqueued = 0, seen = 0, ran = 0, planned = 0;
        // The residual's factor per sweep, from the last two checks, says
        // how many sweeps the tolerance is away. A unit (a chunk, or a final
        // piece) in flight that the trend says will converge is not followed
        // by another until its record is read: a unit queued behind a
        // converged one still launches every sweep (each returns at once); a
        // wrong guess costs one host round trip, and which sweeps and checks
        // run does not change. With final checks, the chunk the trend says
        // will converge runs as pieces of fine_len sweeps, each checked, so
        // the solve stops fewer sweeps past the tolerance.
        double last_rel = -1.0;
        int last_ran = 0;
        bool hold = false, fine = false;
        for (;;) {
            while (!hold && planned < limit && enqueued - seen < kLookahead) {
                const int len = fine ? std::min(fine_len, limit - planned) : chunk;
                // With graphs, a timed run still launches one unit in
                // kTimedChunkStride on the stream, with probes, to sample the
                // kernels' durations.
                const bool eager = !pol.graphs || (timer.enabled && enqueued % kTimedChunkStride == 0);
                if (eager) {
                    enqueue_chunk(&timer, rec, pol, rhs, buf[cur], buf[cur ^ 1], tol, len);
                } else {
                    check(hipGraphLaunch(chunk_graph(rec, pol, rhs, buf[cur], buf[cur ^ 1], tol, len), stream),
                          "hipGraphLaunch");
                    timer.count(vgx::K_SOR, (uint64_t)(len - 1));
                    timer.count(vgx::K_SOR_CHECK, 1);
                    timer.count(vgx::K_REDUCE, 1);
                }
                check(hipEventRecord(ring_ev[enqueued % kRing], stream), "hipEventRecord");
                ++enqueued;
                planned += len;
            }
            check(hipEventSynchronize(ring_ev[seen % kRing]), "hipEventSynchronize");
            const vgx::ProjState st = ring[seen % kRing];
            ++seen;
            ++checks;
            if (st.nonfinite_x) invalid("Invalid stencil coefficient, boundary entry, or vector value");
            ran = st.sweeps_run;
            rel = st.checked_rel;
            if (st.done || (seen == enqueued && planned >= limit)) break;
            // the residual predicted `ahead` sweeps past this check
            auto predicted = [&](int ahead) {
                return rel * std::pow(rel / last_rel, (double)ahead / (double)(ran - last_ran));
            };
            const bool trend = last_rel > 0.0 && rel < last_rel && ran > last_ran;
            hold = trend && enqueued > seen && predicted(planned - ran) <= tol;
            fine = trend && fine_len > 0 && predicted(planned - ran + chunk) <= tol;
            last_rel = rel;
            last_ran = ran;
        }
        sweeps += (uint64_t)ran;
        return ran;
    }

    // One tolerance chunk: `chunk` sweeps from xa (ending in xa again, chunk
    // is even), the last also checking the residual of its input, then the
    // check, which writes the record into the ring. With a timer the launches
    // carry probes; without one (graph capture) they are plain launches.
    void enqueue_chunk(Timer* timer, int rec, const SolvePolicy& pol, const float* rhs, float* xa, float* xb,
                       double tol, int chunk) {
        int blocks = 0;
        // the chunk's plain sweeps are timed as one run (Timer)
        hipEvent_t run = timer ? timer->range_begin(stream) : nullptr;
        for (int k = 0; k < chunk; ++k) {
            const bool last = k + 1 == chunk;
            float* in = (k & 1) ? xb : xa;
            float* out = (k & 1) ? xa : xb;
            std::unique_ptr<Timed> t;
            if (timer && last) {
                timer->range_end(stream, run, vgx::K_SOR, (uint64_t)(chunk - 1));
                t = std::make_unique<Timed>(*timer, stream, vgx::K_SOR_CHECK);
            }
            check(vgx::launch_sor_rb(stream, n, pol.omega, pol.floor, rhs, in, out, ps_dev + rec, &ps_dev[rec].done,
                                     last ? 3 : 0, partials, &blocks),
                  "sor");
        }
        std::unique_ptr<Timed> t;
        if (timer) t = std::make_unique<Timed>(*timer, stream, vgx::K_REDUCE);
        check(vgx::launch_check(stream, partials, blocks, n, ps_dev + rec, tol, chunk, ring_dev, kRing), "check");
    }

    // Captured chunk graphs, one per (record, right-hand side, buffer pair,
    // chunk length, omega, floor, tolerance): within a solve every chunk is
    // the same sequence of launches (the record's chunk counter picks the
    // ring slot on the device), so one capture serves all of them.
    struct GraphKey {
        int rec;
        const float* rhs;
        const float* xa;
        const float* xb;
        int chunk;
        float omega, floor;
        double tol;
        bool operator<(const GraphKey& o) const {
            return std::tie(rec, rhs, xa, xb, chunk, omega, floor, tol) <
                   std::tie(o.rec, o.rhs, o.xa, o.xb, o.chunk, o.omega, o.floor, o.tol);
        }
    };
    std::map<GraphKey, hipGraphExec_t> graphs;
    static constexpr int kMaxGraphs = 64;
    static constexpr int kTimedChunkStride = 8;

    hipGraphExec_t chunk_graph(int rec, const SolvePolicy& pol, const float* rhs, float* xa, float* xb, double tol,
                               int chunk) {
        const GraphKey key{rec, rhs, xa, xb, chunk, pol.omega, pol.floor, tol};
        auto it = graphs.find(key);
        if (it != graphs.end()) return it->second;
        if ((int)graphs.size() >= kMaxGraphs) release_graphs();
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
            enqueue_chunk(nullptr, rec, pol, rhs, xa, xb, tol, chunk);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check(e, "hipGraphInstantiate");
        graphs.emplace(key, exec);
        return exec;
    }
    // Graph execs are destroyed only after the stream has drained: chunks
    // launched from them may still be queued (the solve keeps kLookahead in
    // flight).
    void release_graphs() {
        if (graphs.empty()) return;
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    // FP64 iterative refinement: x64 += e with A e = float(b' - A x64), each
    // correction solved by FP32 SOR to max(0.5 tol / rel, 1e-4) relative.
    void refine(Timer& timer, int slot, const SolvePolicy& pol, Pending& pend) {
        ensure_refinement_buffers();
        const size_t cells = (size_t)n * n * n;
        float* x = current(slot);
        check(vgx::launch_widen(stream, x64, x, cells), "widen");
        int used = 0;
        double rel = 0.0;
        for (int step = 0;; ++step) {
            int blocks = 0;
            {
                Timed t(timer, stream, vgx::K_RESIDUAL);
                check(vgx::launch_residual64(stream, n, b, x64, ps_dev + slot, r32, partials, &blocks), "residual64");
            }
            {
                Timed t(timer, stream, vgx::K_REDUCE);
                check(vgx::launch_finalize(stream, vgx::Finalize::Residual, partials, blocks, n, ps_dev + slot,
                                           nullptr),
                      "finalize");
            }
            read_record(slot);
            ++pend.checks;
            const vgx::ProjState& r = ps_host[slot];
            if (r.nonfinite_x) invalid("Invalid stencil coefficient, boundary entry, or vector value");
            const double count = (double)cells;
            rel = compat_rel(r.res2, r.res_sum, r.bnorm2, r.b_sum, count);
            if (step == 0) {
                pend.initial_l2 = rel;
                pend.initial_inf_warm = r.res_inf;
            }
            if (rel <= pol.l2_tol || step >= pol.refinement || used >= pol.max_iterations) break;
            {
                Timed t(timer, stream, vgx::K_FILL);
                check(vgx::launch_fill_f32(stream, e32[0], cells, 0.f), "fill");
            }
            // The correction's right-hand side is float(r) minus the float of
            // its mean, which the correction record subtracts on the fly.
            ps_host[kInner].rhs_meanf = (float)(r.res_sum / count);
            check(hipMemcpyAsync(&ps_dev[kInner].rhs_meanf, &ps_host[kInner].rhs_meanf, sizeof(float),
                                 hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync(record)");
            // Each correction is asked for an equal share (in decades) of the
            // reduction still missing over the corrections left, with a factor
            // 2 of margin: the smooth correction equation has an FP32 rounding
            // floor that rises with n (about 1e-4 relative at 128^3), so asking
            // one correction for everything stalls it at large grids.
            const double inner_tol =
                std::max(0.5 * std::pow(pol.l2_tol / rel, 1.0 / (double)(pol.refinement - step)), 1e-4);
            double inner_rel = 0.0;
            used += l2_solve(timer, kInner, pol, r32, e32, 0, inner_tol, pol.max_iterations - used, inner_rel,
                             pend.checks);
            check(vgx::launch_axpy_f64(stream, x64, e32[0], cells), "axpy");
            ++pend.refinements;
        }
        check(vgx::launch_narrow(stream, x, x64, cells), "narrow");
        residual(timer, slot);  // sum(x) of the FP32 pressure for the zero-mean gauge
        pend.iterations = used;
        pend.refined_l2 = rel;
    }

    Pending project(Timer& timer, int slot, const SolvePolicy& pol, const float* const in[3], float* const out[3]) {
        if (!std::isfinite(pol.omega) || pol.omega <= 0.f || pol.omega >= 2.f)
            invalid("SOR relaxation factor must be between 0 and 2");
        if (!std::isfinite(pol.dt) || pol.dt <= 0.f) invalid("Projection time step must be finite and positive");
        const float hh = h();
        const float scale = hh * hh / pol.dt;
        if (!std::isfinite(scale) || scale <= 0.f)
            invalid("Projection scale is outside the supported float range");
        if (pol.max_iterations < 0 || pol.check_interval <= 0 || !std::isfinite(pol.abs_tol) || pol.abs_tol < 0.0 ||
            !std::isfinite(pol.rel_tol) || pol.rel_tol < 0.0 || !(pol.l2_tol >= 0.0) || pol.refinement < 0)
            invalid("Invalid linear solve options");
        if (pol.refinement > 0 && (pol.fixed || !pol.use_l2))
            invalid("FP64 iterative refinement needs the relative L2 tolerance solve");
        const size_t cells = (size_t)n * n * n;
        vgx::ProjState* ps = ps_dev + slot;
        Pending pend{slot, pol, 0, 0.0, false, 0.0, 0, 0, -1.0, pol.warm};
        int blocks = 0;
        {
            Timed t(timer, stream, vgx::K_DIVERGENCE);
            check(vgx::launch_divergence_rhs(stream, n, hh, scale, in[0], in[1], in[2], b, partials, &blocks),
                  "divergence");
        }
        {
            Timed t(timer, stream, vgx::K_REDUCE);
            check(vgx::launch_finalize(stream, vgx::Finalize::Rhs, partials, blocks, n, ps, nullptr), "finalize");
        }
        if (!pol.warm) {
            Timed t(timer, stream, vgx::K_FILL);
            check(vgx::launch_fill_f32(stream, current(slot), cells, 0.f), "fill");
        }
        if (pol.fixed) {
            if (pol.warm) {
                residual(timer, slot);
                pend.initial_from_residual = true;
            }
            if (pol.max_iterations > 0) sweeps_fixed(timer, slot, pol, pol.max_iterations, true);
            else if (!pol.warm) residual(timer, slot);
            pend.iterations = pol.max_iterations;
        } else {
            // Residual-checked solve: the reference inf-norm rule
            // (SOR.cpp:22-33,50-61) or the relative L2 rule.
            residual(timer, slot);
            read_record(slot);
            ++pend.checks;
            vgx::ProjState& r = ps_host[slot];
            if (r.nonfinite_x) invalid("Invalid stencil coefficient, boundary entry, or vector value");
            pend.initial_from_residual = true;
            if (!(compat_norm2(r.bnorm2, r.b_sum, (double)cells) > 0.0) && pol.warm) {
                // Nothing to solve: start from zero, as a cold solve would.
                Timed t(timer, stream, vgx::K_FILL);
                check(vgx::launch_fill_f32(stream, current(slot), cells, 0.f), "fill");
                residual(timer, slot);
                read_record(slot);
            }
            const double count = (double)cells;
            const double bnorm = std::sqrt(std::max(compat_norm2(r.bnorm2, r.b_sum, count), 0.0));
            auto rel_l2 = [&]() { return compat_rel(r.res2, r.res_sum, r.bnorm2, r.b_sum, count); };
            pend.initial_l2 = rel_l2();
            pend.initial_inf_warm = r.res_inf;
            if (pol.refinement > 0 && bnorm > 0.0) {
                --pend.checks;  // refine() re-evaluates the start in FP64
                refine(timer, slot, pol, pend);
            } else if (pol.use_l2) {
                double rel = 0.0;
                // learned start / training pairs: warm solves with work left to do
                double rms = std::sqrt(std::max(compat_norm2(r.res2, r.res_sum, count), 0.0) / count);
                bool learn = net && pol.warm && learnable(net->grid()) && pend.initial_l2 > pol.l2_tol && rms > 0.0;
                bool record = recorder && recordable(recorder->n) && recorder->size < recorder->capacity &&
                              pol.warm && pend.initial_l2 > pol.l2_tol && rms > 0.0;
                int pre = 0;
                // a coarse-grid network decides on the device unless pairs
                // are being recorded (which need the scales on the host)
                const bool on_device = learn && !record && net->grid() > 0 && net->grid() < n &&
                                       pol.max_iterations > kPreSweeps + kJudgeSweeps;
                // a coarse-grid network or dataset reads the residual after
                // kPreSweeps plain sweeps (counted in the solve's sweeps)
                if ((learn && net->grid() > 0 && net->grid() < n) || (record && recorder->n < n)) {
                    pre = std::min(kPreSweeps, pol.max_iterations);
                    sweeps_fixed(timer, slot, pol, pre, false);
                    // on the device path the residual field the network
                    // reads is this check's own pass
                    if (on_device) residual_field(timer, slot, true);
                    else residual(timer, slot);
                    ++pend.checks;
                    if (on_device) {
                        pre += learned_start_device(timer, slot, pol, pend);
                        pend.iterations = pre + l2_solve(timer, slot, pol, b, p[slot], p_cur[slot], pol.l2_tol,
                                                         pol.max_iterations - pre, rel, pend.checks, false, true);
                        learn = record = false;
                    } else {
                        read_record(slot);
                        rms = std::sqrt(std::max(compat_norm2(r.res2, r.res_sum, count), 0.0) / count);
                        if (r.nonfinite_x || rel_l2() <= pol.l2_tol || !(rms > 0.0)) learn = record = false;
                    }
                }
                if (!on_device) {
                if (learn || record) residual_field(timer, slot);
                float* record_prev = nullptr;
                if (record) {
                    const int m = recorder->n;
                    const size_t k = (size_t)recorder->size * recorder->cells();
                    if (m == n) {
                        check(vgx::launch_scale_f32(stream, r32, (float)(1.0 / rms), recorder->inputs + k, cells),
                              "record");
                        record_prev = recorder->prev;
                    } else {
                        check(vgx::restrict_mean(stream, m, n / m, (float)(1.0 / rms), r32, recorder->inputs + k),
                              "record");
                        if (!rec_prev) rec_prev = device_alloc<float>(cells, "hipMalloc(record)");
                        record_prev = rec_prev;
                    }
                    check(hipMemcpyAsync(record_prev, current(slot), cells * sizeof(float), hipMemcpyDeviceToDevice,
                                         stream),
                          "hipMemcpyAsync(record)");
                    // the pre-swept start's mean (its last check): the recorded
                    // correction is then zero-mean, as the gauged solution is
                    if (m != n) check(vgx::wide::shift<float>(stream, rec_prev, r.x_meanf, cells), "record");
                }
                if (learn) pre += learned_start(timer, slot, pol, rms, pend);
                pend.iterations = pre + l2_solve(timer, slot, pol, b, p[slot], p_cur[slot], pol.l2_tol,
                                                 std::max(pol.max_iterations - pre, 0), rel, pend.checks, true);
                if (record) {
                    const int m = recorder->n;
                    const size_t k = (size_t)recorder->size * recorder->cells();
                    if (m == n) {
                        check(vgx::launch_correction(stream, current(slot), record_prev, ps, (float)(1.0 / rms),
                                                     recorder->targets + k, cells),
                              "record");
                    } else {
                        check(vgx::launch_correction(stream, current(slot), record_prev, ps, (float)(1.0 / rms),
                                                     e32[1], cells),
                              "record");
                        // targets in coarse-grid units (divided by f^2), as the learned start reads them
                        const float f = (float)(n / m);
                        check(vgx::restrict_mean(stream, m, n / m, 1.f / (f * f), e32[1], recorder->targets + k),
                              "record");
                    }
                    recorder->size += 1;
                }
                }  // !on_device
                // A converged start keeps the gauge sum of its own check; a
                // solve's last chunk leaves the sum of its result in the record.
            } else {
                // The reference's rule (SOR.cpp:22-33,50-61): inf-norm of the
                // current iterate every check_interval sweeps and at the budget.
                const double target = std::max(pol.abs_tol, pol.rel_tol * r.res_inf);
                bool converged = r.res_inf <= target;
                int sweeps_done = 0;
                while (!converged && sweeps_done < pol.max_iterations) {
                    const int chunk = std::min(pol.check_interval, pol.max_iterations - sweeps_done);
                    sweeps_fixed(timer, slot, pol, chunk, false);
                    sweeps_done += chunk;
                    residual(timer, slot);
                    read_record(slot);
                    ++pend.checks;
                    converged = r.res_inf <= target;
                }
                pend.iterations = sweeps_done;
            }
        }
        float* p_next = p[slot][1 - p_cur[slot]];
        {
            Timed t(timer, stream, vgx::K_GRADIENT);
            check(vgx::launch_gradient(stream, n, hh, pol.dt, in[0], in[1], in[2], current(slot), p_next, b, out[0],
                                       out[1], out[2], ps, partials, &blocks),
                  "gradient");
        }
        {
            Timed t(timer, stream, vgx::K_REDUCE);
            check(vgx::launch_finalize(stream, vgx::Finalize::Final, partials, blocks, n, ps, nullptr), "finalize");
        }
        p_cur[slot] = 1 - p_cur[slot];
        return pend;
    }

    // Turns the device record into a ProjectionResult (after a synchronise and
    // read_record of the slot).
    vgx_projection_result finish(const Pending& pend) const {
        const vgx::ProjState& r = ps_host[pend.slot];
        if (r.nonfinite_in) invalid("Non-finite divergence");
        if (r.nonfinite_out) invalid("Non-finite MAC velocity");
        vgx_projection_result out;
        std::memset(&out, 0, sizeof(out));
        out.iterations = pend.iterations;
        out.initial_residual = pend.initial_from_residual ? pend.initial_inf_warm : r.initial_inf;
        if (pend.pol.fixed && pend.initial_from_residual) out.initial_residual = r.res_inf;
        out.final_residual = r.final_inf;
        out.residual_available = 1;
        out.divergence_before = r.div_before;
        out.divergence_after = r.div_after;
        out.removed_rhs_mean = r.rhs_mean;
        const double count = (double)n * n * n;
        const double bnorm2 = compat_norm2(r.final_bnorm2, r.final_b_sum, count);
        const double bnorm = bnorm2 > 0.0 ? std::sqrt(bnorm2) : 0.0;
        out.final_residual_l2 = compat_rel(r.final_res2, r.final_res_sum, r.final_bnorm2, r.final_b_sum, count);
        // A refined solve reports its FP64 iterate's residual: the stored FP32
        // pressure cannot represent it more accurately.
        if (pend.refined_l2 >= 0.0) out.final_residual_l2 = pend.refined_l2;
        out.initial_residual_l2 = pend.checks > 0 ? pend.initial_l2 : (bnorm > 0.0 ? 1.0 : 0.0);
        out.residual_checks = pend.checks;
        out.refinement_steps = pend.refinements;
        out.warm_started = pend.warm_started;
        out.learned_start = pend.learned_on_device ? r.learned : pend.learned_start;
        if (pend.pol.use_l2) {
            out.converged = out.final_residual_l2 <= pend.pol.l2_tol;
        } else {
            const double target = std::max(pend.pol.abs_tol, pend.pol.rel_tol * out.initial_residual);
            out.converged = out.final_residual <= target;
        }
        return out;
    }
};

inline void params_default(vgx_params* p) {
    std::memset(p, 0, sizeof(*p));
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

// GpuMacSimulation3D::advance checks (GpuMacSimulation3D.cpp:226-233) plus the
// kick multiplier check of MacGridFluidSolver3D::advance (.cpp:324-325) and
// the extension fields.
inline void validate_step(const vgx_params& p, float dt) {
    if (!std::isfinite(p.emitter_kick_multiplier) || p.emitter_kick_multiplier < 0.f)
        invalid("Invalid emitter kick multiplier");
    if (!std::isfinite(dt) || dt <= 0 || !std::isfinite(p.source_strength) || p.source_strength < 0 ||
        !std::isfinite(p.emitter_radius) || p.emitter_radius <= 0 || p.emitter_radius > 1 ||
        !std::isfinite(p.buoyancy) || !std::isfinite(p.smoke_decay) || p.smoke_decay < 0 ||
        !std::isfinite(p.sor_omega) || p.sor_omega <= 0 || p.sor_omega >= 2 || p.project_iterations < 0 ||
        !std::isfinite(p.stir_strength) || !std::isfinite(p.cfl_mc) || p.cfl_mc <= 0)
        invalid("Invalid resident GPU step parameters");
    if (p.emitter_profile != VGX_EMITTER_SPHERE && p.emitter_profile != VGX_EMITTER_GAUSSIAN)
        invalid("Unknown emitter profile");
    if (p.emitter2_enabled &&
        (!std::isfinite(p.emitter2_radius) || p.emitter2_radius <= 0 || p.emitter2_radius > 1))
        invalid("Invalid second emitter radius");
    if (p.solve_mode != VGX_SOLVE_FIXED && p.solve_mode != VGX_SOLVE_TOLERANCE) invalid("Unknown solve mode");
    if (p.warm_start != VGX_WARM_ZERO && p.warm_start != VGX_WARM_PREVIOUS) invalid("Unknown warm start");
    if (p.solve_mode == VGX_SOLVE_TOLERANCE &&
        (!(p.tolerance >= 0.0) || p.max_sweeps < 0 || p.check_interval <= 0 || p.refinement_steps < 0 ||
         p.final_check_interval < 0))
        invalid("Invalid tolerance solve parameters");
}

}  // namespace vgx_impl

using namespace vgx_impl;

// ======================================================================== sim

struct vgx_sim {
    int n = 0;
    float box = 0.f;
    vgx_params params{};
    hipStream_t stream = nullptr;
    size_t cells = 0, faces = 0;
    float* rho[2] = {nullptr, nullptr};
    int rho_cur = 0;
    float* S[3] = {};  // state faces
    float* T[3] = {};  // first advection pass (u tilde)
    float* Q[3] = {};  // second advection pass
    float* F[3] = {};  // forward-trace scratch of MacCormack
    float* noise = nullptr;
    uint32_t* rng = nullptr;
    vgx::DiagState* ds_dev = nullptr;
    vgx::DiagState* ds_host = nullptr;
    ProjectionWorkspace proj;
    Timer timer;
    vgx_projection_result last[2]{};
    uint64_t uploaded = 0, downloaded = 0;
    // Installed stages (MacGridFluidSolver3D::setAdvection / setPressureSolver,
    // MacGridFluidSolver3D.cpp:305-321): host callbacks that replace one stage
    // of the step. Their inputs and outputs pass through pinned host buffers:
    // flow u, v, w, source, output, density.
    vgx_advection_hook advect_hook = nullptr;
    void* advect_user = nullptr;
    vgx_projection_hook project_hook = nullptr;
    void* project_user = nullptr;
    float* stage_host[6] = {};
    bool own_stream = true;    // false for the scenes of a batch, which share the batch's stream
    bool batch_owned = false;  // a scene of a batch: the batch steps, resets and frees it

    float h() const { return box / (float)n; }

    // Allocates every field; `shared` (if given) is used instead of a stream
    // of the simulation's own.
    void allocate(hipStream_t shared = nullptr) {
        cells = (size_t)n * n * n;
        faces = (size_t)(n + 1) * n * n;
        own_stream = shared == nullptr;
        if (own_stream) check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        else stream = shared;
        for (auto& r : rho) r = device_alloc<float>(cells, "hipMalloc(rho)");
        for (auto* set : {S, T, Q, F})
            for (int a = 0; a < 3; ++a) set[a] = device_alloc<float>(faces, "hipMalloc(faces)");
        noise = device_alloc<float>(cells, "hipMalloc(noise)");
        rng = device_alloc<uint32_t>(1, "hipMalloc(rng)");
        ds_dev = device_alloc<vgx::DiagState>(1, "hipMalloc(diag)");
        void* host = nullptr;
        check(hipHostMalloc(&host, sizeof(vgx::DiagState), hipHostMallocDefault), "hipHostMalloc");
        ds_host = static_cast<vgx::DiagState*>(host);
        proj.allocate(n, box, stream);
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& r : rho)
            if (r) (void)hipFree(r);
        for (auto* set : {S, T, Q, F})
            for (int a = 0; a < 3; ++a)
                if (set[a]) (void)hipFree(set[a]);
        for (void* q : {(void*)noise, (void*)rng, (void*)ds_dev})
            if (q) (void)hipFree(q);
        if (ds_host) (void)hipHostFree(ds_host);
        for (auto& h : stage_host)
            if (h) (void)hipHostFree(h);
        proj.release();
        if (stream && own_stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // resetState (MacGridFluidSolver3D.cpp:271-290 / GpuMacSimulation3D.cpp:89-98).
    void reset() {
        {
            Timed t(timer, stream, vgx::K_FILL);
            for (auto& r : rho) check(vgx::launch_fill_f32(stream, r, cells, 0.f), "fill");
            for (auto* set : {S, T, Q, F})
                for (int a = 0; a < 3; ++a) check(vgx::launch_fill_f32(stream, set[a], faces, 0.f), "fill");
            check(vgx::launch_set_u32(stream, rng, params.random_seed), "rng");
        }
        proj.clear(timer);
        rho_cur = 0;
        std::memset(last, 0, sizeof(last));
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        timer.flush();
    }

    // emit for one sphere (MacGridFluidSolver3D.cpp:140-165): patch bounds and
    // every float of the in-sphere test computed here exactly as the reference.
    vgx::EmitArgs emit_args(float dt, float cxf, float cyf, float czf, float radius) const {
        const vgx_params& P = params;
        const float strength = P.source_strength;
        vgx::EmitArgs a{};
        a.n = n;
        a.dx = h();
        a.dt_scale = dt * 60.0f;
        a.ecx = cxf * box;
        a.ecy = cyf * box;
        a.ecz = czf * box;
        a.erad = radius * box;
        const int emitR = (int)(a.erad / a.dx + 1);
        const int c[3] = {(int)(a.ecx / a.dx), (int)(a.ecy / a.dx), (int)(a.ecz / a.dx)};
        for (int k = 0; k < 3; ++k) {
            a.lo[k] = std::max(0, c[k] - emitR);
            a.ext[k] = std::max(0, std::min(n, c[k] + emitR) - a.lo[k]);
        }
        a.profile = P.emitter_profile;
        a.kick_on = P.emitter_kick_multiplier > 0.f;
        a.kick = P.emitter_kick_multiplier * (P.emitter_kick_base + P.emitter_kick_scale * strength);
        a.stir = P.stir_strength;
        a.strength = strength;
        return a;
    }
    // Emitter `which` (0 first, 1 second) of a step, if it emits.
    bool emitter(float dt, int which, vgx::EmitArgs& a) const {
        if (params.source_strength <= 0.f) return false;
        if (which == 0) {
            a = emit_args(dt, params.emitter_center_x, params.emitter_center_y, params.emitter_center_z,
                          params.emitter_radius);
            return true;
        }
        if (!params.emitter2_enabled) return false;
        a = emit_args(dt, params.emitter2_center_x, params.emitter2_center_y, params.emitter2_center_z,
                      params.emitter2_radius);
        return true;
    }

    void emit(float dt) {
        for (int which = 0; which < 2; ++which) {
            vgx::EmitArgs a;
            if (!emitter(dt, which, a)) continue;
            Timed t(timer, stream, vgx::K_EMIT);
            check(vgx::launch_emit(stream, a, rng, noise, rho[rho_cur], S[0], S[1], S[2]), "emit");
        }
    }

    // Velocity advection of the three face fields (CpuAdvectionKernels.cpp:82-269):
    // source S (or 2 S - T when `hat`), flow (fu, fv, fw), into out.
    // Arguments of a velocity advection pass (the MacCormack pass when
    // `mc_pass`, else the semi-Lagrangian pass, which writes the forward
    // trace F under MacCormack).
    vgx::AdvArgs faces_args(float* const src[3], float* const tsrc[3], float* const flow[3], float* const out[3],
                            float dt, bool mc_pass) const {
        const bool mc = params.maccormack_vel != 0;
        vgx::AdvArgs a{};
        a.u = flow[0];
        a.v = flow[1];
        a.w = flow[2];
        a.n = n;
        a.first_field = 0;
        a.h = h();
        a.box = box;
        a.dt = dt;
        a.cfl = params.cfl_mc;
        a.out_scale = 1.f;
        for (int k = 0; k < 3; ++k)
            a.f[k] = {src[k], tsrc ? tsrc[k] : nullptr, F[k], (mc && !mc_pass) ? F[k] : out[k]};
        return a;
    }
    // ---- installed stages ----
    void ensure_stage_buffers() {
        for (auto& h : stage_host) {
            if (h) continue;
            void* q = nullptr;
            check(hipHostMalloc(&q, faces * sizeof(float), hipHostMallocDefault), "hipHostMalloc(stage)");
            h = static_cast<float*>(q);
        }
    }
    void to_host(float* host, const float* dev, size_t count) {
        check(hipMemcpyAsync(host, dev, count * sizeof(float), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(stage)");
    }
    void to_device(float* dev, const float* host, size_t count) {
        check(hipMemcpyAsync(dev, host, count * sizeof(float), hipMemcpyHostToDevice, stream), "hipMemcpyAsync(stage)");
    }
    static void stage_status(int rc, const char* stage) {
        if (rc == VGX_OK) return;
        throw Failure{rc == VGX_EINVAL || rc == VGX_ERANGE ? rc : VGX_EDEVICE,
                      std::string("installed ") + stage + " stage failed"};
    }
    // The three face advections through the installed stage, in the
    // reference's order x, y, z (MacGridFluidSolver3D.cpp:203-208,220-222):
    // flow and source are downloaded, each output uploaded. A reflected source
    // 2 S - T is formed on the device first (into F), as the reference forms
    // its hat fields before advecting them (.cpp:211-218).
    void advect_faces_installed(float* const src[3], float* const tsrc[3], float* const flow[3],
                                float* const out[3], float dt) {
        const float* source[3] = {src[0], src[1], src[2]};
        if (tsrc) {
            for (int k = 0; k < 3; ++k) {
                check(vgx::launch_scale_f32(stream, src[k], 2.f, F[k], faces), "reflect");
                check(vgx::launch_add_scaled_f32(stream, F[k], -1.f, tsrc[k], F[k], faces), "reflect");
                source[k] = F[k];
            }
        }
        for (int k = 0; k < 3; ++k) to_host(stage_host[k], flow[k], faces);
        const int method = params.maccormack_vel ? VGX_MACCORMACK : VGX_SEMI_LAGRANGIAN;
        for (int k = 0; k < 3; ++k) {
            to_host(stage_host[3], source[k], faces);
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            stage_status(advect_hook(advect_user, k, method, dt, params.cfl_mc, stage_host[0], stage_host[1],
                                     stage_host[2], stage_host[3], stage_host[4]),
                         "advection");
            to_device(out[k], stage_host[4], faces);
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    // Smoke advection through the installed stage (in place on the host
    // copy, Advection3D::advectScalar), then applyDissipation on the device
    // (MacGridFluidSolver3D.cpp:183-190: density *= clamp(1 - decay dt, 0, 1)).
    void advect_smoke_installed(float dt) {
        for (int k = 0; k < 3; ++k) to_host(stage_host[k], S[k], faces);
        to_host(stage_host[4], rho[rho_cur], cells);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        const int method = params.maccormack_smoke ? VGX_MACCORMACK : VGX_SEMI_LAGRANGIAN;
        stage_status(advect_hook(advect_user, VGX_TARGET_DENSITY, method, dt, params.cfl_mc, stage_host[0],
                                 stage_host[1], stage_host[2], nullptr, stage_host[4]),
                     "advection");
        float* dst = rho[1 - rho_cur];
        to_device(dst, stage_host[4], cells);
        const float factor = std::clamp(1.0f - params.smoke_decay * dt, 0.0f, 1.0f);
        check(vgx::launch_scale_f32(stream, dst, factor, dst, cells), "decay");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        rho_cur = 1 - rho_cur;
    }
    // One projection through the installed stage: the advected faces `in`
    // (and the density) are downloaded, projected in place by the stage with
    // the options MacGridFluidSolver3D::project builds (.cpp:129-138: dt,
    // relaxation = sorOmega, maxIterations, fixedIterations), and uploaded
    // into `out`.
    vgx_projection_result project_installed(float dt, int iterations, float* const in[3], float* const out[3]) {
        for (int k = 0; k < 3; ++k) to_host(stage_host[k], in[k], faces);
        to_host(stage_host[5], rho[rho_cur], cells);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        vgx_projection_options o{};
        o.dt = dt;
        o.relaxation = params.sor_omega;
        o.max_iterations = iterations;
        o.residual_check_interval = 8;
        o.absolute_tolerance = 1e-6;
        o.relative_tolerance = 1e-5;
        o.fixed_iterations = 1;
        o.diagonal_floor = 6.f;
        o.l2_tolerance = 1e-5;
        vgx_projection_result r{};
        stage_status(project_hook(project_user, &o, stage_host[0], stage_host[1], stage_host[2], stage_host[5], &r),
                     "pressure");
        for (int k = 0; k < 3; ++k) to_device(out[k], stage_host[k], faces);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return r;
    }

    void advect_faces(float* const src[3], float* const tsrc[3], float* const flow[3], float* const out[3],
                      float dt) {
        if (advect_hook) {
            advect_faces_installed(src, tsrc, flow, out, dt);
            return;
        }
        {
            const vgx::AdvArgs a = faces_args(src, tsrc, flow, out, dt, false);
            Timed t(timer, stream, vgx::K_ADVECT_SL_FACES);
            check(vgx::launch_advect_sl(stream, a, 3, tsrc != nullptr), "advect_sl");
        }
        if (params.maccormack_vel != 0) {
            const vgx::AdvArgs a = faces_args(src, tsrc, flow, out, dt, true);
            Timed t(timer, stream, vgx::K_ADVECT_MC_FACES);
            check(vgx::launch_advect_mc(stream, a, 3, tsrc != nullptr), "advect_mc");
        }
    }

    // Smoke advection with the projected flow, then applyDissipation
    // (MacGridFluidSolver3D.cpp:183-190) folded into the output write.
    // Arguments of the smoke advection from rho[cur] into rho[1 - cur].
    vgx::AdvArgs smoke_args(float dt, int cur, bool mc_pass) const {
        const bool mc = params.maccormack_smoke != 0;
        vgx::AdvArgs a{};
        a.u = S[0];
        a.v = S[1];
        a.w = S[2];
        a.n = n;
        a.first_field = 3;
        a.h = h();
        a.box = box;
        a.dt = dt;
        a.cfl = params.cfl_mc;
        a.out_scale = std::clamp(1.0f - params.smoke_decay * dt, 0.0f, 1.0f);
        float* dst = rho[1 - cur];
        a.f[3] = {rho[cur], nullptr, F[0], (mc && !mc_pass) ? F[0] : dst};
        a.scale_out = !mc || mc_pass;
        return a;
    }
    void advect_smoke(float dt) {
        if (advect_hook) {
            advect_smoke_installed(dt);
            return;
        }
        {
            const vgx::AdvArgs a = smoke_args(dt, rho_cur, false);
            Timed t(timer, stream, vgx::K_ADVECT_SL_SCALAR);
            check(vgx::launch_advect_sl(stream, a, 1, false), "advect_sl");
        }
        if (params.maccormack_smoke != 0) {
            const vgx::AdvArgs a = smoke_args(dt, rho_cur, true);
            Timed t(timer, stream, vgx::K_ADVECT_MC_SCALAR);
            check(vgx::launch_advect_mc(stream, a, 1, false), "advect_mc");
        }
        rho_cur = 1 - rho_cur;
    }

    SolvePolicy policy(float dt, int iterations) const {
        SolvePolicy pol;
        pol.dt = dt;
        pol.omega = params.sor_omega;
        pol.floor = 6.f;  // MacGridFluidSolver3D.hpp:118
        pol.max_iterations = iterations;
        pol.fixed = true;
        pol.warm = params.warm_start == VGX_WARM_PREVIOUS;
        pol.graphs = params.launch_mode == VGX_LAUNCH_GRAPH;
        if (params.solve_mode == VGX_SOLVE_TOLERANCE) {
            pol.fixed = false;
            pol.use_l2 = true;
            pol.l2_tol = params.tolerance;
            pol.max_iterations = params.max_sweeps;
            pol.check_interval = params.check_interval;
            pol.final_interval = params.final_check_interval;
            pol.refinement = params.refinement_steps;
        }
        return pol;
    }

    // step / stepNoReflect (MacGridFluidSolver3D.cpp:192-245).
    void advance(float dt) {
        validate_step(params, dt);
        emit(dt);
        {
            Timed t(timer, stream, vgx::K_FORCES);
            check(vgx::launch_forces(stream, n, rho[rho_cur], S[0], S[1], S[2], params.buoyancy, dt,
                                     params.buoyancy_split),
                  "forces");
        }
        std::vector<ProjectionWorkspace::Pending> pending;
        vgx_projection_result installed[2]{};
        int projections = 0;
        // slot 1 is the first projection of a reflection step, 0 the last
        auto project = [&](int slot, float pdt, int iterations, float* const in[3]) {
            if (project_hook) installed[projections] = project_installed(pdt, iterations, in, S);
            else pending.push_back(proj.project(timer, slot, policy(pdt, iterations), in, S));
            ++projections;
        };
        if (params.reflection) {
            const float halfDt = 0.5f * dt;
            const int halfIters = std::max(1, params.project_iterations / 2);
            advect_faces(S, nullptr, S, T, halfDt);
            project(1, halfDt, halfIters, T);
            advect_faces(S, T, S, Q, halfDt);
            project(0, halfDt, halfIters, Q);
        } else {
            advect_faces(S, nullptr, S, T, dt);
            project(0, dt, params.project_iterations, T);
        }
        advect_smoke(dt);
        check(hipMemcpyAsync(proj.ps_host, proj.ps_dev, 2 * sizeof(vgx::ProjState), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(records)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        timer.flush();
        vgx_projection_result results[2];
        for (int k = 0; k < projections; ++k)
            results[k] = project_hook ? installed[k] : proj.finish(pending[k]);
        last[0] = results[projections - 1];
        last[1] = results[0];
    }
};
