// vgx_batch.hip - many independent MAC 3D scenes of uneven sizes stepped
// together (include/vapor_mac3d.h, "Batches of scenes").
//
// The scenes of a batch advance in lockstep on one stream. Every stage of the
// step is one launch over all scenes (vgx_kernels.hip, "batches"): scene s
// owns a contiguous range of the launch's work items, and each item runs the
// single-scene kernel's block with that scene's arguments, so a scene's
// fields are bit-identical to the same scene stepped alone. The per-scene
// argument tables live in device memory and are rebuilt only when dt changes.
//
// Tolerance solves run all scenes' sweeps in one launch per sweep. Each scene
// has its own `done` flag (convergence masking): after every check a small
// kernel compacts the sweep's work list to the tiles of the scenes still
// solving, which the persistent sweep grid reads from device memory, so a
// converged scene costs nothing; its check does nothing, and the host stops
// issuing chunks once every scene is done. The fixed parts of a step (everything but the
// chunk loop) and each chunk are captured as HIP graphs and replayed.
//
// An autotuner picks, at creation, the launch shape of each batched kernel
// family by timing candidates on the batch's own buffers: the items per
// thread of the element-wise kernels (advection, forces), the grid size
// (one block per item, or a persistent grid that loops over items) of the
// advection, red-black sweep and cell-stencil kernels, and the sweep's
// staging of each tile (the right-hand side in registers and the iterate in
// a smaller LDS buffer, or both in LDS). No candidate changes what a kernel
// computes, only which block computes it and how a tile is staged.
#include "vgx_sim.hpp"

#include <array>
#include <numeric>

namespace {

using namespace vgx_impl;

constexpr int kTuneReps = 5;

// The batch's per-scene argument tables, as device pointers into one buffer.
struct Tables {
    vgx::EmitSlot* emit[2][2] = {};  // [rho_cur][first, second emitter]
    vgx::BatchList emit_list[2] = {};
    vgx::ForceArgs* forces[2] = {};  // [rho_cur]
    vgx::AdvArgs* adv[4][2] = {};  // [pass: first, reflected, smoke from rho 0, smoke from rho 1][SL, MC]
    vgx::CellArgs* div[2] = {};
    vgx::CellArgs* res[2][2] = {};   // [slot][p_cur]
    vgx::CellArgs* grad[2][2] = {};  // [slot][p_cur]
    vgx::RbSlot* rb[2][2][2] = {};   // [slot][p_cur][sweep parity]
    vgx::FinArgs* fin_cell[2] = {};
    vgx::FinArgs* fin_rb[2] = {};
    vgx::FillSlot* zero[2][2] = {};
    vgx::ProjState** records = nullptr;
    vgx::BatchList elementwise[3] = {};  // items per thread 1, 2, 4
    vgx::BatchList cells = {}, stencil = {}, tiles = {};
};

// Host image of the tables: sections appended with 256-byte alignment, each
// remembered by offset until the blob is on the device.
struct Blob {
    std::vector<char> bytes;
    template <class T> size_t add(const std::vector<T>& v) {
        const size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

}  // namespace

struct vgx_batch {
    std::vector<vgx_sim*> sims;
    int count = 0;
    hipStream_t stream = nullptr;
    vgx_params shared{};  // the solver fields every scene shares
    Timer timer;
    int rho_cur = 0;
    int p_cur[2] = {0, 0};
    bool vec = true;  // float4 staging in the sweeps (every n % 4 == 0)
    bool reg = true;  // every scene's plain sweeps can take the register path (rb_reg_fits)

    char* dev = nullptr;
    size_t dev_bytes = 0;
    float table_dt = -1.f;
    Tables t;

    static constexpr int kRing = 8, kLookahead = 4;
    vgx::ProjState* ring = nullptr;  // pinned [kRing][count]
    vgx::ProjState* ring_dev = nullptr;
    hipEvent_t ring_ev[kRing] = {};
    vgx::ProjState* recs = nullptr;  // pinned [count][3]: the records at the end of a step
    vgx::ProjState* recs_dev = nullptr;
    long long* stamp = nullptr;  // pinned: device clock at the start of the last step
    long long* stamp_dev = nullptr;
    double clock_khz = 0.0;
    std::vector<double> latency_ms;  // per scene, last step

    vgx::ActiveList act{};  // the tiles of the scenes still solving (rebuilt after every check)

    vgx_batch_tuning tune{};
    std::map<std::array<int, 6>, hipGraphExec_t> graphs;  // key: phase, rho_cur, p_cur[0], p_cur[1], slot/chunk

    // ------------------------------------------------------------ set-up

    void create(int n_scenes, const int* n, const float* box, const vgx_params* params) {
        if (n_scenes <= 0 || n_scenes > 65535) invalid("a batch holds 1 to 65535 scenes");
        if (!n || !box || !params) invalid("null scene description");
        count = n_scenes;
        shared = params[0];
        for (int s = 0; s < count; ++s) {
            validate_grid(n[s], box[s]);
            const vgx_params& p = params[s];
            if (p.reflection != shared.reflection || p.maccormack_vel != shared.maccormack_vel ||
                p.maccormack_smoke != shared.maccormack_smoke || p.solve_mode != shared.solve_mode ||
                p.warm_start != shared.warm_start || p.tolerance != shared.tolerance ||
                p.max_sweeps != shared.max_sweeps || p.check_interval != shared.check_interval ||
                p.sor_omega != shared.sor_omega || p.project_iterations != shared.project_iterations ||
                p.launch_mode != shared.launch_mode)
                invalid("the scenes of a batch must share the step variant and the solver parameters");
            if (p.refinement_steps != 0) invalid("FP64 iterative refinement is not available in a batch");
            if (n[s] % 4 != 0) vec = false;
            if (!vgx::rb_reg_fits(n[s])) reg = false;
        }
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (int s = 0; s < count; ++s) {
            vgx_sim* sim = new vgx_sim();
            sims.push_back(sim);
            sim->n = n[s];
            sim->box = box[s];
            sim->params = params[s];
            sim->batch_owned = true;
            sim->allocate(stream);
        }
        void* host = nullptr;
        check(hipHostMalloc(&host, sizeof(vgx::ProjState) * kRing * count, hipHostMallocDefault), "hipHostMalloc");
        ring = static_cast<vgx::ProjState*>(host);
        check(hipHostGetDevicePointer((void**)&ring_dev, host, 0), "hipHostGetDevicePointer");
        check(hipHostMalloc(&host, sizeof(vgx::ProjState) * 3 * count, hipHostMallocDefault), "hipHostMalloc");
        recs = static_cast<vgx::ProjState*>(host);
        check(hipHostGetDevicePointer((void**)&recs_dev, host, 0), "hipHostGetDevicePointer");
        check(hipHostMalloc(&host, sizeof(long long), hipHostMallocDefault), "hipHostMalloc");
        stamp = static_cast<long long*>(host);
        check(hipHostGetDevicePointer((void**)&stamp_dev, host, 0), "hipHostGetDevicePointer");
        for (auto& e : ring_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        int dev_id = 0, khz = 0;
        check(hipGetDevice(&dev_id), "hipGetDevice");
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev_id), "hipDeviceGetAttribute");
        clock_khz = (double)khz;
        latency_ms.assign(count, -1.0);
        reset();
        build_tables(1.f / 60.f);
        act.owner = device_alloc<uint16_t>((size_t)t.tiles.items, "hipMalloc(active list)");
        act.local = device_alloc<int>((size_t)t.tiles.items, "hipMalloc(active list)");
        act.count = device_alloc<int>(1, "hipMalloc(active list)");
        autotune();
        reset();
    }

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        release_graphs();
        for (vgx_sim* sim : sims) {
            sim->release();
            delete sim;
        }
        sims.clear();
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        for (void* p : {(void*)act.owner, (void*)act.local, (void*)act.count})
            if (p) (void)hipFree(p);
        act = {};
        for (void* p : {(void*)ring, (void*)recs, (void*)stamp})
            if (p) (void)hipHostFree(p);
        ring = recs = nullptr;
        stamp = nullptr;
        for (auto& e : ring_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    void reset() {
        for (vgx_sim* sim : sims) sim->reset();
        rho_cur = 0;
        p_cur[0] = p_cur[1] = 0;
    }

    void release_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    // ------------------------------------------------------------- tables

    // A work list in the blob: the prefix and the item -> scene table, as
    // offsets until the upload.
    static vgx::BatchList list_of(Blob& blob, const std::vector<int>& prefix) {
        std::vector<uint16_t> owner((size_t)prefix.back());
        for (size_t s = 0; s + 1 < prefix.size(); ++s)
            for (int w = prefix[s]; w < prefix[s + 1]; ++w) owner[(size_t)w] = (uint16_t)s;
        vgx::BatchList L;
        L.prefix = reinterpret_cast<const int*>(blob.add(prefix));  // fixed up after upload
        L.owner = reinterpret_cast<const uint16_t*>(blob.add(owner));
        L.scenes = (int)prefix.size() - 1;
        L.items = prefix.back();
        return L;
    }
    std::vector<int> prefix_of(vgx::BatchFamily f, int per_thread) const {
        std::vector<int> p(count + 1, 0);
        for (int s = 0; s < count; ++s) p[s + 1] = p[s] + vgx::batch_items(f, sims[s]->n, per_thread);
        return p;
    }

    // Builds every per-scene argument table for steps of size dt.
    void build_tables(float dt) {
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        release_graphs();
        const bool refl = shared.reflection != 0;
        const float halfDt = 0.5f * dt;
        const float pdt = refl ? halfDt : dt;  // time step of every projection of the step
        Blob blob;
        Tables off;  // offsets, in the pointer fields, until the upload
        auto put = [&](auto** field, const auto& v) {
            *field = reinterpret_cast<std::remove_reference_t<decltype(**field)>*>(blob.add(v));
        };

        for (int which = 0; which < 2; ++which) {
            std::vector<vgx::EmitSlot> e(count);
            std::vector<int> prefix(count + 1, 0);
            for (int s = 0; s < count; ++s) {
                vgx_sim* m = sims[s];
                vgx::EmitSlot& slot = e[s];
                std::memset(&slot, 0, sizeof(slot));
                vgx::EmitArgs a{};
                int points = 0;
                if (m->emitter(dt, which, a) && a.ext[0] * a.ext[1] * a.ext[2] > 0) {
                    slot.active = 1;
                    points = (a.ext[0] + 1) * (a.ext[1] + 1) * (a.ext[2] + 1);
                }
                slot.a = a;
                slot.rng = m->rng;
                slot.noise = m->noise;
                slot.u = m->S[0];
                slot.v = m->S[1];
                slot.w = m->S[2];
                prefix[s + 1] = prefix[s] + (points + vgx::kBlock - 1) / vgx::kBlock;
            }
            for (int rc = 0; rc < 2; ++rc) {
                for (int s = 0; s < count; ++s) e[s].rho = sims[s]->rho[rc];
                put(&off.emit[rc][which], e);
            }
            off.emit_list[which] = list_of(blob, prefix);
        }
        for (int rc = 0; rc < 2; ++rc) {
            std::vector<vgx::ForceArgs> f(count);
            for (int s = 0; s < count; ++s) {
                vgx_sim* m = sims[s];
                f[s] = {m->n, m->rho[rc], m->S[0], m->S[1], m->S[2], m->params.buoyancy, dt, m->params.buoyancy_split};
            }
            put(&off.forces[rc], f);
        }
        for (int pass = 0; pass < 4; ++pass)
            for (int mc = 0; mc < 2; ++mc) {
                std::vector<vgx::AdvArgs> a(count);
                for (int s = 0; s < count; ++s) {
                    vgx_sim* m = sims[s];
                    if (pass == 0) a[s] = m->faces_args(m->S, nullptr, m->S, m->T, refl ? halfDt : dt, mc != 0);
                    else if (pass == 1) a[s] = m->faces_args(m->S, m->T, m->S, m->Q, halfDt, mc != 0);
                    else a[s] = m->smoke_args(dt, pass - 2, mc != 0);
                }
                put(&off.adv[pass][mc], a);
            }
        for (int slot = 0; slot < 2; ++slot) {
            std::vector<vgx::CellArgs> d(count);
            std::vector<vgx::FinArgs> fc(count), fr(count);
            for (int s = 0; s < count; ++s) {
                vgx_sim* m = sims[s];
                ProjectionWorkspace& P = m->proj;
                const float h = m->h();
                float* const* in = (slot == 1 || !refl) ? m->T : m->Q;
                vgx::CellArgs c{};
                c.n = m->n;
                c.h = h;
                c.scale = h * h / pdt;
                c.dt = pdt;
                c.u = in[0];
                c.v = in[1];
                c.w = in[2];
                c.b = P.b;
                c.ps = P.ps_dev + slot;
                c.partials = P.partials;
                d[s] = c;
                const double cells = (double)m->n * m->n * m->n;
                fc[s] = {P.partials, vgx::batch_items(vgx::BatchFamily::Stencil, m->n, 1), cells, P.ps_dev + slot};
                fr[s] = {P.partials, vgx::batch_items(vgx::BatchFamily::Tiles, m->n, 1), cells, P.ps_dev + slot};
            }
            put(&off.div[slot], d);
            put(&off.fin_cell[slot], fc);
            put(&off.fin_rb[slot], fr);
            for (int cur = 0; cur < 2; ++cur) {
                std::vector<vgx::CellArgs> r(count), g(count);
                std::vector<vgx::FillSlot> z(count);
                for (int s = 0; s < count; ++s) {
                    vgx_sim* m = sims[s];
                    ProjectionWorkspace& P = m->proj;
                    r[s] = d[s];
                    r[s].p = P.p[slot][cur];
                    g[s] = d[s];
                    g[s].p = P.p[slot][cur];
                    g[s].p_out = P.p[slot][1 - cur];
                    g[s].uo = m->S[0];
                    g[s].vo = m->S[1];
                    g[s].wo = m->S[2];
                    z[s] = {P.p[slot][cur], m->cells, P.ps_dev + slot};
                }
                put(&off.res[slot][cur], r);
                put(&off.grad[slot][cur], g);
                put(&off.zero[slot][cur], z);
                for (int dir = 0; dir < 2; ++dir) {
                    std::vector<vgx::RbSlot> rb(count);
                    for (int s = 0; s < count; ++s) {
                        vgx_sim* m = sims[s];
                        ProjectionWorkspace& P = m->proj;
                        const int in = cur ^ dir;
                        rb[s] = {m->n, P.b, P.p[slot][in], P.p[slot][in ^ 1], P.ps_dev + slot, P.partials,
                                 vgx::rb_geom(m->n)};
                    }
                    put(&off.rb[slot][cur][dir], rb);
                }
            }
        }
        {
            std::vector<vgx::ProjState*> r(count);
            for (int s = 0; s < count; ++s) r[s] = sims[s]->proj.ps_dev;
            put(&off.records, r);
        }
        for (int k = 0; k < 3; ++k) {
            const auto p = prefix_of(vgx::BatchFamily::Elementwise, 1 << k);
            off.elementwise[k] = list_of(blob, p);
        }
        {
            auto p = prefix_of(vgx::BatchFamily::Cells, 1);
            off.cells = list_of(blob, p);
            p = prefix_of(vgx::BatchFamily::Stencil, 1);
            off.stencil = list_of(blob, p);
            p = prefix_of(vgx::BatchFamily::Tiles, 1);
            off.tiles = list_of(blob, p);
        }
        // upload and relocate
        if (blob.bytes.size() > dev_bytes) {
            if (dev) (void)hipFree(dev);
            dev = nullptr;
            dev_bytes = 0;
            dev = device_alloc<char>(blob.bytes.size(), "hipMalloc(batch tables)");
            dev_bytes = blob.bytes.size();
        }
        check(hipMemcpy(dev, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice), "hipMemcpy(tables)");
        auto rel = [&](auto* p) { return reinterpret_cast<decltype(p)>(dev + reinterpret_cast<size_t>(p)); };
        auto rel_list = [&](vgx::BatchList L) {
            L.prefix = rel(L.prefix);
            L.owner = rel(L.owner);
            return L;
        };
        t = off;
        for (int k = 0; k < 2; ++k) {
            for (int which = 0; which < 2; ++which) t.emit[k][which] = rel(off.emit[k][which]);
            t.forces[k] = rel(off.forces[k]);
            t.emit_list[k] = rel_list(off.emit_list[k]);
            t.div[k] = rel(off.div[k]);
            t.fin_cell[k] = rel(off.fin_cell[k]);
            t.fin_rb[k] = rel(off.fin_rb[k]);
            for (int c = 0; c < 2; ++c) {
                t.res[k][c] = rel(off.res[k][c]);
                t.grad[k][c] = rel(off.grad[k][c]);
                t.zero[k][c] = rel(off.zero[k][c]);
                for (int d = 0; d < 2; ++d) t.rb[k][c][d] = rel(off.rb[k][c][d]);
            }
        }
        for (int pass = 0; pass < 4; ++pass)
            for (int mc = 0; mc < 2; ++mc) t.adv[pass][mc] = rel(off.adv[pass][mc]);
        t.records = rel(off.records);
        for (int k = 0; k < 3; ++k) t.elementwise[k] = rel_list(off.elementwise[k]);
        t.cells = rel_list(off.cells);
        t.stencil = rel_list(off.stencil);
        t.tiles = rel_list(off.tiles);
        table_dt = dt;
    }

    // ------------------------------------------------------------ launches

    vgx::BatchList elementwise() const { return t.elementwise[tune.elementwise_per_thread_log2]; }

    // Advection pass `pass` (see Tables::adv) of fields [first, first + fields).
    void advect(int pass, int fields_first, int fields, bool hat) {
        const bool mc = pass >= 2 ? shared.maccormack_smoke != 0 : shared.maccormack_vel != 0;
        const int kind_sl = pass >= 2 ? vgx::K_ADVECT_SL_SCALAR : vgx::K_ADVECT_SL_FACES;
        {
            Timed tm(timer, stream, kind_sl);
            for (int f = fields_first; f < fields_first + fields; ++f)
                check(vgx::batch_advect(stream, false, f, hat, t.adv[pass][0], elementwise(), tune.advect_grid),
                      "batch advect");
        }
        if (mc) {
            Timed tm(timer, stream, kind_sl + 1);
            for (int f = fields_first; f < fields_first + fields; ++f)
                check(vgx::batch_advect(stream, true, f, hat, t.adv[pass][1], elementwise(), tune.advect_grid),
                      "batch advect");
        }
    }

    SolvePolicy policy(float pdt) const { return sims[0]->policy(pdt, shared.reflection ? std::max(1, shared.project_iterations / 2) : shared.project_iterations); }

    // Projection up to the solve: divergence and RHS mean; a fixed solve
    // runs here whole; a tolerance solve gets its starting check and start.
    void project_begin(int slot, const SolvePolicy& pol) {
        const int cur = p_cur[slot];
        {
            Timed tm(timer, stream, vgx::K_DIVERGENCE);
            check(vgx::batch_cell(stream, vgx::CellOp::Divergence, t.div[slot], t.stencil, tune.stencil_grid),
                  "batch divergence");
        }
        {
            Timed tm(timer, stream, vgx::K_REDUCE);
            check(vgx::batch_finalize(stream, vgx::Finalize::Rhs, t.fin_cell[slot], count), "batch finalize");
        }
        if (!pol.warm) {
            Timed tm(timer, stream, vgx::K_FILL);
            check(vgx::batch_zero(stream, t.zero[slot][cur], t.cells, 0, false), "batch zero");
        }
        auto residual = [&]() {
            {
                Timed tm(timer, stream, vgx::K_RESIDUAL);
                check(vgx::batch_cell(stream, vgx::CellOp::Residual, t.res[slot][cur], t.stencil, tune.stencil_grid),
                      "batch residual");
            }
            Timed tm(timer, stream, vgx::K_REDUCE);
            check(vgx::batch_finalize(stream, vgx::Finalize::Residual, t.fin_cell[slot], count), "batch finalize");
        };
        if (pol.fixed) {
            if (pol.warm) residual();
            if (pol.max_iterations > 0) {
                for (int it = 0; it < pol.max_iterations; ++it) {
                    const bool last = it + 1 == pol.max_iterations;
                    Timed tm(timer, stream, vgx::K_SOR);
                    check(vgx::batch_sor(stream, pol.omega, pol.floor, vec, tune.sweep_register != 0, last ? 2 : 0, nullptr,
                                         t.rb[slot][cur][it & 1], t.tiles, tune.sweep_grid),
                          "batch sor");
                }
                Timed tm(timer, stream, vgx::K_REDUCE);
                check(vgx::batch_finalize(stream, vgx::Finalize::XSum, t.fin_rb[slot], count), "batch finalize");
            } else if (!pol.warm) {
                residual();
            }
            return;
        }
        residual();
        check(vgx::batch_start(stream, t.fin_cell[slot], count, pol.l2_tol, pol.max_iterations, pol.warm, ring_dev,
                               kRing),
              "batch start");
        if (pol.warm) {
            Timed tm(timer, stream, vgx::K_FILL);
            check(vgx::batch_zero(stream, t.zero[slot][cur], t.cells, 0, true), "batch zero");
        }
        check(vgx::batch_active(stream, t.fin_rb[slot], t.tiles, act), "batch active");
    }

    // A fixed (odd count: the iterate ends in the other buffer) number of
    // sweeps leaves the pressure in p[slot][cur ^ (count & 1)].
    void fixed_flip(int slot, const SolvePolicy& pol) {
        if (pol.fixed && (pol.max_iterations & 1)) p_cur[slot] ^= 1;
    }

    void chunk(int slot, const SolvePolicy& pol, int len, bool probes) {
        const int cur = p_cur[slot];
        for (int k = 0; k < len; ++k) {
            const bool last = k + 1 == len;
            std::unique_ptr<Timed> tm;
            if (probes) tm = std::make_unique<Timed>(timer, stream, last ? vgx::K_SOR_CHECK : vgx::K_SOR);
            check(vgx::batch_sor(stream, pol.omega, pol.floor, vec, tune.sweep_register != 0, last ? 3 : 0, &act, t.rb[slot][cur][k & 1],
                                 t.tiles, tune.sweep_grid),
                  "batch sor");
        }
        std::unique_ptr<Timed> tm;
        if (probes) tm = std::make_unique<Timed>(timer, stream, vgx::K_REDUCE);
        check(vgx::batch_check(stream, t.fin_rb[slot], count, pol.l2_tol, len, ring_dev, kRing), "batch check");
        check(vgx::batch_active(stream, t.fin_rb[slot], t.tiles, act), "batch active");
    }

    void project_end(int slot, const SolvePolicy& pol) {
        (void)pol;
        const int cur = p_cur[slot];
        {
            Timed tm(timer, stream, vgx::K_GRADIENT);
            check(vgx::batch_cell(stream, vgx::CellOp::Gradient, t.grad[slot][cur], t.stencil, tune.stencil_grid),
                  "batch gradient");
        }
        Timed tm(timer, stream, vgx::K_REDUCE);
        check(vgx::batch_finalize(stream, vgx::Finalize::Final, t.fin_cell[slot], count), "batch finalize");
    }

    // Runs `body` (which enqueues on the stream) eagerly, or through a graph
    // captured for `key` and replayed.
    template <class F> void run(const std::array<int, 6>& key, bool eager, F&& body) {
        if (eager) {
            body();
            return;
        }
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipGraph_t g = nullptr;
            try {
                body();
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
            it = graphs.emplace(key, exec).first;
        }
        check(hipGraphLaunch(it->second, stream), "hipGraphLaunch");
    }

    // The chunk loop of a tolerance solve; the start record (check 0) is
    // already enqueued. Returns after every scene is done or the budget ends.
    void solve(int slot, const SolvePolicy& pol, bool graphs_on) {
        const int len = ProjectionWorkspace::chunk_of(pol);
        const int max_chunks = pol.max_iterations > 0 ? (pol.max_iterations + len - 1) / len : 0;
        check(hipEventRecord(ring_ev[0], stream), "hipEventRecord");
        int enqueued = 1, seen = 0;
        for (;;) {
            while (enqueued < max_chunks + 1 && enqueued - seen < kLookahead) {
                const bool eager = !graphs_on || (timer.enabled && enqueued % 8 == 1);
                if (eager) {
                    chunk(slot, pol, len, true);
                } else {
                    run({4, slot, p_cur[slot], len, 0, 0}, false, [&] { chunk(slot, pol, len, false); });
                    timer.count(vgx::K_SOR, (uint64_t)(len - 1));
                    timer.count(vgx::K_SOR_CHECK, 1);
                    timer.count(vgx::K_REDUCE, 1);
                }
                check(hipEventRecord(ring_ev[enqueued % kRing], stream), "hipEventRecord");
                ++enqueued;
            }
            check(hipEventSynchronize(ring_ev[seen % kRing]), "hipEventSynchronize");
            const vgx::ProjState* rows = ring + (size_t)(seen % kRing) * count;
            ++seen;
            bool all = true;
            for (int s = 0; s < count && all; ++s) all = rows[s].done != 0;
            if (all || seen == max_chunks + 1) break;
        }
    }

    // ---------------------------------------------------------------- step

    void advance(float dt) {
        const bool refl = shared.reflection != 0;
        const float pdt = refl ? 0.5f * dt : dt;
        for (vgx_sim* sim : sims) {
            validate_step(sim->params, dt);
            const float h = sim->h(), scale = h * h / pdt;
            if (!std::isfinite(scale) || scale <= 0.f) invalid("Projection scale is outside the supported float range");
        }
        if (dt != table_dt) build_tables(dt);
        const SolvePolicy pol = policy(pdt);
        const bool graphs_on = pol.graphs;
        const bool eager = !graphs_on || timer.enabled;
        const int first_slot = refl ? 1 : 0;

        // emit, forces, first advection, first projection up to its solve
        run({0, rho_cur, p_cur[0], p_cur[1], 0, 0}, eager, [&] {
            check(vgx::batch_stamp(stream, stamp_dev), "batch stamp");
            for (int which = 0; which < 2; ++which) {
                if (which == 1 && t.emit_list[1].items == 0) break;
                Timed tm(timer, stream, vgx::K_EMIT);
                check(vgx::batch_emit(stream, t.emit[rho_cur][which], count, t.emit_list[which], 0), "batch emit");
            }
            {
                Timed tm(timer, stream, vgx::K_FORCES);
                check(vgx::batch_forces(stream, t.forces[rho_cur], elementwise(), tune.advect_grid), "batch forces");
            }
            advect(0, 0, 3, false);
            project_begin(first_slot, pol);
        });
        if (!pol.fixed) solve(first_slot, pol, graphs_on);
        fixed_flip(first_slot, pol);
        if (refl) {
            run({1, rho_cur, p_cur[0], p_cur[1], 0, 0}, eager, [&] {
                project_end(1, pol);
                advect(1, 0, 3, true);
                project_begin(0, pol);
            });
            p_cur[1] ^= 1;
            if (!pol.fixed) solve(0, pol, graphs_on);
            fixed_flip(0, pol);
        }
        run({2, rho_cur, p_cur[0], p_cur[1], 0, 0}, eager, [&] {
            project_end(0, pol);
            advect(2 + rho_cur, 3, 1, false);
            check(vgx::batch_gather(stream, t.records, count, 3, recs_dev), "batch gather");
        });
        p_cur[0] ^= 1;
        rho_cur ^= 1;
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        timer.flush();
        finish(pol, refl);
    }

    // Per-scene results of the step, as the single-scene step reports them.
    void finish(const SolvePolicy& pol, bool refl) {
        const int len = ProjectionWorkspace::chunk_of(pol);
        for (int s = 0; s < count; ++s) {
            vgx_sim* m = sims[s];
            m->rho_cur = rho_cur;
            m->proj.p_cur[0] = p_cur[0];
            m->proj.p_cur[1] = p_cur[1];
            std::memcpy(m->proj.ps_host, recs + (size_t)s * 3, 3 * sizeof(vgx::ProjState));
            std::vector<ProjectionWorkspace::Pending> pend;
            for (int slot : refl ? std::vector<int>{1, 0} : std::vector<int>{0}) {
                const vgx::ProjState& r = m->proj.ps_host[slot];
                if (!pol.fixed && r.nonfinite_x)
                    invalid("Invalid stencil coefficient, boundary entry, or vector value");
                ProjectionWorkspace::Pending p{slot, pol, 0, 0.0, false, 0.0, 0, 0, -1.0, pol.warm};
                if (pol.fixed) {
                    p.iterations = pol.max_iterations;
                    p.initial_from_residual = pol.warm;
                } else {
                    p.iterations = r.sweeps_run;
                    p.initial_from_residual = true;
                    p.initial_inf_warm = r.start_inf;
                    p.initial_l2 = r.start_rel;
                    p.checks = 1 + r.sweeps_run / len;
                }
                m->proj.sweeps += (uint64_t)p.iterations;
                pend.push_back(p);
            }
            vgx_projection_result results[2];
            for (size_t k = 0; k < pend.size(); ++k) results[k] = m->proj.finish(pend[k]);
            m->last[0] = results[pend.size() - 1];
            m->last[1] = results[0];
            // the scene's last projection reached its tolerance this long after
            // the step began (-1: fixed sweeps, or the budget ran out)
            const vgx::ProjState& lastrec = m->proj.ps_host[0];
            latency_ms[s] = (!pol.fixed && lastrec.done && clock_khz > 0.0)
                                ? (double)(lastrec.done_clock - *stamp) / clock_khz
                                : -1.0;
        }
    }

    // ----------------------------------------------------------- autotuner

    // Mean milliseconds of `body` over kTuneReps launches after one warm-up.
    template <class F> double time_ms(F&& body) {
        hipEvent_t a, b;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
        body();
        check(hipEventRecord(a, stream), "hipEventRecord");
        for (int r = 0; r < kTuneReps; ++r) body();
        check(hipEventRecord(b, stream), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return (double)ms / kTuneReps;
    }

    // Times the candidates on the batch's zeroed fields: advection reads the
    // state and writes only the first-pass scratch T and F, the sweep reads
    // and writes the (zero) pressures of slot 0, the residual writes only
    // partials. Nothing a step reads is changed.
    void autotune() {
        std::memset(&tune, 0, sizeof(tune));
        const int grids[4] = {0, 4096, 2048, 1024};
        static_assert(sizeof(tune.advect_ms) / sizeof(double) == 12, "candidate table size");
        double best = 1e30;
        for (int k = 0; k < 3; ++k)
            for (int g = 0; g < 4; ++g) {
                const double ms = time_ms([&] {
                    check(vgx::batch_advect(stream, true, 0, false, t.adv[0][1], t.elementwise[k], grids[g]),
                          "tune advect");
                });
                tune.advect_ms[k * 4 + g] = ms;
                if (ms < best) {
                    best = ms;
                    tune.elementwise_per_thread_log2 = k;
                    tune.advect_grid = grids[g];
                }
            }
        // The sweep: the two staging variants (bit-identical) at each grid.
        const int sweep_grids[4] = {0, 2048, 1024, 512};
        best = 1e30;
        for (int variant = reg ? 1 : 0; variant >= 0; --variant)
            for (int g = 0; g < 4; ++g) {
                const double ms = time_ms([&] {
                    for (int c = 0; c < 2; ++c)
                        check(vgx::batch_sor(stream, shared.sor_omega, 6.0f, vec, variant == 1, 0, nullptr,
                                             t.rb[0][0][c], t.tiles, sweep_grids[g]),
                              "tune sor");
                }) / 2.0;
                (variant ? tune.sweep_ms : tune.sweep_lds_ms)[g] = ms;
                if (ms < best) {
                    best = ms;
                    tune.sweep_grid = sweep_grids[g];
                    tune.sweep_register = variant;
                }
            }
        best = 1e30;
        for (int g = 0; g < 4; ++g) {
            const double ms = time_ms([&] {
                check(vgx::batch_cell(stream, vgx::CellOp::Residual, t.res[0][0], t.stencil, grids[g]), "tune residual");
            });
            tune.stencil_ms[g] = ms;
            if (ms < best) {
                best = ms;
                tune.stencil_grid = grids[g];
            }
        }
        tune.tile_items = t.tiles.items;
        tune.stencil_items = t.stencil.items;
        tune.elementwise_items = elementwise().items;
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
};

// ====================================================================== C ABI

extern "C" {

int vgx_batch_create(int count, const int* n, const float* box, const vgx_params* params, vgx_batch** out) {
    if (!out) return fail(VGX_EINVAL, "null output handle");
    *out = nullptr;
    vgx_batch* b = nullptr;
    const int rc = guarded([&] {
        b = new vgx_batch();
        b->create(count, n, box, params);
    });
    if (rc != VGX_OK) {
        if (b) {
            b->release();
            delete b;
        }
        return rc;
    }
    *out = b;
    return VGX_OK;
}

void vgx_batch_destroy(vgx_batch* batch) {
    if (!batch) return;
    batch->release();
    delete batch;
}

int vgx_batch_size(const vgx_batch* batch) { return batch ? batch->count : 0; }

int vgx_batch_scene(vgx_batch* batch, int index, vgx_sim** scene) {
    if (!batch || !scene || index < 0 || index >= batch->count) return fail(VGX_EINVAL, "invalid argument");
    *scene = batch->sims[index];
    return VGX_OK;
}

int vgx_batch_advance(vgx_batch* batch, float dt) {
    if (!batch) return fail(VGX_EINVAL, "null batch");
    return guarded([&] { batch->advance(dt); });
}

int vgx_batch_reset(vgx_batch* batch) {
    if (!batch) return fail(VGX_EINVAL, "null batch");
    return guarded([&] {
        batch->reset();
        check(hipStreamSynchronize(batch->stream), "hipStreamSynchronize");
    });
}

int vgx_batch_scene_latency(const vgx_batch* batch, double* ms) {
    if (!batch || !ms) return fail(VGX_EINVAL, "invalid argument");
    for (int s = 0; s < batch->count; ++s) ms[s] = batch->latency_ms[s];
    return VGX_OK;
}

int vgx_batch_get_tuning(const vgx_batch* batch, vgx_batch_tuning* out) {
    if (!batch || !out) return fail(VGX_EINVAL, "invalid argument");
    *out = batch->tune;
    return VGX_OK;
}

int vgx_batch_set_timing(vgx_batch* batch, int enabled) {
    if (!batch) return fail(VGX_EINVAL, "null batch");
    batch->timer.enabled = enabled != 0;
    return VGX_OK;
}

int vgx_batch_reset_timing(vgx_batch* batch) {
    if (!batch) return fail(VGX_EINVAL, "null batch");
    batch->timer.reset();
    return VGX_OK;
}

int vgx_batch_kernel_timing(vgx_batch* batch, int kind, double* total_ms, uint64_t* launches) {
    if (!batch || kind < 0 || kind >= vgx::K_COUNT) return fail(VGX_EINVAL, "invalid argument");
    if (total_ms) *total_ms = batch->timer.total_ms(kind);
    if (launches) *launches = batch->timer.launches[kind];
    return VGX_OK;
}

int vgx_batch_synchronize(vgx_batch* batch) {
    if (!batch) return fail(VGX_EINVAL, "null batch");
    return guarded([&] { check(hipStreamSynchronize(batch->stream), "hipStreamSynchronize"); });
}

}  // extern "C"
