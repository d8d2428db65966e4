// vgx_shadow.hip - shadow runs: a copy of a simulation stepped by plain
// kernels in a chosen precision (include/vapor_mac3d.h, vgx_shadow_*).
//
// The FP64 shadow is the all-FP64 run of the same step that the FP32 path's
// error is measured against: vapor's CPU step (MacGridFluidSolver3D.cpp:
// 192-245) with every field, every intermediate and every solve in double,
// the SOR solving to the tolerance in FP64 itself (no refinement needed). The
// FP32 shadow runs the same kernels in float: vapor's FP32 step by a second
// route, which pins the shared kernels against the FP32 path bit for bit.
//
// Each operation is one grid-stride kernel over the reference's linear
// layouts, in the reference's operation order; the emitter, forces and
// advection share the FP32 path's device bodies. Reductions leave FP64
// per-block partials that the host folds in block order after each launch
// that needs a value, so every shadow run is deterministic.
#include "vgx_sim.hpp"

#include <array>

using namespace vgx_impl;

namespace {

struct ShadowBase {
    virtual ~ShadowBase() = default;
    virtual void advance(float dt) = 0;
    virtual void download(double* rho, double* u, double* v, double* w, double* p) = 0;
    virtual void compare(vgx_sim* sim, vgx_field_error* errors) = 0;
    vgx_projection_result last[2]{};
    uint64_t sweeps = 0;
};

template <class T> struct Shadow final : ShadowBase {
    int n = 0;
    T box = 0;
    vgx_params params{};
    hipStream_t stream = nullptr;
    size_t cells = 0, faces = 0;
    T* rho[2] = {};
    int rho_cur = 0;
    T* S[3] = {};   // state faces
    T* U0[3] = {};  // flow of the first advection (state after forces)
    T* UT[3] = {};  // first advection result, u tilde (reflection)
    T* UD[3] = {};  // state after the first projection (reflection)
    T* F = nullptr;      // forward trace of MacCormack, one field at a time
    T* noise = nullptr;  // emitter stir values over the patch
    T* p[2] = {};        // per projection slot: 1 the reflection step's first, 0 otherwise
    T* b = nullptr;
    uint32_t* rng = nullptr;
    double* partials = nullptr;
    double* host_partials = nullptr;

    T h() const { return box / (T)n; }

    ~Shadow() override { release(); }

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (T* q : rho)
            if (q) (void)hipFree(q);
        for (auto* set : {S, U0, UT, UD})
            for (int a = 0; a < 3; ++a)
                if (set[a]) (void)hipFree(set[a]);
        for (void* q : {(void*)F, (void*)noise, (void*)p[0], (void*)p[1], (void*)b, (void*)rng, (void*)partials})
            if (q) (void)hipFree(q);
        if (host_partials) (void)hipHostFree(host_partials);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // Copies grid, parameters, fields, emitter RNG state and pressure
    // histories of `src` (FP32 values convert exactly to double).
    void create_from(vgx_sim* src) {
        n = src->n;
        box = (T)src->box;
        params = src->params;
        cells = src->cells;
        faces = src->faces;
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& r : rho) r = device_alloc<T>(cells, "hipMalloc(shadow rho)");
        for (auto* set : {S, U0, UT, UD})
            for (int a = 0; a < 3; ++a) set[a] = device_alloc<T>(faces, "hipMalloc(shadow faces)");
        F = device_alloc<T>(faces, "hipMalloc(shadow forward)");
        noise = device_alloc<T>(cells, "hipMalloc(shadow noise)");
        for (auto& q : p) q = device_alloc<T>(cells, "hipMalloc(shadow pressure)");
        b = device_alloc<T>(cells, "hipMalloc(shadow rhs)");
        rng = device_alloc<uint32_t>(1, "hipMalloc(shadow rng)");
        const size_t cap = vgx::partials_capacity(n);
        partials = device_alloc<double>(cap, "hipMalloc(shadow partials)");
        void* host = nullptr;
        check(hipHostMalloc(&host, cap * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        host_partials = static_cast<double*>(host);
        check(hipStreamSynchronize(src->stream), "hipStreamSynchronize");
        check(vgx::wide::convert(stream, rho[0], src->rho[src->rho_cur], cells), "convert");
        for (int a = 0; a < 3; ++a) check(vgx::wide::convert(stream, S[a], src->S[a], faces), "convert");
        for (int slot = 0; slot < 2; ++slot)
            check(vgx::wide::convert(stream, p[slot], src->proj.current(slot), cells), "convert");
        check(hipMemcpyAsync(rng, src->rng, sizeof(uint32_t), hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync(rng)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::memcpy(last, src->last, sizeof(last));
    }

    // Folds partial k of the last reduction over its blocks in block order.
    void fetch(int nblocks) {
        check(hipMemcpyAsync(host_partials, partials, (size_t)nblocks * vgx::kPartialStride * sizeof(double),
                             hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(partials)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    double fold_sum(int nblocks, int k) const {
        double acc = 0.0;
        for (int i = 0; i < nblocks; ++i) acc += host_partials[(size_t)i * vgx::kPartialStride + k];
        return acc;
    }
    double fold_max(int nblocks, int k) const {
        double acc = host_partials[k];
        for (int i = 1; i < nblocks; ++i) {
            const double v = host_partials[(size_t)i * vgx::kPartialStride + k];
            acc = acc < v ? v : acc;
        }
        return acc;
    }

    void copy3(T* const dst[3], T* const src[3]) {
        for (int a = 0; a < 3; ++a)
            check(hipMemcpyAsync(dst[a], src[a], faces * sizeof(T), hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync(faces)");
    }

    // emit (MacGridFluidSolver3D.cpp:140-165) for one sphere: every quantity
    // of the in-sphere test and the patch bounds in precision T.
    void emit_sphere(T dt, float cxf, float cyf, float czf, float radius) {
        const vgx_params& P = params;
        vgx::EmitArgsT<T> a{};
        a.n = n;
        a.dx = h();
        a.dt_scale = dt * (T)60;
        a.ecx = (T)cxf * box;
        a.ecy = (T)cyf * box;
        a.ecz = (T)czf * box;
        a.erad = (T)radius * box;
        const int emitR = (int)(a.erad / a.dx + 1);
        const int c[3] = {(int)(a.ecx / a.dx), (int)(a.ecy / a.dx), (int)(a.ecz / a.dx)};
        for (int k = 0; k < 3; ++k) {
            a.lo[k] = std::max(0, c[k] - emitR);
            a.ext[k] = std::max(0, std::min(n, c[k] + emitR) - a.lo[k]);
        }
        if ((size_t)a.ext[0] * a.ext[1] * a.ext[2] > cells) invalid("emitter patch exceeds the grid");
        const T strength = (T)P.source_strength;
        a.profile = P.emitter_profile;
        a.kick_on = P.emitter_kick_multiplier > 0.f;
        a.kick = (T)P.emitter_kick_multiplier * ((T)P.emitter_kick_base + (T)P.emitter_kick_scale * strength);
        a.stir = (T)P.stir_strength;
        a.strength = strength;
        check(vgx::wide::emit(stream, a, rng, noise, rho[rho_cur], S[0], S[1], S[2]), "shadow emit");
    }
    void emit(T dt) {
        const vgx_params& P = params;
        if (P.source_strength <= 0.f) return;
        emit_sphere(dt, P.emitter_center_x, P.emitter_center_y, P.emitter_center_z, P.emitter_radius);
        if (P.emitter2_enabled)
            emit_sphere(dt, P.emitter2_center_x, P.emitter2_center_y, P.emitter2_center_z, P.emitter2_radius);
    }

    // Advection of the three faces of src (2 src - tsrc with tsrc) by flow
    // into out (CpuAdvectionKernels.cpp:82-269): per axis the semi-Lagrangian
    // pass, then under MacCormack the corrected pass from its forward trace.
    void advect_faces(T* const src[3], T* const tsrc[3], T* const flow[3], T* const out[3], T dt) {
        const bool mc = params.maccormack_vel != 0;
        for (int a = 0; a < 3; ++a) {
            vgx::AdvArgsT<T> args{};
            args.u = flow[0];
            args.v = flow[1];
            args.w = flow[2];
            args.n = n;
            args.h = h();
            args.box = box;
            args.dt = dt;
            args.cfl = (T)params.cfl_mc;
            args.out_scale = (T)1;
            const T* ts = tsrc ? tsrc[a] : nullptr;
            args.f[a] = {src[a], ts, F, mc ? F : out[a]};
            check(vgx::wide::advect(stream, args, a, false, tsrc != nullptr), "shadow advect");
            if (mc) {
                args.f[a] = {src[a], ts, F, out[a]};
                check(vgx::wide::advect(stream, args, a, true, tsrc != nullptr), "shadow advect");
            }
        }
    }

    // Smoke advection by the final flow, then applyDissipation
    // (MacGridFluidSolver3D.cpp:183-190) folded into the last write.
    void advect_smoke(T dt) {
        const bool mc = params.maccormack_smoke != 0;
        vgx::AdvArgsT<T> args{};
        args.u = S[0];
        args.v = S[1];
        args.w = S[2];
        args.n = n;
        args.h = h();
        args.box = box;
        args.dt = dt;
        args.cfl = (T)params.cfl_mc;
        args.out_scale = vgx::clampv((T)1 - (T)params.smoke_decay * dt, (T)0, (T)1);
        T* dst = rho[1 - rho_cur];
        args.f[3] = {rho[rho_cur], nullptr, F, mc ? F : dst};
        args.scale_out = !mc;
        check(vgx::wide::advect(stream, args, 3, false, false), "shadow advect");
        if (mc) {
            args.f[3] = {rho[rho_cur], nullptr, F, dst};
            args.scale_out = 1;
            check(vgx::wide::advect(stream, args, 3, true, false), "shadow advect");
        }
        rho_cur = 1 - rho_cur;
    }

    void sweep(T omega, T* x) {
        for (int parity = 0; parity < 2; ++parity)
            check(vgx::wide::sor_color(stream, n, omega, (T)6, b, x, parity), "shadow sor");
        ++sweeps;
    }

    struct Residual {
        double sum, sum2, inf, bsum, b2, nonfinite;
    };
    Residual residual(const T* x) {
        int nb = 0;
        check(vgx::wide::residual(stream, n, b, x, partials, &nb), "shadow residual");
        fetch(nb);
        return {fold_sum(nb, 0), fold_sum(nb, 1), fold_max(nb, 2), fold_sum(nb, 3), fold_sum(nb, 4), fold_sum(nb, 5)};
    }
    double rel_of(const Residual& r) const { return compat_rel(r.sum2, r.sum, r.b2, r.bsum, (double)cells); }

    // CpuPressureSolver3D::project (CpuPressureSolver3D.cpp:25-60) on S in
    // place, with the solver's policy (MacGridFluidSolver3D.cpp:131-138): a
    // fixed sweep count, or the relative L2 tolerance checked every
    // check_interval sweeps (rounded up to even) on the iterate a chunk's last
    // sweep starts from, as the FP32 path checks it.
    vgx_projection_result project(int slot, T dt, int iterations) {
        vgx_projection_result res;
        std::memset(&res, 0, sizeof(res));
        const T hh = h();
        const T scale = hh * hh / dt;
        T* x = p[slot];
        check(vgx::wide::walls(stream, n, S[0], S[1], S[2]), "shadow walls");
        int nb = 0;
        check(vgx::wide::divergence(stream, n, hh, scale, S[0], S[1], S[2], b, partials, &nb), "shadow divergence");
        fetch(nb);
        if (fold_sum(nb, 2) > 0.0) invalid("Non-finite divergence");
        res.divergence_before = (float)fold_max(nb, 1);
        const double mean = fold_sum(nb, 0) / (double)cells;
        res.removed_rhs_mean = mean;
        check(vgx::wide::shift(stream, b, (T)mean, cells), "shadow shift");
        const bool warm = params.warm_start == VGX_WARM_PREVIOUS;
        res.warm_started = warm;
        if (!warm) check(hipMemsetAsync(x, 0, cells * sizeof(T), stream), "hipMemsetAsync");
        const T omega = (T)params.sor_omega;
        const bool tolerance = params.solve_mode == VGX_SOLVE_TOLERANCE;
        if (!tolerance) {
            for (int k = 0; k < iterations; ++k) sweep(omega, x);
            res.iterations = iterations;
        } else {
            Residual r = residual(x);
            res.residual_checks = 1;
            if (r.nonfinite > 0.0) invalid("Invalid stencil coefficient, boundary entry, or vector value");
            res.initial_residual = r.inf;
            const double bnorm2 = compat_norm2(r.b2, r.bsum, (double)cells);
            double rel = 0.0;
            if (!(bnorm2 > 0.0)) {
                check(hipMemsetAsync(x, 0, cells * sizeof(T), stream), "hipMemsetAsync");
            } else {
                rel = rel_of(r);
                res.initial_residual_l2 = rel;
                const int chunk = std::max(2, params.check_interval + (params.check_interval & 1));
                const int budget = params.max_sweeps;
                int used = 0;
                while (rel > params.tolerance && used < budget) {
                    for (int k = 0; k + 1 < chunk; ++k) sweep(omega, x);
                    r = residual(x);
                    ++res.residual_checks;
                    rel = rel_of(r);
                    sweep(omega, x);
                    used += chunk;
                }
                res.iterations = used;
            }
        }
        // zero-mean gauge (CpuPressureSolver3D.cpp:50-51)
        check(vgx::wide::sum(stream, x, cells, partials, &nb), "shadow sum");
        fetch(nb);
        check(vgx::wide::shift(stream, x, (T)(fold_sum(nb, 0) / (double)cells), cells), "shadow shift");
        const Residual fin = residual(x);
        res.final_residual = fin.inf;
        res.final_residual_l2 = rel_of(fin);
        res.residual_available = 1;
        res.converged = tolerance ? res.final_residual_l2 <= params.tolerance : 1;
        check(vgx::wide::gradient(stream, n, (T)1 / hh, dt, x, S[0], S[1], S[2]), "shadow gradient");
        check(vgx::wide::divergence(stream, n, hh, scale, S[0], S[1], S[2], (T*)nullptr, partials, &nb),
              "shadow divergence");
        fetch(nb);
        if (fold_sum(nb, 2) > 0.0) invalid("Non-finite MAC velocity");
        res.divergence_after = (float)fold_max(nb, 1);
        return res;
    }

    // step / stepNoReflect (MacGridFluidSolver3D.cpp:192-245) in precision T.
    void advance(float dt_in) override {
        validate_step(params, dt_in);
        const T dt = (T)dt_in;
        emit(dt);
        check(vgx::wide::forces(stream, n, (const T*)rho[rho_cur], S[0], S[1], S[2], (T)params.buoyancy, dt,
                                (T)params.buoyancy_split),
              "shadow forces");
        copy3(U0, S);
        vgx_projection_result results[2];
        int count = 0;
        if (params.reflection) {
            const T halfDt = (T)0.5 * dt;
            const int halfIters = std::max(1, params.project_iterations / 2);
            advect_faces(U0, nullptr, U0, S, halfDt);
            copy3(UT, S);
            results[count++] = project(1, halfDt, halfIters);
            copy3(UD, S);
            advect_faces(UD, UT, UD, S, halfDt);
            results[count++] = project(0, halfDt, halfIters);
        } else {
            advect_faces(U0, nullptr, U0, S, dt);
            results[count++] = project(0, dt, params.project_iterations);
        }
        advect_smoke(dt);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        last[0] = results[count - 1];
        last[1] = results[0];
    }

    void download(double* drho, double* du, double* dv, double* dw, double* dp) override {
        double* dst[5] = {drho, du, dv, dw, dp};
        const T* src[5] = {rho[rho_cur], S[0], S[1], S[2], p[0]};
        std::vector<T> tmp;
        for (int k = 0; k < 5; ++k) {
            if (!dst[k]) continue;
            const size_t count = (k == 0 || k == 4) ? cells : faces;
            tmp.resize(count);
            check(hipMemcpyAsync(tmp.data(), src[k], count * sizeof(T), hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync(download)");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            for (size_t i = 0; i < count; ++i) dst[k][i] = (double)tmp[i];
        }
    }

    void compare(vgx_sim* sim, vgx_field_error* errors) override {
        if (sim->n != n) invalid("shadow and simulation grids differ");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
        const float* got[5] = {sim->rho[sim->rho_cur], sim->S[0], sim->S[1], sim->S[2], sim->proj.current(0)};
        const T* want[5] = {rho[rho_cur], S[0], S[1], S[2], p[0]};
        for (int k = 0; k < 5; ++k) {
            const size_t count = (k == 0 || k == 4) ? cells : faces;
            int nb = 0;
            check(vgx::wide::compare(stream, got[k], want[k], count, partials, &nb), "shadow compare");
            fetch(nb);
            errors[k].max_abs = fold_max(nb, 0);
            errors[k].rms = std::sqrt(fold_sum(nb, 1) / (double)count);
            errors[k].ref_max = fold_max(nb, 2);
            errors[k].ref_rms = std::sqrt(fold_sum(nb, 3) / (double)count);
        }
    }
};

}  // namespace

struct vgx_shadow {
    int precision = 0;
    std::unique_ptr<ShadowBase> impl;
};

extern "C" {

int vgx_shadow_create(vgx_sim* source, int precision, vgx_shadow** out) {
    if (!source || !out) return fail(VGX_EINVAL, "null argument");
    if (precision != VGX_PRECISION_FP32 && precision != VGX_PRECISION_FP64)
        return fail(VGX_EINVAL, "shadow precision must be VGX_PRECISION_FP32 or VGX_PRECISION_FP64");
    *out = nullptr;
    return guarded([&] {
        auto s = std::make_unique<vgx_shadow>();
        s->precision = precision;
        if (precision == VGX_PRECISION_FP64) {
            auto impl = std::make_unique<Shadow<double>>();
            impl->create_from(source);
            s->impl = std::move(impl);
        } else {
            auto impl = std::make_unique<Shadow<float>>();
            impl->create_from(source);
            s->impl = std::move(impl);
        }
        *out = s.release();
    });
}

void vgx_shadow_destroy(vgx_shadow* shadow) { delete shadow; }

int vgx_shadow_advance(vgx_shadow* shadow, float dt) {
    if (!shadow) return fail(VGX_EINVAL, "null shadow");
    return guarded([&] { shadow->impl->advance(dt); });
}

int vgx_shadow_download(vgx_shadow* shadow, double* rho, double* u, double* v, double* w, double* p) {
    if (!shadow) return fail(VGX_EINVAL, "null shadow");
    return guarded([&] { shadow->impl->download(rho, u, v, w, p); });
}

int vgx_shadow_last_projection(const vgx_shadow* shadow, int index, vgx_projection_result* result) {
    if (!shadow || !result) return fail(VGX_EINVAL, "null argument");
    if (index != 0 && index != 1) return fail(VGX_EINVAL, "projection index must be 0 or 1");
    *result = shadow->impl->last[index];
    return VGX_OK;
}

int vgx_shadow_sweep_count(const vgx_shadow* shadow, uint64_t* sweeps) {
    if (!shadow || !sweeps) return fail(VGX_EINVAL, "null argument");
    *sweeps = shadow->impl->sweeps;
    return VGX_OK;
}

int vgx_shadow_compare(vgx_shadow* shadow, vgx_sim* sim, vgx_field_error errors[5]) {
    if (!shadow || !sim || !errors) return fail(VGX_EINVAL, "null argument");
    return guarded([&] { shadow->impl->compare(sim, errors); });
}

}  // extern "C"
