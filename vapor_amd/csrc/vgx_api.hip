// vgx_api.hip - C ABI of the MAC 3D smoke step (include/vapor_mac3d.h).
//
// Host orchestration: one HIP stream per handle, all fields resident in HBM,
// ping-pong buffers instead of copies. A step enqueues every kernel without a
// host round trip; the host reads the small projection records once at the end
// of the step (and at each residual check when solving to a tolerance).
#include "vgx_internal.hpp"
#include "vgx_host.hpp"
#include "vgx_unet.hpp"

#include "../../include/vapor_mac3d.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <new>
#include <string>
#include <tuple>
#include <memory>
#include <map>
#include <mutex>
#include <vector>

namespace vgx_host {
std::string& error_slot() {
    static thread_local std::string message;
    return message;
}
}  // namespace vgx_host

#include "vgx_sim.hpp"

// ==================================================================== backend

struct vgx_backend {
    int n = 0;
    float box = 0.f;
    hipStream_t stream = nullptr;
    size_t cells = 0, faces = 0;
    float* flow[3] = {};
    float* src = nullptr;
    float* out = nullptr;
    float* fwd = nullptr;
    float* in3[3] = {};
    ProjectionWorkspace proj;
    Timer timer;

    void allocate() {
        cells = (size_t)n * n * n;
        faces = (size_t)(n + 1) * n * n;
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& f : flow) f = device_alloc<float>(faces, "hipMalloc");
        for (auto& f : in3) f = device_alloc<float>(faces, "hipMalloc");
        src = device_alloc<float>(faces, "hipMalloc");
        out = device_alloc<float>(faces, "hipMalloc");
        fwd = device_alloc<float>(faces, "hipMalloc");
        proj.allocate(n, box, stream);
        proj.clear(timer);
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& f : flow)
            if (f) (void)hipFree(f);
        for (auto& f : in3)
            if (f) (void)hipFree(f);
        for (void* q : {(void*)src, (void*)out, (void*)fwd})
            if (q) (void)hipFree(q);
        proj.release();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    void upload(float* dst, const float* host, size_t count) {
        check(hipMemcpyAsync(dst, host, count * sizeof(float), hipMemcpyHostToDevice, stream), "hipMemcpyAsync(H2D)");
    }
    void download(float* host, const float* dev, size_t count) {
        check(hipMemcpyAsync(host, dev, count * sizeof(float), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(D2H)");
    }
};

namespace {

bool overlaps(const float* a, size_t na, const float* b, size_t nb) {
    if (!a || !b || !na || !nb) return false;
    const std::less<const float*> less;
    return less(a, b + nb) && less(b, a + na);
}

// AdvectionContext validation (AdvectionContext.cpp:15-67).
void validate_advection(float dt, float cfl) {
    if (!std::isfinite(dt) || dt < 0.f || !std::isfinite(cfl) || cfl <= 0.f)
        invalid("Invalid advection time step, CFL limit, or tracer field sizes");
}

}  // namespace

// ------------------------------------------------------------ snapshots
namespace {

// The snapshot's files: a JSON header beside one raw little-endian FP32 file
// per field, in the reference layouts. The density file holds exactly the
// bytes vapor's raw density dump writes (ReflectionComparison.hpp:74-75).
struct SnapshotField {
    const char* name;
    const char* suffix;
    const char* layout;
};
constexpr SnapshotField kSnapshotFields[6] = {
    {"density", ".density", "x + n*(y + n*z)"},
    {"velocity_x", ".velx", "x + (n+1)*(y + n*z)"},
    {"velocity_y", ".vely", "x + n*(y + (n+1)*z)"},
    {"velocity_z", ".velz", "x + n*(y + n*z), z in [0, n]"},
    {"pressure", ".pressure", "x + n*(y + n*z); the step's last projection"},
    {"pressure_first", ".pressure1", "x + n*(y + n*z); the first projection of a reflection step"},
};

std::string base_name(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    return slash == std::string::npos ? path : path.substr(slash + 1);
}

// The number after "key": in the header (the header is this library's own).
double header_number(const std::string& text, const std::string& key) {
    const size_t at = text.find("\"" + key + "\"");
    if (at == std::string::npos) invalid("snapshot header lacks " + key);
    const size_t colon = text.find(':', at);
    if (colon == std::string::npos) invalid("snapshot header lacks " + key);
    char* end = nullptr;
    const double v = std::strtod(text.c_str() + colon + 1, &end);
    if (end == text.c_str() + colon + 1) invalid("snapshot header: " + key + " is not a number");
    return v;
}

// The device arrays of the snapshot's fields, in kSnapshotFields order.
void snapshot_arrays(vgx_sim* sim, float* out[6], size_t count[6]) {
    out[0] = sim->rho[sim->rho_cur];
    for (int k = 0; k < 3; ++k) out[1 + k] = sim->S[k];
    out[4] = sim->proj.current(0);
    out[5] = sim->proj.current(1);
    for (int k = 0; k < 6; ++k) count[k] = (k == 0 || k >= 4) ? sim->cells : sim->faces;
}

}  // namespace

// --------------------------------------------- no-allocation advection
// The forward trace of vgx_advect_face_noalloc's MacCormack pass lives in this
// file-scope device array (one copy per GPU, held until the process ends):
// the faces of one axis of a 128^3 grid, 8,454,144 bytes. Host code never
// passes the array's name to a kernel; it uses the address
// hipGetSymbolAddress gives for the current device, looked up once per device.
constexpr size_t kNoAllocFaces = (size_t)129 * 128 * 128;
static __device__ __align__(256) float scratch[kNoAllocFaces];

namespace {
std::mutex g_scratch_mutex;
std::map<int, float*> g_scratch_by_device;

float* scratch_of_current_device() {
    int device = 0;
    check(hipGetDevice(&device), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    const auto found = g_scratch_by_device.find(device);
    if (found != g_scratch_by_device.end()) return found->second;
    void* address = nullptr;
    check(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_by_device[device] = static_cast<float*>(address);
    return static_cast<float*>(address);
}
}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* vgx_last_error(void) { return vgx_host::error_slot().c_str(); }

size_t vgx_struct_size(int which) {
    switch (which) {
    case 0: return sizeof(vgx_params);
    case 1: return sizeof(vgx_projection_options);
    case 2: return sizeof(vgx_projection_result);
    case 3: return sizeof(vgx_diagnostics);
    case 4: return sizeof(vgx_batch_tuning);
    case 5: return sizeof(vgx_field_error);
    default: return 0;
    }
}

int vgx_device_name(char* buffer, size_t size) {
    return guarded([&] {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        std::snprintf(buffer, size, "%s (%s)", prop.name, prop.gcnArchName);
    });
}

void vgx_params_default(vgx_params* params) { params_default(params); }

// MacGridParameters3D(MacAlgorithm), MacGridFluidSolver3D.cpp:11-26.
int vgx_params_preset(int algorithm, vgx_params* p) {
    return guarded([&] {
        if (!p) invalid("null parameters");
        params_default(p);
        switch (algorithm) {
        case VGX_SIMPLE_SEMI_LAGRANGIAN: p->reflection = 0; p->maccormack_smoke = 0; break;
        case VGX_SIMPLE_MACCORMACK: p->reflection = 0; p->maccormack_smoke = 1; break;
        case VGX_REFLECTION_SEMI_LAGRANGIAN: p->reflection = 1; p->maccormack_smoke = 0; break;
        case VGX_REFLECTION_MACCORMACK: p->reflection = 1; p->maccormack_smoke = 1; break;
        default: invalid("Unknown MAC algorithm");
        }
        p->maccormack_vel = p->maccormack_smoke;
        p->emitter_radius = .08f;
        p->source_strength = 4.f;
        p->buoyancy = 3.f;
        p->smoke_decay = .08f;
        p->project_iterations = 600;
        p->sor_omega = 1.9f;
    });
}

void vgx_projection_options_default(vgx_projection_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->dt = 1.f / 60.f;
    o->relaxation = 1.95f;
    o->max_iterations = 200;
    o->residual_check_interval = 8;
    o->absolute_tolerance = 1e-6;
    o->relative_tolerance = 1e-5;
    o->fixed_iterations = 0;
    o->diagonal_floor = 6.f;
    o->l2_tolerance = 1e-5;
}

int vgx_sim_create(int n, float box, const vgx_params* params, vgx_sim** out) {
    if (!out) return fail(VGX_EINVAL, "null output handle");
    *out = nullptr;
    vgx_sim* sim = nullptr;
    const int rc = guarded([&] {
        validate_grid(n, box);
        sim = new vgx_sim();
        sim->n = n;
        sim->box = box;
        if (params) {
            sim->params = *params;
        } else {
            params_default(&sim->params);
            sim->params.project_iterations = 2 * n;  // MacGridFluidSolver3D.cpp:43
        }
        sim->allocate();
        sim->reset();
    });
    if (rc != VGX_OK) {
        if (sim) {
            sim->release();
            delete sim;
        }
        return rc;
    }
    *out = sim;
    return VGX_OK;
}

int vgx_sim_create_algorithm(int algorithm, vgx_sim** out) {
    vgx_params p;
    const int rc = vgx_params_preset(algorithm, &p);
    if (rc != VGX_OK) return rc;
    return vgx_sim_create(32, 4.5f, &p, out);
}

void vgx_sim_destroy(vgx_sim* sim) {
    if (!sim || sim->batch_owned) return;
    sim->release();
    delete sim;
}

int vgx_sim_get_params(const vgx_sim* sim, vgx_params* params) {
    if (!sim || !params) return fail(VGX_EINVAL, "null argument");
    *params = sim->params;
    return VGX_OK;
}

int vgx_sim_set_params(vgx_sim* sim, const vgx_params* params) {
    if (!sim || !params) return fail(VGX_EINVAL, "null argument");
    if (sim->batch_owned) return fail(VGX_EINVAL, "the parameters of a batch's scene are fixed at creation");
    sim->params = *params;
    return VGX_OK;
}

int vgx_sim_resolution(const vgx_sim* sim) { return sim ? sim->n : 0; }
float vgx_sim_box_size(const vgx_sim* sim) { return sim ? sim->box : 0.f; }

int vgx_sim_advance(vgx_sim* sim, float dt) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "a scene of a batch advances with its batch");
    return guarded([&] { sim->advance(dt); });
}

int vgx_sim_reset(vgx_sim* sim) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "a scene of a batch resets with its batch");
    return guarded([&] { sim->reset(); });
}

int vgx_sim_set_box(vgx_sim* sim, float box) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "the box of a batch's scene is fixed at creation");
    return guarded([&] {
        if (!std::isfinite(box) || box <= 0.f || box / (float)sim->n <= 0.f) invalid("Invalid GPU box size");
        if (box != sim->box) {
            sim->box = box;
            sim->proj.box = box;
            sim->reset();
        }
    });
}

int vgx_sim_upload(vgx_sim* sim, const float* rho, const float* u, const float* v, const float* w) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    return guarded([&] {
        const float* src[4] = {rho, u, v, w};
        float* dst[4] = {sim->rho[sim->rho_cur], sim->S[0], sim->S[1], sim->S[2]};
        for (int k = 0; k < 4; ++k) {
            if (!src[k]) continue;
            const size_t count = k == 0 ? sim->cells : sim->faces;
            check(hipMemcpyAsync(dst[k], src[k], count * sizeof(float), hipMemcpyHostToDevice, sim->stream),
                  "hipMemcpyAsync(upload)");
            sim->uploaded += count * sizeof(float);
        }
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx_sim_set_unet(vgx_sim* sim, vgx_unet* unet) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "a batch's scenes solve without the learned warm start");
    sim->proj.net = vgx::unet_of(unet);
    return VGX_OK;
}

int vgx_sim_record(vgx_sim* sim, vgx_dataset* data) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "a batch's scenes do not record training pairs");
    vgx::Dataset* d = vgx::dataset_of(data);
    if (d && (d->n > sim->n || sim->n % d->n != 0))
        return fail(VGX_EINVAL, "dataset grid must equal or divide the simulation's");
    sim->proj.recorder = d;
    return VGX_OK;
}

int vgx_sim_add_velocity_noise(vgx_sim* sim, float sigma, uint64_t seed) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (!std::isfinite(sigma) || sigma < 0.f) return fail(VGX_EINVAL, "velocity noise sigma must be finite and nonnegative");
    return guarded([&] {
        for (uint32_t a = 0; a < 3; ++a)
            check(vgx::launch_add_noise(sim->stream, sim->S[a], sim->faces, sigma, seed, a), "noise");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx_sim_download(vgx_sim* sim, float* rho, float* u, float* v, float* w) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    return guarded([&] {
        float* dst[4] = {rho, u, v, w};
        const float* src[4] = {sim->rho[sim->rho_cur], sim->S[0], sim->S[1], sim->S[2]};
        for (int k = 0; k < 4; ++k) {
            if (!dst[k]) continue;
            const size_t count = k == 0 ? sim->cells : sim->faces;
            check(hipMemcpyAsync(dst[k], src[k], count * sizeof(float), hipMemcpyDeviceToHost, sim->stream),
                  "hipMemcpyAsync(download)");
            sim->downloaded += count * sizeof(float);
        }
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx_sim_download_pressure(vgx_sim* sim, float* p) {
    if (!sim || !p) return fail(VGX_EINVAL, "null argument");
    return guarded([&] {
        check(hipMemcpyAsync(p, sim->proj.current(0), sim->cells * sizeof(float), hipMemcpyDeviceToHost,
                             sim->stream),
              "hipMemcpyAsync(pressure)");
        sim->downloaded += sim->cells * sizeof(float);
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx_sim_upload_pressure(vgx_sim* sim, const float* p) {
    if (!sim || !p) return fail(VGX_EINVAL, "null argument");
    return guarded([&] {
        for (int slot = 0; slot < 2; ++slot)
            check(hipMemcpyAsync(sim->proj.current(slot), p, sim->cells * sizeof(float), hipMemcpyHostToDevice,
                                 sim->stream),
                  "hipMemcpyAsync(pressure)");
        sim->uploaded += sim->cells * sizeof(float);
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

int vgx_sim_transfer_bytes(const vgx_sim* sim, uint64_t* uploaded, uint64_t* downloaded) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (uploaded) *uploaded = sim->uploaded;
    if (downloaded) *downloaded = sim->downloaded;
    return VGX_OK;
}

int vgx_sim_last_projection(const vgx_sim* sim, int index, vgx_projection_result* result) {
    if (!sim || !result || index < 0 || index > 1) return fail(VGX_EINVAL, "invalid argument");
    *result = sim->last[index];
    return VGX_OK;
}

int vgx_sim_diagnostics(vgx_sim* sim, vgx_diagnostics* out) {
    if (!sim || !out) return fail(VGX_EINVAL, "null argument");
    return guarded([&] {
        int blocks = 0;
        {
            Timed t(sim->timer, sim->stream, vgx::K_DIAGNOSTICS);
            check(vgx::launch_diagnostics(sim->stream, sim->n, sim->h(), sim->S[0], sim->S[1], sim->S[2],
                                          sim->proj.partials, &blocks),
                  "diagnostics");
        }
        {
            Timed t(sim->timer, sim->stream, vgx::K_REDUCE);
            check(vgx::launch_finalize(sim->stream, vgx::Finalize::Diag, sim->proj.partials, blocks, sim->n, nullptr,
                                       sim->ds_dev),
                  "finalize");
        }
        check(hipMemcpyAsync(sim->ds_host, sim->ds_dev, sizeof(vgx::DiagState), hipMemcpyDeviceToHost, sim->stream),
              "hipMemcpyAsync(diag)");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
        sim->timer.flush();
        if (sim->ds_host->nonfinite > 0) invalid("Non-finite MAC velocity");
        const float dx = sim->h();
        out->kinetic_energy = (float)(0.5 * sim->ds_host->ke_sum * dx * dx * dx);
        out->max_velocity = (float)sim->ds_host->max_vel;
        out->max_divergence = (float)sim->ds_host->max_div;
    });
}

int vgx_sim_device_fields(vgx_sim* sim, const float** rho, const float** u, const float** v, const float** w,
                          const float** p) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (rho) *rho = sim->rho[sim->rho_cur];
    if (u) *u = sim->S[0];
    if (v) *v = sim->S[1];
    if (w) *w = sim->S[2];
    if (p) *p = sim->proj.current(0);
    return VGX_OK;
}

int vgx_sim_save_snapshot(vgx_sim* sim, const char* path) {
    if (!sim || !path || !*path) return fail(VGX_EINVAL, "null argument");
    return guarded([&] {
        float* dev[6];
        size_t count[6];
        snapshot_arrays(sim, dev, count);
        uint32_t rng = 0;
        check(hipMemcpyAsync(&rng, sim->rng, sizeof rng, hipMemcpyDeviceToHost, sim->stream), "hipMemcpyAsync(rng)");
        std::vector<float> host(sim->faces);
        const std::string base(path);
        for (int k = 0; k < 6; ++k) {
            check(hipMemcpyAsync(host.data(), dev[k], count[k] * sizeof(float), hipMemcpyDeviceToHost, sim->stream),
                  "hipMemcpyAsync(snapshot)");
            check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
            sim->downloaded += count[k] * sizeof(float);
            std::ofstream raw(base + kSnapshotFields[k].suffix, std::ios::binary | std::ios::trunc);
            raw.write(reinterpret_cast<const char*>(host.data()), (std::streamsize)(count[k] * sizeof(float)));
            if (!raw) invalid("cannot write " + base + kSnapshotFields[k].suffix);
        }
        std::ostringstream h;
        h.precision(9);
        h << "{\n  \"format\": \"vapor-mac3d-snapshot\",\n  \"version\": 1,\n"
          << "  \"resolution\": " << sim->n << ",\n  \"box_size\": " << sim->box << ",\n"
          << "  \"dtype\": \"float32\",\n  \"byte_order\": \"little\",\n"
          << "  \"random_state\": " << rng << ",\n  \"fields\": [\n";
        for (int k = 0; k < 6; ++k)
            h << "    {\"name\": \"" << kSnapshotFields[k].name << "\", \"file\": \""
              << base_name(base) + kSnapshotFields[k].suffix << "\", \"count\": " << count[k]
              << ", \"layout\": \"" << kSnapshotFields[k].layout << "\"}" << (k < 5 ? ",\n" : "\n");
        h << "  ]\n}\n";
        std::ofstream header(base + ".json", std::ios::trunc);
        header << h.str();
        if (!header) invalid("cannot write " + base + ".json");
    });
}

int vgx_sim_load_snapshot(vgx_sim* sim, const char* path) {
    if (!sim || !path || !*path) return fail(VGX_EINVAL, "null argument");
    if (sim->batch_owned) return fail(VGX_EINVAL, "a scene of a batch loads with its batch");
    return guarded([&] {
        const std::string base(path);
        std::ifstream header(base + ".json");
        if (!header) invalid("cannot read " + base + ".json");
        std::stringstream text;
        text << header.rdbuf();
        const std::string h = text.str();
        if (h.find("\"vapor-mac3d-snapshot\"") == std::string::npos) invalid("not a MAC 3D snapshot header");
        if ((int)header_number(h, "version") != 1) invalid("unsupported snapshot version");
        // the same checks as GpuMacSimulation3D::upload (GpuMacSimulation3D.cpp:110-113)
        if ((int)header_number(h, "resolution") != sim->n) invalid("GPU snapshot resolution mismatch");
        if ((float)header_number(h, "box_size") != sim->box) invalid("GPU snapshot box mismatch");
        const double rng = header_number(h, "random_state");
        if (!(rng >= 0.0 && rng <= 4294967295.0)) invalid("snapshot random state out of range");
        float* dev[6];
        size_t count[6];
        snapshot_arrays(sim, dev, count);
        std::vector<std::vector<float>> host(6);
        for (int k = 0; k < 6; ++k) {  // read and check everything before changing the state
            std::ifstream raw(base + kSnapshotFields[k].suffix, std::ios::binary | std::ios::ate);
            if (!raw) invalid("cannot read " + base + kSnapshotFields[k].suffix);
            if ((size_t)raw.tellg() != count[k] * sizeof(float))
                invalid(std::string("snapshot field ") + kSnapshotFields[k].name + " has the wrong size");
            raw.seekg(0);
            host[k].resize(count[k]);
            raw.read(reinterpret_cast<char*>(host[k].data()), (std::streamsize)(count[k] * sizeof(float)));
            for (float v : host[k])
                if (!std::isfinite(v)) invalid(std::string("snapshot field ") + kSnapshotFields[k].name + " is not finite");
        }
        for (int k = 0; k < 6; ++k) {
            check(hipMemcpyAsync(dev[k], host[k].data(), count[k] * sizeof(float), hipMemcpyHostToDevice, sim->stream),
                  "hipMemcpyAsync(snapshot)");
            sim->uploaded += count[k] * sizeof(float);
        }
        check(vgx::launch_set_u32(sim->stream, sim->rng, (uint32_t)rng), "rng");
        check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
        std::memset(sim->last, 0, sizeof(sim->last));
    });
}

int vgx_sim_set_advection_hook(vgx_sim* sim, vgx_advection_hook hook, void* user) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "the stages of a batch's scene are fixed");
    return guarded([&] {
        if (hook) sim->ensure_stage_buffers();
        sim->advect_hook = hook;
        sim->advect_user = hook ? user : nullptr;
    });
}

int vgx_sim_set_projection_hook(vgx_sim* sim, vgx_projection_hook hook, void* user) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    if (sim->batch_owned) return fail(VGX_EINVAL, "the stages of a batch's scene are fixed");
    return guarded([&] {
        if (hook) sim->ensure_stage_buffers();
        sim->project_hook = hook;
        sim->project_user = hook ? user : nullptr;
        std::memset(sim->last, 0, sizeof(sim->last));  // setPressureSolver clears lastProjection (.cpp:309)
    });
}

int vgx_sim_synchronize(vgx_sim* sim) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    return guarded([&] { check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize"); });
}

int vgx_sim_set_timing(vgx_sim* sim, int enabled) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    sim->timer.enabled = enabled != 0;
    return VGX_OK;
}

int vgx_sim_reset_timing(vgx_sim* sim) {
    if (!sim) return fail(VGX_EINVAL, "null simulation");
    sim->timer.reset();
    sim->proj.sweeps = 0;
    return VGX_OK;
}

int vgx_kernel_kinds(void) { return vgx::K_COUNT; }

const char* vgx_kernel_name(int kind) { return kind >= 0 && kind < vgx::K_COUNT ? vgx::kKindNames[kind] : ""; }

int vgx_sim_kernel_timing(vgx_sim* sim, int kind, double* total_ms, uint64_t* launches) {
    if (!sim || kind < 0 || kind >= vgx::K_COUNT) return fail(VGX_EINVAL, "invalid argument");
    if (total_ms) *total_ms = sim->timer.total_ms(kind);
    if (launches) *launches = sim->timer.launches[kind];
    return VGX_OK;
}

int vgx_sim_sweep_count(const vgx_sim* sim, uint64_t* sweeps) {
    if (!sim || !sweeps) return fail(VGX_EINVAL, "invalid argument");
    *sweeps = sim->proj.sweeps;
    return VGX_OK;
}

// ------------------------------------------------------------------ backend

int vgx_backend_create(int n, float box, vgx_backend** out) {
    if (!out) return fail(VGX_EINVAL, "null output handle");
    *out = nullptr;
    vgx_backend* be = nullptr;
    const int rc = guarded([&] {
        validate_grid(n, box);
        be = new vgx_backend();
        be->n = n;
        be->box = box;
        be->allocate();
        check(hipStreamSynchronize(be->stream), "hipStreamSynchronize");
    });
    if (rc != VGX_OK) {
        if (be) {
            be->release();
            delete be;
        }
        return rc;
    }
    *out = be;
    return VGX_OK;
}

void vgx_backend_destroy(vgx_backend* be) {
    if (!be) return;
    be->release();
    delete be;
}

int vgx_backend_advect_scalar(vgx_backend* be, int method, float dt, float cfl_limit, const float* u, const float* v,
                              const float* w, float* field) {
    if (!be) return fail(VGX_EINVAL, "null backend");
    return guarded([&] {
        validate_advection(dt, cfl_limit);
        if (!u || !v || !w) invalid("Invalid advection time step, CFL limit, or tracer field sizes");
        if (!field) invalid("Scalar field must match the grid cell count");
        if (overlaps(field, be->cells, u, be->faces) || overlaps(field, be->cells, v, be->faces) ||
            overlaps(field, be->cells, w, be->faces))
            invalid("Advection output must not alias tracer velocities");
        const float* flows[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) be->upload(be->flow[k], flows[k], be->faces);
        be->upload(be->src, field, be->cells);
        const bool mc = method == VGX_MACCORMACK;
        vgx::AdvArgs a{};
        a.u = be->flow[0];
        a.v = be->flow[1];
        a.w = be->flow[2];
        a.n = be->n;
        a.first_field = 3;
        a.h = be->box / (float)be->n;
        a.box = be->box;
        a.dt = dt;
        a.cfl = cfl_limit;
        a.out_scale = 1.f;
        a.f[3] = {be->src, nullptr, be->fwd, mc ? be->fwd : be->out};
        check(vgx::launch_advect_sl(be->stream, a, 1, false), "advect_sl");
        if (mc) {
            a.f[3].out = be->out;
            check(vgx::launch_advect_mc(be->stream, a, 1, false), "advect_mc");
        }
        be->download(field, be->out, be->cells);
        check(hipStreamSynchronize(be->stream), "hipStreamSynchronize");
    });
}

int vgx_backend_advect_face(vgx_backend* be, int axis, int method, float dt, float cfl_limit, const float* u,
                            const float* v, const float* w, const float* source, float* out) {
    if (!be) return fail(VGX_EINVAL, "null backend");
    return guarded([&] {
        validate_advection(dt, cfl_limit);
        if (!u || !v || !w) invalid("Invalid advection time step, CFL limit, or tracer field sizes");
        if (axis < 0 || axis > 2) invalid("Invalid staggered field axis");
        if (!source || !out || overlaps(source, be->faces, out, be->faces))
            invalid("Velocity fields must match their grid layout and not overlap");
        if (overlaps(out, be->faces, u, be->faces) || overlaps(out, be->faces, v, be->faces) ||
            overlaps(out, be->faces, w, be->faces))
            invalid("Advection output must not alias tracer velocities");
        const float* flows[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) be->upload(be->flow[k], flows[k], be->faces);
        be->upload(be->src, source, be->faces);
        const bool mc = method == VGX_MACCORMACK;
        vgx::AdvArgs a{};
        a.u = be->flow[0];
        a.v = be->flow[1];
        a.w = be->flow[2];
        a.n = be->n;
        a.first_field = axis;
        a.h = be->box / (float)be->n;
        a.box = be->box;
        a.dt = dt;
        a.cfl = cfl_limit;
        a.out_scale = 1.f;
        a.f[axis] = {be->src, nullptr, be->fwd, mc ? be->fwd : be->out};
        check(vgx::launch_advect_sl(be->stream, a, 1, false), "advect_sl");
        if (mc) {
            a.f[axis].out = be->out;
            check(vgx::launch_advect_mc(be->stream, a, 1, false), "advect_mc");
        }
        be->download(out, be->out, be->faces);
        check(hipStreamSynchronize(be->stream), "hipStreamSynchronize");
    });
}

int vgx_advect_face_noalloc(int n, float box, int axis, int method, float dt, float cfl_limit, const float* u,
                            const float* v, const float* w, const float* source, float* out, void* stream) {
    return guarded([&] {
        validate_grid(n, box);
        validate_advection(dt, cfl_limit);
        if (!u || !v || !w) invalid("Invalid advection time step, CFL limit, or tracer field sizes");
        if (axis < 0 || axis > 2) invalid("Invalid staggered field axis");
        if (method != VGX_SEMI_LAGRANGIAN && method != VGX_MACCORMACK) invalid("Unknown advection method");
        const size_t faces = (size_t)(n + 1) * n * n;
        if (faces > kNoAllocFaces)
            throw Failure{VGX_ERANGE, "grid exceeds the no-allocation scratch (n <= 128)"};
        if (!source || !out || overlaps(source, faces, out, faces))
            invalid("Velocity fields must match their grid layout and not overlap");
        if (overlaps(out, faces, u, faces) || overlaps(out, faces, v, faces) || overlaps(out, faces, w, faces))
            invalid("Advection output must not alias tracer velocities");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        float* trace = scratch_of_current_device();
        const bool mc = method == VGX_MACCORMACK;
        vgx::AdvArgs a{};
        a.u = u;
        a.v = v;
        a.w = w;
        a.n = n;
        a.first_field = axis;
        a.h = box / (float)n;
        a.box = box;
        a.dt = dt;
        a.cfl = cfl_limit;
        a.out_scale = 1.f;
        // every face of the trace is written by the first pass before the
        // MacCormack pass reads it, so the scratch needs no clearing
        a.f[axis] = {source, nullptr, trace, mc ? trace : out};
        check(vgx::launch_advect_sl(s, a, 1, false), "advect_sl");
        if (mc) {
            a.f[axis].out = out;
            check(vgx::launch_advect_mc(s, a, 1, false), "advect_mc");
        }
    });
}

int vgx_backend_project(vgx_backend* be, const vgx_projection_options* o, float* u, float* v, float* w,
                        float* pressure_inout, vgx_projection_result* result) {
    if (!be) return fail(VGX_EINVAL, "null backend");
    return guarded([&] {
        if (!o || !u || !v || !w) invalid("null argument");
        if (!std::isfinite(o->diagonal_floor) || o->diagonal_floor < 0.f)
            invalid("SOR diagonal floor must be finite and nonnegative");
        if (o->warm_start && !pressure_inout) invalid("Warm start requires a pressure field");
        SolvePolicy pol;
        pol.dt = o->dt;
        pol.omega = o->relaxation;
        pol.floor = o->diagonal_floor;
        pol.max_iterations = o->max_iterations;
        pol.check_interval = o->residual_check_interval;
        pol.abs_tol = o->absolute_tolerance;
        pol.rel_tol = o->relative_tolerance;
        pol.fixed = o->fixed_iterations != 0;
        pol.use_l2 = o->use_l2 != 0;
        pol.l2_tol = o->l2_tolerance;
        pol.warm = o->warm_start != 0;
        pol.refinement = o->refinement_steps;
        float* faces[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) be->upload(be->in3[k], faces[k], be->faces);
        if (pol.warm) be->upload(be->proj.current(0), pressure_inout, be->cells);
        const float* in[3] = {be->in3[0], be->in3[1], be->in3[2]};
        float* outf[3] = {be->flow[0], be->flow[1], be->flow[2]};
        const auto pend = be->proj.project(be->timer, 0, pol, in, outf);
        for (int k = 0; k < 3; ++k) be->download(faces[k], outf[k], be->faces);
        if (pressure_inout) be->download(pressure_inout, be->proj.current(0), be->cells);
        be->proj.read_record(0);
        be->timer.flush();
        const auto r = be->proj.finish(pend);
        if (result) *result = r;
    });
}

}  // extern "C"
