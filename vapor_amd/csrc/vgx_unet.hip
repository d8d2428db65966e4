// vgx_unet.hip - the learned pressure warm start: network orchestration and
// its C ABI. The kernels live in vgx_unet_kernels.hip.
#include "vgx_unet.hpp"
#include "vgx_host.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace vgx {
namespace {

using namespace vgx_host;

struct LayerDef {
    int cin, cout;
};
// Encoder L0-L7 (two per level), decoder L8-L13; decoder inputs are the
// concatenation [upsampled deeper output | skip].
constexpr LayerDef kLayers[kUnetLayers] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 48},  {48, 48}, {48, 64},
                                           {64, 64}, {112, 48}, {48, 48}, {80, 32}, {32, 32}, {48, 16}, {16, 16}};
// Real input channels of each layer (fan-in for the initial scale).
constexpr int kFanInChannels[kUnetLayers] = {2, 16, 16, 32, 32, 48, 48, 64, 112, 48, 80, 32, 48, 16};

constexpr int round32(int k) { return (k + 31) / 32 * 32; }

size_t layer_params(int l) { return (size_t)kLayers[l].cout * 27 * kLayers[l].cin; }
size_t layer_offset(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += layer_params(i);
    return o;
}
size_t head_offset() { return layer_offset(kUnetLayers); }
int kp_of(int l) { return round32(27 * kLayers[l].cin); }
int kpt_of(int l) { return round32(27 * kLayers[l].cout); }
size_t packed_offset(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)kLayers[i].cout * kp_of(i) + (size_t)kLayers[i].cin * kpt_of(i);
    return o;
}
size_t packed_total() { return packed_offset(kUnetLayers); }

template <class T> void free_ptr(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

// Activations and gradients for one grid size n (levels n, n/2, n/4, n/8).
struct UnetNet::Buffers {
    int n = 0;
    size_t V[4] = {};
    // forward
    bf16 *x0, *a0, *c0, *p0, *a1, *c1, *p1, *a2, *c2, *p2, *a3, *b3, *a4, *d2, *a5, *d1, *a6, *d0;
    float* out;
    // backward
    bf16 *dz, *g_d0, *g_a6, *g_c0, *g_d1, *g_a5, *g_c1, *g_d2, *g_a4, *g_c2, *g_b3, *g_a3, *g_p2, *g_a2, *g_p1,
        *g_a1, *g_p0, *g_a0;
    float *dout, *part;
    double* red;
    size_t part_floats = 0;
    std::vector<void*> all;

    template <class T> T* alloc(size_t count) {
        T* p = device_alloc<T>(count, "hipMalloc(unet)");
        all.push_back(p);
        return p;
    }
    explicit Buffers(int n_) : n(n_) {
        for (int l = 0; l < 4; ++l) V[l] = (size_t)(n >> l) * (n >> l) * (n >> l);
        try {
            x0 = alloc<bf16>(V[0] * kUnetIn);
            a0 = alloc<bf16>(V[0] * 16);
            c0 = alloc<bf16>(V[0] * 48);
            p0 = alloc<bf16>(V[1] * 16);
            a1 = alloc<bf16>(V[1] * 32);
            c1 = alloc<bf16>(V[1] * 80);
            p1 = alloc<bf16>(V[2] * 32);
            a2 = alloc<bf16>(V[2] * 48);
            c2 = alloc<bf16>(V[2] * 112);
            p2 = alloc<bf16>(V[3] * 48);
            a3 = alloc<bf16>(V[3] * 64);
            b3 = alloc<bf16>(V[3] * 64);
            a4 = alloc<bf16>(V[2] * 48);
            d2 = alloc<bf16>(V[2] * 48);
            a5 = alloc<bf16>(V[1] * 32);
            d1 = alloc<bf16>(V[1] * 32);
            a6 = alloc<bf16>(V[0] * 16);
            d0 = alloc<bf16>(V[0] * 16);
            out = alloc<float>(V[0]);
            dz = alloc<bf16>(V[0] * 16);
            g_d0 = alloc<bf16>(V[0] * 16);
            g_a6 = alloc<bf16>(V[0] * 16);
            g_c0 = alloc<bf16>(V[0] * 48);
            g_d1 = alloc<bf16>(V[1] * 32);
            g_a5 = alloc<bf16>(V[1] * 32);
            g_c1 = alloc<bf16>(V[1] * 80);
            g_d2 = alloc<bf16>(V[2] * 48);
            g_a4 = alloc<bf16>(V[2] * 48);
            g_c2 = alloc<bf16>(V[2] * 112);
            g_b3 = alloc<bf16>(V[3] * 64);
            g_a3 = alloc<bf16>(V[3] * 64);
            g_p2 = alloc<bf16>(V[3] * 48);
            g_a2 = alloc<bf16>(V[2] * 48);
            g_p1 = alloc<bf16>(V[2] * 32);
            g_a1 = alloc<bf16>(V[1] * 32);
            g_p0 = alloc<bf16>(V[1] * 16);
            g_a0 = alloc<bf16>(V[0] * 16);
            dout = alloc<float>(V[0]);
            // weight-gradient partials: the largest chunks x cout x 27 x cin over the layers
            const int level_of[kUnetLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
            for (int l = 0; l < kUnetLayers; ++l)
                part_floats = std::max(part_floats, (size_t)unet_wgrad_chunks((int)V[level_of[l]]) * layer_params(l));
            part = alloc<float>(part_floats);
            red = alloc<double>((size_t)unet_reduce_blocks(V[0]) * kUnetHead);
        } catch (...) {
            release();
            throw;
        }
    }
    void release() {
        for (void* p : all) (void)hipFree(p);
        all.clear();
    }
    ~Buffers() { release(); }
};

UnetNet::UnetNet(uint64_t seed) {
    const size_t P = param_count();
    try {
        params_ = device_alloc<float>(P, "hipMalloc(unet params)");
        grads_ = device_alloc<float>(P, "hipMalloc(unet grads)");
        m_ = device_alloc<float>(P, "hipMalloc(unet adam)");
        v_ = device_alloc<float>(P, "hipMalloc(unet adam)");
        packed_ = device_alloc<bf16>(packed_total(), "hipMalloc(unet weights)");
        for (int l = 0; l < kUnetLayers; ++l)
            check(unet_init_weights(nullptr, params_ + layer_offset(l), layer_params(l),
                                    std::sqrt(2.f / (27.f * kFanInChannels[l])), seed, (uint32_t)l),
                  "unet init");
        check(unet_init_weights(nullptr, params_ + head_offset(), kUnetHead, std::sqrt(1.f / kUnetHead), seed,
                                (uint32_t)kUnetLayers),
              "unet init");
        check(hipMemset(grads_, 0, P * sizeof(float)), "hipMemset");
        check(hipMemset(m_, 0, P * sizeof(float)), "hipMemset");
        check(hipMemset(v_, 0, P * sizeof(float)), "hipMemset");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (...) {
        free_ptr(params_), free_ptr(grads_), free_ptr(m_), free_ptr(v_), free_ptr(packed_);
        throw;
    }
}

UnetNet::~UnetNet() {
    delete buf_;
    free_ptr(params_), free_ptr(grads_), free_ptr(m_), free_ptr(v_), free_ptr(packed_);
}

size_t UnetNet::param_count() { return head_offset() + kUnetHead; }

void UnetNet::get_params(float* host) {
    check(hipMemcpy(host, params_, param_count() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}
void UnetNet::set_params(const float* host) {
    check(hipMemcpy(params_, host, param_count() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    packed_valid_ = false;
}
void UnetNet::get_grads(float* host) {
    check(hipMemcpy(host, grads_, param_count() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}
void UnetNet::set_grads(const float* host) {
    check(hipMemcpy(grads_, host, param_count() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
}

void UnetNet::ensure(int n) {
    if (n < 8 || n % 8) invalid("U-Net grid size must be a positive multiple of 8");
    if (buf_ && buf_->n == n) return;
    delete buf_;
    buf_ = nullptr;
    buf_ = new Buffers(n);
}

void UnetNet::pack(hipStream_t s) {
    if (packed_valid_) return;
    for (int l = 0; l < kUnetLayers; ++l) {
        bf16* w = packed_ + packed_offset(l);
        bf16* wt = w + (size_t)kLayers[l].cout * kp_of(l);
        check(unet_pack_weights(s, params_ + layer_offset(l), kLayers[l].cout, kLayers[l].cin, kp_of(l), kpt_of(l), w,
                                wt),
              "unet pack");
    }
    packed_valid_ = true;
}

namespace {
const bf16* wf(const bf16* packed, int l) { return packed + packed_offset(l); }
const bf16* wb(const bf16* packed, int l) { return packed + packed_offset(l) + (size_t)kLayers[l].cout * kp_of(l); }
ConvArgs conv_args(int d, const bf16* in, int in_stride, int in_off, const bf16* w, int kp, int cin, bf16* out,
                   int out_stride, int out_off, int cout, bool relu) {
    return ConvArgs{d, d, d, in, in_stride, in_off, cin, w, kp, out, out_stride, out_off, cout, relu ? 1 : 0};
}
}  // namespace

void UnetNet::forward(hipStream_t s, int n, const float* field, float in_scale, float out_scale, float* out) {
    ensure(n);
    pack(s);
    Buffers& b = *buf_;
    const int d0n = n, d1n = n / 2, d2n = n / 4, d3n = n / 8;
    auto conv = [&](int l, int d, const bf16* in, int in_stride, int in_off, bf16* o, int out_stride, int out_off) {
        check(unet_conv(s, conv_args(d, in, in_stride, in_off, wf(packed_, l), kp_of(l), kLayers[l].cin, o,
                                     out_stride, out_off, kLayers[l].cout, true)),
              "unet conv");
    };
    check(unet_input(s, field, in_scale, n, b.x0), "unet input");
    conv(0, d0n, b.x0, kUnetIn, 0, b.a0, 16, 0);
    conv(1, d0n, b.a0, 16, 0, b.c0, 48, 32);  // s0 -> c0[:, 32:48]
    check(unet_pool(s, View{b.c0, 48, 32}, 16, d1n, d1n, d1n, b.p0), "unet pool");
    conv(2, d1n, b.p0, 16, 0, b.a1, 32, 0);
    conv(3, d1n, b.a1, 32, 0, b.c1, 80, 48);  // s1 -> c1[:, 48:80]
    check(unet_pool(s, View{b.c1, 80, 48}, 32, d2n, d2n, d2n, b.p1), "unet pool");
    conv(4, d2n, b.p1, 32, 0, b.a2, 48, 0);
    conv(5, d2n, b.a2, 48, 0, b.c2, 112, 64);  // s2 -> c2[:, 64:112]
    check(unet_pool(s, View{b.c2, 112, 64}, 48, d3n, d3n, d3n, b.p2), "unet pool");
    conv(6, d3n, b.p2, 48, 0, b.a3, 64, 0);
    conv(7, d3n, b.a3, 64, 0, b.b3, 64, 0);
    check(unet_upsample(s, b.b3, 64, d3n, d3n, d3n, View{b.c2, 112, 0}), "unet up");
    conv(8, d2n, b.c2, 112, 0, b.a4, 48, 0);
    conv(9, d2n, b.a4, 48, 0, b.d2, 48, 0);
    check(unet_upsample(s, b.d2, 48, d2n, d2n, d2n, View{b.c1, 80, 0}), "unet up");
    conv(10, d1n, b.c1, 80, 0, b.a5, 32, 0);
    conv(11, d1n, b.a5, 32, 0, b.d1, 32, 0);
    check(unet_upsample(s, b.d1, 32, d1n, d1n, d1n, View{b.c0, 48, 0}), "unet up");
    conv(12, d0n, b.c0, 48, 0, b.a6, 16, 0);
    conv(13, d0n, b.a6, 16, 0, b.d0, 16, 0);
    check(unet_head(s, b.d0, params_ + head_offset(), kUnetHead, b.V[0], out_scale, out), "unet head");
}

void UnetNet::predict(hipStream_t s, int n, const float* field, float in_scale, float out_scale, float* out) {
    forward(s, n, field, in_scale, out_scale, out);
}

double UnetNet::loss_grad(hipStream_t s, int n, const float* field, float in_scale, const float* target,
                          bool accumulate) {
    ensure(n);
    Buffers& b = *buf_;
    forward(s, n, field, in_scale, 1.f, b.out);
    const int nb = unet_reduce_blocks(b.V[0]);
    check(unet_mse(s, b.out, target, b.V[0], b.dout, b.red), "unet mse");
    std::vector<double> loss_part(nb);
    check(hipMemcpyAsync(loss_part.data(), b.red, nb * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
    double loss = 0.0;
    for (double v : loss_part) loss += v;
    loss /= (double)b.V[0];

    const int d0n = n, d1n = n / 2, d2n = n / 4, d3n = n / 8;
    const int acc = accumulate ? 1 : 0;
    // head
    check(unet_head_back(s, b.d0, params_ + head_offset(), b.dout, kUnetHead, b.V[0], b.g_d0, b.red), "unet head");
    {
        std::vector<double> part((size_t)nb * kUnetHead);
        check(hipMemcpyAsync(part.data(), b.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        float gh[kUnetHead];
        check(hipMemcpy(gh, grads_ + head_offset(), sizeof gh, hipMemcpyDeviceToHost), "hipMemcpy");
        for (int c = 0; c < kUnetHead; ++c) {
            double sum = 0.0;
            for (int k = 0; k < nb; ++k) sum += part[(size_t)k * kUnetHead + c];
            gh[c] = acc ? gh[c] + (float)sum : (float)sum;
        }
        check(hipMemcpy(grads_ + head_offset(), gh, sizeof gh, hipMemcpyHostToDevice), "hipMemcpy");
    }
    // layer l: dz = g * relu'(y); weight gradient against input x; optionally
    // the input gradient (backward-data conv with the mirrored weights).
    auto layer = [&](int l, int d, View g, View y, View x, bf16* g_in, int g_in_stride) {
        const int C = kLayers[l].cout, V = d * d * d;
        check(unet_relu_mask(s, g, y, C, V, b.dz), "unet relu");
        WgradArgs w{d, d, d, b.dz, C, 0, x.p, x.stride, x.off, kLayers[l].cin, C, b.part};
        check(unet_wgrad(s, w, grads_ + layer_offset(l), acc), "unet wgrad");
        if (g_in)
            check(unet_conv(s, conv_args(d, b.dz, C, 0, wb(packed_, l), kpt_of(l), C, g_in, g_in_stride, 0,
                                         kLayers[l].cin, false)),
                  "unet dgrad");
    };
    layer(13, d0n, View{b.g_d0, 16, 0}, View{b.d0, 16, 0}, View{b.a6, 16, 0}, b.g_a6, 16);
    layer(12, d0n, View{b.g_a6, 16, 0}, View{b.a6, 16, 0}, View{b.c0, 48, 0}, b.g_c0, 48);
    check(unet_upsample_back(s, View{b.g_c0, 48, 0}, 32, d1n, d1n, d1n, b.g_d1), "unet up");
    layer(11, d1n, View{b.g_d1, 32, 0}, View{b.d1, 32, 0}, View{b.a5, 32, 0}, b.g_a5, 32);
    layer(10, d1n, View{b.g_a5, 32, 0}, View{b.a5, 32, 0}, View{b.c1, 80, 0}, b.g_c1, 80);
    check(unet_upsample_back(s, View{b.g_c1, 80, 0}, 48, d2n, d2n, d2n, b.g_d2), "unet up");
    layer(9, d2n, View{b.g_d2, 48, 0}, View{b.d2, 48, 0}, View{b.a4, 48, 0}, b.g_a4, 48);
    layer(8, d2n, View{b.g_a4, 48, 0}, View{b.a4, 48, 0}, View{b.c2, 112, 0}, b.g_c2, 112);
    check(unet_upsample_back(s, View{b.g_c2, 112, 0}, 64, d3n, d3n, d3n, b.g_b3), "unet up");
    layer(7, d3n, View{b.g_b3, 64, 0}, View{b.b3, 64, 0}, View{b.a3, 64, 0}, b.g_a3, 64);
    layer(6, d3n, View{b.g_a3, 64, 0}, View{b.a3, 64, 0}, View{b.p2, 48, 0}, b.g_p2, 48);
    check(unet_pool_back_add(s, b.g_p2, 48, d3n, d3n, d3n, View{b.g_c2, 112, 64}), "unet pool");
    layer(5, d2n, View{b.g_c2, 112, 64}, View{b.c2, 112, 64}, View{b.a2, 48, 0}, b.g_a2, 48);
    layer(4, d2n, View{b.g_a2, 48, 0}, View{b.a2, 48, 0}, View{b.p1, 32, 0}, b.g_p1, 32);
    check(unet_pool_back_add(s, b.g_p1, 32, d2n, d2n, d2n, View{b.g_c1, 80, 48}), "unet pool");
    layer(3, d1n, View{b.g_c1, 80, 48}, View{b.c1, 80, 48}, View{b.a1, 32, 0}, b.g_a1, 32);
    layer(2, d1n, View{b.g_a1, 32, 0}, View{b.a1, 32, 0}, View{b.p0, 16, 0}, b.g_p0, 16);
    check(unet_pool_back_add(s, b.g_p0, 16, d1n, d1n, d1n, View{b.g_c0, 48, 32}), "unet pool");
    layer(1, d0n, View{b.g_c0, 48, 32}, View{b.c0, 48, 32}, View{b.a0, 16, 0}, b.g_a0, 16);
    layer(0, d0n, View{b.g_a0, 16, 0}, View{b.a0, 16, 0}, View{b.x0, kUnetIn, 0}, nullptr, 0);
    return loss;
}

void UnetNet::adam(hipStream_t s, float lr, float grad_scale) {
    ++adam_t_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t_), c2 = 1.f - std::pow(b2, (float)adam_t_);
    check(unet_adam(s, params_, grads_, m_, v_, param_count(), grad_scale, lr, b1, b2, 1e-8f, c1, c2), "unet adam");
    packed_valid_ = false;
}

}  // namespace vgx

using vgx_host::fail;
using vgx_host::guarded;

struct vgx_unet {
    vgx::UnetNet net;
    hipStream_t stream = nullptr;
    explicit vgx_unet(uint64_t seed) : net(seed) {}
};

struct vgx_dataset {
    vgx::Dataset d;
    ~vgx_dataset() {
        for (float* q : {d.inputs, d.targets, d.prev})
            if (q) (void)hipFree(q);
    }
};

namespace vgx {
UnetNet* unet_of(vgx_unet* u) { return u ? &u->net : nullptr; }
Dataset* dataset_of(vgx_dataset* d) { return d ? &d->d : nullptr; }
}  // namespace vgx

extern "C" {

int vgx_unet_create(uint64_t seed, vgx_unet** out) {
    if (!out) return fail(VGX_EINVAL, "null output");
    *out = nullptr;
    return guarded([&] {
        auto* u = new vgx_unet(seed);
        if (hipStreamCreate(&u->stream) != hipSuccess) {
            delete u;
            vgx_host::check(hipErrorOutOfMemory, "hipStreamCreate");
        }
        *out = u;
    });
}

void vgx_unet_destroy(vgx_unet* u) {
    if (!u) return;
    if (u->stream) {
        (void)hipStreamSynchronize(u->stream);
        (void)hipStreamDestroy(u->stream);
    }
    delete u;
}

size_t vgx_unet_param_count(void) { return vgx::UnetNet::param_count(); }

int vgx_unet_get_params(vgx_unet* u, float* out) {
    if (!u || !out) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] { u->net.get_params(out); });
}
int vgx_unet_set_params(vgx_unet* u, const float* in) {
    if (!u || !in) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] { u->net.set_params(in); });
}
int vgx_unet_set_grid(vgx_unet* u, int m) {
    if (!u) return fail(VGX_EINVAL, "null network");
    if (m < 0 || (m > 0 && (m < 8 || m % 8 != 0 || m > 1024)))
        return fail(VGX_EINVAL, "network grid must be 0 or a multiple of 8");
    u->net.set_grid(m);
    return VGX_OK;
}
int vgx_unet_get_grid(const vgx_unet* u) { return u ? u->net.grid() : 0; }

int vgx_unet_get_grads(vgx_unet* u, float* out) {
    if (!u || !out) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] { u->net.get_grads(out); });
}
int vgx_unet_set_grads(vgx_unet* u, const float* in) {
    if (!u || !in) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] { u->net.set_grads(in); });
}

namespace {
struct DeviceField {
    float* p = nullptr;
    DeviceField(const float* host, size_t count) {
        vgx_host::check(hipMalloc(&p, count * sizeof(float)), "hipMalloc");
        if (host) vgx_host::check(hipMemcpy(p, host, count * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    }
    ~DeviceField() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

int vgx_unet_predict(vgx_unet* u, int n, const float* field, float in_scale, float out_scale, float* out) {
    if (!u || !field || !out) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] {
        const size_t V = (size_t)n * n * n;
        DeviceField in(field, V), res(nullptr, V);
        u->net.predict(u->stream, n, in.p, in_scale, out_scale, res.p);
        vgx_host::check(hipMemcpyAsync(out, res.p, V * sizeof(float), hipMemcpyDeviceToHost, u->stream), "hipMemcpy");
        vgx_host::check(hipStreamSynchronize(u->stream), "hipStreamSynchronize");
    });
}

int vgx_unet_loss_grad(vgx_unet* u, int n, const float* field, float in_scale, const float* target, int accumulate,
                       double* loss) {
    if (!u || !field || !target) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] {
        const size_t V = (size_t)n * n * n;
        DeviceField in(field, V), tg(target, V);
        const double l = u->net.loss_grad(u->stream, n, in.p, in_scale, tg.p, accumulate != 0);
        vgx_host::check(hipStreamSynchronize(u->stream), "hipStreamSynchronize");
        if (loss) *loss = l;
    });
}

int vgx_dataset_create(int n, int capacity, vgx_dataset** out) {
    if (!out) return fail(VGX_EINVAL, "null output");
    *out = nullptr;
    if (n < 8 || n % 8 || n > 1024 || capacity <= 0) return fail(VGX_EINVAL, "dataset needs n % 8 == 0 and capacity > 0");
    return guarded([&] {
        auto* d = new vgx_dataset();
        try {
            d->d.n = n;
            d->d.capacity = capacity;
            const size_t cells = d->d.cells();
            d->d.inputs = vgx_host::device_alloc<float>(cells * capacity, "hipMalloc(dataset)");
            d->d.targets = vgx_host::device_alloc<float>(cells * capacity, "hipMalloc(dataset)");
            d->d.prev = vgx_host::device_alloc<float>(cells, "hipMalloc(dataset)");
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}

void vgx_dataset_destroy(vgx_dataset* d) { delete d; }

int vgx_dataset_size(const vgx_dataset* d) { return d ? d->d.size : 0; }

int vgx_dataset_get(const vgx_dataset* d, int index, float* input, float* target) {
    if (!d || index < 0 || index >= d->d.size) return fail(VGX_EINVAL, "dataset index out of range");
    return guarded([&] {
        const size_t cells = d->d.cells(), at = (size_t)index * cells;
        if (input)
            vgx_host::check(hipMemcpy(input, d->d.inputs + at, cells * sizeof(float), hipMemcpyDeviceToHost),
                            "hipMemcpy(dataset)");
        if (target)
            vgx_host::check(hipMemcpy(target, d->d.targets + at, cells * sizeof(float), hipMemcpyDeviceToHost),
                            "hipMemcpy(dataset)");
    });
}

int vgx_dataset_clear(vgx_dataset* d) {
    if (!d) return fail(VGX_EINVAL, "null dataset");
    d->d.size = 0;
    return VGX_OK;
}

int vgx_unet_train(vgx_unet* u, const vgx_dataset* data, int epochs, int batch, float lr, uint64_t seed,
                   double* mean_loss) {
    if (!u || !data || epochs < 0 || batch <= 0 || !(lr > 0.f)) return fail(VGX_EINVAL, "invalid argument");
    if (data->d.size == 0) return fail(VGX_EINVAL, "empty dataset");
    return guarded([&] {
        const vgx::Dataset& d = data->d;
        const size_t cells = d.cells();
        std::vector<int> order(d.size);
        uint64_t state = seed * 0x9E3779B97F4A7C15ull + 1;
        double epoch_loss = 0.0;
        for (int e = 0; e < epochs; ++e) {
            for (int i = 0; i < d.size; ++i) order[i] = i;
            for (int i = d.size - 1; i > 0; --i) {  // Fisher-Yates with a 64-bit LCG
                state = state * 6364136223846793005ull + 1442695040888963407ull;
                const int j = (int)((state >> 33) % (uint64_t)(i + 1));
                std::swap(order[i], order[j]);
            }
            epoch_loss = 0.0;
            for (int start = 0; start < d.size; start += batch) {
                const int count = std::min(batch, d.size - start);
                for (int k = 0; k < count; ++k) {
                    const size_t s = (size_t)order[start + k] * cells;
                    epoch_loss += u->net.loss_grad(u->stream, d.n, d.inputs + s, 1.f, d.targets + s, k > 0);
                }
                u->net.adam(u->stream, lr, 1.f / (float)count);
            }
            epoch_loss /= d.size;
        }
        vgx_host::check(hipStreamSynchronize(u->stream), "hipStreamSynchronize");
        if (mean_loss) *mean_loss = epoch_loss;
    });
}

int vgx_unet_adam_step(vgx_unet* u, float lr, float grad_scale) {
    if (!u || !(lr >= 0.f) || !(grad_scale > 0.f)) return fail(VGX_EINVAL, "invalid argument");
    return guarded([&] {
        u->net.adam(u->stream, lr, grad_scale);
        vgx_host::check(hipStreamSynchronize(u->stream), "hipStreamSynchronize");
    });
}

}  // extern "C"
