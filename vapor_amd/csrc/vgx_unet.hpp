// vgx_unet.hpp - types and launchers of the learned pressure warm start.
#pragma once

#include "vgx_device.hpp"
#include "vgx_internal.hpp"

#include <algorithm>

namespace vgx {

using bf16 = __bf16;

constexpr int kUnetIn = 16;    // input channels: scaled field, boundary mask, zero padding
constexpr int kUnetHead = 16;  // channels into the 1x1x1 output head
constexpr int kUnetLayers = 14;

// A channels-last tensor view: element (v, c) at p[v * stride + off + c].
struct View {
    bf16* p;
    int stride, off;
};

// 3^3 convolution (zero padding, no bias), optional ReLU on the output.
struct ConvArgs {
    int X, Y, Z;
    const bf16* in;
    int in_stride, in_off, cin;
    const bf16* w;  // [cout][kp], k = tap * cin + ci, zero beyond 27 cin
    int kp;
    bf16* out;
    int out_stride, out_off, cout;
    int relu;
};

// Weight gradient of a 3^3 convolution: sum over voxels of dz[v][co] x[v + off(tap)][ci].
struct WgradArgs {
    int X, Y, Z;
    const bf16* dz;
    int dz_stride, dz_off;
    const bf16* x;
    int x_stride, x_off;
    int cin, cout;
    float* part;  // [chunks][cout][27][cin]
};

hipError_t unet_conv(hipStream_t s, const ConvArgs& a);
int unet_wgrad_chunks(int V);
hipError_t unet_wgrad(hipStream_t s, const WgradArgs& a, float* grad, int accumulate);
hipError_t unet_relu_mask(hipStream_t s, View g, View y, int C, int V, bf16* out);
hipError_t unet_pool(hipStream_t s, View in, int C, int X, int Y, int Z, bf16* out);
hipError_t unet_pool_back_add(hipStream_t s, const bf16* gl, int C, int X, int Y, int Z, View g);
hipError_t unet_upsample(hipStream_t s, const bf16* in, int C, int X, int Y, int Z, View out);
hipError_t unet_upsample_back(hipStream_t s, View g, int C, int X, int Y, int Z, bf16* gl);
hipError_t unet_input(hipStream_t s, const float* field, float scale, int n, bf16* x0);
hipError_t unet_head(hipStream_t s, const bf16* d0, const float* wh, int C, size_t V, float scale, float* out);
int unet_reduce_blocks(size_t V);
hipError_t unet_mse(hipStream_t s, const float* out, const float* target, size_t V, float* dout, double* part);
hipError_t unet_head_back(hipStream_t s, const bf16* d0, const float* wh, const float* dout, int C, size_t V,
                          bf16* gd0, double* part);
hipError_t unet_adam(hipStream_t s, float* p, const float* g, float* m, float* v, size_t count, float gscale, float lr,
                     float beta1, float beta2, float eps, float c1, float c2);
hipError_t unet_pack_weights(hipStream_t s, const float* w32, int cout, int cin, int kp, int kpt, bf16* w, bf16* wt);
hipError_t unet_init_weights(hipStream_t s, float* w, size_t count, float stdev, uint64_t seed, uint32_t stream);
// coarse[I] = scale * mean of the f^3 block I of the (m f)^3 field fine.
hipError_t restrict_mean(hipStream_t s, int m, int f, float scale, const float* fine, float* coarse);
// Cell-centred trilinear prolongation of an m^3 field to (m f)^3.
hipError_t prolong_linear(hipStream_t s, int m, int f, const float* coarse, float* fine);
// The same with the scales read from a residual record on the device:
// restriction divided by the record's residual RMS, prolongation multiplied
// by that RMS times f^2 (the learned start without a host round trip).
hipError_t restrict_rms(hipStream_t s, int m, int f, const ProjState* ps, const float* fine, float* coarse);
hipError_t prolong_rms(hipStream_t s, int m, int f, const ProjState* ps, const float* coarse, float* fine);

}  // namespace vgx

namespace vgx {

// The 4-level U-Net (16/32/48/64 channels, two 3^3 convolutions with ReLU
// per level and per decoder stage, 2x2x2 average pooling, nearest
// upsampling, skip connections by concatenation, no biases, a 1x1x1 head).
// Without biases every layer is positively homogeneous, so the network
// commutes with positive scaling of its input. Host methods enqueue on the
// given stream; errors throw vgx_host::Failure.
class UnetNet {
public:
    explicit UnetNet(uint64_t seed);
    ~UnetNet();
    UnetNet(const UnetNet&) = delete;
    UnetNet& operator=(const UnetNet&) = delete;

    static size_t param_count();
    void get_params(float* host);
    void set_params(const float* host);
    void get_grads(float* host);
    void set_grads(const float* host);

    // out (n^3) = out_scale * f(in_scale * field) on the device; n % 8 == 0.
    void predict(hipStream_t s, int n, const float* field, float in_scale, float out_scale, float* out);
    // Mean squared error of f(in_scale * field) against target (both n^3 on
    // the device) and its gradient, written into (or added to) the gradient
    // buffer. Returns the loss (synchronises the stream).
    double loss_grad(hipStream_t s, int n, const float* field, float in_scale, const float* target, bool accumulate);
    // One Adam step (beta 0.9, 0.999, eps 1e-8) with gradient * grad_scale.
    void adam(hipStream_t s, float lr, float grad_scale);
    // The grid the network runs on inside a solver's learned start: 0 for
    // the solver's own grid, m > 0 for the residual restricted to m^3 (the
    // solver's n a multiple of m) and the prediction prolonged back.
    int grid() const { return grid_; }
    void set_grid(int m) { grid_ = m; }

private:
    struct Buffers;
    void ensure(int n);
    void forward(hipStream_t s, int n, const float* field, float in_scale, float out_scale, float* out);
    void pack(hipStream_t s);

    float* params_ = nullptr;
    float* grads_ = nullptr;
    float* m_ = nullptr;
    float* v_ = nullptr;
    bf16* packed_ = nullptr;  // forward and backward-data weights of every layer
    bool packed_valid_ = false;
    int adam_t_ = 0;
    int grid_ = 0;
    Buffers* buf_ = nullptr;
};

}  // namespace vgx

struct vgx_unet;
struct vgx_dataset;

namespace vgx {

// Device-resident training pairs of the learned warm start on an n^3 grid:
// inputs[k] is the residual of a projection's starting pressure divided by
// its RMS, targets[k] the solve's gauged correction to that pressure divided
// by the same RMS (both restricted to n^3 when the simulation's grid is a
// multiple of n); prev is scratch for the starting pressure on a grid equal
// to the dataset's.
struct Dataset {
    int n = 0, capacity = 0, size = 0;
    float* inputs = nullptr;
    float* targets = nullptr;
    float* prev = nullptr;
    size_t cells() const { return (size_t)n * n * n; }
};

UnetNet* unet_of(vgx_unet* u);
Dataset* dataset_of(vgx_dataset* d);

}  // namespace vgx
