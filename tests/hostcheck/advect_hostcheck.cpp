// advect_hostcheck.cpp - CPU build of the device advection code, with every
// field read bounds-checked (test infrastructure).
//
// The advection kernels call vgx::sl_value / vgx::mc_value, which are
// __host__ __device__ templates over the field accessor. Here they run on the
// host over accessors that count out-of-range indices instead of reading
// them, so the kernels' indexing is checked on the CPU for any input
// (including NaN, infinite and huge velocities) before it runs on a GPU.
#include "../../vapor_amd/csrc/vgx_device.hpp"

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

template <int AXIS>
long run(int n, float box, int method, float dt, float cfl, const float* u, const float* v, const float* w,
         const float* src, const float* tsrc, float* out) {
    long oob = 0;
    const long faces = (long)(n + 1) * n * n, cells = (long)n * n * n;
    const long count = AXIS < 0 ? cells : faces;
    const float h = box / (float)n;
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
}
