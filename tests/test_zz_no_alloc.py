"""vgx_advect_face_noalloc: face advection on caller-owned device buffers with
no allocation (the MacCormack forward trace in the library's fixed device
array), against vgx_backend_advect_face on the same inputs, bit for bit. Two
calls in a row share the array, the second on a smaller grid; a grid beyond
the array is refused before any launch."""
import ctypes

import numpy as np
import pytest

from helpers import bits_equal
from vapor_amd import abi, mac3d

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.owned = []

    def upload(self, array):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), array.nbytes) == 0
        self.owned.append(p.value)
        assert self.lib.hipMemcpy(p.value, array.ctypes.data, array.nbytes, H2D) == 0
        return p.value

    def empty(self, count):
        return self.upload(np.zeros(count, np.float32))

    def download(self, ptr, count):
        out = np.empty(count, np.float32)
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def close(self):
        for p in self.owned:
            self.lib.hipFree(p)
        self.owned = []


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    faces = (n + 1) * n * n
    flow = [(rng.standard_normal(faces) * 0.6).astype(np.float32) for _ in range(3)]
    src = rng.standard_normal(faces).astype(np.float32)
    return flow, src


def staged(n, box, axis, method, dt, cfl, flow, src):
    backend = mac3d.GpuMacBackend(n, box)
    out = np.zeros_like(src)
    mac3d.GpuAdvection(backend).advectFace(flow, axis, src, out, dt, method, cfl)
    backend.close()
    return out


def test_noalloc_face_advection_matches_the_staged_entry_point():
    lib = abi.load_library()
    hip = Hip()
    try:
        # MacCormack (uses the scratch) then a smaller grid (the array keeps
        # the first call's contents), then semi-Lagrangian
        for n, axis, method, seed in ((96, 1, abi.MACCORMACK, 1), (40, 2, abi.MACCORMACK, 2),
                                      (24, 0, abi.SEMI_LAGRANGIAN, 3)):
            box, dt, cfl = 1.0, 1.0 / 60.0, 3.0
            flow, src = inputs(n, seed)
            want = staged(n, box, axis, method, dt, cfl, flow, src)
            d_flow = [hip.upload(f) for f in flow]
            d_src = hip.upload(src)
            d_out = hip.empty(src.size)
            print(f"[gpu no-alloc] n={n} axis={axis} method={method}", flush=True)
            rc = lib.vgx_advect_face_noalloc(n, box, axis, method, dt, cfl, *d_flow, d_src, d_out, None)
            assert rc == abi.VGX_OK, lib.vgx_last_error()
            hip.sync()
            got = hip.download(d_out, src.size)
            assert bits_equal(got, want), (n, float(np.max(np.abs(got - want))))
            assert np.any(got != src)  # the flow moved the field
        # beyond the fixed array: refused, nothing launched
        big = 129
        rc = lib.vgx_advect_face_noalloc(big, 1.0, 0, abi.MACCORMACK, 0.01, 3.0, d_flow[0], d_flow[1], d_flow[2],
                                         d_src, d_out, None)
        assert rc == abi.VGX_ERANGE and b"128" in lib.vgx_last_error()
        # the staged checks: aliasing and a negative dt
        assert lib.vgx_advect_face_noalloc(24, 1.0, 0, 0, 0.01, 3.0, *d_flow, d_src, d_src, None) == abi.VGX_EINVAL
        assert lib.vgx_advect_face_noalloc(24, 1.0, 0, 0, -1.0, 3.0, *d_flow, d_src, d_out, None) == abi.VGX_EINVAL
    finally:
        hip.sync()
        hip.close()
