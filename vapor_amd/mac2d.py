"""Python mirror of vapor's MAC 2D solvers over the HIP C ABI (include/vapor_mac2d.h).

Names and argument meanings follow the reference:

* ``MacGridParameters2D``        src/solvers/mac2d/MacGridFluidSolver2D.hpp:6-16
* ``MacGridState2D``             MacGridState2D.hpp:8-31 (host copy)
* ``GpuMacFluidSolver2D``        MacGridFluidSolver2D.hpp:19-51, device-resident
* ``SemiLagrangian2D``           SemiLagrangian2D.hpp:15-20, one call on the GPU
* ``GpuPressureSolver2D``        CpuPressureSolver2D.hpp:8-18, one call on the GPU
* ``ReflectionMacFluidSolver2D`` ReflectionMacFluidSolver2D.hpp:19-425, device-resident

Errors raise the Python counterparts of the reference's exceptions: ValueError
for std::invalid_argument, OverflowError for std::length_error, VaporError for
std::runtime_error. Every call runs on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .abi import (DENSITY_2D, VELOCITY_X_2D, VELOCITY_Y_2D, VgxDiagnostics, VgxProjectionOptions,
                  VgxProjectionResult, Vgx2Params, Vgx2rParams, check_status, float_ptr, load_library)

__all__ = [
    "MacGridParameters2D", "MacGridState2D", "GpuMacFluidSolver2D", "SemiLagrangian2D", "GpuPressureSolver2D",
    "ReflectionMacFluidSolver2D", "ReflectionParameters2D", "DENSITY_2D", "VELOCITY_X_2D", "VELOCITY_Y_2D",
]

_f64p = ctypes.POINTER(ctypes.c_double)


def _lib():
    return load_library()


def _alias(field):
    return property(lambda self: getattr(self, field), lambda self, value: setattr(self, field, value))


_PARAM_NAMES = {
    "reflection": "reflection",
    "sourceStrength": "source_strength",
    "emitterKickMultiplier": "emitter_kick_multiplier",
    "emitterRadius": "emitter_radius",
    "emitterCenterX": "emitter_center_x",
    "emitterCenterY": "emitter_center_y",
    "buoyancy": "buoyancy",
    "smokeDecay": "smoke_decay",
    "sorOmega": "sor_omega",
    "projectIterations": "project_iterations",
}


class MacGridParameters2D(Vgx2Params):
    """MacGridParameters2D{} with the reference's attribute names; ``plume=True``
    gives what MacGridFluidSolver2D() sets (MacGridFluidSolver2D.cpp:8-14)."""

    def __init__(self, plume: bool = False, **fields):
        super().__init__()
        lib = _lib()
        (lib.vgx2_params_plume if plume else lib.vgx2_params_default)(ctypes.byref(self))
        for key, value in fields.items():
            setattr(self, _PARAM_NAMES.get(key, key), value)


for _ref_name, _field in _PARAM_NAMES.items():
    if _ref_name != _field:
        setattr(MacGridParameters2D, _ref_name, _alias(_field))


class MacGridState2D:
    """Host copy of the 2D fields (MacGridState2D.hpp:8-31): n^2 density,
    (n+1)n x faces, n(n+1) y faces."""

    def __init__(self, resolution: int, box_size: float = 4.0):
        if resolution <= 0:
            raise ValueError("2D MAC resolution must be positive")
        n = resolution
        if n > (2**31 - 1) // (n + 1):
            raise OverflowError("2D MAC grid exceeds the supported index range")
        self._n = n
        self._box = np.float32(1.0)
        self.setBoxSize(box_size, _reset=False)
        self._rho = np.zeros(n * n, np.float32)
        self._u = np.zeros((n + 1) * n, np.float32)
        self._v = np.zeros(n * (n + 1), np.float32)

    def resolution(self) -> int:
        return self._n

    def boxSize(self) -> float:
        return float(self._box)

    def cellSize(self) -> float:
        return float(np.float32(self._box) / np.float32(self._n))

    def cellCount(self) -> int:
        return self._n * self._n

    def idC(self, x, y):
        return x + self._n * y

    def idX(self, x, y):
        return x + (self._n + 1) * y

    def idY(self, x, y):
        return x + self._n * y

    def density(self) -> np.ndarray:
        return self._rho

    def velocityX(self) -> np.ndarray:
        return self._u

    def velocityY(self) -> np.ndarray:
        return self._v

    def reset(self) -> None:
        for f in (self._rho, self._u, self._v):
            f[:] = 0.0

    def setBoxSize(self, size: float, _reset: bool = True) -> None:
        box = np.float32(size)
        n = np.float32(self._n)
        with np.errstate(all="ignore"):
            bad = (not np.isfinite(box) or box <= 0 or box / n <= 0
                   or not np.isfinite(np.float32(1) / (box / n)))
        if bad:
            raise ValueError("2D MAC requires finite positive size and spacing")
        changed = box != self._box
        self._box = box
        if _reset and changed:
            self.reset()

    def copy(self) -> "MacGridState2D":
        out = MacGridState2D(self._n, self.boxSize())
        out._rho[:] = self._rho
        out._u[:] = self._u
        out._v[:] = self._v
        return out


def _result(lib, fn, handle) -> VgxProjectionResult:
    r = VgxProjectionResult()
    check_status(lib, fn(handle, ctypes.byref(r)))
    return r


class GpuMacFluidSolver2D:
    """MacGridFluidSolver2D on the GPU. ``GpuMacFluidSolver2D()`` is the
    ready-to-run plume (128^2, box 1); ``GpuMacFluidSolver2D(n, box)`` uses
    MacGridParameters2D{}; ``GpuMacFluidSolver2D.from_state(state)`` adopts a
    host state's fields (MacGridFluidSolver2D(MacGridState2D))."""

    def __init__(self, resolution: int | None = None, box_size: float = 4.0,
                 parameters: Vgx2Params | None = None):
        self._lib = _lib()
        h = ctypes.c_void_p()
        if resolution is None:
            check_status(self._lib, self._lib.vgx2_sim_create_default(ctypes.byref(h)))
        else:
            p = ctypes.byref(parameters) if parameters is not None else None
            check_status(self._lib, self._lib.vgx2_sim_create(int(resolution), float(box_size), p, ctypes.byref(h)))
        self._h = h.value
        self._n = int(self._lib.vgx2_sim_resolution(self._h))
        self._params = MacGridParameters2D()
        self._lib.vgx2_sim_get_params(self._h, ctypes.byref(self._params))

    @classmethod
    def from_state(cls, state: MacGridState2D) -> "GpuMacFluidSolver2D":
        sim = cls(state.resolution(), state.boxSize())
        sim.upload(state)
        return sim

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx2_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def parameters(self) -> MacGridParameters2D:
        return self._params

    def resolution(self) -> int:
        return self._n

    def boxSize(self) -> float:
        return float(self._lib.vgx2_sim_box_size(self._h))

    def cellSize(self) -> float:
        return float(np.float32(self.boxSize()) / np.float32(self._n))

    def advance(self, dt: float) -> None:
        check_status(self._lib, self._lib.vgx2_sim_set_params(self._h, ctypes.byref(self._params)))
        check_status(self._lib, self._lib.vgx2_sim_advance(self._h, float(dt)))

    def resetState(self) -> None:
        check_status(self._lib, self._lib.vgx2_sim_reset(self._h))

    def setBoxSize(self, size: float) -> None:
        check_status(self._lib, self._lib.vgx2_sim_set_box(self._h, float(size)))

    def lastProjection(self) -> VgxProjectionResult:
        return _result(self._lib, self._lib.vgx2_sim_last_projection, self._h)

    def midpointProjection(self) -> VgxProjectionResult:
        return _result(self._lib, self._lib.vgx2_sim_midpoint_projection, self._h)

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        check_status(self._lib, self._lib.vgx2_sim_diagnostics(self._h, ctypes.byref(d)))
        return d

    def upload(self, state: MacGridState2D) -> None:
        if state.resolution() != self._n:
            raise ValueError("2D state resolution differs from the solver's")
        check_status(self._lib, self._lib.vgx2_sim_upload(self._h, float_ptr(state.density()),
                                                          float_ptr(state.velocityX()), float_ptr(state.velocityY())))

    def fields(self):
        n = self._n
        rho = np.empty(n * n, np.float32)
        u = np.empty((n + 1) * n, np.float32)
        v = np.empty(n * (n + 1), np.float32)
        check_status(self._lib, self._lib.vgx2_sim_download(self._h, float_ptr(rho), float_ptr(u), float_ptr(v)))
        return rho, u, v

    def state(self) -> MacGridState2D:
        s = MacGridState2D(self._n, self.boxSize())
        rho, u, v = self.fields()
        s._rho[:], s._u[:], s._v[:] = rho, u, v
        return s

    def deviceFields(self) -> dict:
        ptrs = [ctypes.c_void_p() for _ in range(3)]
        check_status(self._lib, self._lib.vgx2_sim_device_fields(self._h, *[ctypes.byref(p) for p in ptrs]))
        return dict(zip(("density", "velocityX", "velocityY"), (p.value for p in ptrs)))

    def synchronize(self) -> None:
        check_status(self._lib, self._lib.vgx2_sim_synchronize(self._h))


class SemiLagrangian2D:
    """SemiLagrangian2D::advect on the GPU (host arrays in and out)."""

    def advect(self, resolution: int, box_size: float, velocity_x, velocity_y, source, output, dt: float,
               field: int) -> None:
        lib = _lib()
        check_status(lib, lib.vgx2_backend_advect(int(resolution), float(box_size), int(field), float(dt),
                                                  float_ptr(velocity_x), float_ptr(velocity_y), float_ptr(source),
                                                  float_ptr(output)))


class GpuPressureSolver2D:
    """CpuPressureSolver2D::project on the GPU; ``pressure()`` is the FP64
    workspace of the last successful solve."""

    def __init__(self):
        self._pressure = np.zeros(0, np.float64)

    def project(self, state: MacGridState2D, options: VgxProjectionOptions) -> VgxProjectionResult:
        lib = _lib()
        n = state.resolution()
        p = np.zeros(n * n, np.float64)
        r = VgxProjectionResult()
        check_status(lib, lib.vgx2_backend_project(n, state.boxSize(), ctypes.byref(options),
                                                   float_ptr(state.velocityX()), float_ptr(state.velocityY()),
                                                   p.ctypes.data_as(_f64p), ctypes.byref(r)))
        self._pressure = p
        return r

    def pressure(self) -> np.ndarray:
        return self._pressure

    def reset(self) -> None:
        self._pressure[:] = 0.0


class ReflectionParameters2D(Vgx2rParams):
    """The historical solver's public parameters (member defaults, or the
    ReflectionMacFluidSolver2D(AdvectionMethod) plume with ``smoke_method``)."""

    def __init__(self, smoke_method: int | None = None, **fields):
        super().__init__()
        lib = _lib()
        if smoke_method is None:
            lib.vgx2r_params_default(ctypes.byref(self))
        else:
            check_status(lib, lib.vgx2r_params_plume(int(smoke_method), ctypes.byref(self)))
        for key, value in fields.items():
            setattr(self, key, value)


class ReflectionMacFluidSolver2D:
    """ReflectionMacFluidSolver2D on the GPU. The reference's public vectors
    are read with ``smoke()``, ``vx()``, ``vy()``, ``pressure()``,
    ``divergence()`` and written with ``upload``; its public methods keep
    their names (``advectScalar`` transports the solver's own smoke)."""

    def __init__(self, resolution: int = 128, parameters: Vgx2rParams | None = None, smoke_method: int | None = None):
        self._lib = _lib()
        if parameters is None:
            parameters = ReflectionParameters2D(smoke_method)
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx2r_sim_create(int(resolution), ctypes.byref(parameters), ctypes.byref(h)))
        self._h = h.value
        self.n = int(resolution)
        self.params = ReflectionParameters2D()
        self._lib.vgx2r_sim_get_params(self._h, ctypes.byref(self.params))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx2r_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _push(self) -> None:
        check_status(self._lib, self._lib.vgx2r_sim_set_params(self._h, ctypes.byref(self.params)))

    def h(self) -> float:
        return float(np.float32(1.0) / np.float32(self.n))

    def advance(self, dt: float) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_advance(self._h, float(dt)))

    def step(self, dt: float, buoyancy: float, dissipation: float, source_strength: float, iterations: int) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_step(self._h, float(dt), float(buoyancy), float(dissipation),
                                                         float(source_strength), int(iterations)))

    def emit(self, strength: float, dt: float) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_emit(self._h, float(strength), float(dt)))

    def applyBoundary(self) -> None:
        check_status(self._lib, self._lib.vgx2r_sim_apply_boundary(self._h))

    def advectScalar(self, dt: float) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_advect_smoke(self._h, float(dt)))

    def project(self, dt: float, iterations: int) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_project(self._h, float(dt), int(iterations)))

    def resetState(self) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx2r_sim_reset(self._h))

    def setBoxSize(self, size: float) -> None:
        check_status(self._lib, self._lib.vgx2r_sim_set_box(self._h, float(size)))

    def boxSize(self) -> float:
        return 1.0

    def upload(self, smoke=None, vx=None, vy=None, pressure=None, divergence=None) -> None:
        check_status(self._lib, self._lib.vgx2r_sim_upload(self._h, float_ptr(smoke), float_ptr(vx), float_ptr(vy),
                                                           float_ptr(pressure), float_ptr(divergence)))

    def fields(self):
        n = self.n
        out = [np.empty(n * n, np.float32), np.empty((n + 1) * n, np.float32), np.empty(n * (n + 1), np.float32),
               np.empty(n * n, np.float32), np.empty(n * n, np.float32)]
        check_status(self._lib, self._lib.vgx2r_sim_download(self._h, *[float_ptr(a) for a in out]))
        return dict(zip(("smoke", "vx", "vy", "pressure", "divergence"), out))

    def smoke(self) -> np.ndarray:
        return self.fields()["smoke"]

    def vx(self) -> np.ndarray:
        return self.fields()["vx"]

    def vy(self) -> np.ndarray:
        return self.fields()["vy"]

    def randomState(self) -> int:
        s = ctypes.c_uint32()
        check_status(self._lib, self._lib.vgx2r_sim_random_state(self._h, ctypes.byref(s)))
        return int(s.value)

    def setRandomState(self, state: int) -> None:
        check_status(self._lib, self._lib.vgx2r_sim_set_random_state(self._h, ctypes.c_uint32(state)))

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        check_status(self._lib, self._lib.vgx2r_sim_diagnostics(self._h, ctypes.byref(d)))
        return d

    def idC(self, x, y):
        return x + self.n * y

    def idX(self, x, y):
        return x + (self.n + 1) * y

    def idY(self, x, y):
        return x + self.n * y
