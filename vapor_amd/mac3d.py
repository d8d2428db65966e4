"""Python mirror of vapor's MAC 3D solver interface over the HIP C ABI.

Names and argument meanings follow the reference classes so that code (and
tests) written against them read the same:

* ``MacGridState3D``        src/solvers/mac3d/MacGridState3D.hpp:11-51
* ``MacGridParameters3D``   src/solvers/mac3d/MacGridFluidSolver3D.hpp:14-36
* ``GpuMacSimulation3D``    src/solvers/mac3d/gpu/GpuMacSimulation3D.hpp:9-27
                            (with the MacGridFluidSolver3D accessors)
* ``GpuAdvection`` /
  ``GpuPressureSolver``     the staged plug-ins, Advection3D.hpp:6-14 and
                            Projection3D.hpp:6-11, over one ``GpuMacBackend``.

Errors raise the Python counterparts of the reference's exceptions:
ValueError for std::invalid_argument, OverflowError for std::length_error,
VaporError for device failures. Every call runs on the GPU; there is no CPU
fallback, and importing a class fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .abi import (VgxBatchTuning, VgxDiagnostics, VgxFieldError, VgxParams, VgxProjectionOptions,
                  VgxProjectionResult, check_status, float_ptr, load_library)

__all__ = [
    "MacGridState3D", "MacGridParameters3D", "ProjectionOptions", "GpuMacSimulation3D", "GpuMacBackend",
    "GpuAdvection", "GpuPressureSolver", "device_name",
]

_INT_MAX = 2**31 - 1


def _lib():
    return load_library()


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    lib = _lib()
    check_status(lib, lib.vgx_device_name(buf, 256))
    return buf.value.decode()


class MacGridState3D:
    """Host copy of the physical fields (MacGridState3D.hpp:11-51)."""

    def __init__(self, resolution: int, box_size: float = 4.5):
        if resolution <= 0:
            raise ValueError("MAC grid resolution must be positive")
        n = resolution
        if n > _INT_MAX // n or n * n > _INT_MAX // (n + 1):
            raise OverflowError("MAC grid exceeds the supported index range")
        self._n = n
        self._box = np.float32(1.0)
        self.setBoxSize(box_size, _reset=False)
        self._rho = np.zeros(n * n * n, np.float32)
        self._u = np.zeros((n + 1) * n * n, np.float32)
        self._v = np.zeros((n + 1) * n * n, np.float32)
        self._w = np.zeros((n + 1) * n * n, np.float32)

    def resolution(self) -> int:
        return self._n

    def boxSize(self) -> float:
        return float(self._box)

    def cellSize(self) -> float:
        return float(np.float32(self._box) / np.float32(self._n))

    def cellCount(self) -> int:
        return self._n ** 3

    def idC(self, x, y, z):
        n = self._n
        return x + n * (y + n * z)

    def idX(self, x, y, z):
        n = self._n
        return x + (n + 1) * (y + n * z)

    def idY(self, x, y, z):
        n = self._n
        return x + n * (y + (n + 1) * z)

    def idZ(self, x, y, z):
        return self.idC(x, y, z)

    def density(self) -> np.ndarray:
        return self._rho

    def velocityX(self) -> np.ndarray:
        return self._u

    def velocityY(self) -> np.ndarray:
        return self._v

    def velocityZ(self) -> np.ndarray:
        return self._w

    def reset(self) -> None:
        for f in (self._rho, self._u, self._v, self._w):
            f[:] = 0.0

    def setBoxSize(self, box_size: float, _reset: bool = True) -> None:
        box = np.float32(box_size)
        if not np.isfinite(box) or box <= 0 or box / np.float32(self._n) <= 0:
            raise ValueError("MAC grid requires a finite positive box size and spacing")
        changed = box != self._box
        self._box = box
        if _reset and changed:
            self.reset()

    def copy(self) -> "MacGridState3D":
        out = MacGridState3D(self._n, self.boxSize())
        for dst, src in zip((out._rho, out._u, out._v, out._w), (self._rho, self._u, self._v, self._w)):
            dst[:] = src
        return out


# reference attribute name -> vgx_params field
_PARAM_NAMES = {
    "emitterRadius": "emitter_radius",
    "emitterCenterX": "emitter_center_x",
    "emitterCenterY": "emitter_center_y",
    "emitterCenterZ": "emitter_center_z",
    "emitterKickMultiplier": "emitter_kick_multiplier",
    "emitterKickBase": "emitter_kick_base",
    "emitterKickScale": "emitter_kick_scale",
    "stirStrength": "stir_strength",
    "randomSeed": "random_seed",
    "sorOmega": "sor_omega",
    "buoyancySplit": "buoyancy_split",
    "cflMc": "cfl_mc",
    "macCormackSmoke": "maccormack_smoke",
    "macCormackVel": "maccormack_vel",
    "smokeDecay": "smoke_decay",
    "reflection": "reflection",
    "buoyancy": "buoyancy",
    "sourceStrength": "source_strength",
    "projectIterations": "project_iterations",
}


def _alias(field):
    return property(lambda self: getattr(self, field), lambda self, value: setattr(self, field, value))


class MacGridParameters3D(VgxParams):
    """MacGridParameters3D with the reference's attribute names as aliases of
    the C struct fields. ``MacGridParameters3D()`` gives the defaults,
    ``MacGridParameters3D(algorithm)`` the MacAlgorithm presets."""

    def __init__(self, algorithm: int | None = None, **fields):
        super().__init__()
        lib = _lib()
        if algorithm is None:
            lib.vgx_params_default(ctypes.byref(self))
        else:
            check_status(lib, lib.vgx_params_preset(int(algorithm), ctypes.byref(self)))
        for key, value in fields.items():
            setattr(self, _PARAM_NAMES.get(key, key), value)

    def copy(self) -> "MacGridParameters3D":
        out = MacGridParameters3D.__new__(MacGridParameters3D)
        VgxParams.__init__(out)
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(VgxParams))
        return out


for _ref_name, _field in _PARAM_NAMES.items():
    if _ref_name != _field:  # identical names are the struct fields themselves
        setattr(MacGridParameters3D, _ref_name, _alias(_field))


class ProjectionOptions(VgxProjectionOptions):
    """ProjectionOptions + LinearSolveOptions defaults (diagonal floor 6)."""

    def __init__(self, **fields):
        super().__init__()
        _lib().vgx_projection_options_default(ctypes.byref(self))
        for key, value in fields.items():
            setattr(self, key, value)


class GpuMacSimulation3D:
    """GPU-resident MAC 3D simulation (GpuMacSimulation3D.hpp:9-27).

    ``GpuMacSimulation3D(n, box, parameters)``; ``parameters=None`` gives the
    reference defaults with projectIterations = 2n (MacGridFluidSolver3D(int)).
    ``GpuMacSimulation3D.from_algorithm(MacAlgorithm)`` gives 32^3, box 4.5 and
    the preset. ``parameters()`` returns the live parameter struct; edits take
    effect at the next ``advance``.
    """

    def __init__(self, resolution: int, box_size: float = 4.5, parameters: VgxParams | None = None):
        self._lib = _lib()
        h = ctypes.c_void_p()
        p = ctypes.byref(parameters) if parameters is not None else None
        check_status(self._lib, self._lib.vgx_sim_create(int(resolution), float(box_size), p, ctypes.byref(h)))
        self._h = h.value
        self._n = int(resolution)
        self._params = MacGridParameters3D()
        self._lib.vgx_sim_get_params(self._h, ctypes.byref(self._params))

    @classmethod
    def from_algorithm(cls, algorithm: int) -> "GpuMacSimulation3D":
        return cls(32, 4.5, MacGridParameters3D(algorithm))

    @classmethod
    def _borrowed(cls, lib, handle: int, owner) -> "GpuMacSimulation3D":
        """A view of a scene a batch owns: never destroyed from here."""
        sim = cls.__new__(cls)
        sim._lib = lib
        sim._h = handle
        sim._n = int(lib.vgx_sim_resolution(handle))
        sim._params = MacGridParameters3D()
        lib.vgx_sim_get_params(handle, ctypes.byref(sim._params))
        sim._owner = owner
        return sim

    def close(self) -> None:
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                self._lib.vgx_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _push(self) -> None:
        check_status(self._lib, self._lib.vgx_sim_set_params(self._h, ctypes.byref(self._params)))

    def parameters(self) -> MacGridParameters3D:
        return self._params

    def resolution(self) -> int:
        return self._n

    def boxSize(self) -> float:
        return float(self._lib.vgx_sim_box_size(self._h))

    def advance(self, dt: float) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx_sim_advance(self._h, float(dt)))

    def resetState(self) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx_sim_reset(self._h))

    def setBoxSize(self, size: float) -> None:
        self._push()
        check_status(self._lib, self._lib.vgx_sim_set_box(self._h, float(size)))

    def lastProjection(self, index: int = 0) -> VgxProjectionResult:
        r = VgxProjectionResult()
        check_status(self._lib, self._lib.vgx_sim_last_projection(self._h, index, ctypes.byref(r)))
        return r

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        check_status(self._lib, self._lib.vgx_sim_diagnostics(self._h, ctypes.byref(d)))
        return d

    def kineticEnergy(self) -> float:
        return self.diagnostics().kinetic_energy

    def maxVelocity(self) -> float:
        return self.diagnostics().max_velocity

    def computeDivergenceNorm(self) -> float:
        return self.diagnostics().max_divergence

    def _check_state(self, state: MacGridState3D) -> None:
        if state.resolution() != self._n:
            raise ValueError("GPU snapshot resolution mismatch")
        if np.float32(state.boxSize()) != np.float32(self.boxSize()):
            raise ValueError("GPU snapshot box mismatch")

    def upload(self, state: MacGridState3D) -> None:
        self._check_state(state)
        check_status(self._lib, self._lib.vgx_sim_upload(
            self._h, float_ptr(state.density()), float_ptr(state.velocityX()),
            float_ptr(state.velocityY()), float_ptr(state.velocityZ())))

    def download(self, state: MacGridState3D) -> None:
        self._check_state(state)
        check_status(self._lib, self._lib.vgx_sim_download(
            self._h, float_ptr(state.density()), float_ptr(state.velocityX()),
            float_ptr(state.velocityY()), float_ptr(state.velocityZ())))

    def setUnet(self, net: "UNet | None") -> None:
        """Extension: learned warm start from ``net`` (None disables); the
        network is kept alive while set."""
        check_status(self._lib, self._lib.vgx_sim_set_unet(self._h, net._h if net is not None else None))
        self._net = net

    def record(self, data: "Dataset | None") -> None:
        """Extension: append training pairs from each warm tolerance
        projection to ``data`` (None stops)."""
        check_status(self._lib, self._lib.vgx_sim_record(self._h, data._h if data is not None else None))
        self._data = data

    def addVelocityNoise(self, sigma: float, seed: int) -> None:
        """Extension: add sigma * N(0, 1) to every face velocity (counter-based,
        Philox4x32-10 + Box-Muller; the same bits as the oracle's)."""
        check_status(self._lib, self._lib.vgx_sim_add_velocity_noise(self._h, float(sigma), int(seed)))

    def state(self) -> MacGridState3D:
        s = MacGridState3D(self._n, self.boxSize())
        self.download(s)
        return s

    def fields(self):
        s = self.state()
        return s.density(), s.velocityX(), s.velocityY(), s.velocityZ()

    def pressure(self) -> np.ndarray:
        out = np.empty(self._n ** 3, np.float32)
        check_status(self._lib, self._lib.vgx_sim_download_pressure(self._h, float_ptr(out)))
        return out

    def setPressure(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self._n ** 3:
            raise ValueError("pressure must have n^3 values")
        check_status(self._lib, self._lib.vgx_sim_upload_pressure(self._h, float_ptr(p)))

    def saveSnapshot(self, path) -> None:
        """Extension: write the state as raw FP32 files plus a JSON header
        (vgx_sim_save_snapshot); read_snapshot() reads them back on the host."""
        check_status(self._lib, self._lib.vgx_sim_save_snapshot(self._h, str(path).encode()))

    def loadSnapshot(self, path) -> None:
        """Extension: restore fields, pressure histories and RNG state from a
        snapshot of the same grid; the trajectory continues bit for bit."""
        check_status(self._lib, self._lib.vgx_sim_load_snapshot(self._h, str(path).encode()))

    def deviceFields(self) -> dict:
        """Device addresses of the current density, faces and pressure (no
        download; RenderData's hand-off), valid until the next advance/reset.
        The handle's stream is synchronised first."""
        self.synchronize()
        ptrs = [ctypes.c_void_p() for _ in range(5)]
        check_status(self._lib, self._lib.vgx_sim_device_fields(self._h, *(ctypes.byref(q) for q in ptrs)))
        return dict(zip(("rho", "u", "v", "w", "p"), (q.value for q in ptrs)))

    def uploadedBytes(self) -> int:
        up = ctypes.c_uint64()
        self._lib.vgx_sim_transfer_bytes(self._h, ctypes.byref(up), None)
        return up.value

    def downloadedBytes(self) -> int:
        down = ctypes.c_uint64()
        self._lib.vgx_sim_transfer_bytes(self._h, None, ctypes.byref(down))
        return down.value

    def synchronize(self) -> None:
        check_status(self._lib, self._lib.vgx_sim_synchronize(self._h))

    # --- timing (HIP events on the simulation's stream) ---
    def set_timing(self, enabled: bool) -> None:
        check_status(self._lib, self._lib.vgx_sim_set_timing(self._h, int(bool(enabled))))

    def reset_timing(self) -> None:
        check_status(self._lib, self._lib.vgx_sim_reset_timing(self._h))

    def kernel_timing(self) -> dict:
        out = {}
        for kind in range(self._lib.vgx_kernel_kinds()):
            ms = ctypes.c_double()
            launches = ctypes.c_uint64()
            check_status(self._lib, self._lib.vgx_sim_kernel_timing(self._h, kind, ctypes.byref(ms),
                                                                    ctypes.byref(launches)))
            out[self._lib.vgx_kernel_name(kind).decode()] = (ms.value, launches.value)
        return out

    def sweep_count(self) -> int:
        s = ctypes.c_uint64()
        check_status(self._lib, self._lib.vgx_sim_sweep_count(self._h, ctypes.byref(s)))
        return s.value


def read_snapshot(path) -> dict:
    """Host reader of a snapshot written by saveSnapshot: the JSON header and
    one float32 array per field (named as in the header), without a GPU."""
    import json
    from pathlib import Path

    base = Path(str(path))
    header = json.loads(Path(str(base) + ".json").read_text())
    if header.get("format") != "vapor-mac3d-snapshot" or header.get("version") != 1:
        raise ValueError("not a MAC 3D snapshot header")
    out = {"header": header}
    for f in header["fields"]:
        data = np.fromfile(base.parent / f["file"], dtype="<f4")
        if data.size != f["count"]:
            raise ValueError(f"snapshot field {f['name']} has the wrong size")
        out[f["name"]] = data.astype(np.float32, copy=False)
    return out


class MacBatch3D:
    """Extension: independent scenes of uneven sizes stepped together, one
    launch per stage over all scenes (vgx_batch_*).

    ``MacBatch3D(resolutions, boxes, parameters)`` with one entry per scene.
    ``scene(i)`` is a borrowed GpuMacSimulation3D view for downloads,
    diagnostics and projection results; only the batch advances it.
    """

    def __init__(self, resolutions, boxes, parameters):
        self._lib = _lib()
        count = len(resolutions)
        if len(boxes) != count or len(parameters) != count:
            raise ValueError("one resolution, box and parameter set per scene")
        n = (ctypes.c_int * count)(*[int(r) for r in resolutions])
        box = (ctypes.c_float * count)(*[float(b) for b in boxes])
        params = (VgxParams * count)()
        for k, p in enumerate(parameters):
            ctypes.memmove(ctypes.byref(params[k]), ctypes.byref(p), ctypes.sizeof(VgxParams))
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx_batch_create(count, n, box, params, ctypes.byref(h)))
        self._h = h.value
        self._scenes = []
        for k in range(count):
            sh = ctypes.c_void_p()
            check_status(self._lib, self._lib.vgx_batch_scene(self._h, k, ctypes.byref(sh)))
            self._scenes.append(GpuMacSimulation3D._borrowed(self._lib, sh.value, self))

    def close(self) -> None:
        if getattr(self, "_h", None):
            for s in self._scenes:
                s._h = None
            self._scenes = []
            self._lib.vgx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __len__(self) -> int:
        return len(self._scenes)

    def scene(self, index: int) -> GpuMacSimulation3D:
        return self._scenes[index]

    def advance(self, dt: float) -> None:
        check_status(self._lib, self._lib.vgx_batch_advance(self._h, float(dt)))

    def resetState(self) -> None:
        check_status(self._lib, self._lib.vgx_batch_reset(self._h))

    def sceneLatency(self) -> np.ndarray:
        """Per scene: ms from the last step's start until its last projection
        met the tolerance (-1 if it did not, or in fixed mode)."""
        out = np.empty(len(self._scenes), np.float64)
        check_status(self._lib, self._lib.vgx_batch_scene_latency(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def tuning(self) -> dict:
        t = VgxBatchTuning()
        check_status(self._lib, self._lib.vgx_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"elementwise_per_thread": 1 << t.elementwise_per_thread_log2, "advect_grid": t.advect_grid,
                "sweep_grid": t.sweep_grid, "sweep_staging": "registers" if t.sweep_register else "lds",
                "stencil_grid": t.stencil_grid,
                "elementwise_items": t.elementwise_items, "tile_items": t.tile_items,
                "stencil_items": t.stencil_items, "advect_ms": list(t.advect_ms), "sweep_ms": list(t.sweep_ms),
                "sweep_lds_ms": list(t.sweep_lds_ms), "stencil_ms": list(t.stencil_ms)}

    def synchronize(self) -> None:
        check_status(self._lib, self._lib.vgx_batch_synchronize(self._h))

    def set_timing(self, enabled: bool) -> None:
        check_status(self._lib, self._lib.vgx_batch_set_timing(self._h, int(bool(enabled))))

    def reset_timing(self) -> None:
        check_status(self._lib, self._lib.vgx_batch_reset_timing(self._h))

    def kernel_timing(self) -> dict:
        out = {}
        for kind in range(self._lib.vgx_kernel_kinds()):
            ms = ctypes.c_double()
            launches = ctypes.c_uint64()
            check_status(self._lib, self._lib.vgx_batch_kernel_timing(self._h, kind, ctypes.byref(ms),
                                                                      ctypes.byref(launches)))
            out[self._lib.vgx_kernel_name(kind).decode()] = (ms.value, launches.value)
        return out


class ShadowRun3D:
    """Extension: a copy of a simulation stepped by plain kernels in a chosen
    precision (vgx_shadow_*). ``ShadowRun3D(sim, 64)`` is the all-FP64 run of
    the same scene; ``ShadowRun3D(sim, 32)`` steps it in FP32 by a second route.
    The copy takes the simulation's grid, parameters, fields, RNG state and
    pressure histories at creation; afterwards the two advance independently.
    """

    FIELDS = ("rho", "u", "v", "w", "p")

    def __init__(self, sim: GpuMacSimulation3D, precision: int = 64):
        self._lib = _lib()
        self.n = sim.resolution()
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx_shadow_create(sim._h, int(precision), ctypes.byref(h)))
        self._h = h.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx_shadow_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def advance(self, dt: float) -> None:
        check_status(self._lib, self._lib.vgx_shadow_advance(self._h, float(dt)))

    def fields(self):
        """(rho, u, v, w, p) as float64 arrays."""
        n = self.n
        out = [np.empty(n ** 3), *(np.empty((n + 1) * n * n) for _ in range(3)), np.empty(n ** 3)]
        dp = ctypes.POINTER(ctypes.c_double)
        check_status(self._lib, self._lib.vgx_shadow_download(self._h, *(a.ctypes.data_as(dp) for a in out)))
        return tuple(out)

    def lastProjection(self, index: int = 0) -> VgxProjectionResult:
        r = VgxProjectionResult()
        check_status(self._lib, self._lib.vgx_shadow_last_projection(self._h, int(index), ctypes.byref(r)))
        return r

    def sweep_count(self) -> int:
        s = ctypes.c_uint64()
        check_status(self._lib, self._lib.vgx_shadow_sweep_count(self._h, ctypes.byref(s)))
        return s.value

    def compare(self, sim: GpuMacSimulation3D) -> dict:
        """Max and RMS error of sim's density, u, v, w and pressure against this
        run's, with this run's max and RMS magnitude, per field."""
        errs = (VgxFieldError * 5)()
        check_status(self._lib, self._lib.vgx_shadow_compare(self._h, sim._h, errs))
        return {name: {"max_abs": e.max_abs, "rms": e.rms, "ref_max": e.ref_max, "ref_rms": e.ref_rms}
                for name, e in zip(self.FIELDS, errs)}


class UNet:
    """Extension: the learned warm start's 3D U-Net on the GPU (BF16 on the
    matrix cores, FP32 masters, Adam)."""

    def __init__(self, seed: int = 0):
        self._lib = _lib()
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx_unet_create(int(seed), ctypes.byref(h)))
        self._h = h.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx_unet_destroy(self._h)
            self._h = None

    __del__ = close

    def param_count(self) -> int:
        return int(self._lib.vgx_unet_param_count())

    def params(self) -> np.ndarray:
        out = np.empty(self.param_count(), np.float32)
        check_status(self._lib, self._lib.vgx_unet_get_params(self._h, float_ptr(out)))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.param_count():
            raise ValueError("U-Net parameter count mismatch")
        check_status(self._lib, self._lib.vgx_unet_set_params(self._h, float_ptr(p)))

    def save(self, path) -> None:
        self.params().tofile(str(path))

    def load(self, path) -> None:
        self.set_params(np.fromfile(str(path), np.float32))

    def set_grid(self, m: int) -> None:
        """The grid the network runs on inside a solver's learned start (0: the
        solver's own; m: the residual restricted to m^3, prediction prolonged)."""
        check_status(self._lib, self._lib.vgx_unet_set_grid(self._h, int(m)))

    def grid(self) -> int:
        return int(self._lib.vgx_unet_get_grid(self._h))

    def predict(self, n: int, field, in_scale: float = 1.0, out_scale: float = 1.0) -> np.ndarray:
        f = np.ascontiguousarray(field, np.float32)
        out = np.empty(n ** 3, np.float32)
        check_status(self._lib, self._lib.vgx_unet_predict(self._h, int(n), float_ptr(f), float(in_scale),
                                                           float(out_scale), float_ptr(out)))
        return out

    def train(self, data: "Dataset", epochs: int, batch: int, lr: float, seed: int = 0) -> float:
        loss = ctypes.c_double()
        check_status(self._lib, self._lib.vgx_unet_train(self._h, data._h, int(epochs), int(batch), float(lr),
                                                         int(seed), ctypes.byref(loss)))
        return loss.value


class Dataset:
    """Extension: device-resident (residual, correction) training pairs."""

    def __init__(self, n: int, capacity: int):
        self._lib = _lib()
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx_dataset_create(int(n), int(capacity), ctypes.byref(h)))
        self._h = h.value
        self.n = int(n)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx_dataset_destroy(self._h)
            self._h = None

    __del__ = close

    def __len__(self) -> int:
        return int(self._lib.vgx_dataset_size(self._h))

    def clear(self) -> None:
        check_status(self._lib, self._lib.vgx_dataset_clear(self._h))

    def pair(self, index: int):
        """(input, target) of pair `index` as host arrays of n^3 floats."""
        cells = self.n ** 3
        x, y = np.empty(cells, np.float32), np.empty(cells, np.float32)
        check_status(self._lib, self._lib.vgx_dataset_get(self._h, int(index), float_ptr(x), float_ptr(y)))
        return x, y


class GpuMacBackend:
    """Staged device backend shared by the plug-ins (GpuMacBackend.hpp:7-48):
    each call uploads host fields, runs on the GPU and downloads the result."""

    def __init__(self, resolution: int, box_size: float):
        self._lib = _lib()
        h = ctypes.c_void_p()
        check_status(self._lib, self._lib.vgx_backend_create(int(resolution), float(box_size), ctypes.byref(h)))
        self._h = h.value
        self.n = int(resolution)
        self.box = float(box_size)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgx_backend_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h


def _faces(flow):
    u, v, w = flow
    return float_ptr(u), float_ptr(v), float_ptr(w)


class GpuAdvection:
    """Advection3D over a GpuMacBackend (Advection3D.hpp:6-14). ``flow`` is the
    tracer velocity (u, v, w) of the AdvectionContext."""

    def __init__(self, backend: GpuMacBackend):
        self._b = backend
        self._lib = backend._lib

    def advectScalar(self, flow, field: np.ndarray, dt: float, method: int, cfl_limit: float = 3.0) -> None:
        if field.size != self._b.n ** 3:
            raise ValueError("Scalar field must match the grid cell count")
        check_status(self._lib, self._lib.vgx_backend_advect_scalar(
            self._b.handle, int(method), float(dt), float(cfl_limit), *_faces(flow), float_ptr(field)))

    def advectFace(self, flow, axis: int, source: np.ndarray, output: np.ndarray, dt: float, method: int,
                   cfl_limit: float = 3.0) -> None:
        faces = (self._b.n + 1) * self._b.n ** 2
        if source.size != faces or output.size != faces:
            raise ValueError("Velocity fields must match their grid layout and not overlap")
        check_status(self._lib, self._lib.vgx_backend_advect_face(
            self._b.handle, int(axis), int(method), float(dt), float(cfl_limit), *_faces(flow),
            float_ptr(source), float_ptr(output)))


class GpuPressureSolver:
    """Projection3D over a GpuMacBackend (Projection3D.hpp:6-11). project()
    projects the state's faces in place and returns the ProjectionResult."""

    def __init__(self, backend: GpuMacBackend):
        self._b = backend
        self._lib = backend._lib
        self._pressure = np.zeros(backend.n ** 3, np.float32)

    def project(self, state: MacGridState3D, options: VgxProjectionOptions) -> VgxProjectionResult:
        if state.resolution() != self._b.n or np.float32(state.boxSize()) != np.float32(self._b.box):
            raise ValueError("state does not match the backend grid")
        r = VgxProjectionResult()
        check_status(self._lib, self._lib.vgx_backend_project(
            self._b.handle, ctypes.byref(options), float_ptr(state.velocityX()), float_ptr(state.velocityY()),
            float_ptr(state.velocityZ()), float_ptr(self._pressure), ctypes.byref(r)))
        return r

    def pressure(self) -> np.ndarray:
        return self._pressure

    def reset(self) -> None:
        self._pressure[:] = 0.0


# MacAlgorithm names
SimpleSemiLagrangian = abi.SIMPLE_SEMI_LAGRANGIAN
SimpleMacCormack = abi.SIMPLE_MACCORMACK
ReflectionSemiLagrangian = abi.REFLECTION_SEMI_LAGRANGIAN
ReflectionMacCormack = abi.REFLECTION_MACCORMACK
