"""Python bindings of the checkers (TEST INFRASTRUCTURE ONLY).

Two libraries, both built by oracle/Makefile:
  * ``liboracle_mac3d.so`` - the C restatement (mac3d_oracle.c), always built;
  * ``_ref/libvapor_ref.so`` - vapor's own CPU solver compiled from the
    reference tree, present only where that tree was available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package. The product (vapor_amd) never does.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from vapor_amd.abi import VgxDiagnostics, VgxParams, VgxProjectionOptions, VgxProjectionResult

ORACLE_DIR = Path(__file__).resolve().parent
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle_mac3d.so"
ORACLE_OMP_LIB = ORACLE_DIR / "_build" / "liboracle_mac3d_omp.so"
REF_LIB = ORACLE_DIR / "_ref" / "libvapor_ref.so"

_P = ctypes.POINTER
_f = _P(ctypes.c_float)
_c = ctypes.c_int
_fl = ctypes.c_float

_ORACLE_SIGS = [
    ("orc_params_default", None, [_P(VgxParams)]),
    ("orc_projection_options_default", None, [_P(VgxProjectionOptions)]),
    ("orc_sample", _fl, [_c, _fl, _c, _f, _fl, _fl, _fl]),
    ("orc_advect_scalar", _c, [_c, _fl, _f, _f, _f, _fl, _fl, _c, _f, _f]),
    ("orc_advect_face", _c, [_c, _fl, _c, _f, _f, _f, _f, _fl, _fl, _c, _f, _f]),
    ("orc_enforce_walls", None, [_c, _f, _f, _f]),
    ("orc_divergence", None, [_c, _fl, _f, _f, _f, _f]),
    ("orc_max_divergence", _c, [_c, _fl, _f, _f, _f, _f]),
    ("orc_apply_gradient", None, [_c, _fl, _f, _fl, _f, _f, _f]),
    ("orc_residual_inf", ctypes.c_double, [_c, _f, _f]),
    ("orc_residual_l2", ctypes.c_double, [_c, _f, _f]),
    ("orc_norm_l2", ctypes.c_double, [ctypes.c_size_t, _f]),
    ("orc_sor_solve", _c, [_c, _fl, _fl, _f, _f, _P(VgxProjectionOptions), _P(VgxProjectionResult)]),
    ("orc_sor_sweep", None, [_c, _fl, _fl, _f, _f]),
    ("orc_project", _c, [_c, _fl, _P(VgxProjectionOptions), _f, _f, _f, _f, _f, _f, _P(VgxProjectionResult)]),
    ("orc_sim_create", ctypes.c_void_p, [_c, _fl, _P(VgxParams)]),
    ("orc_sim_destroy", None, [ctypes.c_void_p]),
    ("orc_sim_set_params", None, [ctypes.c_void_p, _P(VgxParams)]),
    ("orc_sim_reset", None, [ctypes.c_void_p]),
    ("orc_sim_advance", _c, [ctypes.c_void_p, _fl]),
    ("orc_sim_field", _f, [ctypes.c_void_p, _c]),
    ("orc_sim_random_state", ctypes.c_uint32, [ctypes.c_void_p]),
    ("orc_sim_last_projection", None, [ctypes.c_void_p, _c, _P(VgxProjectionResult)]),
    ("orc_sim_diagnostics", None, [ctypes.c_void_p, _P(VgxDiagnostics)]),
    ("orc_kinetic_energy", _fl, [_c, _fl, _f, _f, _f]),
    ("orc_max_velocity", _fl, [_c, _f, _f, _f]),
    ("orc_emit", ctypes.c_uint32, [_c, _fl, _P(VgxParams), _fl, ctypes.c_uint32, _f, _f, _f, _f]),
    ("orc_philox4x32_10", None, [_P(ctypes.c_uint32), _P(ctypes.c_uint32), _P(ctypes.c_uint32)]),
    ("orc_profile_exp", _fl, [_fl]),
    ("orc_noise_log", _fl, [_fl]),
    ("orc_noise_cos2pi", _fl, [_fl]),
    ("orc_noise_normal", _fl, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]),
    ("orc_add_velocity_noise", None, [_c, _fl, ctypes.c_uint64, _f, _f, _f]),
    ("orc_sim_add_velocity_noise", None, [ctypes.c_void_p, _fl, ctypes.c_uint64]),
]

_REF_SIGS = [
    ("ref_params_default", _c, [_P(VgxParams)]),
    ("ref_params_preset", _c, [_c, _P(VgxParams)]),
    ("ref_sim_create", _c, [_c, _fl, _P(VgxParams), _P(ctypes.c_void_p)]),
    ("ref_sim_create_from", _c, [_c, _fl, _P(VgxParams), _f, _f, _f, _f, _P(ctypes.c_void_p)]),
    ("ref_sim_create_algorithm", _c, [_c, _P(ctypes.c_void_p)]),
    ("ref_sim_destroy", None, [ctypes.c_void_p]),
    ("ref_sim_get_params", _c, [ctypes.c_void_p, _P(VgxParams)]),
    ("ref_sim_set_params", _c, [ctypes.c_void_p, _P(VgxParams)]),
    ("ref_sim_advance", _c, [ctypes.c_void_p, _fl]),
    ("ref_sim_reset", _c, [ctypes.c_void_p]),
    ("ref_sim_set_box", _c, [ctypes.c_void_p, _fl]),
    ("ref_sim_resolution", _c, [ctypes.c_void_p]),
    ("ref_sim_box", _fl, [ctypes.c_void_p]),
    ("ref_sim_download", _c, [ctypes.c_void_p, _c, _f]),
    ("ref_sim_last_projection", _c, [ctypes.c_void_p, _P(VgxProjectionResult)]),
    ("ref_sim_diagnostics", _c, [ctypes.c_void_p, _P(VgxDiagnostics)]),
    ("ref_sample", _fl, [_c, _fl, _c, _f, _fl, _fl, _fl]),
    ("ref_advect_face", _c, [_c, _fl, _c, _c, _fl, _fl, _f, _f, _f, _f, _f]),
    ("ref_advect_scalar", _c, [_c, _fl, _c, _fl, _fl, _f, _f, _f, _f]),
    ("ref_project", _c, [_c, _fl, _P(VgxProjectionOptions), _f, _f, _f, _f, _f, _P(VgxProjectionResult)]),
    ("ref_divergence", _c, [_c, _fl, _f, _f, _f, _f]),
    ("ref_max_divergence", _c, [_c, _fl, _f, _f, _f, _f]),
    ("ref_apply_gradient", _c, [_c, _fl, _f, _fl, _f, _f, _f]),
    ("ref_sor", _c, [_c, _fl, _fl, _f, _f, _P(VgxProjectionOptions), _P(VgxProjectionResult)]),
]


def _bind(path: Path, sigs):
    lib = ctypes.CDLL(str(path))
    for name, restype, argtypes in sigs:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


_ORACLE = None
_ORACLE_OMP = None
_REF = None


def oracle_lib(parallel: bool = False):
    """The serial checker, or (parallel) its OpenMP build for timing."""
    global _ORACLE, _ORACLE_OMP
    if parallel:
        if _ORACLE_OMP is None:
            if not ORACLE_OMP_LIB.exists():
                raise ImportError(f"oracle not built: {ORACLE_OMP_LIB} (run make -C oracle oracle)")
            _ORACLE_OMP = _bind(ORACLE_OMP_LIB, _ORACLE_SIGS)
        return _ORACLE_OMP
    if _ORACLE is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"oracle not built: {ORACLE_LIB} (run make -C oracle oracle)")
        _ORACLE = _bind(ORACLE_LIB, _ORACLE_SIGS)
    return _ORACLE


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib():
    global _REF
    if _REF is None:
        if not REF_LIB.exists():
            raise ImportError(f"reference build missing: {REF_LIB}")
        _REF = _bind(REF_LIB, _REF_SIGS)
    return _REF


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_f)


def sizes(n: int):
    return n * n * n, (n + 1) * n * n


def default_params() -> VgxParams:
    p = VgxParams()
    oracle_lib().orc_params_default(ctypes.byref(p))
    return p


def default_options() -> VgxProjectionOptions:
    o = VgxProjectionOptions()
    oracle_lib().orc_projection_options_default(ctypes.byref(o))
    return o


class OracleSim:
    """The C restatement of MacGridFluidSolver3D."""

    def __init__(self, n: int, box: float, params: VgxParams, parallel: bool = False):
        self.n, self.box = n, box
        self._lib = oracle_lib(parallel)
        self._h = self._lib.orc_sim_create(n, box, ctypes.byref(params))

    def close(self):
        if self._h:
            self._lib.orc_sim_destroy(self._h)
            self._h = None

    __del__ = close

    def set_params(self, params: VgxParams):
        self._lib.orc_sim_set_params(self._h, ctypes.byref(params))

    def reset(self):
        self._lib.orc_sim_reset(self._h)

    def advance(self, dt: float):
        if self._lib.orc_sim_advance(self._h, dt) != 0:
            raise ValueError("oracle rejected the step")

    def field(self, which: int) -> np.ndarray:
        cells, faces = sizes(self.n)
        count = cells if which in (0, 4) else faces
        p = self._lib.orc_sim_field(self._h, which)
        return np.ctypeslib.as_array(p, shape=(count,)).copy()

    def fields(self):
        return tuple(self.field(k) for k in range(4))

    def set_fields(self, rho=None, u=None, v=None, w=None, p=None):
        for k, a in enumerate((rho, u, v, w, p)):
            if a is not None:
                cells, faces = sizes(self.n)
                count = cells if k in (0, 4) else faces
                dst = np.ctypeslib.as_array(self._lib.orc_sim_field(self._h, k), shape=(count,))
                dst[:] = a

    def random_state(self) -> int:
        return self._lib.orc_sim_random_state(self._h)

    def add_velocity_noise(self, sigma: float, seed: int):
        self._lib.orc_sim_add_velocity_noise(self._h, sigma, seed)

    def last_projection(self, index: int = 0) -> VgxProjectionResult:
        r = VgxProjectionResult()
        self._lib.orc_sim_last_projection(self._h, index, ctypes.byref(r))
        return r

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        self._lib.orc_sim_diagnostics(self._h, ctypes.byref(d))
        return d


class RefSim:
    """vapor's own MacGridFluidSolver3D (oracle/_ref build)."""

    def __init__(self, n: int, box: float, params: VgxParams | None = None, fields=None):
        self.n, self.box = n, box
        self._lib = ref_lib()
        h = ctypes.c_void_p()
        if fields is None:
            rc = self._lib.ref_sim_create(n, box, ctypes.byref(params) if params is not None else None,
                                          ctypes.byref(h))
        else:
            rho, u, v, w = fields
            rc = self._lib.ref_sim_create_from(n, box, ctypes.byref(params) if params is not None else None,
                                               ptr(rho), ptr(u), ptr(v), ptr(w), ctypes.byref(h))
        if rc != 0:
            raise ValueError(f"reference rejected construction (code {rc})")
        self._h = h.value

    def close(self):
        if self._h:
            self._lib.ref_sim_destroy(self._h)
            self._h = None

    __del__ = close

    def params(self) -> VgxParams:
        p = VgxParams()
        self._lib.ref_sim_get_params(self._h, ctypes.byref(p))
        return p

    def set_params(self, params: VgxParams):
        self._lib.ref_sim_set_params(self._h, ctypes.byref(params))

    def advance(self, dt: float):
        rc = self._lib.ref_sim_advance(self._h, dt)
        if rc != 0:
            raise ValueError(f"reference rejected the step (code {rc})")

    def reset(self):
        self._lib.ref_sim_reset(self._h)

    def field(self, which: int) -> np.ndarray:
        cells, faces = sizes(self.n)
        out = np.empty(cells if which == 0 else faces, np.float32)
        self._lib.ref_sim_download(self._h, which, ptr(out))
        return out

    def fields(self):
        return tuple(self.field(k) for k in range(4))

    def last_projection(self) -> VgxProjectionResult:
        r = VgxProjectionResult()
        self._lib.ref_sim_last_projection(self._h, ctypes.byref(r))
        return r

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        rc = self._lib.ref_sim_diagnostics(self._h, ctypes.byref(d))
        if rc != 0:
            raise ValueError("reference diagnostics rejected a non-finite state")
        return d
