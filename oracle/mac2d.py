"""Python bindings of the MAC 2D checkers (TEST INFRASTRUCTURE ONLY).

* ``_build/liboracle_mac2d.so`` - the C restatement of both 2D solvers
  (mac2d_oracle.c), always built;
* ``_ref/libvapor_ref2d.so`` - vapor's own 2D solvers behind ref2d_capi.cpp,
  present only where the reference tree was available at build time.

Only tests/ and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import contextlib
import ctypes
from pathlib import Path

import numpy as np

from vapor_amd.abi import VgxDiagnostics, VgxProjectionOptions, VgxProjectionResult, Vgx2Params, Vgx2rParams

ORACLE_DIR = Path(__file__).resolve().parent
ORACLE2D_LIB = ORACLE_DIR / "_build" / "liboracle_mac2d.so"
REF2D_LIB = ORACLE_DIR / "_ref" / "libvapor_ref2d.so"

_P = ctypes.POINTER
_f = _P(ctypes.c_float)
_d = _P(ctypes.c_double)
_c = ctypes.c_int
_fl = ctypes.c_float
_vp = ctypes.c_void_p

_ORACLE_SIGS = [
    ("o2_set_libm", None, [_c]),
    ("o2_det_exp", ctypes.c_double, [ctypes.c_double]),
    ("o2_det_log", ctypes.c_double, [ctypes.c_double]),
    ("o2_expf", _fl, [_fl]),
    ("o2_powf", _fl, [_fl, _fl]),
    ("o2_advect", _c, [_c, _fl, _c, _fl, _f, _f, _f, _f]),
    ("o2_project", _c, [_c, _fl, _P(VgxProjectionOptions), _f, _f, _d, _P(VgxProjectionResult)]),
    ("o2_sim_create", _vp, [_c, _fl, _P(Vgx2Params)]),
    ("o2_sim_destroy", None, [_vp]),
    ("o2_sim_set_params", None, [_vp, _P(Vgx2Params)]),
    ("o2_sim_reset", None, [_vp]),
    ("o2_sim_advance", _c, [_vp, _fl]),
    ("o2_sim_field", _f, [_vp, _c]),
    ("o2_sim_pressure", _d, [_vp]),
    ("o2_sim_projection", None, [_vp, _c, _P(VgxProjectionResult)]),
    ("o2_sim_diagnostics", _c, [_vp, _P(VgxDiagnostics)]),
    ("o2h_create", _vp, [_c, _P(Vgx2rParams)]),
    ("o2h_destroy", None, [_vp]),
    ("o2h_set_params", None, [_vp, _P(Vgx2rParams)]),
    ("o2h_advance", _c, [_vp, _fl]),
    ("o2h_step", None, [_vp, _fl, _fl, _fl, _fl, _c]),
    ("o2h_emit", None, [_vp, _fl]),
    ("o2h_apply_boundary", None, [_vp]),
    ("o2h_advect_smoke", None, [_vp, _fl]),
    ("o2h_project", None, [_vp, _fl, _c]),
    ("o2h_reset", None, [_vp]),
    ("o2h_field", _f, [_vp, _c]),
    ("o2h_random_state", ctypes.c_uint32, [_vp]),
    ("o2h_set_random_state", None, [_vp, ctypes.c_uint32]),
    ("o2h_diagnostics", None, [_vp, _P(VgxDiagnostics)]),
]

_REF_SIGS = [
    ("r2_params_default", _c, [_P(Vgx2Params)]),
    ("r2_params_plume", _c, [_P(Vgx2Params)]),
    ("r2_sim_create", _c, [_c, _fl, _P(Vgx2Params), _f, _f, _f, _P(_vp)]),
    ("r2_sim_destroy", None, [_vp]),
    ("r2_sim_set_params", _c, [_vp, _P(Vgx2Params)]),
    ("r2_sim_advance", _c, [_vp, _fl]),
    ("r2_sim_reset", _c, [_vp]),
    ("r2_sim_set_box", _c, [_vp, _fl]),
    ("r2_sim_download", _c, [_vp, _f, _f, _f]),
    ("r2_sim_projection", _c, [_vp, _c, _P(VgxProjectionResult)]),
    ("r2_sim_diagnostics", _c, [_vp, _P(VgxDiagnostics)]),
    ("r2_advect", _c, [_c, _fl, _c, _fl, _f, _f, _f, _f]),
    ("r2_project", _c, [_c, _fl, _P(VgxProjectionOptions), _f, _f, _d, _P(VgxProjectionResult)]),
    ("r2h_params", _c, [_c, _c, _P(Vgx2rParams)]),
    ("r2h_create", _c, [_c, _P(Vgx2rParams), _P(_vp)]),
    ("r2h_destroy", None, [_vp]),
    ("r2h_set_params", _c, [_vp, _P(Vgx2rParams)]),
    ("r2h_advance", _c, [_vp, _fl]),
    ("r2h_step", _c, [_vp, _fl, _fl, _fl, _fl, _c]),
    ("r2h_emit", _c, [_vp, _fl, _fl]),
    ("r2h_apply_boundary", _c, [_vp]),
    ("r2h_advect_smoke", _c, [_vp, _fl]),
    ("r2h_project", _c, [_vp, _fl, _c]),
    ("r2h_reset", _c, [_vp]),
    ("r2h_get_field", _c, [_vp, _c, _f]),
    ("r2h_set_field", _c, [_vp, _c, _f]),
    ("r2h_random_state", ctypes.c_uint32, [_vp]),
    ("r2h_diagnostics", _c, [_vp, _P(VgxDiagnostics)]),
]

_ORACLE = None
_REF = None


def _bind(path: Path, sigs):
    lib = ctypes.CDLL(str(path))
    for name, restype, argtypes in sigs:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def oracle2d_lib():
    global _ORACLE
    if _ORACLE is None:
        if not ORACLE2D_LIB.exists():
            raise ImportError(f"oracle not built: {ORACLE2D_LIB} (run make -C oracle oracle)")
        _ORACLE = _bind(ORACLE2D_LIB, _ORACLE_SIGS)
    return _ORACLE


def ref2d_available() -> bool:
    return REF2D_LIB.exists()


def ref2d_lib():
    global _REF
    if _REF is None:
        if not REF2D_LIB.exists():
            raise ImportError(f"reference 2D build missing: {REF2D_LIB} (make -C oracle ref)")
        _REF = _bind(REF2D_LIB, _REF_SIGS)
    return _REF


@contextlib.contextmanager
def libm_mode(on: bool = True):
    """Within the block the oracle evaluates e^v and x^y with the C library
    (vapor's own functions) instead of the GPU's IEEE sequences."""
    lib = oracle2d_lib()
    lib.o2_set_libm(1 if on else 0)
    try:
        yield
    finally:
        lib.o2_set_libm(0)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_f)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_d)


def sizes(n: int):
    return n * n, n * (n + 1)


class OracleSim2D:
    """The restated MacGridFluidSolver2D."""

    def __init__(self, n: int, box: float, params: Vgx2Params):
        self.lib = oracle2d_lib()
        self.n = n
        self.h = self.lib.o2_sim_create(n, box, ctypes.byref(params))
        if not self.h:
            raise ValueError("invalid 2D domain")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.o2_sim_destroy(self.h)
            self.h = None

    def set_params(self, params: Vgx2Params) -> None:
        self.lib.o2_sim_set_params(self.h, ctypes.byref(params))

    def advance(self, dt: float) -> int:
        return self.lib.o2_sim_advance(self.h, dt)

    def reset(self) -> None:
        self.lib.o2_sim_reset(self.h)

    def fields(self):
        N, F = sizes(self.n)
        return tuple(np.ctypeslib.as_array(self.lib.o2_sim_field(self.h, k), shape=(m,)).copy()
                     for k, m in ((0, N), (1, F), (2, F)))

    def set_fields(self, rho=None, u=None, v=None) -> None:
        N, F = sizes(self.n)
        for k, m, a in ((0, N, rho), (1, F, u), (2, F, v)):
            if a is not None:
                np.ctypeslib.as_array(self.lib.o2_sim_field(self.h, k), shape=(m,))[:] = a

    def pressure(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.lib.o2_sim_pressure(self.h), shape=(self.n * self.n,)).copy()

    def projection(self, midpoint: bool = False) -> VgxProjectionResult:
        r = VgxProjectionResult()
        self.lib.o2_sim_projection(self.h, 1 if midpoint else 0, ctypes.byref(r))
        return r

    def diagnostics(self):
        d = VgxDiagnostics()
        rc = self.lib.o2_sim_diagnostics(self.h, ctypes.byref(d))
        return rc, d


class OracleHist2D:
    """The restated ReflectionMacFluidSolver2D."""

    FIELDS = ("smoke", "vx", "vy", "pressure", "divergence")

    def __init__(self, n: int, params: Vgx2rParams):
        self.lib = oracle2d_lib()
        self.n = n
        self.h = self.lib.o2h_create(n, ctypes.byref(params))
        if not self.h:
            raise ValueError("2D reflection resolution must be 2..512")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.o2h_destroy(self.h)
            self.h = None

    def set_params(self, params: Vgx2rParams) -> None:
        self.lib.o2h_set_params(self.h, ctypes.byref(params))

    def _size(self, k: int) -> int:
        N, F = sizes(self.n)
        return F if k in (1, 2) else N

    def fields(self) -> dict:
        return {name: np.ctypeslib.as_array(self.lib.o2h_field(self.h, k), shape=(self._size(k),)).copy()
                for k, name in enumerate(self.FIELDS)}

    def set_field(self, name: str, values: np.ndarray) -> None:
        k = self.FIELDS.index(name)
        np.ctypeslib.as_array(self.lib.o2h_field(self.h, k), shape=(self._size(k),))[:] = values

    def __getattr__(self, name):
        # advance, step, emit, apply_boundary, advect_smoke, project, reset,
        # random_state, set_random_state pass straight through.
        fn = getattr(self.__dict__["lib"], "o2h_" + name)
        return lambda *args: fn(self.__dict__["h"], *args)

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        self.lib.o2h_diagnostics(self.h, ctypes.byref(d))
        return d


class RefSim2D:
    """vapor's own MacGridFluidSolver2D (reference build)."""

    def __init__(self, n: int, box: float, params: Vgx2Params | None, rho=None, u=None, v=None):
        self.lib = ref2d_lib()
        self.n = n
        h = ctypes.c_void_p()
        p = ctypes.byref(params) if params is not None else None
        rc = self.lib.r2_sim_create(n, box, p, fptr(rho) if rho is not None else None,
                                    fptr(u) if u is not None else None, fptr(v) if v is not None else None,
                                    ctypes.byref(h))
        if rc:
            raise ValueError(f"r2_sim_create failed ({rc})")
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.r2_sim_destroy(self.h)
            self.h = None

    def set_params(self, params: Vgx2Params) -> int:
        return self.lib.r2_sim_set_params(self.h, ctypes.byref(params))

    def advance(self, dt: float) -> int:
        return self.lib.r2_sim_advance(self.h, dt)

    def fields(self):
        N, F = sizes(self.n)
        rho, u, v = np.empty(N, np.float32), np.empty(F, np.float32), np.empty(F, np.float32)
        self.lib.r2_sim_download(self.h, fptr(rho), fptr(u), fptr(v))
        return rho, u, v

    def projection(self, midpoint: bool = False) -> VgxProjectionResult:
        r = VgxProjectionResult()
        self.lib.r2_sim_projection(self.h, 1 if midpoint else 0, ctypes.byref(r))
        return r

    def diagnostics(self):
        d = VgxDiagnostics()
        rc = self.lib.r2_sim_diagnostics(self.h, ctypes.byref(d))
        return rc, d


class RefHist2D:
    """vapor's own ReflectionMacFluidSolver2D (reference build)."""

    FIELDS = OracleHist2D.FIELDS

    def __init__(self, n: int, params: Vgx2rParams | None = None):
        self.lib = ref2d_lib()
        self.n = n
        h = ctypes.c_void_p()
        rc = self.lib.r2h_create(n, ctypes.byref(params) if params is not None else None, ctypes.byref(h))
        if rc:
            raise ValueError(f"r2h_create failed ({rc})")
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.r2h_destroy(self.h)
            self.h = None

    def _size(self, k: int) -> int:
        N, F = sizes(self.n)
        return F if k in (1, 2) else N

    def fields(self) -> dict:
        out = {}
        for k, name in enumerate(self.FIELDS):
            a = np.empty(self._size(k), np.float32)
            self.lib.r2h_get_field(self.h, k, fptr(a))
            out[name] = a
        return out

    def set_field(self, name: str, values: np.ndarray) -> None:
        k = self.FIELDS.index(name)
        self.lib.r2h_set_field(self.h, k, fptr(np.ascontiguousarray(values, np.float32)))

    def __getattr__(self, name):
        fn = getattr(self.__dict__["lib"], "r2h_" + name)
        return lambda *args: fn(self.__dict__["h"], *args)

    def diagnostics(self) -> VgxDiagnostics:
        d = VgxDiagnostics()
        self.lib.r2h_diagnostics(self.h, ctypes.byref(d))
        return d


def hist_params(n: int = 128, smoke_method: int | None = None) -> Vgx2rParams:
    """The reference's own parameter defaults (member defaults for
    ``smoke_method=None``), read from the oracle-independent C ABI defaults."""
    from vapor_amd.abi import load_library  # the C ABI's defaults; no device work

    lib = load_library()
    p = Vgx2rParams()
    if smoke_method is None:
        lib.vgx2r_params_default(ctypes.byref(p))
    else:
        lib.vgx2r_params_plume(int(smoke_method), ctypes.byref(p))
    return p
