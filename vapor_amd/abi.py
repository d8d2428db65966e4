"""ctypes mirror of include/vapor_mac3d.h (structs, constants, library loading).

The layouts below must match the C header byte for byte; the loaded library
reports its own struct sizes through ``vgx_struct_size`` and ``load_library``
refuses a library whose sizes differ.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

VGX_OK, VGX_EINVAL, VGX_ERANGE, VGX_EDEVICE = 0, 1, 2, 3
VGX_PRECISION_FP32, VGX_PRECISION_FP64 = 32, 64

# MacAlgorithm (src/common/MacAlgorithm.hpp:3-9)
SIMPLE_SEMI_LAGRANGIAN, SIMPLE_MACCORMACK, REFLECTION_SEMI_LAGRANGIAN, REFLECTION_MACCORMACK = 0, 1, 2, 3
# AdvectionMethod
SEMI_LAGRANGIAN, MACCORMACK = 0, 1
# FaceAxis
AXIS_X, AXIS_Y, AXIS_Z = 0, 1, 2
# extensions
EMITTER_SPHERE, EMITTER_GAUSSIAN = 0, 1
SOLVE_FIXED, SOLVE_TOLERANCE = 0, 1
WARM_ZERO, WARM_PREVIOUS = 0, 1


class VgxParams(ctypes.Structure):
    """MacGridParameters3D (MacGridFluidSolver3D.hpp:14-36) + extension fields."""

    _fields_ = [
        ("emitter_radius", ctypes.c_float),
        ("emitter_center_x", ctypes.c_float),
        ("emitter_center_y", ctypes.c_float),
        ("emitter_center_z", ctypes.c_float),
        ("emitter_kick_multiplier", ctypes.c_float),
        ("emitter_kick_base", ctypes.c_float),
        ("emitter_kick_scale", ctypes.c_float),
        ("stir_strength", ctypes.c_float),
        ("random_seed", ctypes.c_uint32),
        ("sor_omega", ctypes.c_float),
        ("buoyancy_split", ctypes.c_float),
        ("cfl_mc", ctypes.c_float),
        ("maccormack_smoke", ctypes.c_int32),
        ("maccormack_vel", ctypes.c_int32),
        ("smoke_decay", ctypes.c_float),
        ("reflection", ctypes.c_int32),
        ("buoyancy", ctypes.c_float),
        ("source_strength", ctypes.c_float),
        ("project_iterations", ctypes.c_int32),
        ("emitter_profile", ctypes.c_int32),
        ("emitter2_enabled", ctypes.c_int32),
        ("emitter2_radius", ctypes.c_float),
        ("emitter2_center_x", ctypes.c_float),
        ("emitter2_center_y", ctypes.c_float),
        ("emitter2_center_z", ctypes.c_float),
        ("solve_mode", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("launch_mode", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("max_sweeps", ctypes.c_int32),
        ("check_interval", ctypes.c_int32),
        ("refinement_steps", ctypes.c_int32),
        ("final_check_interval", ctypes.c_int32),
    ]

    def copy(self) -> "VgxParams":
        out = VgxParams()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(self))
        return out

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("reserved")}

    def update(self, **values) -> "VgxParams":
        for key, value in values.items():
            if not hasattr(self, key):
                raise AttributeError(f"vgx_params has no field {key!r}")
            setattr(self, key, value)
        return self


class VgxProjectionOptions(ctypes.Structure):
    """ProjectionOptions + LinearSolveOptions (ProjectionOptions.hpp:5-9)."""

    _fields_ = [
        ("dt", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("max_iterations", ctypes.c_int32),
        ("residual_check_interval", ctypes.c_int32),
        ("absolute_tolerance", ctypes.c_double),
        ("relative_tolerance", ctypes.c_double),
        ("fixed_iterations", ctypes.c_int32),
        ("diagonal_floor", ctypes.c_float),
        ("use_l2", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("l2_tolerance", ctypes.c_double),
        ("refinement_steps", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
    ]


class VgxProjectionResult(ctypes.Structure):
    """ProjectionResult + LinearSolveResult (ProjectionOptions.hpp:11-17)."""

    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("initial_residual", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("residual_available", ctypes.c_int32),
        ("divergence_before", ctypes.c_float),
        ("divergence_after", ctypes.c_float),
        ("learned_start", ctypes.c_int32),
        ("removed_rhs_mean", ctypes.c_double),
        ("initial_residual_l2", ctypes.c_double),
        ("final_residual_l2", ctypes.c_double),
        ("residual_checks", ctypes.c_int32),
        ("refinement_steps", ctypes.c_int32),
        ("warm_started", ctypes.c_int32),
        ("reserved1", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("reserved")}


class VgxDiagnostics(ctypes.Structure):
    """FluidDiagnostics (FluidDiagnostics.hpp:3-7)."""

    _fields_ = [
        ("kinetic_energy", ctypes.c_float),
        ("max_velocity", ctypes.c_float),
        ("max_divergence", ctypes.c_float),
    ]


class VgxBatchTuning(ctypes.Structure):
    """The batch autotuner's choices and candidate timings (vgx_batch_tuning)."""

    _fields_ = [
        ("elementwise_per_thread_log2", ctypes.c_int32),
        ("advect_grid", ctypes.c_int32),
        ("sweep_grid", ctypes.c_int32),
        ("stencil_grid", ctypes.c_int32),
        ("elementwise_items", ctypes.c_int32),
        ("tile_items", ctypes.c_int32),
        ("stencil_items", ctypes.c_int32),
        ("sweep_register", ctypes.c_int32),
        ("advect_ms", ctypes.c_double * 12),
        ("sweep_ms", ctypes.c_double * 4),
        ("stencil_ms", ctypes.c_double * 4),
        ("sweep_lds_ms", ctypes.c_double * 4),
    ]


class VgxFieldError(ctypes.Structure):
    """vgx_field_error: a field's error against a shadow run's."""

    _fields_ = [
        ("max_abs", ctypes.c_double),
        ("rms", ctypes.c_double),
        ("ref_max", ctypes.c_double),
        ("ref_rms", ctypes.c_double),
    ]


STRUCTS = (VgxParams, VgxProjectionOptions, VgxProjectionResult, VgxDiagnostics, VgxBatchTuning, VgxFieldError)

# MacField2D (SemiLagrangian2D.hpp:4)
DENSITY_2D, VELOCITY_X_2D, VELOCITY_Y_2D = 0, 1, 2


class Vgx2Params(ctypes.Structure):
    """MacGridParameters2D (MacGridFluidSolver2D.hpp:6-16): vgx2_params."""

    _fields_ = [
        ("reflection", ctypes.c_int32),
        ("source_strength", ctypes.c_float),
        ("emitter_kick_multiplier", ctypes.c_float),
        ("emitter_radius", ctypes.c_float),
        ("emitter_center_x", ctypes.c_float),
        ("emitter_center_y", ctypes.c_float),
        ("buoyancy", ctypes.c_float),
        ("smoke_decay", ctypes.c_float),
        ("sor_omega", ctypes.c_float),
        ("project_iterations", ctypes.c_int32),
    ]


class Vgx2rParams(ctypes.Structure):
    """ReflectionMacFluidSolver2D's public parameters: vgx2r_params."""

    _fields_ = [
        ("maccormack_smoke", ctypes.c_int32),
        ("source_strength", ctypes.c_float),
        ("buoyancy", ctypes.c_float),
        ("smoke_decay", ctypes.c_float),
        ("project_iterations", ctypes.c_int32),
        ("emitter_radius", ctypes.c_float),
        ("emitter_falloff", ctypes.c_float),
        ("emitter_center_x", ctypes.c_float),
        ("emitter_center_y", ctypes.c_float),
        ("emitter_smoke_scale", ctypes.c_float),
        ("emitter_kick_multiplier", ctypes.c_float),
        ("emitter_kick_base", ctypes.c_float),
        ("emitter_kick_scale", ctypes.c_float),
        ("emitter_height", ctypes.c_int32),
        ("sor_omega", ctypes.c_float),
        ("buoyancy_split", ctypes.c_float),
        ("buoyancy_exp", ctypes.c_float),
        ("stir_strength", ctypes.c_float),
        ("random_seed", ctypes.c_uint32),
    ]


STRUCTS_2D = (Vgx2Params, Vgx2rParams)

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "_lib" / "libvapor_mac3d.so"

_P = ctypes.POINTER
_f32p = _P(ctypes.c_float)

# (name, restype, argtypes) of every entry point declared in include/vapor_mac3d.h
# Installed-stage callbacks (vgx_advection_hook / vgx_projection_hook).
TARGET_DENSITY = 3
AdvectionHook = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_float, _f32p, _f32p, _f32p, _f32p, _f32p)
ProjectionHook = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _P(VgxProjectionOptions), _f32p, _f32p, _f32p,
                                  _f32p, _P(VgxProjectionResult))

SIGNATURES = [
    ("vgx_last_error", ctypes.c_char_p, []),
    ("vgx_struct_size", ctypes.c_size_t, [ctypes.c_int]),
    ("vgx_device_name", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    ("vgx_params_default", None, [_P(VgxParams)]),
    ("vgx_params_preset", ctypes.c_int, [ctypes.c_int, _P(VgxParams)]),
    ("vgx_projection_options_default", None, [_P(VgxProjectionOptions)]),
    ("vgx_sim_create", ctypes.c_int, [ctypes.c_int, ctypes.c_float, _P(VgxParams), _P(ctypes.c_void_p)]),
    ("vgx_sim_create_algorithm", ctypes.c_int, [ctypes.c_int, _P(ctypes.c_void_p)]),
    ("vgx_sim_destroy", None, [ctypes.c_void_p]),
    ("vgx_sim_get_params", ctypes.c_int, [ctypes.c_void_p, _P(VgxParams)]),
    ("vgx_sim_set_params", ctypes.c_int, [ctypes.c_void_p, _P(VgxParams)]),
    ("vgx_sim_resolution", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_sim_box_size", ctypes.c_float, [ctypes.c_void_p]),
    ("vgx_sim_advance", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("vgx_sim_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_sim_set_box", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("vgx_sim_upload", ctypes.c_int, [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p]),
    ("vgx_sim_save_snapshot", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("vgx_sim_load_snapshot", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("vgx_sim_set_advection_hook", ctypes.c_int, [ctypes.c_void_p, AdvectionHook, ctypes.c_void_p]),
    ("vgx_sim_set_projection_hook", ctypes.c_int, [ctypes.c_void_p, ProjectionHook, ctypes.c_void_p]),
    ("vgx_sim_download", ctypes.c_int, [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p]),
    ("vgx_unet_create", ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("vgx_unet_destroy", None, [ctypes.c_void_p]),
    ("vgx_unet_param_count", ctypes.c_size_t, []),
    ("vgx_unet_get_params", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_unet_set_params", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_unet_set_grid", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("vgx_unet_get_grid", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_unet_get_grads", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_unet_set_grads", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_unet_predict", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_float, ctypes.c_float, _f32p]),
    ("vgx_unet_loss_grad", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_float, _f32p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]),
    ("vgx_unet_adam_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float]),
    ("vgx_dataset_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("vgx_dataset_destroy", None, [ctypes.c_void_p]),
    ("vgx_dataset_size", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_dataset_get", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _f32p, _f32p]),
    ("vgx_dataset_clear", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_unet_train", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    ("vgx_sim_record", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("vgx_sim_set_unet", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("vgx_sim_add_velocity_noise", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_uint64]),
    ("vgx_sim_download_pressure", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_sim_upload_pressure", ctypes.c_int, [ctypes.c_void_p, _f32p]),
    ("vgx_sim_transfer_bytes", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    ("vgx_sim_last_projection", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(VgxProjectionResult)]),
    ("vgx_sim_diagnostics", ctypes.c_int, [ctypes.c_void_p, _P(VgxDiagnostics)]),
    ("vgx_sim_device_fields", ctypes.c_int, [ctypes.c_void_p] + [_P(ctypes.c_void_p)] * 5),
    ("vgx_sim_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_sim_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("vgx_sim_reset_timing", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_kernel_kinds", ctypes.c_int, []),
    ("vgx_kernel_name", ctypes.c_char_p, [ctypes.c_int]),
    ("vgx_sim_kernel_timing", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_uint64)]),
    ("vgx_sim_sweep_count", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64)]),
    ("vgx_batch_create", ctypes.c_int,
     [ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_float), _P(VgxParams), _P(ctypes.c_void_p)]),
    ("vgx_batch_destroy", None, [ctypes.c_void_p]),
    ("vgx_batch_size", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_batch_scene", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_void_p)]),
    ("vgx_batch_advance", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("vgx_batch_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_batch_scene_latency", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double)]),
    ("vgx_batch_get_tuning", ctypes.c_int, [ctypes.c_void_p, _P(VgxBatchTuning)]),
    ("vgx_batch_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("vgx_batch_reset_timing", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_batch_kernel_timing", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_uint64)]),
    ("vgx_batch_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("vgx_shadow_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_void_p)]),
    ("vgx_shadow_destroy", None, [ctypes.c_void_p]),
    ("vgx_shadow_advance", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("vgx_shadow_download", ctypes.c_int, [ctypes.c_void_p] + [_P(ctypes.c_double)] * 5),
    ("vgx_shadow_last_projection", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(VgxProjectionResult)]),
    ("vgx_shadow_sweep_count", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64)]),
    ("vgx_shadow_compare", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(VgxFieldError)]),
    ("vgx_backend_create", ctypes.c_int, [ctypes.c_int, ctypes.c_float, _P(ctypes.c_void_p)]),
    ("vgx_backend_destroy", None, [ctypes.c_void_p]),
    ("vgx_backend_advect_scalar", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _f32p, _f32p]),
    ("vgx_backend_advect_face", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
      _f32p, _f32p, _f32p, _f32p, _f32p]),
    ("vgx_advect_face_noalloc", ctypes.c_int,
     [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("vgx_backend_project", ctypes.c_int,
     [ctypes.c_void_p, _P(VgxProjectionOptions), _f32p, _f32p, _f32p, _f32p, _P(VgxProjectionResult)]),
]


_vp = ctypes.c_void_p
_f64p = _P(ctypes.c_double)
_cf = _P(ctypes.POINTER(ctypes.c_float))

# (name, restype, argtypes) of every entry point declared in include/vapor_mac2d.h
SIGNATURES_2D = [
    ("vgx2_struct_size", ctypes.c_size_t, [ctypes.c_int]),
    ("vgx2_params_default", None, [_P(Vgx2Params)]),
    ("vgx2_params_plume", None, [_P(Vgx2Params)]),
    ("vgx2_sim_create", ctypes.c_int, [ctypes.c_int, ctypes.c_float, _P(Vgx2Params), _P(_vp)]),
    ("vgx2_sim_create_default", ctypes.c_int, [_P(_vp)]),
    ("vgx2_sim_destroy", None, [_vp]),
    ("vgx2_sim_get_params", ctypes.c_int, [_vp, _P(Vgx2Params)]),
    ("vgx2_sim_set_params", ctypes.c_int, [_vp, _P(Vgx2Params)]),
    ("vgx2_sim_resolution", ctypes.c_int, [_vp]),
    ("vgx2_sim_box_size", ctypes.c_float, [_vp]),
    ("vgx2_sim_advance", ctypes.c_int, [_vp, ctypes.c_float]),
    ("vgx2_sim_reset", ctypes.c_int, [_vp]),
    ("vgx2_sim_set_box", ctypes.c_int, [_vp, ctypes.c_float]),
    ("vgx2_sim_upload", ctypes.c_int, [_vp, _f32p, _f32p, _f32p]),
    ("vgx2_sim_download", ctypes.c_int, [_vp, _f32p, _f32p, _f32p]),
    ("vgx2_sim_last_projection", ctypes.c_int, [_vp, _P(VgxProjectionResult)]),
    ("vgx2_sim_midpoint_projection", ctypes.c_int, [_vp, _P(VgxProjectionResult)]),
    ("vgx2_sim_diagnostics", ctypes.c_int, [_vp, _P(VgxDiagnostics)]),
    ("vgx2_sim_device_fields", ctypes.c_int, [_vp, _P(_vp), _P(_vp), _P(_vp)]),
    ("vgx2_sim_synchronize", ctypes.c_int, [_vp]),
    ("vgx2_backend_advect", ctypes.c_int,
     [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, _f32p, _f32p, _f32p, _f32p]),
    ("vgx2_backend_project", ctypes.c_int,
     [ctypes.c_int, ctypes.c_float, _P(VgxProjectionOptions), _f32p, _f32p, _f64p, _P(VgxProjectionResult)]),
    ("vgx2r_params_default", None, [_P(Vgx2rParams)]),
    ("vgx2r_params_plume", ctypes.c_int, [ctypes.c_int, _P(Vgx2rParams)]),
    ("vgx2r_sim_create", ctypes.c_int, [ctypes.c_int, _P(Vgx2rParams), _P(_vp)]),
    ("vgx2r_sim_destroy", None, [_vp]),
    ("vgx2r_sim_get_params", ctypes.c_int, [_vp, _P(Vgx2rParams)]),
    ("vgx2r_sim_set_params", ctypes.c_int, [_vp, _P(Vgx2rParams)]),
    ("vgx2r_sim_resolution", ctypes.c_int, [_vp]),
    ("vgx2r_sim_advance", ctypes.c_int, [_vp, ctypes.c_float]),
    ("vgx2r_sim_step", ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_int]),
    ("vgx2r_sim_emit", ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_float]),
    ("vgx2r_sim_apply_boundary", ctypes.c_int, [_vp]),
    ("vgx2r_sim_advect_smoke", ctypes.c_int, [_vp, ctypes.c_float]),
    ("vgx2r_sim_project", ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_int]),
    ("vgx2r_sim_reset", ctypes.c_int, [_vp]),
    ("vgx2r_sim_set_box", ctypes.c_int, [_vp, ctypes.c_float]),
    ("vgx2r_sim_upload", ctypes.c_int, [_vp, _f32p, _f32p, _f32p, _f32p, _f32p]),
    ("vgx2r_sim_download", ctypes.c_int, [_vp, _f32p, _f32p, _f32p, _f32p, _f32p]),
    ("vgx2r_sim_random_state", ctypes.c_int, [_vp, _P(ctypes.c_uint32)]),
    ("vgx2r_sim_set_random_state", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("vgx2r_sim_diagnostics", ctypes.c_int, [_vp, _P(VgxDiagnostics)]),
    ("vgx2r_sim_device_fields", ctypes.c_int, [_vp, _P(_vp), _P(_vp), _P(_vp)]),
]


class VaporError(RuntimeError):
    """Device or library failure (the reference's std::runtime_error)."""


def check_status(lib, status: int) -> None:
    """Raise the exception type the reference would throw for a status code."""
    if status == VGX_OK:
        return
    message = lib.vgx_last_error().decode(errors="replace")
    if status == VGX_EINVAL:
        raise ValueError(message)
    if status == VGX_ERANGE:
        raise OverflowError(message)
    raise VaporError(message)


_LIB = None


def load_library(path: os.PathLike | None = None):
    """Load the HIP library. Raises if it is missing: there is no CPU fallback."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    lib_path = Path(path) if path is not None else LIBRARY_PATH
    if not lib_path.exists():
        raise ImportError(
            f"vapor_amd HIP library not built: {lib_path} is missing "
            "(run __graft_entry__.build() or python -m vapor_amd.build)")
    lib = ctypes.CDLL(str(lib_path))
    for name, restype, argtypes in SIGNATURES + SIGNATURES_2D:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    for size_of, structs in ((lib.vgx_struct_size, STRUCTS), (lib.vgx2_struct_size, STRUCTS_2D)):
        for index, struct in enumerate(structs):
            if size_of(index) != ctypes.sizeof(struct):
                raise ImportError(f"{struct.__name__} layout differs from the library "
                                  f"({ctypes.sizeof(struct)} vs {size_of(index)})")
    if path is None:
        _LIB = lib
    return lib


def float_ptr(array):
    """Pointer to a C-contiguous float32 numpy array (or None)."""
    if array is None:
        return None
    import numpy as np

    if array.dtype != np.float32 or not array.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous float32 array")
    return array.ctypes.data_as(_f32p)
