"""Argument types of the entry points declared in include/clmsoil_ext.h (no
counterpart in CLM.jl's interfaces)."""
from __future__ import annotations

import ctypes as C

from ._abi import HeatArgs, HeatConfig, Levels


def declare(lib) -> None:
    P, I = C.c_void_p, C.c_int
    lib.clmsoil_device_errors.argtypes = [P, C.POINTER(C.c_uint32)]
    lib.clmsoil_device_errors.restype = I
    lib.clmsoil_workspace_errors.argtypes = [P, P, C.POINTER(C.c_uint32)]
    lib.clmsoil_workspace_errors.restype = I
    lib.clmsoil_set_tuning.argtypes = [I, I, I, I]
    lib.clmsoil_set_tuning.restype = I
    lib.clmsoil_get_tuning.argtypes = [C.POINTER(I)] * 4
    lib.clmsoil_get_tuning.restype = I
    lib.clmsoil_set_richards_schedule.argtypes = [I]
    lib.clmsoil_set_richards_schedule.restype = I
    lib.clmsoil_set_richards_residency.argtypes = [I]
    lib.clmsoil_set_richards_residency.restype = I
    lib.clmsoil_set_richards_refill.argtypes = [I, I]
    lib.clmsoil_set_richards_refill.restype = I
    lib.clmsoil_set_richards_lds_cache.argtypes = [I]
    lib.clmsoil_set_richards_lds_cache.restype = I
    lib.clmsoil_set_zd_block.argtypes = [I]
    lib.clmsoil_set_zd_block.restype = I
    lib.clmsoil_math_eval.argtypes = [I, P, P, P, C.c_int64, P]
    lib.clmsoil_math_eval.restype = I
    _declare_learned(lib)
    _declare_adjoint(lib)
    _declare_day(lib)


def _declare_learned(lib) -> None:
    """include/clmsoil_learned.h"""
    P, I, L, F, D = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
    lib.clmsoil_richards_features.argtypes = [P, P, D, P, P, L, P]
    lib.clmsoil_richards_features.restype = I
    lib.clmsoil_mlp_predict_dtsub0.argtypes = [P, P, D, P, P, P, P]
    lib.clmsoil_mlp_predict_dtsub0.restype = I
    lib.clmsoil_soilwm_moisture_form_solve_learned.argtypes = [P, P, D, P, P, P, P]
    lib.clmsoil_soilwm_moisture_form_solve_learned.restype = I
    lib.clmsoil_mlp_forward.argtypes = [P, P, L, L, P, P]
    lib.clmsoil_mlp_forward.restype = I
    lib.clmsoil_mlp_grad.argtypes = [P, P, P, L, L, L, P, P, P]
    lib.clmsoil_mlp_grad.restype = I
    lib.clmsoil_mlp_adam.argtypes = [P, P, P, P, L, L, F, F, F, F, P, P]
    lib.clmsoil_mlp_adam.restype = I


def _declare_adjoint(lib) -> None:
    """include/clmsoil_adjoint.h"""
    P, I, L, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
    lib.clmsoil_adjoint_workspace_bytes.argtypes = [L]
    lib.clmsoil_adjoint_workspace_bytes.restype = L
    lib.clmsoil_soilwm_moisture_form_adjoint.argtypes = [P, P, D, P, P, P, P, P, P, P, P, L, P]
    lib.clmsoil_soilwm_moisture_form_adjoint.restype = I
    lib.clmsoil_soil_temperature_adjoint.argtypes = [P, P, D] + [P] * 7
    lib.clmsoil_soil_temperature_adjoint.restype = I
    lib.clmsoil_tridiagonal_multi_vjp.argtypes = [P] * 7 + [L, L, L] + [P] * 5
    lib.clmsoil_tridiagonal_multi_vjp.restype = I
    lib.clmsoil_batched_band_solve_vjp.argtypes = [P] * 6 + [I, I, I, L, L, L, P, P, P]
    lib.clmsoil_batched_band_solve_vjp.restype = I


def _declare_day(lib) -> None:
    """include/clmsoil_day.h"""
    P, I, D = C.c_void_p, C.c_int, C.c_double
    lib.clmsoil_day_plan_create.argtypes = [P, P, P, D, P, P, P, P, P, C.c_int32, C.POINTER(P)]
    lib.clmsoil_day_plan_create.restype = I
    lib.clmsoil_day_plan_launch.argtypes = [P, P]
    lib.clmsoil_day_plan_launch.restype = I
    lib.clmsoil_day_plan_destroy.argtypes = [P]
    lib.clmsoil_day_plan_destroy.restype = I
