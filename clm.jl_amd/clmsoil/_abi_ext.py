"""Argument types of the entry points declared in include/clmsoil_ext.h (no
counterpart in CLM.jl's interfaces)."""
from __future__ import annotations

import ctypes as C

from ._abi import HeatArgs, HeatConfig, Levels


def declare(lib) -> None:
    P, I = C.c_void_p, C.c_int
    lib.clmsoil_device_errors.argtypes = [P, C.POINTER(C.c_uint32)]
    lib.clmsoil_device_errors.restype = I
    lib.clmsoil_set_tuning.argtypes = [I, I, I, I]
    lib.clmsoil_set_tuning.restype = I
    lib.clmsoil_get_tuning.argtypes = [C.POINTER(I)] * 4
    lib.clmsoil_get_tuning.restype = I
    lib.clmsoil_set_richards_schedule.argtypes = [I]
    lib.clmsoil_set_richards_schedule.restype = I
