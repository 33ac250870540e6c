"""Reverse-mode entry points of the path (include/clmsoil_adjoint.h).

CLM.jl differentiates the soil solves with Enzyme, with hand-written rules for
the solvers (src/infrastructure/enzyme_rules.jl:72-107 band_solve!,
:152-202 tridiagonal_solve!). These functions are the library's hand-written
adjoints of the same operations. Like Enzyme shadows, every gradient argument
accumulates (+=); the incoming adjoint of the solution is consumed.
"""
from __future__ import annotations

import ctypes as C

from . import _abi
from .api import MfIn, MfState, SoilWaterMovementConfig, _mf_dict, _stream, MOISTURE_FORM, CLMSoilError
from .bind import byref, heat_args, levels, mf_args

MAX_SUBSTEPS = 64


def _lib():
    return _abi.load_library()


def _p(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def workspace_bytes(ld: int) -> int:
    return int(_lib().clmsoil_adjoint_workspace_bytes(int(ld)))


def new_workspace(ld: int, device="cuda"):
    import torch
    return torch.empty(workspace_bytes(ld) // 8, dtype=torch.float64, device=device)


def soilwm_moisture_form_adjoint(st: MfState, mfin: MfIn, mask, nbedrock, nlevsno: int,
                                 nlevsoi: int, nlevgrnd: int, dtime: float,
                                 cfg: SoilWaterMovementConfig, liq_bar, qcharge_bar, workspace,
                                 grad_bsw=None, grad_hksat=None, grad_watsat=None,
                                 grad_scale=None) -> None:
    """Reverse of one moisture-form step from the state `st` at its start.

    liq_bar (shape of h2osoi_liq) holds dL/dliq after the step on entry and
    dL/dliq before it on return; qcharge_bar [nc] is dL/dqcharge. grad_scale
    [3, nc] accumulates sum_j p_j dL/dp_j for p = bsw, hksat, watsat."""
    if cfg.lower_boundary_condition not in (2, 1):
        raise CLMSoilError("the adjoint supports BC_FLUX and BC_ZERO_FLUX lower boundaries")
    c = cfg.to_abi()
    c.soilwater_movement_method = MOISTURE_FORM
    d = _mf_dict(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd)
    if d["nc"] == 0:
        return
    L = _lib()
    a, lev = mf_args(d), levels(d)
    nbytes = workspace.numel() * workspace.element_size()
    _abi.check(L, L.clmsoil_soilwm_moisture_form_adjoint(
        byref(a), byref(lev), float(dtime), byref(c), _p(liq_bar), _p(qcharge_bar),
        _p(grad_bsw), _p(grad_hksat), _p(grad_watsat), _p(grad_scale), _p(workspace), nbytes,
        _stream()), "soilwm_moisture_form_adjoint")


def soil_temperature_adjoint(state: dict, dtime: float, t_bar, liq_bar, ice_bar,
                             nlevsno: int = 12, nlevgrnd: int = 25, nlevsoi: int = 20,
                             grad_bsw=None, grad_watsat=None, grad_scale=None) -> None:
    """Reverse of one heat step (api.soil_temperature) from the state at its
    start, on soil columns without snow, surface water or excess ice.

    t_bar, liq_bar, ice_bar (shape of t_soisno) hold dL/d(t_soisno, h2osoi_liq,
    h2osoi_ice) after the step on entry and before it on return. grad_scale
    [3, nc] accumulates sum_j p_j dL/dp_j in rows 0 (bsw) and 2 (watsat)."""
    d = dict(state)
    d.setdefault("nc", d["snl"].shape[0])
    d.update(nlevsno=nlevsno, nlevgrnd=nlevgrnd, nlevsoi=nlevsoi)
    if d["nc"] == 0:
        return
    L = _lib()
    a, lev = heat_args(d), levels(d)
    _abi.check(L, L.clmsoil_soil_temperature_adjoint(
        byref(a), byref(lev), float(dtime), _p(t_bar), _p(liq_bar), _p(ice_bar), _p(grad_bsw),
        _p(grad_watsat), _p(grad_scale), _stream()), "soil_temperature_adjoint")


def tridiagonal_multi_vjp(ubar, u, a, b, c, jtop, mask, ncols: int, nlevs: int,
                          abar=None, bbar=None, cbar=None, rbar=None) -> None:
    """Reverse of tridiagonal_multi! (tridiagonal.jl:129-142)."""
    L = _lib()
    _abi.check(L, L.clmsoil_tridiagonal_multi_vjp(
        _p(ubar), _p(u), _p(a), _p(b), _p(c), _p(jtop), _p(mask), int(ncols), int(nlevs),
        int(u.shape[1]), _p(abar), _p(bbar), _p(cbar), _p(rbar), _stream()),
        "tridiagonal_multi_vjp")


def batched_band_solve_vjp(tbar, tvector, bmatrix, jtop, jbot, mask, kl: int, ku: int,
                           nlevsno_off: int, bmatrix_bar=None, rvector_bar=None) -> None:
    """Reverse of batched_band_solve! (band_diagonal.jl:179-189)."""
    nlev, nc = tvector.shape
    L = _lib()
    _abi.check(L, L.clmsoil_batched_band_solve_vjp(
        _p(tbar), _p(tvector), _p(bmatrix), _p(jtop), _p(jbot), _p(mask), int(kl), int(ku),
        int(nlevsno_off), int(nc), int(nlev), int(nc), _p(bmatrix_bar), _p(rvector_bar),
        _stream()), "batched_band_solve_vjp")
