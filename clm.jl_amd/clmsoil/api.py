"""Host-side mirror of CLM.jl's interfaces for the soil solves.

Function names, argument meaning and error behaviour follow the reference:

  SoilWaterMovementConfig, init_soilwater_movement, use_aquifer_layer
      src/biogeophys/soil_water_movement.jl:654-703
  soilwm_moisture_form_solve      soil_water_movement.jl:1498-1517
  soilwater_moisture_form         soil_water_movement.jl:1532-1601
  soil_water                      soil_water_movement.jl:1780-1846
  soil_temperature                src/biogeophys/soil_temperature.jl:358-593
  tridiagonal_multi               src/infrastructure/tridiagonal.jl:129-142
  batched_band_solve              src/infrastructure/band_diagonal.jl:179-189

State arrays are torch tensors on the GPU in CLM.jl's memory layout: a Julia
`Matrix{Float64}` of size (nc, nlev) is a contiguous tensor of shape (nlev, nc)
(column index fastest); the 3-D `bmatrix[c, band, row]` is a (nlev, nband, nc)
tensor. Every call runs on torch's current stream and returns without
synchronizing. The compute is the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _abi
from .bind import (byref, fractions_args, ghf_args, heat_args, heat_config, levels, mf_args,
                   ch4_tran_args, ch4_tran_params, lake_args, lvt_args, lvt_params, rootsink_args,
                   seb_args, surfhyd_args, surfhyd_params, swm_config, theta_wt_args,
                   water_table_args, zd_args)

ZENGDECKER_2009, MOISTURE_FORM = _abi.ZENGDECKER_2009, _abi.MOISTURE_FORM
BC_HEAD, BC_FLUX, BC_ZERO_FLUX, BC_WATERTABLE, BC_AQUIFER = (
    _abi.BC_HEAD, _abi.BC_FLUX, _abi.BC_ZERO_FLUX, _abi.BC_WATERTABLE, _abi.BC_AQUIFER)


class CLMSoilError(RuntimeError):
    """Raised for invalid inputs (the reference raises `error(...)`)."""


@dataclass
class SoilWaterMovementConfig:
    """SoilWaterMovementConfig (soil_water_movement.jl:654-671)."""
    soilwater_movement_method: int = ZENGDECKER_2009
    upper_boundary_condition: int = BC_FLUX
    lower_boundary_condition: int = BC_AQUIFER
    dtmin: float = 60.0
    verySmall: float = 1.0e-8
    xTolerUpper: float = 1.0e-1
    xTolerLower: float = 1.0e-2
    expensive: int = 42
    inexpensive: int = 1
    flux_calculation: int = 1
    e_ice: float = 6.0

    def to_abi(self):
        return swm_config(lower_bc=self.lower_boundary_condition,
                          upper_bc=self.upper_boundary_condition,
                          method=self.soilwater_movement_method,
                          flux_calculation=self.flux_calculation, dtmin=self.dtmin,
                          verySmall=self.verySmall, xTolerUpper=self.xTolerUpper,
                          xTolerLower=self.xTolerLower, e_ice=self.e_ice)


def init_soilwater_movement(*, use_bedrock: bool = False, **kwargs) -> SoilWaterMovementConfig:
    """init_soilwater_movement (soil_water_movement.jl:681-698)."""
    cfg = SoilWaterMovementConfig(**kwargs)
    if (cfg.soilwater_movement_method == ZENGDECKER_2009 and
            cfg.lower_boundary_condition != BC_AQUIFER):
        raise CLMSoilError("init_soilwater_movement: ZD09 must use bc_aquifer lower boundary condition")
    if use_bedrock and cfg.lower_boundary_condition != BC_ZERO_FLUX:
        raise CLMSoilError("init_soilwater_movement: use_bedrock requires bc_zero_flux lower boundary condition")
    return cfg


def use_aquifer_layer(cfg: SoilWaterMovementConfig) -> bool:
    """use_aquifer_layer (soil_water_movement.jl:700-703)."""
    return cfg.lower_boundary_condition in (BC_AQUIFER, BC_WATERTABLE)


@dataclass
class MfState:
    """MfState (soil_water_movement.jl:1230-1233): read/write column state."""
    h2osoi_liq: object
    smp_l: object
    hk_l: object
    qcharge: object
    nsubsteps: object
    nrejected: object = None


@dataclass
class MfIn:
    """MfIn (soil_water_movement.jl:1234-1238): read-only inputs."""
    z: object
    zi: object
    dz: object
    qflx_rootsoi: object
    watsat: object
    hksat: object
    bsw: object
    sucsat: object
    icefrac: object
    zwt: object
    qflx_infl: object
    smpmin: object = None
    dtsub0: object = None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _lib():
    return _abi.load_library()


def richards_workspace(device=None):
    """A fresh workspace for one Richards call at a time (clmsoil_mf_args.workspace)."""
    import torch
    n = int(_lib().clmsoil_richards_workspace_bytes())
    return torch.empty(n, dtype=torch.uint8, device=device or "cuda")


def _mf_dict(st: MfState, mfin: MfIn, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd):
    nc = nbedrock.shape[0]
    d = dict(nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, mask=mask,
             nbedrock=nbedrock)
    d.update({k: getattr(st, k) for k in ("h2osoi_liq", "smp_l", "hk_l", "qcharge",
                                          "nsubsteps", "nrejected")})
    d.update({k: getattr(mfin, k) for k in ("z", "zi", "dz", "qflx_rootsoi", "watsat",
                                            "hksat", "bsw", "sucsat", "icefrac", "zwt",
                                            "qflx_infl", "smpmin", "dtsub0")})
    return d


def _check_moisture_form_bc(cfg: SoilWaterMovementConfig):
    # soil_water_movement.jl:1500-1505
    if cfg.upper_boundary_condition != BC_FLUX:
        raise CLMSoilError("soilwm_moisture_form_solve!: upper boundary condition must be BC_FLUX")
    if cfg.lower_boundary_condition not in (BC_WATERTABLE, BC_FLUX, BC_ZERO_FLUX):
        raise CLMSoilError("soilwm_moisture_form_solve!: unsupported lower boundary condition")


def soilwm_moisture_form_solve(st: MfState, mfin: MfIn, mask, nbedrock, joff: int, joff_zi: int,
                               nlevsoi: int, dtime: float, cfg: SoilWaterMovementConfig,
                               nlevgrnd: int | None = None, workspace=None) -> None:
    """One adaptive-substep moisture-form Richards step for every masked column.
    `workspace` (richards_workspace()) is the call's scratch; a fresh one is
    taken from torch's allocator when omitted (the reference allocates its
    scratch per call too, soil_water_movement.jl:1555-1561)."""
    _check_moisture_form_bc(cfg)
    if joff_zi != joff + 1:
        raise CLMSoilError("joff_zi must equal joff + 1 (zi has one more level)")
    c = cfg.to_abi()
    c.soilwater_movement_method = MOISTURE_FORM
    d = _mf_dict(st, mfin, mask, nbedrock, joff, nlevsoi, nlevgrnd or nlevsoi + 5)
    if d["nc"] == 0:
        return
    L = _lib()
    d["workspace"] = workspace if workspace is not None else richards_workspace(nbedrock.device)
    a, lev = mf_args(d), levels(d)
    _abi.check(L, L.clmsoil_soilwm_moisture_form_solve(byref(a), byref(lev), float(dtime), byref(c),
                                                       _stream()), "soilwm_moisture_form_solve")


def soilwater_moisture_form(st: MfState, mfin: MfIn, mask, nbedrock, nlevsno: int, nlevsoi: int,
                            nlevgrnd: int, dtime: float, cfg: SoilWaterMovementConfig,
                            workspace=None) -> None:
    """The solve followed by the aquifer-recharge pass when use_aquifer_layer(cfg)."""
    _check_moisture_form_bc(cfg)
    c = cfg.to_abi()
    c.soilwater_movement_method = MOISTURE_FORM
    d = _mf_dict(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd)
    if d["nc"] == 0:
        return
    L = _lib()
    d["workspace"] = workspace if workspace is not None else richards_workspace(nbedrock.device)
    a, lev = mf_args(d), levels(d)
    _abi.check(L, L.clmsoil_soilwater_moisture_form(byref(a), byref(lev), float(dtime), byref(c),
                                                    _stream()), "soilwater_moisture_form")


def soilwater_zengdecker2009(state: dict, dtime: float, cfg: SoilWaterMovementConfig,
                             nlevsno: int = 12, nlevsoi: int = 20, nlevgrnd: int = 25) -> None:
    """One Zeng-Decker 2009 step (soilwater_zengdecker2009!,
    soil_water_movement.jl:1616-1764) for every masked column.

    `state` maps the field names of clmsoil_zd_args (include/clmsoil_zd09.h) to
    float64 device tensors (`mask` uint8 or absent); h2osoi_liq is updated and
    icefrac, smp_l, hk_l (soil layers 1..nlevsoi), qcharge and qflx_deficit are
    written."""
    c = cfg.to_abi()
    if c.soilwater_movement_method != ZENGDECKER_2009:
        raise CLMSoilError("soilwater_zengdecker2009!: soilwater_movement_method must be ZENGDECKER_2009")
    d = dict(state)
    d.setdefault("nc", d["zwt"].shape[0])
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if d["nc"] == 0:
        return
    L = _lib()
    a, lev = zd_args(d), levels(d)
    _abi.check(L, L.clmsoil_soilwater_zengdecker2009(byref(a), byref(lev), float(dtime), byref(c),
                                                     _stream()), "soilwater_zengdecker2009")


def compute_ground_heat_flux_and_deriv(state: dict, nlevsno: int = 12, nlevgrnd: int = 25,
                                       building_temp_method: int = 0) -> None:
    """Ground heat flux and its temperature derivative per column
    (compute_ground_heat_flux_and_deriv!, soil_temperature.jl:1200-1263).

    `state` maps the field names of clmsoil_ghf_args (include/clmsoil_ghf.h) to
    device tensors (index arrays int64 and 1-based, masks uint8) plus the counts
    `nc`, `np`, `nl`; the patch and column outputs are written in place."""
    d = dict(state)
    d.update(nlevsno=nlevsno, nlevgrnd=nlevgrnd, nlevsoi=min(20, nlevgrnd),
             building_temp_method=building_temp_method)
    if int(d["nc"]) == 0 and int(d["np"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_ground_heat_flux(byref(ghf_args(d)), byref(levels(d)), _stream()),
               "ground_heat_flux")


def soil_energy_balance(state: dict, dtime: float, nlevsno: int = 12, nlevgrnd: int = 25,
                        nlevurb: int = 0) -> None:
    """errsoi_patch and errsoi_col of soil_fluxes! (soil_fluxes.jl:380-447, 499-525).

    `state` maps the field names of clmsoil_seb_args (include/clmsoil_ghf.h) to
    device tensors plus the counts `nc` and `np`."""
    d = dict(state)
    d.update(nlevsno=nlevsno, nlevgrnd=nlevgrnd, nlevsoi=min(20, nlevgrnd), nlevurb=nlevurb)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_soil_energy_balance(byref(seb_args(d)), byref(levels(d)), float(dtime),
                                                _stream()), "soil_energy_balance")


def set_soil_water_fractions(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                             nlevgrnd: int = 25) -> None:
    """eff_porosity and icefrac per soil layer (set_soil_water_fractions!,
    soil_hydrology.jl:394-415). `state` maps the clmsoil_fractions_args fields
    (include/clmsoil_hydro.h) to device tensors, plus `nc`."""
    d = dict(state)
    d.setdefault("nc", d["watsat"].shape[-1])
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_set_soil_water_fractions(byref(fractions_args(d)), byref(levels(d)),
                                                     _stream()), "set_soil_water_fractions")


def water_table(state: dict, dtime: float, nlevsno: int = 12, nlevsoi: int = 20,
                nlevgrnd: int = 25, aq_sp_yield_min: float = 0.01,
                recompute_frost_table: bool = True) -> None:
    """water_table! (soil_hydrology.jl:849-902): the step's aquifer recharge moves
    the water table; frost and perched water tables. `state` maps the
    clmsoil_water_table_args fields to device tensors, plus `nc`."""
    d = dict(state)
    d.setdefault("nc", d["zwt"].shape[0])
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_water_table(byref(water_table_args(d)), byref(levels(d)), float(dtime),
                                        float(aq_sp_yield_min), int(bool(recompute_frost_table)),
                                        _stream()), "water_table")


def surface_hydrology(state: dict, dtime: float, nlevsno: int = 12, nlevsoi: int = 20,
                      nlevgrnd: int = 25, **params) -> None:
    """The surface hydrology before the Richards step (clm_driver.jl:1939-2012):
    set_floodc!, saturated_excess_runoff!, set_qflx_inputs!,
    infiltration_excess_runoff!, route_infiltration_excess!, update_h2osfc!,
    infiltration!, total_surface_runoff!, update_urban_ponding!, in one launch.
    `state` maps the clmsoil_surfhyd_args fields (include/clmsoil_surfhyd.h) to
    device tensors, plus `nc`; `params` are the keywords of bind.surfhyd_params
    (steps=bind.SH_ALL for the whole sequence)."""
    d = dict(state)
    d.setdefault("nc", d["h2osfc"].shape[0])
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_surface_hydrology(byref(surfhyd_args(d)), byref(levels(d)),
                                              float(dtime), byref(surfhyd_params(**params)),
                                              _stream()), "surface_hydrology")


def litter_vert_transp(state: dict, dtime: float, **params) -> None:
    """litter_vert_transp! without the matrix-CN path (litter_vert_transp.jl:479-639):
    the vertical advection-diffusion of every pool of every tracer type given.
    `state` maps the clmsoil_lvt_args fields (include/clmsoil_lvt.h) to device
    tensors, `conc`/`source`/`tendency` to lists of [npools, nlevdecomp_full, nc]
    tensors, plus `nc`; `params` are the keywords of bind.lvt_params."""
    d = dict(state)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_litter_vert_transp(byref(lvt_args(d)), float(dtime),
                                               byref(lvt_params(**params)), _stream()),
               "litter_vert_transp")


def ch4_tran(state: dict, dtime: float, **params) -> None:
    """ch4_tran! (methane.jl:2005-2092): the CH4 and O2 reaction-diffusion step of one
    fraction (sat 0 unsaturated, 1 saturated; lake for the lake call). `state` maps
    the clmsoil_ch4_tran_args fields (include/clmsoil_ch4.h) to device tensors,
    [nlevsoi, nc] for the level arrays and [3, ldg] for c_atm, plus `nc` and `ldg`;
    `params` are the keywords of bind.ch4_tran_params (nlevsoi required)."""
    d = dict(state)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_ch4_tran(byref(ch4_tran_args(d)), float(dtime),
                                     byref(ch4_tran_params(**params)), _stream()), "ch4_tran")


def lake_temperature_solve(state: dict, dtime: float, nlevsno: int = 12, nlevsoi: int = 20,
                           nlevgrnd: int = 25) -> None:
    """Steps 6-7 of lake_temperature! and the unpack (lake_temperature.jl:1205-1225):
    assemble the Crank-Nicolson system of the snow + lake + soil column and solve
    it in place into t_soisno and t_lake. `state` maps the clmsoil_lake_solve_args
    fields (include/clmsoil_lake.h) to device tensors, plus `nc` and `nlevlak`."""
    d = dict(state)
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_lake_temperature_solve(byref(lake_args(d)), byref(levels(d)),
                                                   float(dtime), _stream()),
               "lake_temperature_solve")


def compute_effec_rootfrac_and_vert_tran_sink(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                                              nlevgrnd: int = 25) -> None:
    """Root water uptake with use_hydrstress = false
    (compute_effec_rootfrac_and_vert_tran_sink!, soil_water_plant_sink.jl:341-451):
    the transpiration-weighted column root fraction and the Richards step's
    root sink qflx_rootsoi_col. `state` maps the clmsoil_rootsink_args fields
    (include/clmsoil_surfhyd.h) to device tensors, plus `nc` and `np`."""
    d = dict(state)
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, L.clmsoil_compute_effec_rootfrac_and_vert_tran_sink(
        byref(rootsink_args(d)), byref(levels(d)), _stream()),
        "compute_effec_rootfrac_and_vert_tran_sink")


def _theta_wt(entry: str, state: dict, nlevsno: int, nlevsoi: int, nlevgrnd: int) -> None:
    d = dict(state)
    d.setdefault("nc", d["zwt"].shape[0])
    d.update(nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    if int(d["nc"]) == 0:
        return
    L = _lib()
    _abi.check(L, getattr(L, "clmsoil_" + entry)(byref(theta_wt_args(d)), byref(levels(d)),
                                                 _stream()), entry)


def perched_water_table(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                        nlevgrnd: int = 25) -> None:
    """perched_water_table! (soil_hydrology.jl:1604-1653): frost table at the top
    of the first frozen layer under an unfrozen one, and the perched water table
    above it. `state` maps the clmsoil_theta_wt_args fields
    (include/clmsoil_hydro.h) to device tensors, plus `nc`."""
    _theta_wt("perched_water_table", state, nlevsno, nlevsoi, nlevgrnd)


def theta_based_water_table(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                            nlevgrnd: int = 25) -> None:
    """theta_based_water_table! (soil_hydrology.jl:1732-1765): the water table
    where the saturation profile above bedrock falls to 90 %."""
    _theta_wt("theta_based_water_table", state, nlevsno, nlevsoi, nlevgrnd)


def clamp_zwt_to_bedrock(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                         nlevgrnd: int = 25) -> None:
    """clamp_zwt_to_bedrock! (src/infrastructure/kernels.jl:209-227)."""
    _theta_wt("clamp_zwt_to_bedrock", state, nlevsno, nlevsoi, nlevgrnd)


def water_table_moisture_form(state: dict, nlevsno: int = 12, nlevsoi: int = 20,
                              nlevgrnd: int = 25) -> None:
    """The water tables after a step without an aquifer layer
    (clm_driver.jl:2076-2092): perched, theta-based, then the bedrock clamp, in
    one launch."""
    _theta_wt("water_table_moisture_form", state, nlevsno, nlevsoi, nlevgrnd)


def soil_water(st, mfin, mask, nbedrock, nlevsno: int, nlevsoi: int, nlevgrnd: int,
               cfg: SoilWaterMovementConfig, dtime: float) -> None:
    """soil_water! dispatch (soil_water_movement.jl:1780-1810). For the moisture
    form `st`/`mfin` are MfState/MfIn; for Zeng-Decker 2009 `st` is the field
    dict of soilwater_zengdecker2009 (`mfin` and `nbedrock` are unused, a
    non-None `mask` replaces the dict's)."""
    if cfg.soilwater_movement_method == MOISTURE_FORM:
        soilwater_moisture_form(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd, dtime, cfg)
    elif cfg.soilwater_movement_method == ZENGDECKER_2009:
        d = dict(st)
        if mask is not None:
            d["mask"] = mask
        soilwater_zengdecker2009(d, dtime, cfg, nlevsno, nlevsoi, nlevgrnd)
    else:
        raise CLMSoilError("soil_water!: a SoilWater implementation must be specified!")


def soil_temperature(state: dict, dtime: float, nlevsno: int = 12, nlevgrnd: int = 25,
                     nlevsoi: int = 20, snow_thermal_cond_method: str = "Jordan1991") -> None:
    """Non-urban soil_temperature! for columns with given ground heat flux terms.

    `state` maps the field names of clmsoil_heat_args (include/clmsoil.h) to
    tensors; outputs that are absent are not written."""
    code = {"Jordan1991": 1, "Sturm1997": 2}.get(snow_thermal_cond_method)
    if code is None:
        raise CLMSoilError(f"Unknown snow_thermal_cond_method: {snow_thermal_cond_method}")
    d = dict(state)
    d.setdefault("nc", d["snl"].shape[0])
    d.update(nlevsno=nlevsno, nlevgrnd=nlevgrnd, nlevsoi=nlevsoi)
    if d["nc"] == 0:
        return
    L = _lib()
    a, lev, hc = heat_args(d), levels(d), heat_config(code)
    _abi.check(L, L.clmsoil_soil_temperature(byref(a), byref(lev), float(dtime), byref(hc), _stream()),
               "soil_temperature")


def tridiagonal_multi(u, a, b, c, r, jtop, mask, ncols: int, nlevs: int) -> None:
    """Batched Thomas solve; tensors (nlevs, ncols), jtop int64 (ncols,)."""
    L = _lib()
    ld = u.shape[-1]
    _abi.check(L, L.clmsoil_tridiagonal_multi(u.data_ptr(), a.data_ptr(), b.data_ptr(), c.data_ptr(),
                                              r.data_ptr(), jtop.data_ptr(),
                                              None if mask is None else mask.data_ptr(),
                                              int(ncols), int(nlevs), int(ld), _stream()),
               "tridiagonal_multi")


def batched_band_solve(tvector, bmatrix, rvector, jtop, jbot, mask, kl: int, ku: int,
                       nlevsno_off: int) -> None:
    """Batched banded solve; tvector/rvector (nlev, nc), bmatrix (nlev, nband, nc)."""
    L = _lib()
    nlev, nc = tvector.shape
    _abi.check(L, L.clmsoil_batched_band_solve(tvector.data_ptr(), bmatrix.data_ptr(),
                                               rvector.data_ptr(), jtop.data_ptr(), jbot.data_ptr(),
                                               None if mask is None else mask.data_ptr(),
                                               int(kl), int(ku), int(nlevsno_off), int(nc), int(nlev),
                                               int(nc), _stream()), "batched_band_solve")


def workspace_errors(workspace) -> int:
    """Synchronize the current stream and return the error bits of the last
    Richards call that used `workspace`."""
    L = _lib()
    flags = C.c_uint32(0)
    _abi.check(L, L.clmsoil_workspace_errors(workspace.data_ptr(), _stream(), C.byref(flags)),
               "workspace_errors")
    return int(flags.value)


def device_errors() -> int:
    """Synchronize the current stream and return/clear the device error bits."""
    L = _lib()
    flags = C.c_uint32(0)
    _abi.check(L, L.clmsoil_device_errors(_stream(), C.byref(flags)), "device_errors")
    return int(flags.value)


def set_tuning(richards_g: int, richards_block: int, heat_g: int, heat_block: int) -> None:
    L = _lib()
    _abi.check(L, L.clmsoil_set_tuning(richards_g, richards_block, heat_g, heat_block), "set_tuning")


def set_richards_schedule(persistent: bool) -> None:
    L = _lib()
    _abi.check(L, L.clmsoil_set_richards_schedule(1 if persistent else 0), "set_richards_schedule")


def set_richards_residency(blocks_per_cu: int) -> None:
    """Blocks per CU of the resident Richards grid (0 = occupancy calculator)."""
    L = _lib()
    _abi.check(L, L.clmsoil_set_richards_residency(int(blocks_per_cu)), "set_richards_residency")


def set_richards_refill(plain_sixteenths: int, learned_sixteenths: int) -> None:
    """Resident Richards grid: refill a wavefront's idle lane groups once n/16 of
    them are idle (plain step, step with the fused first-substep network)."""
    L = _lib()
    _abi.check(L, L.clmsoil_set_richards_refill(int(plain_sixteenths), int(learned_sixteenths)),
               "set_richards_refill")


def set_richards_lds_cache(fields: int) -> None:
    """Per-layer fields of a Richards column kept in the kernel's LDS cache (0, 5 or 7;
    the cache exists at G = 4 with 256-thread blocks)."""
    L = _lib()
    _abi.check(L, L.clmsoil_set_richards_lds_cache(int(fields)), "set_richards_lds_cache")


def set_zd_block(block: int) -> None:
    """Workgroup size of the Zeng-Decker 2009 step (64, 128 or 256)."""
    L = _lib()
    _abi.check(L, L.clmsoil_set_zd_block(int(block)), "set_zd_block")


def get_tuning() -> tuple:
    L = _lib()
    v = [C.c_int() for _ in range(4)]
    _abi.check(L, L.clmsoil_get_tuning(*[C.byref(x) for x in v]), "get_tuning")
    return tuple(int(x.value) for x in v)


_MATH_OPS = {"pow": 0, "log": 1, "log10": 2, "exp": 3, "jl_max": 4, "jl_min": 5,
             "pow_w": 6, "pow_base_w": 7, "log10_w": 8}


def math_eval(op: str, x, y=None):
    """The kernels' FP64 pow/log/log10/exp (clm_math.h), their wavefront-
    dispatched fast-path forms (pow_w, pow_base_w with x the fixed base, log10_w)
    and Julia max/min evaluated on the device, element-wise over float64 CUDA
    tensors; returns a new tensor."""
    import torch
    x = x.contiguous()
    if y is not None:
        y = y.contiguous()
        if y.shape != x.shape or y.dtype != torch.float64:
            raise ValueError("y must be a float64 tensor shaped like x")
    if x.dtype != torch.float64 or not x.is_cuda:
        raise ValueError("x must be a float64 CUDA tensor")
    out = torch.empty_like(x)
    L = _lib()
    _abi.check(L, L.clmsoil_math_eval(_MATH_OPS[op], x.data_ptr(),
                                      None if y is None else y.data_ptr(), out.data_ptr(),
                                      x.numel(), _stream()), "math_eval")
    return out


def to_device(d: dict, device="cuda"):
    """Copy the numpy arrays of a state dict to GPU tensors (same keys)."""
    import numpy as np
    import torch
    out = {}
    for k, v in d.items():
        out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device) if isinstance(v, np.ndarray) else v
    return out


def to_host(d: dict):
    import numpy as np
    out = {}
    for k, v in d.items():
        out[k] = v.detach().cpu().numpy() if hasattr(v, "detach") else v
    return out


__all__ = [n for n in dir() if not n.startswith("_")] + ["field"]
