"""Build ABI argument structs from numpy arrays (host) or torch tensors (device).

Arrays follow the CLM layout described in `clmsoil.inputs`: shape (nlev, nc),
C-contiguous, so the address of element (level k, column c) is
`base + 8*(c + k*nc)` — the Julia `[c, j]` column-major offset with ld = nc.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (CH4_FIELDS, CH4_INT_FIELDS, Ch4TranArgs, Ch4TranParams, FRAC_FIELDS, GHF_FIELDS, GHF_INT_FIELDS, GHF_U8_FIELDS, HEAT_FIELDS, MF_FIELDS,
                   LAKE_FIELDS, LVT_FIELDS, LVT_MAXPOOLS, LVT_MAXTYPES, LvtArgs, LvtParams,
                   RS_FIELDS, RS_INT_FIELDS, RS_U8_FIELDS, SEB_FIELDS, SH_ALL,
                   SH_FIELDS, LakeSolveArgs,
                   SH_INT_FIELDS, SH_U8_FIELDS, THETA_WT_FIELDS, RootSinkArgs,
                   WT_FIELDS, ZD_FIELDS, FractionsArgs, GhfArgs, HeatArgs, HeatConfig, Levels,
                   MfArgs, SebArgs, SurfHydArgs, SurfHydParams, SwmConfig, ThetaWtArgs,
                   WaterTableArgs, ZdArgs)

_INT_FIELDS = ({"nbedrock", "imelt", "frac_veg_nosno", "col_gridcell"} | set(GHF_INT_FIELDS)
               | set(SH_INT_FIELDS) | set(RS_INT_FIELDS) | set(CH4_INT_FIELDS))
_U8_FIELDS = ({"mask", "mask_urbanp", "mask_lakec", "mask_bgc_soilc"} | set(GHF_U8_FIELDS)
              | set(SH_U8_FIELDS) | set(RS_U8_FIELDS))


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous (CLM layout)")
        return x.ctypes.data
    # torch tensor
    if not x.is_contiguous():
        raise ValueError("tensors must be contiguous (CLM layout)")
    return x.data_ptr()


def _check_dtype(name, x):
    if x is None:
        return
    dt = str(x.dtype)
    if name in _INT_FIELDS:
        ok = dt in ("int64", "torch.int64")
    elif name in _U8_FIELDS:
        ok = dt in ("uint8", "bool", "torch.uint8", "torch.bool")
    else:
        ok = dt in ("float64", "torch.float64")
    if not ok:
        raise TypeError(f"{name}: unexpected dtype {dt}")


def levels(d) -> Levels:
    return Levels(int(d["nlevsno"]), int(d["nlevsoi"]), int(d["nlevgrnd"]), 0)


def mf_args(d: dict) -> MfArgs:
    """clmsoil_mf_args from a dict of arrays; `workspace` (optional) is a device
    buffer of at least clmsoil_richards_workspace_bytes() bytes."""
    nc = int(d["nc"])
    a = MfArgs()
    a.nc = nc
    a.ld = int(d.get("ld", nc))
    for f in MF_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    ws = d.get("workspace")
    if ws is not None:
        a.workspace = _ptr(ws)
        a.workspace_bytes = int(ws.numel() * ws.element_size()) if hasattr(ws, "numel") else ws.nbytes
    return a


def heat_args(d: dict) -> HeatArgs:
    nc = int(d["nc"])
    a = HeatArgs()
    a.nc = nc
    a.ld = int(d.get("ld", nc))
    for f in HEAT_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def zd_args(d: dict) -> ZdArgs:
    nc = int(d["nc"])
    a = ZdArgs()
    a.nc = nc
    a.ld = int(d.get("ld", nc))
    for f in ZD_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def ghf_args(d: dict) -> GhfArgs:
    """clmsoil_ghf_args from a dict of arrays; `nc`, `np`, `nl` and optionally
    `ldc`/`ldp` and `building_temp_method` are integers."""
    a = GhfArgs()
    a.nc, a.np, a.nl = int(d["nc"]), int(d["np"]), int(d["nl"])
    a.ldc = int(d.get("ldc", a.nc))
    a.ldp = int(d.get("ldp", a.np))
    for f in GHF_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    a.building_temp_method = int(d.get("building_temp_method", 0))
    return a


def seb_args(d: dict) -> SebArgs:
    """clmsoil_seb_args from a dict of arrays (`nc`, `np`, `nlevurb` integers)."""
    a = SebArgs()
    a.nc, a.np = int(d["nc"]), int(d["np"])
    a.ldc = int(d.get("ldc", a.nc))
    a.nlevurb = int(d.get("nlevurb", 0))
    for f in SEB_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def _simple_args(cls, fields, d: dict):
    nc = int(d["nc"])
    a = cls()
    a.nc = nc
    a.ld = int(d.get("ld", nc))
    for f in fields:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def fractions_args(d: dict) -> FractionsArgs:
    return _simple_args(FractionsArgs, FRAC_FIELDS, d)


def water_table_args(d: dict) -> WaterTableArgs:
    return _simple_args(WaterTableArgs, WT_FIELDS, d)


def theta_wt_args(d: dict) -> ThetaWtArgs:
    return _simple_args(ThetaWtArgs, THETA_WT_FIELDS, d)


def lvt_args(d: dict) -> LvtArgs:
    """clmsoil_lvt_args from a dict of arrays: the LVT_FIELDS plus lists `conc`,
    `source`, `tendency` (one array per tracer type) and `nc`."""
    a = LvtArgs()
    a.nc = int(d["nc"])
    a.ld = int(d.get("ld", a.nc))
    for f in LVT_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    for key in ("conc", "source", "tendency"):
        xs = list(d[key])
        if len(xs) > LVT_MAXTYPES:
            raise ValueError(f"at most {LVT_MAXTYPES} tracer types")
        arr = getattr(a, key)
        for i, x in enumerate(xs):
            _check_dtype(key, x)
            arr[i] = _ptr(x)
    return a


def lvt_params(nlevdecomp, nlevdecomp_full, npools, ntypes, is_cwd, spinup_factor=None,
               spinup_state=0, som_diffus=1.0e-4 / (365.0 * 86400.0),
               cryoturb_diffusion_k=5.0e-4 / (365.0 * 86400.0), max_altdepth_cryoturbation=2.0,
               som_adv_flux=0.0, max_depth_cryoturb=3.0, use_soil_matrixcn=False) -> LvtParams:
    """clmsoil_lvt_params. The reference reads som_diffus, cryoturb_diffusion_k and
    max_altdepth_cryoturbation from its parameter file (litter_vert_transp.jl:29-33);
    the defaults here are 1e-4 and 5e-4 m2/yr and 2 m. som_adv_flux 0 and
    max_depth_cryoturb 3 m are the reference's (:36-37)."""
    p = LvtParams()
    p.nlevdecomp, p.nlevdecomp_full, p.npools, p.ntypes = (int(nlevdecomp), int(nlevdecomp_full),
                                                           int(npools), int(ntypes))
    p.spinup_state, p.use_soil_matrixcn = int(spinup_state), int(bool(use_soil_matrixcn))
    p.som_diffus, p.cryoturb_diffusion_k = float(som_diffus), float(cryoturb_diffusion_k)
    p.max_altdepth_cryoturbation, p.som_adv_flux = float(max_altdepth_cryoturbation), float(som_adv_flux)
    p.max_depth_cryoturb = float(max_depth_cryoturb)
    sf = [1.0] * npools if spinup_factor is None else list(spinup_factor)
    for i in range(npools):
        p.spinup_factor[i] = float(sf[i])
        p.is_cwd[i] = int(bool(is_cwd[i]))
    return p


def ch4_tran_args(d: dict) -> Ch4TranArgs:
    """clmsoil_ch4_tran_args from a dict of arrays: the CH4_FIELDS plus `nc` and
    optionally `ld` and `ldg` (the gridcell count, c_atm's leading dimension)."""
    a = Ch4TranArgs()
    a.nc = int(d["nc"])
    a.ld = int(d.get("ld", a.nc))
    a.ldg = int(d["ldg"])
    for f in CH4_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def ch4_tran_params(nlevsoi, sat=0, lake=False, ch4frzout=False, use_aereoxid_prog=True,
                    satpow=2.0, scale_factor_gasdiff=1.0, scale_factor_liqdiff=1.0,
                    capthick=100.0, aereoxid=0.5, om_frac_sf=1.0,
                    organic_max=130.0) -> Ch4TranParams:
    """clmsoil_ch4_tran_params. The defaults are CH4Params' (methane.jl:214-236),
    CH4VarCon's (ch4frzout false, use_aereoxid_prog true) and the driver's
    organic_max of 130 kg/m3 (clm_initialize.jl:586)."""
    p = Ch4TranParams()
    p.nlevsoi, p.sat, p.lake = int(nlevsoi), int(sat), int(bool(lake))
    p.ch4frzout, p.use_aereoxid_prog = int(bool(ch4frzout)), int(bool(use_aereoxid_prog))
    p.satpow, p.scale_factor_gasdiff = float(satpow), float(scale_factor_gasdiff)
    p.scale_factor_liqdiff, p.capthick = float(scale_factor_liqdiff), float(capthick)
    p.aereoxid, p.om_frac_sf, p.organic_max = float(aereoxid), float(om_frac_sf), float(organic_max)
    return p


def lake_args(d: dict) -> LakeSolveArgs:
    """clmsoil_lake_solve_args from a dict of arrays (`nc`, `nlevlak` integers)."""
    a = LakeSolveArgs()
    a.nc = int(d["nc"])
    a.ld = int(d.get("ld", a.nc))
    a.nlevlak = int(d["nlevlak"])
    for f in LAKE_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def rootsink_args(d: dict) -> RootSinkArgs:
    """clmsoil_rootsink_args from a dict of arrays (`nc`, `np` integers)."""
    a = RootSinkArgs()
    a.nc, a.np = int(d["nc"]), int(d["np"])
    a.ldc = int(d.get("ldc", a.nc))
    a.ldp = int(d.get("ldp", a.np))
    for f in RS_FIELDS:
        x = d.get(f)
        _check_dtype(f, x)
        setattr(a, f, _ptr(x))
    return a


def surfhyd_args(d: dict) -> SurfHydArgs:
    return _simple_args(SurfHydArgs, SH_FIELDS, d)


def surfhyd_params(steps=SH_ALL, fsat_method=1, qinmax_method=1, h2osfcflag=1,
                   h2osfcflag_update=1, crop_fsat_equals_zero=False,
                   hillslope_fsat_equals_zero=False, fff=0.5, e_ice=6.0, pc=0.4, mu=0.13889,
                   pondmx_urban=1.0) -> SurfHydParams:
    """The driver's settings: TOPMODEL fsat, hksat qinmax, surface water on
    (soil_hydrology.jl:20, surface_water.jl:19-20, infilt_excess_runoff.jl:25,
    varcon.jl:68). The reference reads fff from its parameter file
    (read_params.jl:22-23); the default here is 0.5 /m."""
    return SurfHydParams(int(steps), int(fsat_method), int(qinmax_method), int(h2osfcflag),
                         int(h2osfcflag_update), int(bool(crop_fsat_equals_zero)),
                         int(bool(hillslope_fsat_equals_zero)), 0, float(fff), float(e_ice),
                         float(pc), float(mu), float(pondmx_urban))


def swm_config(lower_bc=2, upper_bc=1, method=1, flux_calculation=1, dtmin=60.0,
               verySmall=1.0e-8, xTolerUpper=1.0e-1, xTolerLower=1.0e-2, e_ice=6.0) -> SwmConfig:
    """SoilWaterMovementConfig defaults (soil_water_movement.jl:654-671), moisture form."""
    return SwmConfig(method, upper_bc, lower_bc, flux_calculation, dtmin, verySmall,
                     xTolerUpper, xTolerLower, e_ice)


def heat_config(snow_thermal_cond_method=1) -> HeatConfig:
    return HeatConfig(snow_thermal_cond_method, 0)


def byref(x):
    return C.byref(x)
