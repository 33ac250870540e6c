"""ctypes mirror of include/clmsoil.h and the loader of the HIP library.

The structs here are byte-for-byte the C structs of the ABI; `load_library()`
opens the in-tree `libclmsoil.so` (built from `clm.jl_amd/csrc`) and fails
loudly when it is missing — there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libclmsoil.so"

# status codes
(OK, ERR_NULL, ERR_BC, ERR_METHOD, ERR_SIZE, ERR_HIP, ERR_WORKSPACE, ERR_ARG, ERR_NOT_TRAINED,
 ERR_INIT) = range(10)

# soil_water_movement.jl:23-34
ZENGDECKER_2009, MOISTURE_FORM, MIXED_FORM, HEAD_FORM = 0, 1, 2, 3
BC_HEAD, BC_FLUX, BC_ZERO_FLUX, BC_WATERTABLE, BC_AQUIFER = 0, 1, 2, 3, 4
FLUX_INEXPENSIVE, FLUX_EXPENSIVE = 1, 42

_p = C.c_void_p


class Levels(C.Structure):
    _fields_ = [("nlevsno", C.c_int32), ("nlevsoi", C.c_int32),
                ("nlevgrnd", C.c_int32), ("reserved", C.c_int32)]


class SwmConfig(C.Structure):
    _fields_ = [("soilwater_movement_method", C.c_int32),
                ("upper_boundary_condition", C.c_int32),
                ("lower_boundary_condition", C.c_int32),
                ("flux_calculation", C.c_int32),
                ("dtmin", C.c_double), ("verySmall", C.c_double),
                ("xTolerUpper", C.c_double), ("xTolerLower", C.c_double),
                ("e_ice", C.c_double)]


MF_FIELDS = ["mask", "nbedrock", "h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps",
             "z", "zi", "dz", "qflx_rootsoi", "watsat", "hksat", "bsw", "sucsat",
             "icefrac", "zwt", "qflx_infl", "smpmin", "nrejected", "dtsub0"]


class MfArgs(C.Structure):
    _fields_ = ([("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in MF_FIELDS] +
                [("workspace", C.c_void_p), ("workspace_bytes", C.c_int64)])


HEAT_FIELDS = ["mask", "snl", "nbedrock", "z", "dz", "zi", "t_soisno", "h2osoi_liq",
               "h2osoi_ice", "excess_ice", "watsat", "tkmg", "tkdry", "csol", "bsw",
               "sucsat", "frac_sno_eff", "frac_h2osfc", "hs_top_snow", "hs_soil",
               "hs_top", "hs_h2osfc", "dhsdT", "sabg_lyr_col", "eflx_bot", "t_h2osfc",
               "h2osfc", "h2osno_no_layers", "snow_depth", "int_snow", "t_grnd", "xmf",
               "xmf_h2osfc", "qflx_snomelt", "qflx_snofrz", "qflx_snow_drain",
               "qflx_h2osfc_to_ice", "eflx_fgr12", "qflx_snomelt_lyr", "qflx_snofrz_lyr",
               "fact", "c_h2osfc", "imelt", "thk", "bw", "eflx_fgr", "exice_subs",
               "eflx_snomelt", "eflx_snomelt_r", "eflx_h2osfc_to_snow", "snomelt_accum"]


class HeatArgs(C.Structure):
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in HEAT_FIELDS]


ZD_FIELDS = ["mask", "z", "dz", "zi", "watsat", "hksat", "bsw", "sucsat", "smpmin",
             "h2osoi_liq", "h2osoi_ice", "h2osoi_vol", "t_soisno", "zwt", "qflx_infl",
             "qflx_rootsoi", "icefrac", "smp_l", "hk_l", "qcharge", "qflx_deficit"]


class ZdArgs(C.Structure):
    """clmsoil_zd_args (include/clmsoil_zd09.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in ZD_FIELDS]


GHF_INT_FIELDS = ["patchi", "patchf", "patch_column", "patch_landunit", "col_itype", "snl"]
GHF_U8_FIELDS = ["urbpoi", "mask_nolakec", "mask_nolakep"]
GHF_FIELDS = (GHF_INT_FIELDS + GHF_U8_FIELDS + ["frac_veg_nosno"] +
              ["emg", "t_grnd", "t_h2osfc", "forc_lwrad", "htvp", "frac_sno_eff", "t_soisno",
               "sabg", "sabg_soil", "sabg_snow", "dlrad", "cgrnd", "eflx_sh_grnd", "eflx_sh_snow",
               "eflx_sh_soil", "eflx_sh_h2osfc", "qflx_evap_soi", "qflx_ev_snow", "qflx_ev_soil",
               "qflx_ev_h2osfc", "eflx_lwrad_net", "qflx_tran_veg", "wtcol", "sabg_lyr_patch",
               "eflx_wasteheat_lun", "eflx_ventilation_lun", "eflx_heat_from_ac_lun",
               "eflx_traffic_lun", "wtlunit_roof",
               "eflx_gnet", "dgnetdT", "sabg_chk", "eflx_wasteheat", "eflx_ventilation",
               "eflx_heat_from_ac", "eflx_traffic", "eflx_anthro",
               "hs_top", "hs_top_snow", "hs_soil", "hs_h2osfc", "dhsdT", "sabg_lyr_col"])


class GhfArgs(C.Structure):
    """clmsoil_ghf_args (include/clmsoil_ghf.h)."""
    _fields_ = ([(f, C.c_int64) for f in ("nc", "ldc", "np", "ldp", "nl")] +
                [(f, _p) for f in GHF_FIELDS] +
                [("building_temp_method", C.c_int32), ("reserved", C.c_int32)])


SEB_FIELDS = ["patchi", "patchf", "patch_column", "col_itype", "snl", "mask_nolakec",
              "mask_nolakep", "mask_urbanp", "eflx_soil_grnd", "wtcol", "xmf", "xmf_h2osfc",
              "frac_h2osfc", "t_h2osfc", "t_h2osfc_bef", "c_h2osfc", "eflx_h2osfc_to_snow",
              "eflx_building_heat_errsoi", "frac_sno_eff", "t_soisno", "t_ssbef", "fact",
              "errsoi_patch", "errsoi_col"]


class SebArgs(C.Structure):
    """clmsoil_seb_args (include/clmsoil_ghf.h)."""
    _fields_ = ([(f, C.c_int64) for f in ("nc", "ldc", "np")] +
                [("nlevurb", C.c_int32), ("reserved", C.c_int32)] +
                [(f, _p) for f in SEB_FIELDS])


FRAC_FIELDS = ["mask", "watsat", "dz", "h2osoi_ice", "excess_ice", "eff_porosity", "icefrac"]
WT_FIELDS = ["mask", "dz", "z", "zi", "t_soisno", "watsat", "sucsat", "bsw", "qcharge",
             "h2osoi_liq", "h2osoi_ice", "zwt", "wa", "zwt_perched", "frost_table", "h2osoi_vol",
             "qflx_drain", "qflx_rsub_sat", "qflx_drain_perched"]

THETA_WT_FIELDS = ["mask", "nbedrock", "dz", "z", "zi", "t_soisno", "watsat", "h2osoi_liq",
                   "h2osoi_ice", "zwt", "zwt_perched", "frost_table", "h2osoi_vol"]

SH_INT_FIELDS = ["col_gridcell", "col_landunit", "lun_itype", "col_itype", "snl", "col_cold"]
SH_U8_FIELDS = ["mask_nolakec", "mask_hydrologyc", "mask_urbanc", "col_urbpoi",
                "is_hillslope_column", "col_active"]
SH_FIELDS = ["mask_nolakec", "mask_hydrologyc", "mask_urbanc", "col_gridcell", "col_landunit",
             "lun_itype", "col_itype", "snl", "col_cold", "col_urbpoi", "is_hillslope_column",
             "col_active", "qflx_floodg", "frost_table", "zwt", "zwt_perched", "wtfact", "b_infil",
             "top_max_moist", "top_moist_limited", "qflx_rain_plus_snomelt", "qflx_snow_h2osfc",
             "qflx_liqevap_from_top_layer", "qflx_ev_soil", "qflx_ev_h2osfc", "frac_sno_eff",
             "frac_h2osfc", "frac_h2osfc_nosnow", "icefrac", "hksat", "topo_slope",
             "h2osfc_thresh", "h2osoi_liq", "h2osfc", "xs_urban", "qflx_floodc", "fsat", "fcov",
             "qflx_sat_excess_surf", "qflx_top_soil", "qflx_in_soil", "qflx_top_soil_to_h2osfc",
             "qinmax", "qflx_infl_excess", "qflx_in_soil_limited", "qflx_in_h2osfc",
             "qflx_infl_excess_surf", "qflx_h2osfc_surf", "qflx_h2osfc_drain", "qflx_infl",
             "qflx_surf"]
SH_STEPS = {"floodc": 0x001, "sat_excess": 0x002, "qflx_inputs": 0x004, "infil_excess": 0x008,
            "route": 0x010, "h2osfc": 0x020, "infiltration": 0x040, "surf_runoff": 0x080,
            "urban_pond": 0x100}
SH_ALL = 0x1ff

RS_INT_FIELDS = ["patchi", "npatches"]
RS_U8_FIELDS = ["mask_hydrologyc", "active"]
RS_FIELDS = ["mask_hydrologyc", "patchi", "npatches", "active", "wtcol", "rootr_patch",
             "qflx_tran_veg_patch", "qflx_tran_veg_col", "rootr_col", "qflx_rootsoi_col"]

LAKE_FIELDS = ["mask_lakec", "snl", "z", "cv", "tk", "sabg_lyr_col", "z_lake", "dz_lake",
               "cv_lake", "tk_lake", "phi", "phi_soil", "fin", "tktopsoillay", "t_soisno",
               "t_lake"]

LVT_FIELDS = ["mask_bgc_soilc", "nbedrock", "col_gridcell", "latdeg", "altmax", "altmax_lastyear",
              "zsoi", "dzsoi_decomp", "zisoi", "som_diffus_coef", "som_adv_coef"]
LVT_MAXPOOLS, LVT_MAXTYPES = 16, 4

CH4_FIELDS = ["mask", "col_gridcell", "jwt", "c_atm", "watsat", "h2osoi_vol", "h2osoi_liq",
              "h2osoi_ice", "bsw", "cellorg", "t_soisno", "dz", "h2osfc", "t_grnd", "t_h2osfc",
              "frac_h2osfc", "grnd_ch4_cond", "ch4_prod_depth", "o2_aere_depth", "o2_decomp_depth",
              "o2_oxid_depth", "ch4_oxid_depth", "ch4_aere_depth", "ch4_ebul_depth", "o2stress",
              "ch4stress", "conc_ch4", "conc_o2", "ch4_ebul_total", "ch4_surf_aere", "ch4_surf_ebul",
              "ch4_surf_diff"]
CH4_INT_FIELDS = ["jwt"]


class FractionsArgs(C.Structure):
    """clmsoil_fractions_args (include/clmsoil_hydro.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in FRAC_FIELDS]


class WaterTableArgs(C.Structure):
    """clmsoil_water_table_args (include/clmsoil_hydro.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in WT_FIELDS]


class ThetaWtArgs(C.Structure):
    """clmsoil_theta_wt_args (include/clmsoil_hydro.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in THETA_WT_FIELDS]


class SurfHydArgs(C.Structure):
    """clmsoil_surfhyd_args (include/clmsoil_surfhyd.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in SH_FIELDS]


class SurfHydParams(C.Structure):
    """clmsoil_surfhyd_params (include/clmsoil_surfhyd.h)."""
    _fields_ = [("steps", C.c_int32), ("fsat_method", C.c_int32), ("qinmax_method", C.c_int32),
                ("h2osfcflag", C.c_int32), ("h2osfcflag_update", C.c_int32),
                ("crop_fsat_equals_zero", C.c_int32), ("hillslope_fsat_equals_zero", C.c_int32),
                ("reserved", C.c_int32), ("fff", C.c_double), ("e_ice", C.c_double),
                ("pc", C.c_double), ("mu", C.c_double), ("pondmx_urban", C.c_double)]


class RootSinkArgs(C.Structure):
    """clmsoil_rootsink_args (include/clmsoil_surfhyd.h)."""
    _fields_ = [("nc", C.c_int64), ("ldc", C.c_int64), ("np", C.c_int64), ("ldp", C.c_int64)] + \
        [(f, _p) for f in RS_FIELDS]


class LakeSolveArgs(C.Structure):
    """clmsoil_lake_solve_args (include/clmsoil_lake.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64), ("nlevlak", C.c_int64)] + \
        [(f, _p) for f in LAKE_FIELDS]


class LvtArgs(C.Structure):
    """clmsoil_lvt_args (include/clmsoil_lvt.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64)] + [(f, _p) for f in LVT_FIELDS] + \
        [("conc", _p * LVT_MAXTYPES), ("source", _p * LVT_MAXTYPES),
         ("tendency", _p * LVT_MAXTYPES)]


class LvtParams(C.Structure):
    """clmsoil_lvt_params (include/clmsoil_lvt.h)."""
    _fields_ = [("nlevdecomp", C.c_int32), ("nlevdecomp_full", C.c_int32), ("npools", C.c_int32),
                ("ntypes", C.c_int32), ("spinup_state", C.c_int32),
                ("use_soil_matrixcn", C.c_int32), ("som_diffus", C.c_double),
                ("cryoturb_diffusion_k", C.c_double), ("max_altdepth_cryoturbation", C.c_double),
                ("som_adv_flux", C.c_double), ("max_depth_cryoturb", C.c_double),
                ("spinup_factor", C.c_double * LVT_MAXPOOLS), ("is_cwd", C.c_uint8 * LVT_MAXPOOLS)]


class Ch4TranArgs(C.Structure):
    """clmsoil_ch4_tran_args (include/clmsoil_ch4.h)."""
    _fields_ = [("nc", C.c_int64), ("ld", C.c_int64), ("ldg", C.c_int64)] + \
        [(f, _p) for f in CH4_FIELDS]


class Ch4TranParams(C.Structure):
    """clmsoil_ch4_tran_params (include/clmsoil_ch4.h)."""
    _fields_ = [("nlevsoi", C.c_int32), ("sat", C.c_int32), ("lake", C.c_int32),
                ("ch4frzout", C.c_int32), ("use_aereoxid_prog", C.c_int32), ("reserved", C.c_int32),
                ("satpow", C.c_double), ("scale_factor_gasdiff", C.c_double),
                ("scale_factor_liqdiff", C.c_double), ("capthick", C.c_double),
                ("aereoxid", C.c_double), ("om_frac_sf", C.c_double), ("organic_max", C.c_double)]


class HeatConfig(C.Structure):
    _fields_ = [("snow_thermal_cond_method", C.c_int32), ("reserved", C.c_int32)]


def declare_common(lib: C.CDLL, prefix: str) -> C.CDLL:
    """Declare the argument types of the entry points shared by the product
    library (prefix "clmsoil_") and the oracle (prefix "orc_")."""
    P = C.c_void_p
    i64, i32, dbl = C.c_int64, C.c_int32, C.c_double
    getattr(lib, prefix + "tridiagonal_multi").argtypes = [P, P, P, P, P, P, P, i64, i64, i64] + \
        ([P] if prefix == "clmsoil_" else [])
    getattr(lib, prefix + "batched_band_solve").argtypes = [P, P, P, P, P, P, i32, i32, i32, i64, i64, i64] + \
        ([P] if prefix == "clmsoil_" else [])
    for f in ("tridiagonal_multi", "batched_band_solve", "soilwater_moisture_form",
              "soil_temperature", "soilwater_zengdecker2009", "ground_heat_flux"):
        getattr(lib, prefix + f).restype = C.c_int
    getattr(lib, prefix + "ground_heat_flux").argtypes = [C.POINTER(GhfArgs), C.POINTER(Levels)] + \
        ([P] if prefix == "clmsoil_" else [])
    f = getattr(lib, prefix + "set_soil_water_fractions")
    f.argtypes = [C.POINTER(FractionsArgs), C.POINTER(Levels)] + ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    f = getattr(lib, prefix + "water_table")
    f.argtypes = [C.POINTER(WaterTableArgs), C.POINTER(Levels), dbl, dbl, C.c_int] + \
        ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    for name in ("perched_water_table", "theta_based_water_table", "clamp_zwt_to_bedrock"):
        f = getattr(lib, prefix + name)
        f.argtypes = [C.POINTER(ThetaWtArgs), C.POINTER(Levels)] + ([P] if prefix == "clmsoil_" else [])
        f.restype = C.c_int
    if prefix == "clmsoil_":
        f = lib.clmsoil_water_table_moisture_form
        f.argtypes = [C.POINTER(ThetaWtArgs), C.POINTER(Levels), P]
        f.restype = C.c_int
    f = getattr(lib, prefix + "surface_hydrology")
    # last argument: the stream (product) or 2*nc doubles of scratch (oracle)
    f.argtypes = [C.POINTER(SurfHydArgs), C.POINTER(Levels), dbl, C.POINTER(SurfHydParams), P]
    f.restype = C.c_int
    f = getattr(lib, prefix + "compute_effec_rootfrac_and_vert_tran_sink")
    f.argtypes = [C.POINTER(RootSinkArgs), C.POINTER(Levels)] + ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    f = getattr(lib, prefix + "lake_temperature_solve")
    # last argument: the stream (product) or the scratch copy-out, may be NULL (oracle)
    f.argtypes = [C.POINTER(LakeSolveArgs), C.POINTER(Levels), dbl, P]
    f.restype = C.c_int
    f = getattr(lib, prefix + "litter_vert_transp")
    f.argtypes = [C.POINTER(LvtArgs), dbl, C.POINTER(LvtParams)] + ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    f = getattr(lib, prefix + "ch4_tran")
    f.argtypes = [C.POINTER(Ch4TranArgs), dbl, C.POINTER(Ch4TranParams)] + \
        ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    f = getattr(lib, prefix + "soil_energy_balance")
    f.argtypes = [C.POINTER(SebArgs), C.POINTER(Levels), dbl] + ([P] if prefix == "clmsoil_" else [])
    f.restype = C.c_int
    lib_zd = getattr(lib, prefix + "soilwater_zengdecker2009")
    lib_zd.argtypes = [C.POINTER(ZdArgs), C.POINTER(Levels), dbl, C.POINTER(SwmConfig)] + \
        ([P] if prefix == "clmsoil_" else [])
    lib_sw = getattr(lib, prefix + "soilwater_moisture_form")
    lib_sw.argtypes = [C.POINTER(MfArgs), C.POINTER(Levels), dbl, C.POINTER(SwmConfig)] + \
        ([P] if prefix == "clmsoil_" else [])
    lib_st = getattr(lib, prefix + "soil_temperature")
    lib_st.argtypes = [C.POINTER(HeatArgs), C.POINTER(Levels), dbl, C.POINTER(HeatConfig)] + \
        ([P] if prefix == "clmsoil_" else [])
    return lib


_LIB = None


class LibraryMissing(RuntimeError):
    pass


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Open libclmsoil.so (the HIP library). Raises LibraryMissing if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else Path(os.environ.get("CLMSOIL_LIB", LIB_PATH))
    if not p.exists():
        raise LibraryMissing(
            f"{p} is not built; run `make -C clm.jl_amd/csrc` (or __graft_entry__.build()). "
            "The clmsoil product path has no CPU fallback.")
    lib = C.CDLL(str(p))
    declare_common(lib, "clmsoil_")
    lib.clmsoil_init.restype = C.c_int
    lib.clmsoil_richards_workspace_bytes.restype = C.c_int64
    lib.clmsoil_abi_version.restype = C.c_int
    lib.clmsoil_strerror.restype = C.c_char_p
    lib.clmsoil_strerror.argtypes = [C.c_int]
    lib.clmsoil_device_count.restype = C.c_int
    lib.clmsoil_sync.argtypes = [C.c_void_p]
    lib.clmsoil_sync.restype = C.c_int
    lib.clmsoil_soilwm_moisture_form_solve.argtypes = [C.POINTER(MfArgs), C.POINTER(Levels),
                                                       C.c_double, C.POINTER(SwmConfig), C.c_void_p]
    lib.clmsoil_soilwm_moisture_form_solve.restype = C.c_int
    lib.clmsoil_soilwm_moisture_form_solve_static.argtypes = [
        C.POINTER(MfArgs), C.POINTER(Levels), C.c_double, C.POINTER(SwmConfig), C.c_void_p]
    lib.clmsoil_soilwm_moisture_form_solve_static.restype = C.c_int
    _declare_extended(lib)
    # the library's one-time device set-up (its only allocation); no-op without a GPU
    check(lib, lib.clmsoil_init(), "init")
    if path is None:
        _LIB = lib
    return lib


def _declare_extended(lib: C.CDLL) -> None:
    """Entry points beyond the reference's interfaces (batched day, adjoint, MLP)."""
    from . import _abi_ext
    _abi_ext.declare(lib)


def check(lib: C.CDLL, status: int, what: str) -> None:
    if status != OK:
        msg = lib.clmsoil_strerror(status).decode() if hasattr(lib, "clmsoil_strerror") else str(status)
        raise RuntimeError(f"{what}: {msg} (status {status})")
