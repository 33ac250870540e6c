"""ctypes loader of the CPU oracle (liboracle_clmsoil.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the clmsoil package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
REPO = ORACLE_DIR.parent
sys.path.insert(0, str(REPO / "clm.jl_amd"))

from clmsoil import _abi  # noqa: E402
from clmsoil.bind import (byref, fractions_args, ghf_args, heat_args, levels, mf_args,  # noqa: E402
                          ch4_tran_args, ch4_tran_params, lake_args, lvt_args, lvt_params, rootsink_args, seb_args, surfhyd_args, surfhyd_params, theta_wt_args,
                          water_table_args, zd_args)

LIB_PATH = ORACLE_DIR / "liboracle_clmsoil.so"
_lib = None


def build(force: bool = False) -> None:
    cmd = ["make", "-C", str(ORACLE_DIR)] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        _abi.declare_common(L, "orc_")
        L.orc_soilwm_moisture_form_solve.argtypes = [C.POINTER(_abi.MfArgs), C.POINTER(_abi.Levels),
                                                     C.c_double, C.POINTER(_abi.SwmConfig),
                                                     C.c_void_p, C.c_void_p]
        L.orc_soilwm_moisture_form_solve.restype = C.c_int
        L.orc_jl_max.argtypes = [C.c_double, C.c_double]
        L.orc_jl_max.restype = C.c_double
        L.orc_jl_min.argtypes = [C.c_double, C.c_double]
        L.orc_jl_min.restype = C.c_double
        D = C.c_double
        PD = C.POINTER(C.c_double)
        L.orc_soil_hk.argtypes = [D, D, D, D, PD, PD]
        L.orc_soil_suction.argtypes = [D, D, D, PD, PD]
        L.orc_soil_suction_inverse.argtypes = [D, D, D]
        L.orc_soil_suction_inverse.restype = D
        L.orc_ice_impedance.argtypes = [D, D]
        L.orc_ice_impedance.restype = D
        L.orc_tridiagonal_col.argtypes = [C.c_void_p] * 5 + [C.c_int] * 3
        L.orc_compute_rhs_moisture_form.argtypes = [C.c_int] + [C.c_void_p] * 5
        L.orc_compute_lhs_moisture_form.argtypes = [C.c_int] + [C.c_void_p] * 8
        L.orc_tridiagonal_solve.argtypes = [C.c_void_p] * 5 + [C.c_int] * 2
        L.orc_num_threads.restype = C.c_int
        L.orc_clm_math.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_set_mf_trace.argtypes = [C.c_int64, C.c_void_p, C.c_int]
        L.orc_mf_trace_count.restype = C.c_int
        L.orc_heat_assemble_probe.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                              C.c_int64] + [C.c_void_p] * 5
        L.orc_phase_change_h2osfc.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int64,
                                              C.c_double, C.c_double]
        L.orc_phase_change_beta.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int64,
                                            C.c_void_p]
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"oracle {what} failed with status {status}")


def soilwm_moisture_form_solve(d: dict, cfg, dbg=None) -> None:
    """Run the moisture-form step on the host arrays of `d` (updated in place)."""
    a = mf_args(d)
    lev = levels(d)
    qin = qout = None
    if dbg is not None:
        qin, qout = dbg["qin"], dbg["qout"]
    st = lib().orc_soilwm_moisture_form_solve(byref(a), byref(lev), float(d["dtime"]), byref(cfg),
                                              None if qin is None else qin.ctypes.data,
                                              None if qout is None else qout.ctypes.data)
    check(st, "soilwm_moisture_form_solve")


def soilwater_moisture_form(d: dict, cfg) -> None:
    a = mf_args(d)
    lev = levels(d)
    check(lib().orc_soilwater_moisture_form(byref(a), byref(lev), float(d["dtime"]), byref(cfg)),
          "soilwater_moisture_form")


def soil_temperature(d: dict, hcfg) -> None:
    a = heat_args(d)
    lev = levels(d)
    check(lib().orc_soil_temperature(byref(a), byref(lev), float(d["dtime"]), byref(hcfg)),
          "soil_temperature")


def ground_heat_flux(d: dict) -> None:
    """compute_ground_heat_flux_and_deriv! on the host arrays of `d` (outputs in place)."""
    L = lib()
    check(L.orc_ground_heat_flux(byref(ghf_args(d)), byref(levels(d))), "ground_heat_flux")


def soil_energy_balance(d: dict, dtime: float) -> None:
    """errsoi_patch / errsoi_col of soil_fluxes! on the host arrays of `d`."""
    L = lib()
    check(L.orc_soil_energy_balance(byref(seb_args(d)), byref(levels(d)), float(dtime)),
          "soil_energy_balance")


def set_soil_water_fractions(d: dict) -> None:
    L = lib()
    check(L.orc_set_soil_water_fractions(byref(fractions_args(d)), byref(levels(d))),
          "set_soil_water_fractions")


def water_table(d: dict, dtime: float, aq_sp_yield_min: float = 0.01,
                recompute_frost_table: bool = True) -> None:
    L = lib()
    check(L.orc_water_table(byref(water_table_args(d)), byref(levels(d)), float(dtime),
                            float(aq_sp_yield_min), int(bool(recompute_frost_table))),
          "water_table")


def jl_pow4(x: float) -> float:
    L = lib()
    L.orc_jl_pow4.argtypes = [C.c_double]
    L.orc_jl_pow4.restype = C.c_double
    return L.orc_jl_pow4(float(x))


def set_zd_pow_jitter(seed: int) -> None:
    """Non-zero: every pow() of the Zeng-Decker step moves one ulp (direction
    from a hash of its operands and `seed`); 0: exact pow()."""
    L = lib()
    L.orc_set_zd_pow_jitter.argtypes = [C.c_uint64]
    L.orc_set_zd_pow_jitter.restype = None
    L.orc_set_zd_pow_jitter(int(seed))


def soilwater_zengdecker2009(d: dict, cfg) -> None:
    """Zeng-Decker 2009 step on the host arrays of `d` (updated in place)."""
    L = lib()
    a = zd_args(d)
    lev = levels(d)
    check(L.orc_soilwater_zengdecker2009(byref(a), byref(lev), float(d["dtime"]), byref(cfg)),
          "soilwater_zengdecker2009")


def num_threads() -> int:
    return int(lib().orc_num_threads())


os.environ.setdefault("OMP_NUM_THREADS", str(os.cpu_count() or 1))


def perched_water_table(d: dict) -> None:
    L = lib()
    check(L.orc_perched_water_table(byref(theta_wt_args(d)), byref(levels(d))), "perched_water_table")


def theta_based_water_table(d: dict) -> None:
    L = lib()
    check(L.orc_theta_based_water_table(byref(theta_wt_args(d)), byref(levels(d))),
          "theta_based_water_table")


def clamp_zwt_to_bedrock(d: dict) -> None:
    L = lib()
    check(L.orc_clamp_zwt_to_bedrock(byref(theta_wt_args(d)), byref(levels(d))), "clamp_zwt_to_bedrock")


def water_table_moisture_form(d: dict) -> None:
    """The non-aquifer sequence of clm_driver.jl:2076-2092: perched, theta-based, clamp."""
    perched_water_table(d)
    theta_based_water_table(d)
    clamp_zwt_to_bedrock(d)


def surface_hydrology(d: dict, dtime: float, **params) -> None:
    """The surface hydrology steps of clm_driver.jl:1939-2012 on the host arrays
    of `d`, each step over all columns in turn; `params` as bind.surfhyd_params."""
    import numpy as np
    L = lib()
    scratch = np.zeros(2 * max(int(d["nc"]), 1))
    check(L.orc_surface_hydrology(byref(surfhyd_args(d)), byref(levels(d)), float(dtime),
                                  byref(surfhyd_params(**params)),
                                  scratch.ctypes.data_as(C.c_void_p)), "surface_hydrology")


def set_surfhyd_jitter(seed: int) -> None:
    """Move every exp/pow/sin of the surface hydrology one ulp (0 = exact)."""
    L = lib()
    L.orc_set_surfhyd_jitter.argtypes = [C.c_uint64]
    L.orc_set_surfhyd_jitter.restype = None
    L.orc_set_surfhyd_jitter(int(seed))


def compute_effec_rootfrac_and_vert_tran_sink(d: dict) -> None:
    L = lib()
    check(L.orc_compute_effec_rootfrac_and_vert_tran_sink(byref(rootsink_args(d)), byref(levels(d))),
          "compute_effec_rootfrac_and_vert_tran_sink")


LAKE_DBG = ("zx", "cvx", "phix", "tx", "tkix", "factx", "fnx", "a", "b", "c1", "r")


def lake_temperature_solve(d: dict, dtime: float, dbg: bool = False):
    """The lake temperature solve on the host arrays of `d`; with `dbg` also the
    assembly's scratch arrays, each (N, nc) over the extended levels."""
    import numpy as np
    L = lib()
    nc = int(d["nc"])
    n = int(d["nlevsno"]) + int(d["nlevlak"]) + int(d["nlevgrnd"])
    out = np.zeros((len(LAKE_DBG), n, max(nc, 1))) if dbg else None
    check(L.orc_lake_temperature_solve(byref(lake_args(d)), byref(levels(d)), float(dtime),
                                       None if out is None else out.ctypes.data_as(C.c_void_p)),
          "lake_temperature_solve")
    return dict(zip(LAKE_DBG, out)) if dbg else None


def litter_vert_transp(d: dict, dtime: float, **params) -> None:
    L = lib()
    check(L.orc_litter_vert_transp(byref(lvt_args(d)), float(dtime), byref(lvt_params(**params))),
          "litter_vert_transp")


def ch4_tran(d: dict, dtime: float, dbg: bool = False, **params):
    """ch4_tran! on the host arrays of `d` (clmsoil_ch4_tran_args fields plus `nc`,
    `ldg`); with `dbg` also the CH4 balance error before its correction (nc,), the
    rows at, bt, ct, rt and the solved relative concentration before the clip
    (nc, 2, nlevsoi+1, 5) and k_h_cc (nc, 2, nlevsoi+1)."""
    import numpy as np
    L = lib()
    L.orc_ch4_debug.argtypes = [C.c_void_p] * 3
    L.orc_ch4_debug.restype = None
    nc, nl = int(d["nc"]), int(params["nlevsoi"])
    out = None
    if dbg:
        out = dict(err_raw=np.full(nc, np.nan), rows=np.full((nc, 2, nl + 1, 5), np.nan),
                   khcc=np.full((nc, 2, nl + 1), np.nan))
        L.orc_ch4_debug(*(out[k].ctypes.data_as(C.c_void_p) for k in ("err_raw", "rows", "khcc")))
    try:
        check(L.orc_ch4_tran(byref(ch4_tran_args(d)), float(dtime), byref(ch4_tran_params(**params))),
              "ch4_tran")
    finally:
        if dbg:
            L.orc_ch4_debug(None, None, None)
    return out


def set_ch4_jitter(seed: int) -> None:
    """Move every exp/log/pow of the methane transport one ulp (0 = exact)."""
    L = lib()
    L.orc_set_ch4_jitter.argtypes = [C.c_uint64]
    L.orc_set_ch4_jitter.restype = None
    L.orc_set_ch4_jitter(int(seed))


def ch4_jl_pow(x: float, y: float) -> float:
    L = lib()
    L.orc_ch4_jl_pow.argtypes = [C.c_double, C.c_double]
    L.orc_ch4_jl_pow.restype = C.c_double
    return float(L.orc_ch4_jl_pow(float(x), float(y)))


def ch4_diffusivity(species: int, gas: bool, t: float, watsat: float, vol_min: float,
                    cellorg: float = 0.0, bsw: float = 5.0, liq: float = 0.0, ice: float = 0.0,
                    **params) -> float:
    """The methane transport's diffusivity (m2/s) of `species` (0 CH4, 1 O2) at one
    level: gaseous (`gas`, at or above the water table) or aqueous."""
    L = lib()
    L.orc_ch4_diffusivity.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 7
    L.orc_ch4_diffusivity.restype = C.c_double
    params.setdefault("nlevsoi", 1)
    return float(L.orc_ch4_diffusivity(byref(ch4_tran_params(**params)), int(species), int(bool(gas)),
                                       float(t), float(watsat), float(vol_min), float(cellorg),
                                       float(bsw), float(liq), float(ice)))


def jl_pow5(x: float) -> float:
    L = lib()
    L.orc_jl_pow5.argtypes = [C.c_double]
    L.orc_jl_pow5.restype = C.c_double
    return float(L.orc_jl_pow5(float(x)))


_MATH_OPS = {"pow": 0, "log": 1, "log10": 2, "exp": 3, "jl_max": 4, "jl_min": 5,
             "pow_pos": 6, "pow_base": 7, "log10_pos": 8}


def clm_math(op: str, x, y=None):
    """The library's FP64 pow/log/log10/exp (clm_math.h, compiled into the
    oracle) and Julia's max/min element-wise over float64 arrays."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(x if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_clm_math(_MATH_OPS[op], x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out


def mf_trace(d: dict, cfg, col: int, cap: int = 4096):
    """Run the moisture-form step on `d` (in place) and return the substep
    iterations of column `col` as rows (dtsub, errorMax, accepted)."""
    import numpy as np
    buf = np.zeros((cap, 3))
    L = lib()
    L.orc_set_mf_trace(int(col), buf.ctypes.data, int(cap))
    try:
        soilwm_moisture_form_solve(d, cfg)
        n = L.orc_mf_trace_count()
    finally:
        L.orc_set_mf_trace(-1, None, 0)
    return buf[:n].copy()


def heat_assemble_probe(d: dict, hcfg, col: int):
    """soil_therm_prop! through the banded solve for column `col` of `d`:
    returns tk, cv [nlev], bmatrix [nlev_total, 5] (row, band), rvector and
    tk_h2osfc. t_soisno of the column is updated in place."""
    import numpy as np
    nlev = d["nlevsno"] + d["nlevgrnd"]
    tk, cv = np.zeros(nlev), np.zeros(nlev)
    bm, rv = np.zeros((nlev + 1, 5)), np.zeros(nlev + 1)
    tkh = C.c_double(0.0)
    a, lev = heat_args(d), levels(d)
    check(lib().orc_heat_assemble_probe(byref(a), byref(lev), float(d["dtime"]), byref(hcfg),
                                        int(col), tk.ctypes.data, cv.ctypes.data, bm.ctypes.data,
                                        rv.ctypes.data, C.addressof(tkh)), "heat_assemble_probe")
    return dict(tk=tk, cv=cv, bmatrix=bm, rvector=rv, tk_h2osfc=tkh.value)


def phase_change_h2osfc(d: dict, col: int, c_h2osfc: float, fact_top: float) -> None:
    a, lev = heat_args(d), levels(d)
    check(lib().orc_phase_change_h2osfc(byref(a), byref(lev), float(d["dtime"]), int(col),
                                        float(c_h2osfc), float(fact_top)), "phase_change_h2osfc")


def phase_change_beta(d: dict, col: int, fact) -> None:
    import numpy as np
    f = np.ascontiguousarray(fact, dtype=np.float64)
    a, lev = heat_args(d), levels(d)
    check(lib().orc_phase_change_beta(byref(a), byref(lev), float(d["dtime"]), int(col),
                                      f.ctypes.data), "phase_change_beta")
