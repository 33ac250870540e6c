"""The C-ABI library loads, exports every declared entry point, and validates
arguments on the host (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

import clmsoil
from clmsoil import _abi, bind, inputs

REPO = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(clmsoil_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = clmsoil.load_library()
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_strerror():
    lib = clmsoil.load_library()
    assert lib.clmsoil_abi_version() == 2
    assert lib.clmsoil_strerror(2) == b"unsupported boundary condition"
    assert lib.clmsoil_strerror(_abi.ERR_INIT) == b"clmsoil_init() has not set up this device"
    assert lib.clmsoil_richards_workspace_bytes() == 256


def test_struct_sizes_match_header():
    # 2 int64 + pointer fields, 8-byte aligned
    assert C.sizeof(_abi.MfArgs) == 16 + 8 * len(_abi.MF_FIELDS) + 16
    assert C.sizeof(_abi.HeatArgs) == 16 + 8 * len(_abi.HEAT_FIELDS)
    assert C.sizeof(_abi.SwmConfig) == 16 + 5 * 8
    assert C.sizeof(_abi.Levels) == 16


def _mf(d):
    return bind.mf_args(d), bind.levels(d)


def test_host_validation_without_gpu():
    lib = clmsoil.load_library()
    d = inputs.richards_config0(nc=8)
    a, lev = _mf(d)
    bad_upper = bind.swm_config(lower_bc=2, upper_bc=0)
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a), C.byref(lev), 1800.0, C.byref(bad_upper), None) == _abi.ERR_BC
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a), C.byref(lev), 1800.0,
                                                  C.byref(bind.swm_config(lower_bc=4)), None) == _abi.ERR_BC
    zd = bind.swm_config(lower_bc=2, method=0)
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a), C.byref(lev), 1800.0, C.byref(zd), None) == _abi.ERR_METHOD
    ok = bind.swm_config(lower_bc=2)
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a), C.byref(lev), float("nan"), C.byref(ok), None) == _abi.ERR_ARG
    a.h2osoi_liq = None
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a), C.byref(lev), 1800.0, C.byref(ok), None) == _abi.ERR_NULL
    big = _abi.Levels(12, 30, 35, 0)
    a2, _ = _mf(d)
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a2), C.byref(big), 1800.0, C.byref(ok), None) == _abi.ERR_SIZE
    a2.nc = 0
    assert lib.clmsoil_soilwm_moisture_form_solve(C.byref(a2), C.byref(lev), 1800.0, C.byref(ok), None) == _abi.OK
    assert lib.clmsoil_batched_band_solve(None, None, None, None, None, None, 3, 2, 12, 4, 38, 4, None) == _abi.ERR_ARG
    assert lib.clmsoil_tridiagonal_multi(None, None, None, None, None, None, None, 4, 100, 4, None) == _abi.ERR_SIZE


def test_api_config_validation():
    import pytest
    from clmsoil import api
    with pytest.raises(api.CLMSoilError):
        api.init_soilwater_movement(soilwater_movement_method=api.ZENGDECKER_2009,
                                    lower_boundary_condition=api.BC_FLUX)
    with pytest.raises(api.CLMSoilError):
        api.init_soilwater_movement(use_bedrock=True, soilwater_movement_method=api.MOISTURE_FORM,
                                    lower_boundary_condition=api.BC_FLUX)
    cfg = api.init_soilwater_movement(use_bedrock=True, soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    assert not api.use_aquifer_layer(cfg)
    assert api.use_aquifer_layer(api.SoilWaterMovementConfig(lower_boundary_condition=api.BC_WATERTABLE))


def test_transport_entries_validate_on_the_host():
    """The litter and methane entries reject bad sizes, methods and missing arrays
    before touching a device (an empty call returns at once)."""
    import numpy as np
    from clmsoil import inputs
    from clmsoil._abi import load_library
    lib = load_library()
    ok = _abi.OK if hasattr(_abi, "OK") else 0
    # litter
    d = inputs.lvt_columns(nc=8, seed=1)
    a = bind.lvt_args(d | {"nc": 8})
    p = bind.lvt_params(d["nlevdecomp"], d["nlevdecomp_full"], d["npools"], d["ntypes"], d["is_cwd"])
    assert lib.clmsoil_litter_vert_transp(C.byref(a), 0.0, C.byref(p), None) == _abi.ERR_ARG
    pm = bind.lvt_params(d["nlevdecomp"], d["nlevdecomp_full"], d["npools"], d["ntypes"], d["is_cwd"],
                         use_soil_matrixcn=True)
    assert lib.clmsoil_litter_vert_transp(C.byref(a), 1800.0, C.byref(pm), None) == _abi.ERR_METHOD
    pbig = bind.lvt_params(63, 63, d["npools"], d["ntypes"], d["is_cwd"])
    assert lib.clmsoil_litter_vert_transp(C.byref(a), 1800.0, C.byref(pbig), None) == _abi.ERR_SIZE
    a0 = bind.lvt_args(d | {"nc": 8})
    a0.zisoi = None
    assert lib.clmsoil_litter_vert_transp(C.byref(a0), 1800.0, C.byref(p), None) == _abi.ERR_NULL
    ae = bind.lvt_args(d | {"nc": 0, "ld": 8})
    assert lib.clmsoil_litter_vert_transp(C.byref(ae), 1800.0, C.byref(p), None) == ok
    # methane
    m = inputs.ch4_columns(nc=8, seed=2)
    b = bind.ch4_tran_args(m)
    q = bind.ch4_tran_params(m["nlevsoi"])
    assert lib.clmsoil_ch4_tran(C.byref(b), -1.0, C.byref(q), None) == _abi.ERR_ARG
    assert lib.clmsoil_ch4_tran(C.byref(b), 1800.0, C.byref(bind.ch4_tran_params(61)), None) == _abi.ERR_SIZE
    assert lib.clmsoil_ch4_tran(C.byref(b), 1800.0, C.byref(bind.ch4_tran_params(m["nlevsoi"], sat=2)),
                                None) == _abi.ERR_ARG
    b0 = bind.ch4_tran_args(m)
    b0.conc_o2 = None
    assert lib.clmsoil_ch4_tran(C.byref(b0), 1800.0, C.byref(q), None) == _abi.ERR_NULL
    bl = bind.ch4_tran_args(m | {"ld": 4})
    assert lib.clmsoil_ch4_tran(C.byref(bl), 1800.0, C.byref(q), None) == _abi.ERR_SIZE
    be = bind.ch4_tran_args(m | {"nc": 0, "ld": 8})
    assert lib.clmsoil_ch4_tran(C.byref(be), 1800.0, C.byref(q), None) == ok


def test_tuning_setters_validate_on_the_host():
    """Launch-tuning setters accept only the shapes the kernels are built for
    (G = 8 stays rejected) and refill thresholds of 1..16 sixteenths."""
    lib = clmsoil.load_library()
    assert lib.clmsoil_set_tuning(8, 256, 4, 256) == _abi.ERR_ARG
    assert lib.clmsoil_set_tuning(4, 96, 4, 256) == _abi.ERR_ARG
    assert lib.clmsoil_set_richards_refill(0, 7) == _abi.ERR_ARG
    assert lib.clmsoil_set_richards_refill(7, 17) == _abi.ERR_ARG
    assert lib.clmsoil_set_richards_refill(16, 1) == 0
    assert lib.clmsoil_set_richards_refill(7, 7) == 0
    assert lib.clmsoil_set_richards_residency(33) == _abi.ERR_ARG
    assert lib.clmsoil_set_richards_lds_cache(6) == _abi.ERR_ARG
    for f in (0, 5, 7):
        assert lib.clmsoil_set_richards_lds_cache(f) == 0
