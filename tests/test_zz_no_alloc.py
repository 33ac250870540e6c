"""The Richards entry point that allocates nothing
(clmsoil_soilwm_moisture_form_solve_static: the library's per-GPU static scratch
instead of a caller workspace) against the workspace entry point it varies:
equal results bit for bit, on two calls in a row, the second on fewer columns
(the scratch keeps the first call's counters and must be cleared by the
second), with no error reported."""
import ctypes as C

import numpy as np
import pytest

from _util import assert_bitwise
from clmsoil import _abi, api, bind, inputs

OUT = ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps", "nrejected")


def _state(nc, seed):
    d = inputs.richards_config0(nc=nc, seed=seed, nbedrock="random", lower_bc=api.BC_WATERTABLE)
    d["nrejected"] = np.full(nc, np.nan)
    return d


@pytest.mark.gpu
def test_static_scratch_entry_matches_workspace_entry(gpu):
    import torch
    import clmsoil
    L = clmsoil.load_library()
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    c = cfg.to_abi()
    api.set_tuning(4, 256, 4, 256)
    api.set_richards_schedule(True)
    for nc, seed in ((1 << 17, 81), (3001, 82)):
        d = _state(nc, seed)
        lev = bind.levels(d)
        # the existing entry point, with a workspace
        t1 = api.to_device(d)
        ws = api.richards_workspace()
        d1 = dict(t1, nc=nc, workspace=ws)
        st = L.clmsoil_soilwm_moisture_form_solve(C.byref(bind.mf_args(d1)), C.byref(lev),
                                                  d["dtime"], C.byref(c), None)
        assert st == _abi.OK
        # the static-scratch entry point
        t2 = api.to_device(d)
        d2 = dict(t2, nc=nc)
        st = L.clmsoil_soilwm_moisture_form_solve_static(C.byref(bind.mf_args(d2)), C.byref(lev),
                                                         d["dtime"], C.byref(c), None)
        assert st == _abi.OK
        torch.cuda.synchronize()
        assert api.device_errors() == 0
        assert api.workspace_errors(ws) == 0
        g1, g2 = api.to_host(t1), api.to_host(t2)
        assert_bitwise(g2, g1, OUT)
        assert np.isfinite(g2["h2osoi_liq"][d["nlevsno"]:]).all()


def test_static_entry_host_validation():
    import clmsoil
    L = clmsoil.load_library()
    d = _state(8, 1)
    a, lev = bind.mf_args(d), bind.levels(d)
    bad = bind.swm_config(lower_bc=4)
    assert L.clmsoil_soilwm_moisture_form_solve_static(C.byref(a), C.byref(lev), 1800.0,
                                                       C.byref(bad), None) == _abi.ERR_BC
    ok = bind.swm_config(lower_bc=2)
    a.nc = (8 << 31) + 1
    a.ld = a.nc
    assert L.clmsoil_soilwm_moisture_form_solve_static(C.byref(a), C.byref(lev), 1800.0,
                                                       C.byref(ok), None) == _abi.ERR_SIZE
