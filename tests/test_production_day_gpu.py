"""The production day's path end to end against the oracle: each step the heat
step, then the moisture-form Richards step with the aquifer recharge pass
(soilwater_moisture_form, as the day graph runs it), several steps in a row on
production-day columns (0-12 snow layers, bedrock at levels 5-20, water table
0.5-10 m), for the water-table lower boundary and for the zero-flux one with
bedrock. The oracle runs the same steps on the same inputs.

Bar: after every step, every state and output array of both steps equal bit for
bit, substep counts identical (the heat and Richards kernels match the oracle
bit for bit on their own, so the coupled steps must too)."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise, clone
from clmsoil import api, bind, inputs, tuning

HEAT_OUT = ["t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
            "h2osno_no_layers", "snow_depth", "int_snow", "t_grnd", "xmf", "xmf_h2osfc",
            "qflx_snomelt", "qflx_snofrz", "qflx_snow_drain", "qflx_h2osfc_to_ice", "eflx_fgr12",
            "qflx_snomelt_lyr", "qflx_snofrz_lyr"]
RICHARDS_OUT = ["smp_l", "hk_l", "qcharge"]


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [api.BC_WATERTABLE, api.BC_ZERO_FLUX])
def test_production_day_steps_vs_oracle(gpu, bc):
    import torch
    d = inputs.production_day(nc=65536, seed=71 + bc, fill=0.0)
    d.update(inputs.heat_outputs(d, diagnostics=False, fill=0.0))
    nc, ns, nsoi, ngrnd, dtime = d["nc"], d["nlevsno"], d["nlevsoi"], d["nlevgrnd"], d["dtime"]
    assert (ns, nsoi, ngrnd) == (12, 20, 25)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=bc,
                                      use_bedrock=bc == api.BC_ZERO_FLUX)
    ref = clone(d)
    t = api.to_device(d)
    api.set_tuning(4, 256, 4, 256)
    api.set_richards_schedule(True)
    st = api.MfState(t["h2osoi_liq"], t["smp_l"], t["hk_l"], t["qcharge"], t["nsubsteps"])
    mi = api.MfIn(t["z"], t["zi"], t["dz"], t["qflx_rootsoi"], t["watsat"], t["hksat"], t["bsw"],
                  t["sucsat"], t["icefrac"], t["zwt"], t["qflx_infl"], t["smpmin"])
    for step in range(4):
        oracle.soil_temperature(ref, bind.heat_config())
        oracle.soilwater_moisture_form(ref, cfg.to_abi())
        api.soil_temperature({k: t[k] for k in tuning.HEAT_KEYS} | {"nc": nc}, dtime)
        api.soilwater_moisture_form(st, mi, None, t["nbedrock"], ns, nsoi, ngrnd, dtime, cfg)
        torch.cuda.synchronize()
        assert api.device_errors() == 0
        got = api.to_host(t)
        assert_bitwise(got, ref, HEAT_OUT + RICHARDS_OUT)
        assert np.array_equal(got["nsubsteps"], ref["nsubsteps"]), step
        if bc == api.BC_ZERO_FLUX:
            assert np.all(got["qcharge"] == 0.0)
