"""The graph-captured production day against the same steps launched one by one.

Every launch shape and schedule computes the same arithmetic, and the graph
replays the same launches, so results must be bit-identical. The graph is
replayed twice, with eager work in between, and each replay is checked."""
import numpy as np
import pytest

from clmsoil import api, day, learned, tuning

pytestmark = pytest.mark.gpu

STATE = ("t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
         "h2osno_no_layers", "snow_depth", "int_snow")
CHECK = STATE + ("t_grnd", "qcharge", "nsubsteps", "smp_l", "hk_l", "qflx_snomelt",
                 "eflx_fgr12")


def _views(dev):
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"])
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"],
                  dev["smpmin"])
    return st, mi


@pytest.mark.parametrize("mode", ["plain", "fused", "two_launch"])
def test_day_plan_matches_eager_steps(gpu, mode):
    import torch
    nsteps = 3
    dev = tuning.production_state(20000, seed=9)
    snap = {k: dev[k].clone() for k in CHECK}
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    pred = learned.DtPredictor.load() if mode != "plain" else None
    # eager reference
    for _ in range(nsteps):
        tuning.heat_step(dev)
        dt0 = None
        if pred is not None:
            st, mi = _views(dev)
            dt0 = pred.predict_dtsub0(st, mi, None, dev["nbedrock"], 12, 20, 25, 1800.0)
        tuning.richards_step(dev, cfg, dtsub0=dt0)
    torch.cuda.synchronize()
    want = {k: dev[k].clone() for k in CHECK}
    # graph
    for k in CHECK:
        dev[k].copy_(snap[k])
    st, mi = _views(dev)
    plan = day.DayPlan({k: dev[k] for k in tuning.HEAT_KEYS} | {"nc": dev["nc"]}, st, mi, None,
                       dev["nbedrock"], 12, 20, 25, 1800.0, cfg, nsteps, predictor=pred,
                       fused=mode != "two_launch")
    for replay in range(2):
        for k in CHECK:
            dev[k].copy_(snap[k])
        plan.launch()
        torch.cuda.synchronize()
        for k in CHECK:
            a, b = dev[k].cpu().numpy(), want[k].cpu().numpy()
            assert np.array_equal(a, b, equal_nan=True), (replay, k)
        tuning.heat_step({k: v.clone() if hasattr(v, "clone") else v for k, v in dev.items()})
    plan.close()
    assert api.device_errors() == 0


def test_autotune_small(gpu, tmp_path):
    entry = tuning.autotune(nc=8192, reps=1, warm_steps=1)
    assert entry["richards"]["g"] in (1, 2, 4) and entry["heat"]["block"] in (64, 128, 256)
    assert len(entry["candidates"]["richards"]) == 3 * 3 * (1 + 5)
    if entry["richards"]["persistent"]:
        assert len(entry["candidates"]["refill"]) == len(tuning.REFILLS)
    assert entry["richards"]["refill"] in tuning.REFILLS + (tuning.DEFAULT_REFILL,)
    assert entry["richards"]["lds_fields"] in tuning.LDS_FIELDS
    p = tmp_path / "table.json"
    tuning.save_entry(entry, p)
    assert tuning.load_table(p)[0]["columns"] == 8192
    # restore the defaults for the tests that follow
    api.set_tuning(4, 256, 4, 256)
    api.set_richards_schedule(True)
    api.set_richards_residency(0)
    api.set_richards_refill(tuning.DEFAULT_REFILL, 7)
    api.set_richards_lds_cache(tuning.DEFAULT_LDS_FIELDS)
