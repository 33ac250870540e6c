"""Pins of the oracle's surface hydrology before the Richards step
(orc_surface_hydrology, surfhyd_oracle.c, following clm_driver.jl:1939-2012).

* the reference's known answers, transcribed with their inputs:
  saturated excess runoff (test_sat_excess_runoff.jl:54-110, 115-176, 180-420:
  TOPMODEL branches, VIC, crop and hillslope zeroing), infiltration excess
  (test_infilt_excess_runoff.jl:72-150, 150-240: qinmax from hksat and ice, the
  HKSAT and NONE methods, masking), set_floodc! and infiltration!
  (test_soil_hydrology_mod.jl:117-150), the relative truncation of the surface
  water store (:152-172), urban ponding with its water balance (:495-590) and
  the default root water uptake (test_soil_water_plant_sink.jl:118-276, 479-530);
* the water balance of the whole sequence on a soil or crop column, which the
  mathematics fixes: infiltration + surface runoff + d(h2osfc)/dt equals the
  water arriving at the surface minus the evaporation drawn from soil and
  ponded water (a dropped term, a wrong sign or a wrong fraction fails it);
* the root uptake distributes the column transpiration: with every patch's
  root profile summing to 1, the column profile sums to 1 and the root sink to
  the column transpiration.
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs
from clmsoil.bind import SH_ALL
from clmsoil._abi import SH_STEPS

ISPVAL = -9999
ROOF, SUNWALL, SHADEWALL, ROAD_IMPERV = 71, 72, 73, 74
OUT = ("qflx_floodc", "fsat", "fcov", "qflx_sat_excess_surf", "qflx_top_soil", "qflx_in_soil",
       "qflx_top_soil_to_h2osfc", "qinmax", "qflx_infl_excess", "qflx_in_soil_limited",
       "qflx_in_h2osfc", "qflx_infl_excess_surf", "qflx_h2osfc_surf", "qflx_h2osfc_drain",
       "qflx_infl", "qflx_surf", "xs_urban")


def cols(nc, nlevsno=12, nlevgrnd=15, **over):
    """nc soil columns on one soil landunit; every output NaN (untouched shows)."""
    z = np.zeros(nc)
    d = dict(nc=nc, nlevsno=nlevsno, nlevsoi=10, nlevgrnd=nlevgrnd,
             mask_nolakec=np.ones(nc, np.uint8), mask_hydrologyc=np.ones(nc, np.uint8),
             mask_urbanc=np.zeros(nc, np.uint8), col_gridcell=np.ones(nc, np.int64),
             col_landunit=np.ones(nc, np.int64), lun_itype=np.ones(1, np.int64),
             col_itype=np.ones(nc, np.int64), snl=np.zeros(nc, np.int64),
             col_cold=np.full(nc, ISPVAL, np.int64), col_urbpoi=np.zeros(nc, np.uint8),
             is_hillslope_column=np.zeros(nc, np.uint8), col_active=np.ones(nc, np.uint8),
             qflx_floodg=np.zeros(1), frost_table=z.copy(), zwt=z.copy(), zwt_perched=z.copy(),
             wtfact=z.copy(), b_infil=z.copy(), top_max_moist=np.ones(nc),
             top_moist_limited=z.copy(), qflx_rain_plus_snomelt=z.copy(),
             qflx_snow_h2osfc=z.copy(), qflx_liqevap_from_top_layer=z.copy(),
             qflx_ev_soil=z.copy(), qflx_ev_h2osfc=z.copy(), frac_sno_eff=z.copy(),
             frac_h2osfc=z.copy(), frac_h2osfc_nosnow=z.copy(),
             icefrac=np.zeros((nlevgrnd, nc)), hksat=np.full((nlevgrnd, nc), 0.01),
             topo_slope=z.copy(), h2osfc_thresh=z.copy(),
             h2osoi_liq=np.zeros((nlevsno + nlevgrnd, nc)), h2osfc=z.copy())
    for k in OUT:
        d[k] = np.full(nc, np.nan)
    d["qflx_floodc"] = z.copy()
    for k, v in over.items():
        d[k] = np.asarray(v, dtype=d[k].dtype) if k in d else v
    return d


def run(d, dtime=1800.0, step=None, **params):
    oracle.surface_hydrology(d, dtime, steps=SH_STEPS[step] if step else SH_ALL, **params)
    return d


# ---- saturated excess runoff (test_sat_excess_runoff.jl) ----

def test_fsat_topmodel_branches():
    d = run(cols(3, frost_table=[5.0, 1.5, 0.5], zwt=[2.0] * 3, zwt_perched=[1.0] * 3,
                 wtfact=[0.4] * 3, qflx_rain_plus_snomelt=[5.0, 10.0, 2.0]), step="sat_excess",
            fff=0.5)
    fs = [0.4 * np.exp(-0.5 * 0.5 * 2.0), 0.4 * np.exp(-0.5 * 0.5 * 1.0), 0.4 * np.exp(-0.5)]
    np.testing.assert_allclose(d["fsat"], fs, rtol=1e-15)
    np.testing.assert_allclose(d["qflx_sat_excess_surf"], np.array(fs) * [5.0, 10.0, 2.0],
                               rtol=1e-15)
    assert np.array_equal(d["fcov"], d["fsat"])
    d = run(cols(2, frost_table=[5.0, 5.0], zwt=[0.0, 0.1], zwt_perched=[10.0, 10.0],
                 wtfact=[0.5, 0.8]), step="sat_excess", fff=1.0)
    np.testing.assert_allclose(d["fsat"], [0.5, 0.8 * np.exp(-0.05)], rtol=1e-15)


def test_fsat_vic():
    d = run(cols(3, mask_hydrologyc=[1, 1, 0], b_infil=[0.5, 2.0, 1.0],
                 top_max_moist=[100.0, 200.0, 100.0], top_moist_limited=[50.0, 180.0, 50.0]),
            step="sat_excess", fsat_method=2)
    np.testing.assert_allclose(d["fsat"][:2], [1.0 - 0.5 ** (0.5 / 1.5), 1.0 - 0.1 ** (2.0 / 3.0)],
                               rtol=1e-14)
    assert np.isnan(d["fsat"][2])
    d = run(cols(2, b_infil=[1.0, 1.0], top_max_moist=[100.0, 100.0],
                 top_moist_limited=[100.0, 0.0]), step="sat_excess", fsat_method=2)
    assert d["fsat"][0] == 1.0 and d["fsat"][1] == 0.0
    d = run(cols(2, b_infil=[1.0, 2.0], top_max_moist=[100.0, 200.0],
                 top_moist_limited=[50.0, 150.0], qflx_rain_plus_snomelt=[8.0, 4.0]),
            step="sat_excess", fsat_method=2)
    fs = [1.0 - 0.5 ** 0.5, 1.0 - 0.25 ** (2.0 / 3.0)]
    np.testing.assert_allclose(d["fsat"], fs, rtol=1e-14)
    np.testing.assert_allclose(d["qflx_sat_excess_surf"], np.array(fs) * [8.0, 4.0], rtol=1e-14)


def test_fsat_crop_and_hillslope_zero():
    base = dict(frost_table=[5.0, 5.0], zwt=[1.0, 1.0], zwt_perched=[10.0, 10.0],
                wtfact=[0.4, 0.4], qflx_rain_plus_snomelt=[10.0, 10.0])
    d = run(cols(2, col_landunit=[1, 2], lun_itype=[1, 2], **base), step="sat_excess",
            crop_fsat_equals_zero=True, fff=0.5)
    fs = 0.4 * np.exp(-0.5 * 0.5 * 1.0)
    np.testing.assert_allclose(d["fsat"], [fs, 0.0], rtol=1e-15)
    np.testing.assert_allclose(d["qflx_sat_excess_surf"], [fs * 10.0, 0.0], rtol=1e-15)
    b3 = {k: v + v[:1] for k, v in base.items()}
    d = run(cols(3, is_hillslope_column=[0, 1, 1], col_cold=[ISPVAL, 3, ISPVAL], **b3),
            step="sat_excess", hillslope_fsat_equals_zero=True, fff=0.5)
    np.testing.assert_allclose(d["fsat"], [fs, 0.0, fs], rtol=1e-15)


def test_fsat_urban_adds_flood_water():
    d = run(cols(2, wtfact=[0.3, 0.3], zwt=[1.0, 1.0], qflx_rain_plus_snomelt=[2.0, 2.0],
                 col_urbpoi=[0, 1], qflx_floodc=[0.5, 0.5]), step="sat_excess", fff=0.5)
    fs = 0.3 * np.exp(-0.25)
    np.testing.assert_allclose(d["qflx_sat_excess_surf"], [fs * 2.0, fs * 2.0 + 0.5], rtol=1e-15)


# ---- infiltration excess (test_infilt_excess_runoff.jl) ----

def test_qinmax_from_hksat_and_ice():
    ice = np.zeros((15, 3))
    ice[:3, 0] = [0.1, 0.2, 0.3]
    ice[:3, 2] = 0.5
    hk = np.full((15, 3), 0.01)
    hk[:3, 0] = [0.01, 0.02, 0.03]
    hk[:3, 1] = 0.05
    d = run(cols(3, icefrac=ice, hksat=hk, fsat=[0.0] * 3, qflx_in_soil=[0.0] * 3),
            step="infil_excess", e_ice=6.0)
    e1 = min(10.0 ** (-0.6) * 0.01, 10.0 ** (-1.2) * 0.02, 10.0 ** (-1.8) * 0.03)
    np.testing.assert_allclose(d["qinmax"], [e1, 0.05, 10.0 ** (-3.0) * 0.01], rtol=1e-14)


def test_infiltration_excess_methods_and_mask():
    d = run(cols(2, fsat=[0.1, 0.2], qflx_in_soil=[0.02, 0.005]), step="infil_excess")
    np.testing.assert_allclose(d["qinmax"], [0.9 * 0.01, 0.8 * 0.01], rtol=1e-15)
    np.testing.assert_allclose(d["qflx_infl_excess"], [0.02 - 0.009, 0.0], atol=1e-17)
    d = run(cols(2, fsat=[0.0, 0.0], qflx_in_soil=[100.0, 200.0]), step="infil_excess",
            qinmax_method=0)
    assert np.all(d["qinmax"] == 1.0e200) and np.all(d["qflx_infl_excess"] == 0.0)
    d = run(cols(3, mask_hydrologyc=[0, 1, 0], fsat=[0.0] * 3, qflx_in_soil=[0.02] * 3),
            step="infil_excess")
    assert np.isnan(d["qinmax"][0]) and np.isnan(d["qinmax"][2])
    assert abs(d["qinmax"][1] - 0.01) < 1e-15
    # ponded water shields its area: only (1 - frac_h2osfc) of qinmax applies
    d = run(cols(1, fsat=[0.0], qflx_in_soil=[0.02], frac_h2osfc=[0.5]), step="infil_excess")
    assert abs(d["qflx_infl_excess"][0] - (0.02 - 0.5 * 0.01)) < 1e-17


# ---- flood water, infiltration, the surface water store (test_soil_hydrology_mod.jl) ----

def test_floodc_and_infiltration():
    d = run(cols(4, qflx_floodg=[0.5, 0.3], col_gridcell=[1, 1, 2, 2],
                 col_itype=[1, SUNWALL, 1, SHADEWALL]), step="floodc")
    np.testing.assert_array_equal(d["qflx_floodc"], [0.5, 0.0, 0.3, 0.0])
    d = run(cols(3, qflx_in_soil_limited=[0.1, 0.2, 0.3], qflx_h2osfc_drain=[0.01, 0.02, 0.03]),
            step="infiltration")
    np.testing.assert_allclose(d["qflx_infl"], [0.11, 0.22, 0.33], rtol=1e-15)


def test_surface_water_truncation_is_relative():
    """A 5e-11 mm remainder of a 1 mm store survives, a 5e-14 mm one is set to
    0 (test_soil_hydrology_mod.jl:152-172; the drainage update uses the same
    rule as the runoff update tested there)."""
    for rest, survives in ((5.0e-11, True), (5.0e-14, False)):
        d = run(cols(1, h2osfc=[1.0], h2osfc_thresh=[10.0], qflx_in_h2osfc=[0.0],
                     frac_h2osfc=[1.0], qinmax=[(1.0 - rest) / 1800.0]), step="h2osfc")
        if survives:
            assert d["h2osfc"][0] > 0.0 and abs(d["h2osfc"][0] - rest) < 1e-15
        else:
            assert d["h2osfc"][0] == 0.0


def test_urban_ponding_and_runoff():
    """test_soil_hydrology_mod.jl:495-590: roof and impervious road, no snow."""
    dt = 3600.0
    ns = 12

    def urban(rain, evap=0.0, pond=(0.0, 0.0)):
        d = cols(2, mask_hydrologyc=[0, 0], mask_urbanc=[1, 1], col_itype=[ROOF, ROAD_IMPERV],
                 qflx_rain_plus_snomelt=[rain] * 2, qflx_liqevap_from_top_layer=[evap] * 2,
                 xs_urban=np.zeros(2))
        d["h2osoi_liq"][ns] = pond
        oracle.surface_hydrology(d, dt, steps=SH_STEPS["surf_runoff"] | SH_STEPS["urban_pond"])
        return d

    d = urban(1.0e-4)
    assert np.all(np.abs(d["xs_urban"]) < 1e-12) and np.all(np.abs(d["qflx_surf"]) < 1e-12)
    np.testing.assert_allclose(d["h2osoi_liq"][ns], 1.0e-4 * dt, rtol=1e-15)
    assert np.all(d["h2osoi_liq"][ns - 1] == 0.0)
    d = urban(1.0e-2)
    np.testing.assert_allclose(d["xs_urban"], 1.0e-2 - 1.0 / dt, rtol=1e-10)
    np.testing.assert_allclose(d["qflx_surf"], 1.0e-2 - 1.0 / dt, rtol=1e-10)
    assert np.all(d["h2osoi_liq"][ns] == 1.0)
    pond0 = np.array([0.2, 0.5])
    d = urban(5.0e-4, 1.0e-4, pond0)
    np.testing.assert_allclose(d["h2osoi_liq"][ns] - pond0,
                               (5.0e-4 - 1.0e-4 - d["qflx_surf"]) * dt, atol=1e-10)


# ---- the whole sequence ----

def test_surface_water_balance_closes():
    d = inputs.surface_hydrology_columns(nc=6000, seed=62)
    out = clone(d)
    dt = 1800.0
    oracle.surface_hydrology(out, dt)
    soil = (d["mask_hydrologyc"] == 1) & ((d["lun_itype"] == 1) | (d["lun_itype"] == 2))
    assert soil.sum() > 3000
    fh = d["frac_h2osfc"]
    nosnow = d["snl"] >= 0
    fsno = np.where(nosnow, 0.0, d["frac_sno_eff"])
    evap = np.where(nosnow, d["qflx_liqevap_from_top_layer"], d["qflx_ev_soil"])
    arriving = d["qflx_rain_plus_snomelt"] + d["qflx_snow_h2osfc"] + out["qflx_floodc"]
    expect = arriving - (1.0 - fsno - fh) * evap - fh * d["qflx_ev_h2osfc"]
    got = out["qflx_infl"] + out["qflx_surf"] + (out["h2osfc"] - d["h2osfc"]) / dt
    scale = np.abs(arriving) + np.abs(evap) + np.abs(d["h2osfc"]) / dt + 1e-12
    err = np.abs(got - expect) / scale
    assert err[soil].max() < 1e-12, err[soil].max()
    # the pieces are where the reference puts them
    assert np.all(out["qflx_infl_excess_surf"][soil] == 0.0)          # h2osfcflag 1
    assert np.all(out["qflx_surf"][d["col_itype"] == SUNWALL] == 0.0)
    lake = d["mask_nolakec"] == 0
    assert np.all(out["qflx_floodc"][lake] == 7.0)                    # untouched
    assert np.sum(out["qflx_infl_excess"][soil] > 0) > 300
    assert np.sum(out["qflx_h2osfc_surf"][soil] > 0) > 300


def test_unknown_method_rejected():
    d = cols(2)
    with pytest.raises(RuntimeError):
        oracle.surface_hydrology(d, 1800.0, fsat_method=3)
    with pytest.raises(RuntimeError):
        oracle.surface_hydrology(d, 1800.0, steps=0x200)


# ---- root water uptake (test_soil_water_plant_sink.jl) ----

def sink(rootr, tran_patch, tran_col, wtcol, active, npatches, nlevsoi=4, mask=None):
    """Columns in order, their patches contiguous (the reference test's layout)."""
    rootr = np.asarray(rootr, dtype=float)
    np_, nc = rootr.shape[0], len(npatches)
    rp = np.zeros((10, np_))
    rp[:nlevsoi] = rootr.T
    return dict(nc=nc, np=np_, nlevsno=12, nlevsoi=nlevsoi, nlevgrnd=10,
                mask_hydrologyc=np.ones(nc, np.uint8) if mask is None else np.asarray(mask, np.uint8),
                patchi=(np.concatenate([[0], np.cumsum(npatches)[:-1]]) + 1).astype(np.int64),
                npatches=np.asarray(npatches, np.int64), active=np.asarray(active, np.uint8),
                wtcol=np.asarray(wtcol, float), rootr_patch=rp,
                qflx_tran_veg_patch=np.asarray(tran_patch, float),
                qflx_tran_veg_col=np.asarray(tran_col, float),
                rootr_col=np.full((10, nc), np.nan), qflx_rootsoi_col=np.full((nlevsoi, nc), np.nan))


def test_root_sink_reference_kats():
    r = [[0.4, 0.3, 0.2, 0.1]]
    d = sink(r, [0.002], [0.002], [1.0], [1], [1])
    oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
    np.testing.assert_allclose(d["rootr_col"][:4, 0], r[0], atol=1e-12)
    np.testing.assert_allclose(d["qflx_rootsoi_col"][:, 0], np.array(r[0]) * 0.002, atol=1e-14)
    r2 = [[0.1, 0.2, 0.3, 0.4], [0.5, 0.3, 0.15, 0.05]]
    tp, wt = [0.001, 0.003], [0.4, 0.6]
    tc = tp[0] * wt[0] + tp[1] * wt[1]
    d = sink(r2, tp, [tc], wt, [1, 1], [2])
    oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
    exp = [(r2[0][j] * tp[0] * wt[0] + r2[1][j] * tp[1] * wt[1]) / tc for j in range(4)]
    np.testing.assert_allclose(d["rootr_col"][:4, 0], exp, atol=1e-12)
    np.testing.assert_allclose(d["qflx_rootsoi_col"][:, 0], np.array(exp) * tc, atol=1e-14)
    d = sink([[0.25] * 4], [0.0], [0.0], [1.0], [1], [1])          # zero transpiration
    oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
    assert np.all(d["rootr_col"][:4, 0] == 0.0) and np.all(d["qflx_rootsoi_col"][:, 0] == 0.0)
    d = sink([[0.4, 0.3, 0.2, 0.1], [0.1, 0.1, 0.1, 0.7]], [0.002, 0.005], [0.001], [0.5, 0.5],
             [1, 0], [2])                                          # inactive patch skipped
    oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
    np.testing.assert_allclose(d["rootr_col"][:4, 0], [0.4, 0.3, 0.2, 0.1], atol=1e-12)
    r3 = [[0.25] * 4, [0.4, 0.3, 0.2, 0.1], [0.1, 0.2, 0.3, 0.4]]  # road, crop, natural veg
    d = sink(r3, [0.001, 0.002, 0.003], [0.001, 0.002, 0.003], [1.0] * 3, [1] * 3, [1, 1, 1])
    oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
    np.testing.assert_allclose(d["rootr_col"][:4], np.array(r3).T, atol=1e-12)
    np.testing.assert_allclose(d["qflx_rootsoi_col"], np.array(r3).T * [0.001, 0.002, 0.003],
                               atol=1e-14)


def test_root_sink_distributes_the_transpiration():
    """Each patch's root profile sums to 1 over the soil layers, so the
    transpiration-weighted column profile does too, and the root sink sums to
    the column transpiration; masked-off columns stay untouched."""
    for scatter in (False, True):
        d = inputs.root_sink_columns(nc=4000, seed=82, scatter=scatter)
        oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
        m = d["mask_hydrologyc"] == 1
        live = np.zeros(d["nc"], bool)
        for c in np.flatnonzero(m):
            p = slice(d["patchi"][c] - 1, d["patchi"][c] - 1 + d["npatches"][c])
            live[c] = np.sum(d["qflx_tran_veg_patch"][p] * d["wtcol"][p] * d["active"][p]) != 0
        n = d["nlevsoi"]
        np.testing.assert_allclose(d["rootr_col"][:n, live].sum(0), 1.0, rtol=1e-14)
        assert np.all(d["rootr_col"][:n, m & ~live] == 0.0)
        np.testing.assert_allclose(d["qflx_rootsoi_col"][:, m].sum(0), d["qflx_tran_veg_col"][m],
                                   rtol=1e-13, atol=1e-19)
        assert np.all(d["rootr_col"][:, ~m] == 7.0) and np.all(d["rootr_col"][n:] == 7.0)
        assert live.sum() > 2000 and (m & ~live).sum() > 50


def test_root_sink_bad_patch_range():
    d = inputs.root_sink_columns(nc=64, seed=83)
    d["mask_hydrologyc"][:] = 1
    d["npatches"][5] = d["np"] + 1
    with pytest.raises(RuntimeError):
        oracle.compute_effec_rootfrac_and_vert_tran_sink(d)
