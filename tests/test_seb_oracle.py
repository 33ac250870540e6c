"""Pins of the oracle's soil energy balance error (orc_soil_energy_balance,
following soil_fluxes.jl:380-447 and 499-525) — and, through it, of the heat
step's energy conservation.

The reference has no test for errsoi. Physics fixes it instead: the heat step
solves, layer by layer, (T_new - T_old)/fact = net conductive flux + boundary
flux (Crank-Nicolson interior fluxes, implicit ground flux hs + dhsdT*dT), and
converts any heat that crosses freezing into the phase-change terms xmf and
xmf_h2osfc. Summing over the layers, the interior fluxes cancel, so with the
ground heat flux the step actually applied (area-weighted over snow, bare soil
and surface water) the error is zero up to rounding. The test chains the heat
oracle into the energy balance oracle on columns with partial snow cover,
surface water and phase change in most layers. A wrong sign, layer index, snow
weighting or phase-change term in either restatement breaks the closure; the
negative controls show the test can see such a mistake.
"""
import numpy as np

import oracle
from _util import clone
from clmsoil import bind, inputs


def heat_then_balance(nc, seed, snow_weight=None):
    d = inputs.heat_config1(nc=nc, seed=seed, h2osfc_variant=True)
    d.update(inputs.heat_outputs(d))
    rng = np.random.default_rng(seed)
    d["frac_sno_eff"] = np.where(d["snl"] < 0, rng.uniform(0.3, 0.7, nc), 0.0)
    t0, th0 = d["t_soisno"].copy(), d["t_h2osfc"].copy()
    h = clone(d)
    oracle.soil_temperature(h, bind.heat_config())
    ns, cols = d["nlevsno"], np.arange(nc)
    top = d["snl"] + ns
    dt_top = h["t_soisno"][top, cols] - t0[top, cols]
    dt_1 = h["t_soisno"][ns] - t0[ns]
    dt_h = h["t_h2osfc"] - th0
    fs, fw, g = d["frac_sno_eff"], d["frac_h2osfc"], d["dhsdT"]
    # the ground heat flux the step applied, by surface fraction
    eflx = (fs * (d["hs_top_snow"] + g * dt_top) + (1.0 - fs - fw) * (d["hs_soil"] + g * dt_1) +
            fw * (d["hs_h2osfc"] + g * dt_h))
    one = np.arange(1, nc + 1, dtype=np.int64)
    seb = dict(nc=nc, np=nc, nlevsno=ns, nlevsoi=d["nlevsoi"], nlevgrnd=d["nlevgrnd"], nlevurb=0,
               patchi=one, patchf=one.copy(), patch_column=one.copy(),
               col_itype=np.ones(nc, dtype=np.int64), snl=d["snl"],
               mask_nolakec=np.ones(nc, np.uint8), mask_nolakep=np.ones(nc, np.uint8),
               mask_urbanp=np.zeros(nc, np.uint8), eflx_soil_grnd=eflx, wtcol=np.ones(nc),
               xmf=h["xmf"], xmf_h2osfc=h["xmf_h2osfc"], frac_h2osfc=h["frac_h2osfc"],
               t_h2osfc=h["t_h2osfc"], t_h2osfc_bef=th0, c_h2osfc=h["c_h2osfc"],
               eflx_h2osfc_to_snow=h["eflx_h2osfc_to_snow"],
               eflx_building_heat_errsoi=np.zeros(nc),
               frac_sno_eff=fs if snow_weight is None else np.full(nc, snow_weight),
               t_soisno=np.nan_to_num(h["t_soisno"]), t_ssbef=np.nan_to_num(t0),
               fact=np.where(np.isnan(h["fact"]), 1.0, h["fact"]),
               errsoi_patch=np.zeros(nc), errsoi_col=np.zeros(nc))
    oracle.soil_energy_balance(seb, d["dtime"])
    return seb, h, d


def test_heat_step_conserves_energy():
    seb, h, d = heat_then_balance(3000, 11)
    assert np.sum(h["xmf"] != 0.0) > 2000              # phase change in most columns
    assert np.sum(h["xmf_h2osfc"] != 0.0) > 500
    assert np.sum((d["snl"] < 0) & (d["frac_h2osfc"] > 0)) > 2000
    assert np.max(np.abs(seb["errsoi_col"])) < 1e-8        # W/m2; fluxes are O(1e2..1e6)
    assert np.array_equal(seb["errsoi_col"], seb["errsoi_patch"])   # one patch, weight 1


def test_closure_negative_controls():
    # snow layers counted at full area instead of frac_sno_eff
    seb, _, d = heat_then_balance(1000, 12, snow_weight=1.0)
    snow = d["snl"] < 0
    assert np.median(np.abs(seb["errsoi_col"][snow])) > 1e-3
    # the stored heat of one layer dropped (fact of that layer set huge)
    seb, h, d = heat_then_balance(1000, 13)
    e0 = seb["errsoi_col"].copy()
    seb["fact"][d["nlevsno"] + 3] = 1e30
    oracle.soil_energy_balance(seb, d["dtime"])
    assert np.median(np.abs(seb["errsoi_col"] - e0)) > 1e-3


def test_patch_average_and_urban_terms():
    d = inputs.energy_balance_inputs(400, 14)
    nc, np_, ns, ng, nu = d["nc"], d["np"], d["nlevsno"], d["nlevgrnd"], d["nlevurb"]
    col, itype, mask_p, mask_u = d["patch_column"], d["col_itype"], d["mask_nolakep"], d["mask_urbanp"]
    wall = np.isin(itype, (71, 72, 73))
    prior = d["errsoi_patch"].copy()
    out = clone(d)
    oracle.soil_energy_balance(out, 1800.0)
    c = col - 1
    # closed form per patch
    base = (d["eflx_soil_grnd"] - d["xmf"][c] - d["xmf_h2osfc"][c] -
            d["frac_h2osfc"][c] * (d["t_h2osfc"][c] - d["t_h2osfc_bef"][c]) * (d["c_h2osfc"][c] / 1800.0)
            + d["eflx_h2osfc_to_snow"][c] + np.where(wall[c], d["eflx_building_heat_errsoi"][c], 0.0))
    jf = np.arange(-ns + 1, max(ng, nu) + 1)[:, None]
    stored = (d["t_soisno"] - d["t_ssbef"]) / d["fact"]
    snow_w = np.where((jf >= d["snl"][None, :] + 1) & (jf < 1), d["frac_sno_eff"][None, :], 0.0)
    soil_w = (jf >= 1).astype(float)
    lim_nonurban = (jf <= ng).astype(float)
    lim_urban = (jf <= nu).astype(float)
    per_col_nonurban = ((snow_w + soil_w) * lim_nonurban * stored).sum(0)
    per_col_urban = ((snow_w + soil_w) * lim_urban * stored).sum(0)
    expect = np.where(mask_p == 1, base, prior)
    expect -= np.where((mask_p == 1) & ~wall[c], per_col_nonurban[c], 0.0)
    expect -= np.where((mask_u == 1) & wall[c], per_col_urban[c], 0.0)
    np.testing.assert_allclose(out["errsoi_patch"], expect, rtol=1e-9, atol=1e-9)
    untouched = (mask_p == 0) & ~((mask_u == 1) & wall[c])
    assert np.array_equal(out["errsoi_patch"][untouched], prior[untouched])
    # column average: every patch in the range, masked or not, weighted
    colsum = np.bincount(c, weights=out["errsoi_patch"] * d["wtcol"], minlength=nc)
    on = d["mask_nolakec"] == 1
    np.testing.assert_allclose(out["errsoi_col"][on], colsum[on], rtol=1e-12, atol=1e-9)
    assert np.all(out["errsoi_col"][~on] == -7.0)      # unmasked columns are not written


def test_ground_flux_heat_step_energy_chain():
    """The ground heat flux of real patch forcing (several patches per column,
    absorbed solar per snow layer) into the heat step, and the energy balance
    over the result: the chain closes too."""
    h, g = inputs.coupled_surface_columns(2000, 15)
    oracle.ground_heat_flux(g)
    for k in ("hs_top_snow", "hs_soil", "hs_top", "hs_h2osfc", "dhsdT", "sabg_lyr_col"):
        h[k] = g[k].copy()
    assert np.max(g["sabg_lyr_col"]) > 100.0
    h1 = clone(h)
    oracle.soil_temperature(h1, bind.heat_config())
    seb = inputs.energy_balance_from_heat(h, h1)
    oracle.soil_energy_balance(seb, h["dtime"])
    assert np.max(np.abs(seb["errsoi_col"])) < 1e-8
