"""Pins of the oracle's ground heat flux (orc_ground_heat_flux, following
compute_ground_heat_flux_and_deriv!, soil_temperature.jl:944-1263).

The reference's own test of this function (test_soil_temperature.jl:307-333) only
asks for finite outputs and dhsdT < 0; both are checked here. Beyond that the
oracle is pinned by what the physics and arithmetic fix:

* Julia's x^4 (compensated power by squaring) against the correctly rounded
  fourth power, computed with exact rationals;
* radiative equilibrium: a bare patch with no turbulent flux under downwelling
  longwave sigma*T^4 has zero net ground heat flux;
* the derivative: dhsdT = -sum_p wt_p*cgrnd_p - 4*emg*sigma*T^3 when the weights
  sum to one;
* structural identities: with the snow-side inputs equal to the soil-side inputs,
  a column without snow gets hs_top_snow == hs_soil bit for bit (the top layer is
  soil layer 1), and the per-layer absorbed solar sums to the weighted patch sums;
* the zero-weight guard of the reference (NaN fluxes on weight-0 patches do not
  reach hs_soil/hs_h2osfc/dhsdT; the SNICAR sums have no such guard);
* urban patch terms (waste heat, A/C, traffic spread over the non-roof area)
  and the building-temperature switches.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs

SB = 5.67e-8
NS = inputs.NLEVSNO


def run(d):
    out = clone(d)
    oracle.ground_heat_flux(out)
    return out


def test_jl_pow4_is_compensated():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(150.0, 350.0, 3000), rng.uniform(1e-3, 1.0, 500),
                         10.0 ** rng.uniform(-70, 70, 500)])
    exact_match = naive_match = 0
    for x in xs:
        cr = float(Fraction(float(x)) ** 4)          # correctly rounded x^4
        got = oracle.jl_pow4(x)
        assert abs(got - cr) <= np.spacing(cr)
        exact_match += got == cr
        naive_match += (x * x) * (x * x) == cr
    assert exact_match >= 0.999 * xs.size
    assert naive_match < 0.9 * xs.size      # the compensation is doing something
    assert oracle.jl_pow4(0.0) == 0.0 and oracle.jl_pow4(1e200) == np.inf


def single_patch_columns(nc, seed):
    d = inputs.ground_heat_flux_inputs(nc=nc, seed=seed, max_patches=1)
    keep = d["col_itype"] <= 2                  # soil and crop columns only
    d["mask_nolakec"][:] = keep
    d["mask_nolakep"][:] = keep[d["patch_column"] - 1]
    return d


def test_reference_smoke_properties():
    d = inputs.ground_heat_flux_inputs(nc=2000, seed=2)
    out = run(d)
    on = d["mask_nolakec"] == 1
    for k in ("hs_top", "hs_top_snow", "hs_soil", "hs_h2osfc", "dhsdT"):
        assert np.all(np.isfinite(out[k]))
    assert np.all(out["dhsdT"][on] < 0.0)
    p_on = d["mask_nolakep"] == 1
    assert np.all(np.isfinite(out["eflx_gnet"][p_on]))
    # lake columns: every sum stays zero
    for k in ("hs_top", "hs_top_snow", "hs_soil", "hs_h2osfc", "dhsdT"):
        assert np.all(out[k][~on] == 0.0)
    assert np.all(out["sabg_lyr_col"][:, ~on] == 0.0)


def test_radiative_equilibrium():
    d = single_patch_columns(500, 3)
    rural = d["mask_nolakep"] == 1
    for k in ("sabg", "dlrad", "eflx_sh_grnd", "qflx_evap_soi"):
        d[k][:] = 0.0
    d["frac_veg_nosno"][:] = 0
    d["emg"][:] = 1.0
    d["forc_lwrad"] = SB * d["t_grnd"] ** 4
    out = run(d)
    g = out["eflx_gnet"][rural]
    scale = SB * d["t_grnd"][d["patch_column"][rural] - 1] ** 4
    assert np.max(np.abs(g) / scale) < 1e-14
    # control: 1 % more longwave in is 1 % of sigma T^4 of net flux
    d["forc_lwrad"] *= 1.01
    g2 = run(d)["eflx_gnet"][rural]
    np.testing.assert_allclose(g2 / scale, 0.01, rtol=1e-12)


def test_derivative_closed_form():
    d = inputs.ground_heat_flux_inputs(nc=3000, seed=4)
    out = run(d)
    on = (d["mask_nolakec"] == 1)
    pc = d["patch_column"] - 1
    rural_p = (d["mask_nolakep"] == 1) & (d["urbpoi"][d["patch_landunit"] - 1] == 0)
    wc = np.bincount(pc[rural_p], weights=(d["wtcol"] * d["cgrnd"])[rural_p], minlength=d["nc"])
    urban_p = (d["mask_nolakep"] == 1) & ~rural_p
    wc += np.bincount(pc[urban_p], weights=(d["wtcol"] * d["cgrnd"])[urban_p], minlength=d["nc"])
    expect = -wc - 4.0 * d["emg"] * SB * d["t_grnd"] ** 3
    np.testing.assert_allclose(out["dhsdT"][on], expect[on], rtol=1e-13)


def test_snow_side_equals_soil_side_without_snow():
    d = inputs.ground_heat_flux_inputs(nc=2000, seed=5, zero_weight_fraction=0.0)
    d["snl"][:] = 0
    d["sabg_lyr_patch"][NS] = d["sabg_soil"]          # layer 1 (index lyr_top + joff)
    d["eflx_sh_snow"] = d["eflx_sh_soil"].copy()
    d["qflx_ev_snow"] = d["qflx_ev_soil"].copy()
    out = run(d)
    rural_c = (d["mask_nolakec"] == 1) & (d["col_itype"] <= 2)
    assert rural_c.sum() > 1000
    assert np.array_equal(out["hs_top_snow"][rural_c], out["hs_soil"][rural_c])
    # control: hs_top uses the grid-cell ground temperature, not layer 1
    assert not np.array_equal(out["hs_top"][rural_c], out["hs_soil"][rural_c])


def test_absorbed_solar_per_layer():
    d = inputs.ground_heat_flux_inputs(nc=2000, seed=6)
    out = run(d)
    pc = d["patch_column"] - 1
    lyr_top = d["snl"][pc] + 1
    on = d["mask_nolakep"] == 1
    rural = d["urbpoi"][d["patch_landunit"] - 1] == 0
    layers = np.arange(-NS + 1, 2)[:, None]                    # j = -nlevsno+1 .. 1
    take = (layers >= lyr_top[None, :]) & (on & rural)[None, :]
    per_patch = np.where(take, d["sabg_lyr_patch"] * d["wtcol"][None, :], 0.0).sum(0)
    urb = on & ~rural
    per_patch[urb] = d["sabg"][urb] * d["wtcol"][urb]
    expect = np.bincount(pc, weights=per_patch, minlength=d["nc"])
    np.testing.assert_allclose(out["sabg_lyr_col"].sum(0), expect, rtol=1e-12, atol=1e-9)
    # layers above the column's snow pack stay zero
    col_layers = np.arange(-NS + 1, 2)[:, None]
    above = col_layers < (d["snl"] + 1)[None, :]
    assert np.all(out["sabg_lyr_col"][above] == 0.0)


def test_zero_weight_guard():
    d = inputs.ground_heat_flux_inputs(nc=1000, seed=7, zero_weight_fraction=0.3)
    zw = (d["wtcol"] == 0.0) & (d["mask_nolakep"] == 1)
    assert zw.sum() > 50
    for k in ("eflx_sh_grnd", "eflx_sh_soil", "eflx_sh_h2osfc", "cgrnd"):
        d[k][zw] = np.nan
    out = run(d)
    on = d["mask_nolakec"] == 1
    for k in ("hs_soil", "hs_h2osfc", "dhsdT"):
        assert np.all(np.isfinite(out[k][on])), k
    # the SNICAR top-layer sum multiplies without a guard, as in the reference
    cols = np.unique(d["patch_column"][zw] - 1)
    assert np.all(np.isnan(out["hs_top"][cols]))


@pytest.mark.parametrize("method", [0, 1, 2])
def test_urban_terms(method):
    d = inputs.ground_heat_flux_inputs(nc=3000, seed=8, building_temp_method=method)
    out = run(d)
    l = d["patch_landunit"] - 1
    c = d["patch_column"] - 1
    urb = (d["mask_nolakep"] == 1) & (d["urbpoi"][l] == 1)
    road = urb & np.isin(d["col_itype"][c], (inputs.ICOL_ROAD_IMPERV, inputs.ICOL_ROAD_PERV))
    other = urb & ~road
    assert road.sum() > 50 and other.sum() > 50
    denom = 1.0 - d["wtlunit_roof"][l]
    np.testing.assert_array_equal(out["eflx_wasteheat"][road], (d["eflx_wasteheat_lun"][l] / denom)[road])
    np.testing.assert_array_equal(out["eflx_traffic"][road], (d["eflx_traffic_lun"][l] / denom)[road])
    assert np.all(out["eflx_wasteheat"][other] == 0.0) and np.all(out["eflx_traffic"][other] == 0.0)
    if method == 0:
        assert np.all(out["eflx_ventilation"][urb] == 0.0)
        np.testing.assert_array_equal(out["eflx_anthro"][urb],
                                      (out["eflx_wasteheat"] + out["eflx_traffic"])[urb])
    elif method == 1:
        np.testing.assert_array_equal(out["eflx_ventilation"][road],
                                      (d["eflx_ventilation_lun"][l] / denom)[road])
        assert np.all(out["eflx_anthro"][urb] == 0.0)       # only written by the simple method
    else:
        np.testing.assert_array_equal(out["eflx_ventilation"][road], d["eflx_ventilation"][road])
    # urban patches put the same net flux into every column sum
    uc = np.unique(c[urb])
    np.testing.assert_array_equal(out["hs_soil"][uc], out["hs_h2osfc"][uc])
