"""Pins of the oracle's methane and oxygen transport (orc_ch4_tran, following
ch4_tran!, methane.jl:101-129 and 1663-2092).

The reference's tests cover only the parameter tables and the full driver
(test/test_methane.jl); the transport step is pinned by what the physics and
the mathematics fix:
* the discrete CH4 budget closes before the step's own balance correction: the
  column's change equals production minus oxidation minus the aerenchyma,
  ebullition and diffusive surface fluxes, to rounding (both fractions, water
  tables at every level, ponding, frozen levels, negative-CH4 clips);
* the new relative concentrations solve the assembled rows (dense solve);
* a column at atmospheric equilibrium without sources stays there;
* Henry's-law solubilities at 25 C are the tabulated ones (CH4 0.0342, O2
  0.0318 dimensionless) and rise as the water cools;
* the O2/CH4 competition rules on hand-computed single-level cases;
* Julia's Float64 ^ with integer-valued exponents (x^2.0 is x*x);
* the gas and aqueous diffusivities (orc_ch4_diffusivity, the transport's own
  routine) on states where the reference's closed form reduces to a number fixed
  by hand: the organic term at f_a = 1 and 1/8, the mineral term at b = 3 and 1.5,
  the organic weighting, the temperature polynomials of both species, the scale
  factors, the frozen liquid share and the 1e-12 floor.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs

FIELDS_2D = ("watsat", "h2osoi_vol", "h2osoi_liq", "h2osoi_ice", "bsw", "cellorg", "t_soisno", "dz",
             "ch4_prod_depth", "o2_aere_depth", "o2_decomp_depth", "o2_oxid_depth", "ch4_oxid_depth",
             "ch4_aere_depth", "ch4_ebul_depth", "o2stress", "ch4stress", "conc_ch4", "conc_o2")


def run(d, dtime=1800.0, dbg=False, **kw):
    out = clone(d)
    kw.setdefault("nlevsoi", d["nlevsoi"])
    r = oracle.ch4_tran(out, dtime, dbg=dbg, **kw)
    return (out, r) if dbg else out


@pytest.mark.parametrize("sat,seed", [(0, 131), (0, 132), (1, 133)])
def test_ch4_budget_closes_before_correction(sat, seed):
    d = inputs.ch4_columns(nc=3000, seed=seed, sat=sat)
    out, dbg = run(d, dbg=True, sat=sat)
    m = d["mask"].astype(bool)
    dz, dt = d["dz"], 1800.0
    # the budget from the outputs, with the diffusive flux before the correction
    # (the correction is applied exactly when |err| < 1e-8)
    err = dbg["err_raw"]
    assert np.all(np.abs(err[m]) < 1e-8)
    raw_diff = out["ch4_surf_diff"] + err / dt
    budget = ((out["conc_ch4"] - d["conc_ch4"]) * dz).sum(0) \
        - (out["ch4_prod_depth"] * dz).sum(0) * dt + (out["ch4_oxid_depth"] * dz).sum(0) * dt \
        + (out["ch4_surf_aere"] + out["ch4_surf_ebul"] + raw_diff) * dt
    scale = (np.abs(out["conc_ch4"]) * dz).sum(0) + (np.abs(d["conc_ch4"]) * dz).sum(0) \
        + (d["ch4_prod_depth"] * dz).sum(0) * dt
    assert np.all(np.abs(budget[m]) <= 1e-13 * scale[m]), np.max(np.abs(budget[m]) / scale[m])
    # the corrected budget closes to rounding of the correction itself
    assert np.all(np.abs(err[m]) <= 1e-13 * scale[m])
    # columns outside the filter are untouched
    for k in FIELDS_2D:
        assert np.array_equal(out[k][:, ~m], d[k][:, ~m]), k
    # surface ebullition only without a water table in the column
    jw = d["jwt"]
    assert np.all(out["ch4_surf_ebul"][m & (jw != 0)] == 0.0)
    assert np.array_equal(out["ch4_surf_ebul"][m & (jw == 0)], out["ch4_ebul_total"][m & (jw == 0)])


def test_rows_are_solved():
    d = inputs.ch4_columns(nc=400, seed=134)
    _, dbg = run(d, dbg=True)
    m = np.flatnonzero(d["mask"])
    for c in m[:120]:
        for s in range(2):
            r = dbg["rows"][c, s]
            n = r.shape[0]
            A = np.diag(r[:, 1]) + np.diag(r[1:, 0], -1) + np.diag(r[:-1, 2], 1)
            x = np.linalg.solve(A, r[:, 3])
            np.testing.assert_allclose(r[:, 4], x, rtol=1e-11, atol=1e-14 * np.abs(x).max())
            assert n == d["nlevsoi"] + 1


@pytest.mark.parametrize("sat", [0, 1])
def test_atmospheric_equilibrium_without_sources(sat):
    """Relative concentrations equal to the atmosphere's (with the ground node's
    Henry partition where the top level is below the water table), no production,
    oxidation, aerenchyma or ebullition: nothing moves."""
    d = inputs.ch4_columns(nc=600, seed=135, sat=sat)
    nl = d["nlevsoi"]
    if sat == 0:
        d["jwt"][:] = nl                                  # the whole column unsaturated
    for k in ("ch4_prod_depth", "o2_aere_depth", "o2_decomp_depth", "o2_oxid_depth",
              "ch4_oxid_depth", "ch4_aere_depth", "ch4_ebul_depth"):
        d[k][:] = 0.0
    d["h2osoi_vol"] = np.minimum(d["h2osoi_vol"], d["watsat"])
    _, dbg = run(d, dbg=True, sat=sat)                    # k_h_cc of these temperatures
    g = d["col_gridcell"] - 1
    catm = d["c_atm"][:, g]                               # (3, nc)
    k = dbg["khcc"]                                       # (nc, 2, nl+1)
    liq = 1.0
    eps = np.empty((2, nl, d["nc"]))
    for s in range(2):
        if sat == 0:
            eps[s] = d["watsat"] - (1.0 - k[:, s, 1:].T) * d["h2osoi_vol"] * liq
        else:
            eps[s] = d["watsat"] * liq
    rel = [catm[s] * (1.0 if sat == 0 else k[:, s, 0]) for s in range(2)]
    d["conc_ch4"] = rel[0][None, :] * eps[0]
    d["conc_o2"] = rel[1][None, :] * eps[1]
    out = run(d, sat=sat)
    m = d["mask"].astype(bool)
    np.testing.assert_allclose(out["conc_ch4"][:, m], d["conc_ch4"][:, m], rtol=1e-12)
    np.testing.assert_allclose(out["conc_o2"][:, m], d["conc_o2"][:, m], rtol=1e-12)
    flux_scale = (d["conc_ch4"] * d["dz"]).sum(0) / 1800.0
    assert np.all(np.abs(out["ch4_surf_diff"][m]) <= 1e-12 * flux_scale[m])


def test_henry_solubility():
    d = inputs.ch4_columns(nc=64, seed=136)
    d["t_grnd"][:] = 298.15
    d["t_soisno"][:] = np.linspace(273.15, 303.15, d["nlevsoi"])[:, None]
    _, dbg = run(d, dbg=True)
    k = dbg["khcc"][0]
    assert abs(k[0, 0] / 0.0342 - 1.0) < 0.01          # CH4 at 25 C
    assert abs(k[1, 0] / 0.0318 - 1.0) < 0.01          # O2 at 25 C
    assert np.all(np.diff(k[:, 1:], axis=1) < 0.0)       # less soluble as the water warms
    assert 1.4 < k[0, 1] / k[0, -1] < 2.0                # 0 C vs 30 C


def single_level(conc_ch4, conc_o2, prod, ch4_oxid, aere, ebul, o2_decomp, o2_aere=0.0, dtime=1.0):
    d = inputs.ch4_columns(nc=1, seed=137, nlevsoi=1, sat=1)
    d["mask"][:] = 1
    for k, v in dict(conc_ch4=conc_ch4, conc_o2=conc_o2, ch4_prod_depth=prod,
                     ch4_oxid_depth=ch4_oxid, o2_oxid_depth=2.0 * ch4_oxid, ch4_aere_depth=aere,
                     ch4_ebul_depth=ebul, o2_decomp_depth=o2_decomp, o2_aere_depth=o2_aere).items():
        d[k][:] = v
    return run(d, dtime=dtime, nlevsoi=1, sat=1)


def test_competition_rules():
    # ample supply: no stress, rates unchanged
    o = single_level(1.0, 10.0, 0.0, 0.1, 0.1, 0.1, 1.0)
    assert o["ch4stress"][0, 0] == 1.0 and o["o2stress"][0, 0] == 1.0
    assert o["ch4_oxid_depth"][0, 0] == 0.1 and o["o2_decomp_depth"][0, 0] == 1.0
    # CH4 limited: supply 0.15 of demand 0.3 -> everything CH4-consuming halves
    o = single_level(0.1, 10.0, 0.05, 0.1, 0.1, 0.1, 1.0)
    assert o["ch4stress"][0, 0] == 0.5 and o["o2stress"][0, 0] == 1.0
    assert o["ch4_oxid_depth"][0, 0] == 0.05 and o["ch4_aere_depth"][0, 0] == 0.05
    assert o["ch4_ebul_depth"][0, 0] == 0.05 and o["o2_oxid_depth"][0, 0] == 0.1
    # O2 limited: supply 0.3 of demand 1.2 (0.2 oxidation + 1.0 decomposition)
    o = single_level(1.0, 0.3, 0.0, 0.1, 0.0, 0.0, 1.0)
    assert o["o2stress"][0, 0] == 0.25 and o["ch4stress"][0, 0] == 1.0
    assert o["ch4_oxid_depth"][0, 0] == 0.025 and o["o2_decomp_depth"][0, 0] == 0.25
    # both limited, CH4 the more: O2 shared again after oxidation at the CH4 rate
    o = single_level(0.05, 0.5, 0.0, 0.1, 0.0, 0.0, 1.0)   # ch4s 0.5, o2s 0.5/1.2
    # ch4s 0.5 > o2s 0.41667: oxidation at the O2 rate, CH4 stress recomputed
    # without oxidation demand (none left: aere + ebul = 0 -> 1)
    assert o["o2stress"][0, 0] == 0.5 / 1.2
    assert o["ch4stress"][0, 0] == 1.0
    o = single_level(0.02, 0.5, 0.0, 0.1, 0.0, 0.0, 1.0)   # ch4s 0.2 <= o2s 0.41667
    # O2 stress recomputed on decomposition after oxidation at the CH4 rate:
    # (0.5 - 0.2 * 0.2) / 1.0 = 0.46
    assert abs(o["ch4stress"][0, 0] - 0.2) < 1e-15
    assert abs(o["o2stress"][0, 0] - 0.46) < 1e-15
    assert abs(o["ch4_oxid_depth"][0, 0] - 0.02) < 1e-16


def test_aerenchyma_oxidation_without_prognostic_switch():
    d = inputs.ch4_columns(nc=200, seed=138)
    a = run(d)
    b = run(d, use_aereoxid_prog=False, aereoxid=0.5)
    m = d["mask"].astype(bool)
    np.testing.assert_allclose(b["ch4_aere_depth"][:, m], 0.5 * a["ch4_aere_depth"][:, m], rtol=1e-15)
    np.testing.assert_allclose(b["ch4_oxid_depth"][:, m],
                               a["ch4_oxid_depth"][:, m] + 0.5 * a["ch4_aere_depth"][:, m], rtol=1e-14)


def test_julia_pow():
    rng = np.random.default_rng(5)
    for x in rng.uniform(0.0, 1.5, 500):
        assert oracle.ch4_jl_pow(x, 2.0) == x * x
        assert oracle.ch4_jl_pow(x, 3.0) == x * x * x
        exact = Fraction(float(x)) ** 7
        got = oracle.ch4_jl_pow(x, 7.0)
        assert abs(Fraction(got) - exact) <= Fraction(float(np.spacing(float(exact))))
        assert oracle.ch4_jl_pow(x, 10.0 / 3.0) == x ** (10.0 / 3.0)
    assert oracle.ch4_jl_pow(1.0, 0.37) == 1.0 and oracle.ch4_jl_pow(0.3, 0.0) == 1.0


def test_bad_jwt_rejected():
    d = inputs.ch4_columns(nc=50, seed=139)
    d["mask"][:] = 1
    d["jwt"][7] = d["nlevsoi"] + 1
    with pytest.raises(RuntimeError):
        run(d)


# ---- diffusivities (methane.jl:1818-1840; D_CON_G, D_CON_W from varcon.jl) ----
# The gas diffusivity at or above the water table is
#   (g0 + g1 tsc) 1e-4 [om f_a^(10/3) / ws^2 + (1 - om) eps^2 f_a^(3/b)] sfgas,
# f_a = 1 - vol/ws, eps = ws - vol, om = min(om_frac_sf cellorg / organic_max, 1);
# below it, ws^satpow (w0 + w1 tsc + w2 tsc^2) 1e-9 sfliq, times the liquid share
# (liq/1000 + sn) / (liq/1000 + ice/917 + sn) when frozen; never below sn = 1e-12.
# Each case below picks a state where that form reduces to a number fixed by
# hand, so a dropped term, a swapped constant or a wrong exponent moves it.
G_CH4, G_O2 = (0.1875, 0.0013), (0.1759, 0.00117)          # varcon.jl D_CON_G
W_CH4, W_O2 = (0.9798, 0.02986, 0.0004381), (1.172, 0.03443, 0.0005048)   # D_CON_W
TFRZ = 273.15


def _rel(a, b):
    return abs(a - b) / abs(b)


def test_gas_diffusivity_organic_term():
    # om = 1 (cellorg above organic_max): f_a = 1 gives g0 1e-4 / ws^2
    d = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.0, cellorg=200.0)
    assert _rel(d, 0.1875e-4 / 0.25) < 1e-15
    # f_a = 1/8 (vol = 7/8 ws): f_a^(10/3) = 2^-10 (3 or 3/10 as the exponent give 2^-9, 2^-0.9)
    d = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.4375, cellorg=200.0)
    assert _rel(d, 0.1875e-4 * 2.0 ** -10 / 0.25) < 1e-14
    # organic_max = 0 means all organic (om = 1) whatever cellorg
    assert oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.4375, cellorg=0.0, organic_max=0.0) == d


def test_gas_diffusivity_mineral_term():
    # om = 0 (cellorg 0): eps^2 f_a^(3/b); b = 3 gives f_a, b = 1.5 gives f_a^2
    # (b/3 as the exponent would give f_a^0.5)
    d3 = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.25, cellorg=0.0, bsw=3.0)
    assert _rel(d3, 0.1875e-4 * 0.0625 * 0.5) < 1e-15
    d15 = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.25, cellorg=0.0, bsw=1.5)
    assert _rel(d15, 0.1875e-4 * 0.0625 * 0.25) < 1e-14
    # a half-organic level is the mean of the two terms (om = 65 / 130)
    d_org = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.25, cellorg=200.0, bsw=3.0)
    d_half = oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.25, cellorg=65.0, bsw=3.0)
    assert _rel(d_half, 0.5 * (d_org + d3)) < 1e-15


def test_gas_diffusivity_temperature_species_scale():
    base = dict(cellorg=200.0)
    for s, (g0, g1) in ((0, G_CH4), (1, G_O2)):
        d0 = oracle.ch4_diffusivity(s, True, TFRZ, 0.5, 0.0, **base)
        d10 = oracle.ch4_diffusivity(s, True, TFRZ + 10.0, 0.5, 0.0, **base)
        assert _rel(d0, g0 * 1e-4 / 0.25) < 1e-15
        assert _rel(d10 / d0, (g0 + 10.0 * g1) / g0) < 1e-12
        # the scale factor multiplies (exactly, for a power of two)
        assert oracle.ch4_diffusivity(s, True, TFRZ + 10.0, 0.5, 0.0, scale_factor_gasdiff=2.0,
                                      **base) == 2.0 * d10
    # a saturated level above the water table (f_a = 0) sits at the floor
    assert oracle.ch4_diffusivity(0, True, TFRZ, 0.5, 0.5, cellorg=200.0) == 1e-12


def test_aqueous_diffusivity():
    for s, (w0, w1, w2) in ((0, W_CH4), (1, W_O2)):
        # thawed: ws^2 (w0 + w1 tsc + w2 tsc^2) 1e-9 at tsc = 10 (satpow 2 is ws*ws)
        d = oracle.ch4_diffusivity(s, False, TFRZ + 10.0, 0.5, 0.5)
        assert _rel(d, 0.25 * (w0 + 10.0 * w1 + 100.0 * w2) * 1e-9) < 1e-12
        # frozen without ice: the liquid share is exactly 1
        d = oracle.ch4_diffusivity(s, False, TFRZ - 10.0, 0.5, 0.5, liq=1.0, ice=0.0)
        assert _rel(d, 0.25 * (w0 - 10.0 * w1 + 100.0 * w2) * 1e-9) < 1e-12
        # frozen, liquid/1000 = ice/917: the share is 1/2 (to sn)
        d_half = oracle.ch4_diffusivity(s, False, TFRZ - 10.0, 0.5, 0.5, liq=1.0, ice=0.917)
        assert _rel(d_half, 0.5 * d) < 1e-8
    # all ice: at the floor
    assert oracle.ch4_diffusivity(0, False, TFRZ - 10.0, 0.5, 0.5, liq=0.0, ice=10.0) == 1e-12
    # a non-integer satpow goes through pow, a scale factor multiplies
    d = oracle.ch4_diffusivity(0, False, TFRZ + 10.0, 0.5, 0.5, satpow=1.7, scale_factor_liqdiff=2.0)
    assert _rel(d, 0.5 ** 1.7 * (0.9798 + 0.2986 + 0.04381) * 1e-9 * 2.0) < 1e-12
