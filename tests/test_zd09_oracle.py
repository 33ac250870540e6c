"""Pins of the oracle's Zeng-Decker 2009 step (orc_soilwater_zengdecker2009,
following soil_water_movement.jl:97-643 and 1616-1764).

The reference's test suite holds no vectors for this path (its tests only check
the configuration rules, test_soil_water_movement.jl:18-258), so the oracle is
pinned by what the mathematics fixes:

* water balance — the linearised fluxes are shared by neighbouring rows, so the
  change of soil water equals infiltration minus root uptake (minus the flux into
  the aquifer layer when the water table is below the soil column). A wrong
  sign, index or derivative in one of the flux terms breaks it;
* hydrostatic equilibrium — a column whose water content is the layer average of
  the Clapp-Hornberger profile above the water table (computed here by
  Gauss-Legendre quadrature, not by the oracle's closed form) has zero flux
  everywhere, so the step leaves it unchanged and the recharge is zero. A wrong
  equilibrium integral (exponent, interface index, water-table layer) or a
  mismatched suction fails it; a profile with the wrong exponent is the
  negative control;
* closed forms of the diagnostics (Clapp-Hornberger suction, conductivity with
  ice impedance, ice fraction, deficit) evaluated in numpy from the inputs;
* the direction of the recharge: wetter than equilibrium above the water table
  drains down (qcharge > 0), drier draws water up (qcharge < 0).
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, bind, inputs

NS, N = inputs.NLEVSNO, inputs.NLEVSOI
CFG = api.init_soilwater_movement().to_abi()   # ZENGDECKER_2009, BC_AQUIFER, e_ice 6


def layer_above_water_table(d):
    """jwt of _soilwm_jwt_vwczwt_kernel!: index of the last soil interface above zwt."""
    zi_soil = d["zi"][NS + 1:NS + 1 + N]        # zi(joff_zi + j), j = 1..nlevsoi
    below = d["zwt"][None, :] <= zi_soil
    return np.where(below.any(0), below.argmax(0), N)


def step(d):
    out = clone(d)
    oracle.soilwater_zengdecker2009(out, CFG)
    return out


def test_water_balance_both_water_table_cases():
    d = inputs.zengdecker_columns(nc=3000, seed=21)
    out = step(d)
    jw = layer_above_water_table(d)
    assert (jw == N).sum() > 200 and (jw < N).sum() > 200
    dliq = (out["h2osoi_liq"][NS:NS + N] - d["h2osoi_liq"][NS:NS + N]).sum(0)
    flux = (d["qflx_infl"] - d["qflx_rootsoi"].sum(0)) * d["dtime"]
    expect = np.where(jw == N, flux - out["qcharge"] * d["dtime"], flux)
    scale = np.abs(d["h2osoi_liq"][NS:NS + N]).sum(0)
    assert np.max(np.abs(dliq - expect) / scale) < 1e-13
    # the soil actually changed, so the balance is not trivially met
    assert np.max(np.abs(dliq)) > 1.0
    # below-column water table: the recharge is the flux that leaves layer nlevsoi
    assert np.any(np.abs(out["qcharge"][jw == N]) > 1e-6)


def test_water_balance_shallow_water_table():
    """Water table above the first interface (jwt = 0) and in the top layers."""
    d = inputs.zengdecker_columns(nc=1000, seed=22, zwt_range=(0.001, 0.2))
    out = step(d)
    jw = layer_above_water_table(d)
    assert (jw == 0).sum() > 20 and (jw > 0).sum() > 20
    dliq = (out["h2osoi_liq"][NS:NS + N] - d["h2osoi_liq"][NS:NS + N]).sum(0)
    flux = (d["qflx_infl"] - d["qflx_rootsoi"].sum(0)) * d["dtime"]
    scale = np.abs(d["h2osoi_liq"][NS:NS + N]).sum(0)
    assert np.max(np.abs(dliq - flux) / scale) < 1e-13


def _gauss_layer_average(a, b, zwt, ws, suc, bsw, expo=1.0, npts=40):
    """Mean over [a, b] (mm) of ws*((suc + zwt - z)/suc)^(-expo/bsw) above zwt and
    ws below it, by Gauss-Legendre quadrature on each side of zwt."""
    x, w = np.polynomial.legendre.leggauss(npts)

    def integral(lo, hi):
        lo, hi = np.broadcast_arrays(lo, hi)
        h = np.maximum(hi - lo, 0.0)
        z = lo[..., None] + 0.5 * h[..., None] * (x + 1.0)
        f = ws[..., None] * ((suc[..., None] + zwt[..., None] - z) / suc[..., None]) ** \
            (-expo / bsw[..., None])
        return 0.5 * h * (f * w).sum(-1)

    top = integral(np.minimum(a, zwt), np.minimum(b, zwt))
    sat = ws * np.maximum(b - np.maximum(a, zwt), 0.0)
    return (top + sat) / (b - a)


def equilibrium_columns(nc, seed, expo=1.0):
    d = inputs.zengdecker_columns(nc=nc, seed=seed, frozen_fraction=0.0, zwt_range=(0.3, 6.0))
    zwt = d["zwt"] * 1.0e3
    zimm = d["zi"][NS + 1:NS + 1 + N] * 1.0e3
    zabove = np.vstack([np.zeros((1, nc)), zimm[:-1]])
    v = _gauss_layer_average(zabove, zimm, zwt[None, :], d["watsat"][:N], d["sucsat"][:N],
                             d["bsw"][:N], expo)
    dz = d["dz"][NS:NS + N]
    d["h2osoi_liq"][NS:NS + N] = v * dz * 1000.0
    d["h2osoi_ice"][:] = 0.0
    d["h2osoi_vol"][:N] = v
    d["qflx_infl"][:] = 0.0
    d["qflx_rootsoi"][:] = 0.0
    assert np.all(layer_above_water_table(d) < N)
    return d


def test_hydrostatic_equilibrium_is_steady():
    d = equilibrium_columns(800, seed=23)
    out = step(d)
    liq0 = d["h2osoi_liq"][NS:NS + N]
    dl = np.abs(out["h2osoi_liq"][NS:NS + N] - liq0) / liq0
    assert dl.max() < 1e-11
    assert np.max(np.abs(out["qcharge"])) < 1e-14
    assert np.all(out["qflx_deficit"] == 0.0)
    # negative control: the profile with a 5 % different exponent is not steady
    bad = equilibrium_columns(800, seed=23, expo=1.05)
    outb = step(bad)
    lb = bad["h2osoi_liq"][NS:NS + N]
    assert (np.abs(outb["h2osoi_liq"][NS:NS + N] - lb) / lb).max() > 1e-4
    assert np.max(np.abs(outb["qcharge"])) > 1e-7


def test_recharge_direction():
    """Above the water table, wetter than equilibrium drains down and drier draws up."""
    d = equilibrium_columns(400, seed=24)
    jw = layer_above_water_table(d)
    jm = np.maximum(jw, 1)          # the layer whose suction drives the recharge
    cols = np.arange(d["nc"])
    for factor, sign in ((1.03, 1.0), (0.97, -1.0)):
        e = clone(d)
        e["h2osoi_liq"][NS + jm - 1, cols] *= factor
        out = step(e)
        q = out["qcharge"]
        # only columns where the perturbed layer stays unsaturated move off zero
        unsat = e["h2osoi_liq"][NS + jm - 1, cols] / (e["dz"][NS + jm - 1, cols] * 1000.0) < \
            e["watsat"][jm - 1, cols]
        assert unsat.sum() > 100
        assert np.all(np.sign(q[unsat]) == sign)


def test_diagnostics_closed_forms():
    d = inputs.zengdecker_columns(nc=1000, seed=25)
    out = step(d)
    dz = d["dz"][NS:NS + N]
    ws, b, suc, hks = d["watsat"][:N], d["bsw"][:N], d["sucsat"][:N], d["hksat"][:N]
    theta = np.maximum(d["h2osoi_liq"][NS:NS + N], 1e-6) / (dz * 1000.0)
    ice = np.minimum(1.0, np.minimum(ws, d["h2osoi_ice"][NS:NS + N] / (dz * 917.0)) / ws)
    np.testing.assert_allclose(out["icefrac"][:N], ice, rtol=1e-15, atol=0)
    # Clapp-Hornberger suction at the start-of-step water content
    smp = np.maximum(d["smpmin"], -suc * np.exp(-b * np.log(np.clip(theta / ws, 0.01, 1.0))))
    np.testing.assert_allclose(out["smp_l"][:N], smp, rtol=1e-13)
    # interface conductivity with ice impedance 10^(-e_ice * icefrac)
    up = np.minimum(np.arange(1, N + 1), N - 1)
    s1 = np.minimum(1.0, (theta + theta[up]) / (ws + ws[up]))
    imp = 10.0 ** (-6.0 * 0.5 * (ice + ice[up]))
    hk = imp * hks * s1 ** (2.0 * b + 3.0)
    np.testing.assert_allclose(out["hk_l"][:N], hk, rtol=1e-12)
    # layers below nlevsoi are not written
    assert np.all(out["smp_l"][N:] == 0.0) and np.all(out["hk_l"][N:] == 0.0)
    # deficit: the negative liquid water left after the update
    neg = np.maximum(-out["h2osoi_liq"][NS:NS + N], 0.0).sum(0)
    np.testing.assert_allclose(out["qflx_deficit"], neg, rtol=1e-14, atol=0)


def test_deficit_from_a_drying_step():
    """Root uptake larger than the water held drives layers negative; the deficit
    is the water that is missing."""
    d = inputs.zengdecker_columns(nc=300, seed=26, frozen_fraction=0.0)
    d["h2osoi_liq"][NS:NS + 3] *= 0.02
    d["qflx_rootsoi"][:3] = 0.05
    d["qflx_infl"][:] = 0.0
    out = step(d)
    neg = np.maximum(-out["h2osoi_liq"][NS:NS + N], 0.0).sum(0)
    assert np.all(neg > 0.0)
    assert np.array_equal(out["qflx_deficit"], out["qflx_deficit"])
    np.testing.assert_allclose(out["qflx_deficit"], neg, rtol=1e-14)


def test_mask_leaves_columns_untouched():
    d = inputs.zengdecker_columns(nc=200, seed=27)
    d["mask"] = (np.arange(200) % 3 != 0).astype(np.uint8)
    out = step(d)
    off = d["mask"] == 0
    for k in ("h2osoi_liq", "icefrac", "smp_l", "hk_l"):
        assert np.array_equal(out[k][:, off], d[k][:, off])
    assert np.all(out["qcharge"][off] == 0.0)
    assert np.any(out["h2osoi_liq"][:, ~off] != d["h2osoi_liq"][:, ~off])


def test_host_checks():
    d = inputs.zengdecker_columns(nc=4, seed=28)
    mf = bind.swm_config(method=api.MOISTURE_FORM, lower_bc=api.BC_ZERO_FLUX)
    with pytest.raises(RuntimeError):
        oracle.soilwater_zengdecker2009(clone(d), mf)
    e = clone(d)
    e["nlevgrnd"] = N   # the aquifer row needs a ground layer below the soil column
    with pytest.raises(RuntimeError):
        oracle.soilwater_zengdecker2009(e, CFG)


def test_pow_jitter_hook():
    """The hook the GPU parity bar uses: off, it changes nothing; on, it moves
    results by pow()'s last bit only, and the water balance still holds."""
    d = inputs.zengdecker_columns(nc=500, seed=29)
    base = step(d)
    oracle.set_zd_pow_jitter(0)
    assert all(np.array_equal(step(d)[k], base[k]) for k in ("h2osoi_liq", "qcharge", "hk_l"))
    try:
        oracle.set_zd_pow_jitter(7)
        j = step(d)
    finally:
        oracle.set_zd_pow_jitter(0)
    liq0, liq1 = base["h2osoi_liq"][NS:NS + N], j["h2osoi_liq"][NS:NS + N]
    assert np.any(liq1 != liq0)
    assert np.max(np.abs(liq1 - liq0) / np.abs(liq0)) < 1e-8
    jw = layer_above_water_table(d)
    dliq = (liq1 - d["h2osoi_liq"][NS:NS + N]).sum(0)
    flux = (d["qflx_infl"] - d["qflx_rootsoi"].sum(0)) * d["dtime"]
    expect = np.where(jw == N, flux - j["qcharge"] * d["dtime"], flux)
    assert np.max(np.abs(dliq - expect) / np.abs(d["h2osoi_liq"][NS:NS + N]).sum(0)) < 1e-13
