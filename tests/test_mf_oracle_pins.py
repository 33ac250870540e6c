"""Pins of the moisture-form Richards oracle (soil_water_movement.jl:1250-1489
restated in oracle/clmsoil_oracle.c) beyond the reference's own known answers
(test_oracle_kat.py):

  * hydrostatic equilibrium: a column whose suction rises with depth by exactly
    the depth (smp[j+1] - smp[j] = 1000 (z[j+1] - z[j]) mm) has zero interface
    fluxes (:1334-1342), so with no infiltration, no root sink and a zero-flux
    bottom the step leaves it unchanged in one accepted substep. Negative
    controls: the mirrored profile (gravity sign flipped) and a 5 % different
    Clapp-Hornberger exponent both move water by orders of magnitude more;
  * the column's smp_l / hk_l equal the SWRC helper forms (orc_soil_suction /
    orc_soil_hk with the ice impedance, pinned by the reference's SWRC KATs)
    at the node and interface saturations (:1281-1310);
  * the adaptive controller (:1460-1483), from the oracle's per-iteration trace:
    accepted substeps add up to dtime, iterations = nsubsteps, rejections =
    nrejected, halving on rejection, doubling below xTolerLower, the dtmin
    floor accepting a substep whose error exceeds xTolerUpper;
  * Julia's Float64 max/min (smooth_ad.jl:91-149 in the default mode): NaN
    propagates from either argument, max(-0.0, 0.0) = 0.0, min(0.0, -0.0) = -0.0.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle
from _util import clone
from clmsoil import api, bind, inputs


def mf_cfg(lower_bc=api.BC_ZERO_FLUX):
    return api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                       lower_boundary_condition=lower_bc).to_abi()


def hydrostatic_columns(nc=256, seed=5, sign=1.0, b_scale=1.0, icefrac=0.0):
    """Config-0 texture and grid; liquid water set so that node suctions satisfy
    smp[j] = smp[1] + sign * 1000 (z[j] - z[1]) with each layer's own SWRC."""
    d = inputs.richards_config0(nc=nc, seed=seed, lower_bc=api.BC_ZERO_FLUX)
    rng = np.random.default_rng(seed)
    ns, n = d["nlevsno"], d["nlevsoi"]
    z = d["z"][ns:ns + n]
    smp1 = -rng.uniform(15000.0, 40000.0, nc)
    smp = smp1[None, :] + sign * 1000.0 * (z - z[0:1])
    b = d["bsw"][:n] * b_scale
    s = (-smp / d["sucsat"][:n]) ** (-1.0 / b)          # inverse Clapp-Hornberger
    assert (s > 0.02).all() and (s < 0.99).all()
    d["h2osoi_liq"][ns:ns + n] = s * d["watsat"][:n] * d["dz"][ns:ns + n] * 1000.0
    d["qflx_infl"][:] = 0.0
    d["icefrac"][:n] = icefrac
    d["nrejected"] = np.full(nc, np.nan)
    return d


def test_hydrostatic_equilibrium_is_steady():
    d = hydrostatic_columns(icefrac=0.3)
    ns, n = d["nlevsno"], d["nlevsoi"]
    liq0 = d["h2osoi_liq"][ns:ns + n].copy()
    dbg = dict(qin=np.zeros((n, d["nc"])), qout=np.zeros((n, d["nc"])))
    oracle.soilwm_moisture_form_solve(d, mf_cfg(), dbg)
    assert np.all(d["nsubsteps"] == 1) and np.all(d["nrejected"] == 0)
    moved = np.abs(d["h2osoi_liq"][ns:ns + n] - liq0) / liq0
    assert moved.max() < 1e-12
    hk = d["hk_l"][:n]
    assert np.all(np.abs(dbg["qout"][:-1]) <= 1e-9 * hk[:-1])        # fluxes vanish
    assert np.all(d["qcharge"] == 0.0)


def test_hydrostatic_negative_controls():
    """A flipped gravity term or a wrong exponent is not an equilibrium: the
    same check would fail by many orders of magnitude."""
    for kw in (dict(sign=-1.0), dict(b_scale=1.05)):
        d = hydrostatic_columns(**kw)
        ns, n = d["nlevsno"], d["nlevsoi"]
        liq0 = d["h2osoi_liq"][ns:ns + n].copy()
        oracle.soilwm_moisture_form_solve(d, mf_cfg())
        moved = np.abs(d["h2osoi_liq"][ns:ns + n] - liq0) / liq0
        assert moved.max() > 1e-6, kw


def test_column_swrc_matches_helper_forms():
    """smp_l and hk_l of a column (one accepted substep, so they belong to the
    initial state) against the helper forms at the same saturations."""
    d = hydrostatic_columns(nc=64, seed=9, icefrac=0.4)
    ns, n = d["nlevsno"], d["nlevsoi"]
    d["icefrac"][:n] = np.random.default_rng(1).uniform(0.0, 0.8, (n, d["nc"]))
    liq = d["h2osoi_liq"][ns:ns + n].copy()
    oracle.soilwm_moisture_form_solve(d, mf_cfg())
    assert np.all(d["nsubsteps"] == 1)
    L = oracle.lib()
    smp, dsmp, hk, dhk = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    e_ice = 6.0
    for c in range(d["nc"]):
        vwc = np.maximum(liq[:, c], 1e-6) / (d["dz"][ns:ns + n, c] * 1000.0)
        s2 = np.maximum(0.01, np.minimum(vwc / d["watsat"][:n, c], 1.0))
        for j in range(n):
            L.orc_soil_suction(d["bsw"][j, c], d["sucsat"][j, c], s2[j], C.byref(smp), C.byref(dsmp))
            assert d["smp_l"][j, c] == smp.value
            if j < n - 1:
                s1 = 0.5 * (s2[j] + s2[j + 1])
                imp = L.orc_ice_impedance(0.5 * (d["icefrac"][j, c] + d["icefrac"][j + 1, c]), e_ice)
            else:
                s1 = s2[j]
                imp = L.orc_ice_impedance(d["icefrac"][j, c], e_ice)
            s1 = max(0.01, min(s1, 1.0))
            L.orc_soil_hk(d["hksat"][j, c], d["bsw"][j, c], s1, imp, C.byref(hk), C.byref(dhk))
            assert d["hk_l"][j, c] == hk.value


def replay_controller(rows, dtime, dtmin=60.0, xtu=0.1, xtl=0.01, verysmall=1e-8):
    """Check the trace against the substep rules of :1460-1483; returns the sum
    of accepted substeps."""
    dtsub, dtdone = dtime, 0.0
    for k, (dt, err, acc) in enumerate(rows):
        assert dt == dtsub, (k, dt, dtsub)
        reject = err > xtu and dt > dtmin
        assert bool(acc) == (not reject), k
        if reject:
            dtsub = max(dtsub / 2.0, dtmin)
            continue
        dtdone += dt
        if abs(dtime - dtdone) < verysmall:
            assert k == len(rows) - 1
            return dtdone
        if err < xtl:
            dtsub = dtsub * 2.0
        dtsub = min(dtsub, dtime - dtdone)
    raise AssertionError("the trace ended without completing the step")


def test_controller_invariants():
    d = inputs.richards_config0(nc=512, seed=41, nbedrock="random", lower_bc=api.BC_FLUX)
    d["nrejected"] = np.full(d["nc"], np.nan)
    cfg = mf_cfg(api.BC_FLUX)
    seen_reject = seen_double = 0
    for col in range(0, 512, 9):
        r = oracle.mf_trace(clone(d), cfg, col)
        ref = clone(d)
        oracle.soilwm_moisture_form_solve(ref, cfg)
        assert len(r) == ref["nsubsteps"][col]
        assert (r[:, 2] == 0).sum() == ref["nrejected"][col]
        total = replay_controller(r, d["dtime"])
        assert abs(total - d["dtime"]) < 1e-8
        seen_reject += int((r[:, 2] == 0).any())
        seen_double += int(any(r[k + 1, 0] > r[k, 0] for k in range(len(r) - 1)))
    assert seen_reject > 5 and seen_double > 5


def test_controller_dtmin_floor():
    """Wet columns under 1e-2 mm/s of infiltration with free drainage keep
    failing the error test; the substep halves down to dtmin = 60 s and is then
    accepted although its error exceeds xTolerUpper (a last substep may be
    shorter than dtmin: the remainder of the step)."""
    d = inputs.richards_config0(nc=64, seed=3, lower_bc=api.BC_FLUX)
    d["qflx_infl"][:] = 1.0e-2
    r = oracle.mf_trace(d, mf_cfg(api.BC_FLUX), 4)
    floor_accepts = (r[:, 0] == 60.0) & (r[:, 2] == 1) & (r[:, 1] > 0.1)
    assert floor_accepts.sum() >= 5
    assert not ((r[:, 0] <= 60.0) & (r[:, 2] == 0)).any()
    assert abs(replay_controller(r, d["dtime"]) - d["dtime"]) < 1e-8


def test_julia_max_min_semantics():
    nan, inf = np.nan, np.inf
    x = np.array([nan, 1.0, nan, -0.0, 0.0, -0.0, 0.0, -inf, 2.0, -3.0, inf])
    y = np.array([1.0, nan, nan, 0.0, -0.0, -0.0, 0.0, inf, 2.0, -1.0, -inf])
    mx = oracle.clm_math("jl_max", x, y)
    mn = oracle.clm_math("jl_min", x, y)
    assert np.isnan(mx[:3]).all() and np.isnan(mn[:3]).all()
    # signed zeros: max picks +0, min picks -0, whatever the order
    assert mx[3] == 0.0 and not np.signbit(mx[3]) and not np.signbit(mx[4])
    assert np.signbit(mn[3]) and np.signbit(mn[4])
    assert np.signbit(mx[5]) and np.signbit(mn[5])          # both -0.0
    assert not np.signbit(mx[6]) and not np.signbit(mn[6])  # both +0.0
    np.testing.assert_array_equal(mx[7:], [inf, 2.0, -1.0, inf])
    np.testing.assert_array_equal(mn[7:], [-inf, 2.0, -3.0, -inf])
    # C's fmax/fmin drop the NaN instead: the semantics differ where it matters
    assert not np.isnan(np.fmax(x[:2], y[:2])).any()
    assert bind is not None


def test_aquifer_pass_zero_recharge_quirk():
    """soilwater_moisture_form! (soil_water_movement.jl:1575-1598) calls
    compute_qcharge! with all-zero smp and imped rows, so its conductivity ka is
    0 and the recharge it writes is zero whenever the water table lies inside the
    soil column (jwt < nlevsoi, :1153-1217); with the table below the column the
    solve's own bottom drainage stands."""
    d = inputs.richards_config0(nc=2048, seed=44, nbedrock="random", lower_bc=api.BC_WATERTABLE,
                                zwt_range=(0.3, 14.0))
    cfg = mf_cfg(api.BC_WATERTABLE)
    solve_only = clone(d)
    oracle.soilwm_moisture_form_solve(solve_only, cfg)
    full = clone(d)
    oracle.soilwater_moisture_form(full, cfg)
    ns, n = d["nlevsno"], d["nlevsoi"]
    zi = d["zi"][ns + 1:ns + 1 + n]                       # soil interfaces 1..nlevsoi
    inside = (d["zwt"][None, :] <= zi).any(axis=0)        # jwt < nlevsoi
    assert inside.mean() > 0.3 and (~inside).mean() > 0.1
    assert np.all(full["qcharge"][inside] == 0.0)
    np.testing.assert_array_equal(full["qcharge"][~inside], solve_only["qcharge"][~inside])
    np.testing.assert_array_equal(full["h2osoi_liq"], solve_only["h2osoi_liq"])
    assert np.abs(solve_only["qcharge"][inside]).max() > 0.0   # the pass does overwrite
