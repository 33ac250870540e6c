"""Pins of the oracle's soil water fractions and water table
(orc_set_soil_water_fractions / orc_water_table, following
soil_hydrology.jl:110-136 and 676-893).

* the reference's known-answer test for the fractions
  (test_soil_hydrology_mod.jl:67-112) and its bounds;
* the aquifer specific yield is what the mathematics fixes it to be: the
  storage of a hydrostatic Clapp-Hornberger column drops by s_y per unit fall
  of the water table, d(storage)/d(zwt) = -s_y. The storage is computed here by
  quadrature of the equilibrium profile and differentiated numerically; the
  oracle's move of a water table below the soil column, -qcharge*dt/1000/s_y,
  must imply the same s_y (a wrong exponent, sign or unit fails);
* the aquifer store gains exactly qcharge*dt when the table is below the column;
* moves inside the soil column in closed form (one layer; a deepening that
  crosses one interface), the frost table definition, and the perched water
  table at the 90 % saturation level of the linear profile between two layers.
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs

NS, N = inputs.NLEVSNO, inputs.NLEVSOI
DT = 1800.0


def test_fractions_reference_kat():
    nc, nlevsno, nlevsoi, nlevgrnd = 3, 12, 10, 15
    d = dict(nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd,
             watsat=np.full((nlevgrnd, nc), 0.4), dz=np.full((nlevsno + nlevgrnd, nc), 0.1),
             h2osoi_ice=np.full((nlevsno + nlevgrnd, nc), 2.0), excess_ice=np.zeros((nlevgrnd, nc)),
             eff_porosity=np.full((nlevgrnd, nc), 0.35), icefrac=np.zeros((nlevgrnd, nc)))
    oracle.set_soil_water_fractions(d)
    assert abs(d["eff_porosity"][0, 0] - max(0.01, 0.4 - 2.0 / (0.1 * 917.0))) < 1e-10
    assert abs(d["icefrac"][0, 0] - min(1.0, (2.0 / (0.1 * 917.0)) / 0.4)) < 1e-10
    assert np.all((d["eff_porosity"][:nlevsoi] >= 0.01) & (d["eff_porosity"][:nlevsoi] <= 0.4))
    assert np.all((d["icefrac"][:nlevsoi] >= 0.0) & (d["icefrac"][:nlevsoi] <= 1.0))
    assert np.all(d["eff_porosity"][nlevsoi:] == 0.35)     # below nlevsoi untouched


def test_fractions_with_excess_ice():
    d = inputs.water_table_columns(nc=2000, seed=31)
    out = clone(d)
    oracle.set_soil_water_fractions(out)
    ws, dz = d["watsat"][:N], d["dz"][NS:NS + N]
    ice, exc = d["h2osoi_ice"][NS:NS + N], d["excess_ice"][:N]
    vol = np.minimum(ws, (ice + exc) / ((dz + exc / 917.0) * 917.0))
    np.testing.assert_allclose(out["eff_porosity"][:N], np.maximum(0.01, ws - vol), rtol=1e-15)
    np.testing.assert_allclose(out["icefrac"][:N], np.minimum(1.0, vol / ws), rtol=1e-15)
    assert np.sum(exc.sum(0) > 0) > 200 and np.sum(out["icefrac"][:N] == 1.0) > 0


def _layer_above(d, zwt):
    zi = d["zi"][NS + 1:NS + 1 + N]
    below = zwt[None, :] <= zi
    return np.where(below.any(0), below.argmax(0), N)


def _storage(zwt_mm, ws, suc, b, depth_mm, npts=200):
    """Water in [0, depth] of a hydrostatic column with its table at zwt_mm."""
    x, w = np.polynomial.legendre.leggauss(npts)
    z = 0.5 * zwt_mm * (x + 1.0)
    theta = ws * ((suc + zwt_mm - z) / suc) ** (-1.0 / b)
    return 0.5 * zwt_mm * (theta * w).sum() + ws * (depth_mm - zwt_mm)


def test_specific_yield_is_the_storage_derivative():
    d = inputs.water_table_columns(nc=600, seed=32, zwt_range=(9.0, 30.0))
    d["qcharge"] = np.abs(d["qcharge"]) + 1e-5        # below the column: moves without clamps
    out = clone(d)
    oracle.water_table(out, DT)
    assert np.all(_layer_above(d, d["zwt"]) == N)
    rous = -(d["qcharge"] * DT) / 1000.0 / (out["zwt"] - d["zwt"])
    checked = 0
    for c in range(0, d["nc"], 7):
        ws, suc, b = d["watsat"][N - 1, c], d["sucsat"][N - 1, c], d["bsw"][N - 1, c]
        z0 = d["zwt"][c] * 1000.0
        h = 1e-3 * z0
        dsdz = (_storage(z0 + h, ws, suc, b, 5e4) - _storage(z0 - h, ws, suc, b, 5e4)) / (2 * h)
        if -dsdz > 0.011:                               # above the 0.01 floor
            assert abs(rous[c] - (-dsdz)) < 1e-6 * abs(dsdz), (c, rous[c], -dsdz)
            checked += 1
    assert checked > 40
    # the aquifer store takes exactly the recharge
    np.testing.assert_array_equal(out["wa"], d["wa"] + d["qcharge"] * DT)
    assert np.all(out["qflx_drain"] == 0.0) and np.all(out["qflx_drain_perched"] == 0.0)


def _sy(d, c, j, zwt):
    ws, suc, b = d["watsat"][j - 1, c], d["sucsat"][j - 1, c], d["bsw"][j - 1, c]
    return max(ws * (1.0 - (1.0 + 1.0e3 * zwt / suc) ** (-1.0 / b)), 0.01)


def test_moves_inside_the_column():
    d = inputs.water_table_columns(nc=3000, seed=33, zwt_range=(0.5, 6.0))
    out = clone(d)
    oracle.water_table(out, DT)
    jw = _layer_above(d, d["zwt"])
    zi = d["zi"][NS:NS + 1 + N]                         # zi[0] = surface .. zi[N]
    n_one = n_cross = 0
    for c in range(d["nc"]):
        if jw[c] >= N:
            continue
        q = d["qcharge"][c] * DT
        z0 = d["zwt"][c]
        j = jw[c] + 1                                   # layer holding the table
        sy = _sy(d, c, j, z0)
        if q > 0 and sy * (z0 - zi[j - 1, c]) * 1e3 >= q:      # rise within layer j
            assert abs(out["zwt"][c] - (z0 - q / sy / 1000.0)) <= 1e-12 * z0
            n_one += 1
        elif q < 0 and -(sy * (zi[j, c] - z0) * 1e3) < q:      # deepening within layer j
            assert abs(out["zwt"][c] - (z0 - q / sy / 1000.0)) <= 1e-12 * z0
            n_one += 1
        elif q < 0 and j < N:                           # crosses into layer j+1
            q1 = -(sy * (zi[j, c] - z0) * 1e3)
            rest = q - q1
            sy2 = _sy(d, c, j + 1, zi[j, c])
            if -(sy2 * (zi[j + 1, c] - zi[j, c]) * 1e3) < rest:
                assert abs(out["zwt"][c] - (zi[j, c] - rest / sy2 / 1000.0)) <= 1e-12 * z0
                n_cross += 1
    assert n_one > 500 and n_cross > 20


def test_frost_and_perched_tables():
    d = inputs.water_table_columns(nc=3000, seed=34)
    out = clone(d)
    oracle.water_table(out, DT)
    t = d["t_soisno"][NS:NS + N]
    z = d["z"][NS:NS + N]
    for c in range(d["nc"]):
        warm = t[:, c] > 273.15
        k = N if warm[0] else 1
        for kk in range(2, N + 1):
            if warm[kk - 2] and not warm[kk - 1]:
                k = kk
                break
        assert out["frost_table"][c] == z[k - 1, c]
    # perched: where it differs from the frost table, the linear saturation
    # profile through the two layers around it reaches 0.9 at zwt_perched
    liq, ice, dz = d["h2osoi_liq"][NS:NS + N], d["h2osoi_ice"][NS:NS + N], d["dz"][NS:NS + N]
    sat = (liq / (dz * 1000.0) + ice / (dz * 917.0)) / d["watsat"][:N]
    moved = np.flatnonzero((out["zwt_perched"] != out["frost_table"]) & (out["zwt_perched"] > 0))
    assert moved.size > 20
    for c in moved:
        zp = out["zwt_perched"][c]
        found = False
        for kp in range(N - 1):
            s1, s2 = sat[kp, c], sat[kp + 1, c]
            m = (z[kp + 1, c] - z[kp, c]) / (s2 - s1)
            if abs(m * 0.9 + (z[kp + 1, c] - m * s2) - zp) <= 1e-12 * max(1.0, zp):
                found = True
                break
        assert found, c


# ---- water tables from the moisture profile (no aquifer layer) ----
#
# theta_based_water_table!: the reference's known answers
# (test_soil_hydrology_mod.jl:175-226: saturated -> the bottom of layer 1, dry
# -> the bottom of the column), then the definition: the water table is where
# the linear saturation profile between the nodes of the deepest unsaturated
# layer above bedrock and the layer below it reaches 90 %, checked by
# evaluating that line at the result. perched_water_table!: a hand-computed
# case, then the definitions of the frost table (top of the first frozen layer
# under an unfrozen one) and of the perched table (the 90 % level above it).
# The bedrock clamp: zwt' = min(zwt, zi[nbedrock]).

def _sat(d):
    ns, n = d["nlevsno"], d["nlevsoi"]
    liq, ice, dz = d["h2osoi_liq"][ns:ns + n], d["h2osoi_ice"][ns:ns + n], d["dz"][ns:ns + n]
    vol = liq / (dz * 1000.0) + ice / (dz * 917.0)
    return vol, vol / d["watsat"][:n]


def _ref_geometry(nc, nlevsoi, nlevsno=12, nlevgrnd=15):
    """The column of test_soil_hydrology_mod.jl:175-198: dz 0.1, z at mid-layer,
    zi at j*0.1, watsat 0.4."""
    nlev = nlevsno + nlevgrnd
    dz = np.full((nlev, nc), 0.1)
    zi = np.zeros((nlev + 1, nc))
    zi[nlevsno:] = (np.arange(nlevgrnd + 1) * 0.1)[:, None]
    z = np.full((nlev, nc), np.nan)
    z[nlevsno:] = (0.05 + np.arange(nlevgrnd) * 0.1)[:, None]
    return dict(nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dz=dz, z=z, zi=zi,
                watsat=np.full((nlevgrnd, nc), 0.4), t_soisno=np.full((nlev, nc), 280.0),
                h2osoi_ice=np.zeros((nlev, nc)), nbedrock=np.full(nc, nlevsoi, dtype=np.int64),
                zwt=np.full(nc, 3.0), zwt_perched=np.zeros(nc), frost_table=np.zeros(nc),
                h2osoi_vol=np.zeros((nlevgrnd, nc)))


def test_theta_based_reference_kat():
    nl = 10
    d = _ref_geometry(2, nl)
    d["h2osoi_liq"] = np.full_like(d["dz"], 0.4 * 0.1 * 1000.0)        # saturated
    oracle.theta_based_water_table(d)
    np.testing.assert_allclose(d["zwt"], 0.1, rtol=1e-15)               # bottom of layer 1
    d["h2osoi_liq"] = np.full_like(d["dz"], 0.01)                       # dry
    oracle.theta_based_water_table(d)
    np.testing.assert_allclose(d["zwt"], nl * 0.1, rtol=1e-15)          # bottom of the column


def test_theta_based_is_the_90_percent_level():
    d = inputs.moisture_water_table_columns(nc=3000, seed=46)
    out = clone(d)
    oracle.theta_based_water_table(out)
    ns = d["nlevsno"]
    vol, sat = _sat(d)
    z, zi = d["z"][ns:], d["zi"][ns:]
    seen = {"top": 0, "bedrock": 0, "between": 0}
    for c in range(d["nc"]):
        nb = d["nbedrock"][c]
        dry = [k for k in range(nb, 0, -1) if sat[k - 1, c] <= 0.9]
        k = dry[0] if dry else 1
        assert np.array_equal(out["h2osoi_vol"][k - 1:nb, c], vol[k - 1:nb, c])
        assert np.all(out["h2osoi_vol"][:k - 1, c] == -1.0) and np.all(out["h2osoi_vol"][nb:, c] == -1.0)
        if k == 1:
            assert out["zwt"][c] == zi[1, c]
            seen["top"] += 1
        elif k == nb:
            assert out["zwt"][c] == zi[nb, c]
            seen["bedrock"] += 1
        else:
            w = out["zwt"][c]
            z1, z2, s1, s2 = z[k - 1, c], z[k, c], sat[k - 1, c], sat[k, c]
            assert z1 <= w <= z2
            assert abs(s1 + (s2 - s1) * (w - z1) / (z2 - z1) - 0.9) < 1e-12
            seen["between"] += 1
    assert min(seen.values()) > 100, seen


def test_perched_hand_case():
    """5 soil layers, 1-3 unfrozen, 4-5 frozen; saturation 0.5, 0.8, 0.95, 0.99,
    1.0; water table below the frost table. The frost table is the top of layer 4
    (0.3 m); the deepest layer above it at or below 90 % is layer 2, so the
    perched table is where the line from (z2 = 0.15, 0.8) to (z3 = 0.25, 0.95)
    reaches 0.9: 0.15 + 0.1 * (0.1 / 0.15)."""
    d = _ref_geometry(1, 5, nlevgrnd=8)
    ns = 12
    sat = np.array([0.5, 0.8, 0.95, 0.99, 1.0, 0.5, 0.5, 0.5])
    d["h2osoi_liq"] = np.zeros_like(d["dz"])
    d["h2osoi_liq"][ns:, 0] = sat * 0.4 * 0.1 * 1000.0
    d["t_soisno"][ns + 3:, 0] = 265.0
    d["zwt"][:] = 0.6
    oracle.perched_water_table(d)
    assert abs(d["frost_table"][0] - 0.3) < 1e-15
    assert abs(d["zwt_perched"][0] - (0.15 + 0.1 * (0.1 / 0.15))) < 1e-14
    # a water table above the frost table leaves the perched table there
    d["zwt"][:] = 0.2
    oracle.perched_water_table(d)
    assert d["zwt_perched"][0] == d["frost_table"][0]
    # a wetter layer above a drier one: the perched table is the top of the drier
    d["zwt"][:] = 0.6
    d["h2osoi_liq"][ns + 1, 0] = 0.85 * 0.4 * 0.1 * 1000.0              # layer 2: 0.85
    d["h2osoi_liq"][ns + 2, 0] = 0.84 * 0.4 * 0.1 * 1000.0              # layer 3: 0.84 <= 0.9
    oracle.perched_water_table(d)
    # layer 3 is now the deepest at or below 90 % and layer 4 is wetter: line crossing
    s1, s2 = 0.84, 0.99
    assert abs(d["zwt_perched"][0] - (0.25 + 0.1 * (0.9 - s1) / (s2 - s1))) < 1e-14


def test_perched_definitions():
    d = inputs.moisture_water_table_columns(nc=3000, seed=47)
    out = clone(d)
    oracle.perched_water_table(out)
    ns, n = d["nlevsno"], d["nlevsoi"]
    _, sat = _sat(d)
    t = d["t_soisno"][ns:ns + n]
    z, zi = d["z"][ns:], d["zi"][ns:]
    seen = {"above": 0, "cross": 0, "ft": 0}
    for c in range(d["nc"]):
        warm = t[:, c] > 273.15
        k_frz = n if warm[0] else 1
        for k in range(2, n + 1):
            if warm[k - 2] and not warm[k - 1]:
                k_frz = k
                break
        ft = zi[k_frz - 1, c]
        assert out["frost_table"][c] == ft
        zp = out["zwt_perched"][c]
        frozen = not warm[k_frz - 1]
        if (d["zwt"][c] < ft and frozen) or k_frz == 1 or not frozen:
            assert zp == ft
            seen["ft"] += 1
            continue
        dry = [k for k in range(k_frz, 0, -1) if sat[k - 1, c] <= 0.9]
        kp = dry[0] if dry else 1
        if kp == k_frz:
            assert zp == ft
        elif sat[kp - 1, c] > sat[kp, c]:
            assert zp == zi[kp - 1, c]
            seen["above"] += 1
        else:
            z1, z2, s1, s2 = z[kp - 1, c], z[kp, c], sat[kp - 1, c], sat[kp, c]
            line = s1 + (s2 - s1) * (zp - z1) / (z2 - z1)
            if zp == 0.0:            # the 90 % level lies above the surface: floored at 0
                assert kp == 1 and line >= 0.9 - 1e-12
                seen["surface"] = seen.get("surface", 0) + 1
            else:
                if dry:              # layer kp at or below 90 %: the crossing is between the nodes
                    assert z1 - 1e-12 <= zp <= z2 + 1e-12
                else:                # all wet up to layer 1: extrapolated above its node
                    assert 0.0 < zp <= z1
                assert abs(line - 0.9) < 1e-12
            seen["cross"] += 1
    assert seen["ft"] > 100 and seen["cross"] > 100 and seen["above"] > 5, seen


def test_clamp_to_bedrock_and_sequence():
    d = inputs.moisture_water_table_columns(nc=2000, seed=48)
    out = clone(d)
    oracle.clamp_zwt_to_bedrock(out)
    zb = d["zi"][d["nlevsno"] + d["nbedrock"], np.arange(d["nc"])]
    assert np.array_equal(out["zwt"], np.minimum(d["zwt"], zb))
    assert np.sum(d["zwt"] > zb) > 100
    # the sequence ends at or above bedrock; perched reads the zwt from before
    seq = clone(d)
    oracle.water_table_moisture_form(seq)
    assert np.all(seq["zwt"] <= zb)
    alone = clone(d)
    oracle.perched_water_table(alone)
    assert np.array_equal(seq["zwt_perched"], alone["zwt_perched"])


def test_bad_nbedrock_is_rejected():
    d = inputs.moisture_water_table_columns(nc=16, seed=49)
    d["nbedrock"][3] = d["nlevsoi"] + 1
    with pytest.raises(RuntimeError):
        oracle.theta_based_water_table(d)
