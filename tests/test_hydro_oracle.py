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
