"""Pins of the heat-step oracle (soil_temperature.jl restated in
oracle/clmsoil_oracle.c) by the reference's own tests and by closed forms.

Reference known answers (test/test_soil_temperature.jl), on the column its
setup_soil_column! builds (nlevsno = 5, nlevgrnd = 10, dz = 0.1 j, 280 K, 5 kg/m2
of liquid per layer, watsat 0.45, tkmg 2.0, tkdry 0.25, csol 2e6):
  * :228-259  soil_therm_prop!: interface conductivity > 0 above the bottom,
    tk[nlevgrnd] == 0, heat capacity > 0, tk_h2osfc finite and > 0;
  * :436-476  set_matrix!: every soil row's diagonal > 1;
  * :478-496  phase_change_h2osfc! without surface water: no flux;
  * :498-531  phase_change_h2osfc! snow depth from the effective snow cover;
  * :533-564  phase_change_beta! above freezing: no phase change;
  * :566-598  phase_change_beta! melting soil ice: imelt = 1, T toward TFRZ,
    ice < 5.
Closed forms (the published parameterisations, evaluated in numpy), each with
a negative control that a plausible mistake would fail:
  * Johansen/Kersten soil conductivity, with its saturated, dry and bedrock
    limits; Jordan (1991) and Sturm (1997) snow conductivity;
  * heat capacity of soil, bedrock, snow and unresolved snow, through
    fact = dtime / cv;
  * harmonic interface conductivity of equal layer conductivities;
  * the 5-band rows: each row of the implicit diffusion sums to 1 (a uniform
    temperature stays uniform), minus the surface flux derivative on the top
    rows.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import bind, inputs

TFRZ, HFUS, CPICE, CPLIQ = 273.15, 3.337e5, 2117.27, 4188.0
TKAIR, TKICE, TKWAT, DENICE, DENH2O = 0.023, 2.290, 0.57, 917.0, 1000.0


def reference_column(nlevsno=5, nlevsoi=8, nlevgrnd=10):
    """setup_soil_column! (test/test_soil_temperature.jl:90-226) for one column."""
    nlev = nlevsno + nlevgrnd
    z = np.full((nlev, 1), np.nan)
    dz = np.full((nlev, 1), np.nan)
    zi = np.full((nlev + 1, 1), np.nan)
    zi[nlevsno, 0] = 0.0
    for j in range(1, nlevgrnd + 1):
        r = j + nlevsno - 1
        dz[r, 0] = 0.1 * j
        z[r, 0] = sum(0.1 * k for k in range(1, j + 1)) - 0.5 * 0.1 * j
        zi[r + 1, 0] = zi[r, 0] + dz[r, 0]
    t = np.full((nlev, 1), np.nan)
    liq = np.full((nlev, 1), np.nan)
    ice = np.full((nlev, 1), np.nan)
    t[nlevsno:] = 280.0
    liq[nlevsno:] = 5.0
    ice[nlevsno:] = 0.0
    g = lambda v: np.full((nlevgrnd, 1), v)
    d = dict(nc=1, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dtime=1800.0,
             snl=np.zeros(1, np.int64), nbedrock=np.full(1, nlevgrnd, np.int64), z=z, dz=dz, zi=zi,
             t_soisno=t, h2osoi_liq=liq, h2osoi_ice=ice, excess_ice=g(0.0), watsat=g(0.45),
             tkmg=g(2.0), tkdry=g(0.25), csol=g(2.0e6), bsw=g(5.0), sucsat=g(100.0),
             frac_sno_eff=np.zeros(1), frac_h2osfc=np.zeros(1), hs_top_snow=np.zeros(1),
             hs_soil=np.zeros(1), hs_top=np.zeros(1), hs_h2osfc=np.zeros(1),
             dhsdT=np.full(1, -5.0), sabg_lyr_col=np.zeros((nlevsno + 1, 1)), eflx_bot=np.zeros(1),
             t_h2osfc=np.full(1, 280.0), h2osfc=np.zeros(1), h2osno_no_layers=np.zeros(1),
             snow_depth=np.zeros(1), int_snow=np.zeros(1))
    d.update(inputs.heat_outputs(d))
    return d


def test_kat_soil_therm_prop():
    """test_soil_temperature.jl:228-259."""
    d = reference_column()
    ns, ng = d["nlevsno"], d["nlevgrnd"]
    p = oracle.heat_assemble_probe(d, bind.heat_config(), 0)
    tk, cv = p["tk"], p["cv"]
    assert np.all(tk[ns:ns + ng - 1] > 0.0)
    assert tk[ns + ng - 1] == 0.0
    assert np.all(cv[ns:ns + ng] > 0.0)
    assert np.isfinite(p["tk_h2osfc"]) and p["tk_h2osfc"] > 0.0
    # closed forms on this column: layers 2.. are drier than satw = 0.1 (dke = 0,
    # thk = tkdry), so their interfaces carry tkdry; cv = csol(1-ws)dz + CPLIQ liq
    np.testing.assert_allclose(tk[ns + 1:ns + ng - 1], 0.25, rtol=1e-15)
    dzs = 0.1 * np.arange(1, ng + 1)
    np.testing.assert_allclose(cv[ns:ns + ng], 2.0e6 * 0.55 * dzs + 5.0 * CPLIQ, rtol=1e-15)


def test_kat_set_matrix_diagonal():
    """test_soil_temperature.jl:436-476: diagonal > 1 on every soil row."""
    d = reference_column()
    ns, ng = d["nlevsno"], d["nlevgrnd"]
    bm = oracle.heat_assemble_probe(d, bind.heat_config(), 0)["bmatrix"]
    for j in range(1, ng + 1):
        assert bm[j + ns, 2] > 1.0


def test_kat_phase_change_h2osfc_without_water():
    """test_soil_temperature.jl:478-496."""
    d = reference_column()
    oracle.phase_change_h2osfc(d, 0, 1.0e-6, 1.0)
    assert d["xmf_h2osfc"][0] == 0.0 and d["qflx_h2osfc_to_ice"][0] == 0.0


def test_kat_phase_change_h2osfc_snow_depth():
    """test_soil_temperature.jl:498-531: the new snow depth uses the effective
    snow cover, (h2osno - xm) / (rho_avg * frac_sno_eff)."""
    d = reference_column()
    joff = d["nlevsno"] - 1        # row of snow level 0
    d["snl"][0] = -1
    d["h2osoi_ice"][joff, 0] = 40.0
    d["h2osoi_liq"][joff, 0] = 0.0
    d["t_soisno"][joff, 0] = TFRZ
    d["dz"][joff, 0] = 0.1
    d["snow_depth"][0] = 0.10
    d["frac_sno_eff"][0] = 1.0
    d["frac_h2osfc"][0] = 0.5
    d["h2osfc"][0] = 20.0
    d["t_h2osfc"][0] = TFRZ - 1.0
    d["dhsdT"][0] = -1000.0
    dtime = d["dtime"]
    xm = 0.5 * (-1000.0 * 1.0 - 1.0 * 0.0 / dtime) * dtime / HFUS
    rho_avg = min(800.0, 40.0 / (1.0 * 0.10))
    expected = (40.0 - xm) / (rho_avg * 1.0)
    oracle.phase_change_h2osfc(d, 0, 0.0, 1.0)
    np.testing.assert_allclose(d["snow_depth"][0], expected, rtol=1e-14)
    # the freezing moves xm from the pond to the top snow layer
    np.testing.assert_allclose(d["h2osfc"][0], 20.0 + xm, rtol=1e-14)
    np.testing.assert_allclose(d["h2osoi_ice"][joff, 0], 40.0 - xm, rtol=1e-14)
    np.testing.assert_allclose(d["qflx_h2osfc_to_ice"][0], -xm / dtime, rtol=1e-14)
    assert d["t_h2osfc"][0] == TFRZ


def soil_fact(d):
    ns = d["nlevsno"]
    f = np.ones(ns + d["nlevgrnd"])
    return f, ns


def test_kat_phase_change_beta_above_freezing():
    """test_soil_temperature.jl:533-564."""
    d = reference_column()
    f, _ = soil_fact(d)
    oracle.phase_change_beta(d, 0, f)
    assert d["xmf"][0] == 0.0 and d["qflx_snomelt"][0] == 0.0 and d["qflx_snofrz"][0] == 0.0
    ns = d["nlevsno"]
    assert np.all(d["imelt"][ns:, 0] == 0)


def test_kat_phase_change_beta_melting_soil_ice():
    """test_soil_temperature.jl:566-598, and the melt it implies: the energy
    -tinc/fact (2 K above freezing) melts hm*dtime/hfus of the 5 kg of ice."""
    d = reference_column()
    ns, dtime = d["nlevsno"], d["dtime"]
    r = ns + 2 - 1
    d["h2osoi_ice"][r, 0] = 5.0
    d["h2osoi_liq"][r, 0] = 2.0
    d["t_soisno"][r, 0] = TFRZ + 2.0
    fact = np.ones(ns + d["nlevgrnd"])
    fact[ns:] = dtime / (2.0e6 * d["dz"][ns:, 0])
    oracle.phase_change_beta(d, 0, fact)
    assert d["imelt"][r, 0] == 1
    assert abs(d["t_soisno"][r, 0] - TFRZ) <= 2.0
    assert d["h2osoi_ice"][r, 0] < 5.0
    hm = 2.0 / fact[r]
    xm = hm * dtime / HFUS
    np.testing.assert_allclose(d["h2osoi_ice"][r, 0], 5.0 - xm, rtol=1e-13)
    np.testing.assert_allclose(d["h2osoi_liq"][r, 0], 2.0 + xm, rtol=1e-13)
    np.testing.assert_allclose(d["t_soisno"][r, 0], TFRZ, rtol=0, atol=1e-9)
    np.testing.assert_allclose(d["xmf"][0], HFUS * xm / dtime, rtol=1e-12)


# ---------------------------------------------------------------- closed forms

def random_columns(seed, nc=3000, snow=True, sturm=False):
    h = inputs.heat_config1(nc=nc, seed=seed, h2osfc_variant=True, nbedrock="random")
    rng = np.random.default_rng(seed)
    ns = h["nlevsno"]
    # excess ice in a third of the columns, unfrozen and frozen layers mixed
    h["excess_ice"] = np.where(rng.uniform(size=h["excess_ice"].shape) < 0.3,
                               rng.uniform(0, 20, h["excess_ice"].shape), 0.0)
    if not snow:
        h["snl"][:] = 0
        h["t_soisno"][:ns] = np.nan
        h["h2osoi_liq"][:ns] = np.nan
        h["h2osoi_ice"][:ns] = np.nan
        h["frac_sno_eff"][:] = 0.0
    h.update(inputs.heat_outputs(h))
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config(2 if sturm else 1))
    return h, ref


def johansen(h, tkwat=TKWAT, tkice=TKICE, frozen_kersten=False):
    """Johansen (1975) with Kersten numbers as CLM5 parameterises it
    (soil_temperature.jl:662-690): thk = Ke dksat + (1 - Ke) tkdry."""
    ns = h["nlevsno"]
    liq, ice = h["h2osoi_liq"][ns:], h["h2osoi_ice"][ns:]
    dz, ws, exc = h["dz"][ns:], h["watsat"], h["excess_ice"]
    t = h["t_soisno"][ns:]
    satw = np.minimum(1.0, (liq / DENH2O + ice / DENICE + exc / DENICE) / (dz * ws))
    with np.errstate(divide="ignore", invalid="ignore"):
        unfrozen = np.maximum(0.0, np.log10(satw) + 1.0)
        dke = np.where((t >= TFRZ) | frozen_kersten, unfrozen, satw)
        vl, vi = liq / (DENH2O * dz), (ice + exc) / (DENICE * dz)
        fl = vl / (vl + vi)
        dksat = h["tkmg"] * tkwat ** (fl * ws) * tkice ** ((1.0 - fl) * ws)
    thk = np.where(satw > 1.0e-7, dke * dksat + (1.0 - dke) * h["tkdry"], h["tkdry"])
    j = np.arange(1, h["nlevgrnd"] + 1)[:, None]
    return np.where(j > h["nbedrock"][None, :], 3.0, thk)


def test_johansen_soil_conductivity():
    h, ref = random_columns(71)
    ns = h["nlevsno"]
    got = ref["thk"][ns:]
    want = johansen(h)
    np.testing.assert_allclose(got, want, rtol=1e-13)
    # negative controls: water and ice conductivities swapped, or the unfrozen
    # Kersten number used below freezing, move most layers far beyond the bar
    for bad in (johansen(h, tkwat=TKICE, tkice=TKWAT), johansen(h, frozen_kersten=True)):
        moved = np.abs(bad - got) > 1e-6 * np.abs(got)
        assert moved.mean() > 0.2


def test_johansen_limits():
    """Saturated unfrozen soil: tkmg 0.57^watsat; saturated frozen (all ice):
    tkmg 2.29^watsat; dry soil: tkdry; below bedrock: 3 W/m/K."""
    d = reference_column()
    ns, ng = d["nlevsno"], d["nlevgrnd"]
    dzs = d["dz"][ns:, 0]
    d["nbedrock"][0] = 8
    d["h2osoi_liq"][ns:ns + 3, 0] = 0.45 * dzs[:3] * DENH2O        # saturated, unfrozen
    d["h2osoi_liq"][ns + 3:ns + 5, 0] = 0.0                         # saturated, frozen
    d["h2osoi_ice"][ns + 3:ns + 5, 0] = 0.45 * dzs[3:5] * DENICE
    d["t_soisno"][ns + 3:ns + 5, 0] = 263.0
    d["h2osoi_liq"][ns + 5:, 0] = 0.0                               # dry
    ref = clone(d)
    oracle.soil_temperature(ref, bind.heat_config())
    thk = ref["thk"][ns:, 0]
    np.testing.assert_allclose(thk[:3], 2.0 * 0.57 ** 0.45, rtol=2e-16)
    np.testing.assert_allclose(thk[3:5], 2.0 * 2.29 ** 0.45, rtol=2e-16)
    assert np.all(thk[5:8] == 0.25)
    assert np.all(thk[8:] == 3.0)


def test_snow_conductivity_jordan_and_sturm():
    for sturm in (False, True):
        h, ref = random_columns(72, sturm=sturm)
        ns = h["nlevsno"]
        act = np.isfinite(h["h2osoi_ice"][:ns])
        bw = (h["h2osoi_ice"][:ns] + h["h2osoi_liq"][:ns]) / (
            np.maximum(h["frac_sno_eff"], 1e-6) * np.maximum(h["dz"][:ns], 1e-6))
        jordan = TKAIR + (7.75e-5 * bw + 1.105e-6 * (bw * bw)) * (TKICE - TKAIR)
        b = bw / 1000.0
        sturm_v = np.where(bw <= 156.0, 0.023 + 0.234 * b, 0.138 - 1.01 * b + 3.233 * (b * b))
        want, other = (sturm_v, jordan) if sturm else (jordan, sturm_v)
        got = ref["thk"][:ns]
        assert act.sum() > 1000
        np.testing.assert_allclose(got[act], want[act], rtol=1e-14)
        np.testing.assert_allclose(ref["bw"][:ns][act], bw[act], rtol=1e-15)
        assert (np.abs(other[act] - got[act]) > 1e-6).mean() > 0.9     # negative control


def heat_capacity(h, cpice=CPICE, cpliq=CPLIQ):
    """cv (soil_temperature.jl:781-872): soil csol(1-ws)dz + ice cpice + liq
    cpliq + excess ice cpice; bedrock 2e6 dz; snow max(1e-6, (cpliq liq + cpice
    ice)/frac_sno); unresolved snow adds cpice h2osno_no_layers to layer 1."""
    ns, ng = h["nlevsno"], h["nlevgrnd"]
    dz = h["dz"]
    cv = np.full(dz.shape, np.nan)
    soil = h["csol"] * (1.0 - h["watsat"]) * dz[ns:] + (h["h2osoi_ice"][ns:] * cpice +
                                                       h["h2osoi_liq"][ns:] * cpliq) + h["excess_ice"] * cpice
    j = np.arange(1, ng + 1)[:, None]
    cv[ns:] = np.where(j > h["nbedrock"][None, :], 2.0e6 * dz[ns:], soil)
    cv[ns] = np.where(h["h2osno_no_layers"] > 0.0, cv[ns] + cpice * h["h2osno_no_layers"], cv[ns])
    with np.errstate(divide="ignore", invalid="ignore"):
        snow = np.maximum(1e-6, (cpliq * h["h2osoi_liq"][:ns] + cpice * h["h2osoi_ice"][:ns]) /
                          h["frac_sno_eff"])
    cv[:ns] = np.where(h["frac_sno_eff"] > 0, snow, 1e-6)
    return cv


def test_heat_capacity_through_fact():
    """fact = dtime / cv below the top layer (the top layer's fact carries the
    CAPR half-layer factor, checked separately)."""
    h, ref = random_columns(73)
    ns, ng = h["nlevsno"], h["nlevgrnd"]
    cv = heat_capacity(h)
    rows = np.arange(ns + ng)[:, None]
    top = (h["snl"] + ns)[None, :]          # row of the top active layer
    below = rows > top
    got = ref["fact"][below]
    np.testing.assert_allclose(got, h["dtime"] / cv[below], rtol=1e-15)
    bad = h["dtime"] / heat_capacity(h, cpice=CPLIQ, cpliq=CPICE)[below]
    assert (np.abs(bad - got) > 1e-6 * got).mean() > 0.5              # negative control
    # top layer: dtime/cv * dz / (0.5 (z - zi + CAPR (z[+1] - zi)))
    c = np.arange(h["nc"])
    tr = top[0]
    z, zi, dz = h["z"], h["zi"], h["dz"]
    want = h["dtime"] / cv[tr, c] * dz[tr, c] / (
        0.5 * (z[tr, c] - zi[tr, c] + 0.34 * (z[tr + 1, c] - zi[tr, c])))
    np.testing.assert_allclose(ref["fact"][tr, c], want, rtol=1e-15)


def test_band_rows_sum_to_one():
    """Implicit diffusion preserves a uniform temperature: every row of the heat
    system (set_matrix!, soil_temperature.jl:1549-1748) sums to 1, except for
    the surface flux derivative: -fact dhsdT on the top layer's row (weighted by
    the bare fraction 1 - frac_sno on soil layer 1 under snow, with
    frac_h2osfc fact dhsdT moved to the pond), -dtime/c_h2osfc dhsdT on the
    surface water row. A wrong band, sign or missing conductance breaks it."""
    h, ref = random_columns(74, nc=600)
    hc = bind.heat_config()
    ns, ng = h["nlevsno"], h["nlevgrnd"]
    checked = 0
    for c in range(0, 600, 5):
        p = oracle.heat_assemble_probe(clone(h), hc, c)
        bm = p["bmatrix"]
        snl = int(h["snl"][c])
        fact = ref["fact"][:, c]
        dhsdT = h["dhsdT"][c]
        fs, fh = h["frac_sno_eff"][c], h["frac_h2osfc"][c]
        rows = {ns: 1.0 - (h["dtime"] / ref["c_h2osfc"][c]) * dhsdT}      # surface water row
        for j in range(snl + 1, ng + 1):
            lv = j + ns - 1                                                # level row (0-based)
            row = lv + 1 if j >= 1 else lv                                 # system row
            if j == snl + 1:
                want = 1.0 - fact[lv] * dhsdT
            elif j == 1:
                want = 1.0 - (1.0 - fs) * fact[lv] * dhsdT
            else:
                want = 1.0
            if j == 1 and fh != 0.0:
                want = want + fh * fact[lv] * dhsdT
            rows[row] = want
        for row, want in rows.items():
            scale = 1.0 + np.abs(bm[row]).sum()
            assert abs(bm[row].sum() - want) <= 4e-16 * scale * 8, (c, row, bm[row], want)
            checked += 1
    assert checked > 3000


def test_ground_temperature_blend():
    """_tgrnd_kernel! (soil_temperature.jl:224-247): t_grnd is the area-weighted
    temperature of the top snow layer, the top soil layer and the surface water."""
    h, _ = random_columns(75)
    h["frac_h2osfc"][::3] = 0.0                # no surface water on a third of the columns
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    ns = h["nlevsno"]
    c = np.arange(h["nc"])
    t = ref["t_soisno"]
    fs, fh = h["frac_sno_eff"], h["frac_h2osfc"]
    top = (h["snl"] + ns).astype(int)
    snow = h["snl"] < 0
    want = np.where(snow,
                    np.where(fh != 0.0, fs * t[top, c] + (1.0 - fs - fh) * t[ns, c] + fh * ref["t_h2osfc"],
                             fs * t[top, c] + (1.0 - fs) * t[ns, c]),
                    np.where(fh != 0.0, (1.0 - fh) * t[ns, c] + fh * ref["t_h2osfc"], t[ns, c]))
    np.testing.assert_allclose(ref["t_grnd"], want, rtol=1e-15)
    bare = ~snow & (fh == 0.0)
    assert bare.any() and np.array_equal(ref["t_grnd"][bare], t[ns, bare])


def test_soil_heat_flux_closes_the_subcolumn_energy():
    """eflx_fgr[j] (soil_temperature.jl:250-269) is the Crank-Nicolson heat flux
    across the interface below soil layer j. Below a layer without phase change
    the layers' heat gain, sum cv dT / dt with cv from its closed form, equals
    that flux (zero flux at the bottom). A sign or layer slip in the flux, or a
    wrong heat capacity, breaks it."""
    h, ref = random_columns(76, snow=False)
    ns, ng, dt = h["nlevsno"], h["nlevgrnd"], h["dtime"]
    # keep every ground layer above freezing and ice-free: no phase change
    h["h2osoi_liq"][ns:] += h["h2osoi_ice"][ns:]
    h["h2osoi_ice"][ns:] = 0.0
    h["excess_ice"][:] = 0.0
    h["t_soisno"][ns:] = np.random.default_rng(3).uniform(275.0, 300.0, (ng, h["nc"]))
    h["h2osno_no_layers"][:] = 0.0
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    assert np.all(ref["imelt"][ns + 1:] == 0)
    cv = heat_capacity(h)[ns:]
    gain = cv * (ref["t_soisno"][ns:] - h["t_soisno"][ns:]) / dt     # W/m2 per layer
    below = np.cumsum(gain[::-1], axis=0)[::-1]                       # sum over layers >= j
    # the temperature changes are differences of ~290 K values: the sums carry
    # rounding of cv * 290 K * 2^-52 / dt per layer
    rnd = np.cumsum((cv * 300.0 / dt)[::-1], axis=0)[::-1] * 2.0 ** -52
    for j in range(1, ng):                                            # interface below layer j
        want = below[j]
        got = ref["eflx_fgr"][j - 1]
        scale = np.abs(gain[j:]).sum(0) + 1e-3
        assert np.all(np.abs(got - want) <= 1e-10 * scale + 8.0 * rnd[j]), j
    np.testing.assert_array_equal(ref["eflx_fgr12"], ref["eflx_fgr"][0])
    assert np.all(ref["eflx_fgr"][ng - 1] == 0.0)
    # negative control: the flux one interface off does not close
    wrong = np.abs(ref["eflx_fgr"][2] - below[1]) / (np.abs(gain[1:]).sum(0) + 1e-3)
    assert np.median(wrong) > 1e-3
