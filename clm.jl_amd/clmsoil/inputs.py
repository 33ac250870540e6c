"""Seeded synthetic column inputs for the soil solves.

Arrays use CLM.jl's memory layout: a Julia `Matrix{Float64}` of size
`(nc, nlev)` is column-major, so it is stored here as a C-ordered numpy array of
shape `(nlev, nc)` — row `k` holds level `k` of every column. Snow+soil arrays
have `nlevsno + nlevgrnd` rows (soil level j at row `j + nlevsno - 1`), `zi`
one more, soil-only arrays `nlevgrnd` rows.

Grid: the CLM5 "20SL_8.5m" layer structure (src/constants/varcon.jl:241-283).
Soil properties: the CLM5 pedotransfer functions of init_soil_properties!
(src/infrastructure/cold_start.jl:276-376) with the default SoilStateParams
(src/types/soil_state.jl:19-33).
"""
from __future__ import annotations

import numpy as np

NLEVSNO, NLEVSOI, NLEVGRND = 12, 20, 25
SB = 5.67e-8
DENH2O, DENICE = 1000.0, 917.0
TFRZ = 273.15


def clm5_grid(nlevsoi: int = NLEVSOI, nlevgrnd: int = NLEVGRND):
    """dzsoi, zsoi, zisoi of the 20SL_8.5m structure (varcon.jl:254-283)."""
    dz = np.zeros(nlevgrnd)
    for j in range(1, min(4, nlevsoi) + 1):
        dz[j - 1] = j * 0.02
    for j in range(5, min(13, nlevsoi) + 1):
        dz[j - 1] = dz[3] + (j - 4) * 0.04
    for j in range(14, nlevsoi + 1):
        dz[j - 1] = dz[min(13, nlevsoi) - 1] + (j - 13) * 0.10
    for j in range(nlevsoi + 1, nlevgrnd + 1):
        dz[j - 1] = dz[nlevsoi - 1] + (((j - nlevsoi) * 25.0) ** 1.5) / 100.0
    zi = np.zeros(nlevgrnd + 1)
    for j in range(1, nlevgrnd + 1):
        zi[j] = zi[j - 1] + dz[j - 1]
    z = 0.5 * (zi[:-1] + zi[1:])
    return dz, z, zi


def pedotransfer(sand, clay, om, zsoi):
    """Per-(column, layer) hydraulic and thermal soil properties.

    cold_start.jl:276-376 with every *_sf scale factor = 1 and zsapric = 0.5.
    sand/clay/om are per column (shape (nc,)); zsoi per layer (nlevgrnd,).
    Returns dict of (nlevgrnd, nc) arrays.
    """
    sand = sand[None, :]
    clay = clay[None, :]
    om = om[None, :]
    dr = (zsoi / 0.5)[:, None]
    watsat_min = 0.489 - 0.00126 * sand
    bsw_min = 2.91 + 0.159 * clay
    sucsat_min = 10.0 * 10.0 ** (1.88 - 0.0131 * sand)
    xksat_min = 0.0070556 * 10.0 ** (-0.884 + 0.0153 * sand)
    watsat_org = np.maximum(0.93 - 0.1 * dr, 0.83)
    bsw_org = np.minimum(2.7 + 9.3 * dr, 12.0)
    sucsat_org = np.minimum(10.3 - 0.2 * dr, 10.1)
    xksat_org = np.maximum(0.28 - 0.2799 * dr, xksat_min)
    watsat = np.minimum((1.0 - om) * watsat_min + om * watsat_org, 0.93)
    bsw = (1.0 - om) * bsw_min + om * bsw_org
    sucsat = (1.0 - om) * sucsat_min + om * sucsat_org
    pcalpha, pcbeta = 0.5, 0.139
    perc_frac = np.where(om > pcalpha,
                         (1.0 - pcalpha) ** (-pcbeta) * np.maximum(om - pcalpha, 0.0) ** pcbeta, 0.0)
    uncon_frac = (1.0 - om) + (1.0 - perc_frac) * om
    uncon_hksat = uncon_frac / ((1.0 - om) / xksat_min + ((1.0 - perc_frac) * om) / xksat_org)
    hksat = uncon_frac * uncon_hksat + (perc_frac * om) * xksat_org
    pd = 2700.0
    bd = (1.0 - watsat_min) * pd
    sc = sand + clay
    tkm = (1.0 - om) * (8.8 * sand + 2.92 * clay) / sc + 0.25 * om
    tkmg = tkm ** (1.0 - watsat)
    tkdry = ((0.135 * bd + 64.7) / (pd - 0.947 * bd)) * (1.0 - om) + 0.05 * om
    csol_min = (2.128e6 * sand + 2.385e6 * clay) / sc
    csol = (1.0 - om) * csol_min + 2.5e6 * om
    nl = zsoi.shape[0]
    out = dict(watsat=watsat, bsw=bsw, sucsat=sucsat, hksat=hksat, tkmg=tkmg,
               tkdry=np.broadcast_to(tkdry, (nl, sand.shape[1])).copy(),
               csol=np.broadcast_to(csol, (nl, sand.shape[1])).copy())
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in out.items()}


def _texture(rng, nc):
    sand = rng.uniform(5.0, 90.0, nc)
    clay = rng.uniform(5.0, 60.0, nc)
    om = rng.uniform(0.0, 0.3, nc)
    return sand, clay, om


def _geometry(nc, snl, snow_dz, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
              fill=np.nan):
    """z, dz, zi in CLM layout. Soil levels from the CLM5 grid; snow layers
    stacked above the surface (zi(j-1) = zi(j) - dz(j)); inactive snow slots
    hold `fill`."""
    dzs, zs, zis = clm5_grid(nlevsoi, nlevgrnd)
    nlev = nlevsno + nlevgrnd
    dz = np.full((nlev, nc), fill)
    z = np.full((nlev, nc), fill)
    zi = np.full((nlev + 1, nc), fill)
    dz[nlevsno:, :] = dzs[:, None]
    z[nlevsno:, :] = zs[:, None]
    zi[nlevsno:, :] = zis[:, None]
    # snow: level j (−nlevsno+1..0) at row j + nlevsno − 1; interface j at row j + nlevsno
    for j in range(0, -nlevsno, -1):
        active = snl <= j - 1          # j >= snl + 1
        r = j + nlevsno - 1
        dz[r, active] = snow_dz[r, active]
        zi[r, active] = zi[r + 1, active] - dz[r, active]
        z[r, active] = 0.5 * (zi[r, active] + zi[r + 1, active])
    return z, dz, zi


def richards_config0(nc=4096, seed=0, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                     nbedrock=None, lower_bc=2, zwt_range=(0.5, 10.0), fill=np.nan):
    """Config 0: one 1800 s moisture-form step over `nc` columns x 20 soil layers.

    Draw order (numpy PCG64(seed)): sand, clay, om (per column), then the
    per-(layer, column) liquid fraction U(0.3, 0.95), then qflx_infl U(0, 2e-3),
    then (if nbedrock == "random") nbedrock U{5..20}, then zwt.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sand, clay, om = _texture(rng, nc)
    dzs, zs, zis = clm5_grid(nlevsoi, nlevgrnd)
    props = pedotransfer(sand, clay, om, zs)
    snl = np.zeros(nc, dtype=np.int64)
    z, dz, zi = _geometry(nc, snl, np.zeros((nlevsno, nc)), nlevsno, nlevsoi, nlevgrnd, fill)
    nlev = nlevsno + nlevgrnd
    frac = rng.uniform(0.3, 0.95, (nlevgrnd, nc))
    liq = np.full((nlev, nc), fill)
    liq[nlevsno:, :] = frac * props["watsat"] * dzs[:, None] * DENH2O
    qflx_infl = rng.uniform(0.0, 2.0e-3, nc)
    if nbedrock is None:
        nb = np.full(nc, nlevsoi, dtype=np.int64)
    elif isinstance(nbedrock, str) and nbedrock == "random":
        nb = rng.integers(5, nlevsoi + 1, nc).astype(np.int64)
    else:
        nb = np.full(nc, int(nbedrock), dtype=np.int64)
    zwt = rng.uniform(zwt_range[0], zwt_range[1], nc)
    return dict(
        nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dtime=1800.0,
        lower_bc=lower_bc, nbedrock=nb, h2osoi_liq=liq,
        smp_l=np.full((nlevgrnd, nc), fill), hk_l=np.full((nlevgrnd, nc), fill),
        qcharge=np.full(nc, fill), nsubsteps=np.full(nc, fill),
        z=z, zi=zi, dz=dz, qflx_rootsoi=np.zeros((nlevsoi, nc)),
        watsat=props["watsat"], hksat=props["hksat"], bsw=props["bsw"],
        sucsat=props["sucsat"], icefrac=np.zeros((nlevgrnd, nc)), zwt=zwt,
        qflx_infl=qflx_infl, smpmin=np.full(nc, -1.0e8),
        sand=sand, clay=clay, om=om)


def heat_config1(nc=65536, seed=1, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                 h2osfc_variant=False, nbedrock=None, fill=np.nan):
    """Config 1: one 1800 s snow+soil heat step.

    Per column: texture as config 0; ns ~ U{0..12} snow layers (snl = -ns) with
    dz ~ U(0.01, 0.5) m, density ~ U(80, 500) kg/m3 split into ice/liquid by an
    ice fraction ~ U(0, 1); ground water theta ~ U(0.05, 0.95) x watsat split by
    an ice fraction ~ U(0, 1); t_soisno ~ U(250, 300) K; hs_top_snow = hs_soil =
    hs_top = hs_h2osfc ~ N(50, 100) W/m2; dhsdT = -4*0.97*sigma*T_top^3;
    frac_sno_eff = 1 when snow exists. The h2osfc variant adds frac_h2osfc ~
    U(0, 0.3), h2osfc ~ U(0, 40) kg/m2, t_h2osfc ~ U(265, 285) K, and unresolved
    snow h2osno_no_layers ~ U(0, 5) kg/m2 on half of the snow-free columns.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sand, clay, om = _texture(rng, nc)
    dzs, zs, zis = clm5_grid(nlevsoi, nlevgrnd)
    props = pedotransfer(sand, clay, om, zs)
    ns = rng.integers(0, nlevsno + 1, nc)
    snl = (-ns).astype(np.int64)
    snow_dz = rng.uniform(0.01, 0.5, (nlevsno, nc))
    rho = rng.uniform(80.0, 500.0, (nlevsno, nc))
    fice_snow = rng.uniform(0.0, 1.0, (nlevsno, nc))
    z, dz, zi = _geometry(nc, snl, snow_dz, nlevsno, nlevsoi, nlevgrnd, fill)
    nlev = nlevsno + nlevgrnd
    liq = np.full((nlev, nc), fill)
    ice = np.full((nlev, nc), fill)
    t = np.full((nlev, nc), fill)
    rows = np.arange(nlevsno)[:, None]
    active_snow = (rows - nlevsno + 1) >= (snl[None, :] + 1)
    mass = rho * snow_dz
    liq[:nlevsno] = np.where(active_snow, (1.0 - fice_snow) * mass, fill)
    ice[:nlevsno] = np.where(active_snow, fice_snow * mass, fill)
    theta = rng.uniform(0.05, 0.95, (nlevgrnd, nc)) * props["watsat"]
    fice = rng.uniform(0.0, 1.0, (nlevgrnd, nc))
    liq[nlevsno:] = (1.0 - fice) * theta * dzs[:, None] * DENH2O
    ice[nlevsno:] = fice * theta * dzs[:, None] * DENICE
    tt = rng.uniform(250.0, 300.0, (nlev, nc))
    t[nlevsno:] = tt[nlevsno:]
    t[:nlevsno] = np.where(active_snow, tt[:nlevsno], fill)
    hs = rng.normal(50.0, 100.0, nc)
    top = (snl + nlevsno).astype(np.int64)          # row of the top active layer
    t_top = t[top, np.arange(nc)]
    dhsdT = -4.0 * 0.97 * SB * t_top ** 3
    frac_sno_eff = (ns > 0).astype(np.float64)
    snow_depth = np.where(active_snow, dz[:nlevsno], 0.0).sum(axis=0)
    if nbedrock is None:
        nb = np.full(nc, nlevsoi, dtype=np.int64)
    elif isinstance(nbedrock, str) and nbedrock == "random":
        nb = rng.integers(5, nlevsoi + 1, nc).astype(np.int64)
    else:
        nb = np.full(nc, int(nbedrock), dtype=np.int64)
    frac_h2osfc = np.zeros(nc)
    h2osfc = np.zeros(nc)
    t_h2osfc = t[nlevsno].copy()
    h2osno_nl = np.zeros(nc)
    if h2osfc_variant:
        frac_h2osfc = rng.uniform(0.0, 0.3, nc)
        h2osfc = rng.uniform(0.0, 40.0, nc)
        t_h2osfc = rng.uniform(265.0, 285.0, nc)
        h2osno_nl = np.where((ns == 0) & (rng.uniform(0, 1, nc) < 0.5),
                             rng.uniform(0.0, 5.0, nc), 0.0)
    int_snow = np.where(active_snow, liq[:nlevsno] + ice[:nlevsno], 0.0).sum(axis=0) + h2osno_nl
    return dict(
        nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dtime=1800.0,
        snl=snl, nbedrock=nb, z=z, dz=dz, zi=zi, t_soisno=t, h2osoi_liq=liq,
        h2osoi_ice=ice, excess_ice=np.zeros((nlevgrnd, nc)),
        watsat=props["watsat"], tkmg=props["tkmg"], tkdry=props["tkdry"],
        csol=props["csol"], bsw=props["bsw"], sucsat=props["sucsat"],
        hksat=props["hksat"],
        frac_sno_eff=frac_sno_eff, frac_h2osfc=frac_h2osfc,
        hs_top_snow=hs.copy(), hs_soil=hs.copy(), hs_top=hs.copy(), hs_h2osfc=hs.copy(),
        dhsdT=dhsdT, sabg_lyr_col=np.zeros((nlevsno + 1, nc)), eflx_bot=np.zeros(nc),
        t_h2osfc=t_h2osfc, h2osfc=h2osfc, h2osno_no_layers=h2osno_nl,
        snow_depth=snow_depth, int_snow=int_snow)


HEAT_OUTPUT_SHAPES = {
    # name: (rows or None for per-column)
    "t_grnd": None, "xmf": None, "xmf_h2osfc": None, "qflx_snomelt": None,
    "qflx_snofrz": None, "qflx_snow_drain": None, "qflx_h2osfc_to_ice": None,
    "eflx_fgr12": None, "qflx_snomelt_lyr": "nlevsno", "qflx_snofrz_lyr": "nlevsno",
}
HEAT_DIAG_SHAPES = {
    "fact": "nlev", "c_h2osfc": None, "thk": "nlev", "bw": "nlevsno",
    "eflx_fgr": "nlevgrnd", "exice_subs": "nlevgrnd", "eflx_snomelt": None,
    "eflx_snomelt_r": None, "eflx_h2osfc_to_snow": None, "snomelt_accum": None,
}


def heat_outputs(cfg: dict, diagnostics: bool = True, fill=np.nan) -> dict:
    """Allocate the heat step's output arrays (host, NaN-filled)."""
    nc, ns, ng = cfg["nc"], cfg["nlevsno"], cfg["nlevgrnd"]
    rows = {"nlevsno": ns, "nlevgrnd": ng, "nlev": ns + ng}
    out = {}
    shapes = dict(HEAT_OUTPUT_SHAPES)
    if diagnostics:
        shapes.update(HEAT_DIAG_SHAPES)
    for k, r in shapes.items():
        out[k] = np.full(nc if r is None else (rows[r], nc), fill)
    if diagnostics:
        out["snomelt_accum"][:] = 0.0
        out["imelt"] = np.full((ns + ng, nc), -1, dtype=np.int64)
    return out


def production_day(nc=1 << 20, seed=2, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                   fill=0.0):
    """Config 2: uneven production columns for a 48-step day of Richards + heat.

    Snow layers ns ~ U{0..12}, bedrock index nbedrock ~ U{5..20}, water-table
    depth zwt ~ U(0.5, 10) m (the Richards lower boundary is BC_WATERTABLE so
    zwt takes effect), otherwise drawn as the heat-step configuration. icefrac is
    the initial ice fraction of the pore space, min(1, vol_ice / watsat).
    """
    h = heat_config1(nc=nc, seed=seed, nlevsno=nlevsno, nlevsoi=nlevsoi,
                     nlevgrnd=nlevgrnd, nbedrock="random", fill=fill)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    dzs, _, _ = clm5_grid(nlevsoi, nlevgrnd)
    vol_ice = h["h2osoi_ice"][nlevsno:] / (dzs[:, None] * DENICE)
    icefrac = np.minimum(1.0, vol_ice / h["watsat"])
    h.update(
        lower_bc=3, qflx_infl=rng.uniform(0.0, 2.0e-3, nc), zwt=rng.uniform(0.5, 10.0, nc),
        icefrac=np.ascontiguousarray(icefrac), qflx_rootsoi=np.zeros((nlevsoi, nc)),
        smpmin=np.full(nc, -1.0e8), smp_l=np.full((nlevgrnd, nc), fill),
        hk_l=np.full((nlevgrnd, nc), fill), qcharge=np.full(nc, fill),
        nsubsteps=np.full(nc, fill))
    return h


def calibration_columns(nc=262144, seed=3, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                        freezing=False):
    """Config 3: snow-free columns for the differentiable calibration.

    Texture and pedotransfer as config 0, nbedrock ~ U{5..20}, liquid water
    theta ~ U(0.3, 0.95) x watsat with no ice, t_soisno ~ U(283, 303) K, ground
    heat flux forcing hs ~ U(0, 40) W/m2 (warming: no layer reaches freezing
    within 30 days), dhsdT = -4*0.97*sigma*T_top^3. The Richards lower boundary
    is BC_FLUX (free drainage), whose qcharge is the drainage the calibration
    loss uses.

    freezing=True draws columns around the freezing point instead, so that
    layers freeze and thaw: t_soisno ~ U(266, 280) K, a quarter of the water of
    each layer below 273.15 K frozen, hs ~ U(-60, 30) W/m2. Still no snow, no
    standing surface water and no excess ice.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sand, clay, om = _texture(rng, nc)
    dzs, zs, zis = clm5_grid(nlevsoi, nlevgrnd)
    props = pedotransfer(sand, clay, om, zs)
    snl = np.zeros(nc, dtype=np.int64)
    z, dz, zi = _geometry(nc, snl, np.zeros((nlevsno, nc)), nlevsno, nlevsoi, nlevgrnd, 0.0)
    nlev = nlevsno + nlevgrnd
    liq = np.zeros((nlev, nc))
    liq[nlevsno:] = rng.uniform(0.3, 0.95, (nlevgrnd, nc)) * props["watsat"] * dzs[:, None] * DENH2O
    t = np.zeros((nlev, nc))
    ice = np.zeros((nlev, nc))
    if freezing:
        t[nlevsno:] = rng.uniform(266.0, 280.0, (nlevgrnd, nc))
        frozen = t[nlevsno:] < TFRZ
        ice[nlevsno:] = np.where(frozen, 0.25 * liq[nlevsno:], 0.0)
        liq[nlevsno:] = np.where(frozen, 0.75 * liq[nlevsno:], liq[nlevsno:])
    else:
        t[nlevsno:] = rng.uniform(283.0, 303.0, (nlevgrnd, nc))
    nb = rng.integers(5, nlevsoi + 1, nc).astype(np.int64)
    hs = rng.uniform(-60.0, 30.0, nc) if freezing else rng.uniform(0.0, 40.0, nc)
    dhsdT = -4.0 * 0.97 * SB * t[nlevsno] ** 3
    return dict(
        nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dtime=1800.0,
        snl=snl, nbedrock=nb, z=z, dz=dz, zi=zi, t_soisno=t, h2osoi_liq=liq,
        h2osoi_ice=ice, excess_ice=np.zeros((nlevgrnd, nc)),
        watsat=props["watsat"], tkmg=props["tkmg"], tkdry=props["tkdry"],
        csol=props["csol"], bsw=props["bsw"], sucsat=props["sucsat"], hksat=props["hksat"],
        frac_sno_eff=np.zeros(nc), frac_h2osfc=np.zeros(nc),
        hs_top_snow=hs.copy(), hs_soil=hs.copy(), hs_top=hs.copy(), hs_h2osfc=hs.copy(),
        dhsdT=dhsdT, sabg_lyr_col=np.zeros((nlevsno + 1, nc)), eflx_bot=np.zeros(nc),
        t_h2osfc=t[nlevsno].copy(), h2osfc=np.zeros(nc), h2osno_no_layers=np.zeros(nc),
        snow_depth=np.zeros(nc), int_snow=np.zeros(nc),
        lower_bc=1, qflx_infl=rng.uniform(0.0, 2.0e-3, nc), zwt=np.full(nc, 5.0),
        icefrac=np.zeros((nlevgrnd, nc)), qflx_rootsoi=np.zeros((nlevsoi, nc)),
        smpmin=np.full(nc, -1.0e8), smp_l=np.zeros((nlevgrnd, nc)),
        hk_l=np.zeros((nlevgrnd, nc)), qcharge=np.zeros(nc), nsubsteps=np.zeros(nc),
        nrejected=np.zeros(nc))


def zengdecker_columns(nc=4096, seed=4, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                       frozen_fraction=0.3, zwt_range=(0.3, 12.0)):
    """Columns for the Zeng-Decker 2009 step (the reference's default method with
    the BC_AQUIFER lower boundary): texture and pedotransfer as config 0, liquid
    water theta ~ U(0.2, 0.95) x watsat, ice on a `frozen_fraction` of the
    columns (ice fraction ~ U(0, 0.5) of the pore water, those columns' soil at
    U(260, 273) K, others U(274, 300) K), water table ~ U(zwt_range) m (so both
    'inside the soil column' and 'below it' occur), infiltration ~ U(0, 2e-3)
    mm/s, root sink ~ U(0, 1e-5) mm/s per layer, smpmin = -1e8 mm. h2osoi_vol
    is the layers' total volumetric water content."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sand, clay, om = _texture(rng, nc)
    dzs, zs, zis = clm5_grid(nlevsoi, nlevgrnd)
    props = pedotransfer(sand, clay, om, zs)
    snl = np.zeros(nc, dtype=np.int64)
    z, dz, zi = _geometry(nc, snl, np.zeros((nlevsno, nc)), nlevsno, nlevsoi, nlevgrnd, 0.0)
    nlev = nlevsno + nlevgrnd
    theta = rng.uniform(0.2, 0.95, (nlevgrnd, nc)) * props["watsat"]
    frozen = rng.uniform(0.0, 1.0, nc) < frozen_fraction
    fice = np.where(frozen[None, :], rng.uniform(0.0, 0.5, (nlevgrnd, nc)), 0.0)
    liq = np.zeros((nlev, nc))
    ice = np.zeros((nlev, nc))
    liq[nlevsno:] = (1.0 - fice) * theta * dzs[:, None] * DENH2O
    ice[nlevsno:] = fice * theta * dzs[:, None] * DENICE
    t = np.zeros((nlev, nc))
    t[nlevsno:] = np.where(frozen[None, :], rng.uniform(260.0, 273.0, (nlevgrnd, nc)),
                           rng.uniform(274.0, 300.0, (nlevgrnd, nc)))
    vol = liq[nlevsno:] / (dzs[:, None] * DENH2O) + ice[nlevsno:] / (dzs[:, None] * DENICE)
    return dict(
        nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, dtime=1800.0,
        z=z, dz=dz, zi=zi, watsat=props["watsat"], hksat=props["hksat"], bsw=props["bsw"],
        sucsat=props["sucsat"], smpmin=np.full(nc, -1.0e8), h2osoi_liq=liq, h2osoi_ice=ice,
        h2osoi_vol=np.ascontiguousarray(vol), t_soisno=t,
        zwt=rng.uniform(zwt_range[0], zwt_range[1], nc),
        qflx_infl=rng.uniform(0.0, 2.0e-3, nc),
        qflx_rootsoi=rng.uniform(0.0, 1.0e-5, (nlevsoi, nc)),
        icefrac=np.zeros((nlevgrnd, nc)), smp_l=np.zeros((nlevgrnd, nc)),
        hk_l=np.zeros((nlevgrnd, nc)), qcharge=np.zeros(nc), qflx_deficit=np.zeros(nc))


# column_varcon.jl:7-11 and landunit_varcon.jl: soil/crop/urban/lake landunit kinds
ICOL_ROOF, ICOL_SUNWALL, ICOL_SHADEWALL, ICOL_ROAD_IMPERV, ICOL_ROAD_PERV = 71, 72, 73, 74, 75
ISTSOIL, ISTCROP, ISTDLAK = 1, 2, 5


def ground_heat_flux_inputs(nc=4096, seed=6, nlevsno=NLEVSNO, nlevgrnd=NLEVGRND,
                            max_patches=12, zero_weight_fraction=0.05,
                            building_temp_method=0):
    """Subgrid structure and forcing for the ground heat flux.

    Columns are laid out as CLM's subgrid initialisation does (patches of a
    column contiguous, in column order): 70 % natural-vegetation columns with
    1..max_patches patches, 12 % crop columns with 1-2 patches, 10 % urban
    columns (landunits of five columns: roof, sunwall, shadewall, impervious
    and pervious road, one patch each) and 8 % lake columns (masked out). Patch
    weights sum to one per column; a `zero_weight_fraction` of the natural
    patches get weight 0. Snow layers 0..nlevsno per column, temperatures
    250-300 K, radiative and turbulent fluxes in physical ranges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = rng.choice(4, size=nc, p=[0.70, 0.12, 0.10, 0.08])
    col_itype = np.zeros(nc, dtype=np.int64)
    col_lun = np.zeros(nc, dtype=np.int64)
    lun_urb, lun_type = [], []
    npatch = np.zeros(nc, dtype=np.int64)
    n_urban = 0
    for c in range(nc):
        k = kinds[c]
        if k == 2:
            if n_urban % 5 == 0 or c == 0 or kinds[c - 1] != 2:
                lun_urb.append(True)
                lun_type.append(8)
            col_itype[c] = ICOL_ROOF + n_urban % 5
            n_urban += 1
            npatch[c] = 1
        else:
            lun_urb.append(False)
            lun_type.append({0: ISTSOIL, 1: ISTCROP, 3: ISTDLAK}[k])
            col_itype[c] = {0: ISTSOIL, 1: ISTCROP, 3: ISTDLAK}[k]
            npatch[c] = (rng.integers(1, max_patches + 1) if k == 0 else
                         rng.integers(1, 3) if k == 1 else 1)
        col_lun[c] = len(lun_urb)
    nl = len(lun_urb)
    npt = int(npatch.sum())
    patchf = np.cumsum(npatch)
    patchi = patchf - npatch + 1
    patch_column = np.repeat(np.arange(1, nc + 1, dtype=np.int64), npatch)
    patch_landunit = col_lun[patch_column - 1].astype(np.int64)
    mask_nolakec = (kinds != 3).astype(np.uint8)
    mask_nolakep = mask_nolakec[patch_column - 1].copy()
    # weights: random simplex per column, some natural patches at zero weight
    w = rng.gamma(1.0, 1.0, npt)
    nat = kinds[patch_column - 1] == 0
    zw = nat & (rng.uniform(size=npt) < zero_weight_fraction) & (npatch[patch_column - 1] > 1)
    w[zw] = 0.0
    sums = np.bincount(patch_column - 1, weights=w, minlength=nc)
    w[patchi[sums == 0] - 1] = 1.0                 # keep one weighted patch per column
    sums = np.bincount(patch_column - 1, weights=w, minlength=nc)
    wtcol = w / np.where(sums[patch_column - 1] > 0, sums[patch_column - 1], 1.0)
    snl = -rng.integers(0, nlevsno + 1, nc).astype(np.int64)
    snl[kinds >= 2] = 0
    nlev = nlevsno + nlevgrnd
    t_soisno = rng.uniform(250.0, 300.0, (nlev, nc))
    u = lambda lo, hi, n: rng.uniform(lo, hi, n)  # noqa: E731
    d = dict(
        nc=nc, np=npt, nl=nl, nlevsno=nlevsno, nlevsoi=min(20, nlevgrnd), nlevgrnd=nlevgrnd,
        building_temp_method=building_temp_method,
        patchi=patchi.astype(np.int64), patchf=patchf.astype(np.int64),
        patch_column=patch_column, patch_landunit=patch_landunit, col_itype=col_itype, snl=snl,
        urbpoi=np.array(lun_urb, dtype=np.uint8), mask_nolakec=mask_nolakec,
        mask_nolakep=mask_nolakep,
        frac_veg_nosno=rng.integers(0, 2, npt).astype(np.int64),
        emg=u(0.9, 1.0, nc), t_grnd=u(250.0, 300.0, nc), t_h2osfc=u(265.0, 300.0, nc),
        forc_lwrad=u(200.0, 420.0, nc), htvp=np.where(rng.uniform(size=nc) < 0.3, 2.8440e6, 2.501e6),
        frac_sno_eff=np.where(snl < 0, u(0.2, 1.0, nc), u(0.0, 0.2, nc)), t_soisno=t_soisno,
        sabg=u(0.0, 600.0, npt), sabg_soil=u(0.0, 600.0, npt), sabg_snow=u(0.0, 400.0, npt),
        dlrad=u(0.0, 120.0, npt), cgrnd=u(0.0, 60.0, npt),
        eflx_sh_grnd=u(-100.0, 250.0, npt), eflx_sh_snow=u(-100.0, 250.0, npt),
        eflx_sh_soil=u(-100.0, 250.0, npt), eflx_sh_h2osfc=u(-100.0, 250.0, npt),
        qflx_evap_soi=u(-5e-5, 2e-4, npt), qflx_ev_snow=u(-5e-5, 2e-4, npt),
        qflx_ev_soil=u(-5e-5, 2e-4, npt), qflx_ev_h2osfc=u(-5e-5, 2e-4, npt),
        eflx_lwrad_net=u(0.0, 150.0, npt), qflx_tran_veg=u(0.0, 1e-4, npt), wtcol=wtcol,
        sabg_lyr_patch=rng.uniform(0.0, 300.0, (nlevsno + 1, npt)),
        eflx_wasteheat_lun=u(0.0, 30.0, nl), eflx_ventilation_lun=u(-20.0, 20.0, nl),
        eflx_heat_from_ac_lun=u(0.0, 10.0, nl), eflx_traffic_lun=u(0.0, 20.0, nl),
        wtlunit_roof=u(0.1, 0.6, nl))
    for k in ("eflx_gnet", "dgnetdT", "sabg_chk", "eflx_wasteheat", "eflx_heat_from_ac",
              "eflx_traffic", "eflx_anthro"):
        d[k] = np.zeros(npt)
    d["eflx_ventilation"] = u(-5.0, 5.0, npt)     # kept when neither building method is set
    for k in ("hs_top", "hs_top_snow", "hs_soil", "hs_h2osfc", "dhsdT"):
        d[k] = np.zeros(nc)
    d["sabg_lyr_col"] = np.zeros((nlevsno + 1, nc))
    return d


def energy_balance_inputs(nc=400, seed=14, nlevsno=NLEVSNO, nlevgrnd=NLEVGRND, nlevurb=10):
    """Random multi-patch columns for the soil energy balance error alone: 1-4
    patches per column, 30 % urban column types (71-75), 10 % of the patches
    outside the non-lake mask (their prior errsoi is kept), 10 % of the columns
    unmasked, random fluxes, temperatures, factors and prior errsoi values."""
    rng = np.random.default_rng(seed)
    npc = rng.integers(1, 5, nc)
    patchf = np.cumsum(npc).astype(np.int64)
    patchi = (patchf - npc + 1).astype(np.int64)
    np_ = int(patchf[-1])
    col = np.repeat(np.arange(1, nc + 1), npc).astype(np.int64)
    itype = np.where(rng.uniform(size=nc) < 0.3, rng.integers(71, 76, nc), 1).astype(np.int64)
    mask_p = np.ones(np_, np.uint8)
    mask_p[rng.uniform(size=np_) < 0.1] = 0
    mask_u = (np.isin(itype[col - 1], (71, 72, 73, 74, 75))).astype(np.uint8)
    nl = nlevsno + max(nlevgrnd, nlevurb)
    return dict(
        nc=nc, np=np_, nlevsno=nlevsno, nlevsoi=min(20, nlevgrnd), nlevgrnd=nlevgrnd,
        nlevurb=nlevurb, patchi=patchi, patchf=patchf, patch_column=col, col_itype=itype,
        snl=-rng.integers(0, nlevsno + 1, nc).astype(np.int64),
        mask_nolakec=(rng.uniform(size=nc) < 0.9).astype(np.uint8),
        mask_nolakep=mask_p, mask_urbanp=mask_u,
        eflx_soil_grnd=rng.normal(50, 100, np_), wtcol=rng.uniform(0, 1, np_),
        xmf=rng.normal(0, 50, nc), xmf_h2osfc=rng.normal(0, 5, nc),
        frac_h2osfc=rng.uniform(0, 0.3, nc), t_h2osfc=rng.uniform(270, 280, nc),
        t_h2osfc_bef=rng.uniform(270, 280, nc), c_h2osfc=rng.uniform(1e3, 1e5, nc),
        eflx_h2osfc_to_snow=rng.normal(0, 3, nc),
        eflx_building_heat_errsoi=rng.normal(0, 10, nc),
        frac_sno_eff=rng.uniform(0, 1, nc), t_soisno=rng.uniform(260, 290, (nl, nc)),
        t_ssbef=rng.uniform(260, 290, (nl, nc)), fact=rng.uniform(1e-4, 1e-2, (nl, nc)),
        errsoi_patch=rng.normal(0, 1, np_), errsoi_col=np.full(nc, -7.0))


def coupled_surface_columns(nc=2000, seed=15, nlevsno=NLEVSNO, nlevgrnd=NLEVGRND):
    """Heat-step columns (config 1 with surface water, partial snow cover) and a
    patch layer over them for the ground heat flux: the subgrid and patch
    forcing of ground_heat_flux_inputs with every column vegetated or crop and
    unmasked, and the column temperatures, snow layers and fractions taken from
    the heat columns (t_grnd the area-weighted surface temperature). Returns
    (heat, ghf) dicts that share the column arrays."""
    h = heat_config1(nc=nc, seed=seed, nlevsno=nlevsno, nlevgrnd=nlevgrnd, h2osfc_variant=True)
    h.update(heat_outputs(h))
    rng = np.random.default_rng(seed + 7)
    ns = nlevsno
    h["frac_sno_eff"] = np.where(h["snl"] < 0, rng.uniform(0.3, 0.7, nc), 0.0)
    g = ground_heat_flux_inputs(nc=nc, seed=seed, nlevsno=nlevsno, nlevgrnd=nlevgrnd)
    g["urbpoi"][:] = 0
    g["col_itype"][:] = ISTSOIL
    g["mask_nolakec"][:] = 1
    g["mask_nolakep"][:] = 1
    cols = np.arange(nc)
    top = h["snl"] + ns
    fs, fw = h["frac_sno_eff"], h["frac_h2osfc"]
    g["snl"] = h["snl"]
    g["t_soisno"] = np.nan_to_num(h["t_soisno"])
    g["t_h2osfc"] = h["t_h2osfc"]
    g["frac_sno_eff"] = fs
    g["t_grnd"] = (fs * g["t_soisno"][top, cols] + (1.0 - fs - fw) * g["t_soisno"][ns] +
                   fw * h["t_h2osfc"])
    return h, g


def applied_ground_flux(h0, h1):
    """The ground heat flux a heat step applied, per column, from its inputs h0
    and outputs h1: over the snow fraction the top-layer flux hs_top_snow +
    dhsdT*dT_top plus the solar absorbed in the lower snow layers and soil layer
    1; over bare soil hs_soil + dhsdT*dT_1; over surface water hs_h2osfc +
    dhsdT*dT_h2osfc (soil_temperature.jl set_rhs_vec!, :1365-1542). Without snow
    layers the top layer is soil layer 1 and takes hs_top_snow."""
    ns, nc = h0["nlevsno"], h0["nc"]
    cols = np.arange(nc)
    top = h0["snl"] + ns
    t0, t1 = h0["t_soisno"], h1["t_soisno"]
    g = h0["dhsdT"]
    dt_top = t1[top, cols] - t0[top, cols]
    dt_1 = t1[ns] - t0[ns]
    dt_h = h1["t_h2osfc"] - h0["t_h2osfc"]
    fs, fw = h0["frac_sno_eff"], h0["frac_h2osfc"]
    sab = h0["sabg_lyr_col"]
    rows = np.arange(ns + 1)[:, None]                  # level k+1 is layer j = k+1-ns
    below_top = (rows > top[None, :]) & (rows <= ns)
    sab_lower = np.where(below_top, sab, 0.0).sum(0)
    snow = (fs * (h0["hs_top_snow"] + g * dt_top + sab_lower) +
            (1.0 - fs - fw) * (h0["hs_soil"] + g * dt_1) + fw * (h0["hs_h2osfc"] + g * dt_h))
    bare = (h0["hs_top_snow"] + g * dt_1 - fw * (h0["hs_soil"] + g * dt_1) +
            fw * (h0["hs_h2osfc"] + g * dt_h))
    return np.where(h0["snl"] < 0, snow, bare)


def energy_balance_from_heat(h0, h1):
    """clmsoil_seb_args fields for one patch of weight 1 per column, the applied
    ground flux as eflx_soil_grnd and the heat step's own outputs."""
    nc, ns = h0["nc"], h0["nlevsno"]
    one = np.arange(1, nc + 1, dtype=np.int64)
    return dict(
        nc=nc, np=nc, nlevsno=ns, nlevsoi=h0["nlevsoi"], nlevgrnd=h0["nlevgrnd"], nlevurb=0,
        patchi=one, patchf=one.copy(), patch_column=one.copy(),
        col_itype=np.full(nc, ISTSOIL, dtype=np.int64), snl=h0["snl"],
        mask_nolakec=np.ones(nc, np.uint8), mask_nolakep=np.ones(nc, np.uint8),
        mask_urbanp=np.zeros(nc, np.uint8), eflx_soil_grnd=applied_ground_flux(h0, h1),
        wtcol=np.ones(nc), xmf=h1["xmf"], xmf_h2osfc=h1["xmf_h2osfc"],
        frac_h2osfc=h1["frac_h2osfc"], t_h2osfc=h1["t_h2osfc"], t_h2osfc_bef=h0["t_h2osfc"],
        c_h2osfc=h1["c_h2osfc"], eflx_h2osfc_to_snow=h1["eflx_h2osfc_to_snow"],
        eflx_building_heat_errsoi=np.zeros(nc), frac_sno_eff=h0["frac_sno_eff"],
        t_soisno=np.nan_to_num(h1["t_soisno"]), t_ssbef=np.nan_to_num(h0["t_soisno"]),
        fact=np.where(np.isnan(h1["fact"]), 1.0, h1["fact"]),
        errsoi_patch=np.zeros(nc), errsoi_col=np.zeros(nc))


def water_table_columns(nc=4096, seed=16, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                        zwt_range=(0.3, 12.0), frozen_fraction=0.3):
    """Columns for the water-table update after an aquifer step: the soil and
    water state of zengdecker_columns, a recharge qcharge ~ U(-2e-4, 2e-4) mm/s
    (half of the columns draining, half filling, enough to move the table across
    layer interfaces over 1800 s), aquifer water wa ~ U(3000, 5000) mm, excess
    ice on a fifth of the columns, and frozen layers below unfrozen ones on the
    frozen columns (so a frost table exists)."""
    d = zengdecker_columns(nc=nc, seed=seed, nlevsno=nlevsno, nlevsoi=nlevsoi,
                           nlevgrnd=nlevgrnd, frozen_fraction=frozen_fraction,
                           zwt_range=zwt_range)
    rng = np.random.default_rng(seed + 3)
    ns = nlevsno
    t = d["t_soisno"]
    frozen = d["h2osoi_ice"][ns:ns + nlevsoi].sum(0) > 0
    kf = rng.integers(1, nlevsoi, nc)
    lay = np.arange(1, nlevgrnd + 1)[:, None]
    t[ns:] = np.where(frozen[None, :] & (lay >= kf[None, :]), rng.uniform(260.0, 273.0, (nlevgrnd, nc)),
                      rng.uniform(274.0, 290.0, (nlevgrnd, nc)))
    exc = np.zeros((nlevgrnd, nc))
    has = rng.uniform(size=nc) < 0.2
    exc[:, has] = rng.uniform(0.0, 30.0, (nlevgrnd, int(has.sum())))
    d.update(qcharge=rng.uniform(-2e-4, 2e-4, nc) * rng.uniform(0.0, 20.0, nc),
             wa=rng.uniform(3000.0, 5000.0, nc), zwt_perched=np.zeros(nc), frost_table=np.zeros(nc),
             h2osoi_vol=d["h2osoi_vol"].copy(), qflx_drain=np.full(nc, 7.0),
             qflx_rsub_sat=np.full(nc, 7.0), qflx_drain_perched=np.full(nc, 7.0),
             excess_ice=exc, eff_porosity=np.zeros((nlevgrnd, nc)))
    return d


def moisture_water_table_columns(nc=4096, seed=40, nlevsno=NLEVSNO, nlevsoi=NLEVSOI,
                                 nlevgrnd=NLEVGRND):
    """Columns for the water tables diagnosed from the moisture profile after a
    step without an aquifer layer (perched, theta-based, bedrock clamp): the
    geometry, soil and temperature state of water_table_columns (frozen layers
    below unfrozen ones on 30 % of the columns), bedrock at layer nbedrock ~
    U{3..nlevsoi}, and a saturation profile that is wet (saturation ~ U(0.91, 1))
    from a per-column layer k_sat ~ U{1..nlevsoi+1} down and drier (~ U(0.3, 0.89))
    above it, so the 90 % level falls inside, above and below the column. Frozen
    layers hold 30-100 % of their water as ice."""
    d = water_table_columns(nc=nc, seed=seed, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd)
    rng = np.random.default_rng(seed + 7)
    ns = nlevsno
    lay = np.arange(1, nlevgrnd + 1)[:, None]
    k_sat = rng.integers(1, nlevsoi + 2, nc)
    sat = np.where(lay >= k_sat[None, :], rng.uniform(0.91, 1.0, (nlevgrnd, nc)),
                   rng.uniform(0.3, 0.89, (nlevgrnd, nc)))
    vol = sat * d["watsat"]
    dz = d["dz"][ns:]
    frozen = d["t_soisno"][ns:] <= TFRZ
    f_ice = np.where(frozen, rng.uniform(0.3, 1.0, (nlevgrnd, nc)), 0.0)
    d["h2osoi_ice"][ns:] = f_ice * vol * dz * DENICE
    d["h2osoi_liq"][ns:] = (1.0 - f_ice) * vol * dz * DENH2O
    d["nbedrock"] = rng.integers(3, nlevsoi + 1, nc).astype(np.int64)
    d["zwt_perched"] = np.full(nc, -1.0)
    d["frost_table"] = np.full(nc, -1.0)
    d["h2osoi_vol"] = np.full((nlevgrnd, nc), -1.0)
    return d


def surface_hydrology_columns(nc=4096, seed=60, nlevsno=NLEVSNO, nlevgrnd=NLEVGRND):
    """Columns for the surface hydrology before the Richards step: one landunit
    per column, 55 % soil, 15 % crop, 15 % urban (roof, sunwall, shadewall,
    impervious and pervious road in turn), 8 % deep lake, 7 % wetland, on
    gridcells of about 4 columns. Masks as CLM's filters: no-lake, hydrology
    (soil, crop, pervious road), urban. Forcing: rain plus snowmelt ~ U(0, 5e-3)
    mm/s with a fifth of the columns at U(0, 5e-2) (beyond hksat, so
    infiltration excess occurs); flood water on a tenth of the gridcells; 30 %
    snow-covered columns; surface water fractions with 40 % of the columns dry;
    ponded h2osfc above and below its threshold (a few slightly negative);
    frozen top layers on 30 % of the columns; water, perched and frost tables
    that take both TOPMODEL branches; VIC inputs; hillslope flags."""
    rng = np.random.default_rng(seed)
    ng = max(1, nc // 4)
    u = rng.uniform(size=nc)
    lun = np.where(u < 0.55, 1, np.where(u < 0.70, 2, np.where(u < 0.85, 8,
                   np.where(u < 0.93, 5, 6)))).astype(np.int64)
    urban = lun == 8
    itype = lun.copy()
    itype[urban] = 71 + (np.cumsum(urban)[urban] - 1) % 5
    hyd = (lun == 1) | (lun == 2) | (itype == 75)
    nlev = nlevsno + nlevgrnd
    wet = rng.uniform(size=nc) < 0.2
    rain = np.where(wet, rng.uniform(0.0, 5e-2, nc), rng.uniform(0.0, 5e-3, nc))
    snl = np.where(rng.uniform(size=nc) < 0.3, -rng.integers(1, 4, nc), 0).astype(np.int64)
    fh = np.where(rng.uniform(size=nc) < 0.4, 0.0, rng.uniform(0.0, 0.3, nc))
    frozen = rng.uniform(size=nc) < 0.3
    icefrac = np.zeros((nlevgrnd, nc))
    icefrac[:3, frozen] = rng.uniform(0.0, 0.8, (3, int(frozen.sum())))
    zwt = rng.uniform(0.1, 5.0, nc)
    h2osfc = np.where(rng.uniform(size=nc) < 0.3, 0.0, rng.uniform(0.0, 100.0, nc))
    h2osfc[rng.uniform(size=nc) < 0.02] = -rng.uniform(0.0, 1e-3)
    liq = np.zeros((nlev, nc))
    liq[nlevsno] = rng.uniform(0.0, 1.5, nc)
    top_max = rng.uniform(100.0, 500.0, nc)
    sentinel = lambda: np.full(nc, 7.0)
    d = dict(nc=nc, ng=ng, nl=nc, nlevsno=nlevsno, nlevsoi=20, nlevgrnd=nlevgrnd,
             mask_nolakec=(lun != 5).astype(np.uint8), mask_hydrologyc=hyd.astype(np.uint8),
             mask_urbanc=urban.astype(np.uint8),
             col_gridcell=rng.integers(1, ng + 1, nc).astype(np.int64),
             col_landunit=np.arange(1, nc + 1, dtype=np.int64), lun_itype=lun, col_itype=itype,
             snl=snl, col_cold=np.where(rng.uniform(size=nc) < 0.5, -9999,
                                        rng.integers(1, nc + 1, nc)).astype(np.int64),
             col_urbpoi=urban.astype(np.uint8),
             is_hillslope_column=(rng.uniform(size=nc) < 0.3).astype(np.uint8),
             col_active=(rng.uniform(size=nc) < 0.9).astype(np.uint8),
             qflx_floodg=np.where(rng.uniform(size=ng) < 0.1, rng.uniform(0.0, 1e-3, ng), 0.0),
             frost_table=rng.uniform(0.0, 6.0, nc), zwt=zwt,
             zwt_perched=zwt * rng.uniform(0.0, 1.0, nc), wtfact=rng.uniform(0.2, 0.5, nc),
             b_infil=rng.uniform(0.01, 0.5, nc), top_max_moist=top_max,
             top_moist_limited=top_max * rng.uniform(0.0, 1.0, nc),
             qflx_rain_plus_snomelt=rain, qflx_snow_h2osfc=rng.uniform(0.0, 1e-4, nc),
             qflx_liqevap_from_top_layer=rng.normal(1e-5, 2e-5, nc),
             qflx_ev_soil=rng.normal(1e-5, 2e-5, nc), qflx_ev_h2osfc=rng.normal(1e-5, 2e-5, nc),
             frac_sno_eff=np.where(snl < 0, rng.uniform(0.2, 1.0, nc) * (1.0 - fh), 0.0),
             frac_h2osfc=fh, frac_h2osfc_nosnow=np.minimum(1.0, fh + rng.uniform(0.0, 0.7, nc)),
             icefrac=icefrac, hksat=rng.uniform(0.001, 0.05, (nlevgrnd, nc)),
             topo_slope=rng.uniform(0.0, 30.0, nc), h2osfc_thresh=rng.uniform(0.0, 50.0, nc),
             h2osoi_liq=liq, h2osfc=h2osfc, xs_urban=np.zeros(nc))
    for k in ("qflx_floodc", "fsat", "fcov", "qflx_sat_excess_surf", "qflx_top_soil",
              "qflx_in_soil", "qflx_top_soil_to_h2osfc", "qinmax", "qflx_infl_excess",
              "qflx_in_soil_limited", "qflx_in_h2osfc", "qflx_infl_excess_surf",
              "qflx_h2osfc_surf", "qflx_h2osfc_drain", "qflx_infl", "qflx_surf"):
        d[k] = sentinel()
    return d


def root_sink_columns(nc=4096, seed=80, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND, max_patches=12,
                      scatter=False):
    """Columns and patches for the root water uptake: 0..max_patches patches per
    column (CLM layout: a column's patches contiguous, columns in order; with
    `scatter` the columns' patch blocks are shuffled over the patch arrays),
    90 % of the patches active, weights summing to 1 over each column, root
    profiles rootr ~ exp(-z/d) with d ~ U(0.2, 1.5) m normalised over the soil
    layers, transpiration ~ U(0, 5e-4) mm/s with a fifth of the patches at 0,
    and the column transpiration as the weighted sum over active patches. 85 %
    of the columns are in the hydrology filter."""
    rng = np.random.default_rng(seed)
    npc = rng.integers(0, max_patches + 1, nc).astype(np.int64)
    np_ = int(npc.sum())
    order = rng.permutation(nc) if scatter else np.arange(nc)
    start = np.zeros(nc, dtype=np.int64)
    start[order] = np.concatenate([[0], np.cumsum(npc[order])[:-1]])
    patchi = start + 1
    owner = np.repeat(order, npc[order]).astype(np.int64)      # column of each patch
    active = (rng.uniform(size=np_) < 0.9).astype(np.uint8)
    w = rng.uniform(0.05, 1.0, np_)
    tot = np.bincount(owner, weights=w, minlength=nc)
    wtcol = w / tot[owner]
    _, zs, _ = clm5_grid(nlevsoi, nlevgrnd)
    depth = rng.uniform(0.2, 1.5, np_)
    prof = np.exp(-zs[:nlevsoi, None] / depth[None, :])
    rootr = np.zeros((nlevgrnd, np_))
    rootr[:nlevsoi] = prof / prof.sum(0)
    qt = np.where(rng.uniform(size=np_) < 0.2, 0.0, rng.uniform(0.0, 5e-4, np_))
    qcol = np.bincount(owner, weights=qt * wtcol * active, minlength=nc)
    return dict(nc=nc, np=np_, nlevsno=NLEVSNO, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd,
                mask_hydrologyc=(rng.uniform(size=nc) < 0.85).astype(np.uint8), patchi=patchi,
                npatches=npc, active=active, wtcol=wtcol, rootr_patch=rootr,
                qflx_tran_veg_patch=qt, qflx_tran_veg_col=qcol,
                rootr_col=np.full((nlevgrnd, nc), 7.0), qflx_rootsoi_col=np.full((nlevsoi, nc), 7.0))


# varcon.jl:291 (the default 10 lake levels of a 50 m lake)
DZLAK_10 = np.array([0.1, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0, 7.0, 10.45, 10.45])


def lake_columns(nc=4096, seed=90, nlevsno=NLEVSNO, nlevsoi=NLEVSOI, nlevgrnd=NLEVGRND,
                 max_snow=6, uniform_tk=None):
    """Lake columns for the lake temperature solve: 10 lake levels of the
    default grid scaled to a depth ~ U(2, 60) m, 0..max_snow snow layers (dz ~
    U(0.02, 0.2) m stacked above the surface), the CLM5 soil grid below the lake;
    heat capacities ~ 2-4e6 J/m3/K times thickness, conductivities snow ~
    U(0.1, 0.5), soil ~ U(0.5, 3) with 0 at the bottom interface, lake ~ U(0.5, 50)
    (molecular plus eddy); temperatures snow 250-273 K, lake 274-295 K, soil
    275-290 K; solar sources in snow and lake, flux into the top ~ N(0, 100) W/m2.
    `uniform_tk` sets every conductivity to one value. 90 % of the columns are
    in the lake filter."""
    rng = np.random.default_rng(seed)
    nlev = nlevsno + nlevgrnd
    ns = rng.integers(0, max_snow + 1, nc)
    snl = (-ns).astype(np.int64)
    snow_dz = rng.uniform(0.02, 0.2, (nlevsno, nc))
    z, dz, zi = _geometry(nc, snl, snow_dz, nlevsno, nlevsoi, nlevgrnd, fill=0.0)
    depth = rng.uniform(2.0, 60.0, nc)
    dzl = DZLAK_10[:, None] * (depth / DZLAK_10.sum())[None, :]
    zl = np.cumsum(dzl, 0) - 0.5 * dzl
    cv = rng.uniform(2e6, 4e6, (nlev, nc)) * dz
    cv[:nlevsno] = rng.uniform(1e5, 2e6, (nlevsno, nc)) * dz[:nlevsno]
    tk = np.concatenate([rng.uniform(0.1, 0.5, (nlevsno, nc)), rng.uniform(0.5, 3.0, (nlevgrnd, nc))])
    tk[nlev - 1] = 0.0
    tkl = rng.uniform(0.5, 50.0, (10, nc))
    tts = rng.uniform(0.5, 3.0, nc)
    if uniform_tk is not None:
        tk[:] = uniform_tk
        tkl[:] = uniform_tk
        tts[:] = uniform_tk
    t = np.concatenate([rng.uniform(250.0, 273.0, (nlevsno, nc)), rng.uniform(275.0, 290.0, (nlevgrnd, nc))])
    phi = np.sort(rng.uniform(0.0, 100.0, (10, nc)), axis=0)[::-1].copy()
    return dict(nc=nc, nlevsno=nlevsno, nlevsoi=nlevsoi, nlevgrnd=nlevgrnd, nlevlak=10,
                mask_lakec=(rng.uniform(size=nc) < 0.9).astype(np.uint8), snl=snl, z=z, dz=dz,
                cv=cv, tk=tk, sabg_lyr_col=rng.uniform(0.0, 50.0, (nlevsno + 1, nc)),
                z_lake=zl, dz_lake=dzl, cv_lake=rng.uniform(3.5e6, 4.2e6, (10, nc)) * dzl,
                tk_lake=tkl, phi=phi, phi_soil=rng.uniform(0.0, 10.0, nc),
                fin=rng.normal(0.0, 100.0, nc), tktopsoillay=tts, t_soisno=t,
                t_lake=rng.uniform(274.0, 295.0, (10, nc)))


def lvt_columns(nc=4096, seed=110, nlevdecomp=NLEVSOI, nlevdecomp_full=NLEVGRND, npools=7,
                ntypes=2, cwd_pools=(3,), ng=None):
    """Columns for the vertical transport of the soil organic matter pools: the
    CLM5 grid's first nlevdecomp levels, bedrock at layer U{3..nlevdecomp}, active
    layer thickness ~ U(0, 4) m this year and last year (a fifth of the columns
    0: frozen, no mixing; those below 2 m mix by cryoturbation), pools
    [npools, nlevdecomp_full, nc] (the reference's [c, j, s]) with carbon
    ~ U(0.5, 1.5) * 1000 * exp(-z) g/m3 and nitrogen at C/N ~ U(10, 30), sources
    ~ U(-0.01, 0.05) g/m3 per step; pool indices in cwd_pools are coarse woody
    debris (CENTURY's fourth pool). 90 % of the columns in the filter."""
    rng = np.random.default_rng(seed)
    dzs, zs, zis = clm5_grid(NLEVSOI, NLEVGRND)
    nd = nlevdecomp
    ng = ng or max(1, nc // 4)
    alt = np.where(rng.uniform(size=nc) < 0.2, 0.0, rng.uniform(0.0, 4.0, nc))
    shape = (npools, nlevdecomp_full, nc)
    prof = np.exp(-zs[:nlevdecomp_full])[None, :, None]
    cpool = rng.uniform(0.5, 1.5, shape) * 1000.0 * prof
    conc = [cpool]
    if ntypes >= 2:
        conc.append(cpool / rng.uniform(10.0, 30.0, shape))
    for _ in range(2, ntypes):
        conc.append(cpool * rng.uniform(0.9, 1.1, shape))
    source = [rng.uniform(-0.01, 0.05, shape) for _ in range(ntypes)]
    is_cwd = [i in cwd_pools for i in range(npools)]
    return dict(nc=nc, nlevdecomp=nd, nlevdecomp_full=nlevdecomp_full, npools=npools,
                ntypes=ntypes, is_cwd=is_cwd,
                mask_bgc_soilc=(rng.uniform(size=nc) < 0.9).astype(np.uint8),
                nbedrock=rng.integers(3, nd + 1, nc).astype(np.int64),
                col_gridcell=rng.integers(1, ng + 1, nc).astype(np.int64),
                latdeg=rng.uniform(-80.0, 80.0, ng),
                altmax=alt, altmax_lastyear=np.where(alt > 0, rng.uniform(0.0, 4.0, nc), 0.0),
                zsoi=zs[:nd].copy(), dzsoi_decomp=dzs[:nd].copy(), zisoi=zis[:nd + 2].copy(),
                som_diffus_coef=np.full((nd + 1, nc), 7.0), som_adv_coef=np.full((nd + 1, nc), 7.0),
                conc=conc, source=source, tendency=[np.full(shape, 7.0) for _ in range(ntypes)])


def ch4_columns(nc=4096, seed=130, nlevsoi=NLEVSOI, sat=0, ng=None):
    """Columns for the methane and oxygen transport step (one fraction of ch4!):
    the CLM5 grid's first nlevsoi levels; porosity U(0.4, 0.6), volumetric water
    U(0.1, 1.05) of it (some above saturation), a tenth of the levels frozen
    (ice share U(0.3, 1)), Clapp-Hornberger b U(3, 12), organic matter
    U(0, 160) kg/m3 (some above the peat density 130), soil temperatures
    U(265, 300) K; ponded water U(0, 200) mm on U(0, 1) of the area (a third
    dry); ground conductance U(0, 0.05) m/s (a tenth 0). Atmospheric CH4 1.6-2.0
    ppm and O2 20.9 % at U(250, 305) K and 1 atm. Concentrations: CH4
    U(0, 0.5) mol/m3, O2 U(0.1, 8) mol/m3; process rates (mol/m3/s): production
    U(0, 1e-6), oxidation U(0, 5e-7) with twice that O2, aerenchyma U(0, 2e-7)
    (O2 in U(0, 1e-6)), ebullition U(0, 1e-7) on half the levels, decomposition
    O2 demand U(0, 1e-5); a fifth of the levels starved (concentrations
    scaled by 1e-4) so that the O2/CH4 competition binds. jwt U{0..nlevsoi} for
    the unsaturated fraction (sat 0), 0 for the saturated one. 90 % of the columns
    in the filter."""
    rng = np.random.default_rng(seed)
    dzs, _, _ = clm5_grid(NLEVSOI, NLEVGRND)
    nl = nlevsoi
    ng = ng or max(1, nc // 4)
    shp = (nl, nc)
    u = lambda lo, hi: rng.uniform(lo, hi, shp)
    dz = np.repeat(dzs[:nl, None], nc, axis=1).copy()
    watsat = u(0.4, 0.6)
    vol = watsat * u(0.1, 1.05)
    frozen = rng.uniform(size=shp) < 0.1
    icef = np.where(frozen, u(0.3, 1.0), 0.0)
    t_soisno = np.where(frozen, u(263.0, 273.15), u(265.0, 300.0))
    liq = np.minimum(vol, watsat) * dz * DENH2O * (1.0 - icef)
    ice = np.minimum(vol, watsat) * dz * DENICE * icef
    starved = rng.uniform(size=shp) < 0.2
    scale = np.where(starved, 1e-4, 1.0)
    tatm = rng.uniform(250.0, 305.0, ng)
    rgas = 8.3145
    c_atm = np.stack([rng.uniform(1.6e-6, 2.0e-6, ng) * 101325.0 / rgas / tatm,
                      0.209 * 101325.0 / rgas / tatm,
                      4.0e-4 * 101325.0 / rgas / tatm])
    frac = np.where(rng.uniform(size=nc) < 1.0 / 3.0, 0.0, rng.uniform(0.0, 1.0, nc))
    gcond = np.where(rng.uniform(size=nc) < 0.1, 0.0, rng.uniform(0.0, 0.05, nc))
    oxid = u(0.0, 5e-7)
    jwt = rng.integers(0, nl + 1, nc) if sat == 0 else np.zeros(nc, np.int64)
    return dict(nc=nc, ldg=ng, nlevsoi=nl,
                mask=(rng.uniform(size=nc) < 0.9).astype(np.uint8),
                col_gridcell=rng.integers(1, ng + 1, nc).astype(np.int64),
                jwt=jwt.astype(np.int64), c_atm=c_atm,
                watsat=watsat, h2osoi_vol=vol, h2osoi_liq=liq, h2osoi_ice=ice, bsw=u(3.0, 12.0),
                cellorg=u(0.0, 160.0), t_soisno=t_soisno, dz=dz,
                h2osfc=rng.uniform(0.0, 200.0, nc), t_grnd=rng.uniform(265.0, 305.0, nc),
                t_h2osfc=rng.uniform(268.0, 300.0, nc), frac_h2osfc=frac, grnd_ch4_cond=gcond,
                ch4_prod_depth=u(0.0, 1e-6), o2_aere_depth=u(0.0, 1e-6),
                o2_decomp_depth=u(0.0, 1e-5), o2_oxid_depth=2.0 * oxid, ch4_oxid_depth=oxid,
                ch4_aere_depth=u(0.0, 2e-7),
                ch4_ebul_depth=np.where(rng.uniform(size=shp) < 0.5, u(0.0, 1e-7), 0.0),
                o2stress=np.full(shp, 7.0), ch4stress=np.full(shp, 7.0),
                conc_ch4=u(0.0, 0.5) * scale, conc_o2=u(0.1, 8.0) * scale,
                ch4_ebul_total=np.full(nc, 7.0), ch4_surf_aere=np.full(nc, 7.0),
                ch4_surf_ebul=np.full(nc, 7.0), ch4_surf_diff=np.full(nc, 7.0))
