"""HIP surface hydrology (one fused launch) vs the oracle's step-by-step sequence.

exp, pow and sin (fsat, qinmax, the surface water connectivity and slope) are the
operations whose last bit may differ between the GPU and the host library. The
bar per element is 1e-12 relative plus 4x the band the oracle itself moves when
every one of those calls is moved one ulp (8 jitter seeds), and at least 99.9 %
of the written elements must meet plain 1e-12 relative. Everything else in the
sequence (sums, products, min/max, the truncation rule) carries no such freedom:
with the transcendentals removed (steps that use none) the results are
bit-identical. The root water uptake has no transcendental and is bit-exact.
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, inputs
from clmsoil._abi import SH_FIELDS, SH_STEPS

TOL = 1e-12
OUTS = ("qflx_floodc", "fsat", "fcov", "qflx_sat_excess_surf", "qflx_top_soil", "qflx_in_soil",
        "qflx_top_soil_to_h2osfc", "qinmax", "qflx_infl_excess", "qflx_in_soil_limited",
        "qflx_in_h2osfc", "qflx_infl_excess_surf", "qflx_h2osfc_surf", "qflx_h2osfc_drain",
        "qflx_infl", "qflx_surf", "xs_urban", "h2osfc", "h2osoi_liq")


def run_gpu(d, dtime, **params):
    import torch
    t = api.to_device(d)
    api.surface_hydrology({k: t[k] for k in SH_FIELDS if k in t} | {"nc": d["nc"]}, dtime,
                          d["nlevsno"], d["nlevsoi"], d["nlevgrnd"], **params)
    torch.cuda.synchronize()
    return api.to_host(t)


def oracle_band(d, dtime, draws=8, **params):
    base = clone(d)
    oracle.surface_hydrology(base, dtime, **params)
    band = {k: np.zeros_like(base[k]) for k in OUTS}
    try:
        for seed in range(1, draws + 1):
            oracle.set_surfhyd_jitter(seed)
            out = clone(d)
            oracle.surface_hydrology(out, dtime, **params)
            for k in OUTS:
                band[k] = np.maximum(band[k], np.abs(out[k] - base[k]))
    finally:
        oracle.set_surfhyd_jitter(0)
    return base, band


def compare(got, ref, band):
    for k in OUTS:
        a, b = got[k], ref[k]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        m = ~np.isnan(a)
        a, b, bd = a[m], b[m], band[k][m]
        diff = np.abs(a - b)
        scale = np.maximum(np.abs(a), np.abs(b))
        assert np.all(diff <= TOL * scale + 4.0 * bd), (k, float(np.max(diff - TOL * scale - 4 * bd)))
        written = (a != 7.0) | (b != 7.0)
        if written.any():
            assert (diff <= TOL * scale)[written].mean() >= 0.999, k


@pytest.mark.gpu
@pytest.mark.parametrize("seed,params", [
    (70, {}),                                                    # the driver's settings
    (71, {"fsat_method": 2, "qinmax_method": 0}),                # VIC, unlimited qinmax
    (72, {"h2osfcflag": 0, "h2osfcflag_update": 0}),             # surface water off
    (73, {"crop_fsat_equals_zero": True, "hillslope_fsat_equals_zero": True, "fff": 2.5}),
])
def test_surface_hydrology_vs_oracle(gpu, seed, params):
    d = inputs.surface_hydrology_columns(nc=5003, seed=seed)
    got = run_gpu(d, 1800.0, **params)
    ref, band = oracle_band(d, 1800.0, **params)
    compare(got, ref, band)


@pytest.mark.gpu
@pytest.mark.parametrize("step", ["floodc", "qflx_inputs", "route", "infiltration",
                                  "surf_runoff", "urban_pond"])
def test_single_steps_bit_exact(gpu, step):
    """Steps without a transcendental, run alone on the state the full sequence
    leaves: bit-identical to the oracle."""
    d = inputs.surface_hydrology_columns(nc=3001, seed=74)
    oracle.surface_hydrology(d, 1800.0)               # fill every intermediate
    got = run_gpu(d, 1800.0, steps=SH_STEPS[step])
    ref = clone(d)
    oracle.surface_hydrology(ref, 1800.0, steps=SH_STEPS[step])
    for k in OUTS:
        assert np.array_equal(got[k], ref[k], equal_nan=True), (step, k)


@pytest.mark.gpu
def test_surface_hydrology_then_richards_infiltration(gpu):
    """The surface hydrology's qflx_infl feeds the moisture-form Richards step
    on the device: the soil takes in exactly what the surface passed down."""
    import torch
    nc = 2048
    r = inputs.richards_config0(nc=nc, seed=75)
    s = inputs.surface_hydrology_columns(nc=nc, seed=75)
    s["mask_hydrologyc"][:] = 1
    s["mask_urbanc"][:] = 0
    s["lun_itype"][:] = 1
    s["col_itype"][:] = 1
    t = api.to_device(s)
    api.surface_hydrology({k: t[k] for k in SH_FIELDS if k in t} | {"nc": nc}, 1800.0)
    rt = api.to_device(r)
    rt["qflx_infl"].copy_(t["qflx_infl"])
    st = api.MfState(rt["h2osoi_liq"], rt["smp_l"], rt["hk_l"], rt["qcharge"], rt["nsubsteps"])
    mi = api.MfIn(rt["z"], rt["zi"], rt["dz"], rt["qflx_rootsoi"], rt["watsat"], rt["hksat"],
                  rt["bsw"], rt["sucsat"], rt["icefrac"], rt["zwt"], rt["qflx_infl"], rt["smpmin"])
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    api.soilwm_moisture_form_solve(st, mi, None, rt["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
    torch.cuda.synchronize()
    h = api.to_host(rt)
    infl = api.to_host(t)["qflx_infl"]
    dliq = (h["h2osoi_liq"][12:32] - r["h2osoi_liq"][12:32]).sum(0)
    np.testing.assert_allclose(dliq, infl * 1800.0, rtol=1e-10, atol=1e-9)


RS = ("mask_hydrologyc", "patchi", "npatches", "active", "wtcol", "rootr_patch",
      "qflx_tran_veg_patch", "qflx_tran_veg_col", "rootr_col", "qflx_rootsoi_col")


@pytest.mark.gpu
@pytest.mark.parametrize("seed,scatter,max_patches", [(84, False, 12), (85, True, 12),
                                                      (86, False, 40)])
def test_root_sink_vs_oracle(gpu, seed, scatter, max_patches):
    """Bit-exact: the same products and sums in ascending patch order. Ordered
    columns take the LDS-staged path; shuffled patch blocks and 40-patch
    columns (windows wider than the LDS stage) take the direct one."""
    import torch
    d = inputs.root_sink_columns(nc=3001, seed=seed, scatter=scatter, max_patches=max_patches)
    t = api.to_device(d)
    api.device_errors()
    api.compute_effec_rootfrac_and_vert_tran_sink({k: t[k] for k in RS} | {"nc": d["nc"],
                                                                          "np": d["np"]})
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    got = api.to_host(t)
    ref = clone(d)
    oracle.compute_effec_rootfrac_and_vert_tran_sink(ref)
    assert np.array_equal(got["rootr_col"], ref["rootr_col"])
    assert np.array_equal(got["qflx_rootsoi_col"], ref["qflx_rootsoi_col"])


@pytest.mark.gpu
def test_root_sink_bad_range_flags(gpu):
    import torch
    d = inputs.root_sink_columns(nc=500, seed=87)
    d["mask_hydrologyc"][:] = 1
    d["npatches"][9] = d["np"] + 3
    t = api.to_device(d)
    api.device_errors()
    api.compute_effec_rootfrac_and_vert_tran_sink({k: t[k] for k in RS} | {"nc": d["nc"],
                                                                          "np": d["np"]})
    torch.cuda.synchronize()
    assert api.device_errors() & 16                      # CLMSOIL_DEVERR_SUBGRID
    got = api.to_host(t)
    assert np.all(got["rootr_col"][:, 9] == 7.0)
    ref = clone(d)
    ref["mask_hydrologyc"][9] = 0
    oracle.compute_effec_rootfrac_and_vert_tran_sink(ref)
    ok = np.arange(d["nc"]) != 9
    assert np.array_equal(got["rootr_col"][:, ok], ref["rootr_col"][:, ok])
