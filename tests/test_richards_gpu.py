"""HIP moisture-form Richards step vs the oracle (same seeded inputs).

Bar: h2osoi_liq, smp_l, hk_l and qcharge equal to the oracle's bit for bit
(both sides evaluate pow with clm_math.h, pinned in test_clm_math.py, and
neither fuses a multiply-add), so the north star's 1e-11 relative holds with
nothing to spare; nsubsteps identical (a differing accept/reject decision would
show here)."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise, clone
from clmsoil import api, bind, inputs

NS = 12


def run_gpu(d, cfg: api.SoilWaterMovementConfig, aquifer=False):
    import torch
    t = api.to_device(d)
    st = api.MfState(t["h2osoi_liq"], t["smp_l"], t["hk_l"], t["qcharge"], t["nsubsteps"],
                     t.get("nrejected"))
    mi = api.MfIn(t["z"], t["zi"], t["dz"], t["qflx_rootsoi"], t["watsat"], t["hksat"], t["bsw"],
                  t["sucsat"], t["icefrac"], t["zwt"], t["qflx_infl"], t.get("smpmin"), t.get("dtsub0"))
    if aquifer:
        api.soilwater_moisture_form(st, mi, t.get("mask"), t["nbedrock"], d["nlevsno"], d["nlevsoi"],
                                    d["nlevgrnd"], d["dtime"], cfg)
    else:
        api.soilwm_moisture_form_solve(st, mi, t.get("mask"), t["nbedrock"], d["nlevsno"],
                                       d["nlevsno"] + 1, d["nlevsoi"], d["dtime"], cfg,
                                       nlevgrnd=d["nlevgrnd"])
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    return api.to_host(t)


def compare(got, ref, d):
    assert_bitwise(got, ref, ("h2osoi_liq", "smp_l", "hk_l", "qcharge"))
    assert np.array_equal(got["nsubsteps"], ref["nsubsteps"], equal_nan=True)


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("G", [1, 2, 4])
@pytest.mark.parametrize("bc", [api.BC_FLUX, api.BC_ZERO_FLUX, api.BC_WATERTABLE])
def test_config0_vs_oracle(gpu, G, bc, persistent):
    nb = "random" if bc == api.BC_WATERTABLE else None
    d = inputs.richards_config0(nc=4096, seed=10 + bc, nbedrock=nb, lower_bc=bc)
    d["nrejected"] = np.full(d["nc"], np.nan)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=bc)
    ref = clone(d)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    api.set_tuning(G, 256, 4, 256)
    api.set_richards_schedule(persistent)
    got = run_gpu(d, cfg)
    api.set_richards_schedule(True)
    compare(got, ref, d)
    assert np.array_equal(got["nrejected"], ref["nrejected"], equal_nan=True)


@pytest.mark.gpu
def test_aquifer_pass_and_expensive(gpu):
    d = inputs.richards_config0(nc=2048, seed=21, nbedrock="random", lower_bc=api.BC_WATERTABLE)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    ref = clone(d)
    oracle.soilwater_moisture_form(ref, cfg.to_abi())
    api.set_tuning(4, 256, 4, 256)
    got = run_gpu(d, cfg, aquifer=True)
    compare(got, ref, d)
    cfg.flux_calculation = cfg.expensive
    d2 = inputs.richards_config0(nc=2048, seed=22, lower_bc=api.BC_FLUX)
    cfg.lower_boundary_condition = api.BC_FLUX
    ref = clone(d2)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    compare(run_gpu(d2, cfg), ref, d2)


@pytest.mark.gpu
def test_conservation_kat_on_gpu(gpu):
    """The reference's 12x10 zero-flux conservation case on the HIP path."""
    from test_oracle_kat import conservation_case
    d = conservation_case()
    liq0 = d["h2osoi_liq"].copy()
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    for G in (1, 4):
        api.set_tuning(G, 256, 4, 256)
        got = run_gpu(clone(d), cfg)
        dl = (got["h2osoi_liq"][5:15] - liq0[5:15]).sum(0)
        np.testing.assert_allclose(dl, d["qflx_infl"] * d["dtime"], rtol=1e-10)
        assert np.abs(dl - d["qflx_infl"] * d["dtime"]).max() < 1e-12


@pytest.mark.gpu
def test_mask_and_bad_nbedrock(gpu):
    d = inputs.richards_config0(nc=1000, seed=23)
    mask = np.zeros(d["nc"], dtype=np.uint8)
    mask[::3] = 1
    d["mask"] = mask
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    ref = clone(d)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    api.set_tuning(4, 256, 4, 256)
    got = run_gpu(d, cfg)
    compare(got, ref, d)
    assert np.all(np.isnan(got["qcharge"][1::3]))          # unmasked columns untouched
    d["nbedrock"][5] = 25                                  # > nlevsoi: the reference errors
    import torch
    t = api.to_device(d)
    st = api.MfState(t["h2osoi_liq"], t["smp_l"], t["hk_l"], t["qcharge"], t["nsubsteps"])
    mi = api.MfIn(t["z"], t["zi"], t["dz"], t["qflx_rootsoi"], t["watsat"], t["hksat"], t["bsw"],
                  t["sucsat"], t["icefrac"], t["zwt"], t["qflx_infl"])
    api.soilwm_moisture_form_solve(st, mi, None, t["nbedrock"], 12, 13, 20, 1800.0, cfg)
    torch.cuda.synchronize()
    assert api.device_errors() & 1


@pytest.mark.gpu
def test_learned_first_substep_input(gpu):
    """dtsub0 replaces the initial dtsub = dtime; clamped as the oracle clamps it."""
    d = inputs.richards_config0(nc=2048, seed=24)
    rng = np.random.default_rng(0)
    d["dtsub0"] = rng.choice([np.nan, 5.0, 60.0, 225.0, 450.0, 1000.0, 5000.0], d["nc"])
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    ref = clone(d)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    api.set_tuning(4, 256, 4, 256)
    compare(run_gpu(d, cfg), ref, d)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [api.BC_ZERO_FLUX, api.BC_WATERTABLE])
def test_water_balance_1m_columns(gpu, bc):
    """The production day's size (2^20 columns, bedrock at levels 5-20), persistent
    launch at the production tuning: every column's water change equals infiltration
    minus the root sink minus the bottom flux over the step, to 1e-13 of its water
    (a size-independent property of the moisture-form solve, which the oracle meets
    to 1.3e-16), and a strided sample of 1,052 columns matches the oracle."""
    d = inputs.richards_config0(nc=1 << 20, seed=60 + bc, nbedrock="random", lower_bc=bc)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=bc)
    api.set_tuning(4, 256, 4, 256)
    api.set_richards_schedule(True)
    got = run_gpu(clone(d), cfg)
    ns, n = d["nlevsno"], d["nlevsoi"]
    dliq = (got["h2osoi_liq"][ns:ns + n] - d["h2osoi_liq"][ns:ns + n]).sum(0)
    expect = (d["qflx_infl"] - d["qflx_rootsoi"][:n].sum(0) - got["qcharge"]) * d["dtime"]
    scale = np.abs(d["h2osoi_liq"][ns:ns + n]).sum(0)
    assert np.all(np.isfinite(dliq))
    assert np.max(np.abs(dliq - expect) / scale) < 1e-13
    if bc == api.BC_ZERO_FLUX:
        assert np.all(got["qcharge"] == 0.0)
    idx = np.arange(0, 1 << 20, 997)
    sub = {k: (v[..., idx].copy() if isinstance(v, np.ndarray) and v.shape[-1] == d["nc"] else v)
           for k, v in d.items()}
    sub["nc"] = idx.size
    ref = clone(sub)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    compare({k: got[k][..., idx] for k in ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps")},
            ref, d)


@pytest.mark.gpu
def test_nan_and_signed_zero_inputs(gpu):
    """NaN inputs propagate as in the reference (its kernels never raise) and
    -0.0 is an ordinary value: columns with a NaN layer, a NaN infiltration, a
    NaN Clapp-Hornberger exponent and -0.0 liquid water come out of the HIP step
    exactly as out of the oracle (NaN where the oracle has NaN)."""
    d = inputs.richards_config0(nc=1024, seed=25)
    ns = d["nlevsno"]
    d["h2osoi_liq"][ns + 2, 5] = np.nan
    d["qflx_infl"][7] = np.nan
    d["bsw"][4, 9] = np.nan
    d["h2osoi_liq"][ns + 1, 11] = -0.0
    d["h2osoi_liq"][ns:, 13] = -0.0
    d["nrejected"] = np.full(d["nc"], np.nan)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_FLUX)
    ref = clone(d)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    assert np.isnan(ref["h2osoi_liq"][ns:, 5]).any() and np.isnan(ref["qcharge"][7])
    for G in (1, 4):
        api.set_tuning(G, 256, 4, 256)
        got = run_gpu(clone(d), cfg)
        compare(got, ref, d)
        assert np.array_equal(got["nrejected"], ref["nrejected"], equal_nan=True)


def column_block_args(a_full, h, ncb):
    """clmsoil_mf_args of columns h*ncb .. (h+1)*ncb-1 of the arrays a_full
    points to: pointers moved by the block's first column, leading dimension
    kept (the arrays are column-major [c, j])."""
    import ctypes as C
    from clmsoil import _abi
    b = _abi.MfArgs()
    C.memmove(C.byref(b), C.byref(a_full), C.sizeof(_abi.MfArgs))
    b.nc = ncb
    for f in _abi.MF_FIELDS:
        p = getattr(a_full, f)
        if p:
            setattr(b, f, p + h * ncb * (1 if f == "mask" else 8))
    return b


@pytest.mark.gpu
def test_two_streams_disjoint_column_blocks(gpu):
    """Re-entrancy: two streams of one GPU run the resident-grid Richards step at
    the same time on disjoint column blocks of the same arrays, each call with its
    own workspace; both blocks match the oracle bit for bit and neither call
    reports an error."""
    import ctypes as C
    import torch
    from clmsoil import _abi
    nc = 1 << 17
    d = inputs.richards_config0(nc=nc, seed=26, nbedrock="random", lower_bc=api.BC_WATERTABLE)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    ref = clone(d)
    oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    t = api.to_device(d)
    api.set_tuning(4, 256, 4, 256)
    api.set_richards_schedule(True)
    L = clmsoil_lib()
    full = bind.mf_args(t)
    lev = bind.levels(d)
    c = cfg.to_abi()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ws = [api.richards_workspace(), api.richards_workspace()]
    torch.cuda.synchronize()
    half = nc // 2
    for h in range(2):
        b = column_block_args(full, h, half)
        b.workspace = ws[h].data_ptr()
        b.workspace_bytes = ws[h].numel()
        st = L.clmsoil_soilwm_moisture_form_solve(C.byref(b), C.byref(lev), d["dtime"], C.byref(c),
                                                  C.c_void_p(streams[h].cuda_stream))
        assert st == _abi.OK
    torch.cuda.synchronize()
    for h in range(2):
        with torch.cuda.stream(streams[h]):
            assert api.workspace_errors(ws[h]) == 0
    assert api.device_errors() == 0
    got = api.to_host(t)
    compare(got, ref, d)


def clmsoil_lib():
    import clmsoil
    return clmsoil.load_library()


@pytest.mark.gpu
@pytest.mark.parametrize("fields", [0, 5])
def test_lds_cache_variants_give_the_same_bits(gpu, fields):
    """The Richards kernels with 0 or 5 per-layer fields in the LDS column cache
    against the default 7 (checked against the oracle by the tests above): same
    bits on ragged production-day columns, resident grid and one-shot grid."""
    import torch
    from clmsoil import tuning
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    keys = ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps")
    try:
        for persistent in (True, False):
            api.set_richards_schedule(persistent)
            out = {}
            for f in (7, fields):
                api.set_richards_lds_cache(f)
                dev = tuning.production_state(20000, seed=13)
                tuning.richards_step(dev, cfg)
                torch.cuda.synchronize()
                out[f] = {k: dev[k].cpu().numpy() for k in keys}
            for k in keys:
                assert np.array_equal(out[7][k], out[fields][k], equal_nan=True), (persistent, k)
    finally:
        api.set_richards_lds_cache(7)
        api.set_richards_schedule(True)
    assert api.device_errors() == 0


@pytest.mark.gpu
def test_residency_gives_the_same_bits(gpu):
    """The resident Richards grid at 1 and 2 blocks per CU and at a request above
    what fits (clamped to the resident count) against the default: same bits on
    more ragged production-day columns than one resident grid holds at a time."""
    import torch
    from clmsoil import tuning
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    keys = ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps")
    out = {}
    try:
        for res in (0, 1, 2, 8):
            api.set_richards_residency(res)
            dev = tuning.production_state(140000, seed=17)
            tuning.richards_step(dev, cfg)
            torch.cuda.synchronize()
            out[res] = {k: dev[k].cpu().numpy() for k in keys}
    finally:
        api.set_richards_residency(0)
    for res in (1, 2, 8):
        for k in keys:
            assert np.array_equal(out[0][k], out[res][k], equal_nan=True), (res, k)
    assert api.device_errors() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("aquifer", [False, True])
def test_small_layout_vs_oracle(gpu, aquifer):
    """The reference tests' smaller layout (5 snow, 8 soil, 10 ground levels) runs the
    kernels whose level counts are read at run time (the CLM5 grid has its own,
    with the counts compiled in): same bits as the oracle, water table and aquifer
    pass included."""
    d = inputs.richards_config0(nc=2048, seed=33, nlevsno=5, nlevsoi=8, nlevgrnd=10,
                                nbedrock="random", lower_bc=api.BC_WATERTABLE)
    d["nrejected"] = np.full(d["nc"], np.nan)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    ref = clone(d)
    if aquifer:
        oracle.soilwater_moisture_form(ref, cfg.to_abi())
    else:
        oracle.soilwm_moisture_form_solve(ref, cfg.to_abi())
    got = run_gpu(d, cfg, aquifer=aquifer)
    compare(got, ref, d)
    assert np.array_equal(got["nrejected"], ref["nrejected"], equal_nan=True)
