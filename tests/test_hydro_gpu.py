"""HIP soil water fractions and water table vs the oracle (same seeded inputs).

Bars: fractions bit-exact (divisions and min/max only). Water table: the
specific yield takes one pow() per layer visited (the only operation whose last
bit may differ between the GPU and the host library), so zwt and wa are compared
at 1e-11 relative, the frost table, the drainage fluxes and the water contents
written on the way bit-exact, and zwt_perched at 1e-11 relative. The water
tables diagnosed from the moisture profile (perched, theta-based, bedrock clamp)
are bit-exact."""
import numpy as np
import pytest

import oracle
from _util import clone, rel_err
from clmsoil import api, inputs

TOL = 1e-11
WT = ("dz", "z", "zi", "t_soisno", "watsat", "sucsat", "bsw", "qcharge", "h2osoi_liq",
      "h2osoi_ice", "zwt", "wa", "zwt_perched", "frost_table", "h2osoi_vol", "qflx_drain",
      "qflx_rsub_sat", "qflx_drain_perched", "mask")


def run_gpu(d, fn, keys, **kw):
    import torch
    t = api.to_device(d)
    st = {k: t[k] for k in keys if k in t} | {"nc": d["nc"]}
    fn(st, **kw)
    torch.cuda.synchronize()
    return api.to_host(t)


@pytest.mark.gpu
def test_fractions_vs_oracle(gpu):
    d = inputs.water_table_columns(nc=4097, seed=41)
    got = run_gpu(d, api.set_soil_water_fractions,
              ("mask", "watsat", "dz", "h2osoi_ice", "excess_ice", "eff_porosity", "icefrac"))
    ref = clone(d)
    oracle.set_soil_water_fractions(ref)
    assert np.array_equal(got["eff_porosity"], ref["eff_porosity"])
    assert np.array_equal(got["icefrac"], ref["icefrac"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,zwt_range,frost", [(42, (0.3, 12.0), True), (43, (0.5, 6.0), True),
                                                  (44, (9.0, 30.0), False)])
def test_water_table_vs_oracle(gpu, seed, zwt_range, frost):
    d = inputs.water_table_columns(nc=4099, seed=seed, zwt_range=zwt_range)
    d["mask"] = (np.arange(d["nc"]) % 5 != 2).astype(np.uint8)
    got = run_gpu(d, lambda st: api.water_table(st, 1800.0, recompute_frost_table=frost), WT)
    ref = clone(d)
    oracle.water_table(ref, 1800.0, recompute_frost_table=frost)
    assert rel_err(got["zwt"], ref["zwt"]) <= TOL
    assert rel_err(got["wa"], ref["wa"]) <= TOL
    assert rel_err(got["zwt_perched"], ref["zwt_perched"]) <= TOL
    for k in ("frost_table", "h2osoi_vol", "qflx_drain", "qflx_rsub_sat", "qflx_drain_perched"):
        assert np.array_equal(got[k], ref[k]), k
    off = d["mask"] == 0
    assert np.array_equal(got["zwt"][off], d["zwt"][off])


@pytest.mark.gpu
def test_zd09_then_water_table_conserves_water(gpu):
    """The Zeng-Decker step and the water table update in sequence on the GPU:
    where the table is below the soil column, soil water plus aquifer water
    changes by exactly the infiltration minus the root uptake."""
    import torch
    d = inputs.water_table_columns(nc=1 << 16, seed=45, zwt_range=(9.0, 20.0))
    t = api.to_device(d)
    zd_keys = ("z", "dz", "zi", "watsat", "hksat", "bsw", "sucsat", "smpmin", "h2osoi_liq",
               "h2osoi_ice", "h2osoi_vol", "t_soisno", "zwt", "qflx_infl", "qflx_rootsoi",
               "icefrac", "smp_l", "hk_l", "qcharge", "qflx_deficit")
    api.soilwater_zengdecker2009({k: t[k] for k in zd_keys}, 1800.0, api.init_soilwater_movement())
    api.water_table({k: t[k] for k in WT if k in t} | {"nc": d["nc"]}, 1800.0)
    torch.cuda.synchronize()
    h = api.to_host(t)
    ns, n = d["nlevsno"], d["nlevsoi"]
    soil = (h["h2osoi_liq"][ns:ns + n] - d["h2osoi_liq"][ns:ns + n]).sum(0)
    dwa = h["wa"] - d["wa"]
    expect = (d["qflx_infl"] - d["qflx_rootsoi"].sum(0)) * 1800.0
    scale = np.abs(d["h2osoi_liq"][ns:ns + n]).sum(0) + d["wa"]
    assert np.max(np.abs(soil + dwa - expect) / scale) < 1e-13


TW = ("mask", "nbedrock", "dz", "z", "zi", "t_soisno", "watsat", "h2osoi_liq", "h2osoi_ice",
      "zwt", "zwt_perched", "frost_table", "h2osoi_vol")


@pytest.mark.gpu
@pytest.mark.parametrize("entry", ["perched_water_table", "theta_based_water_table",
                                   "clamp_zwt_to_bedrock", "water_table_moisture_form"])
def test_moisture_water_tables_vs_oracle(gpu, entry):
    """Bit-exact: divisions, products, sums and max only, no contraction."""
    d = inputs.moisture_water_table_columns(nc=4101, seed=50)
    d["mask"] = (np.arange(d["nc"]) % 7 != 3).astype(np.uint8)
    got = run_gpu(d, getattr(api, entry), TW)
    ref = clone(d)
    getattr(oracle, entry)(ref)
    for k in ("zwt", "zwt_perched", "frost_table", "h2osoi_vol"):
        assert np.array_equal(got[k], ref[k]), (entry, k)
    off = d["mask"] == 0
    assert np.array_equal(got["zwt"][off], d["zwt"][off])


@pytest.mark.gpu
def test_moisture_water_table_bad_nbedrock_flags(gpu):
    d = inputs.moisture_water_table_columns(nc=300, seed=51)
    d["nbedrock"][17] = 0
    api.device_errors()                                   # clear
    got = run_gpu(d, api.water_table_moisture_form, TW)
    assert api.device_errors() & 1                        # CLMSOIL_DEVERR_NBEDROCK
    ref = clone(d)
    ref["nbedrock"][17] = d["nlevsoi"]
    ref["mask"] = (np.arange(d["nc"]) != 17).astype(np.uint8)
    oracle.water_table_moisture_form(ref)
    ok = np.arange(d["nc"]) != 17
    assert np.array_equal(got["zwt"][ok], ref["zwt"][ok])
    assert got["zwt"][17] == d["zwt"][17]                 # left unchanged
