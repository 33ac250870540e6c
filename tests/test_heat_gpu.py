"""HIP snow+soil heat step vs the oracle (same seeded inputs).

Bar: every output equal to the oracle's bit for bit (t_soisno, h2osoi_liq/ice,
the qflx_* outputs, the diagnostics), imelt identical. Both sides evaluate
pow/log10 with the same function (clm_math.h, pinned in test_clm_math.py),
divide where the reference divides and fuse nothing, and the banded solve
performs the dense pivoted elimination's arithmetic on every entry it touches,
so no element needs a tolerance. (This is stronger than the north star's
1e-11 relative, which it implies.)"""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise, clone
from clmsoil import api, bind, inputs

OUT = ["t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
       "h2osno_no_layers", "snow_depth", "int_snow", "t_grnd", "xmf", "xmf_h2osfc", "qflx_snomelt",
       "qflx_snofrz", "qflx_snow_drain", "qflx_h2osfc_to_ice", "eflx_fgr12", "qflx_snomelt_lyr",
       "qflx_snofrz_lyr", "fact", "c_h2osfc", "thk", "bw", "eflx_fgr", "exice_subs",
       "eflx_snomelt", "eflx_snomelt_r", "eflx_h2osfc_to_snow", "snomelt_accum"]
def run_gpu(d, method="Jordan1991"):
    import torch
    t = api.to_device(d)
    st = {k: v for k, v in t.items() if not isinstance(v, (int, float))}
    st["nc"] = d["nc"]
    api.soil_temperature(st, d["dtime"], d["nlevsno"], d["nlevgrnd"], d["nlevsoi"], method)
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    return api.to_host(t)


def compare(got, ref, keys=OUT, h0=None):  # noqa: ARG001 (h0: call sites name the inputs)
    assert_bitwise(got, ref, keys)
    assert np.array_equal(got["imelt"], ref["imelt"])


def heat_case(nc, seed, variant=False, nbedrock=None):
    h = inputs.heat_config1(nc=nc, seed=seed, h2osfc_variant=variant, nbedrock=nbedrock)
    h.update(inputs.heat_outputs(h))
    return h


@pytest.mark.gpu
@pytest.mark.parametrize("G", [1, 2, 4])
@pytest.mark.parametrize("variant", [False, True])
def test_config1_vs_oracle(gpu, G, variant):
    h = heat_case(8192, 30 + int(variant), variant, nbedrock="random" if variant else None)
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    api.set_tuning(4, 256, G, 256)
    compare(run_gpu(clone(h)), ref, h0=h)


@pytest.mark.gpu
def test_config1_full_size_vs_oracle(gpu):
    """The heat workload at its full size: 65,536 columns, 0-12 snow layers on 25
    ground layers, every column against the oracle, at the production tuning."""
    h = heat_case(65536, 37, True, nbedrock="random")
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    api.set_tuning(4, 256, 4, 256)
    compare(run_gpu(clone(h)), ref, h0=h)


@pytest.mark.gpu
def test_sturm_snow_conductivity(gpu):
    h = heat_case(4096, 33)
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config(2))
    api.set_tuning(4, 256, 4, 256)
    compare(run_gpu(clone(h), "Sturm1997"), ref, h0=h)


def pivot_columns(h, every=3):
    """Make every `every`-th column need a row interchange: a soil layer with an
    almost-zero heat capacity gives its row a sub-diagonal far larger than the
    pivot above it (band_diagonal.jl:125-140 then swaps rows)."""
    ns = h["nlevsno"]
    cols = np.arange(0, h["nc"], every)
    for j in (5, 9):
        h["csol"][j - 1, cols] = 1.0e-3
        h["h2osoi_liq"][ns + j - 1, cols] = 0.0
        h["h2osoi_ice"][ns + j - 1, cols] = 0.0
    return cols


@pytest.mark.gpu
def test_pivoting_columns(gpu):
    h = heat_case(2048, 34, variant=True)
    pivot_columns(h)
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    for G in (1, 4):
        api.set_tuning(4, 256, G, 256)
        compare(run_gpu(clone(h)), ref, h0=h)


@pytest.mark.gpu
def test_small_levels_and_mask(gpu):
    """nlevsno = 5, nlevgrnd = 10 (the reference tests' layout) with a mask."""
    h = inputs.heat_config1(nc=1024, seed=35, nlevsno=5, nlevsoi=8, nlevgrnd=10, h2osfc_variant=True)
    h.update(inputs.heat_outputs(h))
    h["mask"] = (np.arange(h["nc"]) % 4 != 1).astype(np.uint8)
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    api.set_tuning(4, 256, 4, 256)
    got = run_gpu(clone(h))
    compare(got, ref, h0=h)
    assert np.all(np.isnan(got["t_grnd"][1::4]))


@pytest.mark.gpu
def test_nan_and_signed_zero_inputs(gpu):
    """NaN in a layer temperature, in the soil porosity and in a heat flux
    propagates through the HIP step exactly as through the oracle; -0.0 ice is
    an ordinary value."""
    h = heat_case(1024, 36, True, nbedrock="random")
    ns = h["nlevsno"]
    h["t_soisno"][ns + 3, 17] = np.nan
    h["watsat"][6, 21] = np.nan
    h["hs_soil"][30] = np.nan
    h["h2osoi_ice"][ns + 1, 40] = -0.0
    ref = clone(h)
    oracle.soil_temperature(ref, bind.heat_config())
    assert np.isnan(ref["t_soisno"][:, 17]).any()
    api.set_tuning(4, 256, 4, 256)
    compare(run_gpu(clone(h)), ref, h0=h)
