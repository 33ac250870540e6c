"""HIP ground heat flux vs the oracle (same seeded inputs).

Bar: bit-exact. The arithmetic is additions and multiplications in the
reference's order plus Julia's compensated x^4 (fma-based, exact on both
sides), and each column adds its patches in ascending patch order, the order of
the reference's sequential CPU path."""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, inputs

PATCH_OUT = ("eflx_gnet", "dgnetdT", "sabg_chk", "eflx_wasteheat", "eflx_ventilation",
             "eflx_heat_from_ac", "eflx_traffic", "eflx_anthro")
COL_OUT = ("hs_top", "hs_top_snow", "hs_soil", "hs_h2osfc", "dhsdT", "sabg_lyr_col")


def run_gpu(d):
    import torch
    t = api.to_device(d)
    api.compute_ground_heat_flux_and_deriv(t | {"nc": d["nc"], "np": d["np"], "nl": d["nl"]},
                                           d["nlevsno"], d["nlevgrnd"],
                                           d.get("building_temp_method", 0))
    torch.cuda.synchronize()
    err = api.device_errors()
    return api.to_host(t), err


def run_oracle(d):
    ref = clone(d)
    oracle.ground_heat_flux(ref)
    return ref


def compare(got, ref):
    for k in PATCH_OUT + COL_OUT:
        assert np.array_equal(got[k], ref[k], equal_nan=True), k


@pytest.mark.gpu
@pytest.mark.parametrize("seed,method", [(51, 0), (52, 1), (53, 2)])
def test_vs_oracle(gpu, seed, method):
    d = inputs.ground_heat_flux_inputs(nc=5003, seed=seed, building_temp_method=method)
    got, err = run_gpu(d)
    assert err == 0
    compare(got, run_oracle(d))


@pytest.mark.gpu
def test_zero_weight_nan_patches(gpu):
    d = inputs.ground_heat_flux_inputs(nc=2000, seed=54, zero_weight_fraction=0.3)
    zw = (d["wtcol"] == 0.0) & (d["mask_nolakep"] == 1)
    for k in ("eflx_sh_grnd", "eflx_sh_soil", "eflx_sh_h2osfc", "cgrnd"):
        d[k][zw] = np.nan
    got, err = run_gpu(d)
    assert err == 0
    compare(got, run_oracle(d))


@pytest.mark.gpu
def test_few_snow_levels_and_edge_sizes(gpu):
    d = inputs.ground_heat_flux_inputs(nc=777, seed=55, nlevsno=5, nlevgrnd=15)
    got, err = run_gpu(d)
    assert err == 0
    compare(got, run_oracle(d))
    for nc in (1, 64, 65):
        d = inputs.ground_heat_flux_inputs(nc=nc, seed=56 + nc)
        got, err = run_gpu(d)
        assert err == 0
        compare(got, run_oracle(d))


@pytest.mark.gpu
def test_inconsistent_subgrid_is_flagged(gpu):
    """A patch whose column index disagrees with its column's patch range is not
    added and is reported in the device error word (CLMSOIL_DEVERR_SUBGRID)."""
    d = inputs.ground_heat_flux_inputs(nc=300, seed=57)
    multi = np.flatnonzero(d["patchf"] > d["patchi"])
    c = int(multi[0])
    p = int(d["patchi"][c])           # 1-based first patch of column c
    d["patch_column"][p - 1] = c + 2 if c + 2 <= d["nc"] else c
    got, err = run_gpu(d)
    assert err & 16
