"""HIP soil energy balance error vs the oracle, and the energy closure of the
GPU chain ground heat flux -> heat step -> energy balance.

Parity bar: bit-exact (subtractions, multiplications and IEEE divisions in the
reference's order; each column adds its patches in ascending patch order).
Closure bar: |errsoi| < 1e-8 W/m2 on columns whose fluxes are O(1e2) W/m2 and
whose phase-change terms reach O(1e5) W/m2 (see test_seb_oracle.py)."""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, inputs

HEAT_FLUX = ("hs_top_snow", "hs_soil", "hs_top", "hs_h2osfc", "dhsdT", "sabg_lyr_col")


def seb_gpu(d):
    import torch
    t = api.to_device(d)
    api.soil_energy_balance(t | {"nc": d["nc"], "np": d["np"]}, 1800.0, d["nlevsno"],
                            d["nlevgrnd"], d["nlevurb"])
    torch.cuda.synchronize()
    err = api.device_errors()
    return api.to_host(t), err


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [61, 62])
def test_vs_oracle(gpu, seed):
    d = inputs.energy_balance_inputs(3001, seed)
    got, err = seb_gpu(d)
    assert err == 0
    ref = clone(d)
    oracle.soil_energy_balance(ref, 1800.0)
    assert np.array_equal(got["errsoi_patch"], ref["errsoi_patch"])
    assert np.array_equal(got["errsoi_col"], ref["errsoi_col"])


@pytest.mark.gpu
def test_gpu_chain_closes_energy(gpu):
    import torch
    h, g = inputs.coupled_surface_columns(4000, 63)
    tg = api.to_device(g)
    api.compute_ground_heat_flux_and_deriv(tg | {"nc": g["nc"], "np": g["np"], "nl": g["nl"]},
                                           g["nlevsno"], g["nlevgrnd"], 0)
    th = api.to_device(h)
    for k in HEAT_FLUX:
        th[k].copy_(tg[k])
    h0 = api.to_host({k: v.clone() if hasattr(v, "clone") else v for k, v in th.items()})
    st = {k: v for k, v in th.items() if not isinstance(v, (int, float))}
    st["nc"] = h["nc"]
    api.soil_temperature(st, h["dtime"], h["nlevsno"], h["nlevgrnd"], h["nlevsoi"])
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    h1 = api.to_host(th)
    h0.update({k: h[k] for k in ("nc", "nlevsno", "nlevgrnd", "nlevsoi", "dtime")})
    seb = inputs.energy_balance_from_heat(h0, h1)
    got, err = seb_gpu(seb | {"nlevurb": 0})
    assert err == 0
    assert np.sum(h1["xmf"] != 0.0) > 1000
    assert np.max(np.abs(got["errsoi_col"])) < 1e-8
    # and the GPU's ground heat flux is the oracle's, bit for bit
    ref = clone(g)
    oracle.ground_heat_flux(ref)
    for k in HEAT_FLUX:
        assert np.array_equal(h0[k], ref[k]), k
