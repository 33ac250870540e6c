"""HIP vertical transport of the soil organic matter pools vs the oracle.

Bar: bit-exact without spinup (the same sums, products, quotients, Julia max/min
and x^5 through fma, no transcendental function). With accelerated spinup the
latitude term calls exp(), whose last bit may differ between the GPU and the host
library: the concentrations are held to 1e-12 relative plus the oracle's own
sensitivity to a one-ulp change of that term's input."""
import numpy as np
import pytest

import oracle
from _util import clone, rel_err
from clmsoil import api, inputs
from clmsoil._abi import LVT_FIELDS


def params(d, **kw):
    p = dict(nlevdecomp=d["nlevdecomp"], nlevdecomp_full=d["nlevdecomp_full"], npools=d["npools"],
             ntypes=d["ntypes"], is_cwd=d["is_cwd"])
    p.update(kw)
    return p


def run_gpu(d, dtime, **kw):
    import torch
    t = api.to_device({k: d[k] for k in LVT_FIELDS})
    conc = [torch.from_numpy(x).cuda() for x in d["conc"]]
    src = [torch.from_numpy(x).cuda() for x in d["source"]]
    tend = [torch.from_numpy(x).cuda() for x in d["tendency"]]
    api.litter_vert_transp(t | {"nc": d["nc"], "ld": d.get("ld", d["nc"]), "conc": conc,
                                "source": src, "tendency": tend},
                           dtime, **params(d, **kw))
    torch.cuda.synchronize()
    out = api.to_host(t)
    out["conc"] = [x.cpu().numpy() for x in conc]
    out["tendency"] = [x.cpu().numpy() for x in tend]
    return out


def run_oracle(d, dtime, **kw):
    ref = clone(d)
    for k in ("conc", "source", "tendency"):
        ref[k] = [x.copy() for x in d[k]]
    oracle.litter_vert_transp(ref, dtime, **params(d, **kw))
    return ref


CASES = [
    (116, 1800.0, {}),                                                    # CLM5 rates
    (117, 3.0e7, {"som_diffus": 1e-8, "cryoturb_diffusion_k": 5e-8}),     # strong mixing
    (118, 3.0e7, {"som_diffus": 1e-8, "som_adv_flux": 3e-10}),            # advection too
]


@pytest.mark.gpu
@pytest.mark.parametrize("seed,dtime,kw", CASES)
def test_lvt_vs_oracle(gpu, seed, dtime, kw):
    d = inputs.lvt_columns(nc=2003, seed=seed)
    api.device_errors()
    got = run_gpu(d, dtime, **kw)
    assert api.device_errors() == 0
    ref = run_oracle(d, dtime, **kw)
    assert np.array_equal(got["som_diffus_coef"], ref["som_diffus_coef"])
    assert np.array_equal(got["som_adv_coef"], ref["som_adv_coef"])
    for t in range(d["ntypes"]):
        assert np.array_equal(got["conc"][t], ref["conc"][t]), t
        assert np.array_equal(got["tendency"][t], ref["tendency"][t]), t


@pytest.mark.gpu
def test_lvt_spinup_and_four_tracers(gpu):
    """The spinup term's exp() may differ in its last bit between the GPU and the
    host library; the bar is 1e-12 relative plus 4x how far the oracle's results
    move when every latitude is moved one ulp up or down (which moves exp()'s
    result by several ulps), and 99.9 % of the elements must meet plain 1e-12."""
    d = inputs.lvt_columns(nc=1501, seed=119, ntypes=4)
    sf = [1.0, 1.0, 1.0, 1.0, 5.0, 10.0, 30.0]
    kw = dict(spinup_state=1, spinup_factor=sf, som_diffus=1e-8, cryoturb_diffusion_k=5e-8)
    got = run_gpu(d, 3.0e7, **kw)
    ref = run_oracle(d, 3.0e7, **kw)
    band = [np.zeros_like(x) for x in ref["conc"]]
    for direction in (np.inf, -np.inf):
        e = clone(d)
        e["latdeg"] = np.nextafter(d["latdeg"], direction)
        alt = run_oracle(e, 3.0e7, **kw)
        band = [np.maximum(b, np.abs(x - y)) for b, x, y in zip(band, alt["conc"], ref["conc"])]
    for t in range(4):
        a, b = got["conc"][t], ref["conc"][t]
        diff = np.abs(a - b)
        scale = np.maximum(np.abs(a), np.abs(b))
        assert np.all(diff <= 1e-12 * scale + 4.0 * band[t]), t
        assert np.mean(diff <= 1e-12 * scale) >= 0.999, t
        assert rel_err(got["tendency"][t], ref["tendency"][t], floor=1e-12) <= 1e-6, t


@pytest.mark.gpu
def test_lvt_bad_nbedrock_flags(gpu):
    d = inputs.lvt_columns(nc=400, seed=120)
    d["mask_bgc_soilc"][:] = 1
    d["nbedrock"][11] = 0
    api.device_errors()
    got = run_gpu(d, 1800.0)
    assert api.device_errors() & 1                       # CLMSOIL_DEVERR_NBEDROCK
    assert np.array_equal(got["conc"][0][:, :, 11], d["conc"][0][:, :, 11])
    ref = clone(d)
    ref["mask_bgc_soilc"] = d["mask_bgc_soilc"].copy()
    ref["mask_bgc_soilc"][11] = 0
    ref = run_oracle(ref, 1800.0)
    assert np.array_equal(got["conc"][0], ref["conc"][0])


@pytest.mark.gpu
@pytest.mark.parametrize("nd", [10, 7])
def test_lvt_other_level_counts(gpu, nd):
    """10 levels take the shared-matrix kernel (one factorisation per column for all
    pools and tracers), 7 the per-system kernel; both bit-exact, with and without a
    mix of spinup factors (the pools with factor 1 stay on the shared matrix)."""
    d = inputs.lvt_columns(nc=1003, seed=121 + nd, nlevdecomp=nd)
    for kw in ({}, dict(spinup_state=1, spinup_factor=[1.0, 1.0, 1.0, 1.0, 5.0, 1.0, 30.0])):
        kw = dict(kw, som_diffus=1e-8, cryoturb_diffusion_k=5e-8, som_adv_flux=3e-10)
        got = run_gpu(d, 3.0e7, **kw)
        ref = run_oracle(d, 3.0e7, **kw)
        spin = "spinup_state" in kw
        for t in range(d["ntypes"]):
            for s in range(d["npools"]):
                a, b = got["conc"][t][s], ref["conc"][t][s]
                if spin and s in (4, 6):                 # exp() of the latitude term
                    assert rel_err(a, b, floor=1e-12) <= 1e-9, (t, s)
                else:
                    assert np.array_equal(a, b), (t, s)
                    assert np.array_equal(got["tendency"][t][s], ref["tendency"][t][s]), (t, s)


@pytest.mark.gpu
def test_lvt_padded_leading_dimension(gpu):
    """nc below the arrays' leading dimension: the columns past nc stay as they were
    and the others match the oracle bit for bit."""
    d = inputs.lvt_columns(nc=700, seed=124)
    d["ld"], d["nc"] = 700, 611
    got = run_gpu(d, 1800.0)
    ref = run_oracle(d, 1800.0)
    for t in range(d["ntypes"]):
        assert np.array_equal(got["conc"][t][:, :, 611:], d["conc"][t][:, :, 611:]), t
        assert np.array_equal(got["conc"][t], ref["conc"][t]), t
        assert np.array_equal(got["tendency"][t], ref["tendency"][t]), t


@pytest.mark.gpu
def test_lvt_sixteen_pools_three_tracers(gpu):
    """The most pools the interface takes (16, two of them coarse woody debris) and
    three tracer types: 48 systems per column, so each of the shared-matrix kernel's
    8 wavefronts solves six; bit-exact."""
    d = inputs.lvt_columns(nc=777, seed=125, npools=16, ntypes=3, cwd_pools=(3, 11))
    kw = dict(som_diffus=1e-8, cryoturb_diffusion_k=5e-8, som_adv_flux=3e-10)
    got = run_gpu(d, 3.0e7, **kw)
    ref = run_oracle(d, 3.0e7, **kw)
    for t in range(3):
        assert np.array_equal(got["conc"][t], ref["conc"][t]), t
        assert np.array_equal(got["tendency"][t], ref["tendency"][t]), t


@pytest.mark.gpu
def test_lvt_without_mask(gpu):
    """No filter (a NULL mask): every column is transported; bit-exact."""
    d = inputs.lvt_columns(nc=900, seed=126)
    d["mask_bgc_soilc"][:] = 1
    ref = run_oracle(d, 1800.0)
    e = dict(d)
    e["mask_bgc_soilc"] = None
    got = run_gpu(e, 1800.0)
    for t in range(d["ntypes"]):
        assert np.array_equal(got["conc"][t], ref["conc"][t]), t
        assert np.array_equal(got["tendency"][t], ref["tendency"][t]), t
