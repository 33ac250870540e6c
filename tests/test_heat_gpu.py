"""HIP snow+soil heat step vs the oracle (same seeded inputs).

Bar: <= 1e-11 relative, |a-b| / max(|a|, |b|, scale), on t_soisno, h2osoi_liq/ice,
the qflx_* outputs and every other output; imelt identical. `scale` is 0 except
for quantities the phase change forms as a difference of near-equal terms: a
layer's ice and liquid (ice = wice0 - xm when a layer almost melts out) are
compared relative to the layer's water mass wmass0, the melt/freeze fluxes
relative to wmass0/dtime (their column sums to the largest layer's), and the
latent-heat sums relative to hfus times that. The surface water's freezing
energy and ice flux are formed from tinc = TFRZ - t_h2osfc, which cancels to a
few mK near freezing: they are compared relative to the same expression with
TFRZ in place of tinc, frac_h2osfc * (|dhsdT| + c_h2osfc/dtime) * TFRZ (and that
over hfus for the ice flux)."""
import numpy as np
import pytest

import oracle
from _util import clone, rel_err
from clmsoil import api, bind, inputs

TOL = 1e-11
OUT = ["t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
       "h2osno_no_layers", "snow_depth", "int_snow", "t_grnd", "xmf", "xmf_h2osfc", "qflx_snomelt",
       "qflx_snofrz", "qflx_snow_drain", "qflx_h2osfc_to_ice", "eflx_fgr12", "qflx_snomelt_lyr",
       "qflx_snofrz_lyr", "fact", "c_h2osfc", "thk", "bw", "eflx_fgr", "exice_subs",
       "eflx_snomelt", "eflx_snomelt_r", "eflx_h2osfc_to_snow", "snomelt_accum"]
# sums over layers with both signs (melt and freeze) can cancel: compared
# relative to the magnitude of their largest term (W/m2 scale)
SUMS = {"xmf": 1.0e3, "eflx_fgr12": 1.0, "eflx_fgr": 1.0}


def run_gpu(d, method="Jordan1991"):
    import torch
    t = api.to_device(d)
    st = {k: v for k, v in t.items() if not isinstance(v, (int, float))}
    st["nc"] = d["nc"]
    api.soil_temperature(st, d["dtime"], d["nlevsno"], d["nlevgrnd"], d["nlevsoi"], method)
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    return api.to_host(t)


def scales(h0, ref=None):
    """Per-element magnitude floors from the input state (see module doc)."""
    ns, dt = h0["nlevsno"], h0["dtime"]
    extra = {}
    if ref is not None and "c_h2osfc" in ref:
        hm = h0["frac_h2osfc"] * (np.abs(h0["dhsdT"]) + np.nan_to_num(ref["c_h2osfc"]) / dt) * 273.15
        extra = {"xmf_h2osfc": hm, "qflx_h2osfc_to_ice": hm / 3.337e5}
    m0 = np.nan_to_num(h0["h2osoi_ice"]) + np.nan_to_num(h0["h2osoi_liq"])
    flux = m0[:ns] / dt
    colflux = flux.max(axis=0)
    return {"h2osoi_ice": m0, "h2osoi_liq": m0, "qflx_snomelt_lyr": flux, "qflx_snofrz_lyr": flux,
            "qflx_snomelt": colflux, "qflx_snofrz": colflux, "qflx_snow_drain": colflux,
            "eflx_snomelt": 3.337e5 * colflux, "eflx_snomelt_r": 3.337e5 * colflux,
            "xmf": np.maximum(3.337e5 * (m0 / dt).max(axis=0), 1.0e3)} | extra


def compare(got, ref, keys=OUT, h0=None):
    bad = {}
    sc = scales(h0, ref) if h0 is not None else {}
    for k in keys:
        if k not in ref:
            continue
        if k == "imelt":
            continue
        e = rel_err(got[k], ref[k], floor=sc.get(k, SUMS.get(k, 0.0)))
        if e > TOL:
            bad[k] = e
    assert not bad, bad
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
