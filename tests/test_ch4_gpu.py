"""HIP methane and oxygen transport vs the oracle.

The competition, the scaled process rates, the ebullition and aerenchyma totals,
the surface ebullition and the ground conductance carry no transcendental
function and are bit-exact. Everything downstream of the Henry partition (exp,
log) and the diffusivities (pow with non-integer exponents) may differ in the
last bits between the GPU and the host library: the bar per element is 1e-12
relative plus 4x the band the oracle itself moves when every exp, log and pow is
moved one ulp (8 jitter seeds), and at least 99.9 % of the elements must meet
plain 1e-12 relative. The diffusive surface flux, a difference of nearly equal
terms, is compared on the scale of the column's CH4 content per time step.
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, inputs
from clmsoil._abi import CH4_FIELDS

TOL = 1e-12
EXACT = ("o2stress", "ch4stress", "o2_decomp_depth", "o2_oxid_depth", "ch4_oxid_depth",
         "ch4_aere_depth", "ch4_ebul_depth", "ch4_ebul_total", "ch4_surf_aere", "ch4_surf_ebul",
         "grnd_ch4_cond")
BANDED = ("conc_ch4", "conc_o2", "ch4_surf_diff")


def run_gpu(d, dtime, **params):
    import torch
    t = api.to_device({k: d[k] for k in CH4_FIELDS})
    api.ch4_tran(t | {"nc": d["nc"], "ldg": d["ldg"], "ld": d.get("ld", d["nc"])}, dtime,
                 nlevsoi=d["nlevsoi"], **params)
    torch.cuda.synchronize()
    return api.to_host(t)


def oracle_band(d, dtime, draws=8, **params):
    base = clone(d)
    oracle.ch4_tran(base, dtime, nlevsoi=d["nlevsoi"], **params)
    band = {k: np.zeros_like(base[k]) for k in BANDED}
    try:
        for seed in range(1, draws + 1):
            oracle.set_ch4_jitter(seed)
            out = clone(d)
            oracle.ch4_tran(out, dtime, nlevsoi=d["nlevsoi"], **params)
            for k in BANDED:
                band[k] = np.maximum(band[k], np.abs(out[k] - base[k]))
    finally:
        oracle.set_ch4_jitter(0)
    return base, band


def compare(d, got, ref, band, dtime):
    for k in EXACT:
        assert np.array_equal(got[k], ref[k]), (k, float(np.max(np.abs(got[k] - ref[k]))))
    content = (np.abs(ref["conc_ch4"]) * d["dz"]).sum(0) / dtime
    for k in BANDED:
        a, b, bd = got[k], ref[k], band[k]
        diff = np.abs(a - b)
        scale = np.maximum(np.abs(a), np.abs(b))
        if k == "ch4_surf_diff":
            scale = np.maximum(scale, content)
        assert np.all(diff <= TOL * scale + 4.0 * bd), (k, float(np.max(diff - TOL * scale - 4 * bd)))
        assert (diff <= TOL * scale).mean() >= 0.999, (k, float((diff <= TOL * scale).mean()))


CASES = [
    (140, 0, {}),                                                   # unsaturated, CLM defaults
    (141, 1, {}),                                                   # saturated, ponding
    (142, 0, {"ch4frzout": True, "use_aereoxid_prog": False, "organic_max": 0.0}),
    (143, 1, {"lake": True, "satpow": 1.7, "scale_factor_gasdiff": 0.8}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("seed,sat,kw", CASES)
def test_ch4_tran_vs_oracle(gpu, seed, sat, kw):
    d = inputs.ch4_columns(nc=4003, seed=seed, sat=sat)
    api.device_errors()
    got = run_gpu(d, 1800.0, sat=sat, **kw)
    assert api.device_errors() == 0
    ref, band = oracle_band(d, 1800.0, sat=sat, **kw)
    compare(d, got, ref, band, 1800.0)


@pytest.mark.gpu
def test_ch4_tran_levels_and_padding(gpu):
    """A short column (5 levels) with a leading dimension above nc: the columns
    past nc stay as they were."""
    d = inputs.ch4_columns(nc=700, seed=144, nlevsoi=5)
    d["ld"], d["nc"] = 700, 640
    got = run_gpu(d, 3600.0)
    ref, band = oracle_band(d, 3600.0)
    for k in ("conc_ch4", "o2stress", "ch4_surf_diff"):
        assert np.array_equal(got[k][..., 640:], d[k][..., 640:]), k
    compare(d, got, ref, band, 3600.0)


@pytest.mark.gpu
def test_ch4_tran_bad_jwt_flags(gpu):
    d = inputs.ch4_columns(nc=300, seed=145)
    d["mask"][:] = 1
    d["jwt"][17] = -1
    api.device_errors()
    got = run_gpu(d, 1800.0)
    assert api.device_errors() & 32                       # CLMSOIL_DEVERR_JWT
    assert np.array_equal(got["conc_ch4"][:, 17], d["conc_ch4"][:, 17])
    ref = clone(d)
    ref["mask"] = d["mask"].copy()
    ref["mask"][17] = 0
    oracle.ch4_tran(ref, 1800.0, nlevsoi=d["nlevsoi"])
    assert np.array_equal(got["ch4stress"], ref["ch4stress"])


@pytest.mark.gpu
@pytest.mark.parametrize("nl,sat", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_ch4_tran_one_and_two_levels(gpu, nl, sat):
    """The shortest columns: the level below the ground node is also the bottom
    row, and jwt at 0, 1 or 2 takes every branch of the interface conductances."""
    d = inputs.ch4_columns(nc=3001, seed=146 + 2 * nl + sat, nlevsoi=nl, sat=sat)
    got = run_gpu(d, 1800.0, sat=sat)
    ref, band = oracle_band(d, 1800.0, sat=sat)
    compare(d, got, ref, band, 1800.0)


@pytest.mark.gpu
def test_ch4_tran_without_mask(gpu):
    """No filter (a NULL mask): every column is stepped."""
    d = inputs.ch4_columns(nc=1500, seed=151)
    d["mask"][:] = 1
    ref, band = oracle_band(d, 1800.0)
    e = dict(d)
    e["mask"] = None
    got = run_gpu(e, 1800.0)
    compare(d, got, ref, band, 1800.0)
