"""The heat-step adjoint (clmsoil_soil_temperature_adjoint) against central
finite differences, on snow-free columns around the freezing point where layers
freeze and thaw in the step (phase_change_beta!, soil_temperature.jl:1929-2289).

For a random weighting L = sum_c sum_j (wt t_out + wl liq_out + wi ice_out) and
a random direction v of each input (t_soisno, h2osoi_liq, h2osoi_ice, watsat,
bsw), the adjoint's directional derivative per column must match
(L(x + h v) - L(x - h v)) / 2h with the reference's norm-wise metric
max|adj - fd| / max|fd| <= 1e-6 (test_enzyme_smoke.jl:238). A column whose
path is not smooth inside [x - h, x + h] (a melt/freeze label, a max/min
selection or a row interchange changes) shows different one-sided differences
and is left out; the test requires that most columns remain and that they
include columns with phase change.
"""
import numpy as np
import pytest

from clmsoil import adjoint, api, inputs

pytestmark = pytest.mark.gpu

NC = 2048


def _state(torch, freezing=True, seed=21):
    host = inputs.calibration_columns(NC, seed, freezing=freezing)
    host.update(inputs.heat_outputs(host, diagnostics=False, fill=0.0))
    host["imelt"] = np.zeros((host["nlevsno"] + host["nlevgrnd"], NC), dtype=np.int64)
    return api.to_device(host)


def _run(torch, d, over=None):
    """One heat step on a copy of d (with fields replaced by `over`)."""
    s = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in d.items()}
    if over:
        s.update({k: v.clone() for k, v in over.items()})
    api.soil_temperature(s, float(d["dtime"]), d["nlevsno"], d["nlevgrnd"])
    return s


def _loss(s, w, b):
    """The weighting of the outputs' differences from the unperturbed step (the
    differences are exact, so the finite differences carry no cancellation)."""
    return (w["t"] * (s["t_soisno"] - b["t_soisno"]) + w["l"] * (s["h2osoi_liq"] - b["h2osoi_liq"]) +
            w["i"] * (s["h2osoi_ice"] - b["h2osoi_ice"])).sum(dim=0)


@pytest.mark.parametrize("name", ["t_soisno", "h2osoi_liq", "h2osoi_ice", "watsat", "bsw"])
def test_heat_adjoint_directional_fd(gpu, name):
    import torch
    d = _state(torch)
    ns = d["nlevsno"]
    g = torch.Generator(device="cuda").manual_seed(5)
    rnd = lambda shape: torch.rand(shape, generator=g, dtype=torch.float64, device="cuda") * 2 - 1
    shape = tuple(d["t_soisno"].shape)
    w = {"t": rnd(shape), "l": rnd(shape), "i": rnd(shape)}
    for x in w.values():
        x[:ns] = 0.0
    base = _run(torch, d)
    assert api.device_errors() == 0
    # phase change happens in this step
    changed = (base["h2osoi_ice"] != d["h2osoi_ice"]).any(dim=0)
    assert changed.float().mean().item() > 0.2

    t_bar, l_bar, i_bar = w["t"].clone(), w["l"].clone(), w["i"].clone()
    gw = torch.zeros_like(d["watsat"])
    gb = torch.zeros_like(d["bsw"])
    adjoint.soil_temperature_adjoint(d, float(d["dtime"]), t_bar, l_bar, i_bar,
                                     d["nlevsno"], d["nlevgrnd"], grad_bsw=gb, grad_watsat=gw)
    assert api.device_errors() == 0

    x0 = d[name]
    v = rnd(tuple(x0.shape)) * x0
    if x0.shape[0] == shape[0]:
        v[:ns] = 0.0
    adj_of = {"t_soisno": t_bar, "h2osoi_liq": l_bar, "h2osoi_ice": i_bar, "watsat": gw, "bsw": gb}
    adj = (adj_of[name] * v).sum(dim=0)
    h = 1e-6
    Lp = _loss(_run(torch, d, {name: x0 + h * v}), w, base)
    Lm = _loss(_run(torch, d, {name: x0 - h * v}), w, base)
    L0 = _loss(base, w, base)
    fd = (Lp - Lm) / (2 * h)
    fwd, bwd = (Lp - L0) / h, (L0 - Lm) / h
    scale = fd.abs().max().clamp_min(1e-300)
    smooth = (fwd - bwd).abs() <= 1e-6 * scale + 1e-6 * fd.abs()
    assert smooth.float().mean().item() > 0.8
    assert (smooth & changed).sum().item() > 0.2 * NC
    err = ((adj - fd).abs()[smooth].max() / fd.abs()[smooth].max()).item()
    assert err <= 1e-6, err


def test_heat_adjoint_rejects_snow_columns(gpu):
    """Columns with snow layers are outside the adjoint's scope: flagged, left
    unchanged."""
    import torch
    host = inputs.heat_config1(nc=256, seed=3)
    host.update(inputs.heat_outputs(host, diagnostics=False, fill=0.0))
    d = api.to_device(host)
    tb = torch.ones_like(d["t_soisno"])
    lb, ib = torch.zeros_like(tb), torch.zeros_like(tb)
    api.device_errors()
    adjoint.soil_temperature_adjoint(d, float(d["dtime"]), tb, lb, ib, d["nlevsno"], d["nlevgrnd"])
    flags = api.device_errors()
    assert flags & 64
    snow = (d["snl"] < 0)
    assert torch.all(tb[:, snow] == 1.0)
