"""Differentiable calibration of per-column soil hydraulic parameters.

Workload: `nc` snow-free columns (inputs.calibration_columns) advanced
`days` x 48 steps of 1800 s. Each step runs the heat step and then the
moisture-form Richards step with a free-drainage (BC_FLUX) lower boundary. The
parameters are per-column multipliers m = (m_b, m_ksat, m_porosity) on bsw,
hksat and watsat (all layers). The loss is

    L = sum_c sum_d (D[d, c] - T[d, c])^2,  D[d, c] = sum over day d of qcharge * dtime,

the squared error of daily drainage (mm) against a target T made by the same
model from perturbed multipliers.

Gradient: reverse mode through the library's hand-written adjoints of both
steps. The forward pass keeps one checkpoint of the heat and water state per
day. The reverse pass recomputes each day's 48 steps from its checkpoint
(keeping the state at the start of every heat and Richards step) and then
applies, step by step in reverse order, the Richards adjoint
(clmsoil_soilwm_moisture_form_adjoint, which replays its substeps from a tape)
and the heat-step adjoint (clmsoil_soil_temperature_adjoint). Layers that
freeze or thaw move water between h2osoi_liq and h2osoi_ice in the heat step,
so the temperature, liquid water and ice all carry adjoints from step to step;
watsat and bsw reach the loss through both steps (conductivity, heat capacity
and the supercooled-water limit in the heat step). The ice fraction the
Richards step reads (icefrac) is an input of the workload, held at its initial
value.

The reference computes this kind of gradient with Enzyme over its driver
(src/biogeophys/driver_reverse.jl:169-190 soil temperature, :598-615 soil
water, multi-step checkpointing in canopy_fluxes_reverse.jl:488-542).
"""
from __future__ import annotations

import numpy as np

from . import adjoint, inputs
from .api import (BC_FLUX, MOISTURE_FORM, MfIn, MfState, device_errors,
                  init_soilwater_movement, soil_temperature, soilwm_moisture_form_solve,
                  to_device)

HEAT_STATE = ["t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
              "h2osno_no_layers", "snow_depth", "int_snow"]
HEAT_INPUTS = ["snl", "nbedrock", "z", "dz", "zi", "tkmg", "tkdry", "csol", "sucsat",
               "frac_sno_eff", "frac_h2osfc", "hs_top_snow", "hs_soil", "hs_top", "hs_h2osfc",
               "dhsdT", "sabg_lyr_col", "eflx_bot"]


class Calibration:
    def __init__(self, nc: int, days: int, steps_per_day: int = 48, seed: int = 3,
                 device="cuda", freezing: bool = False):
        import torch
        self.nc, self.days, self.spd, self.seed, self.device = nc, days, steps_per_day, seed, device
        host = inputs.calibration_columns(nc, seed, freezing=freezing)
        self.nlevsno, self.nlevsoi, self.nlevgrnd = host["nlevsno"], host["nlevsoi"], host["nlevgrnd"]
        self.dtime = float(host["dtime"])
        self.base = to_device(host, device)
        self.heat_out = to_device(inputs.heat_outputs(host, diagnostics=False, fill=0.0), device)
        nlev = self.nlevsno + self.nlevgrnd
        # melt/freeze labels of the last heat step: part of the path signature
        self.heat_out["imelt"] = torch.zeros((nlev, nc), dtype=torch.int64, device=device)
        self.cfg = init_soilwater_movement(soilwater_movement_method=MOISTURE_FORM,
                                           lower_boundary_condition=BC_FLUX)
        self.workspace = adjoint.new_workspace(nc, device)
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=device)
        self.scratch = dict(smp_l=z(self.nlevgrnd, nc), hk_l=z(self.nlevgrnd, nc),
                            qcharge=z(nc), nsubsteps=z(nc), nrejected=z(nc))

    # ---- forcing and parameters ----
    def infl(self, step: int):
        """Infiltration of one step, U(0, 2e-3) mm/s, reproducible per step."""
        import torch
        g = torch.Generator(device=self.device)
        g.manual_seed(self.seed * 1000003 + step)
        return torch.rand(self.nc, generator=g, dtype=torch.float64, device=self.device) * 2.0e-3

    def params(self, m):
        b = self.base
        return {"bsw": (b["bsw"] * m[0]).contiguous(), "hksat": (b["hksat"] * m[1]).contiguous(),
                "watsat": (b["watsat"] * m[2]).contiguous()}

    def _richards(self, liq, p, infl):
        b, s = self.base, self.scratch
        st = MfState(liq, s["smp_l"], s["hk_l"], s["qcharge"], s["nsubsteps"], s["nrejected"])
        mi = MfIn(b["z"], b["zi"], b["dz"], b["qflx_rootsoi"], p["watsat"], p["hksat"], p["bsw"],
                  b["sucsat"], b["icefrac"], b["zwt"], infl)
        soilwm_moisture_form_solve(st, mi, None, b["nbedrock"], self.nlevsno, self.nlevsno + 1,
                                   self.nlevsoi, self.dtime, self.cfg, self.nlevgrnd)

    # ---- forward ----
    def _heat_dict(self, state, p):
        b = self.base
        return ({k: b[k] for k in HEAT_INPUTS} | dict(state) | dict(self.heat_out) | p |
                {"nc": self.nc})

    def _heat(self, heat):
        soil_temperature(heat, self.dtime, self.nlevsno, self.nlevgrnd)

    def forward(self, m, keep_checkpoints: bool = False):
        """Daily drainage D [days, nc], the per-day checkpoints of the heat and
        water state and a per-column signature of the path (substep iterations
        and melt/freeze labels, weighted by step)."""
        import torch
        b = self.base
        state = {k: b[k].clone() for k in HEAT_STATE}
        p = self.params(m)
        heat = self._heat_dict(state, p)
        D = torch.zeros((self.days, self.nc), dtype=torch.float64, device=self.device)
        ck = [] if keep_checkpoints else None
        sig = torch.zeros(self.nc, dtype=torch.float64, device=self.device)
        imelt = self.heat_out["imelt"]
        levw = torch.arange(1, imelt.shape[0] + 1, dtype=torch.float64,
                            device=self.device)[:, None]
        liq = state["h2osoi_liq"]
        for d in range(self.days):
            if keep_checkpoints:
                ck.append({k: state[k].clone() for k in HEAT_STATE})
            for k in range(self.spd):
                step = d * self.spd + k
                self._heat(heat)
                sig += (imelt.to(torch.float64) * levw).sum(dim=0) * (1.0e4 * (step + 1))
                self._richards(liq, p, self.infl(step))
                D[d] += self.scratch["qcharge"] * self.dtime
                sig += self.scratch["nsubsteps"] * (step + 1) + self.scratch["nrejected"]
        return D, ck, sig

    @staticmethod
    def column_loss(D, T):
        return ((D - T) ** 2).sum(dim=0)

    # ---- reverse ----
    def loss_and_gradient(self, m, T):
        """Per-column loss L_c [nc] and dL/dm [3, nc] at multipliers m [3, nc]."""
        import torch
        D, ck, _ = self.forward(m, keep_checkpoints=True)
        Lc = self.column_loss(D, T)
        p = self.params(m)
        b = self.base
        shape = tuple(b["h2osoi_liq"].shape)
        t_bar = torch.zeros(shape, dtype=torch.float64, device=self.device)
        liq_bar = torch.zeros_like(t_bar)
        ice_bar = torch.zeros_like(t_bar)
        g = torch.zeros((3, self.nc), dtype=torch.float64, device=self.device)
        empty = lambda: torch.empty((self.spd,) + shape, dtype=torch.float64, device=self.device)
        liq_r = empty()                     # water at the start of each Richards step
        t_h, liq_h, ice_h = empty(), empty(), empty()   # state at the start of each heat step
        s = self.scratch
        device_errors()
        for d in reversed(range(self.days)):
            state = {k: ck[d][k].clone() for k in HEAT_STATE}
            heat = self._heat_dict(state, p)
            liq = state["h2osoi_liq"]
            for k in range(self.spd):   # recompute the day's states from its checkpoint
                t_h[k].copy_(state["t_soisno"])
                liq_h[k].copy_(liq)
                ice_h[k].copy_(state["h2osoi_ice"])
                self._heat(heat)
                liq_r[k].copy_(liq)
                self._richards(liq, p, self.infl(d * self.spd + k))
            qbar = (2.0 * self.dtime) * (D[d] - T[d])
            for k in reversed(range(self.spd)):
                st = MfState(liq_r[k], s["smp_l"], s["hk_l"], s["qcharge"], s["nsubsteps"])
                mi = MfIn(b["z"], b["zi"], b["dz"], b["qflx_rootsoi"], p["watsat"], p["hksat"],
                          p["bsw"], b["sucsat"], b["icefrac"], b["zwt"],
                          self.infl(d * self.spd + k))
                adjoint.soilwm_moisture_form_adjoint(
                    st, mi, None, b["nbedrock"], self.nlevsno, self.nlevsoi, self.nlevgrnd,
                    self.dtime, self.cfg, liq_bar, qbar, self.workspace, grad_scale=g)
                hstate = dict(heat, t_soisno=t_h[k], h2osoi_liq=liq_h[k], h2osoi_ice=ice_h[k])
                adjoint.soil_temperature_adjoint(hstate, self.dtime, t_bar, liq_bar, ice_bar,
                                                 self.nlevsno, self.nlevgrnd, self.nlevsoi,
                                                 grad_scale=g)
        flags = device_errors()
        if flags:
            raise RuntimeError(f"device error bits {flags:#x} during the gradient (64: a column "
                               "outside the heat adjoint's scope or with a row interchange)")
        # grad_scale holds sum_j p_j dL/dp_j = m dL/dm for p = m * p0
        return Lc, g / m

    def make_target(self, sigma=(0.2, 0.2, 0.05), seed: int = 17):
        """Multipliers exp(sigma_p N(0,1)) per column and parameter, clipped to
        [0.6, 1.6] and so that porosity stays <= 0.9, and the daily drainage they
        produce (the calibration target)."""
        import torch
        rng = np.random.default_rng(seed)
        z = rng.standard_normal((3, self.nc))
        m = np.clip(np.exp(np.asarray(sigma)[:, None] * z), 0.6, 1.6)
        wmax = self.base["watsat"][:self.nlevgrnd].max(dim=0).values.cpu().numpy()
        m[2] = np.minimum(m[2], 0.9 / wmax)
        mstar = torch.from_numpy(m).to(self.device)
        T, _, _ = self.forward(mstar)
        return mstar, T


def fd_check(nc: int = 64, days: int = 1, rel_step: float = 1e-6, seed: int = 5, device="cuda",
             freezing: bool = False):
    """Adjoint gradient against central finite differences (step rel_step * m)
    for every column and parameter. Columns are independent, so one perturbed run
    per parameter and sign covers all columns; a column whose substep path
    changes under the perturbation is excluded and counted."""
    import torch
    cal = Calibration(nc, days, seed=seed, device=device, freezing=freezing)
    _, T = cal.make_target()
    m0 = torch.ones((3, nc), dtype=torch.float64, device=device)
    Lc, g = cal.loss_and_gradient(m0, T)
    _, _, sig0 = cal.forward(m0)
    rel = np.zeros((3, nc))
    normwise = np.zeros(3)
    valid = np.ones((3, nc), dtype=bool)
    fd_all = np.zeros((3, nc))
    for q in range(3):
        mp, mm = m0.clone(), m0.clone()
        mp[q] *= 1.0 + rel_step
        mm[q] *= 1.0 - rel_step
        Dp, _, sp = cal.forward(mp)
        Dm, _, sm = cal.forward(mm)
        fd = (cal.column_loss(Dp, T) - cal.column_loss(Dm, T)) / (2.0 * rel_step * m0[q])
        ok = (sp == sig0) & (sm == sig0)
        fd, gq = fd.cpu().numpy(), g[q].cpu().numpy()
        valid[q] = ok.cpu().numpy()
        v = valid[q]
        scale = max(np.max(np.abs(fd[v])), 1e-300) if v.any() else 1.0
        rel[q] = np.abs(gq - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
        normwise[q] = np.max(np.abs(gq - fd)[v]) / scale if v.any() else np.nan
        fd_all[q] = fd
    return {"columns": nc, "days": days, "rel_step": rel_step,
            # max |adjoint - fd| / max |fd| per parameter (the reference's metric)
            "max_rel_err": float(np.nanmax(normwise)),
            "max_elementwise_rel_err": float(rel[valid].max()) if valid.any() else None,
            "median_rel_err": float(np.median(rel[valid])) if valid.any() else None,
            "excluded_columns": int((~valid).any(axis=0).sum()),
            "loss": float(Lc.sum()), "grad": g.cpu().numpy(), "fd": fd_all, "valid": valid}


def main(argv=None) -> int:
    import argparse
    import json
    import time
    import torch
    ap = argparse.ArgumentParser(description="Calibration gradient through the adjoint.")
    ap.add_argument("--nc", type=int, default=262144)
    ap.add_argument("--days", type=int, default=30)
    ap.add_argument("--fd-columns", type=int, default=64)
    ap.add_argument("--fd-days", type=int, default=30)
    args = ap.parse_args(argv)
    out = {}
    keys = ("columns", "days", "rel_step", "max_rel_err", "max_elementwise_rel_err",
            "median_rel_err", "excluded_columns")
    r = fd_check(args.fd_columns, args.fd_days)
    out["fd_check"] = {k: r[k] for k in keys}
    r = fd_check(args.fd_columns, 2, freezing=True)
    out["fd_check_freezing"] = {k: r[k] for k in keys}
    print(json.dumps(out), flush=True)
    cal = Calibration(args.nc, args.days)
    _, T = cal.make_target()
    m0 = torch.ones((3, args.nc), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    D, _, _ = cal.forward(m0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    Lc, g = cal.loss_and_gradient(m0, T)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["gradient"] = {"columns": args.nc, "steps": args.days * 48, "forward_s": t1 - t0,
                       "loss_and_gradient_s": t2 - t1, "loss": float(Lc.sum()),
                       "grad_norm": float(g.norm())}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
