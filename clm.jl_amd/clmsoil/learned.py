"""Learned first-substep controller for the moisture-form Richards step.

CLM.jl starts every column's substep loop at dtsub = dtime
(src/biogeophys/soil_water_movement.jl:1259) and halves it until the error test
(:1441-1463) accepts. Here a 24-64-64-1 GELU network predicts each column's
first substep instead. The plain controller's error test still judges every
substep, and a missing model means the plain start (dtsub0 = None).

Everything numeric runs in the HIP library (include/clmsoil_learned.h):
features, inference, the loss gradient and the Adam update. torch supplies
device memory, random permutations and gathers.

Training data come from the plain controller itself. For every logged
(column, step) the label is the first substep, among dtime * 2^(-k/2) for
k = 0..11, that completes the step in the fewest substep iterations (ties go
to the longer substep). The step is run once per candidate through the
product Richards kernel (clmsoil_mf_args.dtsub0). The state is then advanced
with the plain controller.

Train (on a GPU, from the repository root):
    PYTHONPATH=clm.jl_amd python -m clmsoil.learned --nc 131072 --steps 40 --epochs 8
which writes the weights that DtPredictor.load() reads by default. With
--workload production_day the data are logged on production-day columns
(heat + Richards, BC_WATERTABLE) and the weights go to
dt_predictor_production_day.npz in the working directory; no such weights
ship, because on held-out production-day columns they did not cut substep
iterations.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _abi
from .api import (MfIn, MfState, SoilWaterMovementConfig, _mf_dict, _stream, richards_workspace,
                  soilwm_moisture_form_solve)
from .bind import byref, levels, mf_args

NIN, NHID, NPARAM, NNORM, GRAD_TILE = 24, 64, 5825, 50, 64
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 1536, 1600, 5696, 5760, 5824
N_CANDIDATES = 12
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "dt_predictor.npz"


def _lib():
    return _abi.load_library()


def _p(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-normal weights, zero biases, packed W1|b1|W2|b2|w3|b3 (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OW1:OB1] = rng.normal(0, math.sqrt(2.0 / (NIN + NHID)), NHID * NIN)
    p[OW2:OB2] = rng.normal(0, math.sqrt(2.0 / (2 * NHID)), NHID * NHID)
    p[OW3:OB3] = rng.normal(0, math.sqrt(2.0 / (NHID + 1)), NHID)
    return p


def richards_features(st: MfState, mfin: MfIn, mask, nbedrock, nlevsno: int, nlevsoi: int,
                      nlevgrnd: int, dtime: float, norm=None, out=None):
    """[24, nc] float32 features of every column (standardised when `norm` is given)."""
    import torch
    d = _mf_dict(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd)
    nc = int(d["nc"])
    if out is None:
        out = torch.empty((NIN, nc), dtype=torch.float32, device=nbedrock.device)
    if out.dtype != torch.float32 or out.shape != (NIN, nc) or out.stride(1) != 1:
        raise ValueError("out must be a float32 [24, nc] view with unit column stride")
    L = _lib()
    a, lev = mf_args(d), levels(d)
    _abi.check(L, L.clmsoil_richards_features(byref(a), byref(lev), float(dtime), _p(norm),
                                              _p(out), out.stride(0), _stream()),
               "richards_features")
    return out


def _check_x(x):
    if x.dim() != 2 or x.shape[0] != NIN or x.stride(1) != 1 or str(x.dtype) != "torch.float32":
        raise ValueError("x must be a float32 [24, n] tensor with unit column stride")


def mlp_forward(params, x):
    import torch
    _check_x(x)
    n = x.shape[1]
    y = torch.empty(n, dtype=torch.float32, device=x.device)
    L = _lib()
    _abi.check(L, L.clmsoil_mlp_forward(_p(params), _p(x), n, x.stride(0), _p(y), _stream()),
               "mlp_forward")
    return y


@dataclass
class DtPredictor:
    """Trained network and its normalisation (both float32 device tensors)."""
    params: object
    norm: object

    def predict_dtsub0(self, st: MfState, mfin: MfIn, mask, nbedrock, nlevsno: int, nlevsoi: int,
                       nlevgrnd: int, dtime: float, out=None):
        import torch
        d = _mf_dict(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd)
        nc = int(d["nc"])
        if out is None:
            out = torch.full((nc,), float(dtime), dtype=torch.float64, device=nbedrock.device)
        L = _lib()
        a, lev = mf_args(d), levels(d)
        _abi.check(L, L.clmsoil_mlp_predict_dtsub0(byref(a), byref(lev), float(dtime),
                                                   _p(self.params), _p(self.norm), _p(out),
                                                   _stream()), "mlp_predict_dtsub0")
        return out

    def solve(self, st: MfState, mfin: MfIn, mask, nbedrock, joff: int, joff_zi: int,
              nlevsoi: int, dtime: float, cfg: SoilWaterMovementConfig,
              nlevgrnd: int | None = None, workspace=None) -> None:
        """soilwm_moisture_form_solve with the network fused into the Richards
        kernel: each column's first substep is predicted as it is loaded
        (clmsoil_soilwm_moisture_form_solve_learned). Same arguments as
        api.soilwm_moisture_form_solve; mfin.dtsub0 must be None."""
        from .api import MOISTURE_FORM, CLMSoilError, _check_moisture_form_bc, richards_workspace
        _check_moisture_form_bc(cfg)
        if joff_zi != joff + 1:
            raise CLMSoilError("joff_zi must equal joff + 1 (zi has one more level)")
        c = cfg.to_abi()
        c.soilwater_movement_method = MOISTURE_FORM
        d = _mf_dict(st, mfin, mask, nbedrock, joff, nlevsoi, nlevgrnd or nlevsoi + 5)
        if d["nc"] == 0:
            return
        d["workspace"] = workspace if workspace is not None else richards_workspace(nbedrock.device)
        L = _lib()
        a, lev = mf_args(d), levels(d)
        _abi.check(L, L.clmsoil_soilwm_moisture_form_solve_learned(
            byref(a), byref(lev), float(dtime), byref(c), _p(self.params), _p(self.norm),
            _stream()), "soilwm_moisture_form_solve_learned")

    def save(self, path) -> None:
        np.savez(path, params=self.params.detach().cpu().numpy(),
                 norm=self.norm.detach().cpu().numpy())

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device="cuda") -> "DtPredictor":
        import torch
        z = np.load(path)   # plain arrays; no pickled objects
        params, norm = z["params"], z["norm"]
        if params.shape != (NPARAM,) or norm.shape != (NNORM,):
            raise ValueError(f"{path}: not a {NPARAM}-parameter first-substep model")
        return cls(torch.from_numpy(params.astype(np.float32)).to(device),
                   torch.from_numpy(norm.astype(np.float32)).to(device))


class AdamTrainer:
    """Mean-squared-error training of the network with the library's kernels."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, max_batch=8192):
        import torch
        self.params = params
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.t = 0
        self.lr, self.betas, self.eps = lr, betas, eps
        nt = (max_batch + GRAD_TILE - 1) // GRAD_TILE
        self.max_batch = max_batch
        self.partial = torch.empty(nt * NPARAM, dtype=torch.float32, device=params.device)
        self.loss_partial = torch.empty(nt, dtype=torch.float32, device=params.device)

    def _grad(self, x, t):
        _check_x(x)
        n = x.shape[1]
        if n > self.max_batch:
            raise ValueError("batch larger than max_batch")
        if t.shape != (n,) or not t.is_contiguous():
            raise ValueError("t must be a contiguous [n] tensor")
        L = _lib()
        _abi.check(L, L.clmsoil_mlp_grad(_p(self.params), _p(x), _p(t), n, x.stride(0), n,
                                         _p(self.partial), _p(self.loss_partial), _stream()),
                   "mlp_grad")
        return (n + GRAD_TILE - 1) // GRAD_TILE

    def step(self, x, t):
        """One Adam step on the batch (x [24, n] standardised, t [n]); returns the
        batch loss as a device scalar."""
        nparts = self._grad(x, t)
        self.t += 1
        L = _lib()
        _abi.check(L, L.clmsoil_mlp_adam(_p(self.params), _p(self.m), _p(self.v), _p(self.partial),
                                         nparts, self.t, self.lr, self.betas[0], self.betas[1],
                                         self.eps, None, _stream()), "mlp_adam")
        return self.loss_partial[:nparts].sum()

    def gradient(self, x, t):
        """(loss, gradient) of the batch without changing the parameters: the Adam
        kernel's reduction with a zero learning rate on scratch moments."""
        import torch
        nparts = self._grad(x, t)
        g = torch.empty_like(self.params)
        m, v = torch.zeros_like(self.params), torch.zeros_like(self.params)
        L = _lib()
        _abi.check(L, L.clmsoil_mlp_adam(_p(self.params), _p(m), _p(v), _p(self.partial), nparts,
                                         1, 0.0, 0.9, 0.999, 1e-8, _p(g), _stream()), "mlp_adam")
        return self.loss_partial[:nparts].sum(), g


def candidate_dtsub0(dtime: float, k: int) -> float:
    return dtime * 2.0 ** (-0.5 * k)


def config0_day(nc: int, seed: int, device="cuda"):
    """Columns of the config-0 distribution (20 soil layers, BC_ZERO_FLUX, no
    roots) as device tensors, plus a generator of per-step infiltration."""
    import torch
    from . import inputs
    from .api import to_device
    d = inputs.richards_config0(nc=nc, seed=seed, fill=0.0)
    d["nrejected"] = np.zeros(nc)
    dev = to_device(d, device)
    rng = np.random.default_rng(seed + 7919)

    def infl():
        return torch.from_numpy(rng.uniform(0.0, 2.0e-3, nc)).to(device)
    return dev, infl


def mf_views(dev, dtsub0=None):
    st = MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"],
                 dev.get("nrejected"))
    mi = MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
              dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"], None,
              dtsub0)
    return st, mi


def run_step(dev, cfg: SoilWaterMovementConfig, dtsub0=None, pred: "DtPredictor | None" = None,
             workspace=None):
    """One Richards step of the config-0 state `dev`: plain, from given first
    substeps `dtsub0`, or with `pred`'s network fused into the kernel."""
    st, mi = mf_views(dev, dtsub0)
    args = (st, mi, None, dev["nbedrock"], dev["nlevsno"], dev["nlevsno"] + 1, dev["nlevsoi"],
            dev["dtime"], cfg, dev["nlevgrnd"], workspace)
    if pred is not None:
        pred.solve(*args)
    else:
        soilwm_moisture_form_solve(*args)


def log_plain_controller(nc: int, steps: int, seed: int, cfg: SoilWaterMovementConfig,
                         device="cuda", progress=None):
    """Features (raw, [24, nc*steps]) and labels log(dt_best / dtime) logged while
    the plain controller advances `nc` config-0 columns through `steps` steps."""
    import torch
    dev, infl = config0_day(nc, seed, device)
    dtime = float(dev["dtime"])
    X = torch.empty((NIN, nc * steps), dtype=torch.float32, device=device)
    Y = torch.empty(nc * steps, dtype=torch.float32, device=device)
    work = dict(dev)
    for s in range(steps):
        dev["qflx_infl"] = infl()
        work["qflx_infl"] = dev["qflx_infl"]
        st, mi = mf_views(dev)
        richards_features(st, mi, None, dev["nbedrock"], dev["nlevsno"], dev["nlevsoi"],
                          dev["nlevgrnd"], dtime, out=X[:, s * nc:(s + 1) * nc])
        best_n = torch.full((nc,), float("inf"), dtype=torch.float64, device=device)
        best_dt = torch.full((nc,), dtime, dtype=torch.float64, device=device)
        for k in range(N_CANDIDATES):
            dt0 = candidate_dtsub0(dtime, k)
            for key in ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps", "nrejected"):
                work[key] = dev[key].clone()
            run_step(work, cfg, torch.full((nc,), dt0, dtype=torch.float64, device=device))
            better = work["nsubsteps"] < best_n
            best_n = torch.where(better, work["nsubsteps"], best_n)
            best_dt = torch.where(better, torch.full_like(best_dt, max(dt0, cfg.dtmin)), best_dt)
        Y[s * nc:(s + 1) * nc] = torch.log(best_dt / dtime).float()
        run_step(dev, cfg)   # the plain controller advances the state
        if progress:
            progress(s, float(best_n.mean()), float(dev["nsubsteps"].mean()))
    return X, Y


def production_day_state(nc: int, seed: int, device="cuda"):
    """Columns of the production day (0-12 snow layers, bedrock index 5-20,
    water table 0.5-10 m, heat and Richards inputs) as device tensors."""
    import torch
    from .tuning import production_state
    dev = production_state(nc, seed)
    dev["nrejected"] = torch.zeros(nc, dtype=torch.float64, device=device)
    return dev


def log_production_day(nc: int, steps: int, seed: int, cfg: SoilWaterMovementConfig,
                       device="cuda", progress=None):
    """Features and labels as log_plain_controller, logged on production-day
    columns: each step runs the heat step, logs the Richards state it leaves
    (features, best candidate first substep) and advances it with the plain
    controller (BC_WATERTABLE in `cfg`)."""
    import torch
    from .tuning import heat_step
    dev = production_day_state(nc, seed, device)
    dtime = float(dev["dtime"])
    X = torch.empty((NIN, nc * steps), dtype=torch.float32, device=device)
    Y = torch.empty(nc * steps, dtype=torch.float32, device=device)
    work = dict(dev)
    for s in range(steps):
        heat_step(dev, dtime)
        st, mi = mf_views(dev)
        richards_features(st, mi, None, dev["nbedrock"], dev["nlevsno"], dev["nlevsoi"],
                          dev["nlevgrnd"], dtime, out=X[:, s * nc:(s + 1) * nc])
        best_dt, best_n = best_candidate_dtsub0(dev, cfg, work)
        Y[s * nc:(s + 1) * nc] = torch.log(best_dt / dtime).float()
        run_step(dev, cfg)   # the plain controller advances the state
        if progress:
            progress(s, float(best_n.mean()), float(dev["nsubsteps"].mean()))
    return X, Y


def fit_normalisation(X, Y):
    """norm[50] = mean[24], scale[24], y_mean, y_scale (unit scale for constants)."""
    import torch
    mu = X.mean(dim=1)
    sd = X.std(dim=1)
    sd = torch.where(sd > 1e-6, sd, torch.ones_like(sd))
    ym, ys = Y.mean(), Y.std()
    ys = ys if float(ys) > 1e-6 else torch.ones_like(ys)
    return torch.cat([mu, sd, ym.reshape(1), ys.reshape(1)]).float().contiguous()


def train(X, Y, epochs=6, batch=4096, lr=2e-3, seed=0, progress=None) -> DtPredictor:
    """Fit the network to (X raw features, Y log-ratio labels) on the GPU."""
    import torch
    device = X.device
    norm = fit_normalisation(X, Y)
    Xs = ((X - norm[:NIN, None]) / norm[NIN:2 * NIN, None]).contiguous()
    Ys = ((Y - norm[2 * NIN]) / norm[2 * NIN + 1]).contiguous()
    params = torch.from_numpy(init_params(seed)).to(device)
    tr = AdamTrainer(params, lr=lr, max_batch=batch)
    n = X.shape[1]
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device=device)
        tot = torch.zeros((), dtype=torch.float32, device=device)
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            tot += tr.step(Xs[:, idx].contiguous(), Ys[idx].contiguous())
            nb += 1
        if progress:
            progress(ep, float(tot) / max(nb, 1))
    return DtPredictor(params, norm)


def column_water(dev):
    """Soil liquid water of every column (kg/m2) over the Richards layers."""
    ns, n = int(dev["nlevsno"]), int(dev["nlevsoi"])
    return dev["h2osoi_liq"][ns:ns + n].sum(dim=0)


def best_candidate_dtsub0(dev, cfg: SoilWaterMovementConfig, work=None):
    """Per column, the candidate first substep dtime 2^(-k/2) (k = 0..11) that
    completes the step from the current state in the fewest substep iterations,
    found by running the step once per candidate on a copy (the labels the
    network is trained on)."""
    import torch
    nc = int(dev["nc"]) if "nc" in dev else dev["nbedrock"].shape[0]
    dtime = float(dev["dtime"])
    device = dev["nbedrock"].device
    work = dict(dev) if work is None else work
    best_n = torch.full((nc,), float("inf"), dtype=torch.float64, device=device)
    best_dt = torch.full((nc,), dtime, dtype=torch.float64, device=device)
    for k in range(N_CANDIDATES):
        dt0 = candidate_dtsub0(dtime, k)
        for key in ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps", "nrejected"):
            work[key] = dev[key].clone()
        run_step(work, cfg, torch.full((nc,), dt0, dtype=torch.float64, device=device))
        better = work["nsubsteps"] < best_n
        best_n = torch.where(better, work["nsubsteps"], best_n)
        best_dt = torch.where(better, torch.full_like(best_dt, max(dt0, cfg.dtmin)), best_dt)
    return best_dt, best_n


def compare_controllers(pred: DtPredictor | None, nc: int, steps: int, seed: int,
                        cfg: SoilWaterMovementConfig, device="cuda", hindsight=False,
                        two_launch=False) -> dict:
    """Advance the same config-0 columns through `steps` steps with the plain
    controller and with the learned first substep (when `pred` is given), and,
    with `hindsight`, with each column's best candidate first substep chosen by
    trying them all first (untimed): the ceiling any first-substep predictor can
    reach under the unchanged controller. The learned controller runs the
    network fused into the Richards kernel; with `two_launch`, also as a
    prediction launch before each step (`learned_two_launch`, the same first
    substeps). Returns, per controller:
    column-steps/s of the timed GPU work (prediction included), mean substep iterations,
    rejected-substep rate, and the largest water-balance residual
    |dW - (infl - qcharge - roots) dt| in kg/m2."""
    import torch
    out = {}
    modes = (["plain"] + (["learned"] if pred is not None else []) +
             (["learned_two_launch"] if pred is not None and two_launch else []) +
             (["hindsight"] if hindsight else []))
    for mode in modes:
        dev, infl = config0_day(nc, seed, device)
        dtime = float(dev["dtime"])
        dt0 = (torch.empty(nc, dtype=torch.float64, device=device)
               if mode == "learned_two_launch" else None)
        ws = richards_workspace(device)
        work = dict(dev)
        evs = []
        iters = torch.zeros((), dtype=torch.float64, device=device)
        rej = torch.zeros((), dtype=torch.float64, device=device)
        bal = torch.zeros((), dtype=torch.float64, device=device)
        for _ in range(steps):
            dev["qflx_infl"] = infl()
            w0 = column_water(dev)
            if mode == "hindsight":
                work["qflx_infl"] = dev["qflx_infl"]
                dt0, _ = best_candidate_dtsub0(dev, cfg, work)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            if mode == "learned":
                run_step(dev, cfg, pred=pred, workspace=ws)   # inference fused in the kernel
            elif mode == "learned_two_launch":
                st, mi = mf_views(dev)
                pred.predict_dtsub0(st, mi, None, dev["nbedrock"], dev["nlevsno"], dev["nlevsoi"],
                                    dev["nlevgrnd"], dtime, out=dt0)
                run_step(dev, cfg, dt0, workspace=ws)
            else:
                run_step(dev, cfg, dt0, workspace=ws)
            e1.record()
            evs.append((e0, e1))
            roots = dev["qflx_rootsoi"][:int(dev["nlevsoi"])].sum(dim=0)
            resid = column_water(dev) - w0 - (dev["qflx_infl"] - dev["qcharge"] - roots) * dtime
            bal = torch.maximum(bal, resid.abs().max())
            iters += dev["nsubsteps"].sum()
            rej += dev["nrejected"].sum()
        torch.cuda.synchronize()
        ms = sum(a.elapsed_time(b) for a, b in evs)
        out[mode] = {"column_steps_per_s": nc * steps / (ms / 1e3), "ms_per_step": ms / steps,
                     "mean_iterations": float(iters) / (nc * steps),
                     "rejected_rate": float(rej) / float(iters),
                     "max_balance_residual": float(bal)}
    if pred is not None:
        out["speedup"] = out["learned"]["column_steps_per_s"] / out["plain"]["column_steps_per_s"]
    if pred is not None and two_launch:
        out["speedup_two_launch"] = (out["learned_two_launch"]["column_steps_per_s"] /
                                     out["plain"]["column_steps_per_s"])
    if hindsight:
        out["hindsight_speedup"] = (out["hindsight"]["column_steps_per_s"] /
                                    out["plain"]["column_steps_per_s"])
    return out


def main(argv=None) -> int:
    import argparse
    import json
    import time
    import torch
    from .api import init_soilwater_movement, MOISTURE_FORM, BC_WATERTABLE, BC_ZERO_FLUX
    ap = argparse.ArgumentParser(description="Train the first-substep predictor on the GPU.")
    ap.add_argument("--workload", choices=("config0", "production_day"), default="config0",
                    help="columns the training data are logged on: config-0 columns "
                         "(Richards only, BC_ZERO_FLUX) or production-day columns (heat + "
                         "Richards, BC_WATERTABLE, ragged snow and bedrock)")
    ap.add_argument("--nc", type=int, default=131072)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--out", default=None,
                    help="weights file (default: the workload's committed weights)")
    args = ap.parse_args(argv)
    prod = args.workload == "production_day"
    out = args.out or ("dt_predictor_production_day.npz" if prod else str(DEFAULT_WEIGHTS))
    cfg = init_soilwater_movement(soilwater_movement_method=MOISTURE_FORM,
                                  lower_boundary_condition=BC_WATERTABLE if prod else BC_ZERO_FLUX)
    t0 = time.time()
    logger = log_production_day if prod else log_plain_controller
    X, Y = logger(args.nc, args.steps, args.seed, cfg,
                  progress=lambda s, b, p: print(
                      f"# log step {s}: best-candidate iters {b:.2f}, plain {p:.2f}", flush=True))
    torch.cuda.synchronize()
    t1 = time.time()
    pred = train(X, Y, epochs=args.epochs,
                 progress=lambda e, l: print(f"# epoch {e}: loss {l:.4f}", flush=True))
    torch.cuda.synchronize()
    pred.save(out)
    print(json.dumps({"workload": args.workload, "samples": int(X.shape[1]), "log_s": t1 - t0,
                      "train_s": time.time() - t1, "out": out}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
