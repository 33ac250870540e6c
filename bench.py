"""Benchmark: the production day of per-column soil solves on one MI355X.

Workload (one "step" = one 1800 s model step of every column): 1,048,576 uneven
columns (0..12 snow layers, bedrock index 5..20, water table 0.5..10 m), each
running the snow+soil heat step then the moisture-form Richards step with a
water-table lower boundary, FP64, synthetic inputs already resident in HBM.
Prints one JSON line: column-steps/s, the HBM roofline of the dominant kernel,
and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "clm.jl_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch at 1,048,576 columns (seed 2), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE counters, each scaled by what the counter reads for a plain 2 GiB
# device copy (fetch x2.0, write x1.0); recorded in
# profiles/hbm_traffic_production_day_1M.md. FP64 FLOP per launch (fma = 2) from
# the SQ_INSTS_VALU_*_F64 counters of the same kernels, recorded in
# profiles/sq_counters_production_day_1M.md. Both were measured on the kernel
# sources of commit PROFILED_AT, whose hash is PROFILED_SOURCES: when the sources
# differ, the bench line reports the traffic as null (the profile is stale).
MEASURED_TRAFFIC_1M = {"heat": 6.287e9, "richards": 3.076e9}
FP64_FLOP_1M = {"heat": 2.938e10, "richards": 4.111e10}
PROFILED_AT = "b6b0cde"
PROFILED_SOURCES = "c292c36b66038861"
KERNEL_SOURCES = ("common.hpp", "clm_math.h", "clm_math_tables.h", "mlp_wave.hpp", "rcolumn.hpp",
                  "richards.hpp", "richards.hip", "heat.hpp", "heat.hip", "Makefile")


def kernel_sources_hash() -> str:
    """sha256 (first 16 hex digits) of the heat and Richards kernel sources and
    the build flags."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((REPO / "clm.jl_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:16]


FP64_PEAK_TFLOPS = 78.6
METRIC = "column-steps/s @1M cols FP64 soil solves; HBM GB/s frac; learned vs plain dt"


def algorithmic_bytes(nb, snl, bc_watertable=True):
    """Bytes per column-step, each input read once and each output written once
    (FP64 / Int64): Richards 8*(12n+4) (+8*(n+2) for zwt and zi under
    BC_WATERTABLE); heat 8*(9L + 3ns + 198), L = ns + 25."""
    import numpy as np
    n = nb.astype(np.float64)
    ns = (-snl).astype(np.float64)
    br = 8.0 * (12.0 * n + 4.0) + (8.0 * (n + 2.0) if bc_watertable else 0.0)
    L = ns + 25.0
    bh = 8.0 * (9.0 * L + 3.0 * ns + 198.0)
    return float(br.sum()), float(bh.sum())


def make_state(nc, seed):
    import torch
    from clmsoil import api, inputs
    t0 = time.time()
    d = inputs.production_day(nc=nc, seed=seed, fill=0.0)
    outs = inputs.heat_outputs(d, diagnostics=False, fill=0.0)
    d.update(outs)
    host = d
    dev = api.to_device(d)
    torch.cuda.synchronize()
    return host, dev, time.time() - t0


def gpu_step(dev, nc, cfg, heat_keys, events=None):
    from clmsoil import api
    if events:
        events[0].record()
    api.soil_temperature({k: dev[k] for k in heat_keys} | {"nc": nc}, 1800.0)
    if events:
        events[1].record()
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"])
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"], dev["smpmin"])
    api.soilwater_moisture_form(st, mi, None, dev["nbedrock"], 12, 20, 25, 1800.0, cfg)
    if events:
        events[2].record()


HEAT_KEYS = ["snl", "nbedrock", "z", "dz", "zi", "t_soisno", "h2osoi_liq", "h2osoi_ice",
             "excess_ice", "watsat", "tkmg", "tkdry", "csol", "bsw", "sucsat", "frac_sno_eff",
             "frac_h2osfc", "hs_top_snow", "hs_soil", "hs_top", "hs_h2osfc", "dhsdT",
             "sabg_lyr_col", "eflx_bot", "t_h2osfc", "h2osfc", "h2osno_no_layers", "snow_depth",
             "int_snow", "t_grnd", "xmf", "xmf_h2osfc", "qflx_snomelt", "qflx_snofrz",
             "qflx_snow_drain", "qflx_h2osfc_to_ice", "eflx_fgr12", "qflx_snomelt_lyr",
             "qflx_snofrz_lyr"]


def host_cpus():
    """CPUs this process may run on (nproc) and the CPU model (lscpu's)."""
    n = len(os.sched_getaffinity(0))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(host, sample, steps, reps=5):
    """The oracle (a C restatement of CLM.jl's CPU path, OpenMP over columns on
    OMP_NUM_THREADS threads, -O3) on the first `sample` columns for `steps`
    steps of the production day: one warm-up pass, then the best of `reps`."""
    import numpy as np
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle
    from clmsoil import bind
    sub = {}
    for k, v in host.items():
        if isinstance(v, np.ndarray):
            sub[k] = np.ascontiguousarray(v[..., :sample])
        else:
            sub[k] = v
    sub["nc"] = sample
    cfg = bind.swm_config(lower_bc=3)
    hcfg = bind.heat_config()
    best = float("inf")
    for rep in range(reps + 1):
        s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sub.items()}
        t0 = time.perf_counter()
        for _ in range(steps):
            oracle.soil_temperature(s, hcfg)
            oracle.soilwater_moisture_form(s, cfg)
        if rep:                                  # pass 0 is the warm-up
            best = min(best, time.perf_counter() - t0)
    return sample * steps / best, oracle.num_threads(), best


def zero_flux_day(dev, snap, nc, steps):
    """The same production day with CLM.jl's live lower boundary (moisture form
    with BC_ZERO_FLUX, clm_driver.jl:2043-2047, which accepts ragged bedrock):
    from the same warmed-up state, the day as one graph launch, then the same
    steps launched one by one with HIP events around each kernel."""
    import torch
    from clmsoil import api, day
    for k, v in snap.items():
        dev[k].copy_(v)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX, use_bedrock=True)
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"])
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"],
                  dev["smpmin"])
    plan = day.DayPlan({k: dev[k] for k in HEAT_KEYS} | {"nc": nc}, st, mi, None, dev["nbedrock"],
                       12, 20, 25, 1800.0, cfg, steps)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    plan.launch()
    b.record()
    torch.cuda.synchronize()
    plan.close()
    day_ms = a.elapsed_time(b)
    for k, v in snap.items():
        dev[k].copy_(v)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for e in ev:
        gpu_step(dev, nc, cfg, HEAT_KEYS, e)
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    assert bool((dev["qcharge"] == 0.0).all())
    return {"workload": f"{nc} columns, the production day's columns with BC_ZERO_FLUX "
                        f"(use_bedrock), {steps} x 1800 s steps of heat + Richards",
            "column_steps_per_s": nc * steps / (day_ms / 1e3), "ms_per_step": day_ms / steps,
            "kernels_ms": {"heat": sum(e[0].elapsed_time(e[1]) for e in ev) / steps,
                           "richards": sum(e[1].elapsed_time(e[2]) for e in ev) / steps},
            "mean_substep_iterations": dev["nsubsteps"].float().mean().item()}


def production_hindsight(dev, snap, nc, steps, st, nrej):
    """The ceiling of any first-substep predictor on the production day: at every
    step (heat, then Richards) each column's best candidate first substep
    dtime 2^(-k/2), k = 0..11 (fewest substep iterations), found by trying them
    all on a copy of the state (untimed). The Richards solve from those first
    substeps is timed against the plain solve of the same state (both eager);
    the day follows the plain controller's states."""
    import torch
    from clmsoil import api, learned
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    dt0 = torch.empty(nc, dtype=torch.float64, device="cuda")
    mi = {}
    for tag, d in (("plain", None), ("best", dt0)):
        mi[tag] = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"],
                           dev["hksat"], dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"],
                           dev["qflx_infl"], dev["smpmin"], d)
    keys = ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps")

    def solve(tag):
        api.soilwm_moisture_form_solve(st, mi[tag], None, dev["nbedrock"], 12, 13, 20, 1800.0,
                                       cfg, 25)

    for k, v in snap.items():
        dev[k].copy_(v)
    ev = {"heat": [], "plain": [], "best": []}
    it = {"plain": torch.zeros((), dtype=torch.float64, device="cuda"),
          "best": torch.zeros((), dtype=torch.float64, device="cuda")}

    def timed(tag, fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        ev[tag].append((a, b))

    for _ in range(steps):
        timed("heat", lambda: api.soil_temperature({k: dev[k] for k in HEAT_KEYS} | {"nc": nc},
                                                   1800.0))
        save = {k: dev[k].clone() for k in keys}
        best_n = torch.full((nc,), float("inf"), dtype=torch.float64, device="cuda")
        best_dt = torch.full((nc,), 1800.0, dtype=torch.float64, device="cuda")
        for kk in range(learned.N_CANDIDATES):
            d0 = learned.candidate_dtsub0(1800.0, kk)
            for k in keys:
                dev[k].copy_(save[k])
            dt0.fill_(d0)
            solve("best")
            better = dev["nsubsteps"] < best_n
            best_n = torch.where(better, dev["nsubsteps"], best_n)
            best_dt = torch.where(better, torch.full_like(best_dt, max(d0, cfg.dtmin)), best_dt)
        dt0.copy_(best_dt)
        for k in keys:
            dev[k].copy_(save[k])
        timed("best", lambda: solve("best"))
        it["best"] += dev["nsubsteps"].sum()
        for k in keys:
            dev[k].copy_(save[k])
        timed("plain", lambda: solve("plain"))
        it["plain"] += dev["nsubsteps"].sum()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = {t: sum(a.elapsed_time(b) for a, b in e) / steps for t, e in ev.items()}
    return {"richards_ms_plain": ms["plain"], "richards_ms_best": ms["best"],
            "heat_ms": ms["heat"],
            "mean_iterations_plain": float(it["plain"]) / (nc * steps),
            "mean_iterations_best": float(it["best"]) / (nc * steps),
            "speedup_richards": ms["plain"] / ms["best"],
            "speedup_step": (ms["heat"] + ms["plain"]) / (ms["heat"] + ms["best"])}


def learned_production_day(dev, snap, nc, steps):
    """The committed first-substep network (trained on config-0 columns) on the
    production day itself (heat + Richards, BC_WATERTABLE): the day as one graph
    launch with the network fused into the Richards kernel, against the plain day
    from the same warmed-up state; then both again launched step by step to
    count substep iterations and rejections over the whole day."""
    import torch
    from clmsoil import api, day, learned
    pred = learned.DtPredictor.load()
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    nrej = torch.zeros(nc, dtype=torch.float64, device="cuda")
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"],
                     nrej)
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"],
                  dev["smpmin"])
    out = {}
    for mode in ("plain", "learned"):
        p = pred if mode == "learned" else None
        for k, v in snap.items():
            dev[k].copy_(v)
        plan = day.DayPlan({k: dev[k] for k in HEAT_KEYS} | {"nc": nc}, st, mi, None,
                           dev["nbedrock"], 12, 20, 25, 1800.0, cfg, steps, predictor=p)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        plan.launch()
        b.record()
        torch.cuda.synchronize()
        plan.close()
        ms = a.elapsed_time(b)
        for k, v in snap.items():
            dev[k].copy_(v)
        iters = torch.zeros((), dtype=torch.float64, device="cuda")
        rej = torch.zeros((), dtype=torch.float64, device="cuda")
        for _ in range(steps):
            api.soil_temperature({k: dev[k] for k in HEAT_KEYS} | {"nc": nc}, 1800.0)
            if p is not None:
                p.solve(st, mi, None, dev["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
            else:
                api.soilwm_moisture_form_solve(st, mi, None, dev["nbedrock"], 12, 13, 20, 1800.0,
                                               cfg, 25)
            iters += dev["nsubsteps"].sum()
            rej += nrej.sum()
        torch.cuda.synchronize()
        assert api.device_errors() == 0
        out[mode] = {"column_steps_per_s": nc * steps / (ms / 1e3), "ms_per_step": ms / steps,
                     "mean_iterations": float(iters) / (nc * steps),
                     "rejected_rate": float(rej) / float(iters)}
    out["speedup"] = out["learned"]["column_steps_per_s"] / out["plain"]["column_steps_per_s"]
    out["hindsight"] = production_hindsight(dev, snap, nc, steps, st, nrej)
    out["workload"] = (f"{nc} production-day columns x {steps} steps of heat + Richards "
                       f"(BC_WATERTABLE), each day one graph launch; the network fused into "
                       f"the Richards kernel; iteration counts from a step-by-step rerun "
                       f"(solve without the aquifer recharge pass)")
    return out


def learned_vs_plain(nc, steps):
    """The committed first-substep network against the plain controller: the same
    config-0 columns (20 soil layers, BC_ZERO_FLUX, per-step infiltration
    U(0, 2e-3) mm/s, held-out seed) advanced `steps` steps by each."""
    from clmsoil import api, learned
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    pred = learned.DtPredictor.load()
    r = learned.compare_controllers(pred, nc, steps, 2024, cfg, hindsight=True, two_launch=True)
    r["workload"] = (f"{nc} config-0 columns x {steps} steps, Richards only (BC_ZERO_FLUX), "
                     f"the network fused into the Richards kernel ('learned') or run as a "
                     f"prediction launch before each step ('learned_two_launch', the same "
                     f"first substeps, prediction included in the timing); 'hindsight' runs each "
                     f"column-step with the best of the 12 candidate first substeps, found by "
                     f"trying them all first (untimed): the ceiling of any first-substep "
                     f"predictor under the unchanged controller")
    return r


def calibration_section(nc=262144, days=30, fd_columns=64, fd_days=30):
    """The differentiable-calibration workload at its size: 262,144 columns x 30
    days (1,440 steps) of heat + Richards (BC_FLUX), the loss of daily drainage
    against a target from perturbed parameters, and its gradient w.r.t. per-column
    multipliers of b, ksat and porosity through the hand-written adjoints of both
    steps with one checkpoint per day; plus the adjoint against central
    differences (relative step 1e-6) on 64 columns over the same 30 days, and on
    64 freezing columns over 2 days."""
    import torch
    from clmsoil import calibration
    keys = ("columns", "days", "rel_step", "max_rel_err", "median_rel_err", "excluded_columns")
    r = calibration.fd_check(fd_columns, fd_days)
    out = {"fd_check": {k: r[k] for k in keys}}
    # columns around the freezing point (the heat-step adjoint with phase change);
    # two days: gradients through layers held at their supercooled-water limit grow
    # by orders of magnitude per day, so longer horizons are not a useful check
    r = calibration.fd_check(fd_columns, 2, freezing=True)
    out["fd_check_freezing"] = {k: r[k] for k in keys}
    cal = calibration.Calibration(nc, days)
    _, T = cal.make_target()
    m0 = torch.ones((3, nc), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cal.forward(m0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    Lc, g = cal.loss_and_gradient(m0, T)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    assert bool(torch.isfinite(g).all())
    out["gradient"] = {"columns": nc, "steps": days * 48, "forward_s": t1 - t0,
                       "loss_and_gradient_s": t2 - t1, "loss": float(Lc.sum()),
                       "grad_norm": float(g.norm())}
    return out


def zd09_bytes_per_column(n):
    """Algorithmic bytes of one Zeng-Decker 2009 column-step, each input read once
    and each output written once (FP64): reads z, dz, zi, watsat, hksat, bsw,
    sucsat, h2osoi_liq, h2osoi_ice, qflx_rootsoi per soil layer, smpmin, zwt,
    qflx_infl and the two water-table-layer values (t_soisno, h2osoi_vol); writes
    icefrac, smp_l, hk_l, h2osoi_liq per layer, qcharge and qflx_deficit."""
    return 8 * (14 * n + 7)


def zengdecker_2009(nc, steps, cpu_sample):
    """The aquifer configuration of the soil water: per 1800 s step the soil
    water fractions, the Zeng-Decker 2009 step (CLM.jl's default soil water
    method) and the water table update, on `nc` columns with 20 soil layers, the
    water table at 0.3-12 m and 30 % of the columns frozen; `steps` consecutive
    steps on the evolving state, each launch bracketed by HIP events on its
    stream; the oracle is timed on the Zeng-Decker step of the first
    `cpu_sample` columns."""
    import numpy as np
    import torch
    from clmsoil import api, inputs
    d = inputs.water_table_columns(nc=nc, seed=5)
    t = api.to_device(d)
    keys = ("z", "dz", "zi", "watsat", "hksat", "bsw", "sucsat", "smpmin", "h2osoi_liq",
            "h2osoi_ice", "h2osoi_vol", "t_soisno", "zwt", "qflx_infl", "qflx_rootsoi",
            "icefrac", "smp_l", "hk_l", "qcharge", "qflx_deficit")
    state = {k: t[k] for k in keys}
    frac = {k: t[k] for k in ("watsat", "dz", "h2osoi_ice", "excess_ice", "eff_porosity",
                              "icefrac")} | {"nc": nc}
    wt = {k: t[k] for k in ("dz", "z", "zi", "t_soisno", "watsat", "sucsat", "bsw", "qcharge",
                            "h2osoi_liq", "h2osoi_ice", "zwt", "wa", "zwt_perched",
                            "frost_table", "h2osoi_vol", "qflx_drain", "qflx_rsub_sat",
                            "qflx_drain_perched")} | {"nc": nc}
    cfg = api.init_soilwater_movement()

    def step(ev=None):
        if ev:
            ev[0].record()
        api.set_soil_water_fractions(frac)
        if ev:
            ev[1].record()
        api.soilwater_zengdecker2009(state, 1800.0, cfg)
        if ev:
            ev[2].record()
        api.water_table(wt, 1800.0)
        if ev:
            ev[3].record()
    for _ in range(2):
        step()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    torch.cuda.synchronize()
    for e in ev:
        step(e)
    torch.cuda.synchronize()
    ms_f = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    ms = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    ms_w = sum(e[2].elapsed_time(e[3]) for e in ev) / steps
    assert bool(torch.isfinite(t["h2osoi_liq"]).all()) and bool(torch.isfinite(t["zwt"]).all())
    by = zd09_bytes_per_column(d["nlevsoi"]) * nc
    by_f = 8 * 6 * d["nlevsoi"] * nc      # watsat, dz, ice, excess ice in; two fractions out
    out = {"workload": f"{nc} columns x {steps} steps, 20 soil layers, zwt 0.3-12 m, "
                       f"30% frozen columns, 1800 s",
           "ms_per_step": ms, "column_steps_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_column": zd09_bytes_per_column(d["nlevsoi"])},
           "soil_water_fractions": {
               "ms_per_launch": ms_f,
               "roofline": {"bound": "hbm", "achieved": by_f / (ms_f / 1e3) / 1e9,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": by_f / (ms_f / 1e3) / 1e9 / HBM_PEAK_GBS,
                            "bytes_per_column": 8 * 6 * d["nlevsoi"]}},
           "water_table": {"ms_per_launch": ms_w,
                           "note": "per-column searches and layer-by-layer moves: the bytes read "
                                   "depend on the state (0.96 GB of HBM traffic per "
                                   "launch at 1M columns measured with FETCH_SIZE/"
                                   "WRITE_SIZE)"},
           "aquifer_step_ms": ms_f + ms + ms_w}
    if cpu_sample:
        sys.path.insert(0, str(REPO / "oracle"))
        import oracle
        m = min(cpu_sample, nc)
        sub = {k: (np.ascontiguousarray(v[..., :m]) if isinstance(v, np.ndarray) else v)
               for k, v in d.items()}
        sub["nc"] = m
        best = float("inf")
        for rep in range(6):
            s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sub.items()}
            t0 = time.perf_counter()
            for _ in range(2):
                oracle.soilwater_zengdecker2009(s, cfg.to_abi())
            if rep:                              # pass 0 is the warm-up
                best = min(best, time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": 2 * m / best, "unit": "column-steps/s",
                               "cores": oracle.num_threads(), "kind": "port",
                               "sample": f"first {m} columns x 2 steps, one warm-up pass then "
                                         f"best of 5 ({best:.2f} s)"}
    return out


def ghf_bytes(d):
    """Algorithmic bytes of one ground-heat-flux launch, each input read once and
    each output written once: per non-lake patch 14 float64 forcing values and
    its weight, 3 int64 indices, the absorbed solar of its snow layers and the
    top soil layer, and 3 float64 outputs (urban patches: 8 inputs, 5 landunit
    values, 8 outputs); per lake patch its mask and column index; per column 6
    float64 inputs, two layer temperatures, 4 int64 indices and a mask, and
    5 + nlevsno + 1 float64 outputs."""
    import numpy as np
    pc = d["patch_column"] - 1
    on = d["mask_nolakep"] == 1
    urb = on & (d["urbpoi"][d["patch_landunit"] - 1] == 1)
    rur = on & ~urb
    lyr = 1 - d["snl"][pc]
    b = (rur * (8 * (15 + 3) + 1 + 8 * lyr + 8 * 3) + urb * (8 * (8 + 5 + 3) + 1 + 8 * 8) +
         (~on) * (1 + 8))
    cols = d["nc"] * (8 * 6 + 8 * 2 + 8 * 4 + 1 + 8 * (5 + d["nlevsno"] + 1))
    return float(np.sum(b)) + cols


def seb_bytes(d):
    """Algorithmic bytes of one energy-balance launch: per column 4 int64 and a
    mask, 9 float64 inputs, t_soisno, t_ssbef and fact over its snow and ground
    layers, one output; per patch its index, masks, flux, weight, prior and
    result."""
    import numpy as np
    ns = -d["snl"]
    layers = ns + d["nlevgrnd"]
    cols = np.sum(8 * 4 + 1 + 8 * 9 + 3 * 8 * layers + 8)
    patches = d["np"] * (8 + 2 + 8 * 4)
    return float(cols + patches)


def timed_launches(fn, steps):
    """Average ms per launch of `fn`, HIP events around each launch."""
    import torch
    fn()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / steps


def surface_energy(nc, steps):
    """The ground heat flux (subgrid patches summed per column) and the soil
    energy balance error at `nc` columns, each launched `steps` times."""
    from clmsoil import api, inputs
    g = inputs.ground_heat_flux_inputs(nc=nc, seed=9)
    tg = api.to_device(g)
    ga = tg | {"nc": g["nc"], "np": g["np"], "nl": g["nl"]}
    ms_g = timed_launches(lambda: api.compute_ground_heat_flux_and_deriv(ga, 12, 25, 0), steps)
    assert api.device_errors() == 0
    bg = ghf_bytes(g)
    del tg, ga
    e = inputs.energy_balance_inputs(nc=nc, seed=10)
    te = api.to_device(e)
    ea = te | {"nc": e["nc"], "np": e["np"]}
    ms_e = timed_launches(lambda: api.soil_energy_balance(ea, 1800.0, 12, 25, e["nlevurb"]), steps)
    assert api.device_errors() == 0
    be = seb_bytes(e)

    def line(ms, by, what):
        return {"workload": what, "ms_per_launch": ms, "column_steps_per_s": nc / (ms / 1e3),
                "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                             "bytes_per_launch": by}}
    return {"ground_heat_flux": line(ms_g, bg, f"{nc} columns, {g['np']} patches "
                                                "(70% vegetated with 1-12 patches, 12% crop, "
                                                "10% urban, 8% lake), 0-12 snow layers"),
            "soil_energy_balance": line(ms_e, be, f"{nc} columns, {e['np']} patches (1-4 per "
                                                  "column), 0-12 snow layers + 25 ground layers")}


# clmsoil_surface_hydrology with the driver's settings: bytes each column
# reads (masks 3 x 1 B, 6 integer indices, 1 flag, 25 doubles: forcing,
# fractions, tables, 3 layers of icefrac and hksat, slope, threshold, h2osfc,
# gathered flood water) and writes (17 fluxes and states)
SURFHYD_READ_BYTES = 3 * 1 + 6 * 8 + 1 + 25 * 8
SURFHYD_WRITE_BYTES = 17 * 8


def surface_hydrology_step(nc, steps, cpu_sample=0):
    """The surface hydrology before the Richards step (flood water, saturated
    and infiltration excess runoff, surface water, infiltration, total runoff,
    urban ponding) in one launch on `nc` columns of mixed landunits (55 % soil,
    15 % crop, 15 % urban, 8 % lake, 7 % wetland); `steps` launches on the same
    inputs, each bracketed by HIP events on its stream."""
    import torch
    from clmsoil import api, inputs
    from clmsoil._abi import SH_FIELDS
    d = inputs.surface_hydrology_columns(nc=nc, seed=53)
    t = api.to_device(d)
    st = {k: t[k] for k in SH_FIELDS if k in t} | {"nc": nc}
    h0 = t["h2osfc"].clone()
    for _ in range(2):
        api.surface_hydrology(st, 1800.0)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for e in ev:
        t["h2osfc"].copy_(h0)
        e[0].record()
        api.surface_hydrology(st, 1800.0)
        e[1].record()
    torch.cuda.synchronize()
    ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    by = (SURFHYD_READ_BYTES + SURFHYD_WRITE_BYTES) * nc
    out = {"workload": f"{nc} columns, mixed soil/crop/urban/lake/wetland landunits",
           "ms_per_launch": ms, "columns_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_column": SURFHYD_READ_BYTES + SURFHYD_WRITE_BYTES}}
    if cpu_sample:
        import oracle
        out["cpu_baseline"] = _oracle_baseline(
            d, nc, cpu_sample, lambda s: oracle.surface_hydrology(s, 1800.0))
    return out


def root_sink_step(nc, steps, cpu_sample=0):
    """Root water uptake (default method) on `nc` columns of 0-12 patches each
    (CLM's patch layout); `steps` launches, each bracketed by HIP events."""
    import torch
    from clmsoil import api, inputs
    from clmsoil._abi import RS_FIELDS
    d = inputs.root_sink_columns(nc=nc, seed=54)
    t = api.to_device(d)
    st = {k: t[k] for k in RS_FIELDS} | {"nc": nc, "np": d["np"]}
    for _ in range(2):
        api.compute_effec_rootfrac_and_vert_tran_sink(st)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for e in ev:
        e[0].record()
        api.compute_effec_rootfrac_and_vert_tran_sink(st)
        e[1].record()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    n = d["nlevsoi"]
    # per column: mask, patchi, npatches, column transpiration in; rootr_col and
    # qflx_rootsoi_col out over the soil layers. Per patch: active, wtcol,
    # transpiration and the root profile over the soil layers in.
    by = nc * (1 + 3 * 8 + 2 * 8 * n) + d["np"] * (1 + 2 * 8 + 8 * n)
    out = {"workload": f"{nc} columns, {d['np']} patches (0-12 per column)",
           "ms_per_launch": ms, "columns_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_launch": by}}
    if cpu_sample:
        import oracle
        # the patch arrays stay whole: the first columns index their own patches
        out["cpu_baseline"] = _oracle_baseline(
            d, nc, cpu_sample, lambda s: oracle.compute_effec_rootfrac_and_vert_tran_sink(s))
    return out


def lake_solve_step(nc, steps, cpu_sample=0):
    """The lake temperature solve (assembly, tridiagonal solve and unpack in one
    launch) on `nc` lake columns: 0-6 snow layers, 10 lake levels, 25 soil
    levels; `steps` launches on the same inputs, each bracketed by HIP events."""
    import numpy as np
    import torch
    from clmsoil import api, inputs
    from clmsoil._abi import LAKE_FIELDS
    d = inputs.lake_columns(nc=nc, seed=55)
    t = api.to_device(d)
    st = {k: t[k] for k in LAKE_FIELDS} | {"nc": nc, "nlevlak": d["nlevlak"]}
    t0s, t0l = t["t_soisno"].clone(), t["t_lake"].clone()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps + 2)]
    for e in ev:
        t["t_soisno"].copy_(t0s)
        t["t_lake"].copy_(t0l)
        e[0].record()
        api.lake_temperature_solve(st, 1800.0)
        e[1].record()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = sum(e[0].elapsed_time(e[1]) for e in ev[2:]) / steps
    on = d["mask_lakec"] == 1
    ns = -d["snl"][on]
    # per lake column: snl, fin, phi_soil, tktopsoillay; per snow level z, cv, tk,
    # sabg_lyr_col and T in, T out; per soil level z, cv, tk and T in, T out; per
    # lake level z, dz, cv, tk, phi and T in, T out; plus the mask of every column
    by = nc + int(on.sum()) * (4 * 8 + 25 * 5 * 8 + 10 * 7 * 8) + int(ns.sum()) * 6 * 8
    out = {"workload": f"{nc} lake columns: 0-6 snow, 10 lake, 25 soil levels",
           "ms_per_launch": ms, "columns_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_launch": by}}
    if cpu_sample:
        import oracle
        out["cpu_baseline"] = _oracle_baseline(
            d, nc, cpu_sample, lambda s: oracle.lake_temperature_solve(s, 1800.0))
    return out


def _first_columns(d, nc, m):
    """The first m columns of a section's host inputs: arrays whose last axis is the
    column axis (and lists of them) are cut, everything else is kept."""
    import numpy as np

    def cut(v):
        if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[-1] == nc:
            return np.ascontiguousarray(v[..., :m])
        if isinstance(v, list) and v and isinstance(v[0], np.ndarray):
            return [cut(x) for x in v]
        return v
    out = {k: cut(v) for k, v in d.items()}
    out["nc"] = m
    if "ld" in out:
        out["ld"] = m
    return out


def _oracle_baseline(d, nc, m, run):
    """The oracle (one thread) timed on one step of the first m columns: one
    warm-up pass, then the best of 5."""
    import numpy as np
    sys.path.insert(0, str(REPO / "oracle"))
    sub = _first_columns(d, nc, min(m, nc))

    def fresh():
        return {k: (v.copy() if isinstance(v, np.ndarray)
                    else [x.copy() if isinstance(x, np.ndarray) else x for x in v]
                    if isinstance(v, list) else v)
                for k, v in sub.items()}
    best = float("inf")
    for rep in range(6):
        s = fresh()
        t0 = time.perf_counter()
        run(s)
        if rep:
            best = min(best, time.perf_counter() - t0)
    return {"value": sub["nc"] / best, "unit": "column-steps/s", "cores": 1, "kind": "port",
            "sample": f"first {sub['nc']} columns, one step, one warm-up pass then best of 5 "
                      f"({best:.2f} s)"}


def ch4_transport_step(nc, steps, cpu_sample=0):
    """Methane and oxygen transport (the unsaturated fraction of ch4!, 20 soil
    levels, water tables at every level) on `nc` columns; `steps` launches, each
    from the same state and bracketed by HIP events."""
    import torch
    from clmsoil import api, inputs
    from clmsoil._abi import CH4_FIELDS
    d = inputs.ch4_columns(nc=nc, seed=57)
    t = api.to_device({k: d[k] for k in CH4_FIELDS})
    st = t | {"nc": nc, "ldg": d["ldg"]}
    t0 = {k: v.clone() for k, v in t.items() if hasattr(v, "clone")}
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps + 2)]
    for e in ev:
        for k in t0:
            t[k].copy_(t0[k])
        e[0].record()
        api.ch4_tran(st, 1800.0, nlevsoi=d["nlevsoi"])
        e[1].record()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = sum(e[0].elapsed_time(e[1]) for e in ev[2:]) / steps
    on = int(d["mask"].sum())
    nl = d["nlevsoi"]
    # per filtered column and level: 8 forcing and 9 state arrays in, the 7 scaled
    # rates and stresses and the 2 concentrations out; per filtered column: the
    # gridcell and jwt indices, 4 surface values and the ground conductance in, c_atm
    # (CH4 and O2), 5 column values out; plus the mask
    by = nc + on * (nl * 8 * (17 + 9) + 8 * (2 + 5) + 16 + 8 * 5)
    out = {"workload": f"{nc} columns x {nl} levels, unsaturated fraction, jwt 0-{nl}",
           "ms_per_launch": ms, "column_steps_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_launch": by}}
    if cpu_sample:
        import oracle
        out["cpu_baseline"] = _oracle_baseline(
            d, nc, cpu_sample, lambda s: oracle.ch4_tran(s, 1800.0, nlevsoi=nl))
    return out


def litter_transport_step(nc, steps, cpu_sample=0):
    """Vertical transport of the soil organic matter pools (the CENTURY cascade's
    7 pools of carbon and nitrogen, 20 decomposition levels) on `nc` columns;
    `steps` launches on the same inputs, each bracketed by HIP events."""
    import torch
    from clmsoil import api, inputs
    from clmsoil._abi import LVT_FIELDS
    d = inputs.lvt_columns(nc=nc, seed=56)
    t = api.to_device({k: d[k] for k in LVT_FIELDS})
    conc = [torch.from_numpy(x).cuda() for x in d["conc"]]
    conc0 = [x.clone() for x in conc]
    src = [torch.from_numpy(x).cuda() for x in d["source"]]
    tend = [torch.from_numpy(x).cuda() for x in d["tendency"]]
    st = t | {"nc": nc, "conc": conc, "source": src, "tendency": tend}
    kw = dict(nlevdecomp=d["nlevdecomp"], nlevdecomp_full=d["nlevdecomp_full"],
              npools=d["npools"], ntypes=d["ntypes"], is_cwd=d["is_cwd"])
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps + 2)]
    for e in ev:
        for x, x0 in zip(conc, conc0):
            x.copy_(x0)
        e[0].record()
        api.litter_vert_transp(st, 1800.0, **kw)
        e[1].record()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = sum(e[0].elapsed_time(e[1]) for e in ev[2:]) / steps
    on = int(d["mask_bgc_soilc"].sum())
    nd, npools, nt = d["nlevdecomp"], d["npools"], d["ntypes"]
    # per filtered column: nbedrock, altmax twice, the two coefficient rows written
    # and read back once; per (tracer, pool, level): concentration and source in,
    # concentration and tendency out (coarse woody debris: no tendency); plus the mask
    ncwd = int(sum(bool(x) for x in d["is_cwd"]))
    by = (nc + on * (3 * 8 + 2 * 2 * 8 * (nd + 1))
          + on * nt * nd * 8 * (4 * (npools - ncwd) + 3 * ncwd))
    out = {"workload": f"{nc} columns x {npools} pools x {nt} tracers x {nd} levels",
           "ms_per_launch": ms, "column_steps_per_s": nc / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": by / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_launch": by}}
    if cpu_sample:
        import oracle
        out["cpu_baseline"] = _oracle_baseline(
            d, nc, cpu_sample, lambda s: oracle.litter_vert_transp(s, 1800.0, **kw))
    return out


def water_tables_moisture_form(nc, steps, cpu_sample=0):
    """The water tables after a step without an aquifer layer (perched,
    theta-based, bedrock clamp in one launch) on `nc` columns of 20 soil layers
    with frozen layers under unfrozen ones on 30 % of the columns and the 90 %
    saturation level anywhere in the column; `steps` launches on the same state,
    each bracketed by HIP events on its stream."""
    import torch
    from clmsoil import api, inputs
    d = inputs.moisture_water_table_columns(nc=nc, seed=52)
    t = api.to_device(d)
    st = {k: t[k] for k in ("nbedrock", "dz", "z", "zi", "t_soisno", "watsat", "h2osoi_liq",
                            "h2osoi_ice", "zwt", "zwt_perched", "frost_table", "h2osoi_vol")}
    st["nc"] = nc
    zwt0 = t["zwt"].clone()
    for _ in range(2):
        t["zwt"].copy_(zwt0)
        api.water_table_moisture_form(st)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for e in ev:
        t["zwt"].copy_(zwt0)
        e[0].record()
        api.water_table_moisture_form(st)
        e[1].record()
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    out = {"workload": f"{nc} columns, 20 soil layers, nbedrock 3-20, 30% frozen columns",
           "ms_per_launch": ms, "columns_per_s": nc / (ms / 1e3),
           "note": "per-column searches up and down the saturation and temperature "
                   "profiles: the bytes read depend on the state"}
    if cpu_sample:
        import oracle
        out["cpu_baseline"] = _oracle_baseline(d, nc, cpu_sample, oracle.water_table_moisture_form)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--nc", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain first-substep comparison")
    ap.add_argument("--no-calibration", action="store_true",
                    help="skip the calibration-gradient section (262,144 columns x 30 days)")
    ap.add_argument("--no-zero-flux", action="store_true",
                    help="skip the production day with the BC_ZERO_FLUX lower boundary")
    ap.add_argument("--no-zd", action="store_true",
                    help="skip the Zeng-Decker 2009 step section")
    ap.add_argument("--no-surface", action="store_true",
                    help="skip the ground heat flux / energy balance section")
    ap.add_argument("--surface-only", action="store_true",
                    help="run only the ground heat flux / energy balance section")
    ap.add_argument("--zd-only", action="store_true",
                    help="run only the Zeng-Decker 2009 section (for profiling its kernel)")
    ap.add_argument("--hydro-only", action="store_true",
                    help="run only the soil hydrology sections (aquifer step, water tables "
                         "without an aquifer, surface hydrology, root uptake, lake solve, "
                         "litter transport, methane transport)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 4 times (the known-size kernel "
                         "that scales FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the per-column solves need no exchange between "
              f"GPUs, so --gpus {args.gpus} is not run (columns shard across GPUs unchanged).",
              file=sys.stderr)
        return 2
    import numpy as np
    import torch
    from clmsoil import api

    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8, "copies": 4}), flush=True)
        return 0
    nc = args.nc
    # the oracle is timed beside each hydrology section on its first 65,536 columns
    sec_cpu = 0 if args.no_cpu_baseline else 65536
    if args.surface_only:
        print(json.dumps({"surface_energy": surface_energy(nc, args.steps)}), flush=True)
        return 0
    if args.zd_only:
        print(json.dumps({"zengdecker_2009": zengdecker_2009(
            nc, args.steps, 0 if args.no_cpu_baseline else args.cpu_sample)}), flush=True)
        return 0
    if args.hydro_only:
        print(json.dumps({"zengdecker_2009": zengdecker_2009(
            nc, args.steps, 0 if args.no_cpu_baseline else args.cpu_sample),
            "water_tables_moisture_form": water_tables_moisture_form(nc, args.steps, sec_cpu),
            "surface_hydrology": surface_hydrology_step(nc, args.steps, sec_cpu),
            "root_sink": root_sink_step(nc, args.steps, sec_cpu),
            "lake_solve": lake_solve_step(nc, args.steps, sec_cpu),
            "litter_transport": litter_transport_step(nc, args.steps, sec_cpu),
            "ch4_transport": ch4_transport_step(nc, args.steps, sec_cpu)}), flush=True)
        return 0
    host, dev, gen_s = make_state(nc, args.seed)
    print(f"# state ready: {nc} columns in {gen_s:.1f} s", file=sys.stderr, flush=True)
    from clmsoil import day, tuning
    tuned = tuning.lookup(nc)
    tuning.apply(tuned)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    for _ in range(args.warmup):
        gpu_step(dev, nc, cfg, HEAT_KEYS)
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    mutable = ("t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
               "h2osno_no_layers", "snow_depth", "int_snow")
    snap = {k: dev[k].clone() for k in mutable}
    # timed region 1: the day as one graph launch (the production configuration)
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"])
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"],
                  dev["smpmin"])
    plan = day.DayPlan({k: dev[k] for k in HEAT_KEYS} | {"nc": nc}, st, mi, None, dev["nbedrock"],
                       12, 20, 25, 1800.0, cfg, args.steps)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    plan.launch()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(end)
    plan.close()
    assert api.device_errors() == 0
    # timed region 2: the same steps from the same state, launched one by one with
    # HIP events around every kernel on the launch stream (per-kernel durations
    # for the roofline)
    for k in mutable:
        dev[k].copy_(snap[k])
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for k in range(args.steps):
        gpu_step(dev, nc, cfg, HEAT_KEYS, ev[k])
    e1.record()
    torch.cuda.synchronize()
    eager_ms = e0.elapsed_time(e1)
    heat_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps
    rich_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps
    assert api.device_errors() == 0
    value = nc * args.steps / (total_ms / 1e3)
    nsub = dev["nsubsteps"].float().mean().item()
    b_r, b_h = algorithmic_bytes(host["nbedrock"], host["snl"])
    kern = {"richards": (rich_ms, b_r), "heat": (heat_ms, b_h)}
    dom = max(kern, key=lambda k: kern[k][0])
    ms, by = kern[dom]
    achieved = by / (ms / 1e3) / 1e9
    # the counter figures belong to this workload and to the profiled kernel sources
    fresh = kernel_sources_hash() == PROFILED_SOURCES
    profiled = fresh and nc == 1 << 20 and args.seed == 2
    out = {
        "metric": METRIC, "value": value, "unit": "column-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": total_ms / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "production day: 1,048,576 uneven columns (0-12 snow layers, "
                               "nbedrock 5-20, zwt 0.5-10 m), heat step + moisture-form Richards "
                               "(BC_WATERTABLE) per 1800 s step" if nc == 1 << 20 else
                   f"production day, {nc} columns",
                   "columns": nc, "layers": "12 snow + 25 ground (20 soil)",
                   "step_seconds": 1800.0, "launch": "one HIP graph per day of steps",
                   "tuning": ({"richards": tuned["richards"], "heat": tuned["heat"]} if tuned
                              else "defaults")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC_1M[dom] if profiled else None,
                     "traffic_profiled_at": PROFILED_AT, "profile_current": fresh},
        # the same kernel against the FP64 vector pipe (FLOP from its counters)
        "fp64_view": ({"kernel": dom, "achieved": FP64_FLOP_1M[dom] / (ms / 1e3) / 1e12,
                       "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": FP64_FLOP_1M[dom] / (ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS,
                       "profiled_at": PROFILED_AT}
                      if profiled else None),
        "kernels_ms": {"heat": heat_ms, "richards": rich_ms},
        "eager_column_steps_per_s": nc * args.steps / (eager_ms / 1e3),
        "fused_step_hbm_frac": (b_r + b_h) / (total_ms / args.steps / 1e3) / 1e9 / HBM_PEAK_GBS,
        "mean_substep_iterations": nsub,
        "wall_s": wall,
    }
    if not args.no_zero_flux:
        out["zero_flux_day"] = zero_flux_day(dev, snap, nc, args.steps)
    if not args.no_learned:
        out["learned_production_day"] = learned_production_day(dev, snap, nc, args.steps)
        out["learned_vs_plain"] = learned_vs_plain(nc, args.steps)
    if not args.no_calibration and nc == 1 << 20:
        out["calibration"] = calibration_section()
    if not args.no_surface:
        out["surface_energy"] = surface_energy(nc, args.steps)
    if not args.no_zd:
        out["zengdecker_2009"] = zengdecker_2009(
            nc, args.steps, 0 if args.no_cpu_baseline else args.cpu_sample)
        out["water_tables_moisture_form"] = water_tables_moisture_form(nc, args.steps, sec_cpu)
        out["surface_hydrology"] = surface_hydrology_step(nc, args.steps, sec_cpu)
        out["root_sink"] = root_sink_step(nc, args.steps, sec_cpu)
        out["lake_solve"] = lake_solve_step(nc, args.steps, sec_cpu)
        out["litter_transport"] = litter_transport_step(nc, args.steps, sec_cpu)
        out["ch4_transport"] = ch4_transport_step(nc, args.steps, sec_cpu)
    if not args.no_cpu_baseline:
        v, cores, secs = cpu_baseline(host, min(args.cpu_sample, nc), args.cpu_steps)
        ncpu, model = host_cpus()
        out["cpu_baseline"] = {"value": v, "unit": "column-steps/s", "cores": cores,
                               "kind": "port", "host_cpus": ncpu, "cpu_model": model,
                               "sample": f"first {min(args.cpu_sample, nc)} columns x {args.cpu_steps} "
                                         f"steps of the same workload (heat + Richards, BC_WATERTABLE), "
                                         f"oracle -O3 -march=x86-64-v3 on {cores} OpenMP threads, one "
                                         f"warm-up pass then best of 5 ({secs:.2f} s)"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
