"""Autotuner for the launch shapes of the two fused kernels.

Searched, on the workload itself (same columns, same state):
  * Richards: lanes per column G in {1, 2, 4} (64/G columns per wavefront),
    workgroup size in {64, 128, 256}, schedule (one-shot grid or resident grid
    with column refill) and, for the resident grid, blocks per CU (0 = occupancy
    calculator, or 1..4); then, for the best resident-grid shape, the refill
    threshold (a wavefront refills its idle lane groups once n/16 of them are
    idle: fewer refill passes against more idle lanes) and, at G = 4 with
    256-thread blocks, the LDS tile of the column: how many per-layer fields live
    in the lane-private LDS cache (0, 5 or 7: 0, 50 or 70 KB per block) rather
    than in registers or memory;
  * heat: G in {1, 2, 4} and workgroup size in {64, 128, 256}. The per-wavefront
    LDS tile of the pivoting fallback (38 x 7 band entries + right-hand side)
    comes with the workgroup size: 64 / 128 / 256 threads hold 1 / 2 / 4 tiles.
Every candidate is timed with HIP events on the launch stream after resetting
the state to the same starting point, best of `reps`. The winner is stored
per GPU architecture and workload size in tuning_table.json next to this file,
and apply() installs it with clmsoil_set_tuning & co. All candidates give
identical results; only speed differs.
"""
from __future__ import annotations

import json
from pathlib import Path

from . import api, inputs

TABLE = Path(__file__).resolve().parent / "tuning_table.json"
GS = (1, 2, 4)
BLOCKS = (64, 128, 256)
RESIDENCY = (0, 1, 2, 3, 4)
REFILLS = (4, 6, 7, 8, 9, 10, 12)
DEFAULT_REFILL = 7
LDS_FIELDS = (0, 5, 7)
DEFAULT_LDS_FIELDS = 7

_RICHARDS_STATE = ("h2osoi_liq",)
_HEAT_STATE = ("t_soisno", "h2osoi_liq", "h2osoi_ice", "excess_ice", "t_h2osfc", "h2osfc",
               "h2osno_no_layers", "snow_depth", "int_snow")


def device_key() -> str:
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or p.name
    return arch.split(":")[0]


def production_state(nc: int, seed: int = 2):
    """Device state of the production-day workload (heat + Richards inputs)."""
    d = inputs.production_day(nc=nc, seed=seed, fill=0.0)
    d.update(inputs.heat_outputs(d, diagnostics=False, fill=0.0))
    return api.to_device(d)


HEAT_KEYS = ["snl", "nbedrock", "z", "dz", "zi", "t_soisno", "h2osoi_liq", "h2osoi_ice",
             "excess_ice", "watsat", "tkmg", "tkdry", "csol", "bsw", "sucsat", "frac_sno_eff",
             "frac_h2osfc", "hs_top_snow", "hs_soil", "hs_top", "hs_h2osfc", "dhsdT",
             "sabg_lyr_col", "eflx_bot", "t_h2osfc", "h2osfc", "h2osno_no_layers", "snow_depth",
             "int_snow", "t_grnd", "xmf", "xmf_h2osfc", "qflx_snomelt", "qflx_snofrz",
             "qflx_snow_drain", "qflx_h2osfc_to_ice", "eflx_fgr12", "qflx_snomelt_lyr",
             "qflx_snofrz_lyr"]


def heat_step(dev, dtime=1800.0):
    api.soil_temperature({k: dev[k] for k in HEAT_KEYS} | {"nc": dev["nc"]}, dtime)


def richards_step(dev, cfg, dtime=1800.0, dtsub0=None):
    st = api.MfState(dev["h2osoi_liq"], dev["smp_l"], dev["hk_l"], dev["qcharge"], dev["nsubsteps"])
    mi = api.MfIn(dev["z"], dev["zi"], dev["dz"], dev["qflx_rootsoi"], dev["watsat"], dev["hksat"],
                  dev["bsw"], dev["sucsat"], dev["icefrac"], dev["zwt"], dev["qflx_infl"],
                  dev["smpmin"], dtsub0)
    api.soilwater_moisture_form(st, mi, None, dev["nbedrock"], dev["nlevsno"], dev["nlevsoi"],
                                dev["nlevgrnd"], dtime, cfg)


def _time(fn, reset, reps: int) -> float:
    import torch
    best = float("inf")
    for _ in range(reps):
        reset()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def autotune(nc: int = 1 << 20, seed: int = 2, reps: int = 3, warm_steps: int = 2,
             log=None) -> dict:
    """Tune on the production day after `warm_steps` steps (a typical mid-day
    state) and return the table entry."""
    import torch
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_WATERTABLE)
    dev = production_state(nc, seed)
    for _ in range(warm_steps):
        heat_step(dev)
        richards_step(dev, cfg)
    torch.cuda.synchronize()
    snap = {k: dev[k].clone() for k in set(_HEAT_STATE) | set(_RICHARDS_STATE)}

    def reset(keys):
        def f():
            for k in keys:
                dev[k].copy_(snap[k])
        return f

    results = {"richards": [], "heat": []}
    g0 = api.get_tuning()
    for g in GS:
        for blk in BLOCKS:
            for pers in (False, True):
                for res in (RESIDENCY if pers else (0,)):
                    api.set_tuning(g, blk, g0[2], g0[3])
                    api.set_richards_schedule(pers)
                    api.set_richards_residency(res)
                    ms = _time(lambda: richards_step(dev, cfg), reset(_RICHARDS_STATE), reps)
                    results["richards"].append({"g": g, "block": blk, "persistent": pers,
                                                "blocks_per_cu": res, "ms": ms})
                    if log:
                        log(results["richards"][-1])
    rbest = min(results["richards"], key=lambda r: r["ms"])
    results["refill"] = []
    if rbest["persistent"]:
        api.set_tuning(rbest["g"], rbest["block"], g0[2], g0[3])
        api.set_richards_schedule(True)
        api.set_richards_residency(rbest["blocks_per_cu"])
        for n in REFILLS:
            api.set_richards_refill(n, 7)
            ms = _time(lambda: richards_step(dev, cfg), reset(_RICHARDS_STATE), reps)
            results["refill"].append({"refill": n, "ms": ms})
            if log:
                log(results["refill"][-1])
        best_refill = min(results["refill"], key=lambda r: r["ms"])
        rbest = dict(rbest, refill=best_refill["refill"], ms=best_refill["ms"])
    else:
        rbest = dict(rbest, refill=DEFAULT_REFILL)
    results["lds_fields"] = []
    if rbest["g"] == 4 and rbest["block"] == 256:
        api.set_richards_refill(rbest["refill"], 7)
        for f in LDS_FIELDS:
            api.set_richards_lds_cache(f)
            ms = _time(lambda: richards_step(dev, cfg), reset(_RICHARDS_STATE), reps)
            results["lds_fields"].append({"lds_fields": f, "ms": ms})
            if log:
                log(results["lds_fields"][-1])
        best_f = min(results["lds_fields"], key=lambda r: r["ms"])
        rbest = dict(rbest, lds_fields=best_f["lds_fields"], ms=best_f["ms"])
        api.set_richards_lds_cache(best_f["lds_fields"])
    else:
        rbest = dict(rbest, lds_fields=DEFAULT_LDS_FIELDS)
    for g in GS:
        for blk in BLOCKS:
            api.set_tuning(rbest["g"], rbest["block"], g, blk)
            ms = _time(lambda: heat_step(dev), reset(_HEAT_STATE), reps)
            results["heat"].append({"g": g, "block": blk, "ms": ms})
            if log:
                log(results["heat"][-1])
    hbest = min(results["heat"], key=lambda r: r["ms"])
    assert api.device_errors() == 0
    entry = {"device": device_key(), "workload": "production_day", "columns": nc,
             "richards": {k: rbest[k] for k in ("g", "block", "persistent", "blocks_per_cu",
                                                 "refill", "lds_fields")},
             "heat": {k: hbest[k] for k in ("g", "block")},
             "richards_ms": rbest["ms"], "heat_ms": hbest["ms"], "candidates": results}
    apply(entry)
    return entry


def load_table(path: Path = TABLE) -> list:
    if not Path(path).exists():
        return []
    return json.loads(Path(path).read_text())


def save_entry(entry: dict, path: Path = TABLE) -> None:
    table = [e for e in load_table(path)
             if not (e["device"] == entry["device"] and e["workload"] == entry["workload"]
                     and e["columns"] == entry["columns"])]
    table.append({k: v for k, v in entry.items() if k != "candidates"})
    Path(path).write_text(json.dumps(table, indent=1) + "\n")


def lookup(columns: int, workload: str = "production_day", device: str | None = None):
    """The table entry for this device and workload whose size is nearest."""
    device = device or device_key()
    cands = [e for e in load_table() if e["device"] == device and e["workload"] == workload]
    if not cands:
        return None
    return min(cands, key=lambda e: abs(e["columns"] - columns))


def apply(entry: dict | None) -> bool:
    """Install an entry's launch shapes; False (defaults kept) when there is none."""
    if entry is None:
        return False
    r, h = entry["richards"], entry["heat"]
    api.set_tuning(r["g"], r["block"], h["g"], h["block"])
    api.set_richards_schedule(bool(r["persistent"]))
    api.set_richards_residency(r["blocks_per_cu"])
    api.set_richards_refill(int(r.get("refill", DEFAULT_REFILL)), 7)
    api.set_richards_lds_cache(int(r.get("lds_fields", DEFAULT_LDS_FIELDS)))
    return True


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="Autotune the kernels' launch shapes on this GPU.")
    ap.add_argument("--nc", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--save", action="store_true", help="write the result to tuning_table.json")
    args = ap.parse_args(argv)
    entry = autotune(args.nc, reps=args.reps, log=lambda r: print(json.dumps(r), flush=True))
    print(json.dumps({k: v for k, v in entry.items() if k != "candidates"}), flush=True)
    if args.save:
        save_entry(entry)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
