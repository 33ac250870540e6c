"""A day of production steps captured as one HIP graph (include/clmsoil_day.h).

CLM.jl launches every kernel of a step separately and synchronises after each
(src/infrastructure/kernels.jl:85-91). A DayPlan records `nsteps` steps on
fixed arrays once and replays them with one graph launch. Each step is the heat
step, then the Richards step with its aquifer pass. With a predictor, the
learned network sets each column's first substep inside the Richards kernel
(fused=True), or in a launch of its own before it (fused=False; same results).
"""
from __future__ import annotations

import ctypes as C

from . import _abi
from .api import (CLMSoilError, MfIn, MfState, SoilWaterMovementConfig, _mf_dict, _stream,
                  MOISTURE_FORM)
from .bind import byref, heat_args, heat_config, levels, mf_args


class DayPlan:
    def __init__(self, heat_state: dict, st: MfState, mfin: MfIn, mask, nbedrock, nlevsno: int,
                 nlevsoi: int, nlevgrnd: int, dtime: float, cfg: SoilWaterMovementConfig,
                 nsteps: int, predictor=None, snow_thermal_cond_method: str = "Jordan1991",
                 fused: bool = True):
        import torch
        code = {"Jordan1991": 1, "Sturm1997": 2}.get(snow_thermal_cond_method)
        if code is None:
            raise CLMSoilError(f"Unknown snow_thermal_cond_method: {snow_thermal_cond_method}")
        if mfin.dtsub0 is not None and predictor is not None:
            raise CLMSoilError("give either a predictor or MfIn.dtsub0, not both")
        hd = dict(heat_state)
        hd.setdefault("nc", hd["snl"].shape[0])
        hd.update(nlevsno=nlevsno, nlevgrnd=nlevgrnd, nlevsoi=nlevsoi)
        md = _mf_dict(st, mfin, mask, nbedrock, nlevsno, nlevsoi, nlevgrnd)
        self.dtsub0 = None
        if predictor is not None and not fused:
            self.dtsub0 = torch.full((int(md["nc"]),), float(dtime), dtype=torch.float64,
                                     device=nbedrock.device)
        c = cfg.to_abi()
        c.soilwater_movement_method = MOISTURE_FORM
        self._keep = (hd, md, predictor)       # the graph holds raw device pointers
        L = _abi.load_library()
        self._L = L
        ha, ma, lev, hc = heat_args(hd), mf_args(md), levels(md), heat_config(code)
        p = C.c_void_p()
        pp = (C.c_void_p(predictor.params.data_ptr()) if predictor else None,
              C.c_void_p(predictor.norm.data_ptr()) if predictor else None,
              C.c_void_p(self.dtsub0.data_ptr()) if self.dtsub0 is not None else None)
        _abi.check(L, L.clmsoil_day_plan_create(byref(ha), byref(ma), byref(lev), float(dtime),
                                                byref(hc), byref(c), *pp, int(nsteps), C.byref(p)),
                   "day_plan_create")
        self._plan = p
        self.nsteps = nsteps

    def launch(self) -> None:
        _abi.check(self._L, self._L.clmsoil_day_plan_launch(self._plan, _stream()), "day_plan_launch")

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            self._L.clmsoil_day_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
