// day.cpp — capture a day of production steps as one HIP graph (clmsoil_day.h).
//
// The capture runs on a private non-blocking stream (default priority), so
// callers may pass any stream, including the null stream, to the launch. The
// plan owns the Richards workspace its graph uses (allocated when the plan is
// created), so replays share nothing with eager calls or other plans; every
// step clears its work counters with a small kernel, never a memset node, so
// the graph replays correctly every time.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "../../include/clmsoil.h"
#include "../../include/clmsoil_day.h"
#include "../../include/clmsoil_learned.h"
#include "heat.hpp"
#include "richards.hpp"
#include "state.hpp"

namespace clmsoil {
int validate_mf(const clmsoil_mf_args *a, const clmsoil_levels *lev, double dtime,
                const clmsoil_swm_config *cfg);
int validate_heat(const clmsoil_heat_args *a, const clmsoil_levels *lev);
RichardsParams richards_params(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                               double dtime, const clmsoil_swm_config *cfg, bool aquifer,
                               unsigned *err);
int run_richards(const RichardsParams &P, hipStream_t stream);
int run_richards_learned(const RichardsParams &P, hipStream_t stream);
int run_heat(const clmsoil_heat_args *a, const clmsoil_levels *lev, double dtime,
             const clmsoil_heat_config *cfg, hipStream_t stream, unsigned *err);
}  // namespace clmsoil

using namespace clmsoil;

struct clmsoil_day_plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *workspace = nullptr;   // the Richards steps' workspace (plan-owned)
    int nsteps = 0;
};

static void free_plan(clmsoil_day_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->workspace) (void)hipFree(p->workspace);
    delete p;
}

extern "C" {

int clmsoil_day_plan_create(const clmsoil_heat_args *heat, const clmsoil_mf_args *mf,
                            const clmsoil_levels *lev, double dtime,
                            const clmsoil_heat_config *heat_cfg,
                            const clmsoil_swm_config *swm_cfg, const float *mlp_params,
                            const float *mlp_norm, double *dtsub0, int32_t nsteps,
                            clmsoil_day_plan **plan) {
    if (!plan) return CLMSOIL_ERR_NULL;
    *plan = nullptr;
    if (nsteps < 1 || nsteps > 100000) return CLMSOIL_ERR_ARG;
    int st = validate_heat(heat, lev);
    if (st) return st;
    if ((st = validate_mf(mf, lev, dtime, swm_cfg))) return st;
    if (heat->nc != mf->nc) return CLMSOIL_ERR_SIZE;
    if (heat_cfg && heat_cfg->snow_thermal_cond_method != 1 &&
        heat_cfg->snow_thermal_cond_method != 2)
        return CLMSOIL_ERR_ARG;
    const bool learned = mlp_params || mlp_norm || dtsub0;
    if (learned && (!mlp_params || !mlp_norm)) return CLMSOIL_ERR_NOT_TRAINED;
    if (learned && mf->dtsub0) return CLMSOIL_ERR_ARG;
    // the network inside the Richards kernel, unless the caller wants the
    // predictions in dtsub0 (a separate launch before each step)
    const bool fused = learned && !dtsub0;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    auto *p = new (std::nothrow) clmsoil_day_plan();
    if (!p) return CLMSOIL_ERR_HIP;
    p->nsteps = nsteps;
    if (hipMalloc(&p->workspace, RICHARDS_WORKSPACE_BYTES) != hipSuccess) {
        p->workspace = nullptr;
        free_plan(p);
        return CLMSOIL_ERR_HIP;
    }
    clmsoil_mf_args m = *mf;
    if (learned && !fused) m.dtsub0 = dtsub0;
    m.workspace = p->workspace;
    m.workspace_bytes = RICHARDS_WORKSPACE_BYTES;
    const int bc = swm_cfg->lower_boundary_condition;
    const bool aquifer = bc == CLMSOIL_BC_WATERTABLE || bc == CLMSOIL_BC_AQUIFER;
    RichardsParams P = richards_params(&m, lev, dtime, swm_cfg, aquifer, s->err);
    if (fused) {
        P.mlp = mlp_params;
        P.mlp_norm = mlp_norm;
        P.refill_sixteenths = g_tuning.richards_refill_learned.load(std::memory_order_relaxed);
    }

    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        free_plan(p);
        return CLMSOIL_ERR_HIP;
    }
    st = CLMSOIL_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        st = CLMSOIL_ERR_HIP;
    } else {
        for (int k = 0; k < nsteps && st == CLMSOIL_OK; ++k) {
            st = run_heat(heat, lev, dtime, heat_cfg, cs, s->err);
            if (st == CLMSOIL_OK && learned && !fused)
                st = clmsoil_mlp_predict_dtsub0(&m, lev, dtime, mlp_params, mlp_norm, dtsub0, cs);
            if (st == CLMSOIL_OK) st = fused ? run_richards_learned(P, cs) : run_richards(P, cs);
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (e != hipSuccess && st == CLMSOIL_OK) st = CLMSOIL_ERR_HIP;
        p->graph = g;
    }
    if (st == CLMSOIL_OK && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
        st = CLMSOIL_ERR_HIP;
    (void)hipStreamDestroy(cs);
    if (st != CLMSOIL_OK) {
        free_plan(p);
        return st;
    }
    *plan = p;
    return CLMSOIL_OK;
}

int clmsoil_day_plan_launch(clmsoil_day_plan *plan, void *stream) {
    if (!plan || !plan->exec) return CLMSOIL_ERR_NULL;
    return hipGraphLaunch(plan->exec, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK
                                                                          : CLMSOIL_ERR_HIP;
}

int clmsoil_day_plan_destroy(clmsoil_day_plan *plan) {
    if (!plan) return CLMSOIL_OK;
    // the graph may still run on the caller's stream: wait before freeing
    if (hipDeviceSynchronize() != hipSuccess) return CLMSOIL_ERR_HIP;
    free_plan(plan);
    return CLMSOIL_OK;
}

}  // extern "C"
