// capi.cpp — C ABI entry points of libclmsoil (see include/clmsoil.h).
//
// Host-side validation mirrors the reference's host checks
// (soilwm_moisture_form_solve!, src/biogeophys/soil_water_movement.jl:1500-1505);
// everything else is asynchronous on the caller's stream. Conditions the
// reference only meets inside its kernels (nbedrock or snl out of range) are
// recorded in a per-device error word that clmsoil_device_errors() reads back.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <initializer_list>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/clmsoil.h"
#include "../../include/clmsoil_ext.h"
#include "../../include/clmsoil_adjoint.h"
#include "../../include/clmsoil_learned.h"
#include "../../include/clmsoil_zd09.h"
#include "../../include/clmsoil_ghf.h"
#include "../../include/clmsoil_hydro.h"
#include "../../include/clmsoil_surfhyd.h"
#include "../../include/clmsoil_lake.h"
#include "../../include/clmsoil_lvt.h"
#include "common.hpp"
#include "heat.hpp"
#include "richards.hpp"
#include "solvers.hpp"
#include "state.hpp"
#include "zd09.hpp"
#include "ghf.hpp"
#include "hydro.hpp"
#include "surfhyd.hpp"
#include "lake.hpp"
#include "lvt.hpp"
#include "ch4.hpp"

using namespace clmsoil;

namespace clmsoil {

Tuning g_tuning;   // set by clmsoil_set_tuning (the autotuner), read at launch

// Scratch of clmsoil_soilwm_moisture_form_solve_static, the entry point for
// callers that cannot allocate: one Richards workspace (the call's error word
// and the per-XCD work counters) per GPU, held until the process ends. Every
// call on a GPU uses the same array, so calls that overlap on one GPU race.
// Host code must not pass this name to a kernel (it is a host-side stand-in):
// clmsoil_init() resolves each device's copy with hipGetSymbolAddress.
static __device__ __align__(256) unsigned scratch[RICHARDS_WORKSPACE_BYTES / sizeof(unsigned)];

namespace {
std::mutex g_mu;                     // serialises clmsoil_init
std::vector<DeviceState> g_dev;      // one entry per device, filled by clmsoil_init
std::atomic<bool> g_ready{false};    // g_dev is complete and no longer changes
}  // namespace

int init_devices() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ready.load(std::memory_order_acquire)) return CLMSOIL_OK;
    int n = 0;
    const hipError_t ec = hipGetDeviceCount(&n);
    if (ec == hipErrorNoDevice) n = 0;
    else if (ec != hipSuccess) return CLMSOIL_ERR_HIP;
    int cur = 0;
    if (n > 0 && hipGetDevice(&cur) != hipSuccess) return CLMSOIL_ERR_HIP;
    std::vector<DeviceState> dev((size_t)n);
    int st = CLMSOIL_OK;
    for (int d = 0; d < n && st == CLMSOIL_OK; ++d) {
        void *p = nullptr;
        if (hipSetDevice(d) != hipSuccess || hipMalloc(&p, 256) != hipSuccess ||
            hipMemset(p, 0, 256) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            st = CLMSOIL_ERR_HIP;
            if (p) (void)hipFree(p);
            break;
        }
        dev[(size_t)d].err = static_cast<unsigned *>(p);
        void *sp = nullptr;
        if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
            st = CLMSOIL_ERR_HIP;
            break;
        }
        dev[(size_t)d].scratch = static_cast<unsigned *>(sp);
    }
    if (n > 0 && hipSetDevice(cur) != hipSuccess) st = CLMSOIL_ERR_HIP;
    if (st != CLMSOIL_OK) {
        for (int d = 0; d < n; ++d)
            if (dev[(size_t)d].err && hipSetDevice(d) == hipSuccess) (void)hipFree(dev[(size_t)d].err);
        if (n > 0) (void)hipSetDevice(cur);
        return st;
    }
    g_dev.swap(dev);
    g_ready.store(true, std::memory_order_release);
    return CLMSOIL_OK;
}

// The state clmsoil_init() set up for the current device; allocates nothing.
int current_device_state(DeviceState **out) {
    if (!g_ready.load(std::memory_order_acquire)) return CLMSOIL_ERR_INIT;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CLMSOIL_ERR_HIP;
    if (dev < 0 || (size_t)dev >= g_dev.size() || !g_dev[(size_t)dev].err) return CLMSOIL_ERR_INIT;
    *out = &g_dev[(size_t)dev];
    return CLMSOIL_OK;
}

int validate_levels(const clmsoil_levels *lev) {
    if (!lev) return CLMSOIL_ERR_NULL;
    if (lev->nlevsno < 1 || lev->nlevsno > MAX_NLEVSNO) return CLMSOIL_ERR_SIZE;
    if (lev->nlevsoi < 2 || lev->nlevsoi > MAX_NLEVSOI) return CLMSOIL_ERR_SIZE;
    if (lev->nlevgrnd < lev->nlevsoi || lev->nlevgrnd > MAX_NLEVGRND) return CLMSOIL_ERR_SIZE;
    return CLMSOIL_OK;
}

int validate_mf(const clmsoil_mf_args *a, const clmsoil_levels *lev, double dtime,
                const clmsoil_swm_config *cfg) {
    if (!a || !cfg) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (cfg->soilwater_movement_method != CLMSOIL_MOISTURE_FORM) return CLMSOIL_ERR_METHOD;
    if (cfg->upper_boundary_condition != CLMSOIL_BC_FLUX) return CLMSOIL_ERR_BC;
    const int bc = cfg->lower_boundary_condition;
    if (bc != CLMSOIL_BC_WATERTABLE && bc != CLMSOIL_BC_FLUX && bc != CLMSOIL_BC_ZERO_FLUX)
        return CLMSOIL_ERR_BC;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (!(cfg->dtmin > 0.0) || !(cfg->verySmall > 0.0) || !std::isfinite(cfg->dtmin))
        return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->nbedrock || !a->h2osoi_liq || !a->smp_l || !a->hk_l || !a->qcharge ||
        !a->nsubsteps || !a->z || !a->dz || !a->qflx_rootsoi || !a->watsat || !a->hksat ||
        !a->bsw || !a->sucsat || !a->icefrac || !a->qflx_infl)
        return CLMSOIL_ERR_NULL;
    if (bc == CLMSOIL_BC_WATERTABLE && (!a->zwt || !a->zi)) return CLMSOIL_ERR_NULL;
    if (a->workspace && (a->workspace_bytes < RICHARDS_WORKSPACE_BYTES ||
                         (reinterpret_cast<uintptr_t>(a->workspace) & 255u) != 0))
        return CLMSOIL_ERR_WORKSPACE;
    // the per-XCD work counters are 32-bit: every XCD's range stays below 2^31
    if (a->nc > (int64_t)NUM_XCD << 31) return CLMSOIL_ERR_SIZE;
    return CLMSOIL_OK;
}

RichardsParams richards_params(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                               double dtime, const clmsoil_swm_config *cfg, bool aquifer,
                               unsigned *err) {
    RichardsParams P{};
    P.nc = a->nc; P.ld = a->ld; P.mask = a->mask; P.nbedrock = a->nbedrock;
    P.h2osoi_liq = a->h2osoi_liq; P.smp_l = a->smp_l; P.hk_l = a->hk_l;
    P.qcharge = a->qcharge; P.nsubsteps = a->nsubsteps; P.nrejected = a->nrejected;
    P.z = a->z; P.zi = a->zi; P.dz = a->dz; P.qflx_rootsoi = a->qflx_rootsoi;
    P.watsat = a->watsat; P.hksat = a->hksat; P.bsw = a->bsw; P.sucsat = a->sucsat;
    P.icefrac = a->icefrac; P.zwt = a->zwt; P.qflx_infl = a->qflx_infl;
    P.smpmin = a->smpmin; P.dtsub0 = a->dtsub0;
    P.nlevsno = lev->nlevsno; P.nlevsoi = lev->nlevsoi;
    P.lower_bc = cfg->lower_boundary_condition;
    P.expensive = cfg->flux_calculation == CLMSOIL_FLUX_EXPENSIVE;
    P.aquifer_pass = aquifer ? 1 : 0;
    P.max_iter = 100000;
    P.refill_sixteenths = g_tuning.richards_refill.load(std::memory_order_relaxed);
    P.dtime = dtime; P.e_ice = cfg->e_ice; P.dtmin = cfg->dtmin;
    P.xtu = cfg->xTolerUpper; P.xtl = cfg->xTolerLower; P.verysmall = cfg->verySmall;
    P.err = err;
    // the workspace (if any) holds the call's error word and the work counters
    unsigned *ws = static_cast<unsigned *>(a->workspace);
    P.call_err = ws;
    P.work_counters = ws ? ws + 8 : nullptr;
    return P;
}

int validate_heat(const clmsoil_heat_args *a, const clmsoil_levels *lev) {
    if (!a) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->snl || !a->nbedrock || !a->z || !a->dz || !a->zi || !a->t_soisno ||
        !a->h2osoi_liq || !a->h2osoi_ice || !a->excess_ice || !a->watsat || !a->tkmg ||
        !a->tkdry || !a->csol || !a->bsw || !a->sucsat || !a->frac_sno_eff ||
        !a->frac_h2osfc || !a->hs_top_snow || !a->hs_soil || !a->hs_top || !a->hs_h2osfc ||
        !a->dhsdT || !a->t_h2osfc || !a->h2osfc || !a->h2osno_no_layers || !a->snow_depth ||
        !a->t_grnd || !a->xmf || !a->xmf_h2osfc || !a->qflx_snomelt || !a->qflx_snofrz ||
        !a->qflx_snow_drain || !a->qflx_h2osfc_to_ice || !a->eflx_fgr12 ||
        !a->qflx_snomelt_lyr || !a->qflx_snofrz_lyr)
        return CLMSOIL_ERR_NULL;
    return CLMSOIL_OK;
}

__global__ void zero_words_kernel(unsigned *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// The resident-grid launch needs the workspace's counters; without a workspace
// the step runs as a plain launch (identical results).
int run_richards(const RichardsParams &P, hipStream_t stream) {
    const bool persistent =
        g_tuning.richards_persistent.load(std::memory_order_relaxed) != 0 && P.work_counters;
    if (P.call_err) {
        // clear the call's error word and the counters (a kernel, not a memset,
        // so that a captured graph replays it)
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, P.call_err, 16);
        if (hipGetLastError() != hipSuccess) return CLMSOIL_ERR_HIP;
    }
    return launch_richards(P, g_tuning.richards_g.load(std::memory_order_relaxed),
                           g_tuning.richards_block.load(std::memory_order_relaxed), persistent,
                           g_tuning.richards_blocks_per_cu.load(std::memory_order_relaxed),
                           g_tuning.richards_lds_fields.load(std::memory_order_relaxed),
                           stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// The step with the first-substep network fused into the column load (G = 4,
// 256-thread blocks whatever the tuning says; the persistent grid when the
// workspace is there).
int run_richards_learned(const RichardsParams &P, hipStream_t stream) {
    const bool persistent =
        g_tuning.richards_persistent.load(std::memory_order_relaxed) != 0 && P.work_counters;
    if (P.call_err) {
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, P.call_err, 16);
        if (hipGetLastError() != hipSuccess) return CLMSOIL_ERR_HIP;
    }
    return launch_richards_learned(P, persistent,
                                   g_tuning.richards_blocks_per_cu.load(std::memory_order_relaxed),
                                   stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

int run_heat(const clmsoil_heat_args *a, const clmsoil_levels *lev, double dtime,
             const clmsoil_heat_config *cfg, hipStream_t stream, unsigned *err) {
    HeatParams P{};
    P.a = *a;
    P.nlevsno = lev->nlevsno;
    P.nlevgrnd = lev->nlevgrnd;
    P.snow_method = cfg ? cfg->snow_thermal_cond_method : 1;
    P.dtime = dtime;
    P.err = err;
    return launch_heat(P, g_tuning.heat_g.load(std::memory_order_relaxed),
                       g_tuning.heat_block.load(std::memory_order_relaxed), stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

}  // namespace clmsoil

extern "C" {

int clmsoil_init(void) { return init_devices(); }

int64_t clmsoil_richards_workspace_bytes(void) { return RICHARDS_WORKSPACE_BYTES; }

int clmsoil_abi_version(void) { return CLMSOIL_ABI_VERSION; }

const char *clmsoil_strerror(int status) {
    switch (status) {
        case CLMSOIL_OK: return "ok";
        case CLMSOIL_ERR_NULL: return "a required pointer is NULL";
        case CLMSOIL_ERR_BC: return "unsupported boundary condition";
        case CLMSOIL_ERR_METHOD: return "unsupported soil water movement method";
        case CLMSOIL_ERR_SIZE: return "column or level count out of the supported range";
        case CLMSOIL_ERR_HIP: return "HIP runtime error";
        case CLMSOIL_ERR_WORKSPACE: return "workspace missing or too small";
        case CLMSOIL_ERR_ARG: return "invalid argument";
        case CLMSOIL_ERR_NOT_TRAINED: return "learned controller has no weights";
        case CLMSOIL_ERR_INIT: return "clmsoil_init() has not set up this device";
        default: return "unknown status";
    }
}

int clmsoil_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int clmsoil_sync(void *stream) {
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

int clmsoil_device_errors(void *stream, uint32_t *flags) {
    if (!flags) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return CLMSOIL_ERR_HIP;
    unsigned h = 0;
    if (hipMemcpy(&h, s->err, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)
        return CLMSOIL_ERR_HIP;
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, s->err, 1);
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return CLMSOIL_ERR_HIP;
    *flags = h;
    return CLMSOIL_OK;
}

int clmsoil_workspace_errors(const void *workspace, void *stream, uint32_t *flags) {
    if (!workspace || !flags) return CLMSOIL_ERR_NULL;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return CLMSOIL_ERR_HIP;
    unsigned h = 0;
    if (hipMemcpy(&h, workspace, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)
        return CLMSOIL_ERR_HIP;
    *flags = h;
    return CLMSOIL_OK;
}

int clmsoil_set_tuning(int richards_g, int richards_block, int heat_g, int heat_block) {
    auto okg = [](int g) { return g == 1 || g == 2 || g == 4; };
    auto okb = [](int b) { return b == 64 || b == 128 || b == 256; };
    if (!okg(richards_g) || !okg(heat_g) || !okb(richards_block) || !okb(heat_block))
        return CLMSOIL_ERR_ARG;
    g_tuning.richards_g.store(richards_g);
    g_tuning.richards_block.store(richards_block);
    g_tuning.heat_g.store(heat_g);
    g_tuning.heat_block.store(heat_block);
    return CLMSOIL_OK;
}

int clmsoil_set_richards_schedule(int persistent) {
    if (persistent != 0 && persistent != 1) return CLMSOIL_ERR_ARG;
    g_tuning.richards_persistent.store(persistent);
    return CLMSOIL_OK;
}

int clmsoil_set_richards_residency(int blocks_per_cu) {
    if (blocks_per_cu < 0 || blocks_per_cu > 32) return CLMSOIL_ERR_ARG;
    g_tuning.richards_blocks_per_cu.store(blocks_per_cu);
    return CLMSOIL_OK;
}

int clmsoil_set_richards_refill(int plain_sixteenths, int learned_sixteenths) {
    if (plain_sixteenths < 1 || plain_sixteenths > 16 || learned_sixteenths < 1 ||
        learned_sixteenths > 16)
        return CLMSOIL_ERR_ARG;
    g_tuning.richards_refill.store(plain_sixteenths);
    g_tuning.richards_refill_learned.store(learned_sixteenths);
    return CLMSOIL_OK;
}

int clmsoil_set_richards_lds_cache(int fields) {
    if (fields != 0 && fields != 5 && fields != 7) return CLMSOIL_ERR_ARG;
    g_tuning.richards_lds_fields.store(fields);
    return CLMSOIL_OK;
}

int clmsoil_set_zd_block(int block) {
    if (block != 64 && block != 128 && block != 256) return CLMSOIL_ERR_ARG;
    g_tuning.zd_block.store(block);
    return CLMSOIL_OK;
}

int clmsoil_get_tuning(int *richards_g, int *richards_block, int *heat_g, int *heat_block) {
    if (!richards_g || !richards_block || !heat_g || !heat_block) return CLMSOIL_ERR_NULL;
    *richards_g = g_tuning.richards_g.load();
    *richards_block = g_tuning.richards_block.load();
    *heat_g = g_tuning.heat_g.load();
    *heat_block = g_tuning.heat_block.load();
    return CLMSOIL_OK;
}

int clmsoil_soilwm_moisture_form_solve(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                       double dtime, const clmsoil_swm_config *cfg,
                                       void *stream) {
    int st = validate_mf(a, lev, dtime, cfg);
    if (st || a->nc == 0) return st;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    RichardsParams P = richards_params(a, lev, dtime, cfg, false, s->err);
    return run_richards(P, (hipStream_t)stream);
}

int clmsoil_soilwm_moisture_form_solve_static(const clmsoil_mf_args *a,
                                              const clmsoil_levels *lev, double dtime,
                                              const clmsoil_swm_config *cfg, void *stream) {
    int st = validate_mf(a, lev, dtime, cfg);
    if (st || a->nc == 0) return st;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    // the static scratch stands in for the caller's workspace
    clmsoil_mf_args m = *a;
    m.workspace = s->scratch;
    m.workspace_bytes = RICHARDS_WORKSPACE_BYTES;
    RichardsParams P = richards_params(&m, lev, dtime, cfg, false, s->err);
    return run_richards(P, (hipStream_t)stream);
}

int clmsoil_soilwm_moisture_form_solve_learned(const clmsoil_mf_args *a,
                                               const clmsoil_levels *lev, double dtime,
                                               const clmsoil_swm_config *cfg,
                                               const float *params, const float *norm,
                                               void *stream) {
    int st = validate_mf(a, lev, dtime, cfg);
    if (st) return st;
    if (!params || !norm) return CLMSOIL_ERR_NOT_TRAINED;
    if (a->dtsub0) return CLMSOIL_ERR_ARG;   // the network supplies the first substep
    if (a->nc == 0) return CLMSOIL_OK;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    RichardsParams P = richards_params(a, lev, dtime, cfg, false, s->err);
    P.mlp = params;
    P.mlp_norm = norm;
    P.refill_sixteenths = g_tuning.richards_refill_learned.load(std::memory_order_relaxed);
    return run_richards_learned(P, (hipStream_t)stream);
}

int clmsoil_soilwater_moisture_form(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                    double dtime, const clmsoil_swm_config *cfg, void *stream) {
    int st = validate_mf(a, lev, dtime, cfg);
    if (st || a->nc == 0) return st;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    const bool aquifer = cfg->lower_boundary_condition == CLMSOIL_BC_WATERTABLE ||
                         cfg->lower_boundary_condition == CLMSOIL_BC_AQUIFER;
    RichardsParams P = richards_params(a, lev, dtime, cfg, aquifer, s->err);
    return run_richards(P, (hipStream_t)stream);
}

int clmsoil_soil_temperature(const clmsoil_heat_args *a, const clmsoil_levels *lev, double dtime,
                             const clmsoil_heat_config *cfg, void *stream) {
    int st = validate_heat(a, lev);
    if (st || a->nc == 0) return st;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (cfg && cfg->snow_thermal_cond_method != 1 && cfg->snow_thermal_cond_method != 2)
        return CLMSOIL_ERR_ARG;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    return run_heat(a, lev, dtime, cfg, (hipStream_t)stream, s->err);
}

int clmsoil_tridiagonal_multi(double *u, const double *a, const double *b, const double *c,
                              const double *r, const int64_t *jtop, const uint8_t *mask,
                              int64_t ncols, int64_t nlevs, int64_t ld, void *stream) {
    if (ncols < 0 || ld < ncols || nlevs < 1 || nlevs > 64) return CLMSOIL_ERR_SIZE;
    if (ncols == 0) return CLMSOIL_OK;
    if (!u || !a || !b || !c || !r || !jtop) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    TridiagParams P{u, a, b, c, r, jtop, mask, ncols, nlevs, ld, s->err};
    return launch_tridiag_multi(P, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

int clmsoil_batched_band_solve(double *tvector, const double *bmatrix, const double *rvector,
                               const int64_t *jtop, const int64_t *jbot, const uint8_t *mask,
                               int32_t kl, int32_t ku, int32_t nlevsno_off, int64_t nc,
                               int64_t nlev, int64_t ld, void *stream) {
    if (nc < 0 || ld < nc || nlev < 1) return CLMSOIL_ERR_SIZE;
    if (kl < 1 || ku < 1 || kl > 2 || ku > 2) return CLMSOIL_ERR_ARG;
    if (nc == 0) return CLMSOIL_OK;
    if (!tvector || !bmatrix || !rvector || !jtop || !jbot) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    BandParams P{tvector, bmatrix, rvector, jtop, jbot, mask, kl, ku, nlevsno_off,
                 nc, nlev, ld, s->err};
    return launch_band_solve(P, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

int clmsoil_tridiagonal_multi_vjp(double *ubar, const double *u, const double *a, const double *b,
                                  const double *c, const int64_t *jtop, const uint8_t *mask,
                                  int64_t ncols, int64_t nlevs, int64_t ld, double *abar,
                                  double *bbar, double *cbar, double *rbar, void *stream) {
    if (ncols < 0 || ld < ncols || nlevs < 1 || nlevs > 64) return CLMSOIL_ERR_SIZE;
    if (ncols == 0) return CLMSOIL_OK;
    if (!ubar || !u || !a || !b || !c || !jtop) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    TridiagVjpParams P{ubar, u, a, b, c, jtop, mask, ncols, nlevs, ld, abar, bbar, cbar, rbar,
                       s->err};
    return launch_tridiag_vjp(P, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

int clmsoil_batched_band_solve_vjp(double *tbar, const double *tvector, const double *bmatrix,
                                   const int64_t *jtop, const int64_t *jbot, const uint8_t *mask,
                                   int32_t kl, int32_t ku, int32_t nlevsno_off, int64_t nc,
                                   int64_t nlev, int64_t ld, double *bmatrix_bar,
                                   double *rvector_bar, void *stream) {
    if (nc < 0 || ld < nc || nlev < 1) return CLMSOIL_ERR_SIZE;
    if (kl < 1 || ku < 1 || kl > 2 || ku > 2) return CLMSOIL_ERR_ARG;
    if (nc == 0) return CLMSOIL_OK;
    if (!tbar || !tvector || !bmatrix || !jtop || !jbot) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    BandVjpParams P{tbar, tvector, bmatrix, jtop, jbot, mask, kl, ku, nlevsno_off, nc, nlev, ld,
                    bmatrix_bar, rvector_bar, s->err};
    return launch_band_vjp(P, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

// Host checks of soil_water! (soil_water_movement.jl:1780-1810) for the
// Zeng-Decker 2009 method. The aquifer row reads layer nlevsoi+1 of t_soisno and
// zi, so the ground must extend below the soil column.
int clmsoil_soilwater_zengdecker2009(const clmsoil_zd_args *a, const clmsoil_levels *lev,
                                     double dtime, const clmsoil_swm_config *cfg,
                                     void *stream) {
    if (!a || !cfg) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (lev->nlevgrnd < lev->nlevsoi + 1) return CLMSOIL_ERR_SIZE;
    if (cfg->soilwater_movement_method != CLMSOIL_ZENGDECKER_2009) return CLMSOIL_ERR_METHOD;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime) || !std::isfinite(cfg->e_ice))
        return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->z || !a->dz || !a->zi || !a->watsat || !a->hksat || !a->bsw || !a->sucsat ||
        !a->smpmin || !a->h2osoi_liq || !a->h2osoi_ice || !a->h2osoi_vol || !a->t_soisno ||
        !a->zwt || !a->qflx_infl || !a->qflx_rootsoi || !a->icefrac || !a->smp_l ||
        !a->hk_l || !a->qcharge || !a->qflx_deficit)
        return CLMSOIL_ERR_NULL;
    ZdParams P{a->nc, a->ld, a->mask, a->z, a->dz, a->zi, a->watsat, a->hksat, a->bsw,
               a->sucsat, a->smpmin, a->h2osoi_liq, a->h2osoi_ice, a->h2osoi_vol,
               a->t_soisno, a->zwt, a->qflx_infl, a->qflx_rootsoi, a->icefrac, a->smp_l,
               a->hk_l, a->qcharge, a->qflx_deficit, lev->nlevsno, lev->nlevsoi,
               lev->nlevgrnd, dtime, cfg->e_ice};
    return launch_zengdecker2009(P, g_tuning.zd_block.load(std::memory_order_relaxed),
                                 (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// compute_ground_heat_flux_and_deriv! (soil_temperature.jl:1200-1263).
int clmsoil_ground_heat_flux(const clmsoil_ghf_args *a, const clmsoil_levels *lev,
                             void *stream) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    if (lev->nlevsno < 1 || lev->nlevsno > MAX_NLEVSNO || lev->nlevgrnd < 1)
        return CLMSOIL_ERR_SIZE;
    if (a->nc < 0 || a->np < 0 || a->nl < 0 || a->ldc < a->nc || a->ldp < a->np)
        return CLMSOIL_ERR_SIZE;
    if (a->nc == 0 && a->np == 0) return CLMSOIL_OK;
    const void *req[] = {
        a->patchi, a->patchf, a->patch_column, a->patch_landunit, a->col_itype, a->snl,
        a->urbpoi, a->mask_nolakec, a->mask_nolakep, a->frac_veg_nosno, a->emg, a->t_grnd,
        a->t_h2osfc, a->forc_lwrad, a->htvp, a->frac_sno_eff, a->t_soisno, a->sabg,
        a->sabg_soil, a->sabg_snow, a->dlrad, a->cgrnd, a->eflx_sh_grnd, a->eflx_sh_snow,
        a->eflx_sh_soil, a->eflx_sh_h2osfc, a->qflx_evap_soi, a->qflx_ev_snow, a->qflx_ev_soil,
        a->qflx_ev_h2osfc, a->eflx_lwrad_net, a->qflx_tran_veg, a->wtcol, a->sabg_lyr_patch,
        a->eflx_wasteheat_lun, a->eflx_ventilation_lun, a->eflx_heat_from_ac_lun,
        a->eflx_traffic_lun, a->wtlunit_roof, a->eflx_gnet, a->dgnetdT, a->sabg_chk,
        a->eflx_wasteheat, a->eflx_ventilation, a->eflx_heat_from_ac, a->eflx_traffic,
        a->eflx_anthro, a->hs_top, a->hs_top_snow, a->hs_soil, a->hs_h2osfc, a->dhsdT,
        a->sabg_lyr_col};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    GhfParams P{*a, lev->nlevsno, s->err};
    return launch_ground_heat_flux(P, g_tuning.ghf_block.load(std::memory_order_relaxed),
                                   (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// errsoi of soil_fluxes! (soil_fluxes.jl:380-447, 499-525).
int clmsoil_soil_energy_balance(const clmsoil_seb_args *a, const clmsoil_levels *lev,
                                double dtime, void *stream) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    if (lev->nlevsno < 1 || lev->nlevsno > MAX_NLEVSNO || lev->nlevgrnd < 1 || a->nlevurb < 0)
        return CLMSOIL_ERR_SIZE;
    if (a->nc < 0 || a->np < 0 || a->ldc < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->patchi, a->patchf, a->patch_column, a->col_itype, a->snl,
                         a->mask_nolakec, a->mask_nolakep, a->mask_urbanp, a->eflx_soil_grnd,
                         a->wtcol, a->xmf, a->xmf_h2osfc, a->frac_h2osfc, a->t_h2osfc,
                         a->t_h2osfc_bef, a->c_h2osfc, a->eflx_h2osfc_to_snow,
                         a->eflx_building_heat_errsoi, a->frac_sno_eff, a->t_soisno, a->t_ssbef,
                         a->fact, a->errsoi_patch, a->errsoi_col};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    return launch_soil_energy_balance(*a, lev->nlevsno, lev->nlevgrnd, dtime, s->err,
                                      (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// set_soil_water_fractions! (soil_hydrology.jl:394-415).
int clmsoil_set_soil_water_fractions(const clmsoil_fractions_args *a, const clmsoil_levels *lev,
                                     void *stream) {
    if (!a) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->watsat || !a->dz || !a->h2osoi_ice || !a->eff_porosity || !a->icefrac)
        return CLMSOIL_ERR_NULL;
    return launch_soil_water_fractions(*a, lev->nlevsno, lev->nlevsoi, (hipStream_t)stream) ==
                   hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// perched_water_table! / theta_based_water_table! / clamp_zwt_to_bedrock!
// (soil_hydrology.jl:1604-1653, 1732-1765; kernels.jl:209-227), ops as in
// launch_theta_water_table.
static int theta_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev, int ops,
                             void *stream) {
    if (!a) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->zi || !a->zwt) return CLMSOIL_ERR_NULL;
    if ((ops & 3) && (!a->dz || !a->watsat || !a->h2osoi_liq || !a->h2osoi_ice || !a->h2osoi_vol))
        return CLMSOIL_ERR_NULL;
    if ((ops & 1) && (!a->z || !a->t_soisno || !a->zwt_perched || !a->frost_table))
        return CLMSOIL_ERR_NULL;
    if ((ops & 2) && !a->z) return CLMSOIL_ERR_NULL;
    if ((ops & 6) && !a->nbedrock) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    st = current_device_state(&s);
    if (st) return st;
    return launch_theta_water_table(*a, lev->nlevsno, lev->nlevsoi, ops, s->err,
                                    (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

int clmsoil_perched_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                void *stream) {
    return theta_water_table(a, lev, 1, stream);
}

int clmsoil_theta_based_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                    void *stream) {
    return theta_water_table(a, lev, 2, stream);
}

int clmsoil_clamp_zwt_to_bedrock(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                 void *stream) {
    return theta_water_table(a, lev, 4, stream);
}

// The non-aquifer water tables of clm_driver.jl:2076-2092 in one launch.
int clmsoil_water_table_moisture_form(const clmsoil_theta_wt_args *a,
                                      const clmsoil_levels *lev, void *stream) {
    return theta_water_table(a, lev, 7, stream);
}

// The surface hydrology of clm_driver.jl:1939-2012 (include/clmsoil_surfhyd.h).
int clmsoil_surface_hydrology(const clmsoil_surfhyd_args *a, const clmsoil_levels *lev,
                              double dtime, const clmsoil_surfhyd_params *p, void *stream) {
    if (!a || !p) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    const int s = p->steps;
    if ((s & ~CLMSOIL_SH_ALL) != 0) return CLMSOIL_ERR_ARG;
    if (p->fsat_method != CLMSOIL_FSAT_TOPMODEL && p->fsat_method != CLMSOIL_FSAT_VIC)
        return CLMSOIL_ERR_ARG;
    if (p->qinmax_method != CLMSOIL_QINMAX_NONE && p->qinmax_method != CLMSOIL_QINMAX_HKSAT)
        return CLMSOIL_ERR_ARG;
    if (a->nc == 0 || s == 0) return CLMSOIL_OK;
    auto need = [](std::initializer_list<const void *> l) {
        for (const void *q : l)
            if (!q) return false;
        return true;
    };
    bool ok = true;
    if (s & CLMSOIL_SH_FLOODC)
        ok = ok && need({a->col_gridcell, a->col_itype, a->qflx_floodg, a->qflx_floodc});
    if (s & CLMSOIL_SH_SAT_EXCESS) {
        ok = ok && need({a->col_urbpoi, a->qflx_rain_plus_snomelt, a->qflx_floodc, a->fsat,
                         a->fcov, a->qflx_sat_excess_surf});
        ok = ok && (p->fsat_method == CLMSOIL_FSAT_TOPMODEL
                        ? need({a->frost_table, a->zwt, a->zwt_perched, a->wtfact})
                        : need({a->b_infil, a->top_max_moist, a->top_moist_limited}));
        if (p->crop_fsat_equals_zero) ok = ok && need({a->col_landunit, a->lun_itype});
        if (p->hillslope_fsat_equals_zero)
            ok = ok && need({a->is_hillslope_column, a->col_active, a->col_cold});
    }
    if (s & CLMSOIL_SH_QFLX_INPUTS)
        ok = ok && need({a->snl, a->qflx_rain_plus_snomelt, a->qflx_snow_h2osfc, a->qflx_floodc,
                         a->qflx_liqevap_from_top_layer, a->qflx_ev_soil, a->qflx_ev_h2osfc,
                         a->qflx_sat_excess_surf, a->frac_sno_eff, a->frac_h2osfc,
                         a->qflx_top_soil, a->qflx_in_soil, a->qflx_top_soil_to_h2osfc});
    if (s & CLMSOIL_SH_INFIL_EXCESS) {
        ok = ok && need({a->fsat, a->qflx_in_soil, a->frac_h2osfc, a->qinmax,
                         a->qflx_infl_excess});
        if (p->qinmax_method == CLMSOIL_QINMAX_HKSAT) ok = ok && need({a->icefrac, a->hksat});
    }
    if (s & CLMSOIL_SH_ROUTE)
        ok = ok && need({a->col_landunit, a->lun_itype, a->qflx_in_soil, a->qflx_infl_excess,
                         a->qflx_top_soil_to_h2osfc, a->qflx_in_soil_limited,
                         a->qflx_in_h2osfc, a->qflx_infl_excess_surf});
    if (s & CLMSOIL_SH_H2OSFC)
        ok = ok && need({a->h2osfc, a->qflx_in_h2osfc, a->frac_h2osfc, a->qinmax,
                         a->qflx_h2osfc_surf, a->qflx_h2osfc_drain}) &&
             (p->h2osfcflag_update != 1 ||
              need({a->frac_h2osfc_nosnow, a->h2osfc_thresh, a->topo_slope}));
    if (s & CLMSOIL_SH_INFILTRATION)
        ok = ok && need({a->qflx_in_soil_limited, a->qflx_h2osfc_drain, a->qflx_infl});
    if (s & CLMSOIL_SH_SURF_RUNOFF)
        ok = ok && need({a->qflx_sat_excess_surf, a->qflx_infl_excess_surf, a->qflx_h2osfc_surf,
                         a->qflx_surf});
    if (s & (CLMSOIL_SH_SURF_RUNOFF | CLMSOIL_SH_URBAN_POND))
        ok = ok && need({a->col_itype, a->snl, a->qflx_rain_plus_snomelt,
                         a->qflx_liqevap_from_top_layer, a->qflx_floodc, a->h2osoi_liq,
                         a->xs_urban});
    if (!ok) return CLMSOIL_ERR_NULL;
    return launch_surface_hydrology(*a, *p, lev->nlevsno, dtime, (hipStream_t)stream) ==
                   hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// compute_effec_rootfrac_and_vert_tran_sink! with use_hydrstress = false
// (soil_water_plant_sink.jl:341-451).
int clmsoil_compute_effec_rootfrac_and_vert_tran_sink(const clmsoil_rootsink_args *a,
                                                      const clmsoil_levels *lev, void *stream) {
    if (!a) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->np < 0 || a->ldc < a->nc || a->ldp < a->np) return CLMSOIL_ERR_SIZE;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->patchi, a->npatches, a->qflx_tran_veg_col, a->rootr_col,
                         a->qflx_rootsoi_col};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    if (a->np > 0 && (!a->active || !a->wtcol || !a->rootr_patch || !a->qflx_tran_veg_patch))
        return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    st = current_device_state(&s);
    if (st) return st;
    return launch_root_sink(*a, lev->nlevsoi, s->err, (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// Steps 6-7 of lake_temperature! and the unpack (lake_temperature.jl:1205-1225).
int clmsoil_lake_temperature_solve(const clmsoil_lake_solve_args *a, const clmsoil_levels *lev,
                                   double dtime, void *stream) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    if (lev->nlevsno < 1 || lev->nlevsno > MAX_NLEVSNO || lev->nlevgrnd < 1) return CLMSOIL_ERR_SIZE;
    if (a->nlevlak < 1 || lev->nlevsno + a->nlevlak + lev->nlevgrnd > LK_NMAX)
        return CLMSOIL_ERR_SIZE;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->snl, a->z, a->cv, a->tk, a->sabg_lyr_col, a->z_lake, a->dz_lake,
                         a->cv_lake, a->tk_lake, a->phi, a->phi_soil, a->fin, a->tktopsoillay,
                         a->t_soisno, a->t_lake};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    return launch_lake_temperature_solve(*a, lev->nlevsno, lev->nlevgrnd, dtime, s->err,
                                         (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// litter_vert_transp! without the matrix-CN path (litter_vert_transp.jl:479-639).
int clmsoil_litter_vert_transp(const clmsoil_lvt_args *a, double dtime,
                               const clmsoil_lvt_params *p, void *stream) {
    if (!a || !p) return CLMSOIL_ERR_NULL;
    if (p->use_soil_matrixcn) return CLMSOIL_ERR_METHOD;
    if (p->nlevdecomp < 1 || p->nlevdecomp > 62 || p->nlevdecomp_full < p->nlevdecomp ||
        p->npools < 1 || p->npools > CLMSOIL_LVT_MAXPOOLS || p->ntypes < 1 ||
        p->ntypes > CLMSOIL_LVT_MAXTYPES)
        return CLMSOIL_ERR_SIZE;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->nbedrock, a->altmax, a->altmax_lastyear, a->zsoi, a->dzsoi_decomp,
                         a->zisoi, a->som_diffus_coef, a->som_adv_coef};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    for (int t = 0; t < p->ntypes; ++t)
        if (!a->conc[t] || !a->source[t] || !a->tendency[t]) return CLMSOIL_ERR_NULL;
    if (p->spinup_state >= 1) {
        bool lat = false;
        for (int s = 0; s < p->npools; ++s)
            if (!p->is_cwd[s] && std::fabs(p->spinup_factor[s] - 1.0) > 1.0e-6) lat = true;
        if (lat && (!a->col_gridcell || !a->latdeg)) return CLMSOIL_ERR_NULL;
    }
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    return launch_litter_vert_transp(*a, *p, dtime, s->err, (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

// ch4_tran! (methane.jl:2005-2092).
int clmsoil_ch4_tran(const clmsoil_ch4_tran_args *a, double dtime, const clmsoil_ch4_tran_params *p,
                     void *stream) {
    if (!a || !p) return CLMSOIL_ERR_NULL;
    if (p->nlevsoi < 1 || p->nlevsoi > 60) return CLMSOIL_ERR_SIZE;
    if (p->sat != 0 && p->sat != 1) return CLMSOIL_ERR_ARG;
    if (a->nc < 0 || a->ld < a->nc || a->ldg < 1) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->col_gridcell, a->jwt, a->c_atm, a->watsat, a->h2osoi_vol, a->h2osoi_liq,
                         a->h2osoi_ice, a->bsw, a->cellorg, a->t_soisno, a->dz, a->h2osfc, a->t_grnd,
                         a->t_h2osfc, a->frac_h2osfc, a->grnd_ch4_cond, a->ch4_prod_depth,
                         a->o2_aere_depth, a->o2_decomp_depth, a->o2_oxid_depth, a->ch4_oxid_depth,
                         a->ch4_aere_depth, a->ch4_ebul_depth, a->o2stress, a->ch4stress, a->conc_ch4,
                         a->conc_o2, a->ch4_ebul_total, a->ch4_surf_aere, a->ch4_surf_ebul,
                         a->ch4_surf_diff};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    int st = current_device_state(&s);
    if (st) return st;
    return launch_ch4_tran(*a, *p, dtime, s->err, (hipStream_t)stream) == hipSuccess ? CLMSOIL_OK
                                                                                     : CLMSOIL_ERR_HIP;
}

// water_table! (soil_hydrology.jl:849-902).
int clmsoil_water_table(const clmsoil_water_table_args *a, const clmsoil_levels *lev,
                        double dtime, double aq_sp_yield_min, int recompute_frost_table,
                        void *stream) {
    if (!a) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime) || !std::isfinite(aq_sp_yield_min))
        return CLMSOIL_ERR_ARG;
    if (recompute_frost_table != 0 && recompute_frost_table != 1) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    const void *req[] = {a->dz, a->z, a->zi, a->t_soisno, a->watsat, a->sucsat, a->bsw,
                         a->qcharge, a->h2osoi_liq, a->h2osoi_ice, a->zwt, a->wa, a->zwt_perched,
                         a->frost_table, a->h2osoi_vol, a->qflx_drain, a->qflx_rsub_sat,
                         a->qflx_drain_perched};
    for (const void *q : req)
        if (!q) return CLMSOIL_ERR_NULL;
    return launch_water_table(*a, lev->nlevsno, lev->nlevsoi, dtime, aq_sp_yield_min,
                              recompute_frost_table, (hipStream_t)stream) == hipSuccess
               ? CLMSOIL_OK
               : CLMSOIL_ERR_HIP;
}

}  // extern "C"
