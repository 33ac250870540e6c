/*
 * clmsoil.h — C ABI of the MI355X (gfx950) per-column implicit soil solves.
 *
 * Drop-in boundary for CLM.jl's soil-water (moisture-form Richards) and
 * snow/soil heat steps. Every entry point takes plain pointers to device memory
 * laid out exactly as CLM.jl's `Matrix{Float64}` state fields: column-major
 * `[c, j]`, element (c, j) (0-based) at `c + j*ld`, `ld >= nc`. Snow+soil arrays
 * (`z`, `dz`, `t_soisno`, `h2osoi_liq/ice`, `fact`, `thk`, `imelt`) have
 * `nlevsno + nlevgrnd` levels with Fortran level j at 0-based index
 * `j + nlevsno - 1`; `zi` has one more level (interface j at `j + nlevsno`);
 * soil-only arrays (`watsat`, `bsw`, …) are indexed by soil level j at `j - 1`.
 * Integer state (`nbedrock`, `snl`, `imelt`) is Int64 as in Julia; masks are Bool
 * (one byte). Streams are `hipStream_t` passed as `void*` (NULL = default stream).
 *
 * All calls are asynchronous on the given stream, never free caller memory and
 * never allocate: clmsoil_init() allocates the library's only device memory (a
 * 256-byte status block per device) once, before any other call; scratch that
 * a solve needs comes from a caller workspace. They return a status code;
 * `clmsoil_strerror` names it. Calls are re-entrant: calls on different streams
 * (or threads) share nothing but the per-device error flags, which kernels
 * only OR into.
 *
 * Reference interfaces replaced (paths relative to the CLM.jl tree):
 *   soilwm_moisture_form_solve!   src/biogeophys/soil_water_movement.jl:1498-1517
 *   soilwater_moisture_form!      src/biogeophys/soil_water_movement.jl:1532-1601
 *   soil_water!                   src/biogeophys/soil_water_movement.jl:1780-1846
 *   soil_temperature!             src/biogeophys/soil_temperature.jl:358-593
 *   tridiagonal_multi!            src/infrastructure/tridiagonal.jl:129-142
 *   batched_band_solve!           src/infrastructure/band_diagonal.jl:179-189
 * The backend seam they bypass is `_launch!` (src/infrastructure/kernels.jl:85-91)
 * with the AMDGPU backend registered by ext/CLMAMDGPUExt.jl:41-54.
 */
#ifndef CLMSOIL_H
#define CLMSOIL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLMSOIL_ABI_VERSION 2

/* ---- status codes ---------------------------------------------------- */
enum {
    CLMSOIL_OK = 0,
    CLMSOIL_ERR_NULL = 1,        /* a required pointer is NULL                    */
    CLMSOIL_ERR_BC = 2,          /* unsupported boundary condition                */
    CLMSOIL_ERR_METHOD = 3,      /* unsupported solution method                   */
    CLMSOIL_ERR_SIZE = 4,        /* nc/ld/level counts out of the supported range */
    CLMSOIL_ERR_HIP = 5,         /* a HIP runtime call failed                     */
    CLMSOIL_ERR_WORKSPACE = 6,   /* workspace missing or too small                */
    CLMSOIL_ERR_ARG = 7,         /* other invalid argument                        */
    CLMSOIL_ERR_NOT_TRAINED = 8, /* learned controller used without weights       */
    CLMSOIL_ERR_INIT = 9         /* clmsoil_init() has not run on this device      */
};

/* ---- method / boundary-condition codes (soil_water_movement.jl:23-34) --- */
#define CLMSOIL_ZENGDECKER_2009 0
#define CLMSOIL_MOISTURE_FORM   1
#define CLMSOIL_BC_HEAD       0
#define CLMSOIL_BC_FLUX       1
#define CLMSOIL_BC_ZERO_FLUX  2
#define CLMSOIL_BC_WATERTABLE 3
#define CLMSOIL_BC_AQUIFER    4
#define CLMSOIL_FLUX_INEXPENSIVE 1
#define CLMSOIL_FLUX_EXPENSIVE   42

/* Level counts (varpar, src/constants/varpar.jl:112-149). The kernels support
 * nlevsoi <= 20, nlevgrnd <= 25 and nlevsno <= 12 (the CLM5 20SL_8.5m layout). */
typedef struct clmsoil_levels {
    int32_t nlevsno;
    int32_t nlevsoi;
    int32_t nlevgrnd;
    int32_t reserved;
} clmsoil_levels;

/* SoilWaterMovementConfig (soil_water_movement.jl:654-671). */
typedef struct clmsoil_swm_config {
    int32_t soilwater_movement_method;   /* must be CLMSOIL_MOISTURE_FORM        */
    int32_t upper_boundary_condition;    /* must be CLMSOIL_BC_FLUX              */
    int32_t lower_boundary_condition;    /* WATERTABLE, FLUX or ZERO_FLUX        */
    int32_t flux_calculation;            /* 1 inexpensive (default), 42 expensive */
    double dtmin;                        /* 60 s                                 */
    double verySmall;                    /* 1e-8                                 */
    double xTolerUpper;                  /* 0.1                                  */
    double xTolerLower;                  /* 0.01                                 */
    double e_ice;                        /* 6                                    */
} clmsoil_swm_config;

/* MfState + MfIn (soil_water_movement.jl:1230-1240) plus mask/nbedrock. */
typedef struct clmsoil_mf_args {
    int64_t nc;                 /* number of columns                            */
    int64_t ld;                 /* leading dimension of every [c, j] array       */
    const uint8_t *mask;        /* [nc] active columns; NULL = all active        */
    const int64_t *nbedrock;    /* [nc] number of Richards layers (1..nlevsoi)    */
    /* MfState (read/write) */
    double *h2osoi_liq;         /* [ld, nlevsno+nlevgrnd] kg/m2                   */
    double *smp_l;              /* [ld, >=nlevsoi] mm (written)                   */
    double *hk_l;               /* [ld, >=nlevsoi] mm/s (written)                 */
    double *qcharge;            /* [nc] mm/s (written)                            */
    double *nsubsteps;          /* [nc] substep iterations incl. rejected (written) */
    /* MfIn (read only) */
    const double *z;            /* [ld, nlevsno+nlevgrnd] m                       */
    const double *zi;           /* [ld, nlevsno+nlevgrnd+1] m                     */
    const double *dz;           /* [ld, nlevsno+nlevgrnd] m                       */
    const double *qflx_rootsoi; /* [ld, nlevsoi] mm/s                             */
    const double *watsat;       /* [ld, nlevgrnd]                                 */
    const double *hksat;        /* [ld, nlevgrnd] mm/s                            */
    const double *bsw;          /* [ld, nlevgrnd]                                 */
    const double *sucsat;       /* [ld, nlevgrnd] mm                              */
    const double *icefrac;      /* [ld, nlevgrnd]                                 */
    const double *zwt;          /* [nc] m (read for BC_WATERTABLE)                */
    const double *qflx_infl;    /* [nc] mm/s                                      */
    const double *smpmin;       /* [nc] mm, used by the aquifer recharge pass;
                                   NULL = -1e8 (cold_start.jl:763)                */
    /* Optional diagnostics (NULL = not written). */
    double *nrejected;          /* [nc] rejected substeps                         */
    const double *dtsub0;       /* [nc] first substep length (s); NULL = dtime
                                   (soil_water_movement.jl:1259). Used by the
                                   learned controller.                            */
    /* Scratch of the call (the reference allocates its MfScr per call,
     * soil_water_movement.jl:1555-1561): device memory of at least
     * clmsoil_richards_workspace_bytes() bytes, 256-byte aligned, owned by the
     * caller and used by one call at a time. It holds the work counters of the
     * resident-grid launch and the call's own error flags
     * (clmsoil_workspace_errors). NULL: the step runs as a plain launch with one
     * column per lane group (same results, needs no scratch).              */
    void *workspace;
    int64_t workspace_bytes;
} clmsoil_mf_args;

/* Heat step state and forcing (soil_temperature.jl:358-593, non-urban soil
 * columns). Ground-heat-flux terms (hs_*, dhsdT, sabg_lyr_col) are inputs, as
 * produced by compute_ground_heat_flux_and_deriv! (soil_temperature.jl:1200). */
typedef struct clmsoil_heat_args {
    int64_t nc;
    int64_t ld;
    const uint8_t *mask;          /* [nc] NULL = all active                       */
    const int64_t *snl;           /* [nc] minus the number of snow layers         */
    const int64_t *nbedrock;      /* [nc]                                         */
    /* column geometry */
    const double *z, *dz;         /* [ld, nlevsno+nlevgrnd]                       */
    const double *zi;             /* [ld, nlevsno+nlevgrnd+1]                     */
    /* layer state (read/write) */
    double *t_soisno;             /* [ld, nlevsno+nlevgrnd] K                     */
    double *h2osoi_liq;           /* [ld, nlevsno+nlevgrnd] kg/m2                 */
    double *h2osoi_ice;           /* [ld, nlevsno+nlevgrnd] kg/m2                 */
    double *excess_ice;           /* [ld, nlevgrnd] kg/m2                         */
    /* soil properties (read) */
    const double *watsat, *tkmg, *tkdry, *csol, *bsw, *sucsat;   /* [ld, nlevgrnd] */
    /* column forcing / state */
    const double *frac_sno_eff;   /* [nc]                                         */
    const double *frac_h2osfc;    /* [nc]                                         */
    const double *hs_top_snow, *hs_soil, *hs_top, *hs_h2osfc, *dhsdT;  /* [nc]    */
    const double *sabg_lyr_col;   /* [ld, nlevsno+1] W/m2; NULL = zero             */
    const double *eflx_bot;       /* [nc] W/m2; NULL = zero                       */
    double *t_h2osfc;             /* [nc] K (read/write)                          */
    double *h2osfc;               /* [nc] kg/m2 (read/write)                      */
    double *h2osno_no_layers;     /* [nc] kg/m2 (read/write)                      */
    double *snow_depth;           /* [nc] m (read/write)                          */
    double *int_snow;             /* [nc] kg/m2 (read/write); NULL = not kept     */
    /* column outputs (written) */
    double *t_grnd;               /* [nc] K                                       */
    double *xmf;                  /* [nc] W/m2                                    */
    double *xmf_h2osfc;           /* [nc] W/m2                                    */
    double *qflx_snomelt;         /* [nc] mm/s                                    */
    double *qflx_snofrz;          /* [nc] mm/s                                    */
    double *qflx_snow_drain;      /* [nc] mm/s                                    */
    double *qflx_h2osfc_to_ice;   /* [nc] mm/s                                    */
    double *eflx_fgr12;           /* [nc] W/m2                                    */
    double *qflx_snomelt_lyr;     /* [ld, nlevsno] mm/s                           */
    double *qflx_snofrz_lyr;      /* [ld, nlevsno] mm/s                           */
    /* optional diagnostics (NULL = not written) */
    double *fact;                 /* [ld, nlevsno+nlevgrnd] TemperatureData.fact_col */
    double *c_h2osfc;             /* [nc]                                         */
    int64_t *imelt;               /* [ld, nlevsno+nlevgrnd]                       */
    double *thk;                  /* [ld, nlevsno+nlevgrnd] SoilStateData.thk_col */
    double *bw;                   /* [ld, nlevsno] WaterDiagnosticBulk.bw_col     */
    double *eflx_fgr;             /* [ld, nlevgrnd]                               */
    double *exice_subs;           /* [ld, nlevgrnd]                               */
    double *eflx_snomelt;         /* [nc]                                         */
    double *eflx_snomelt_r;       /* [nc] (soil/crop land unit copy)              */
    double *eflx_h2osfc_to_snow;  /* [nc]                                         */
    double *snomelt_accum;        /* [nc] accumulated (read/write)                */
} clmsoil_heat_args;

/* Heat-step switches (varctl.jl:158,215). */
typedef struct clmsoil_heat_config {
    int32_t snow_thermal_cond_method;   /* 1 = Jordan1991 (default), 2 = Sturm1997 */
    int32_t reserved;
} clmsoil_heat_config;

/* ---- library ---------------------------------------------------------- */
/* Allocate the per-device status block on every visible GPU. Call once, before
 * any other call (graph capture included); idempotent and thread-safe. Other
 * entry points return CLMSOIL_ERR_INIT on a device it has not set up. */
int         clmsoil_init(void);
/* Bytes of the workspace clmsoil_mf_args.workspace points to. */
int64_t     clmsoil_richards_workspace_bytes(void);
int         clmsoil_abi_version(void);
const char *clmsoil_strerror(int status);
/* Number of GPUs visible to the library (0 when none). */
int         clmsoil_device_count(void);
/* Block until all work queued on `stream` has finished. */
int         clmsoil_sync(void *stream);

/* ---- moisture-form Richards step --------------------------------------
 * Replaces soilwm_moisture_form_solve! (soil_water_movement.jl:1498-1517):
 * one adaptive-substep moisture-form step of length `dtime` for every active
 * column. Upper BC must be FLUX; lower BC in {WATERTABLE, FLUX, ZERO_FLUX}. */
int clmsoil_soilwm_moisture_form_solve(const clmsoil_mf_args *a,
                                       const clmsoil_levels *lev, double dtime,
                                       const clmsoil_swm_config *cfg, void *stream);

/* The same step for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): no counterpart in
 * the reference. Results equal clmsoil_soilwm_moisture_form_solve's. Instead of
 * a->workspace (ignored) it uses a 256-byte scratch the library keeps in
 * device memory, one per GPU, set up by clmsoil_init(): the resident-grid work
 * counters (32-bit, so nc <= 2^34; larger nc returns CLMSOIL_ERR_SIZE) and the
 * call's error word. It allocates nothing. All calls on one GPU share that
 * scratch: calls that overlap on a GPU race (serialise them, e.g. on one
 * stream). */
int clmsoil_soilwm_moisture_form_solve_static(const clmsoil_mf_args *a,
                                              const clmsoil_levels *lev, double dtime,
                                              const clmsoil_swm_config *cfg, void *stream);

/* Replaces soilwater_moisture_form! (soil_water_movement.jl:1532-1601): the
 * solve above followed, when the lower BC uses an aquifer layer (WATERTABLE or
 * AQUIFER, :700-703), by the recharge pass `compute_qcharge!` (:1153-1217) with
 * the reference's all-zero smp/imped arguments (:1575-1598). */
int clmsoil_soilwater_moisture_form(const clmsoil_mf_args *a,
                                    const clmsoil_levels *lev, double dtime,
                                    const clmsoil_swm_config *cfg, void *stream);

/* ---- snow + soil heat step --------------------------------------------
 * Replaces the non-urban body of soil_temperature! (soil_temperature.jl:358-593):
 * thermal properties, heat-diffusion factor, Crank-Nicolson RHS and 5-band
 * matrix, banded solve (results equal to the reference's partially pivoted
 * elimination, band_diagonal.jl:87-170), h2osfc and layer phase change,
 * t_grnd and soil heat-flux diagnostics — one fused launch. */
int clmsoil_soil_temperature(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                             double dtime, const clmsoil_heat_config *cfg,
                             void *stream);

/* ---- batched solver operators -----------------------------------------
 * tridiagonal_multi! (tridiagonal.jl:86-142): a[c,j]u[c,j-1] + b u + c u[c,j+1]
 * = r from 1-based jtop[c] to nlevs, one system per masked column, 1e-30 pivot
 * guard. Arrays [ld, nlevs]; jtop Int64 [ncols]. */
int clmsoil_tridiagonal_multi(double *u, const double *a, const double *b,
                              const double *c, const double *r,
                              const int64_t *jtop, const uint8_t *mask,
                              int64_t ncols, int64_t nlevs, int64_t ld,
                              void *stream);

/* batched_band_solve! (band_diagonal.jl:179-189): per column the rows
 * jt = jtop[c]+nlevsno_off+1 (1-based) … jt+n-1, n = jbot[c]-jtop[c]+1, of the
 * band matrix bmatrix[c, b, row] = A[row, row+kl+1-b] ([ld, kl+ku+1, nlev]
 * column-major), right-hand side rvector [ld, nlev]; the solution is written to
 * tvector [ld, nlev]. Results equal the reference's densified elimination with
 * partial pivoting. Supports kl = ku = 1 or 2 and n <= 64. */
int clmsoil_batched_band_solve(double *tvector, const double *bmatrix,
                               const double *rvector, const int64_t *jtop,
                               const int64_t *jbot, const uint8_t *mask,
                               int32_t kl, int32_t ku, int32_t nlevsno_off,
                               int64_t nc, int64_t nlev, int64_t ld, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_H */
