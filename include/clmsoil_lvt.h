/* clmsoil_lvt.h — vertical transport of the soil organic matter pools: the
 * advection-diffusion of every decomposing carbon and nitrogen pool down the
 * soil column, one tridiagonal system per (column, pool, tracer).
 *
 * clmsoil_litter_vert_transp replaces litter_vert_transp!
 * (src/biogeochem/litter_vert_transp.jl:479-639) with use_soil_matrixcn = false:
 *   _lvt_coef_kernel!   :93-136   cryoturbation / advection coefficients per column
 *   _lvt_column_kernel! :142-471  per (tracer, pool): Patankar power-law fluxes,
 *                                 the tridiagonal system over levels 0..nlevdecomp+1,
 *                                 its Thomas solve (no pivot guard), the transport
 *                                 tendency and the bedrock-leak write-back
 * for every tracer type given (C, N, and C13 / C14 when present) and every pool.
 * Arrays are the reference's [c, j, s] matrices, column-major with leading
 * dimension ld and level dimension nlevdecomp_full.
 */
#ifndef CLMSOIL_LVT_H
#define CLMSOIL_LVT_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLMSOIL_LVT_MAXPOOLS 16
#define CLMSOIL_LVT_MAXTYPES 4

typedef struct clmsoil_lvt_params {
    int32_t nlevdecomp;               /* vertically resolved decomposition levels (<= 62) */
    int32_t nlevdecomp_full;          /* level dimension of the pool arrays             */
    int32_t npools;                   /* ndecomp_pools (<= CLMSOIL_LVT_MAXPOOLS)        */
    int32_t ntypes;                   /* tracer types present, 1..CLMSOIL_LVT_MAXTYPES  */
    int32_t spinup_state;             /* 0, or >= 1 for accelerated spinup              */
    int32_t use_soil_matrixcn;        /* must be 0 (the matrix-CN path is not built)    */
    double som_diffus;                /* LitterVertTranspParams.som_diffus (m2/s)       */
    double cryoturb_diffusion_k;      /* (m2/s)                                         */
    double max_altdepth_cryoturbation;/* (m)                                            */
    double som_adv_flux;              /* (m/s)                                          */
    double max_depth_cryoturb;        /* (m)                                            */
    double spinup_factor[CLMSOIL_LVT_MAXPOOLS];  /* per pool                            */
    uint8_t is_cwd[CLMSOIL_LVT_MAXPOOLS];        /* per pool: coarse woody debris       */
} clmsoil_lvt_params;

typedef struct clmsoil_lvt_args {
    int64_t nc, ld;                   /* columns; leading dimension of column arrays    */
    const uint8_t *mask_bgc_soilc;    /* [nc] NULL = all columns                        */
    const int64_t *nbedrock;          /* [nc] 1..nlevdecomp                             */
    const int64_t *col_gridcell;      /* [nc] 1-based (spinup only)                     */
    const double *latdeg;             /* [ngridcells] (spinup only)                     */
    const double *altmax, *altmax_lastyear;   /* [nc] active layer thickness (m)       */
    const double *zsoi;               /* [nlevdecomp] node depths (m)                   */
    const double *dzsoi_decomp;       /* [nlevdecomp] layer thicknesses (m)             */
    const double *zisoi;              /* [nlevdecomp + 2] interfaces, zisoi[0] = 0      */
    double *som_diffus_coef;          /* [ld, nlevdecomp + 1] written                   */
    double *som_adv_coef;             /* [ld, nlevdecomp + 1] written                   */
    double *conc[CLMSOIL_LVT_MAXTYPES];         /* [ld, nlevdecomp_full, npools] g/m3, updated */
    const double *source[CLMSOIL_LVT_MAXTYPES]; /* same shape: the step's source/sink   */
    double *tendency[CLMSOIL_LVT_MAXTYPES];     /* same shape: transport tendency       */
} clmsoil_lvt_args;

/* A column with nbedrock outside 1..nlevdecomp is left unchanged and sets
 * CLMSOIL_DEVERR_NBEDROCK. use_soil_matrixcn != 0 returns CLMSOIL_ERR_METHOD. */
int clmsoil_litter_vert_transp(const clmsoil_lvt_args *a, double dtime,
                               const clmsoil_lvt_params *p, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_LVT_H */
