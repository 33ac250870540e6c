/* clmsoil_ch4.h — methane and oxygen transport in the soil column: the
 * reaction-diffusion step of CH4 and O2 for one fraction of every column.
 *
 * clmsoil_ch4_tran replaces ch4_tran! (src/biogeochem/methane.jl:2005-2092):
 *   meth_khcc! / _meth_khcc_kernel!  :101-129    Henry's-law partition per level
 *   _ch4tran_column_kernel!          :1663-2003  O2/CH4 competition, sources,
 *        diffusivities, the Crank-Nicolson system of each species over the ground
 *        node and nlevsoi levels, its Thomas solve, the surface fluxes, the
 *        negative-CH4 clip, the O2 bounds and the CH4 balance correction
 * for one call of ch4! (methane.jl:2413): the unsaturated fraction (sat 0), the
 * saturated fraction (sat 1, jwt 0) or the lakes (lake, sat 1, jwt 0). The state
 * arrays are that fraction's (the *_sat_col or *_unsat_col set). Every [ld, *]
 * array holds soil levels 1..nlevsoi only, column-major with leading dimension ld
 * (the driver passes the soil-layer views, clm_driver.jl:2792-2830); a caller
 * holding snow + soil arrays passes the pointer of soil level 1.
 */
#ifndef CLMSOIL_CH4_H
#define CLMSOIL_CH4_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clmsoil_ch4_tran_params {
    int32_t nlevsoi;                  /* soil levels of the transport (1..60)          */
    int32_t sat;                      /* 0 unsaturated fraction, 1 saturated fraction  */
    int32_t lake;                     /* nonzero for the lake call                     */
    int32_t ch4frzout;                /* CH4VarCon.ch4frzout                           */
    int32_t use_aereoxid_prog;        /* CH4VarCon.use_aereoxid_prog                   */
    int32_t reserved;
    double satpow;                    /* CH4Params.satpow                              */
    double scale_factor_gasdiff;      /* CH4Params.scale_factor_gasdiff                */
    double scale_factor_liqdiff;      /* CH4Params.scale_factor_liqdiff                */
    double capthick;                  /* CH4Params.capthick (mm)                       */
    double aereoxid;                  /* CH4Params.aereoxid                            */
    double om_frac_sf;                /* CH4Params.om_frac_sf                          */
    double organic_max;               /* organic matter density of peat (kg/m3)        */
} clmsoil_ch4_tran_params;

typedef struct clmsoil_ch4_tran_args {
    int64_t nc, ld;                   /* columns; leading dimension of [ld, nlevsoi]   */
    int64_t ldg;                      /* leading dimension of c_atm (gridcells)        */
    const uint8_t *mask;              /* [nc] the call's column filter; NULL = all     */
    const int64_t *col_gridcell;      /* [nc] 1-based                                  */
    const int64_t *jwt;               /* [nc] 0..nlevsoi: deepest unsaturated level    */
    const double *c_atm;              /* [ldg, 3] atmospheric CH4, O2, CO2 (mol/m3)    */
    /* forcing, [ld, nlevsoi] */
    const double *watsat, *h2osoi_vol, *h2osoi_liq, *h2osoi_ice, *bsw, *cellorg;
    const double *t_soisno, *dz;
    /* forcing, [nc] */
    const double *h2osfc, *t_grnd, *t_h2osfc, *frac_h2osfc;
    double *grnd_ch4_cond;            /* [nc] in: patch-aggregated ground conductance
                                         (m/s); out: CH4's, with the ponding resistance */
    /* the fraction's state, [ld, nlevsoi] (mol/m3/s, mol/m3) */
    const double *ch4_prod_depth, *o2_aere_depth;
    double *o2_decomp_depth, *o2_oxid_depth;          /* scaled by the O2 stress      */
    double *ch4_oxid_depth, *ch4_aere_depth, *ch4_ebul_depth;  /* scaled in place    */
    double *o2stress, *ch4stress;     /* written                                       */
    double *conc_ch4, *conc_o2;       /* updated                                       */
    /* the fraction's column fluxes, [nc] (mol/m2/s) and mol/m2: written */
    double *ch4_ebul_total, *ch4_surf_aere, *ch4_surf_ebul, *ch4_surf_diff;
} clmsoil_ch4_tran_args;

/* A column with jwt outside 0..nlevsoi is left unchanged and sets
 * CLMSOIL_DEVERR_JWT. */
int clmsoil_ch4_tran(const clmsoil_ch4_tran_args *a, double dtime,
                     const clmsoil_ch4_tran_params *p, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_CH4_H */
