/* clmsoil_ghf.h — ground heat flux and its temperature derivative, the input of
 * the heat step (clmsoil_soil_temperature), and the soil energy balance error
 * computed after it.
 *
 * Replaces compute_ground_heat_flux_and_deriv! (src/biogeophys/soil_temperature.jl:
 * 1200-1263) and the kernels it launches: longwave emission per column (:944-960),
 * the per-patch net ground heat flux for vegetated/bare and urban patches
 * (:1008-1085) with its patch-to-column sums (compute_gnet_patch!, :1087-1130,
 * scatter_add_1d!, src/infrastructure/kernels.jl:135-170), and the top-layer flux and
 * absorbed solar per snow layer (_gnet_snicar_kernel!, :1135-1196).
 *
 * The reference adds each patch's contribution to its column with an atomic add.
 * Here the column sums are formed by the column itself, over its patches in
 * ascending patch order, which is the order of the reference's CPU path: the result
 * is deterministic and bit-identical to it. This relies on the subgrid layout CLM
 * builds: the patches of column c are the contiguous range patchi[c]..patchf[c]
 * (ColumnData, src/types/column.jl:39-40), each with patch_column[p] == c. A patch
 * in that range with another column index, or a column with patchi[c] > patchf[c],
 * contributes nothing and sets CLMSOIL_DEVERR_SUBGRID in the device error word.
 *
 * Index arrays are the reference's: 1-based Int64 (ISPVAL = -9999 for none).
 * Column matrices are [ldc, ...], patch matrices [ldp, ...], column-major.
 */
#ifndef CLMSOIL_GHF_H
#define CLMSOIL_GHF_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

/* urban_params.jl building_temp_method */
#define CLMSOIL_BUILDING_TEMP_SIMPLE 0
#define CLMSOIL_BUILDING_TEMP_PROG   1

typedef struct clmsoil_ghf_args {
    int64_t nc, ldc;             /* columns, leading dimension of column matrices     */
    int64_t np, ldp;             /* patches, leading dimension of sabg_lyr_patch      */
    int64_t nl;                  /* landunits                                         */
    /* subgrid structure (1-based) */
    const int64_t *patchi, *patchf;      /* [nc] ColumnData.patchi / patchf          */
    const int64_t *patch_column;         /* [np] PatchData.column                    */
    const int64_t *patch_landunit;       /* [np] PatchData.landunit                  */
    const int64_t *col_itype;            /* [nc] ColumnData.itype                    */
    const int64_t *snl;                  /* [nc] ColumnData.snl                      */
    const uint8_t *urbpoi;               /* [nl] LandunitData.urbpoi                 */
    const uint8_t *mask_nolakec;         /* [nc]                                     */
    const uint8_t *mask_nolakep;         /* [np]                                     */
    const int64_t *frac_veg_nosno;       /* [np] CanopyStateData.frac_veg_nosno_patch */
    /* column inputs */
    const double *emg, *t_grnd, *t_h2osfc, *forc_lwrad, *htvp, *frac_sno_eff;  /* [nc] */
    const double *t_soisno;              /* [ldc, nlevsno+nlevgrnd]                  */
    /* patch inputs [np] */
    const double *sabg, *sabg_soil, *sabg_snow, *dlrad, *cgrnd;
    const double *eflx_sh_grnd, *eflx_sh_snow, *eflx_sh_soil, *eflx_sh_h2osfc;
    const double *qflx_evap_soi, *qflx_ev_snow, *qflx_ev_soil, *qflx_ev_h2osfc;
    const double *eflx_lwrad_net, *qflx_tran_veg, *wtcol;
    const double *sabg_lyr_patch;        /* [ldp, nlevsno+1]                         */
    /* landunit inputs [nl] */
    const double *eflx_wasteheat_lun, *eflx_ventilation_lun, *eflx_heat_from_ac_lun;
    const double *eflx_traffic_lun, *wtlunit_roof;
    /* patch outputs [np] (written for masked patches, as the reference) */
    double *eflx_gnet, *dgnetdT, *sabg_chk;
    double *eflx_wasteheat, *eflx_ventilation, *eflx_heat_from_ac, *eflx_traffic;
    double *eflx_anthro;
    /* column outputs [nc] (every column written) */
    double *hs_top, *hs_top_snow, *hs_soil, *hs_h2osfc, *dhsdT;
    double *sabg_lyr_col;                /* [ldc, nlevsno+1]                         */
    int32_t building_temp_method;        /* CLMSOIL_BUILDING_TEMP_*                  */
    int32_t reserved;
} clmsoil_ghf_args;

/* lev: nlevsno and nlevgrnd give the layer counts of t_soisno and the snow-layer
 * matrices; nlevsoi is not used. */
int clmsoil_ground_heat_flux(const clmsoil_ghf_args *a, const clmsoil_levels *lev,
                             void *stream);

/* ---- soil energy balance error --------------------------------------------
 * Replaces the errsoi part of soil_fluxes! (src/biogeophys/soil_fluxes.jl:380-447,
 * 499-525): per patch, the ground heat flux minus the phase-change and
 * surface-water heat terms minus the heat stored in every snow and soil layer,
 * sum_j (t_soisno - t_ssbef)/fact (snow layers weighted by frac_sno_eff; wall and
 * roof columns use their own layer count nlevurb); then the weighted patch-to-
 * column average over masked columns. As above, each column walks its own patch
 * range patchi..patchf in ascending order (no atomics, the order of the
 * reference's CPU path) and writes the errsoi of those patches. */
typedef struct clmsoil_seb_args {
    int64_t nc, ldc, np;
    int32_t nlevurb;                     /* layers of wall and roof columns           */
    int32_t reserved;
    const int64_t *patchi, *patchf;      /* [nc] 1-based patch range of each column   */
    const int64_t *patch_column;         /* [np] 1-based                              */
    const int64_t *col_itype, *snl;      /* [nc]                                      */
    const uint8_t *mask_nolakec;         /* [nc]                                      */
    const uint8_t *mask_nolakep, *mask_urbanp;   /* [np]                              */
    const double *eflx_soil_grnd, *wtcol;        /* [np] W/m2, weight in column       */
    const double *xmf, *xmf_h2osfc;              /* [nc] phase-change heat, W/m2      */
    const double *frac_h2osfc, *t_h2osfc, *t_h2osfc_bef, *c_h2osfc;   /* [nc]        */
    const double *eflx_h2osfc_to_snow, *eflx_building_heat_errsoi;    /* [nc] W/m2   */
    const double *frac_sno_eff;                  /* [nc]                              */
    /* [ldc, nlevsno + max(nlevgrnd, nlevurb)] */
    const double *t_soisno, *t_ssbef, *fact;
    double *errsoi_patch;                /* [np] (patches that no pass touches keep their value) */
    double *errsoi_col;                  /* [nc] written for masked columns           */
} clmsoil_seb_args;

int clmsoil_soil_energy_balance(const clmsoil_seb_args *a, const clmsoil_levels *lev,
                                double dtime, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_GHF_H */
