/* clmsoil_surfhyd.h — the surface hydrology that produces the Richards step's
 * infiltration: flood water, runoff, surface water and infiltration, per column.
 *
 * clmsoil_surface_hydrology replaces the calls of clm_driver.jl:1939-2012 that
 * run between set_soil_water_fractions! and the root water uptake, in this
 * order (each a step bit; CLMSOIL_SH_ALL runs the driver's sequence, a single
 * bit stands in for that one reference function):
 *
 *   CLMSOIL_SH_FLOODC        set_floodc!                 soil_hydrology.jl:139-152, 430-442
 *   CLMSOIL_SH_SAT_EXCESS    saturated_excess_runoff!    sat_excess_runoff.jl:246-300, 329-365
 *   CLMSOIL_SH_QFLX_INPUTS   set_qflx_inputs!            soil_hydrology.jl:160-187, 456-485
 *   CLMSOIL_SH_INFIL_EXCESS  infiltration_excess_runoff! infilt_excess_runoff.jl:142-211
 *                            (compute_qinmax_hksat! :241-274 for qinmax_method 1)
 *   CLMSOIL_SH_ROUTE         route_infiltration_excess!  soil_hydrology.jl:199-225, 500-525
 *   CLMSOIL_SH_H2OSFC        update_h2osfc!              surface_water.jl:279-316, 357-374,
 *                                                        400-443, 462-520
 *   CLMSOIL_SH_INFILTRATION  infiltration!               soil_hydrology.jl:236-243, 538-552
 *   CLMSOIL_SH_SURF_RUNOFF   total_surface_runoff!       soil_hydrology.jl:250-301, 567-604
 *   CLMSOIL_SH_URBAN_POND    update_urban_ponding!       soil_hydrology.jl:316-338, 618-640
 *
 * Every step works on one column at a time, so one pass over the columns runs
 * them in order with results identical to the reference's step-by-step launches.
 * Masks: set_floodc! uses mask_nolakec, the urban runoff and ponding use
 * mask_urbanc, every other step mask_hydrologyc (NULL = all columns).
 */
#ifndef CLMSOIL_SURFHYD_H
#define CLMSOIL_SURFHYD_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLMSOIL_SH_FLOODC        0x001
#define CLMSOIL_SH_SAT_EXCESS    0x002
#define CLMSOIL_SH_QFLX_INPUTS   0x004
#define CLMSOIL_SH_INFIL_EXCESS  0x008
#define CLMSOIL_SH_ROUTE         0x010
#define CLMSOIL_SH_H2OSFC        0x020
#define CLMSOIL_SH_INFILTRATION  0x040
#define CLMSOIL_SH_SURF_RUNOFF   0x080
#define CLMSOIL_SH_URBAN_POND    0x100
#define CLMSOIL_SH_ALL           0x1ff

/* Method codes (sat_excess_runoff.jl:35-36, infilt_excess_runoff.jl:40-41) */
#define CLMSOIL_FSAT_TOPMODEL    1
#define CLMSOIL_FSAT_VIC         2
#define CLMSOIL_QINMAX_NONE      0
#define CLMSOIL_QINMAX_HKSAT     1

typedef struct clmsoil_surfhyd_params {
    int32_t steps;                    /* CLMSOIL_SH_* bits                              */
    int32_t fsat_method;              /* CLMSOIL_FSAT_*                                 */
    int32_t qinmax_method;            /* CLMSOIL_QINMAX_*                               */
    int32_t h2osfcflag;               /* soilhydrology.h2osfcflag (route step), 1       */
    int32_t h2osfcflag_update;        /* update_h2osfc!'s h2osfcflag keyword: the driver
                                         leaves it at its default, 1                    */
    int32_t crop_fsat_equals_zero;    /* saturated_excess_runoff! keywords, 0 and 0     */
    int32_t hillslope_fsat_equals_zero;
    int32_t reserved;
    double fff;                       /* TOPMODEL decay factor (1/m), from the params file */
    double e_ice;                     /* ice impedance of qinmax, 6.0                   */
    double pc, mu;                    /* surface water connectivity, 0.4 and 0.13889    */
    double pondmx_urban;              /* urban ponding depth (mm), 1.0                  */
} clmsoil_surfhyd_params;

typedef struct clmsoil_surfhyd_args {
    int64_t nc, ld;                   /* columns; leading dimension of the [ld, *] arrays */
    const uint8_t *mask_nolakec, *mask_hydrologyc, *mask_urbanc;   /* [nc]             */
    const int64_t *col_gridcell;      /* [nc] 1-based gridcell (set_floodc!)           */
    const int64_t *col_landunit;      /* [nc] 1-based landunit                         */
    const int64_t *lun_itype;         /* [nl] landunit type (ISTSOIL 1, ISTCROP 2, ...) */
    const int64_t *col_itype;         /* [nc] column type (urban 71..75)               */
    const int64_t *snl;               /* [nc] minus the number of snow layers          */
    const int64_t *col_cold;          /* [nc] downhill column or -9999 (hillslope only) */
    const uint8_t *col_urbpoi, *is_hillslope_column, *col_active;  /* [nc]             */
    const double *qflx_floodg;        /* [ng] gridcell flood water (mm/s)              */
    const double *frost_table, *zwt, *zwt_perched, *wtfact;        /* [nc] TOPMODEL fsat */
    const double *b_infil, *top_max_moist, *top_moist_limited;     /* [nc] VIC fsat (NULL
                                                                      unless fsat_method 2) */
    const double *qflx_rain_plus_snomelt, *qflx_snow_h2osfc, *qflx_liqevap_from_top_layer,
        *qflx_ev_soil, *qflx_ev_h2osfc;                            /* [nc] mm/s         */
    const double *frac_sno_eff, *frac_h2osfc, *frac_h2osfc_nosnow; /* [nc]              */
    const double *icefrac, *hksat;    /* [ld, nlevgrnd] (layers 1..3 read)             */
    const double *topo_slope;         /* [nc] degrees                                  */
    const double *h2osfc_thresh;      /* [nc] mm                                       */
    double *h2osoi_liq;               /* [ld, nlevsno+nlevgrnd]: the pond in the top soil
                                         layer of roof and impervious road columns     */
    double *h2osfc;                   /* [nc] mm, surface water (read/write)           */
    double *xs_urban;                 /* [nc] mm/s, urban excess over the pond cap     */
    double *qflx_floodc;              /* [nc] written by SH_FLOODC, read by later steps */
    double *fsat, *fcov, *qflx_sat_excess_surf;                    /* [nc]             */
    double *qflx_top_soil, *qflx_in_soil, *qflx_top_soil_to_h2osfc;
    double *qinmax, *qflx_infl_excess;
    double *qflx_in_soil_limited, *qflx_in_h2osfc, *qflx_infl_excess_surf;
    double *qflx_h2osfc_surf, *qflx_h2osfc_drain;
    double *qflx_infl;                /* [nc] the Richards step's infiltration (mm/s)  */
    double *qflx_surf;                /* [nc] total surface runoff (mm/s)              */
} clmsoil_surfhyd_args;

/* Pointers a step does not touch may be NULL. Returns CLMSOIL_ERR_ARG for an
 * unknown method or step bit, CLMSOIL_ERR_NULL for a missing array of a
 * requested step. */
int clmsoil_surface_hydrology(const clmsoil_surfhyd_args *a, const clmsoil_levels *lev,
                              double dtime, const clmsoil_surfhyd_params *p, void *stream);

/* Root water uptake: compute_effec_rootfrac_and_vert_tran_sink!
 * (soil_water_plant_sink.jl:341-451) with use_hydrstress = false, which runs
 * compute_effec_rootfrac_and_vert_tran_sink_default! (:115-141, kernel :48-81)
 * on the pervious-road, other and natural-vegetation columns of the hydrology
 * filter; the three groups partition that filter, so one pass over it is the
 * same. Per hydrology column c with patches patchi[c] .. patchi[c]+npatches[c]-1:
 *   temp      = sum over active patches of qflx_tran_veg_patch * wtcol
 *   rootr_col = sum over active patches of rootr_patch * qflx_tran_veg_patch * wtcol,
 *               divided by temp when temp != 0
 *   qflx_rootsoi_col = rootr_col * qflx_tran_veg_col
 * for soil layers 1..nlevsoi, summed in ascending patch order. This produces the
 * Richards step's root sink (clm_driver.jl:2016-2022). A column whose patch range
 * leaves 1..np is left unchanged and sets CLMSOIL_DEVERR_SUBGRID. */
typedef struct clmsoil_rootsink_args {
    int64_t nc, ldc;                  /* columns; leading dimension of the column arrays */
    int64_t np, ldp;                  /* patches; leading dimension of rootr_patch     */
    const uint8_t *mask_hydrologyc;   /* [nc] NULL = all columns                       */
    const int64_t *patchi;            /* [nc] first patch (1-based)                    */
    const int64_t *npatches;          /* [nc] number of patches                        */
    const uint8_t *active;            /* [np] patch active                             */
    const double *wtcol;              /* [np] patch weight in its column               */
    const double *rootr_patch;        /* [ldp, nlevgrnd] effective root fraction       */
    const double *qflx_tran_veg_patch;/* [np] transpiration (mm/s)                     */
    const double *qflx_tran_veg_col;  /* [nc] column transpiration (mm/s)              */
    double *rootr_col;                /* [ldc, nlevgrnd] (layers 1..nlevsoi written)   */
    double *qflx_rootsoi_col;         /* [ldc, nlevsoi] (mm/s)                         */
} clmsoil_rootsink_args;

int clmsoil_compute_effec_rootfrac_and_vert_tran_sink(const clmsoil_rootsink_args *a,
                                                      const clmsoil_levels *lev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_SURFHYD_H */
