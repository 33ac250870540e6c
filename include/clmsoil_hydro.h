/* clmsoil_hydro.h — the soil hydrology either side of the Richards step: the ice
 * fraction and effective porosity it reads, and the water table it moves.
 *
 * clmsoil_set_soil_water_fractions replaces set_soil_water_fractions!
 * (src/biogeophys/soil_hydrology.jl:394-415, kernel :110-136): per soil layer,
 * the volumetric ice content (excess ice included), the effective porosity and
 * the ice fraction that the Richards step's impedance uses.
 *
 * clmsoil_water_table replaces water_table! (soil_hydrology.jl:849-902, kernel
 * :676-826), which the reference runs after soil_water! when the lower boundary
 * uses an aquifer layer (clm_driver.jl:2066-2085): the aquifer recharge qcharge
 * of the step moves the water table through the layers' specific yield (or
 * fills the aquifer store wa when the table is below the soil column), then the
 * frost table and the perched water table are located.
 *
 * Without an aquifer layer (the CLM5 default: moisture form, zero-flux bottom)
 * the reference instead diagnoses the water tables from the moisture profile
 * after soil_water! (clm_driver.jl:2076-2092): perched_water_table!
 * (soil_hydrology.jl:1604-1653, kernel :1522-1587), theta_based_water_table!
 * (:1732-1765, kernel :1670-1715), then clamp_zwt_to_bedrock!
 * (src/infrastructure/kernels.jl:209-227), which also follows water_table!.
 * clmsoil_perched_water_table, clmsoil_theta_based_water_table and
 * clmsoil_clamp_zwt_to_bedrock replace them one for one;
 * clmsoil_water_table_moisture_form runs the three in that order in one pass
 * over each column.
 */
#ifndef CLMSOIL_HYDRO_H
#define CLMSOIL_HYDRO_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clmsoil_fractions_args {
    int64_t nc, ld;
    const uint8_t *mask;           /* [nc] NULL = all active                       */
    const double *watsat;          /* [ld, nlevgrnd]                               */
    const double *dz;              /* [ld, nlevsno+nlevgrnd] m                     */
    const double *h2osoi_ice;      /* [ld, nlevsno+nlevgrnd] kg/m2                 */
    const double *excess_ice;      /* [ld, nlevgrnd] kg/m2; NULL = none            */
    double *eff_porosity;          /* [ld, nlevgrnd] (layers 1..nlevsoi written)   */
    double *icefrac;               /* [ld, nlevgrnd] (layers 1..nlevsoi written)   */
} clmsoil_fractions_args;

int clmsoil_set_soil_water_fractions(const clmsoil_fractions_args *a,
                                     const clmsoil_levels *lev, void *stream);

typedef struct clmsoil_water_table_args {
    int64_t nc, ld;
    const uint8_t *mask;           /* [nc] NULL = all active                       */
    const double *dz, *z;          /* [ld, nlevsno+nlevgrnd] m                     */
    const double *zi;              /* [ld, nlevsno+nlevgrnd+1] m                   */
    const double *t_soisno;        /* [ld, nlevsno+nlevgrnd] K                     */
    const double *watsat, *sucsat, *bsw;   /* [ld, nlevgrnd]                       */
    const double *qcharge;         /* [nc] mm/s, the step's aquifer recharge        */
    const double *h2osoi_liq, *h2osoi_ice; /* [ld, nlevsno+nlevgrnd] kg/m2          */
    double *zwt;                   /* [nc] m (read/write)                          */
    double *wa;                    /* [nc] mm, aquifer water (read/write)           */
    double *zwt_perched;           /* [nc] m (written when the frost table is)      */
    double *frost_table;           /* [nc] m (written when the frost table is)      */
    double *h2osoi_vol;            /* [ld, nlevgrnd] (layers visited written)       */
    double *qflx_drain, *qflx_rsub_sat, *qflx_drain_perched;   /* [nc] set to 0    */
} clmsoil_water_table_args;

/* aq_sp_yield_min: minimum aquifer specific yield (SoilHydrologyParams, 0.01 by
 * default, soil_hydrology.jl:35). recompute_frost_table: 1 as in the reference's
 * default; 0 leaves frost_table, zwt_perched and h2osoi_vol untouched. */
int clmsoil_water_table(const clmsoil_water_table_args *a, const clmsoil_levels *lev,
                        double dtime, double aq_sp_yield_min, int recompute_frost_table,
                        void *stream);

typedef struct clmsoil_theta_wt_args {
    int64_t nc, ld;
    const uint8_t *mask;           /* [nc] NULL = all active (hydrology columns)    */
    const int64_t *nbedrock;       /* [nc] 1..nlevsoi (theta-based, clamp)          */
    const double *dz, *z;          /* [ld, nlevsno+nlevgrnd] m                      */
    const double *zi;              /* [ld, nlevsno+nlevgrnd+1] m                    */
    const double *t_soisno;        /* [ld, nlevsno+nlevgrnd] K (perched)            */
    const double *watsat;          /* [ld, nlevgrnd]                                */
    const double *h2osoi_liq, *h2osoi_ice; /* [ld, nlevsno+nlevgrnd] kg/m2          */
    double *zwt;                   /* [nc] m: read by perched, written by
                                      theta-based, lowered to bedrock by the clamp  */
    double *zwt_perched;           /* [nc] m (perched)                              */
    double *frost_table;           /* [nc] m (perched)                              */
    double *h2osoi_vol;            /* [ld, nlevgrnd] (layers visited written)       */
} clmsoil_theta_wt_args;

/* Each needs the fields its reference function reads (perched: dz, z, zi,
 * t_soisno, watsat, h2osoi_liq/ice, zwt, zwt_perched, frost_table, h2osoi_vol;
 * theta-based: nbedrock, dz, z, zi, watsat, h2osoi_liq/ice, zwt, h2osoi_vol;
 * clamp: nbedrock, zi, zwt); the others may be NULL. A column whose nbedrock is
 * outside 1..nlevsoi is left unchanged by the theta-based step and the clamp and
 * sets CLMSOIL_DEVERR_NBEDROCK in the device error word. */
int clmsoil_perched_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                void *stream);
int clmsoil_theta_based_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                    void *stream);
int clmsoil_clamp_zwt_to_bedrock(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev,
                                 void *stream);
int clmsoil_water_table_moisture_form(const clmsoil_theta_wt_args *a,
                                      const clmsoil_levels *lev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_HYDRO_H */
