/* clmsoil_zd09.h — the Zeng-Decker 2009 soil water step (CLM.jl's default
 * soilwater_movement_method, with the BC_AQUIFER lower boundary).
 *
 * Replaces soilwater_zengdecker2009! (src/biogeophys/soil_water_movement.jl:1616-1764)
 * and its kernels (:97-643): ice fractions and geometry in mm, the layer above
 * the water table and the water content there, the equilibrium water content
 * and suction, hydraulic conductivity and suction, the (nlevsoi+1)-row
 * tridiagonal system with the aquifer layer, its solve (tridiagonal_multi!,
 * tridiagonal.jl:86-127), the water update, aquifer recharge and the water
 * deficit, in one launch. Every soil column uses all nlevsoi layers, as in the
 * reference (this method has no bedrock).
 */
#ifndef CLMSOIL_ZD09_H
#define CLMSOIL_ZD09_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clmsoil_zd_args {
    int64_t nc;
    int64_t ld;
    const uint8_t *mask;         /* [nc] NULL = all active                        */
    /* geometry */
    const double *z, *dz;        /* [ld, nlevsno+nlevgrnd] m                      */
    const double *zi;            /* [ld, nlevsno+nlevgrnd+1] m                    */
    /* soil properties */
    const double *watsat, *hksat, *bsw, *sucsat;   /* [ld, nlevgrnd]               */
    const double *smpmin;        /* [nc] mm                                       */
    /* state and forcing */
    double *h2osoi_liq;          /* [ld, nlevsno+nlevgrnd] kg/m2 (read/write)     */
    const double *h2osoi_ice;    /* [ld, nlevsno+nlevgrnd] kg/m2                  */
    const double *h2osoi_vol;    /* [ld, nlevgrnd] m3/m3                          */
    const double *t_soisno;      /* [ld, nlevsno+nlevgrnd] K                      */
    const double *zwt;           /* [nc] m                                        */
    const double *qflx_infl;     /* [nc] mm/s                                     */
    const double *qflx_rootsoi;  /* [ld, nlevsoi] mm/s                            */
    /* outputs */
    double *icefrac;             /* [ld, nlevgrnd] (layers 1..nlevsoi written)    */
    double *smp_l, *hk_l;        /* [ld, nlevgrnd] (layers 1..nlevsoi written)    */
    double *qcharge;             /* [nc] mm/s                                     */
    double *qflx_deficit;        /* [nc] mm                                       */
} clmsoil_zd_args;

/* cfg: soilwater_movement_method must be CLMSOIL_ZENGDECKER_2009; only e_ice is
 * read (the reference's soil_water! dispatch, :1780-1810). */
int clmsoil_soilwater_zengdecker2009(const clmsoil_zd_args *a, const clmsoil_levels *lev,
                                     double dtime, const clmsoil_swm_config *cfg,
                                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_ZD09_H */
