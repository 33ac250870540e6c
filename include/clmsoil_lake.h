/* clmsoil_lake.h — the lake temperature solve: the tridiagonal system of the
 * snow + lake + soil column of a lake, assembled and solved in place.
 *
 * clmsoil_lake_temperature_solve replaces steps 6 and 7 of lake_temperature!
 * and the unpack after them (src/biogeophys/lake_temperature.jl:1205-1225):
 *   _lake_extcol_setup_kernel!    :632-664  extended column z, cv, solar source, T
 *   _lake_tkix_kernel!            :673-703  interface conductivities
 *   _lake_tridiag_factors_kernel! :705-722  dt/cv and the explicit fluxes
 *   _lake_tridiag_coeffs_kernel!  :724-755  Crank-Nicolson rows (CNFAC 0.5)
 *   tridiagonal_multi!            src/infrastructure/tridiagonal.jl:86-142
 *   _lake_unpack_kernel!          :757-785  back into t_soisno and t_lake
 * The extended column runs over snow layers snl+1..0, lake layers 1..nlevlak
 * and soil layers 1..nlevgrnd; the solve starts at its top active level. Only
 * t_soisno (active levels) and t_lake change; the reference's scratch arrays
 * (zx, cvx, phix, tkix, factx, fnx, a, b, c1, r) are not materialised.
 */
#ifndef CLMSOIL_LAKE_H
#define CLMSOIL_LAKE_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clmsoil_lake_solve_args {
    int64_t nc, ld;                   /* columns; leading dimension of every [ld, *] array */
    int64_t nlevlak;                  /* lake layers (10, or 25 with extra layers)      */
    const uint8_t *mask_lakec;        /* [nc] NULL = all columns                        */
    const int64_t *snl;               /* [nc] minus the number of snow layers           */
    const double *z, *cv, *tk;        /* [ld, nlevsno+nlevgrnd] node depth (m), heat
                                         capacity (J/m2/K), interface conductivity (W/m/K)
                                         of the snow and soil layers                     */
    const double *sabg_lyr_col;       /* [ld, nlevsno+1] solar absorbed per snow layer  */
    const double *z_lake, *dz_lake;   /* [ld, nlevlak] m                                */
    const double *cv_lake, *tk_lake;  /* [ld, nlevlak]                                  */
    const double *phi;                /* [ld, nlevlak] solar source in the lake (W/m2)  */
    const double *phi_soil;           /* [nc] solar reaching the top soil layer (W/m2)  */
    const double *fin;                /* [nc] heat flux into the top of the column      */
    const double *tktopsoillay;       /* [nc] top soil layer conductivity               */
    double *t_soisno;                 /* [ld, nlevsno+nlevgrnd] K (read/write)          */
    double *t_lake;                   /* [ld, nlevlak] K (read/write)                   */
} clmsoil_lake_solve_args;

/* snl outside -nlevsno..0 leaves the column unchanged and sets
 * CLMSOIL_DEVERR_SNL; nlevsno + nlevlak + nlevgrnd may be at most 64. */
int clmsoil_lake_temperature_solve(const clmsoil_lake_solve_args *a, const clmsoil_levels *lev,
                                   double dtime, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_LAKE_H */
