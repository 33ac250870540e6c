/* clmsoil_adjoint.h — reverse-mode (vector-Jacobian) entry points of the path.
 *
 * CLM.jl differentiates this path with Enzyme: custom rules for the solvers
 * (src/infrastructure/enzyme_rules.jl:72-107 band_solve!, :152-202
 * tridiagonal_solve!) and Enzyme's own reverse pass through the Richards
 * kernel. These entry points are hand-written adjoints of the same functions.
 * All of them accumulate (+=) into the gradient arrays they are given, as
 * Enzyme's shadows do, and run asynchronously on `stream`.
 */
#ifndef CLMSOIL_ADJOINT_H
#define CLMSOIL_ADJOINT_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Accepted substeps one adjoint call can hold per column. An 1800 s step with a
 * 60 s minimum substep needs at most ~40; a column beyond the limit sets
 * CLMSOIL_DEVERR_TAPE and is left out of the gradient. */
#define CLMSOIL_ADJOINT_MAX_SUBSTEPS 64

/* Workspace (bytes) that clmsoil_soilwm_moisture_form_adjoint needs for
 * leading dimension ld. */
int64_t clmsoil_adjoint_workspace_bytes(int64_t ld);

/* Adjoint of one moisture-form Richards step (clmsoil_soilwm_moisture_form_solve,
 * soil_water_movement.jl:1250-1489), for lower BC FLUX or ZERO_FLUX.
 *
 * `a` describes the state at the START of the step (h2osoi_liq is read, not
 * written; smp_l, hk_l, qcharge and nsubsteps are neither read nor written and
 * may be NULL). The call replays the step with the same arithmetic, so the
 * accept/reject sequence is the primal's; the substep lengths are treated as
 * constants (they are piecewise constant in the inputs).
 *
 *   liq_bar      [ld, nlevsno+nlevgrnd] in: dL/dh2osoi_liq after the step;
 *                out: dL/dh2osoi_liq before the step (soil layers 1..nbedrock)
 *   qcharge_bar  [nc] dL/dqcharge of this step
 *   grad_bsw, grad_hksat, grad_watsat  [ld, nlevgrnd] per-layer dL/dp (+=), each
 *                optional (NULL)
 *   grad_scale   [3, ld] optional (+=): sum over layers of p_j dL/dp_j for
 *                p = bsw, hksat, watsat — the gradient with respect to a
 *                per-column multiplier of each parameter, at multiplier 1.
 */
int clmsoil_soilwm_moisture_form_adjoint(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                         double dtime, const clmsoil_swm_config *cfg,
                                         double *liq_bar, const double *qcharge_bar,
                                         double *grad_bsw, double *grad_hksat,
                                         double *grad_watsat, double *grad_scale,
                                         void *workspace, int64_t workspace_bytes,
                                         void *stream);

/* Adjoint of one heat step (clmsoil_soil_temperature: soil_temperature.jl:358-593
 * with phase_change_beta!, :1929-2289) on soil columns that may freeze and thaw:
 * no snow layers, no standing surface water, no unresolved snow and no excess
 * ice (snl == 0, frac_h2osfc == 0, h2osno_no_layers == 0, excess_ice == 0).
 * Other columns, and columns whose elimination would interchange rows, set
 * CLMSOIL_DEVERR_ADJ and are left unchanged.
 *
 * `a` describes the state at the START of the step (t_soisno, h2osoi_liq and
 * h2osoi_ice are read, nothing is written; the output fields of `a` are not
 * used and may be NULL). The call replays the step with the heat kernel's
 * arithmetic, so the melt/freeze labels are the primal's; they, and the max/min
 * selections, are piecewise constant in the inputs.
 *
 *   t_bar, liq_bar, ice_bar  [ld, nlevsno+nlevgrnd] in: dL/d(t_soisno,
 *                h2osoi_liq, h2osoi_ice) after the step; out: before the step
 *                (soil levels; snow levels pass through unchanged)
 *   grad_bsw, grad_watsat    [ld, nlevgrnd] per-layer dL/dp (+=), optional
 *   grad_scale   [3, ld] optional (+=): rows 0 and 2 get sum_j p_j dL/dp_j for
 *                p = bsw and watsat (row 1, hksat, is untouched: the heat step
 *                does not read it) — the layout of the Richards adjoint's.
 * Reference: Enzyme's reverse pass through soil_temperature! in the reverse
 * driver (driver_reverse.jl:169-190) with the band_solve! rule
 * (enzyme_rules.jl:72-107). */
int clmsoil_soil_temperature_adjoint(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                                     double dtime, double *t_bar, double *liq_bar,
                                     double *ice_bar, double *grad_bsw, double *grad_watsat,
                                     double *grad_scale, void *stream);

/* Reverse of tridiagonal_multi! (tridiagonal.jl:86-142), the batched form of the
 * tridiagonal_solve! rule (enzyme_rules.jl:152-202): with u the forward
 * solution, lambda = A^-T ubar over rows jtop..nlevs; rbar += lambda,
 * abar -= lambda u[j-1], bbar -= lambda u, cbar -= lambda u[j+1]; ubar is
 * consumed (zeroed) over jtop..nlevs. Any of abar..rbar may be NULL. */
int clmsoil_tridiagonal_multi_vjp(double *ubar, const double *u, const double *a,
                                  const double *b, const double *c, const int64_t *jtop,
                                  const uint8_t *mask, int64_t ncols, int64_t nlevs,
                                  int64_t ld, double *abar, double *bbar, double *cbar,
                                  double *rbar, void *stream);

/* Reverse of batched_band_solve! (band_diagonal.jl:87-189), in the form of the
 * band_solve! rule (enzyme_rules.jl:72-107): with t the forward solution,
 * lambda = A^-T tbar (partially pivoted elimination of the transposed band
 * system); rvector_bar += lambda; bmatrix_bar -= lambda t^T on the band entries
 * the forward solve reads; tbar is consumed (zeroed) on the system rows.
 * Same layouts and row convention as clmsoil_batched_band_solve. */
int clmsoil_batched_band_solve_vjp(double *tbar, const double *tvector, const double *bmatrix,
                                   const int64_t *jtop, const int64_t *jbot,
                                   const uint8_t *mask, int32_t kl, int32_t ku,
                                   int32_t nlevsno_off, int64_t nc, int64_t nlev, int64_t ld,
                                   double *bmatrix_bar, double *rvector_bar, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_ADJOINT_H */
