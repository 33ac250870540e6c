/* clmsoil_day.h — a day of production steps captured as one HIP graph.
 *
 * No counterpart in the reference. CLM.jl's driver launches each kernel and
 * synchronises after it (src/infrastructure/kernels.jl:85-91, clm_drv_core!
 * src/main/clm_driver.jl:1801, 2049). A plan records `nsteps` steps on fixed
 * arrays as one graph. Each step is the heat step (clmsoil_soil_temperature)
 * followed by the Richards step (clmsoil_soilwater_moisture_form, which adds the
 * aquifer recharge pass when the lower boundary condition uses one).
 * Optionally the learned first-substep predictor runs before each Richards
 * step. One graph launch then replays the whole day.
 *
 * The graph holds the argument values and launch shapes (clmsoil_set_tuning &
 * co.) of the moment it was created; change them, and you must create a new
 * plan. All state stays in the caller's arrays; creating a plan allocates the
 * 256-byte workspace of its Richards steps (mf->workspace is not used), which
 * destroying it frees after the device is idle. Two launches of one plan, or of
 * plans sharing arrays, must not overlap; plans on disjoint arrays may run
 * concurrently on different streams.
 */
#ifndef CLMSOIL_DAY_H
#define CLMSOIL_DAY_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clmsoil_day_plan clmsoil_day_plan;

/* mlp_params / mlp_norm / dtsub0: all NULL for the plain controller, or all set
 * (see clmsoil_learned.h) to predict dtsub0 [nc] before every Richards step; the
 * plan then passes dtsub0 to the Richards step (mf->dtsub0 must be NULL). */
int clmsoil_day_plan_create(const clmsoil_heat_args *heat, const clmsoil_mf_args *mf,
                            const clmsoil_levels *lev, double dtime,
                            const clmsoil_heat_config *heat_cfg,
                            const clmsoil_swm_config *swm_cfg, const float *mlp_params,
                            const float *mlp_norm, double *dtsub0, int32_t nsteps,
                            clmsoil_day_plan **plan);
int clmsoil_day_plan_launch(clmsoil_day_plan *plan, void *stream);
int clmsoil_day_plan_destroy(clmsoil_day_plan *plan);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_DAY_H */
