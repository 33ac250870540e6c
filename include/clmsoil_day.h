/* clmsoil_day.h — a day of production steps captured as one HIP graph.
 *
 * No counterpart in the reference. CLM.jl's driver launches each kernel and
 * synchronises after it (src/infrastructure/kernels.jl:85-91, clm_drv_core!
 * src/main/clm_driver.jl:1801, 2049). A plan records `nsteps` steps on fixed
 * arrays as one graph. Each step is the heat step (clmsoil_soil_temperature)
 * followed by the Richards step (clmsoil_soilwater_moisture_form, which adds the
 * aquifer recharge pass when the lower boundary condition uses one).
 * Optionally the learned first-substep predictor sets each column's first
 * substep, inside the Richards kernel or as a launch before it. One graph
 * launch then replays the whole day.
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

/* mlp_params / mlp_norm / dtsub0 (see clmsoil_learned.h): all NULL for the plain
 * controller; mlp_params and mlp_norm set with dtsub0 NULL for the network fused
 * into every Richards step (clmsoil_soilwm_moisture_form_solve_learned); all
 * three set to predict dtsub0 [nc] in a launch of its own before every Richards
 * step, which then reads it (same results as the fused form). mf->dtsub0 must
 * be NULL whenever the network is used. */
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
