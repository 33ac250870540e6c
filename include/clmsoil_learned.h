/* clmsoil_learned.h — learned first-substep controller for the moisture-form
 * Richards step.
 *
 * None of these entry points has a counterpart in the reference: CLM.jl starts
 * every column's substep loop at dtsub = dtime (soil_water_movement.jl:1259). A
 * small MLP predicts each column's first substep instead, and the result goes to
 * clmsoil_mf_args.dtsub0. The plain controller's error test
 * (soil_water_movement.jl:1441-1463) still accepts or rejects every substep, so
 * a bad prediction costs iterations, never accuracy. A NaN prediction falls back
 * to dtime.
 *
 * Network: 24 inputs -> 64 GELU -> 64 GELU -> 1, FP32 weights and FP32
 * accumulation, 5,825 parameters packed as
 *   W1[64][24] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3[1].
 * GELU is the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
 *
 * Inputs (feature-major [24][ldx] FP32):
 *   0..19  node saturation s = clamp(max(liq,1e-6)/(dz*1000)/watsat, 0.01, 1) of
 *          soil layers 1..20 (0 below nbedrock);
 *   20     log(hksat) of layer 1 (mm/s);  21  bsw of layer 1;
 *   22     qflx_infl (mm/s);               23  remaining step length (s) = dtime.
 * A normalisation block norm[50] = { mean[24], scale[24], y_mean, y_scale }
 * standardises x -> (x - mean) / scale; the network output y gives
 * log(dtsub0 / dtime) = y_mean + y_scale * y.
 *
 * All calls are asynchronous on `stream` and allocate nothing.
 */
#ifndef CLMSOIL_LEARNED_H
#define CLMSOIL_LEARNED_H
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLMSOIL_MLP_NIN    24
#define CLMSOIL_MLP_NHID   64
#define CLMSOIL_MLP_NPARAM 5825
#define CLMSOIL_MLP_NNORM  50
/* Samples per block of the gradient kernel: clmsoil_mlp_grad writes
 * ceil(n / CLMSOIL_MLP_GRAD_TILE) rows of partial gradients. */
#define CLMSOIL_MLP_GRAD_TILE 64

/* Features of every active column of a moisture-form state (a->mask honoured;
 * inactive columns get zeros). norm == NULL writes raw features, otherwise
 * standardised ones. x is [24][ldx] FP32, ldx >= a->nc. */
int clmsoil_richards_features(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                              double dtime, const float *norm, float *x, int64_t ldx,
                              void *stream);

/* Features + network + inverse transform in one launch: writes
 * dtsub0[c] = dtime * exp(y_mean + y_scale * y(c)) for every active column
 * (the Richards kernel clamps it to [min(dtmin, dtime), dtime]). */
int clmsoil_mlp_predict_dtsub0(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                               double dtime, const float *params, const float *norm,
                               double *dtsub0, void *stream);

/* The moisture-form step of clmsoil_soilwm_moisture_form_solve
 * (soil_water_movement.jl:1498-1517) with this network fused into the kernel:
 * each column's first substep is predicted from its state as the column is
 * loaded (the features above, the network on the matrix cores for the 16
 * columns of a wavefront), instead of dtsub = dtime (:1259). The result is the
 * same as clmsoil_mlp_predict_dtsub0 into a dtsub0 array followed by
 * clmsoil_soilwm_moisture_form_solve with that array, without the extra launch
 * and the nc doubles of memory. a->dtsub0 must be NULL (CLMSOIL_ERR_ARG);
 * params/norm NULL gives CLMSOIL_ERR_NOT_TRAINED. Workspace, masks, tuning and
 * error reporting as in clmsoil_soilwm_moisture_form_solve; the kernel always
 * runs 4 lanes per column. */
int clmsoil_soilwm_moisture_form_solve_learned(const clmsoil_mf_args *a,
                                               const clmsoil_levels *lev, double dtime,
                                               const clmsoil_swm_config *cfg,
                                               const float *params, const float *norm,
                                               void *stream);

/* Raw network on standardised inputs x [24][ldx] -> y [n]. */
int clmsoil_mlp_forward(const float *params, const float *x, int64_t n, int64_t ldx, float *y,
                        void *stream);

/* Loss L = (1/n_total) * sum over these n samples of (y - t)^2, and its gradient.
 * Block b of 64 samples writes its partial gradient to partial[b * 5825 + p] and
 * its partial loss to loss_partial[b]; n_total scales the mean (the full batch
 * may span several calls). partial holds ceil(n/64) * 5825 floats. */
int clmsoil_mlp_grad(const float *params, const float *x, const float *t, int64_t n,
                     int64_t ldx, int64_t n_total, float *partial, float *loss_partial,
                     void *stream);

/* Adam step t (1-based) on params with moments m, v [5825]: the gradient is the
 * sum of the nparts rows of `partial` (summed in row order). grad_out (optional,
 * [5825]) receives that sum. */
int clmsoil_mlp_adam(float *params, float *m, float *v, const float *partial, int64_t nparts,
                     int64_t t, float lr, float beta1, float beta2, float eps, float *grad_out,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLMSOIL_LEARNED_H */
