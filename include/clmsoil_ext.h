/*
 * clmsoil_ext.h — entry points of libclmsoil with no direct counterpart in
 * CLM.jl's interfaces: device status and launch tuning.
 * Conventions as in clmsoil.h.
 */
#ifndef CLMSOIL_EXT_H
#define CLMSOIL_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "clmsoil.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device error bits recorded by the kernels. */
#define CLMSOIL_DEVERR_NBEDROCK 1u   /* nbedrock (or jtop) outside 1..nlevsoi      */
#define CLMSOIL_DEVERR_SNL      2u   /* snl outside -nlevsno..0 (or band range)    */
#define CLMSOIL_DEVERR_ITER     4u   /* substep iteration cap reached              */
#define CLMSOIL_DEVERR_TAPE     8u   /* adjoint: more accepted substeps than the tape holds */
#define CLMSOIL_DEVERR_SUBGRID 16u   /* ground heat flux: patch ranges and patch columns disagree */
#define CLMSOIL_DEVERR_JWT     32u   /* methane transport: jwt outside 0..nlevsoi   */
#define CLMSOIL_DEVERR_ADJ     64u   /* heat adjoint: column outside its scope (snow, surface water,
                                         excess ice) or a row interchange in its solve */

/* Synchronize `stream`, return and clear the current device's error bits (the
 * flags of every call on the device since the last read). The reference raises
 * these conditions as Julia bounds errors inside its kernels. */
int clmsoil_device_errors(void *stream, uint32_t *flags);
/* Synchronize `stream` and return the error bits of the last Richards call that
 * used `workspace` (a call clears them when it starts). */
int clmsoil_workspace_errors(const void *workspace, void *stream, uint32_t *flags);

/* Launch tuning. Lanes per column (1, 2 or 4) and workgroup size (64, 128 or
 * 256) of the fused Richards and heat kernels; process-wide, read once per call
 * (and fixed in a day plan when it is created). Every setting gives the same
 * results bit for bit, so changing them while other threads launch work changes
 * only speed. */
int clmsoil_set_tuning(int richards_g, int richards_block, int heat_g, int heat_block);
int clmsoil_get_tuning(int *richards_g, int *richards_block, int *heat_g, int *heat_block);
/* 1 (default): the Richards step runs as a resident grid whose lane groups each
 * take a new column from a work counter as soon as their column's substeps are
 * done; 0: one column per lane group for the launch. Results are identical. */
int clmsoil_set_richards_schedule(int persistent);
/* Blocks per CU of the resident Richards grid (0, the default: as many as the
 * occupancy calculator says fit; 1..32 ask for fewer; a request above what fits
 * is clamped to it). Results are identical. */
int clmsoil_set_richards_residency(int blocks_per_cu);
/* Resident Richards grid: a wavefront refills its idle lane groups once
 * n/16 of them are idle (or none is busy), n in 1..16; separately for the plain
 * step and for the step with the fused first-substep network
 * (clmsoil_soilwm_moisture_form_solve_learned), whose refills also evaluate the
 * network for the whole wavefront. Defaults 7 and 7. Results are identical. */
int clmsoil_set_richards_refill(int plain_sixteenths, int learned_sixteenths);
/* Per-layer fields of a Richards column the G = 4, 256-thread kernels keep in a
 * lane-private LDS cache instead of registers or memory: 7 (default: bsw, sucsat,
 * root sink, smp, hk, the impedance-scaled conductivity and the node distance;
 * 70 KB per block), 5 (the first five; 50 KB) or 0 (no cache). The step with the
 * fused first-substep network always uses 7. Results are identical. */
int clmsoil_set_richards_lds_cache(int fields);
/* Workgroup size (64, 128 — the default — or 256) of the Zeng-Decker 2009 step
 * (include/clmsoil_zd09.h). Results are identical. */
int clmsoil_set_zd_block(int block);

/* The library's scalar FP64 functions evaluated element-wise on device arrays:
 * the kernels' pow, log, log10 and exp (Tang's table-driven method, one source
 * shared with the CPU restatement) and Julia's NaN-propagating max/min
 * (smooth_ad.jl:91-149 on Float64). out[i] = f(x[i], y[i]) for POW, JL_MAX and
 * JL_MIN, f(x[i]) otherwise (y unused). No counterpart in the reference; it
 * exists so that the device build of these functions can be checked against
 * the host build. */
#define CLMSOIL_MATH_POW    0
#define CLMSOIL_MATH_LOG    1
#define CLMSOIL_MATH_LOG10  2
#define CLMSOIL_MATH_EXP    3
#define CLMSOIL_MATH_JL_MAX 4
#define CLMSOIL_MATH_JL_MIN 5
/* The kernels' forms: the fast path of pow / log10 for a wavefront whose
 * arguments all lie in its domain (positive normal x, finite y), the general
 * function otherwise — the same bits either way. POW_BASE_W takes x[i] as a
 * fixed base (its logarithm computed once, as the kernels do for constant
 * bases). */
#define CLMSOIL_MATH_POW_W      6
#define CLMSOIL_MATH_POW_BASE_W 7
#define CLMSOIL_MATH_LOG10_W    8
int clmsoil_math_eval(int op, const double *x, const double *y, double *out, int64_t n,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif
