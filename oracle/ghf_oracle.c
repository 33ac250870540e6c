/*
 * ghf_oracle.c — CPU restatement of CLM.jl's ground heat flux assembly
 * (compute_ground_heat_flux_and_deriv!, src/biogeophys/soil_temperature.jl:1200-1263).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 *
 * Follows the reference's kernel sequence statement by statement: longwave
 * emission per column, the rural, rural-contribution and urban patch kernels,
 * the patch-to-column sums of scatter_add_1d! in ascending patch order (the order
 * of the reference's sequential CPU path), then the SNICAR top-layer kernel, also
 * in ascending patch order. Index arrays are 1-based as in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "clmsoil_oracle.h"

#define JMAX orc_jl_max
#define JMIN orc_jl_min

/* varcon.jl:16, 24 */
static const double SB = 5.67e-8, HVAP = 2.501e6;
/* column_varcon.jl:10-11 */
static const int64_t ICOL_ROAD_IMPERV = 74, ICOL_ROAD_PERV = 75;

#define CM(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)a->ldc]
#define PM(X, p, j) (X)[(size_t)(p) + ((size_t)(j) - 1u) * (size_t)a->ldp]

/* Julia's x^4 for Float64: ^(x::Float64, n::Integer) takes Base.Math.pow_body,
 * power by squaring with the rounding error of each square carried along
 * (two_mul through fma) and added back at the end. For n = 4 the multiplier
 * `y` stays 1, so the result is X2 + (e2 + 2*X1*e1) with X1 = fl(x*x),
 * e1 = x*x - X1, X2 = fl(X1*X1), e2 = X1*X1 - X2. */
double orc_jl_pow4(double x) {
    const double x1 = x * x, e1 = fma(x, x, -x1);
    const double err = x1 * 2.0 * e1;
    const double x2 = x1 * x1, e2 = fma(x1, x1, -x2);
    const double lo = e2 + err;
    if (!(isfinite(x2) && isfinite(lo))) return x2;
    return x2 + lo;
}

int orc_ground_heat_flux(const clmsoil_ghf_args *a, const clmsoil_levels *lev) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    if (a->nc < 0 || a->np < 0 || a->ldc < a->nc || a->ldp < a->np) return CLMSOIL_ERR_SIZE;
    const int64_t nc = a->nc, np = a->np, nl = a->nl;
    const int nlevsno = lev->nlevsno, joff = lev->nlevsno;
    const int simple = a->building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE;
    const int prog = a->building_temp_method == CLMSOIL_BUILDING_TEMP_PROG;
    double *lwrad_emit = calloc((size_t)nc + 1, sizeof(double));
    double *dlwrad_emit = calloc((size_t)nc + 1, sizeof(double));
    double *lwrad_emit_snow = calloc((size_t)nc + 1, sizeof(double));
    double *lwrad_emit_soil = calloc((size_t)nc + 1, sizeof(double));
    double *lwrad_emit_h2osfc = calloc((size_t)nc + 1, sizeof(double));
    double *hs_patch = calloc((size_t)np + 1, sizeof(double));
    double *dhsdT_patch = calloc((size_t)np + 1, sizeof(double));
    double *hs_soil_patch = calloc((size_t)np + 1, sizeof(double));
    double *hs_h2osfc_patch = calloc((size_t)np + 1, sizeof(double));
    int st = CLMSOIL_OK;

    /* _ground_lwrad_emit_kernel! (:946-960) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!a->mask_nolakec[c]) continue;
        const double eg = a->emg[c];
        const double tg = a->t_grnd[c];
        lwrad_emit[c] = eg * SB * orc_jl_pow4(tg);
        dlwrad_emit[c] = 4.0 * eg * SB * (tg * tg * tg);
        lwrad_emit_snow[c] = eg * SB * orc_jl_pow4(CM(a->t_soisno, c, a->snl[c] + 1 + joff));
        lwrad_emit_soil[c] = eg * SB * orc_jl_pow4(CM(a->t_soisno, c, 1 + joff));
        lwrad_emit_h2osfc[c] = eg * SB * orc_jl_pow4(a->t_h2osfc[c]);
    }

    /* _gnet_patch_rural_kernel! (:1008-1023) */
    for (int64_t p = 0; p < np; ++p) {
        if (!a->mask_nolakep[p] || a->urbpoi[a->patch_landunit[p] - 1]) continue;
        const int64_t c = a->patch_column[p] - 1;
        const double base = a->dlrad[p] + (1.0 - (double)a->frac_veg_nosno[p]) * a->emg[c] *
                                              a->forc_lwrad[c];
        a->eflx_gnet[p] = a->sabg[p] + base - lwrad_emit[c] -
                          (a->eflx_sh_grnd[p] + a->qflx_evap_soi[p] * a->htvp[c]);
        a->sabg_chk[p] = a->frac_sno_eff[c] * a->sabg_snow[p] +
                         (1.0 - a->frac_sno_eff[c]) * a->sabg_soil[p];
        a->dgnetdT[p] = -a->cgrnd[p] - dlwrad_emit[c];
    }
    /* _gnet_patch_rural_contrib_kernel! (:1025-1049) */
    for (int64_t p = 0; p < np; ++p) {
        if (a->mask_nolakep[p] && !a->urbpoi[a->patch_landunit[p] - 1]) {
            const int64_t c = a->patch_column[p] - 1;
            const double wt = a->wtcol[p];
            const double base = a->dlrad[p] + (1.0 - (double)a->frac_veg_nosno[p]) *
                                                  a->emg[c] * a->forc_lwrad[c];
            const double gsoil = a->sabg_soil[p] + base - lwrad_emit_soil[c] -
                                 (a->eflx_sh_soil[p] + a->qflx_ev_soil[p] * a->htvp[c]);
            const double gh2o = a->sabg_soil[p] + base - lwrad_emit_h2osfc[c] -
                                (a->eflx_sh_h2osfc[p] + a->qflx_ev_h2osfc[p] * a->htvp[c]);
            const int wnz = wt != 0.0;
            hs_patch[p] = wnz ? a->eflx_gnet[p] * wt : 0.0;
            dhsdT_patch[p] = wnz ? a->dgnetdT[p] * wt : 0.0;
            hs_soil_patch[p] = wnz ? gsoil * wt : 0.0;
            hs_h2osfc_patch[p] = wnz ? gh2o * wt : 0.0;
        } else {
            hs_patch[p] = dhsdT_patch[p] = hs_soil_patch[p] = hs_h2osfc_patch[p] = 0.0;
        }
    }
    /* _gnet_patch_urban_kernel! (:1051-1085) */
    for (int64_t p = 0; p < np; ++p) {
        if (!a->mask_nolakep[p] || !a->urbpoi[a->patch_landunit[p] - 1]) continue;
        const int64_t c = a->patch_column[p] - 1, l = a->patch_landunit[p] - 1;
        const double wt = a->wtcol[p];
        const int64_t it = a->col_itype[c];
        const double denom = 1.0 - a->wtlunit_roof[l];
        if (it == ICOL_ROAD_PERV || it == ICOL_ROAD_IMPERV) {
            a->eflx_wasteheat[p] = a->eflx_wasteheat_lun[l] / denom;
            if (simple)
                a->eflx_ventilation[p] = 0.0;
            else if (prog)
                a->eflx_ventilation[p] = a->eflx_ventilation_lun[l] / denom;
            a->eflx_heat_from_ac[p] = a->eflx_heat_from_ac_lun[l] / denom;
            a->eflx_traffic[p] = a->eflx_traffic_lun[l] / denom;
        } else {
            a->eflx_wasteheat[p] = 0.0;
            a->eflx_ventilation[p] = 0.0;
            a->eflx_heat_from_ac[p] = 0.0;
            a->eflx_traffic[p] = 0.0;
        }
        a->eflx_gnet[p] = a->sabg[p] + a->dlrad[p] - a->eflx_lwrad_net[p] -
                          (a->eflx_sh_grnd[p] + a->qflx_evap_soi[p] * a->htvp[c] +
                           a->qflx_tran_veg[p] * HVAP) +
                          a->eflx_wasteheat[p] + a->eflx_heat_from_ac[p] + a->eflx_traffic[p] +
                          a->eflx_ventilation[p];
        if (simple) a->eflx_anthro[p] = a->eflx_wasteheat[p] + a->eflx_traffic[p];
        a->dgnetdT[p] = -a->cgrnd[p] - dlwrad_emit[c];
        const int wnz = wt != 0.0;
        hs_patch[p] = wnz ? a->eflx_gnet[p] * wt : 0.0;
        dhsdT_patch[p] = wnz ? a->dgnetdT[p] * wt : 0.0;
        hs_soil_patch[p] = wnz ? a->eflx_gnet[p] * wt : 0.0;
        hs_h2osfc_patch[p] = wnz ? a->eflx_gnet[p] * wt : 0.0;
    }
    /* scatter_add_1d! (kernels.jl:135-170): reset, then out[column[p]] += v[p] for
     * valid destinations, patches in ascending order */
    for (int64_t c = 0; c < nc; ++c) a->dhsdT[c] = a->hs_soil[c] = a->hs_h2osfc[c] = 0.0;
    for (int64_t p = 0; p < np; ++p) {
        const int64_t d = a->patch_column[p];
        if (d < 1 || d > nc) continue;
        a->dhsdT[d - 1] += dhsdT_patch[p];
        a->hs_soil[d - 1] += hs_soil_patch[p];
        a->hs_h2osfc[d - 1] += hs_h2osfc_patch[p];
    }
    /* _gnet_snicar_kernel! (:1135-1162) after zeroing its outputs (:1252-1254) */
    for (int64_t c = 0; c < nc; ++c) {
        a->hs_top[c] = a->hs_top_snow[c] = 0.0;
        for (int j = 1; j <= nlevsno + 1; ++j) CM(a->sabg_lyr_col, c, j) = 0.0;
    }
    for (int64_t p = 0; p < np; ++p) {
        if (!a->mask_nolakep[p]) continue;
        const int64_t c = a->patch_column[p] - 1, l = a->patch_landunit[p] - 1;
        const double wt = a->wtcol[p];
        const int64_t lyr_top = a->snl[c] + 1;
        if (!a->urbpoi[l]) {
            const double base = a->dlrad[p] + (1.0 - (double)a->frac_veg_nosno[p]) *
                                                  a->emg[c] * a->forc_lwrad[c];
            const double sab = PM(a->sabg_lyr_patch, p, lyr_top + joff);
            const double gtop = sab + base - lwrad_emit[c] -
                                (a->eflx_sh_grnd[p] + a->qflx_evap_soi[p] * a->htvp[c]);
            a->hs_top[c] += gtop * wt;
            const double gsnow = sab + base - lwrad_emit_snow[c] -
                                 (a->eflx_sh_snow[p] + a->qflx_ev_snow[p] * a->htvp[c]);
            a->hs_top_snow[c] += gsnow * wt;
            for (int64_t j = lyr_top; j <= 1; ++j)
                CM(a->sabg_lyr_col, c, j + nlevsno) += PM(a->sabg_lyr_patch, p, j + joff) * wt;
        } else {
            const double g = a->eflx_gnet[p];
            a->hs_top[c] += g * wt;
            a->hs_top_snow[c] += g * wt;
            CM(a->sabg_lyr_col, c, lyr_top + nlevsno) += a->sabg[p] * wt;
        }
    }
    (void)nl;
    free(lwrad_emit); free(dlwrad_emit); free(lwrad_emit_snow); free(lwrad_emit_soil);
    free(lwrad_emit_h2osfc); free(hs_patch); free(dhsdT_patch); free(hs_soil_patch);
    free(hs_h2osfc_patch);
    return st;
}

/* Soil energy balance error: _soilflux_errsoi_base_kernel! (soil_fluxes.jl:380-406),
 * _soilflux_errsoi_layers_kernel! for the non-urban and the urban pass (:425-447,
 * launched at :651-660) and the patch-to-column average (:499-525), in that order. */
int orc_soil_energy_balance(const clmsoil_seb_args *a, const clmsoil_levels *lev,
                            double dtime) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    const int64_t nc = a->nc, np = a->np;
    const int nlevsno = lev->nlevsno, nlevgrnd = lev->nlevgrnd, nlevurb = a->nlevurb;
    const int64_t ICOL_ROOF = 71, ICOL_SUNWALL = 72, ICOL_SHADEWALL = 73;
    for (int64_t p = 0; p < np; ++p) {
        if (!a->mask_nolakep[p]) continue;
        const int64_t c = a->patch_column[p] - 1;
        a->errsoi_patch[p] = a->eflx_soil_grnd[p] - a->xmf[c] - a->xmf_h2osfc[c] -
                             a->frac_h2osfc[c] * (a->t_h2osfc[c] - a->t_h2osfc_bef[c]) *
                                 (a->c_h2osfc[c] / dtime);
        a->errsoi_patch[p] += a->eflx_h2osfc_to_snow[c];
        const int64_t it = a->col_itype[c];
        if (it == ICOL_SUNWALL || it == ICOL_SHADEWALL || it == ICOL_ROOF)
            a->errsoi_patch[p] += a->eflx_building_heat_errsoi[c];
    }
    for (int pass = 0; pass < 2; ++pass) {
        const uint8_t *mask = pass == 0 ? a->mask_nolakep : a->mask_urbanp;
        const int jmax = pass == 0 ? nlevgrnd : nlevurb;
        for (int64_t p = 0; p < np; ++p) {
            if (!mask[p]) continue;
            const int64_t c = a->patch_column[p] - 1;
            const int64_t it = a->col_itype[c];
            const int wall = it == ICOL_SUNWALL || it == ICOL_SHADEWALL || it == ICOL_ROOF;
            if (pass != wall) continue;
            for (int jf = -nlevsno + 1; jf <= jmax; ++jf) {
                const int jj = jf + nlevsno;
                const double dt = CM(a->t_soisno, c, jj) - CM(a->t_ssbef, c, jj);
                if (jf >= a->snl[c] + 1 && jf < 1)
                    a->errsoi_patch[p] -= a->frac_sno_eff[c] * dt / CM(a->fact, c, jj);
                if (jf >= 1) a->errsoi_patch[p] -= dt / CM(a->fact, c, jj);
            }
        }
    }
    for (int64_t c = 0; c < nc; ++c)
        if (a->mask_nolakec[c]) a->errsoi_col[c] = 0.0;
    for (int64_t p = 0; p < np; ++p) {
        const int64_t c = a->patch_column[p];
        if (c >= 1 && c <= nc && a->mask_nolakec[c - 1] && a->patchi[c - 1] <= p + 1 &&
            p + 1 <= a->patchf[c - 1])
            a->errsoi_col[c - 1] += a->errsoi_patch[p] * a->wtcol[p];
    }
    return CLMSOIL_OK;
}
