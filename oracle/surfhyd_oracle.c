/*
 * surfhyd_oracle.c — CPU restatement of CLM.jl's surface hydrology before the
 * Richards step (clm_driver.jl:1939-2012): flood water, saturated and
 * infiltration excess runoff, the surface water store, infiltration, total
 * surface runoff and urban ponding.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 * Each function below restates one reference kernel and, like the reference,
 * runs over all columns before the next one starts. Julia's min/max on Float64
 * are orc_jl_min/max (NaN-propagating).
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "clmsoil_oracle.h"

#define JMAX orc_jl_max
#define JMIN orc_jl_min

/* landunit_varcon.jl:7-8, column_varcon.jl:7-11, varcon.jl:85 */
enum { SH_ISTSOIL = 1, SH_ISTCROP = 2 };
enum { SH_ROOF = 71, SH_SUNWALL = 72, SH_SHADEWALL = 73, SH_ROAD_IMPERV = 74 };
static const int64_t SH_ISPVAL = -9999;
/* infilt_excess_runoff.jl:38 */
static const double SH_QINMAX_UNLIMITED = 1.0e200;
/* Julia's pi / 180.0 in Float64 (surface_water.jl:303) */
static const double SH_DEG2RAD = 3.141592653589793 / 180.0;

static int on(const uint8_t *m, int64_t c) { return !m || m[c]; }

/* Test hook: with a non-zero seed every exp(), pow() and sin() below is moved
 * one ulp up or down, the direction taken from a hash of its operands and the
 * seed (deterministic). These are the operations whose last bit can differ
 * between the host libm and the HIP device library; the parity tests use the
 * hook to measure how far such a difference moves the results. Zero (the
 * default) leaves them exact. */
static uint64_t g_sh_jitter = 0;
void orc_set_surfhyd_jitter(uint64_t seed) { g_sh_jitter = seed; }

static double sh_jit(double r, double x, double y) {
    if (!g_sh_jitter || !isfinite(r) || r == 0.0) return r;
    uint64_t bx, by;
    memcpy(&bx, &x, 8);
    memcpy(&by, &y, 8);
    uint64_t h = bx ^ (by * 0x9E3779B97F4A7C15ull) ^ g_sh_jitter;
    h ^= h >> 30; h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27; h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return nextafter(r, (h & 1u) ? INFINITY : -INFINITY);
}
static double sh_exp(double x) { return sh_jit(exp(x), x, 0.5); }
static double sh_pow(double x, double y) { return sh_jit(pow(x, y), x, y); }
static double sh_sin(double x) { return sh_jit(sin(x), x, 1.5); }

/* _soilhyd_floodc_kernel! (soil_hydrology.jl:139-152) */
void orc_sh_floodc(const clmsoil_surfhyd_args *a) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_nolakec, c)) continue;
        const int64_t g = a->col_gridcell[c] - 1;
        if (a->col_itype[c] != SH_SUNWALL && a->col_itype[c] != SH_SHADEWALL)
            a->qflx_floodc[c] = a->qflx_floodg[g];
        else
            a->qflx_floodc[c] = 0.0;
    }
}

/* _sat_excess_runoff_kernel! (sat_excess_runoff.jl:246-300) */
void orc_sh_sat_excess(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        double fs;
        if (p->fsat_method == CLMSOIL_FSAT_TOPMODEL) {
            const double ft = a->frost_table[c], zw = a->zwt[c], zwp = a->zwt_perched[c];
            if (ft > zwp && ft <= zw)
                fs = a->wtfact[c] * sh_exp(-0.5 * p->fff * zwp);
            else
                fs = a->wtfact[c] * sh_exp(-0.5 * p->fff * zw);
        } else {
            const double bi = a->b_infil[c];
            const double ex = bi / (1.0 + bi);
            fs = 1.0 - sh_pow(1.0 - a->top_moist_limited[c] / a->top_max_moist[c], ex);
        }
        if (p->crop_fsat_equals_zero && a->lun_itype[a->col_landunit[c] - 1] == SH_ISTCROP)
            fs = 0.0;
        if (p->hillslope_fsat_equals_zero && a->is_hillslope_column[c] && a->col_active[c] &&
            a->col_cold[c] != SH_ISPVAL)
            fs = 0.0;
        a->fsat[c] = fs;
        double q = fs * a->qflx_rain_plus_snomelt[c];
        a->fcov[c] = fs;
        if (a->col_urbpoi[c]) q = q + a->qflx_floodc[c];
        a->qflx_sat_excess_surf[c] = q;
    }
}

/* _soilhyd_qflx_inputs_kernel! (soil_hydrology.jl:160-187) */
void orc_sh_qflx_inputs(const clmsoil_surfhyd_args *a) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        a->qflx_top_soil[c] = a->qflx_rain_plus_snomelt[c] + a->qflx_snow_h2osfc[c] + a->qflx_floodc[c];
        double fsno, qflx_evap;
        if (a->snl[c] >= 0) {
            fsno = 0.0;
            qflx_evap = a->qflx_liqevap_from_top_layer[c];
        } else {
            fsno = a->frac_sno_eff[c];
            qflx_evap = a->qflx_ev_soil[c];
        }
        const double fh = a->frac_h2osfc[c];
        a->qflx_in_soil[c] = (1.0 - fh) * (a->qflx_top_soil[c] - a->qflx_sat_excess_surf[c]);
        a->qflx_top_soil_to_h2osfc[c] = fh * (a->qflx_top_soil[c] - a->qflx_sat_excess_surf[c]);
        a->qflx_in_soil[c] = a->qflx_in_soil[c] - (1.0 - fsno - fh) * qflx_evap;
        a->qflx_top_soil_to_h2osfc[c] = a->qflx_top_soil_to_h2osfc[c] - fh * a->qflx_ev_h2osfc[c];
    }
}

/* qinmax on the unsaturated area: _qinmax_hksat_kernel! (infilt_excess_runoff.jl:241-256)
 * or the unlimited value (:142-147) */
static double qinmax_unsat(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                           int64_t c) {
    if (p->qinmax_method == CLMSOIL_QINMAX_NONE) return SH_QINMAX_UNLIMITED;
    double qmin = INFINITY;
    for (int j = 1; j <= 3; ++j) {
        const size_t o = (size_t)c + (size_t)(j - 1) * (size_t)a->ld;
        const double val = sh_pow(10.0, -p->e_ice * a->icefrac[o]) * a->hksat[o];
        if (val < qmin) qmin = val;
    }
    return qmin;
}

/* _infexcess_colavg_kernel! (infilt_excess_runoff.jl:154-165) */
void orc_sh_infil_excess(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        a->qinmax[c] = (1.0 - a->fsat[c]) * qinmax_unsat(a, p, c);
        a->qflx_infl_excess[c] =
            JMAX(0.0, a->qflx_in_soil[c] - (1.0 - a->frac_h2osfc[c]) * a->qinmax[c]);
    }
}

/* _soilhyd_route_infl_excess_kernel! (soil_hydrology.jl:199-225) */
void orc_sh_route(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        const int64_t lt = a->lun_itype[a->col_landunit[c] - 1];
        if (lt == SH_ISTSOIL || lt == SH_ISTCROP) {
            a->qflx_in_soil_limited[c] = a->qflx_in_soil[c] - a->qflx_infl_excess[c];
            if (p->h2osfcflag != 0) {
                a->qflx_in_h2osfc[c] = a->qflx_top_soil_to_h2osfc[c] + a->qflx_infl_excess[c];
                a->qflx_infl_excess_surf[c] = 0.0;
            } else {
                a->qflx_in_h2osfc[c] = a->qflx_top_soil_to_h2osfc[c];
                a->qflx_infl_excess_surf[c] = a->qflx_infl_excess[c];
            }
        } else {
            a->qflx_in_soil_limited[c] = a->qflx_in_soil[c];
            a->qflx_in_h2osfc[c] = 0.0;
            a->qflx_infl_excess_surf[c] = 0.0;
        }
    }
}

/* update_h2osfc! (surface_water.jl:462-520): surface runoff (:279-316), partial
 * update (:400-414), drainage (:357-374), final update (:426-443), each over
 * all columns in turn. surf and drain are the per-call scratch arrays. */
void orc_sh_update_h2osfc(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                          double dtime, double *surf, double *drain) {
    const int flag = p->h2osfcflag_update;
    for (int64_t c = 0; c < a->nc; ++c) {
        surf[c] = 0.0;
        drain[c] = 0.0;
    }
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        if (flag != 1) {
            surf[c] = 0.0;
            continue;
        }
        const double fn = a->frac_h2osfc_nosnow[c];
        const double frac_infclust = fn <= p->pc ? 0.0 : sh_pow(fn - p->pc, p->mu);
        double q = 0.0;
        if (a->h2osfc[c] > a->h2osfc_thresh[c] && flag != 0) {
            const double k_wet = 1.0e-4 * sh_sin(SH_DEG2RAD * a->topo_slope[c]);
            q = k_wet * frac_infclust * (a->h2osfc[c] - a->h2osfc_thresh[c]);
            q = JMIN(q, (a->h2osfc[c] - a->h2osfc_thresh[c]) / dtime);
        }
        if (q < 1.0e-8) q = 0.0;
        surf[c] = q;
    }
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        double h = a->h2osfc[c] + (a->qflx_in_h2osfc[c] - surf[c]) * dtime;
        if (fabs(h) < 1.0e-13 * fabs(a->h2osfc[c])) h = 0.0;
        a->h2osfc[c] = h;
    }
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        if (a->h2osfc[c] < 0.0) {
            drain[c] = a->h2osfc[c] / dtime;
        } else {
            double q = JMIN(a->frac_h2osfc[c] * a->qinmax[c], a->h2osfc[c] / dtime);
            if (flag == 0) q = JMAX(0.0, a->h2osfc[c] / dtime);
            drain[c] = q;
        }
    }
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        double h = a->h2osfc[c] - drain[c] * dtime;
        if (fabs(h) < 1.0e-13 * fabs(a->h2osfc[c])) h = 0.0;
        a->h2osfc[c] = h;
        a->qflx_h2osfc_surf[c] = surf[c];
        a->qflx_h2osfc_drain[c] = drain[c];
    }
}

/* _soilhyd_infiltration_kernel! (soil_hydrology.jl:236-243) */
void orc_sh_infiltration(const clmsoil_surfhyd_args *a) {
    for (int64_t c = 0; c < a->nc; ++c)
        if (on(a->mask_hydrologyc, c))
            a->qflx_infl[c] = a->qflx_in_soil_limited[c] + a->qflx_h2osfc_drain[c];
}

/* _soilhyd_surf_runoff_kernel! and _soilhyd_surf_runoff_urban_kernel!
 * (soil_hydrology.jl:250-258, 273-301), in the order of total_surface_runoff! */
void orc_sh_surf_runoff(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                        int nlevsno, double dtime) {
    for (int64_t c = 0; c < a->nc; ++c)
        if (on(a->mask_hydrologyc, c))
            a->qflx_surf[c] = a->qflx_sat_excess_surf[c] + a->qflx_infl_excess_surf[c] +
                              a->qflx_h2osfc_surf[c];
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_urbanc, c)) continue;
        const int64_t it = a->col_itype[c];
        if (it == SH_ROOF || it == SH_ROAD_IMPERV) {
            if (a->snl[c] < 0) {
                a->qflx_surf[c] = JMAX(0.0, a->qflx_rain_plus_snomelt[c]);
            } else {
                const double pond = a->h2osoi_liq[(size_t)c + (size_t)nlevsno * (size_t)a->ld];
                a->xs_urban[c] = JMAX(0.0, pond / dtime + a->qflx_rain_plus_snomelt[c] -
                                                a->qflx_liqevap_from_top_layer[c] -
                                                p->pondmx_urban / dtime);
                a->qflx_surf[c] = a->xs_urban[c];
            }
            a->qflx_surf[c] = a->qflx_surf[c] + a->qflx_floodc[c];
        } else if (it == SH_SUNWALL || it == SH_SHADEWALL) {
            a->qflx_surf[c] = 0.0;
        }
    }
}

/* _soilhyd_urban_ponding_kernel! (soil_hydrology.jl:316-338) */
void orc_sh_urban_pond(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                       int nlevsno, double dtime) {
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_urbanc, c)) continue;
        const int64_t it = a->col_itype[c];
        if ((it == SH_ROOF || it == SH_ROAD_IMPERV) && a->snl[c] >= 0) {
            double *pond = &a->h2osoi_liq[(size_t)c + (size_t)nlevsno * (size_t)a->ld];
            if (a->xs_urban[c] > 0.0)
                *pond = p->pondmx_urban;
            else
                *pond = JMAX(0.0, *pond + (a->qflx_rain_plus_snomelt[c] -
                                           a->qflx_liqevap_from_top_layer[c]) * dtime);
        }
    }
}

/* The requested steps in the driver's order (clm_driver.jl:1939-2012). */
int orc_surface_hydrology(const clmsoil_surfhyd_args *a, const clmsoil_levels *lev, double dtime,
                          const clmsoil_surfhyd_params *p, double *scratch2nc) {
    if (!a || !lev || !p) return CLMSOIL_ERR_NULL;
    if ((p->steps & ~CLMSOIL_SH_ALL) != 0) return CLMSOIL_ERR_ARG;
    if (p->fsat_method != CLMSOIL_FSAT_TOPMODEL && p->fsat_method != CLMSOIL_FSAT_VIC)
        return CLMSOIL_ERR_ARG;
    if (p->qinmax_method != CLMSOIL_QINMAX_NONE && p->qinmax_method != CLMSOIL_QINMAX_HKSAT)
        return CLMSOIL_ERR_ARG;
    const int s = p->steps;
    if (s & CLMSOIL_SH_FLOODC) orc_sh_floodc(a);
    if (s & CLMSOIL_SH_SAT_EXCESS) orc_sh_sat_excess(a, p);
    if (s & CLMSOIL_SH_QFLX_INPUTS) orc_sh_qflx_inputs(a);
    if (s & CLMSOIL_SH_INFIL_EXCESS) orc_sh_infil_excess(a, p);
    if (s & CLMSOIL_SH_ROUTE) orc_sh_route(a, p);
    if (s & CLMSOIL_SH_H2OSFC) orc_sh_update_h2osfc(a, p, dtime, scratch2nc, scratch2nc + a->nc);
    if (s & CLMSOIL_SH_INFILTRATION) orc_sh_infiltration(a);
    if (s & CLMSOIL_SH_SURF_RUNOFF) orc_sh_surf_runoff(a, p, lev->nlevsno, dtime);
    if (s & CLMSOIL_SH_URBAN_POND) orc_sh_urban_pond(a, p, lev->nlevsno, dtime);
    return CLMSOIL_OK;
}

/* _plantsink_default_kernel! (soil_water_plant_sink.jl:48-81) over the hydrology
 * columns, as compute_effec_rootfrac_and_vert_tran_sink! (:341-451) runs it with
 * use_hydrstress = false. A patch range outside 1..np is an error here. */
int orc_compute_effec_rootfrac_and_vert_tran_sink(const clmsoil_rootsink_args *a,
                                                  const clmsoil_levels *lev) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        if (a->npatches[c] < 0 || a->patchi[c] < 1 || a->patchi[c] - 1 + a->npatches[c] > a->np)
            return CLMSOIL_ERR_SIZE;
    }
    for (int64_t c = 0; c < a->nc; ++c) {
        if (!on(a->mask_hydrologyc, c)) continue;
        const int64_t p_lo = a->patchi[c] - 1, p_hi = a->patchi[c] - 1 + a->npatches[c] - 1;
        double temp_c = 0.0;
        for (int64_t p = p_lo; p <= p_hi; ++p)
            if (a->active[p]) temp_c += a->qflx_tran_veg_patch[p] * a->wtcol[p];
        for (int j = 1; j <= lev->nlevsoi; ++j) {
            double acc = 0.0;
            for (int64_t p = p_lo; p <= p_hi; ++p)
                if (a->active[p])
                    acc += a->rootr_patch[(size_t)p + (size_t)(j - 1) * (size_t)a->ldp] *
                           a->qflx_tran_veg_patch[p] * a->wtcol[p];
            if (temp_c != 0.0) acc /= temp_c;
            const size_t o = (size_t)c + (size_t)(j - 1) * (size_t)a->ldc;
            a->rootr_col[o] = acc;
            a->qflx_rootsoi_col[o] = acc * a->qflx_tran_veg_col[c];
        }
    }
    return CLMSOIL_OK;
}
