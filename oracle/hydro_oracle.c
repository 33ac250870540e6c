/*
 * hydro_oracle.c — CPU restatement of CLM.jl's soil water fractions and water
 * table (set_soil_water_fractions!, water_table!; src/biogeophys/soil_hydrology.jl).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 * smooth_min/smooth_max on Float64 are Julia's min/max (orc_jl_min/max).
 */
#include <math.h>
#include <stddef.h>

#include "clmsoil_oracle.h"

#define JMAX orc_jl_max
#define JMIN orc_jl_min
#define A(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)a->ld]

/* varcon.jl:15-70 */
static const double H_DENH2O = 1000.0, H_DENICE = 917.0, H_TFRZ = 273.15;

/* _soilhyd_water_fractions_kernel! (soil_hydrology.jl:110-131) */
int orc_set_soil_water_fractions(const clmsoil_fractions_args *a, const clmsoil_levels *lev) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    const int joff = lev->nlevsno, n = lev->nlevsoi;
    for (int64_t c = 0; c < a->nc; ++c) {
        if (a->mask && !a->mask[c]) continue;
        for (int j = 1; j <= n; ++j) {
            const double exc = a->excess_ice ? A(a->excess_ice, c, j) : 0.0;
            const double dz_ext = A(a->dz, c, j + joff) + exc / H_DENICE;
            const double vol_ice = JMIN(A(a->watsat, c, j),
                                        (A(a->h2osoi_ice, c, j + joff) + exc) / (dz_ext * H_DENICE));
            A(a->eff_porosity, c, j) = JMAX(0.01, A(a->watsat, c, j) - vol_ice);
            A(a->icefrac, c, j) = JMIN(1.0, vol_ice / A(a->watsat, c, j));
        }
    }
    return CLMSOIL_OK;
}

/* specific yield of layer j for a water table at zwt (the analytical aquifer
 * specific yield of :691-693 and :705-707) */
static double spec_yield(const clmsoil_water_table_args *a, int64_t c, int j, double zwt) {
    return A(a->watsat, c, j) *
           (1.0 - pow(1.0 + 1.0e3 * zwt / A(a->sucsat, c, j), -1.0 / A(a->bsw, c, j)));
}

/* _water_table_kernel! (soil_hydrology.jl:676-826) */
int orc_water_table(const clmsoil_water_table_args *a, const clmsoil_levels *lev, double dtime,
                    double aqy, int recompute_frost_table) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    const int n = lev->nlevsoi, joff = lev->nlevsno, joff_zi = lev->nlevsno + 1;
    for (int64_t c = 0; c < a->nc; ++c) {
        if (a->mask && !a->mask[c]) continue;
        a->qflx_drain[c] = 0.0;
        a->qflx_rsub_sat[c] = 0.0;
        a->qflx_drain_perched[c] = 0.0;
        int jwt = n;
        for (int j = 1; j <= n; ++j)
            if (a->zwt[c] <= A(a->zi, c, joff_zi + j)) { jwt = j - 1; break; }
        double rous = spec_yield(a, c, n, a->zwt[c]);
        rous = JMAX(rous, aqy);
        if (jwt == n) {
            a->wa[c] = a->wa[c] + a->qcharge[c] * dtime;
            a->zwt[c] = a->zwt[c] - (a->qcharge[c] * dtime) / 1000.0 / rous;
        } else {
            double qtot = a->qcharge[c] * dtime;
            if (qtot > 0.0) {                                   /* rising water table */
                for (int j = jwt + 1; j >= 1; --j) {
                    double s_y = spec_yield(a, c, j, a->zwt[c]);
                    s_y = JMAX(s_y, aqy);
                    double ql = JMIN(qtot, s_y * (a->zwt[c] - A(a->zi, c, joff_zi + j - 1)) * 1.0e3);
                    ql = JMAX(ql, 0.0);
                    if (s_y > 0.0) a->zwt[c] = a->zwt[c] - ql / s_y / 1000.0;
                    qtot = qtot - ql;
                    if (qtot <= 0.0) break;
                }
            } else {                                            /* deepening water table */
                for (int j = jwt + 1; j <= n; ++j) {
                    double s_y = spec_yield(a, c, j, a->zwt[c]);
                    s_y = JMAX(s_y, aqy);
                    double ql = JMAX(qtot, -(s_y * (A(a->zi, c, joff_zi + j) - a->zwt[c]) * 1.0e3));
                    ql = JMIN(ql, 0.0);
                    qtot = qtot - ql;
                    if (qtot >= 0.0) {
                        a->zwt[c] = a->zwt[c] - ql / s_y / 1000.0;
                        break;
                    } else {
                        a->zwt[c] = A(a->zi, c, joff_zi + j);
                    }
                }
                if (qtot > 0.0) a->zwt[c] = a->zwt[c] - qtot / 1000.0 / rous;
            }
        }
        if (!recompute_frost_table) continue;
        /* frost table and perched water table (:757-823) */
        int k_frz = A(a->t_soisno, c, joff + 1) > H_TFRZ ? n : 1;
        for (int k = 2; k <= n; ++k)
            if (A(a->t_soisno, c, joff + k - 1) > H_TFRZ && A(a->t_soisno, c, joff + k) <= H_TFRZ) {
                k_frz = k;
                break;
            }
        a->frost_table[c] = A(a->z, c, joff + k_frz);
        a->zwt_perched[c] = a->frost_table[c];
        if (a->zwt[c] < a->frost_table[c] && A(a->t_soisno, c, joff + k_frz) <= H_TFRZ) continue;
        const double sat_lev = 0.9;
        int k_perch = 1;
        for (int k = k_frz; k >= 1; --k) {
            A(a->h2osoi_vol, c, k) = A(a->h2osoi_liq, c, joff + k) / (A(a->dz, c, joff + k) * H_DENH2O) +
                                     A(a->h2osoi_ice, c, joff + k) / (A(a->dz, c, joff + k) * H_DENICE);
            if (A(a->h2osoi_vol, c, k) / A(a->watsat, c, k) <= sat_lev) { k_perch = k; break; }
        }
        if (A(a->t_soisno, c, joff + k_frz) > H_TFRZ) k_perch = k_frz;
        if (k_frz > k_perch) {
            const int kp = k_perch;
            const double s1 = (A(a->h2osoi_liq, c, joff + kp) / (A(a->dz, c, joff + kp) * H_DENH2O) +
                               A(a->h2osoi_ice, c, joff + kp) / (A(a->dz, c, joff + kp) * H_DENICE)) /
                              A(a->watsat, c, kp);
            const double s2 = (A(a->h2osoi_liq, c, joff + kp + 1) / (A(a->dz, c, joff + kp + 1) * H_DENH2O) +
                               A(a->h2osoi_ice, c, joff + kp + 1) / (A(a->dz, c, joff + kp + 1) * H_DENICE)) /
                              A(a->watsat, c, kp + 1);
            const double m = (A(a->z, c, joff + kp + 1) - A(a->z, c, joff + kp)) / (s2 - s1);
            const double b = A(a->z, c, joff + kp + 1) - m * s2;
            a->zwt_perched[c] = JMAX(0.0, m * sat_lev + b);
        }
    }
    return CLMSOIL_OK;
}

/* ---- water tables diagnosed from the moisture profile (no aquifer layer) ---- */

#define T_(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)t->ld]

static double t_vol(const clmsoil_theta_wt_args *t, int64_t c, int k, int joff) {
    return T_(t->h2osoi_liq, c, k + joff) / (T_(t->dz, c, k + joff) * H_DENH2O) +
           T_(t->h2osoi_ice, c, k + joff) / (T_(t->dz, c, k + joff) * H_DENICE);
}

static int theta_nbedrock_ok(const clmsoil_theta_wt_args *t, int n) {
    for (int64_t c = 0; c < t->nc; ++c)
        if ((!t->mask || t->mask[c]) && (t->nbedrock[c] < 1 || t->nbedrock[c] > n)) return 0;
    return 1;
}

/* _perched_water_table_kernel! (soil_hydrology.jl:1522-1587) */
int orc_perched_water_table(const clmsoil_theta_wt_args *t, const clmsoil_levels *lev) {
    if (!t || !lev) return CLMSOIL_ERR_NULL;
    const int n = lev->nlevsoi, joff = lev->nlevsno, joff_zi = lev->nlevsno + 1;
    const double sat_lev = 0.9;
    for (int64_t c = 0; c < t->nc; ++c) {
        if (t->mask && !t->mask[c]) continue;
        int k_frz = T_(t->t_soisno, c, 1 + joff) > H_TFRZ ? n : 1;
        for (int k = 2; k <= n; ++k) {
            if (T_(t->t_soisno, c, k - 1 + joff) > H_TFRZ && T_(t->t_soisno, c, k + joff) <= H_TFRZ) {
                k_frz = k;
                break;
            }
        }
        t->frost_table[c] = T_(t->zi, c, k_frz - 1 + joff_zi);
        t->zwt_perched[c] = t->frost_table[c];
        if (t->zwt[c] < t->frost_table[c] && T_(t->t_soisno, c, k_frz + joff) <= H_TFRZ) {
            /* water table above the frost table: nothing to do */
        } else if (k_frz > 1) {
            int k_perch = 1;
            for (int k = k_frz; k >= 1; --k) {
                T_(t->h2osoi_vol, c, k) = t_vol(t, c, k, joff);
                if (T_(t->h2osoi_vol, c, k) / T_(t->watsat, c, k) <= sat_lev) {
                    k_perch = k;
                    break;
                }
            }
            if (T_(t->t_soisno, c, k_frz + joff) > H_TFRZ) k_perch = k_frz;
            if (k_frz > k_perch) {
                const double s1 = t_vol(t, c, k_perch, joff) / T_(t->watsat, c, k_perch);
                const double s2 = t_vol(t, c, k_perch + 1, joff) / T_(t->watsat, c, k_perch + 1);
                if (s1 > s2) {
                    t->zwt_perched[c] = T_(t->zi, c, k_perch - 1 + joff_zi);
                } else {
                    const double m = (T_(t->z, c, k_perch + 1 + joff) - T_(t->z, c, k_perch + joff)) /
                                     (s2 - s1);
                    const double b = T_(t->z, c, k_perch + 1 + joff) - m * s2;
                    t->zwt_perched[c] = JMAX(0.0, m * sat_lev + b);
                }
            }
        }
    }
    return CLMSOIL_OK;
}

/* _theta_based_water_table_kernel! (soil_hydrology.jl:1670-1715) */
int orc_theta_based_water_table(const clmsoil_theta_wt_args *t, const clmsoil_levels *lev) {
    if (!t || !lev) return CLMSOIL_ERR_NULL;
    const int n = lev->nlevsoi, joff = lev->nlevsno, joff_zi = lev->nlevsno + 1;
    const double sat_lev = 0.9;
    if (!theta_nbedrock_ok(t, n)) return CLMSOIL_ERR_SIZE;
    for (int64_t c = 0; c < t->nc; ++c) {
        if (t->mask && !t->mask[c]) continue;
        const int nbed = (int)t->nbedrock[c];
        t->zwt[c] = T_(t->zi, c, n + joff_zi);
        int k_zwt = nbed, sat_flag = 1;
        for (int k = nbed; k >= 1; --k) {
            T_(t->h2osoi_vol, c, k) = t_vol(t, c, k, joff);
            if (T_(t->h2osoi_vol, c, k) / T_(t->watsat, c, k) <= sat_lev) {
                k_zwt = k;
                sat_flag = 0;
                break;
            }
        }
        if (sat_flag == 1) k_zwt = 1;
        if (k_zwt == 1) {
            t->zwt[c] = T_(t->zi, c, 1 + joff_zi);
        } else if (k_zwt < nbed) {
            const double s1 = t_vol(t, c, k_zwt, joff) / T_(t->watsat, c, k_zwt);
            const double s2 = t_vol(t, c, k_zwt + 1, joff) / T_(t->watsat, c, k_zwt + 1);
            const double m = (T_(t->z, c, k_zwt + 1 + joff) - T_(t->z, c, k_zwt + joff)) / (s2 - s1);
            const double b = T_(t->z, c, k_zwt + 1 + joff) - m * s2;
            t->zwt[c] = JMAX(0.0, m * sat_lev + b);
        } else {
            t->zwt[c] = T_(t->zi, c, nbed + joff_zi);
        }
    }
    return CLMSOIL_OK;
}

/* _clamp_zwt_bedrock_kernel! (src/infrastructure/kernels.jl:209-218) */
int orc_clamp_zwt_to_bedrock(const clmsoil_theta_wt_args *t, const clmsoil_levels *lev) {
    if (!t || !lev) return CLMSOIL_ERR_NULL;
    if (!theta_nbedrock_ok(t, lev->nlevsoi)) return CLMSOIL_ERR_SIZE;
    for (int64_t c = 0; c < t->nc; ++c) {
        if (t->mask && !t->mask[c]) continue;
        const double zb = T_(t->zi, c, t->nbedrock[c] + lev->nlevsno + 1);
        if (t->zwt[c] > zb) t->zwt[c] = zb;
    }
    return CLMSOIL_OK;
}
