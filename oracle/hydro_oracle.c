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
