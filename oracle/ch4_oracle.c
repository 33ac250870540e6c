/*
 * ch4_oracle.c — CPU restatement of CLM.jl's methane and oxygen transport step
 * (ch4_tran!, src/biogeochem/methane.jl:2005-2092, with meth_khcc! :101-129 and
 * _ch4tran_column_kernel! :1663-2003).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 * Follows the reference statement by statement, one column at a time, with the
 * reference's scratch rows (k_h_cc, epsilon_t, source, conc_*_rel, diffus,
 * dm1_zm1, dp1_zp1, at/bt/ct/rt, cp/dp) as local arrays. smooth_min/smooth_max
 * on Float64 are min/max (src/infrastructure/smooth_ad.jl:91-133), here Julia's
 * min/max. Julia's Float64 ^ Float64 with an integer-valued exponent is the
 * integer power (pow_body, compensated squaring); other exponents call pow().
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "clmsoil_oracle.h"

#define CH4_MAXLEV 64

static const double TFRZ_ = 273.15, DENH2O_ = 1000.0, DENICE_ = 917.0;
static const double KH_TBASE_ = 298.0, RGASLATM_ = 0.0821;    /* varcon.jl:167, methane.jl:31 */
static const double C_H_INV_[2] = {1600.0, 1500.0};             /* varcon.jl:219 (CH4, O2) */
static const double KH_THETA_[2] = {714.29, 769.23};            /* varcon.jl:220 */
static const double D_CON_W_[2][3] = {{0.9798, 0.02986, 0.0004381},   /* varcon.jl:196-200 */
                                      {1.172, 0.03443, 0.0005048}};
static const double D_CON_G_[2][2] = {{0.1875, 0.0013}, {0.1759, 0.00117}};  /* :205-209 */

/* one-ulp jitter of exp/log/pow results, keyed by the arguments (0 = off) */
static uint64_t g_ch4_jitter = 0;
void orc_set_ch4_jitter(uint64_t seed) { g_ch4_jitter = seed; }

static double ch4_jit(double r, double x, double y) {
    if (!g_ch4_jitter || !isfinite(r) || r == 0.0) return r;
    uint64_t bx, by;
    memcpy(&bx, &x, 8);
    memcpy(&by, &y, 8);
    uint64_t h = bx ^ (by * 0x9E3779B97F4A7C15ull) ^ g_ch4_jitter;
    h ^= h >> 30; h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27; h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return nextafter(r, (h & 1u) ? INFINITY : -INFINITY);
}

/* debug outputs of the next call (NULL = none): the CH4 balance error before its
 * correction [nc]; per column and species the rows at, bt, ct, rt and the solved
 * relative concentration before the clip [nc][2][nlevsoi+1][5]; k_h_cc
 * [nc][2][nlevsoi+1] */
static double *g_dbg_err = NULL, *g_dbg_rows = NULL, *g_dbg_khcc = NULL;
void orc_ch4_debug(double *err_raw, double *rows, double *khcc) {
    g_dbg_err = err_raw;
    g_dbg_rows = rows;
    g_dbg_khcc = khcc;
}

/* Julia's pow_body(x, n) for n >= 0 (Base.Math): compensated power by squaring,
 * muladd fused. */
static double jl_pow_int(double x, int64_t n) {
    if (n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            const double err = fma(y, xnlo, x * ynlo);
            const double yy = x * y;
            ynlo = fma(x, y, -yy);
            y = yy;
            ynlo += err;
        }
        const double err = x * 2.0 * xnlo;
        const double xx = x * x;
        xnlo = fma(x, x, -xx);
        x = xx;
        xnlo += err;
        n >>= 1;
    }
    const double err = fma(y, xnlo, x * ynlo);
    return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

/* Julia's Float64 ^ Float64 for the exponents this step uses (>= 0) */
static double jl_pow(double x, double y) {
    if (x == 1.0) return 1.0;
    if (y == trunc(y) && fabs(y) < 0x1.8p62 && y >= 0.0) return jl_pow_int(x, (int64_t)y);
    return ch4_jit(pow(x, y), x, y);
}

double orc_ch4_jl_pow(double x, double y) { return jl_pow(x, y); }

/* Diffusivity of species s (0 CH4, 1 O2) at a level of temperature t (K) and
 * porosity ws (:1818-1840): gaseous at or above the water table (gas != 0; the
 * Millington-Quirk organic term and the Moldrup mineral term, weighted by the
 * organic fraction, vol_min the level's water content capped at ws), aqueous
 * below it (ws^satpow times the temperature polynomial, reduced by the liquid
 * share of the water when frozen); never below 1e-12. */
static double ch4_diffusivity(const clmsoil_ch4_tran_params *p, int s, int gas, double t,
                              double ws, double vol_min, double cellorg, double bsw, double liq,
                              double ice) {
    const double sn = 1.0e-12;
    const double tsc = t - TFRZ_;
    double d;
    if (gas) {
        const double f_a = 1.0 - vol_min / ws;
        const double e = ws - vol_min;
        const double om_frac = p->organic_max > 0.0
                                   ? orc_jl_min(p->om_frac_sf * cellorg / p->organic_max, 1.0)
                                   : 1.0;
        d = (D_CON_G_[s][0] + D_CON_G_[s][1] * tsc) * 1.0e-4 *
            (om_frac * jl_pow(f_a, 10.0 / 3.0) / (ws * ws) +
             (1.0 - om_frac) * (e * e) * jl_pow(f_a, 3.0 / bsw)) *
            p->scale_factor_gasdiff;
    } else {
        d = jl_pow(ws, p->satpow) *
            (D_CON_W_[s][0] + D_CON_W_[s][1] * tsc + D_CON_W_[s][2] * (tsc * tsc)) * 1.0e-9 *
            p->scale_factor_liqdiff;
        if (t <= TFRZ_) d *= (liq / DENH2O_ + sn) / (liq / DENH2O_ + ice / DENICE_ + sn);
    }
    return orc_jl_max(d, sn);
}

double orc_ch4_diffusivity(const clmsoil_ch4_tran_params *p, int s, int gas, double t, double ws,
                           double vol_min, double cellorg, double bsw, double liq, double ice) {
    return ch4_diffusivity(p, s, gas, t, ws, vol_min, cellorg, bsw, liq, ice);
}

int orc_ch4_tran(const clmsoil_ch4_tran_args *a, double dtime, const clmsoil_ch4_tran_params *p) {
    const int nl = p->nlevsoi;
    if (nl < 1 || nl > CH4_MAXLEV - 4) return CLMSOIL_ERR_SIZE;
    const double sn = 1.0e-12;
    const int64_t ld = a->ld;
    int bad = 0;
    for (int64_t c = 0; c < a->nc; ++c) {
        if (a->mask && !a->mask[c]) continue;
        const int64_t jw = a->jwt[c];
        if (jw < 0 || jw > nl) { bad = 1; continue; }
        const int jwt = (int)jw;
        const int64_t g = a->col_gridcell[c] - 1;
#define AT(arr, j) (arr)[c + (int64_t)((j) - 1) * ld]
        double k_h_cc[CH4_MAXLEV][2], eps_t[CH4_MAXLEV][2], source[CH4_MAXLEV][2];
        double conc_ch4_bef[CH4_MAXLEV], conc_ch4_rel[CH4_MAXLEV], conc_o2_rel[CH4_MAXLEV];
        double conc_ch4_rel_old[CH4_MAXLEV], conc_work[CH4_MAXLEV];
        double vol_min[CH4_MAXLEV], liqfrac[CH4_MAXLEV], diffus[CH4_MAXLEV];
        double dp1_zp1[CH4_MAXLEV], dm1_zm1[CH4_MAXLEV];
        double at[CH4_MAXLEV], bt[CH4_MAXLEV], ct[CH4_MAXLEV], rt[CH4_MAXLEV];
        double cpv[CH4_MAXLEV], dpv[CH4_MAXLEV], spec_grnd_cond[2];

        /* meth_khcc! (:101-117): jj = 1 is the ground (t_grnd), jj = j + 1 level j */
        for (int jj = 1; jj <= nl + 1; ++jj)
            for (int s = 0; s < 2; ++s) {
                const double t = jj == 1 ? a->t_grnd[c] : AT(a->t_soisno, jj - 1);
                const double arg = -C_H_INV_[s] * (1.0 / t - 1.0 / KH_TBASE_) +
                                   ch4_jit(log(KH_THETA_[s]), KH_THETA_[s], 1.0);
                const double k_h_inv = ch4_jit(exp(arg), arg, 0.0);
                k_h_cc[jj][s] = t / k_h_inv * RGASLATM_;
                if (g_dbg_khcc) g_dbg_khcc[(c * 2 + s) * (nl + 1) + (jj - 1)] = k_h_cc[jj][s];
            }

        /* competition for O2 and CH4 (:1671-1724) */
        for (int j = 1; j <= nl; ++j) {
            const double o2demand = AT(a->o2_decomp_depth, j) + AT(a->o2_oxid_depth, j);
            double o2s, ch4s;
            if (o2demand > 0.0) {
                if (AT(a->conc_o2, j) / dtime + AT(a->o2_aere_depth, j) > o2demand)
                    o2s = 1.0;
                else
                    o2s = (AT(a->conc_o2, j) / dtime + AT(a->o2_aere_depth, j)) / o2demand;
            } else {
                o2s = 1.0;
            }
            const double ch4demand =
                AT(a->ch4_oxid_depth, j) + AT(a->ch4_aere_depth, j) + AT(a->ch4_ebul_depth, j);
            if (ch4demand > 0.0)
                ch4s = orc_jl_min((AT(a->conc_ch4, j) / dtime + AT(a->ch4_prod_depth, j)) / ch4demand, 1.0);
            else
                ch4s = 1.0;
            if (o2s < 1.0 || ch4s < 1.0) {
                if (ch4s <= o2s) {
                    if (o2s < 1.0) {
                        const double o2demand2 = AT(a->o2_decomp_depth, j);
                        if (o2demand2 > 0.0)
                            o2s = orc_jl_min((AT(a->conc_o2, j) / dtime + AT(a->o2_aere_depth, j) -
                                              ch4s * AT(a->o2_oxid_depth, j)) / o2demand2, 1.0);
                        else
                            o2s = 1.0;
                    }
                    AT(a->ch4_oxid_depth, j) *= ch4s;
                    AT(a->o2_oxid_depth, j) *= ch4s;
                } else {
                    if (ch4s < 1.0) {
                        const double ch4demand2 = AT(a->ch4_aere_depth, j) + AT(a->ch4_ebul_depth, j);
                        if (ch4demand2 > 0.0)
                            ch4s = orc_jl_min((AT(a->conc_ch4, j) / dtime + AT(a->ch4_prod_depth, j) -
                                               o2s * AT(a->ch4_oxid_depth, j)) / ch4demand2, 1.0);
                        else
                            ch4s = 1.0;
                    }
                    AT(a->ch4_oxid_depth, j) *= o2s;
                    AT(a->o2_oxid_depth, j) *= o2s;
                }
            }
            AT(a->ch4_aere_depth, j) *= ch4s;
            AT(a->ch4_ebul_depth, j) *= ch4s;
            AT(a->o2_decomp_depth, j) *= o2s;
            AT(a->o2stress, j) = o2s;
            AT(a->ch4stress, j) = ch4s;
        }

        /* ebullition total (:1735-1739) */
        double ebul_total = 0.0;
        for (int j = 1; j <= nl; ++j) ebul_total += AT(a->ch4_ebul_depth, j) * AT(a->dz, j);
        a->ch4_ebul_total[c] = ebul_total;

        /* sources and epsilon_t (:1741-1766) */
        for (int j = 1; j <= nl; ++j) {
            vol_min[j] = orc_jl_min(AT(a->watsat, j), AT(a->h2osoi_vol, j));
            if (p->ch4frzout)
                liqfrac[j] = orc_jl_max(0.05, (AT(a->h2osoi_liq, j) / DENH2O_ + sn) /
                                                  (AT(a->h2osoi_liq, j) / DENH2O_ +
                                                   AT(a->h2osoi_ice, j) / DENICE_ + sn));
            else
                liqfrac[j] = 1.0;
            for (int s = 0; s < 2; ++s)
                eps_t[j][s] = j <= jwt ? AT(a->watsat, j) - (1.0 - k_h_cc[j + 1][s]) * vol_min[j] * liqfrac[j]
                                       : AT(a->watsat, j) * liqfrac[j];
            if (!p->use_aereoxid_prog) {
                AT(a->ch4_oxid_depth, j) += p->aereoxid * AT(a->ch4_aere_depth, j);
                AT(a->ch4_aere_depth, j) -= p->aereoxid * AT(a->ch4_aere_depth, j);
            }
            source[j][0] = AT(a->ch4_prod_depth, j) - AT(a->ch4_oxid_depth, j) -
                           AT(a->ch4_aere_depth, j) - AT(a->ch4_ebul_depth, j);
            source[j][1] = -AT(a->o2_oxid_depth, j) - AT(a->o2_decomp_depth, j) + AT(a->o2_aere_depth, j);
            conc_ch4_bef[j] = AT(a->conc_ch4, j);
        }

        /* aerenchyma surface flux (:1770-1774) */
        double surf_aere = 0.0;
        for (int j = 1; j <= nl; ++j) surf_aere += AT(a->ch4_aere_depth, j) * AT(a->dz, j);
        a->ch4_surf_aere[c] = surf_aere;

        /* ebullition into the water-table level (:1777-1779) */
        if (jwt != 0) source[jwt][0] += a->ch4_ebul_total[c] / AT(a->dz, jwt);

        /* relative concentrations, jj = 1 the atmosphere (:1782-1793) */
        const double catm1 = a->c_atm[g], catm2 = a->c_atm[g + a->ldg];
        conc_ch4_rel[1] = catm1;
        conc_o2_rel[1] = catm2;
        for (int j = 1; j <= nl; ++j) {
            conc_ch4_rel[j + 1] = AT(a->conc_ch4, j) / eps_t[j][0];
            conc_o2_rel[j + 1] = AT(a->conc_o2, j) / eps_t[j][1];
        }
        for (int jj = 1; jj <= nl + 1; ++jj) conc_ch4_rel_old[jj] = conc_ch4_rel[jj];

        const int nlevs = nl + 1;
        for (int s = 0; s < 2; ++s) {
            const double catm = s == 0 ? catm1 : catm2;
            for (int jj = 1; jj <= nlevs; ++jj) conc_work[jj] = s == 0 ? conc_ch4_rel[jj] : conc_o2_rel[jj];

            /* snow/pond resistance and ground conductance (:1800-1815) */
            if (a->grnd_ch4_cond[c] < sn && s == 0) a->grnd_ch4_cond[c] = sn;
            const double snowres = 0.0;
            double pondres = 0.0;
            if (!p->lake && p->sat == 1 && a->frac_h2osfc[c] > 0.0) {
                if (a->t_h2osfc[c] >= TFRZ_) {
                    const double tsc = a->t_h2osfc[c] - TFRZ_;
                    const double ponddiff = (D_CON_W_[s][0] + D_CON_W_[s][1] * tsc + D_CON_W_[s][2] * (tsc * tsc)) *
                                            1.0e-9 * p->scale_factor_liqdiff;
                    const double pondz = a->h2osfc[c] / 1000.0 / a->frac_h2osfc[c];
                    pondres = pondz / ponddiff;
                } else if (a->h2osfc[c] / a->frac_h2osfc[c] > p->capthick) {
                    pondres = 1.0 / sn;
                }
            }
            spec_grnd_cond[s] = 1.0 / (1.0 / a->grnd_ch4_cond[c] + snowres + pondres);

            /* gas / liquid diffusivity (:1818-1840) */
            for (int j = 1; j <= nl; ++j)
                diffus[j] = ch4_diffusivity(p, s, j <= jwt, AT(a->t_soisno, j), AT(a->watsat, j),
                                            vol_min[j], AT(a->cellorg, j), AT(a->bsw, j),
                                            AT(a->h2osoi_liq, j), AT(a->h2osoi_ice, j));

            /* dm1_zm1 / dp1_zp1 (:1843-1871) */
            for (int j = 1; j <= nl; ++j) {
                const double dzj = AT(a->dz, j);
                if (j == 1 && j != jwt && j != jwt + 1) {
                    dm1_zm1[j] = 1.0 / (1.0 / spec_grnd_cond[s] + dzj / (diffus[j] * 2.0));
                    dp1_zp1[j] = j < nl ? 2.0 / (dzj / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1]) : 0.0;
                } else if (j == 1 && j == jwt) {
                    dm1_zm1[j] = 1.0 / (1.0 / spec_grnd_cond[s] + dzj / (diffus[j] * 2.0));
                    dp1_zp1[j] = j < nl ? 2.0 / (dzj * k_h_cc[j + 1][s] / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1])
                                        : 0.0;
                } else if (j == 1) {
                    dm1_zm1[j] = 1.0 / (k_h_cc[j][s] / spec_grnd_cond[s] + dzj / (diffus[j] * 2.0));
                    dp1_zp1[j] = j < nl ? 2.0 / (dzj / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1]) : 0.0;
                } else if (j < nl && j != jwt && j != jwt + 1) {
                    dm1_zm1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j - 1) / diffus[j - 1]);
                    dp1_zp1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1]);
                } else if (j < nl && j == jwt) {
                    dm1_zm1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j - 1) / diffus[j - 1]);
                    dp1_zp1[j] = 2.0 / (dzj * k_h_cc[j + 1][s] / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1]);
                } else if (j < nl) {
                    dm1_zm1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j - 1) * k_h_cc[j][s] / diffus[j - 1]);
                    dp1_zp1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j + 1) / diffus[j + 1]);
                } else if (j != jwt + 1) {
                    dm1_zm1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j - 1) / diffus[j - 1]);
                } else {
                    dm1_zm1[j] = 2.0 / (dzj / diffus[j] + AT(a->dz, j - 1) * k_h_cc[j][s] / diffus[j - 1]);
                }
            }

            /* the tridiagonal system, jj = j + 1 (:1874-1924) */
            for (int j = 0; j <= nl; ++j) {
                const int jj = j + 1;
                if (j == 0) {
                    at[jj] = 0.0;
                    bt[jj] = 1.0;
                    ct[jj] = 0.0;
                    rt[jj] = catm;
                    continue;
                }
                const double dzj = AT(a->dz, j), e = eps_t[j][s], src = source[j][s];
                const double dm = dm1_zm1[j];
                if (j < nl && j == jwt) {
                    const double dp = dp1_zp1[j], k1 = k_h_cc[j + 1][s];
                    at[jj] = -0.5 / dzj * dm;
                    bt[jj] = e / dtime + 0.5 / dzj * (dp * k1 + dm);
                    ct[jj] = -0.5 / dzj * dp;
                    rt[jj] = e / dtime * conc_work[jj] +
                             0.5 / dzj * (dp * (conc_work[jj + 1] - conc_work[jj] * k1) -
                                          dm * (conc_work[jj] - conc_work[jj - 1])) + src;
                } else if (j < nl && j == jwt + 1) {
                    const double dp = dp1_zp1[j], k0 = k_h_cc[j][s];
                    at[jj] = -0.5 / dzj * dm * k0;
                    bt[jj] = e / dtime + 0.5 / dzj * (dp + dm);
                    ct[jj] = -0.5 / dzj * dp;
                    rt[jj] = e / dtime * conc_work[jj] +
                             0.5 / dzj * (dp * (conc_work[jj + 1] - conc_work[jj]) -
                                          dm * (conc_work[jj] - conc_work[jj - 1] * k0)) + src;
                } else if (j < nl) {
                    const double dp = dp1_zp1[j];
                    at[jj] = -0.5 / dzj * dm;
                    bt[jj] = e / dtime + 0.5 / dzj * (dp + dm);
                    ct[jj] = -0.5 / dzj * dp;
                    rt[jj] = e / dtime * conc_work[jj] +
                             0.5 / dzj * (dp * (conc_work[jj + 1] - conc_work[jj]) -
                                          dm * (conc_work[jj] - conc_work[jj - 1])) + src;
                } else if (j == nl && j == jwt + 1) {
                    const double k0 = k_h_cc[j][s];
                    at[jj] = -0.5 / dzj * dm * k0;
                    bt[jj] = e / dtime + 0.5 / dzj * dm;
                    ct[jj] = 0.0;
                    rt[jj] = e / dtime * conc_work[jj] +
                             0.5 / dzj * (-dm * (conc_work[jj] - conc_work[jj - 1] * k0)) + src;
                } else {
                    at[jj] = -0.5 / dzj * dm;
                    bt[jj] = e / dtime + 0.5 / dzj * dm;
                    ct[jj] = 0.0;
                    rt[jj] = e / dtime * conc_work[jj] + 0.5 / dzj * (-dm * (conc_work[jj] - conc_work[jj - 1])) + src;
                }
            }

            /* Thomas solve from jj = 1 (:1940-1957) */
            double cp_prev = ct[1] / bt[1], dp_prev = rt[1] / bt[1];
            cpv[1] = cp_prev;
            dpv[1] = dp_prev;
            for (int jj = 2; jj <= nlevs; ++jj) {
                const double denom = bt[jj] - at[jj] * cp_prev;
                cp_prev = ct[jj] / denom;
                dp_prev = (rt[jj] - at[jj] * dp_prev) / denom;
                cpv[jj] = cp_prev;
                dpv[jj] = dp_prev;
            }
            double u_next = dp_prev;
            conc_work[nlevs] = u_next;
            for (int jj = nlevs - 1; jj >= 1; --jj) {
                u_next = dpv[jj] - cpv[jj] * u_next;
                conc_work[jj] = u_next;
            }
            if (g_dbg_rows)
                for (int jj = 1; jj <= nlevs; ++jj) {
                    double *r = g_dbg_rows + (((c * 2 + s) * (int64_t)nlevs) + (jj - 1)) * 5;
                    r[0] = at[jj];
                    r[1] = bt[jj];
                    r[2] = ct[jj];
                    r[3] = rt[jj];
                    r[4] = conc_work[jj];
                }

            if (s == 0) {
                /* surface fluxes and the negative-CH4 clip (:1959-1982) */
                if (jwt != 0) {
                    a->ch4_surf_diff[c] = dm1_zm1[1] * ((conc_work[2] + conc_ch4_rel_old[2]) / 2.0 - catm);
                    a->ch4_surf_ebul[c] = 0.0;
                } else {
                    a->ch4_surf_diff[c] =
                        dm1_zm1[1] * ((conc_work[2] + conc_ch4_rel_old[2]) / 2.0 - catm * k_h_cc[1][0]);
                    a->ch4_surf_ebul[c] = a->ch4_ebul_total[c];
                }
                double deficit_tot = 0.0;
                for (int j = 1; j <= nl; ++j) {
                    const int jj = j + 1;
                    if (conc_work[jj] < 0.0) {
                        deficit_tot += -conc_work[jj] * eps_t[j][0] * AT(a->dz, j);
                        conc_work[jj] = 0.0;
                    }
                }
                a->ch4_surf_diff[c] -= deficit_tot / dtime;
                for (int jj = 1; jj <= nlevs; ++jj) conc_ch4_rel[jj] = conc_work[jj];
            } else {
                /* O2 bounds (:1983-1989) */
                for (int j = 1; j <= nl; ++j) {
                    const int jj = j + 1;
                    conc_work[jj] = orc_jl_max(conc_work[jj], 1.0e-12);
                    conc_work[jj] = orc_jl_min(conc_work[jj], catm2 / eps_t[j][1]);
                }
                for (int jj = 1; jj <= nlevs; ++jj) conc_o2_rel[jj] = conc_work[jj];
            }
        }

        /* absolute concentrations (:1993-1996) */
        for (int j = 1; j <= nl; ++j) {
            AT(a->conc_ch4, j) = conc_ch4_rel[j + 1] * eps_t[j][0];
            AT(a->conc_o2, j) = conc_o2_rel[j + 1] * eps_t[j][1];
        }

        /* CH4 balance and its correction (:1990-2001) */
        double errch4 = 0.0;
        for (int j = 1; j <= nl; ++j) {
            errch4 += (AT(a->conc_ch4, j) - conc_ch4_bef[j]) * AT(a->dz, j);
            errch4 -= AT(a->ch4_prod_depth, j) * AT(a->dz, j) * dtime;
            errch4 += AT(a->ch4_oxid_depth, j) * AT(a->dz, j) * dtime;
        }
        errch4 += (a->ch4_surf_aere[c] + a->ch4_surf_ebul[c] + a->ch4_surf_diff[c]) * dtime;
        if (g_dbg_err) g_dbg_err[c] = errch4;
        if (fabs(errch4) < 1.0e-8) a->ch4_surf_diff[c] -= errch4 / dtime;
        a->grnd_ch4_cond[c] = spec_grnd_cond[0];
#undef AT
    }
    return bad ? CLMSOIL_ERR_ARG : CLMSOIL_OK;
}
