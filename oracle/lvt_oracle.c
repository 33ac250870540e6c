/*
 * lvt_oracle.c — CPU restatement of CLM.jl's vertical transport of the soil
 * organic matter pools (litter_vert_transp!, src/biogeochem/litter_vert_transp.jl).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 * Follows the reference statement by statement: the node geometry built on the
 * host (:566-575), the coefficient kernel over all columns, then for each tracer
 * type and pool the column kernel over all columns, each with its own scratch
 * rows (diffus, adv_flux, conc_trcr, D/F/Pe, a/b/c/r, cp, dp).
 */
#include <math.h>
#include <stdlib.h>

#include "clmsoil_oracle.h"

/* Julia's x^5 for Float64: Base.Math.pow_body (power by squaring with the
 * rounding error of each product carried, two_mul through fma). For n = 5 the
 * odd step sets y = x exactly, two squarings give X2 = fl(X1^2) with error
 * xnlo = e2 + X1*2*e1, and the result is muladd(X2, x, muladd(x, xnlo, X2*0)),
 * muladd fused as on FMA hardware. */
double orc_jl_pow5(double x) {
    const double x1 = x * x, e1 = fma(x, x, -x1);
    const double err1 = x1 * 2.0 * e1;
    const double x2 = x1 * x1, e2 = fma(x1, x1, -x2);
    const double xnlo = e2 + err1;
    const double err = fma(x, xnlo, x2 * 0.0);
    if (!(isfinite(x2) && isfinite(err))) return x2 * x;
    return fma(x2, x, err);
}

/* patankar_A (:66-70): max(0, (1 - 0.1|pe|)^5) */
static double patankar_A(double pe) { return orc_jl_max(0.0, orc_jl_pow5(1.0 - 0.1 * fabs(pe))); }

/* _spinup_lat_term (src/biogeochem/decomp_bgc.jl:400-401) */
static double spinup_lat_term(double lat) { return 1.0 + 50.0 / (1.0 + exp(-0.15 * (fabs(lat) - 60.0))); }

int orc_litter_vert_transp(const clmsoil_lvt_args *a, double dtime, const clmsoil_lvt_params *p) {
    if (!a || !p) return CLMSOIL_ERR_NULL;
    if (p->use_soil_matrixcn) return CLMSOIL_ERR_METHOD;
    const int nd = p->nlevdecomp, nfull = p->nlevdecomp_full, npools = p->npools;
    if (nd < 1 || nd > 62 || nfull < nd || npools < 1 || npools > CLMSOIL_LVT_MAXPOOLS ||
        p->ntypes < 1 || p->ntypes > CLMSOIL_LVT_MAXTYPES)
        return CLMSOIL_ERR_SIZE;
    const int64_t nc = a->nc, ld = a->ld;
    for (int64_t c = 0; c < nc; ++c)
        if ((!a->mask_bgc_soilc || a->mask_bgc_soilc[c]) && (a->nbedrock[c] < 1 || a->nbedrock[c] > nd))
            return CLMSOIL_ERR_SIZE;
    /* 1-based views: zisoi[j] = interface j-1 (zisoi[1] = surface) as the reference's zisoi_vals */
    const double *zisoi = a->zisoi - 1, *dzd = a->dzsoi_decomp - 1;
    double zsoi_ext[64], dz_node[64];                     /* [1..nd+1] */
    for (int j = 1; j <= nd; ++j) zsoi_ext[j] = a->zsoi[j - 1];
    zsoi_ext[nd + 1] = zisoi[nd + 1];
    dz_node[1] = zsoi_ext[1];
    for (int j = 2; j <= nd + 1; ++j) dz_node[j] = zsoi_ext[j] - zsoi_ext[j - 1];
    const int n1 = nd + 1, nj = nd + 2;
#define CM(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)ld]
#define ON(c) (!a->mask_bgc_soilc || a->mask_bgc_soilc[c])
    /* _lvt_coef_kernel! (:93-136) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        const int64_t nbr = a->nbedrock[c];
        const double alt = orc_jl_max(a->altmax[c], a->altmax_lastyear[c]);
        for (int j = 1; j <= nd + 1; ++j) {
            double dif = 0.0, adv = 0.0;
            if (alt <= p->max_altdepth_cryoturbation && alt > 0.0) {
                if (j <= nbr + 1) {
                    if (zisoi[j + 1] < alt) {
                        dif = p->cryoturb_diffusion_k;
                    } else {
                        dif = orc_jl_max(p->cryoturb_diffusion_k *
                                             (1.0 - (zisoi[j + 1] - alt) /
                                                        (orc_jl_min(p->max_depth_cryoturb, zisoi[nbr + 2]) - alt)),
                                         0.0);
                    }
                }
            } else if (alt > 0.0) {
                if (j <= nbr + 1) {
                    adv = p->som_adv_flux;
                    dif = p->som_diffus;
                }
            }
            CM(a->som_diffus_coef, c, j) = dif;
            CM(a->som_adv_coef, c, j) = adv;
        }
    }
    /* scratch rows, [nc, *] as the reference's _LvtScr */
    double *buf = (double *)calloc((size_t)(nc > 0 ? nc : 1) * (size_t)(8 * n1 + 7 * nj), sizeof(double));
    if (!buf) return CLMSOIL_ERR_SIZE;
    double *diffus = buf, *adv_flux = diffus + (size_t)nc * n1, *d_p1_zp1 = adv_flux + (size_t)nc * n1,
           *d_m1_zm1 = d_p1_zp1 + (size_t)nc * n1, *f_p1 = d_m1_zm1 + (size_t)nc * n1,
           *f_m1 = f_p1 + (size_t)nc * n1, *pe_p1 = f_m1 + (size_t)nc * n1, *pe_m1 = pe_p1 + (size_t)nc * n1,
           *conc_trcr = pe_m1 + (size_t)nc * n1, *a_tri = conc_trcr + (size_t)nc * nj,
           *b_tri = a_tri + (size_t)nc * nj, *c_tri = b_tri + (size_t)nc * nj, *r_tri = c_tri + (size_t)nc * nj,
           *cp = r_tri + (size_t)nc * nj, *dp = cp + (size_t)nc * nj;
#define S(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)nc]
#define P3(X, c, j, s) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)ld + ((size_t)(s) - 1u) * (size_t)ld * (size_t)nfull]
    const double epsilon = 1.0e-30;
    for (int t = 0; t < p->ntypes; ++t) {
        double *conc = a->conc[t], *tend = a->tendency[t];
        const double *source = a->source[t];
        for (int s = 1; s <= npools; ++s) {
            /* _lvt_column_kernel! (:142-471), use_soil_matrixcn = false */
            for (int64_t c = 0; c < nc; ++c) {
                if (!ON(c)) continue;
                const int nbr = (int)a->nbedrock[c];
                if (p->is_cwd[s - 1]) {
                    for (int j = 1; j <= nd; ++j) S(conc_trcr, c, j + 1) = P3(conc, c, j, s) + P3(source, c, j, s);
                } else {
                    double spinup_term = 1.0;
                    if (p->spinup_state >= 1) spinup_term = p->spinup_factor[s - 1];
                    if (fabs(spinup_term - 1.0) > 1.0e-6)
                        spinup_term = spinup_term * spinup_lat_term(a->latdeg[a->col_gridcell[c] - 1]);
                    for (int j = 1; j <= nd + 1; ++j) {
                        S(adv_flux, c, j) = fabs(CM(a->som_adv_coef, c, j)) * spinup_term < epsilon
                                                ? epsilon : CM(a->som_adv_coef, c, j) * spinup_term;
                        S(diffus, c, j) = fabs(CM(a->som_diffus_coef, c, j)) * spinup_term < epsilon
                                              ? epsilon : CM(a->som_diffus_coef, c, j) * spinup_term;
                    }
                    S(conc_trcr, c, 1) = 0.0;
                    for (int jj = nbr + 1; jj <= nd + 1; ++jj) S(conc_trcr, c, jj + 1) = 0.0;
                    for (int j = 1; j <= nd + 1; ++j) {
                        if (j <= nd) S(conc_trcr, c, j + 1) = P3(conc, c, j, s);
                        if (j == 1) {
                            S(d_m1_zm1, c, j) = 0.0;
                            const double w_p1 = (zsoi_ext[j + 1] - zisoi[j + 1]) / dz_node[j + 1];
                            double d_p1;
                            if (S(diffus, c, j + 1) > 0.0 && S(diffus, c, j) > 0.0)
                                d_p1 = 1.0 / ((1.0 - w_p1) / S(diffus, c, j) + w_p1 / S(diffus, c, j + 1));
                            else
                                d_p1 = 0.0;
                            S(d_p1_zp1, c, j) = d_p1 / dz_node[j + 1];
                            S(f_m1, c, j) = S(adv_flux, c, j);
                            S(f_p1, c, j) = S(adv_flux, c, j + 1);
                            S(pe_m1, c, j) = 0.0;
                            S(pe_p1, c, j) = S(f_p1, c, j) / S(d_p1_zp1, c, j);
                        } else if (j >= nbr + 1) {
                            const double w_m1 = (zisoi[j] - zsoi_ext[j - 1]) / dz_node[j];
                            double d_m1;
                            if (S(diffus, c, j) > 0.0 && S(diffus, c, j - 1) > 0.0)
                                d_m1 = 1.0 / ((1.0 - w_m1) / S(diffus, c, j) + w_m1 / S(diffus, c, j - 1));
                            else
                                d_m1 = 0.0;
                            S(d_m1_zm1, c, j) = d_m1 / dz_node[j];
                            S(d_p1_zp1, c, j) = S(d_m1_zm1, c, j);
                            S(f_m1, c, j) = S(adv_flux, c, j);
                            S(f_p1, c, j) = 0.0;
                            S(pe_m1, c, j) = S(f_m1, c, j) / S(d_m1_zm1, c, j);
                            S(pe_p1, c, j) = S(f_p1, c, j) / S(d_p1_zp1, c, j);
                        } else {
                            const double w_m1 = (zisoi[j] - zsoi_ext[j - 1]) / dz_node[j];
                            double d_m1, d_p1;
                            if (S(diffus, c, j - 1) > 0.0 && S(diffus, c, j) > 0.0)
                                d_m1 = 1.0 / ((1.0 - w_m1) / S(diffus, c, j) + w_m1 / S(diffus, c, j - 1));
                            else
                                d_m1 = 0.0;
                            const double w_p1 = (zsoi_ext[j + 1] - zisoi[j + 1]) / dz_node[j + 1];
                            if (S(diffus, c, j + 1) > 0.0 && S(diffus, c, j) > 0.0)
                                d_p1 = 1.0 / ((1.0 - w_p1) / S(diffus, c, j) + w_p1 / S(diffus, c, j + 1));
                            else
                                d_p1 = (1.0 - w_m1) * S(diffus, c, j) + w_p1 * S(diffus, c, j + 1);
                            S(d_m1_zm1, c, j) = d_m1 / dz_node[j];
                            S(d_p1_zp1, c, j) = d_p1 / dz_node[j + 1];
                            S(f_m1, c, j) = S(adv_flux, c, j);
                            S(f_p1, c, j) = S(adv_flux, c, j + 1);
                            S(pe_m1, c, j) = S(f_m1, c, j) / S(d_m1_zm1, c, j);
                            S(pe_p1, c, j) = S(f_p1, c, j) / S(d_p1_zp1, c, j);
                        }
                    }
                    for (int j = 0; j <= nd + 1; ++j) {
                        const int jj = j + 1;
                        const double a_p_0 = (j > 0 && j < nd + 1) ? dzd[j] / dtime : 0.0;
                        if (j == 0) {
                            S(a_tri, c, jj) = 0.0;
                            S(b_tri, c, jj) = 1.0;
                            S(c_tri, c, jj) = -1.0;
                            S(r_tri, c, jj) = 0.0;
                        } else if (j < nd + 1) {
                            S(a_tri, c, jj) = -(S(d_m1_zm1, c, j) * patankar_A(S(pe_m1, c, j)) +
                                                orc_jl_max(S(f_m1, c, j), 0.0));
                            S(c_tri, c, jj) = -(S(d_p1_zp1, c, j) * patankar_A(S(pe_p1, c, j)) +
                                                orc_jl_max(-S(f_p1, c, j), 0.0));
                            S(b_tri, c, jj) = -S(a_tri, c, jj) - S(c_tri, c, jj) + a_p_0;
                            if (j == 1)
                                S(r_tri, c, jj) = P3(source, c, j, s) * dzd[j] / dtime +
                                                  (a_p_0 - S(adv_flux, c, j)) * S(conc_trcr, c, jj);
                            else
                                S(r_tri, c, jj) = P3(source, c, j, s) * dzd[j] / dtime +
                                                  a_p_0 * S(conc_trcr, c, jj);
                        } else {
                            S(a_tri, c, jj) = -1.0;
                            S(b_tri, c, jj) = 1.0;
                            S(c_tri, c, jj) = 0.0;
                            S(r_tri, c, jj) = 0.0;
                        }
                    }
                    for (int j = 1; j <= nd; ++j)
                        P3(tend, c, j, s) = 0.0 - (S(conc_trcr, c, j + 1) + P3(source, c, j, s));
                    double cp_prev = S(c_tri, c, 1) / S(b_tri, c, 1);
                    double dp_prev = S(r_tri, c, 1) / S(b_tri, c, 1);
                    S(cp, c, 1) = cp_prev;
                    S(dp, c, 1) = dp_prev;
                    for (int jj = 2; jj <= nj; ++jj) {
                        const double denom = S(b_tri, c, jj) - S(a_tri, c, jj) * cp_prev;
                        cp_prev = S(c_tri, c, jj) / denom;
                        dp_prev = (S(r_tri, c, jj) - S(a_tri, c, jj) * dp_prev) / denom;
                        S(cp, c, jj) = cp_prev;
                        S(dp, c, jj) = dp_prev;
                    }
                    double u_next = dp_prev;
                    S(conc_trcr, c, nj) = u_next;
                    for (int jj = nj - 1; jj >= 1; --jj) {
                        u_next = S(dp, c, jj) - S(cp, c, jj) * u_next;
                        S(conc_trcr, c, jj) = u_next;
                    }
                    for (int j = 1; j <= nd; ++j)
                        P3(tend, c, j, s) = (P3(tend, c, j, s) + S(conc_trcr, c, j + 1)) / dtime;
                }
                for (int j = 1; j <= nd; ++j) {
                    P3(conc, c, j, s) = S(conc_trcr, c, j + 1);
                    if (j > nbr) {
                        P3(conc, c, nbr, s) = P3(conc, c, nbr, s) + S(conc_trcr, c, j + 1) * (dzd[j] / dzd[nbr]);
                        P3(conc, c, j, s) = 0.0;
                    }
                }
            }
        }
    }
#undef CM
#undef ON
#undef S
#undef P3
    free(buf);
    return CLMSOIL_OK;
}
