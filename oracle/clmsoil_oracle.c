/*
 * clmsoil_oracle.c — CPU restatement of CLM.jl's per-column soil solves.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline); see the header.
 *
 * Indexing: reference arrays are Julia column-major Matrix{Float64} `A[c, j]`.
 * Here the column `c` is 0-based and the level `j` keeps the reference's 1-based
 * value, so `A(X, c, j)` reads exactly the element the reference line reads.
 * Constants: src/constants/varcon.jl:15-70.
 */
#include "clmsoil_oracle.h"

#include <math.h>
/* The FP64 log/log10/exp/pow of the HIP library (one source for both sides:
 * Tang's table-driven method); the pins of those functions are in the tests. */
#include "../clm.jl_amd/csrc/clm_math.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define A(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)ld]

/* varcon.jl:15-70 */
static const double GRAV = 9.80616, CPLIQ = 4188.0, CPICE = 2117.27,
                    HFUS = 3.337e5, DENH2O = 1000.0, DENICE = 917.0,
                    TFRZ = 273.15, TKAIR = 0.023, TKICE = 2.290, TKWAT = 0.57,
                    CAPR = 0.34, CNFAC = 0.5, THK_BEDROCK = 3.0,
                    CSOL_BEDROCK = 2.0e6;
static const double M_TO_MM = 1.0e3;          /* soil_water_movement.jl:37 */
static const double THIN_SFCLAYER = 1.0e-6;   /* soil_temperature.jl:26   */

/* The shared FP64 functions, element-wise over arrays (pinned by the tests). */
double orc_jl_max(double x, double y);
double orc_jl_min(double x, double y);
void orc_clm_math(int op, const double *x, const double *y, double *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (op) {
            case 0: out[i] = clm_pow(x[i], y[i]); break;
            case 1: out[i] = clm_log(x[i]); break;
            case 2: out[i] = clm_log10(x[i]); break;
            case 3: out[i] = clm_exp(x[i]); break;
            case 4: out[i] = orc_jl_max(x[i], y[i]); break;
            case 5: out[i] = orc_jl_min(x[i], y[i]); break;
            case 6: out[i] = clm_pow_pos(x[i], y[i]); break;
            case 7: out[i] = clm_pow_base(clm_log_ext(x[i]), y[i]); break;
            default: out[i] = clm_log10_pos(x[i]); break;
        }
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Julia Base.max/min for IEEE floats: diff-based select, NaN propagates through
 * the difference, max(-0.0, 0.0) == 0.0 and min(0.0, -0.0) == -0.0. */
double orc_jl_max(double x, double y) {
    double d = x - y;
    double am = signbit(d) ? y : x;
    return (isnan(x) || isnan(y)) ? d : am;
}
double orc_jl_min(double x, double y) {
    double d = x - y;
    double am = signbit(d) ? x : y;
    return (isnan(x) || isnan(y)) ? d : am;
}
#define JMAX orc_jl_max
#define JMIN orc_jl_min

/* ---------------------------------------------------------------------- */
/* SWRC (swrc_clapp_hornberg.jl)                                           */
/* ---------------------------------------------------------------------- */
void orc_soil_hk(double hksat, double bsw, double s, double imped, double *hk,
                 double *dhkds) {
    /* :47-53  hk = imped*hksat*s^(2b+3); dhkds = (2b+3)*hk/s */
    double h = imped * hksat * clm_pow(s, 2.0 * bsw + 3.0);
    *hk = h;
    *dhkds = (2.0 * bsw + 3.0) * h / s;
}
void orc_soil_suction(double bsw, double sucsat, double s, double *smp,
                      double *dsmpds) {
    /* :80-86  smp = -sucsat*s^(-b); dsmpds = -b*smp/s */
    double v = -sucsat * clm_pow(s, -bsw);
    *smp = v;
    *dsmpds = -bsw * v / s;
}
double orc_soil_suction_inverse(double bsw, double sucsat, double smp_target) {
    /* :113-115 */
    return clm_pow(-smp_target / sucsat, -1.0 / bsw);
}
double orc_ice_impedance(double icefrac, double e_ice) {
    /* soil_water_movement.jl:722-724 */
    return clm_pow(10.0, -e_ice * icefrac);
}

/* ---------------------------------------------------------------------- */
/* scalar helpers                                                          */
/* ---------------------------------------------------------------------- */
void orc_tridiagonal_col(double *u, const double *a, const double *b,
                         const double *c, const double *r, int jtop, int lbj,
                         int ubj) {
    /* soil_water_movement.jl:736-773; arrays hold levels lbj..ubj at 0..n-1 */
#define V(X, j) (X)[(j) - lbj]
    int n = ubj - lbj + 1;
    double *gam = (double *)calloc((size_t)n, sizeof(double));
    double bet = V(b, jtop);
    if (fabs(bet) < 1.0e-30) bet = 1.0e-30;
    for (int j = lbj; j <= ubj; ++j) {
        if (j >= jtop) {
            if (j == jtop) {
                V(u, j) = V(r, j) / bet;
            } else {
                V(gam, j) = V(c, j - 1) / bet;
                bet = V(b, j) - V(a, j) * V(gam, j);
                if (fabs(bet) < 1.0e-30) bet = copysign(1.0e-30, bet == 0.0 ? 1.0 : bet);
                V(u, j) = (V(r, j) - V(a, j) * V(u, j - 1)) / bet;
            }
        }
    }
    for (int j = ubj - 1; j >= lbj; --j)
        if (j >= jtop) V(u, j) = V(u, j) - V(gam, j + 1) * V(u, j + 1);
    free(gam);
#undef V
}

void orc_compute_rhs_moisture_form(int nlayers, const double *vts,
                                   const double *qin, const double *qout,
                                   const double *dt_dz, double *rmx) {
    /* soil_water_movement.jl:1080-1097 */
    for (int j = 0; j < nlayers; ++j) {
        double fluxNet = qin[j] - qout[j] - vts[j];
        rmx[j] = -fluxNet * dt_dz[j];
    }
}

void orc_compute_lhs_moisture_form(int nlayers, const double *dt_dz,
                                   const double *dqidw0, const double *dqidw1,
                                   const double *dqodw1, const double *dqodw2,
                                   double *amx, double *bmx, double *cmx) {
    /* soil_water_movement.jl:1108-1140 */
    int j = 0;
    amx[j] = 0.0;
    bmx[j] = -1.0 - (-dqidw1[j] + dqodw1[j]) * dt_dz[j];
    cmx[j] = -dqodw2[j] * dt_dz[j];
    for (j = 1; j < nlayers - 1; ++j) {
        amx[j] = dqidw0[j] * dt_dz[j];
        bmx[j] = -1.0 - (-dqidw1[j] + dqodw1[j]) * dt_dz[j];
        cmx[j] = -dqodw2[j] * dt_dz[j];
    }
    j = nlayers - 1;
    amx[j] = dqidw0[j] * dt_dz[j];
    bmx[j] = -1.0 - (-dqidw1[j] + dqodw1[j]) * dt_dz[j];
    cmx[j] = 0.0;
}

void orc_tridiagonal_solve(double *u, const double *a, const double *b,
                           const double *c, const double *r, int jtop, int nlevs) {
    /* tridiagonal.jl:35-63 (1-based jtop, arrays 0-based of length nlevs) */
    double *cp = (double *)malloc(sizeof(double) * (size_t)nlevs);
    double *dp = (double *)malloc(sizeof(double) * (size_t)nlevs);
    int j = jtop - 1;
    cp[j] = c[j] / b[j];
    dp[j] = r[j] / b[j];
    for (int jj = j + 1; jj < nlevs; ++jj) {
        double denom = b[jj] - a[jj] * cp[jj - 1];
        cp[jj] = c[jj] / denom;
        dp[jj] = (r[jj] - a[jj] * dp[jj - 1]) / denom;
    }
    u[nlevs - 1] = dp[nlevs - 1];
    for (int jj = nlevs - 2; jj >= j; --jj) u[jj] = dp[jj] - cp[jj] * u[jj + 1];
    free(cp);
    free(dp);
}

/* ---------------------------------------------------------------------- */
/* batched tridiagonal (tridiagonal.jl:86-142)                             */
/* ---------------------------------------------------------------------- */
int orc_tridiagonal_multi(double *u, const double *a, const double *b,
                          const double *c, const double *r, const int64_t *jtop,
                          const uint8_t *mask, int64_t ncols, int64_t nlevs,
                          int64_t ld) {
    if (!u || !a || !b || !c || !r || !jtop) return CLMSOIL_ERR_NULL;
    if (ncols < 0 || nlevs < 1 || ld < ncols) return CLMSOIL_ERR_SIZE;
#pragma omp parallel for schedule(static)
    for (int64_t col = 0; col < ncols; ++col) {
        if (mask && !mask[col]) continue;
        int64_t j = jtop[col];
        if (j < 1 || j > nlevs) continue;
        double *cp = (double *)malloc(sizeof(double) * (size_t)(nlevs + 1));
        double *dp = (double *)malloc(sizeof(double) * (size_t)(nlevs + 1));
        double bet = A(b, col, j);
        const double tiny = 1e-30;
        if (fabs(bet) < tiny) bet = tiny;
        double cp_prev = A(c, col, j) / bet;
        double dp_prev = A(r, col, j) / bet;
        cp[j] = cp_prev;
        dp[j] = dp_prev;
        for (int64_t jj = j + 1; jj <= nlevs; ++jj) {
            double denom = A(b, col, jj) - A(a, col, jj) * cp_prev;
            if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0 ? 1.0 : denom);
            cp_prev = A(c, col, jj) / denom;
            dp_prev = (A(r, col, jj) - A(a, col, jj) * dp_prev) / denom;
            cp[jj] = cp_prev;
            dp[jj] = dp_prev;
        }
        double u_next = dp_prev;
        A(u, col, nlevs) = u_next;
        for (int64_t jj = nlevs - 1; jj >= j; --jj) {
            u_next = dp[jj] - cp[jj] * u_next;
            A(u, col, jj) = u_next;
        }
        free(cp);
        free(dp);
    }
    return CLMSOIL_OK;
}

/* ---------------------------------------------------------------------- */
/* batched band solve: densify + partial-pivot GE (band_diagonal.jl:87-189) */
/* ---------------------------------------------------------------------- */
static void band_solve_column(double *tvector, const double *bmatrix,
                              const double *rvector, int64_t jt, int64_t n, int kl,
                              int ku, int64_t c, int64_t ld, double *Am, double *x) {
    const int64_t nband = kl + ku + 1;
#define BM(cc, bb, rr) bmatrix[(size_t)(cc) + ((size_t)(bb) - 1u) * (size_t)ld + \
                               ((size_t)(rr) - 1u) * (size_t)ld * (size_t)nband]
#define AD(i, j) Am[((size_t)(i) - 1u) + ((size_t)(j) - 1u) * (size_t)n]
    for (int64_t j = 1; j <= n; ++j)
        for (int64_t i = 1; i <= n; ++i) AD(i, j) = 0.0;
    for (int64_t j = 1; j <= n; ++j) {
        int64_t ilo = j - kl > 1 ? j - kl : 1;
        int64_t ihi = j + ku < n ? j + ku : n;
        for (int64_t i = ilo; i <= ihi; ++i) AD(i, j) = BM(c, (i - j) + kl + 1, jt + i - 1);
    }
    for (int64_t i = 1; i <= n; ++i) x[i - 1] = A(rvector, c, jt + i - 1);
    for (int64_t k = 1; k <= n - 1; ++k) {
        int64_t piv = k;
        double maxv = fabs(AD(k, k));
        for (int64_t i = k + 1; i <= n; ++i) {
            double v = fabs(AD(i, k));
            if (v > maxv) { maxv = v; piv = i; }
        }
        double akk = AD(piv, k);
        if (piv != k) {
            for (int64_t j = k; j <= n; ++j) {
                double akj = AD(k, j), apj = AD(piv, j);
                AD(k, j) = apj;
                AD(piv, j) = akj;
            }
            double xk = x[k - 1], xp = x[piv - 1];
            x[k - 1] = xp;
            x[piv - 1] = xk;
        }
        for (int64_t i = k + 1; i <= n; ++i) {
            double f = AD(i, k) / akk;
            for (int64_t j = k; j <= n; ++j) AD(i, j) = AD(i, j) - f * AD(k, j);
            x[i - 1] = x[i - 1] - f * x[k - 1];
        }
    }
    for (int64_t i = n; i >= 1; --i) {
        double s = x[i - 1];
        for (int64_t j = i + 1; j <= n; ++j) s = s - AD(i, j) * x[j - 1];
        x[i - 1] = s / AD(i, i);
    }
    for (int64_t i = 1; i <= n; ++i) A(tvector, c, jt + i - 1) = x[i - 1];
#undef BM
#undef AD
}

int orc_batched_band_solve(double *tvector, const double *bmatrix,
                           const double *rvector, const int64_t *jtop,
                           const int64_t *jbot, const uint8_t *mask, int32_t kl,
                           int32_t ku, int32_t nlevsno_off, int64_t nc,
                           int64_t nlev, int64_t ld) {
    if (!tvector || !bmatrix || !rvector || !jtop || !jbot) return CLMSOIL_ERR_NULL;
    if (nc < 0 || ld < nc || kl < 0 || ku < 0) return CLMSOIL_ERR_SIZE;
#pragma omp parallel
    {
        double *Am = (double *)malloc(sizeof(double) * (size_t)(nlev * nlev));
        double *x = (double *)malloc(sizeof(double) * (size_t)nlev);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            if (mask && !mask[c]) continue;
            int64_t jt = jtop[c] + nlevsno_off + 1;
            int64_t n = jbot[c] - jtop[c] + 1;
            if (n > 0 && jt >= 1 && jt + n - 1 <= nlev)
                band_solve_column(tvector, bmatrix, rvector, jt, n, kl, ku, c, ld, Am, x);
        }
        free(Am);
        free(x);
    }
    return CLMSOIL_OK;
}

/* ---------------------------------------------------------------------- */
/* Moisture-form Richards (soil_water_movement.jl:1250-1489)              */
/* ---------------------------------------------------------------------- */
static int validate_mf(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                       const clmsoil_swm_config *cfg) {
    if (!a || !lev || !cfg) return CLMSOIL_ERR_NULL;
    if (cfg->soilwater_movement_method != CLMSOIL_MOISTURE_FORM) return CLMSOIL_ERR_METHOD;
    /* :1500-1505 */
    if (cfg->upper_boundary_condition != CLMSOIL_BC_FLUX) return CLMSOIL_ERR_BC;
    if (cfg->lower_boundary_condition != CLMSOIL_BC_WATERTABLE &&
        cfg->lower_boundary_condition != CLMSOIL_BC_FLUX &&
        cfg->lower_boundary_condition != CLMSOIL_BC_ZERO_FLUX)
        return CLMSOIL_ERR_BC;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    /* a non-positive or non-finite dtime/dtmin/verySmall never terminates the
       substep loop (:1460-1483); reject it on the host */
    if (!(cfg->dtmin > 0.0) || !(cfg->verySmall > 0.0) || !isfinite(cfg->dtmin))
        return CLMSOIL_ERR_ARG;
    if (!a->nbedrock || !a->h2osoi_liq || !a->smp_l || !a->hk_l || !a->qcharge ||
        !a->nsubsteps || !a->z || !a->dz || !a->qflx_rootsoi || !a->watsat ||
        !a->hksat || !a->bsw || !a->sucsat || !a->icefrac || !a->qflx_infl)
        return CLMSOIL_ERR_NULL;
    if (cfg->lower_boundary_condition == CLMSOIL_BC_WATERTABLE && (!a->zwt || !a->zi))
        return CLMSOIL_ERR_NULL;
    return CLMSOIL_OK;
}

typedef struct {
    double *hk, *smp, *dhkdw, *dsmpdw, *imped, *s2, *vwc_liq, *dt_dz, *qin, *qout,
        *dqidw0, *dqidw1, *dqodw1, *dqodw2, *dwat, *amx, *bmx, *cmx, *rmx, *gam,
        *fluxNet0, *fluxNet1;
} mf_scr;   /* MfScr, soil_water_movement.jl:1241-1247 (one column's rows) */

/* Test hook: every substep iteration of column g_mf_trace_col appends
 * (dtsub, errorMax, accepted) to g_mf_trace_buf (at most g_mf_trace_cap rows). */
static int64_t g_mf_trace_col = -1;
static double *g_mf_trace_buf = NULL;
static int g_mf_trace_cap = 0, g_mf_trace_n = 0;
void orc_set_mf_trace(int64_t col, double *buf, int cap) {
    g_mf_trace_col = col;
    g_mf_trace_buf = buf;
    g_mf_trace_cap = cap;
    g_mf_trace_n = 0;
}
int orc_mf_trace_count(void) { return g_mf_trace_n; }

static void mf_column(const clmsoil_mf_args *a, const clmsoil_levels *lev, int64_t c,
                      double dtime, const clmsoil_swm_config *cfg, mf_scr *S,
                      double *dbg_qin, double *dbg_qout) {
    const int64_t ld = a->ld;
    const int joff = lev->nlevsno, joff_zi = lev->nlevsno + 1, nlevsoi = lev->nlevsoi;
    const int lower_bc = cfg->lower_boundary_condition;
    const int expensive = (cfg->flux_calculation == CLMSOIL_FLUX_EXPENSIVE);
    const double e_ice = cfg->e_ice, dtmin = cfg->dtmin, xtu = cfg->xTolerUpper,
                 xtl = cfg->xTolerLower, verysmall = cfg->verySmall;
    const double denh2o = DENH2O, m_to_mm = M_TO_MM;
    double *liq = a->h2osoi_liq;
    const int nlayers = (int)a->nbedrock[c];               /* :1256 */
#define SC(X, j) S->X[(j) - 1]
    int nsubstep = 0, nrej = 0;                            /* :1258-1261 */
    double dtsub = dtime;
    if (a->dtsub0) {   /* learned first substep, clamped to [min(dtmin,dtime), dtime] */
        double d0 = a->dtsub0[c], lo = JMIN(dtmin, dtime);
        dtsub = isnan(d0) ? dtime : JMIN(JMAX(d0, lo), dtime);
    }
    double dtdone = 0.0, qc = 0.0;
    for (;;) {                                             /* :1263 */
        nsubstep += 1;
        for (int j = 1; j <= nlayers; ++j) {               /* :1273-1279 */
            SC(vwc_liq, j) = JMAX(A(liq, c, joff + j), 1.0e-6) / (A(a->dz, c, joff + j) * denh2o);
            SC(dt_dz, j) = dtsub / (m_to_mm * A(a->dz, c, joff + j));
        }
        for (int j = 1; j <= nlayers; ++j) {               /* :1282-1286 */
            double s = SC(vwc_liq, j) / A(a->watsat, c, j);
            s = JMIN(s, 1.0);
            s = JMAX(0.01, s);
            SC(s2, j) = s;
        }
        for (int j = 1; j <= nlayers; ++j) {               /* :1287-1309 */
            double s1, imp;
            if (j == nlayers) {
                s1 = SC(s2, j);
                imp = clm_pow(10.0, -e_ice * A(a->icefrac, c, j));
            } else {
                s1 = 0.5 * (SC(s2, j) + SC(s2, j + 1));
                imp = clm_pow(10.0, -e_ice * (0.5 * (A(a->icefrac, c, j) + A(a->icefrac, c, j + 1))));
            }
            SC(imped, j) = imp;
            s1 = JMIN(s1, 1.0);
            s1 = JMAX(0.01, s1);
            double b = A(a->bsw, c, j);
            double hkv = imp * A(a->hksat, c, j) * clm_pow(s1, 2.0 * b + 3.0);
            double dhkds = (2.0 * b + 3.0) * hkv / s1;
            double sn = SC(s2, j);
            double smpv = -A(a->sucsat, c, j) * clm_pow(sn, -b);
            double dsmpds = -b * smpv / sn;
            SC(hk, j) = hkv;
            SC(smp, j) = smpv;
            SC(dhkdw, j) = dhkds;
            SC(dsmpdw, j) = dsmpds / A(a->watsat, c, j);
            A(a->smp_l, c, j) = smpv;
            A(a->hk_l, c, j) = hkv;
        }
        for (int j = 1; j <= nlayers; ++j) {               /* :1313-1317 */
            SC(qin, j) = 0.0; SC(qout, j) = 0.0;
            SC(dqidw0, j) = 0.0; SC(dqidw1, j) = 0.0;
            SC(dqodw1, j) = 0.0; SC(dqodw2, j) = 0.0;
        }
        for (int j = 1; j <= nlayers - 1; ++j) {           /* :1334-1342 */
            double dhkds1 = 0.5 * SC(dhkdw, j) / A(a->watsat, c, j);
            double dhkds2 = 0.5 * SC(dhkdw, j) / A(a->watsat, c, j + 1);
            double num = SC(smp, j + 1) - SC(smp, j);
            double den = m_to_mm * (A(a->z, c, joff + j + 1) - A(a->z, c, joff + j));
            SC(qout, j) = -SC(hk, j) * num / den + SC(hk, j);
            SC(dqodw1, j) = (SC(hk, j) * SC(dsmpdw, j) - dhkds1 * num) / den + dhkds1;
            SC(dqodw2, j) = (-SC(hk, j) * SC(dsmpdw, j + 1) - dhkds2 * num) / den + dhkds2;
        }
        {                                                  /* :1344-1367 */
            const int j = nlayers;
            if (lower_bc == CLMSOIL_BC_WATERTABLE) {
                int jwt = nlevsoi;
                for (int jj = 1; jj <= nlevsoi; ++jj)
                    if (a->zwt[c] <= A(a->zi, c, joff_zi + jj)) { jwt = jj - 1; break; }
                if (j > jwt) {
                    SC(qout, j) = 0.0; SC(dqodw1, j) = 0.0; SC(dqodw2, j) = 0.0;
                } else {
                    double dhkds1l = SC(dhkdw, j) / A(a->watsat, c, j);
                    double num = -SC(smp, j);
                    double den = m_to_mm * (a->zwt[c] - A(a->z, c, joff + j));
                    SC(qout, j) = -SC(hk, j) * num / den + SC(hk, j);
                    SC(dqodw1, j) = (SC(hk, j) * SC(dsmpdw, j) - dhkds1l * num) / den + dhkds1l;
                    SC(dqodw2, j) = 0.0;
                }
            } else if (lower_bc == CLMSOIL_BC_FLUX) {
                SC(qout, j) = SC(hk, j);
                SC(dqodw1, j) = SC(dhkdw, j) / A(a->watsat, c, j);
            } else {
                SC(qout, j) = 0.0;
                SC(dqodw1, j) = 0.0;
            }
        }
        for (int j = 2; j <= nlayers; ++j) {               /* :1371-1375 */
            SC(qin, j) = SC(qout, j - 1);
            SC(dqidw0, j) = SC(dqodw1, j - 1);
            SC(dqidw1, j) = SC(dqodw2, j - 1);
        }
        SC(qin, 1) = a->qflx_infl[c];                      /* :1377-1378 */
        SC(dqidw1, 1) = 0.0;
        for (int j = 1; j <= nlayers; ++j) {               /* :1381-1384 */
            double fluxNet = SC(qin, j) - SC(qout, j) - A(a->qflx_rootsoi, c, j);
            SC(rmx, j) = -fluxNet * SC(dt_dz, j);
        }
        SC(amx, 1) = 0.0;                                  /* :1387-1397 */
        SC(bmx, 1) = -1.0 - (-SC(dqidw1, 1) + SC(dqodw1, 1)) * SC(dt_dz, 1);
        SC(cmx, 1) = -SC(dqodw2, 1) * SC(dt_dz, 1);
        for (int j = 2; j <= nlayers - 1; ++j) {
            SC(amx, j) = SC(dqidw0, j) * SC(dt_dz, j);
            SC(bmx, j) = -1.0 - (-SC(dqidw1, j) + SC(dqodw1, j)) * SC(dt_dz, j);
            SC(cmx, j) = -SC(dqodw2, j) * SC(dt_dz, j);
        }
        SC(amx, nlayers) = SC(dqidw0, nlayers) * SC(dt_dz, nlayers);
        SC(bmx, nlayers) = -1.0 - (-SC(dqidw1, nlayers) + SC(dqodw1, nlayers)) * SC(dt_dz, nlayers);
        SC(cmx, nlayers) = 0.0;
        {                                                  /* :1421-1438 */
            double bet = SC(bmx, 1);
            if (fabs(bet) < 1.0e-30) bet = 1.0e-30;
            double dwat_prev = SC(rmx, 1) / bet;
            SC(dwat, 1) = dwat_prev;
            for (int j = 2; j <= nlayers; ++j) {
                SC(gam, j) = SC(cmx, j - 1) / bet;
                bet = SC(bmx, j) - SC(amx, j) * SC(gam, j);
                if (fabs(bet) < 1.0e-30) bet = copysign(1.0e-30, bet == 0.0 ? 1.0 : bet);
                dwat_prev = (SC(rmx, j) - SC(amx, j) * dwat_prev) / bet;
                SC(dwat, j) = dwat_prev;
            }
            double dwat_next = dwat_prev;
            for (int j = nlayers - 1; j >= 1; --j) {
                dwat_next = SC(dwat, j) - SC(gam, j + 1) * dwat_next;
                SC(dwat, j) = dwat_next;
            }
        }
        double errorMax = 0.0;                             /* :1441-1457 */
        for (int j = 1; j <= nlayers; ++j) {
            if (expensive) {
                double qin_test = j == 1 ? SC(qin, j) + SC(dqidw1, j) * SC(dwat, j)
                                         : SC(qin, j) + SC(dqidw0, j) * SC(dwat, j - 1) +
                                               SC(dqidw1, j) * SC(dwat, j);
                double qout_test = j == nlayers
                                       ? SC(qout, j) + SC(dqodw1, j) * SC(dwat, j)
                                       : SC(qout, j) + SC(dqodw1, j) * SC(dwat, j) +
                                             SC(dqodw2, j) * SC(dwat, j + 1);
                SC(fluxNet0, j) = qin_test - qout_test - A(a->qflx_rootsoi, c, j);
            } else {
                SC(fluxNet0, j) = SC(dwat, j) / SC(dt_dz, j);
            }
            SC(fluxNet1, j) = SC(qin, j) - SC(qout, j) - A(a->qflx_rootsoi, c, j);
        }
        for (int j = 1; j <= nlayers; ++j) {
            double ev = fabs(SC(fluxNet1, j) - SC(fluxNet0, j)) * dtsub * 0.5;
            errorMax = JMAX(errorMax, ev);
        }
        if (dbg_qin && dbg_qout) {
            for (int j = 1; j <= nlayers; ++j) {
                A(dbg_qin, c, j) = SC(qin, j);
                A(dbg_qout, c, j) = SC(qout, j);
            }
        }
        if (c == g_mf_trace_col && g_mf_trace_n < g_mf_trace_cap) {
            double *r = g_mf_trace_buf + 3 * g_mf_trace_n++;
            r[0] = dtsub;
            r[1] = errorMax;
            r[2] = (errorMax > xtu && dtsub > dtmin) ? 0.0 : 1.0;
        }
        if (errorMax > xtu && dtsub > dtmin) {             /* :1460-1463 */
            dtsub = JMAX(dtsub / 2.0, dtmin);
            nrej += 1;
            continue;
        }
        for (int j = 1; j <= nlayers; ++j)                 /* :1466-1468 */
            A(liq, c, joff + j) = A(liq, c, joff + j) + SC(dwat, j) * (m_to_mm * A(a->dz, c, joff + j));
        double qcTemp;                                     /* :1471-1478 */
        if (lower_bc == CLMSOIL_BC_FLUX)
            qcTemp = SC(hk, nlayers) + SC(dhkdw, nlayers) * SC(dwat, nlayers);
        else if (lower_bc == CLMSOIL_BC_ZERO_FLUX)
            qcTemp = 0.0;
        else
            qcTemp = SC(qout, nlayers) + SC(dqodw1, nlayers) * SC(dwat, nlayers);
        qc += qcTemp * (dtsub / dtime);
        dtdone += dtsub;                                   /* :1480-1483 */
        if (fabs(dtime - dtdone) < verysmall) break;
        if (errorMax < xtl) dtsub = dtsub * 2.0;
        dtsub = JMIN(dtsub, dtime - dtdone);
    }
    a->qcharge[c] = qc;                                    /* :1486-1487 */
    a->nsubsteps[c] = (double)nsubstep;
    if (a->nrejected) a->nrejected[c] = (double)nrej;
#undef SC
}

int orc_soilwm_moisture_form_solve(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                   double dtime, const clmsoil_swm_config *cfg,
                                   double *dbg_qin, double *dbg_qout) {
    int st = validate_mf(a, lev, cfg);
    if (st) return st;
    if (!(dtime > 0.0) || !isfinite(dtime)) return CLMSOIL_ERR_ARG;
    const int nlevsoi = lev->nlevsoi;
    for (int64_t c = 0; c < a->nc; ++c)
        if ((!a->mask || a->mask[c]) && (a->nbedrock[c] < 1 || a->nbedrock[c] > nlevsoi))
            return CLMSOIL_ERR_SIZE;
#pragma omp parallel
    {
        mf_scr S;
        double *buf = (double *)calloc((size_t)22 * (size_t)nlevsoi, sizeof(double));
        double **f = (double **)&S;
        for (int k = 0; k < 22; ++k) f[k] = buf + (size_t)k * (size_t)nlevsoi;
#pragma omp for schedule(dynamic, 16)
        for (int64_t c = 0; c < a->nc; ++c) {
            if (a->mask && !a->mask[c]) continue;
            mf_column(a, lev, c, dtime, cfg, &S, dbg_qin, dbg_qout);
        }
        free(buf);
    }
    return CLMSOIL_OK;
}

/* compute_qcharge! with the zero smp/imped rows soilwater_moisture_form! passes
 * (soil_water_movement.jl:1153-1217 called from :1575-1598). */
static void qcharge_column(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                           int64_t c, double dtime) {
    const int64_t ld = a->ld;
    const int joff = lev->nlevsno, joff_zi = lev->nlevsno + 1, nlevsoi = lev->nlevsoi;
    const int nlayers = (int)a->nbedrock[c];
    int jwt = nlevsoi;
    for (int j = 1; j <= nlevsoi; ++j)
        if (a->zwt[c] <= A(a->zi, c, joff_zi + j)) { jwt = j - 1; break; }
    if (jwt < nlevsoi) {
        double wh_zwt = -A(a->sucsat, c, jwt + 1) - a->zwt[c] * M_TO_MM;
        /* vwc_2d is filled for j <= min(nlayers, nlevsoi) only (:1586-1592) */
        double vwc = 0.0;
        if (jwt + 1 <= nlayers)
            vwc = JMAX(A(a->h2osoi_liq, c, joff + jwt + 1), 1.0e-6) /
                  (A(a->dz, c, joff + jwt + 1) * DENH2O);
        double s1 = JMAX(vwc / A(a->watsat, c, jwt + 1), 0.01);
        s1 = JMIN(1.0, s1);
        double ka, dummy;
        orc_soil_hk(A(a->hksat, c, jwt + 1), A(a->bsw, c, jwt + 1), s1, 0.0, &ka, &dummy);
        double smpmin = a->smpmin ? a->smpmin[c] : -1.0e8;
        double smp1 = JMAX(smpmin, 0.0);
        int jm = jwt > 1 ? jwt : 1;
        double wh = smp1 - A(a->z, c, joff + jm) * M_TO_MM;
        double q;
        if (jwt == 0)
            q = -ka * (wh_zwt - wh) / ((a->zwt[c] + 1.0e-3) * M_TO_MM);
        else
            q = -ka * (wh_zwt - wh) / ((a->zwt[c] - A(a->z, c, joff + jwt)) * M_TO_MM * 2.0);
        q = JMAX(-10.0 / dtime, q);
        q = JMIN(10.0 / dtime, q);
        a->qcharge[c] = q;
    }
}

int orc_soilwater_moisture_form(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                double dtime, const clmsoil_swm_config *cfg) {
    int st = orc_soilwm_moisture_form_solve(a, lev, dtime, cfg, NULL, NULL);
    if (st) return st;
    int bc = cfg->lower_boundary_condition;
    if (bc == CLMSOIL_BC_AQUIFER || bc == CLMSOIL_BC_WATERTABLE) {   /* :700-703 */
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < a->nc; ++c) {
            if (a->mask && !a->mask[c]) continue;
            qcharge_column(a, lev, c, dtime);
        }
    }
    return CLMSOIL_OK;
}

/* ---------------------------------------------------------------------- */
/* Heat step (soil_temperature.jl), non-urban soil columns                 */
/* ---------------------------------------------------------------------- */
static int validate_heat(const clmsoil_heat_args *a, const clmsoil_levels *lev) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!a->snl || !a->nbedrock || !a->z || !a->dz || !a->zi || !a->t_soisno ||
        !a->h2osoi_liq || !a->h2osoi_ice || !a->excess_ice || !a->watsat ||
        !a->tkmg || !a->tkdry || !a->csol || !a->bsw || !a->sucsat ||
        !a->frac_sno_eff || !a->frac_h2osfc || !a->hs_top_snow || !a->hs_soil ||
        !a->hs_top || !a->hs_h2osfc || !a->dhsdT || !a->t_h2osfc || !a->h2osfc ||
        !a->h2osno_no_layers || !a->snow_depth || !a->t_grnd || !a->xmf ||
        !a->xmf_h2osfc || !a->qflx_snomelt || !a->qflx_snofrz ||
        !a->qflx_snow_drain || !a->qflx_h2osfc_to_ice || !a->eflx_fgr12 ||
        !a->qflx_snomelt_lyr || !a->qflx_snofrz_lyr)
        return CLMSOIL_ERR_NULL;
    return CLMSOIL_OK;
}

typedef struct {
    double *thk, *bw, *tk, *cv, *fact, *fn, *fn1, *rvector, *tvector, *bmatrix, *Am, *x;
    double *hm, *xm, *xm2, *wice0, *wliq0, *wexice0, *wmass0, *tinc, *supercool;
    int64_t *imelt;
} heat_scr;

/* Locals every stage of the heat column reads (nlevmaxurbgrnd = max(nlevurb,
 * nlevgrnd) with nlevurb <= nlevgrnd). */
#define HEAT_LOCALS                                                                   \
    const int64_t ld = a->ld;                                                         \
    const int nlevsno = lev->nlevsno, nlevgrnd = lev->nlevgrnd;                       \
    const int nlevmaxurbgrnd = nlevgrnd;                                              \
    const int joff = nlevsno;                                                         \
    const int nlev = nlevsno + nlevmaxurbgrnd;                                        \
    const int nlev_total = nlevsno + 1 + nlevmaxurbgrnd;                              \
    const int64_t snl = a->snl[c];                                                    \
    const int nbedrock = (int)a->nbedrock[c];                                         \
    double *t = a->t_soisno, *liq = a->h2osoi_liq, *ice = a->h2osoi_ice;              \
    double *exc = a->excess_ice;                                                      \
    const double *z = a->z, *zi = a->zi, *dz = a->dz;                                 \
    const double frac_sno = a->frac_sno_eff[c];                                       \
    const double frac_h2osfc = a->frac_h2osfc[c];                                     \
    const double dhsdT = a->dhsdT[c];                                                 \
    const double cnfac = CNFAC;                                                       \
    (void)ld; (void)nlevmaxurbgrnd; (void)joff; (void)nlev; (void)nlev_total;        \
    (void)snl; (void)nbedrock; (void)t; (void)liq; (void)ice; (void)exc; (void)z;     \
    (void)zi; (void)dz; (void)frac_sno; (void)frac_h2osfc; (void)dhsdT; (void)cnfac
/* local 1-based level vectors */
#define L(X, jj) S->X[(jj) - 1]

/* Stage 1 of soil_temperature! (soil_temperature.jl:480-550): thermal
 * properties, heat-diffusion factor, RHS and 5-band rows, the banded solve, the
 * unpack and fn1. Returns c_h2osfc (the h2osfc heat capacity the later stages
 * use) and tk_h2osfc. */
static void heat_assemble_solve(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                                int64_t c, double dtime, const clmsoil_heat_config *cfg,
                                heat_scr *S, double *c_h2osfc_out, double *tk_h2osfc_out) {
    HEAT_LOCALS;
    for (int jj = 1; jj <= nlev; ++jj) {
        L(thk, jj) = NAN; L(tk, jj) = 0.0; L(cv, jj) = 0.0;
        L(fact, jj) = NAN; L(fn, jj) = 0.0; L(fn1, jj) = 0.0; L(bw, jj) = NAN;
    }
    /* ---- soil_therm_prop! (:874-942) ---- */
    /* _soil_snow_tk_kernel! (:647-716), non-special soil column, Jordan/Sturm snow */
    for (int jj = 1; jj <= nlevsno + nlevgrnd; ++jj) {
        int j = jj - nlevsno;
        if (j >= 1) {
            double satw = (A(liq, c, jj) / DENH2O + A(ice, c, jj) / DENICE +
                           A(exc, c, j) / DENICE) / (A(dz, c, jj) * A(a->watsat, c, j));
            satw = JMIN(1.0, satw);
            if (satw > 1.0e-7) {
                double dke;
                if (A(t, c, jj) >= TFRZ) dke = JMAX(0.0, clm_log10(satw) + 1.0);
                else dke = satw;
                double fl_denom = A(liq, c, jj) / (DENH2O * A(dz, c, jj)) +
                                  A(ice, c, jj) / (DENICE * A(dz, c, jj)) +
                                  A(exc, c, j) / (DENICE * A(dz, c, jj));
                double fl = (A(liq, c, jj) / (DENH2O * A(dz, c, jj))) / fl_denom;
                double dksat = A(a->tkmg, c, j) * clm_pow(TKWAT, fl * A(a->watsat, c, j)) *
                               clm_pow(TKICE, (1.0 - fl) * A(a->watsat, c, j));
                L(thk, jj) = dke * dksat + (1.0 - dke) * A(a->tkdry, c, j);
            } else {
                L(thk, jj) = A(a->tkdry, c, j);
            }
            if (j > nbedrock) L(thk, jj) = THK_BEDROCK;
        }
        if (snl + 1 < 1 && j >= snl + 1 && j <= 0) {
            double denom = JMAX(frac_sno, 1.0e-6) * JMAX(A(dz, c, jj), 1.0e-6);
            double bwv = (A(ice, c, jj) + A(liq, c, jj)) / denom;
            L(bw, jj) = bwv;
            if (cfg->snow_thermal_cond_method == 2) {
                if (bwv <= 156.0)
                    L(thk, jj) = 0.023 + 0.234 * (bwv / 1000.0);
                else
                    L(thk, jj) = 0.138 - 1.01 * (bwv / 1000.0) +
                                 3.233 * ((bwv / 1000.0) * (bwv / 1000.0));
            } else {
                L(thk, jj) = TKAIR + (7.75e-5 * bwv + 1.105e-6 * (bwv * bwv)) * (TKICE - TKAIR);
            }
        }
    }
    /* _soil_tk_interface_kernel! (:605-629) */
    for (int jj = 1; jj <= nlev; ++jj) {
        int j = jj - nlevsno;
        if (j >= snl + 1 && j <= nlevgrnd - 1) {
            double tkj = L(thk, jj), tkj1 = L(thk, jj + 1), zj = A(z, c, jj),
                   zj1 = A(z, c, jj + 1), zij1 = A(zi, c, jj + 1);
            L(tk, jj) = tkj * tkj1 * (zj1 - zj) / (tkj * (zj1 - zij1) + tkj1 * (zij1 - zj));
        } else if (j == nlevgrnd) {
            L(tk, jj) = 0.0;
        }
    }
    /* _tk_h2osfc_kernel! (:762-778) */
    double tk_h2osfc;
    {
        double zh2osfc = 1.0e-3 * (0.5 * a->h2osfc[c]);
        double z1 = A(z, c, 1 + joff), tk1 = L(thk, 1 + joff);
        double denom = TKWAT * z1 + tk1 * zh2osfc;
        tk_h2osfc = denom > 0.0 ? TKWAT * tk1 * (z1 + zh2osfc) / denom : tk1;
    }
    /* _soil_cv_kernel! (:781-813) */
    for (int j = 1; j <= nlevgrnd; ++j) {
        int jj = j + nlevsno;
        L(cv, jj) = A(a->csol, c, j) * (1.0 - A(a->watsat, c, j)) * A(dz, c, jj) +
                    (A(ice, c, jj) * CPICE + A(liq, c, jj) * CPLIQ) + A(exc, c, j) * CPICE;
        if (j > nbedrock) L(cv, jj) = CSOL_BEDROCK * A(dz, c, jj);
    }
    /* _unresolved_snow_cv_kernel! (:842-851) */
    if (a->h2osno_no_layers[c] > 0.0) L(cv, 1 + joff) += CPICE * a->h2osno_no_layers[c];
    /* _snow_cv_kernel! (:857-872) */
    for (int jj = 1; jj <= nlevsno; ++jj) {
        int j = jj - nlevsno;
        if (snl + 1 < 1 && j >= snl + 1) {
            if (frac_sno > 0.0)
                L(cv, jj) = JMAX(THIN_SFCLAYER, (CPLIQ * A(liq, c, jj) + CPICE * A(ice, c, jj)) / frac_sno);
            else
                L(cv, jj) = THIN_SFCLAYER;
        }
    }
    /* ---- ground heat flux: inputs ---- */
    const double hs_top_snow = a->hs_top_snow[c], hs_soil = a->hs_soil[c],
                 hs_h2osfc = a->hs_h2osfc[c];
    const double eflx_bot = a->eflx_bot ? a->eflx_bot[c] : 0.0;
#define SABG(s) (a->sabg_lyr_col ? A(a->sabg_lyr_col, c, (s)) : 0.0)
    /* ---- _heat_diff_kernel! (:1271-1310) ---- */
    for (int jj = 1; jj <= nlev; ++jj) {
        int j = jj - nlevsno;
        if (j <= nlevgrnd && j >= snl + 1) {
            if (j == snl + 1) {
                L(fact, jj) = dtime / L(cv, jj) * A(dz, c, jj) /
                              (0.5 * (A(z, c, jj) - A(zi, c, jj) +
                                      CAPR * (A(z, c, jj + 1) - A(zi, c, jj))));
                L(fn, jj) = L(tk, jj) * (A(t, c, jj + 1) - A(t, c, jj)) /
                            (A(z, c, jj + 1) - A(z, c, jj));
            } else if (j <= nlevgrnd - 1) {
                L(fact, jj) = dtime / L(cv, jj);
                L(fn, jj) = L(tk, jj) * (A(t, c, jj + 1) - A(t, c, jj)) /
                            (A(z, c, jj + 1) - A(z, c, jj));
            } else if (j == nlevgrnd) {
                L(fact, jj) = dtime / L(cv, jj);
                L(fn, jj) = eflx_bot;
            }
        }
    }
    /* ---- _h2osfc_thermprop_kernel! (:116-130) ---- */
    double c_h2osfc, dz_h2osfc;
    if (a->h2osfc[c] > THIN_SFCLAYER && a->frac_h2osfc[c] > THIN_SFCLAYER) {
        c_h2osfc = JMAX(THIN_SFCLAYER, CPLIQ * a->h2osfc[c] / a->frac_h2osfc[c]);
        dz_h2osfc = JMAX(THIN_SFCLAYER, 1.0e-3 * a->h2osfc[c] / a->frac_h2osfc[c]);
    } else {
        c_h2osfc = THIN_SFCLAYER;
        dz_h2osfc = THIN_SFCLAYER;
    }
    /* ---- set_rhs_vec! (:1365-1542), vectors of length nlev_total ---- */
    for (int k = 1; k <= nlev_total; ++k) { L(rvector, k) = NAN; L(tvector, k) = NAN; }
    for (int s = 1; s <= nlevsno; ++s) {                   /* _rhs_snow_kernel! */
        int j = s - nlevsno, jj = s;
        double hs_top_lev = hs_top_snow;
        if (j == snl + 1) {
            L(rvector, s) = A(t, c, jj) + L(fact, jj) * (hs_top_lev - dhsdT * A(t, c, jj) +
                                                         cnfac * L(fn, jj));
        } else if (j > snl + 1) {
            L(rvector, s) = A(t, c, jj) + cnfac * L(fact, jj) * (L(fn, jj) - L(fn, jj - 1));
            L(rvector, s) += L(fact, jj) * SABG(s);
        }
    }
    double fn_h2osfc;                                      /* _rhs_ssw_kernel! */
    {
        double dzm = 0.5 * dz_h2osfc + A(z, c, 1 + joff);
        double fnh = tk_h2osfc * (A(t, c, 1 + joff) - a->t_h2osfc[c]) / dzm;
        fn_h2osfc = fnh;
        L(rvector, nlevsno + 1) = a->t_h2osfc[c] + (dtime / c_h2osfc) *
                                  (hs_h2osfc - dhsdT * a->t_h2osfc[c] + cnfac * fnh);
    }
    for (int j = 1; j <= nlevgrnd; ++j) {                  /* _rhs_soil_kernel! */
        int jj = j + nlevsno, rv = j + nlevsno + 1;
        if (j == snl + 1) {
            L(rvector, rv) = A(t, c, jj) + L(fact, jj) * (hs_top_snow - dhsdT * A(t, c, jj) +
                                                          cnfac * L(fn, jj));
        } else if (j == 1) {
            L(rvector, rv) = A(t, c, jj) + L(fact, jj) *
                             ((1.0 - frac_sno) * (hs_soil - dhsdT * A(t, c, jj)) +
                              cnfac * (L(fn, jj) - frac_sno * L(fn, jj - 1)));
            L(rvector, rv) += frac_sno * L(fact, jj) * SABG(j + nlevsno);
        } else if (j <= nlevgrnd - 1) {
            L(rvector, rv) = A(t, c, jj) + cnfac * L(fact, jj) * (L(fn, jj) - L(fn, jj - 1));
        } else if (j == nlevgrnd) {
            L(rvector, rv) = A(t, c, jj) - cnfac * L(fact, jj) * L(fn, jj - 1) +
                             L(fact, jj) * L(fn, jj);
        }
    }
    if (frac_h2osfc != 0.0) {                              /* _rhs_h2osfc_corr_kernel! */
        int jj = 1 + nlevsno, rv = 1 + nlevsno + 1;
        L(rvector, rv) -= frac_h2osfc * L(fact, jj) *
                          ((hs_soil - dhsdT * A(t, c, jj)) + cnfac * fn_h2osfc);
    }
    /* ---- set_matrix! (:1549-1748); bmatrix local [5, nlev_total] ---- */
#define B(b, r) S->bmatrix[((b) - 1) + 5 * ((r) - 1)]
    for (int k = 0; k < 5 * nlev_total; ++k) S->bmatrix[k] = 0.0;
    const double omc = 1.0 - CNFAC;
    for (int s = 1; s <= nlevsno; ++s) {                   /* _mat_snow_kernel! */
        int j = s - nlevsno, jj = s, bm = s;
        if (j >= snl + 1) {
            double dzp = A(z, c, jj + 1) - A(z, c, jj);
            if (j == snl + 1) {
                B(3, bm) = 1.0 + omc * L(fact, jj) * L(tk, jj) / dzp - L(fact, jj) * dhsdT;
            } else {
                double dzm = A(z, c, jj) - A(z, c, jj - 1);
                B(4, bm) = -omc * L(fact, jj) * L(tk, jj - 1) / dzm;
                B(3, bm) = 1.0 + omc * L(fact, jj) * (L(tk, jj) / dzp + L(tk, jj - 1) / dzm);
            }
            if (j != 0) B(2, bm) = -omc * L(fact, jj) * L(tk, jj) / dzp;
            else B(1, bm) = -omc * L(fact, jj) * L(tk, jj) / dzp;
        }
    }
    {                                                      /* _mat_ssw_kernel! */
        int ssw = nlevsno + 1;
        double dzm = 0.5 * dz_h2osfc + A(z, c, 1 + joff);
        B(3, ssw) = 1.0 + omc * (dtime / c_h2osfc) * tk_h2osfc / dzm - (dtime / c_h2osfc) * dhsdT;
        B(2, ssw) = -omc * (dtime / c_h2osfc) * tk_h2osfc / dzm;
    }
    for (int j = 1; j <= nlevgrnd; ++j) {                  /* _mat_soil_kernel! */
        int jj = j + nlevsno, bm = j + nlevsno + 1;
        if (j == snl + 1) {
            double dzp = A(z, c, jj + 1) - A(z, c, jj);
            B(3, bm) = 1.0 + omc * L(fact, jj) * L(tk, jj) / dzp - L(fact, jj) * dhsdT;
            B(2, bm) = -omc * L(fact, jj) * L(tk, jj) / dzp;
        } else if (j == 1) {
            double dzm = A(z, c, jj) - A(z, c, jj - 1);
            double dzp = A(z, c, jj + 1) - A(z, c, jj);
            B(2, bm) = -omc * L(fact, jj) * L(tk, jj) / dzp;
            B(3, bm) = 1.0 + omc * L(fact, jj) * (L(tk, jj) / dzp + frac_sno * L(tk, jj - 1) / dzm) -
                       (1.0 - frac_sno) * L(fact, jj) * dhsdT;
            B(5, bm) = -frac_sno * omc * L(fact, jj) * L(tk, jj - 1) / dzm;
        } else if (j <= nlevgrnd - 1) {
            double dzm = A(z, c, jj) - A(z, c, jj - 1);
            double dzp = A(z, c, jj + 1) - A(z, c, jj);
            B(2, bm) = -omc * L(fact, jj) * L(tk, jj) / dzp;
            B(3, bm) = 1.0 + omc * L(fact, jj) * (L(tk, jj) / dzp + L(tk, jj - 1) / dzm);
            B(4, bm) = -omc * L(fact, jj) * L(tk, jj - 1) / dzm;
        } else if (j == nlevgrnd) {
            double dzm = A(z, c, jj) - A(z, c, jj - 1);
            B(3, bm) = 1.0 + omc * L(fact, jj) * L(tk, jj - 1) / dzm;
            B(4, bm) = -omc * L(fact, jj) * L(tk, jj - 1) / dzm;
        }
    }
    if (frac_h2osfc != 0.0) {                              /* _mat_h2osfc_corr_kernel! */
        double dzm = 0.5 * dz_h2osfc + A(z, c, 1 + joff);
        int bm = 1 + nlevsno + 1;
        B(3, bm) += frac_h2osfc * (omc * L(fact, 1 + joff) * tk_h2osfc / dzm +
                                  L(fact, 1 + joff) * dhsdT);
        B(4, bm) = -frac_h2osfc * omc * L(fact, 1 + joff) * tk_h2osfc / dzm;
    }
    /* ---- batched_band_solve! (band_diagonal.jl:87-170), kl = ku = 2 ---- */
    {
        const int kl = 2, ku = 2;
        const int64_t jtop = snl, jbot = nlevgrnd;
        const int64_t jt = jtop + nlevsno + 1, n = jbot - jtop + 1;
#define AD(i, j) S->Am[((i) - 1) + ((j) - 1) * n]
        for (int64_t j = 1; j <= n; ++j)
            for (int64_t i = 1; i <= n; ++i) AD(i, j) = 0.0;
        for (int64_t j = 1; j <= n; ++j) {
            int64_t ilo = j - kl > 1 ? j - kl : 1, ihi = j + ku < n ? j + ku : n;
            for (int64_t i = ilo; i <= ihi; ++i) AD(i, j) = B((int)((i - j) + kl + 1), (int)(jt + i - 1));
        }
        for (int64_t i = 1; i <= n; ++i) S->x[i - 1] = L(rvector, jt + i - 1);
        for (int64_t k = 1; k <= n - 1; ++k) {
            int64_t piv = k;
            double maxv = fabs(AD(k, k));
            for (int64_t i = k + 1; i <= n; ++i) {
                double v = fabs(AD(i, k));
                if (v > maxv) { maxv = v; piv = i; }
            }
            double akk = AD(piv, k);
            if (piv != k) {
                for (int64_t j = k; j <= n; ++j) {
                    double akj = AD(k, j), apj = AD(piv, j);
                    AD(k, j) = apj; AD(piv, j) = akj;
                }
                double xk = S->x[k - 1], xp = S->x[piv - 1];
                S->x[k - 1] = xp; S->x[piv - 1] = xk;
            }
            for (int64_t i = k + 1; i <= n; ++i) {
                double f = AD(i, k) / akk;
                for (int64_t j = k; j <= n; ++j) AD(i, j) = AD(i, j) - f * AD(k, j);
                S->x[i - 1] = S->x[i - 1] - f * S->x[k - 1];
            }
        }
        for (int64_t i = n; i >= 1; --i) {
            double s = S->x[i - 1];
            for (int64_t j = i + 1; j <= n; ++j) s = s - AD(i, j) * S->x[j - 1];
            S->x[i - 1] = s / AD(i, i);
        }
        for (int64_t i = 1; i <= n; ++i) L(tvector, jt + i - 1) = S->x[i - 1];
#undef AD
    }
#undef B
    /* ---- _tvector_extract_kernel! (:149-167) ---- */
    for (int64_t j = snl + 1; j <= 0; ++j) A(t, c, j + joff) = L(tvector, j - 1 + nlevsno + 1);
    for (int j = 1; j <= nlevmaxurbgrnd; ++j) A(t, c, j + joff) = L(tvector, j + nlevsno + 1);
    if (frac_h2osfc == 0.0) a->t_h2osfc[c] = A(t, c, 1 + joff);
    else a->t_h2osfc[c] = L(tvector, nlevsno + 1);
    /* ---- _fn1_kernel! (:170-206) ---- */
    for (int jj = 1; jj <= nlev; ++jj) {
        int j = jj - nlevsno;
        if (j >= snl + 1) {
            if (j <= nlevgrnd - 1)
                L(fn1, jj) = L(tk, jj) * (A(t, c, jj + 1) - A(t, c, jj)) / (A(z, c, jj + 1) - A(z, c, jj));
            else if (j == nlevgrnd)
                L(fn1, jj) = 0.0;
        }
    }
    *c_h2osfc_out = c_h2osfc;
    *tk_h2osfc_out = tk_h2osfc;
}

/* phase_change_h2osfc! (soil_temperature.jl:1757-1892) for one column, after
 * the solve; fact_top = fact of the top snow level (joff), used when snow
 * layers exist. */
static void heat_pc_h2osfc(const clmsoil_heat_args *a, const clmsoil_levels *lev, int64_t c,
                           double dtime, double c_h2osfc, double fact_top) {
    HEAT_LOCALS;
    /* ---- phase_change_h2osfc! (:1757-1826), PHASE_CHANGE_MASS_K clamps exact ---- */
    {
        a->xmf_h2osfc[c] = 0.0;
        a->qflx_h2osfc_to_ice[c] = 0.0;
        double eflx_h2osfc_to_snow = 0.0;
        double *t_h2osfc = &a->t_h2osfc[c], *h2osfc = &a->h2osfc[c];
        double *h2osno_nl = &a->h2osno_no_layers[c], *snow_depth = &a->snow_depth[c];
        double int_snow_dummy = 0.0;
        double *int_snow = a->int_snow ? &a->int_snow[c] : &int_snow_dummy;
        if (frac_h2osfc > 0.0 && *t_h2osfc <= TFRZ) {
            double tinc = TFRZ - *t_h2osfc;
            *t_h2osfc = TFRZ;
            double hm = frac_h2osfc * (dhsdT * tinc - tinc * c_h2osfc / dtime);
            double xm = hm * dtime / HFUS;
            double temp1 = *h2osfc + xm;
            double h2osno_total = *h2osno_nl;
            if (snl < 0)
                for (int64_t j = snl + 1; j <= 0; ++j)
                    h2osno_total += A(ice, c, j + joff) + A(liq, c, j + joff);
            double z_avg = frac_sno * *snow_depth;
            double rho_avg = z_avg > 0.0 ? JMIN(800.0, h2osno_total / z_avg) : 200.0;
            if (temp1 >= 0.0) {
                *int_snow -= xm;
                if (snl == 0) *h2osno_nl -= xm;
                else A(ice, c, joff) -= xm;
                h2osno_total -= xm;
                *h2osfc += xm;
                a->xmf_h2osfc[c] = hm;
                a->qflx_h2osfc_to_ice[c] = -xm / dtime;
                if (frac_sno > 0.0 && snl < 0) *snow_depth = h2osno_total / (rho_avg * frac_sno);
                else *snow_depth = h2osno_total / DENICE;
                if (snl == 0) {
                    A(t, c, joff) = *t_h2osfc;
                    eflx_h2osfc_to_snow = 0.0;
                } else {
                    double c1;
                    if (snl == -1) c1 = frac_sno * (dtime / fact_top - dhsdT * dtime);
                    else c1 = frac_sno / fact_top * dtime;
                    double c2 = frac_h2osfc != 0.0 ? (-CPLIQ * xm - frac_h2osfc * dhsdT * dtime) : 0.0;
                    A(t, c, joff) = (c1 * A(t, c, joff) + c2 * *t_h2osfc) / (c1 + c2);
                    eflx_h2osfc_to_snow = (*t_h2osfc - A(t, c, joff)) * c2 / dtime;
                }
            } else {
                rho_avg = (h2osno_total * rho_avg + *h2osfc * DENICE) / (h2osno_total + *h2osfc);
                *int_snow += *h2osfc;
                if (snl == 0) *h2osno_nl += *h2osfc;
                else A(ice, c, joff) += *h2osfc;
                h2osno_total += *h2osfc;
                a->qflx_h2osfc_to_ice[c] = *h2osfc / dtime;
                *t_h2osfc -= temp1 * HFUS / (dtime * dhsdT - c_h2osfc);
                a->xmf_h2osfc[c] = hm - frac_h2osfc * temp1 * HFUS / dtime;
                if (snl == 0) {
                    A(t, c, joff) = *t_h2osfc;
                } else if (snl == -1) {
                    double c1 = frac_sno * (dtime / fact_top - dhsdT * dtime);
                    double c2 = frac_h2osfc != 0.0 ? frac_h2osfc * (c_h2osfc - dtime * dhsdT) : 0.0;
                    A(t, c, joff) = (c1 * A(t, c, joff) + c2 * *t_h2osfc) / (c1 + c2);
                    *t_h2osfc = A(t, c, joff);
                } else {
                    double c1 = frac_sno / fact_top * dtime;
                    double c2 = frac_h2osfc != 0.0 ? frac_h2osfc * (c_h2osfc - dtime * dhsdT) : 0.0;
                    A(t, c, joff) = (c1 * A(t, c, joff) + c2 * *t_h2osfc) / (c1 + c2);
                    *t_h2osfc = A(t, c, joff);
                }
                *h2osfc = 0.0;
                if (frac_sno > 0.0 && snl < 0) *snow_depth = h2osno_total / (rho_avg * frac_sno);
                else *snow_depth = h2osno_total / DENICE;
            }
        }
        if (a->eflx_h2osfc_to_snow) a->eflx_h2osfc_to_snow[c] = eflx_h2osfc_to_snow;
    }
}

/* phase_change_beta! (soil_temperature.jl:1902-2289) for one column, with the
 * heat-diffusion factor fact in S. */
static void heat_pc_beta(const clmsoil_heat_args *a, const clmsoil_levels *lev, int64_t c,
                         double dtime, heat_scr *S) {
    HEAT_LOCALS;
    /* ---- phase_change_beta! (:1929-2289), hard (Float64) branch ---- */
    {
        const double tfrz = TFRZ, hfus = HFUS, grav = GRAV, thou = 1000.0;
        const double frac_h = frac_h2osfc;
        double xmf = 0.0, qflx_snomelt = 0.0, qflx_snofrz = 0.0, qflx_snow_drain = 0.0;
        double snomelt_accum = a->snomelt_accum ? a->snomelt_accum[c] : 0.0;
        double *h2osno_nl = &a->h2osno_no_layers[c], *snow_depth = &a->snow_depth[c];
        for (int j = -nlevsno + 1; j <= nlevmaxurbgrnd; ++j) {
            int jj = j + joff;
            if (j >= snl + 1) {
                L(imelt, jj) = 0;
                L(hm, jj) = 0.0; L(xm, jj) = 0.0; L(xm2, jj) = 0.0;
                L(wice0, jj) = A(ice, c, jj);
                L(wliq0, jj) = A(liq, c, jj);
                L(wexice0, jj) = j >= 1 ? A(exc, c, j) : 0.0;
                L(wmass0, jj) = A(ice, c, jj) + A(liq, c, jj) + L(wexice0, jj);
                if (j >= 1 && a->exice_subs) A(a->exice_subs, c, j) = 0.0;
            }
            if (j <= 0) {
                A(a->qflx_snomelt_lyr, c, jj) = 0.0;
                A(a->qflx_snofrz_lyr, c, jj) = 0.0;
            }
        }
        for (int j = -nlevsno + 1; j <= 0; ++j) {
            int jj = j + joff;
            if (j >= snl + 1) {
                if (A(ice, c, jj) > 0.0 && A(t, c, jj) > tfrz) {
                    L(imelt, jj) = 1; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                }
                if (A(liq, c, jj) > 0.0 && A(t, c, jj) < tfrz) {
                    L(imelt, jj) = 2; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                }
            }
        }
        for (int j = 1; j <= nlevmaxurbgrnd; ++j) {
            int jj = j + joff;
            S->supercool[j - 1] = 0.0;
            if (j <= nlevgrnd) {
                if (A(ice, c, jj) > 0.0 && A(t, c, jj) > tfrz) {
                    L(imelt, jj) = 1; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                }
                if (A(exc, c, j) > 0.0 && A(t, c, jj) > tfrz) {
                    L(imelt, jj) = 1; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                }
                S->supercool[j - 1] = 0.0;
                if (A(t, c, jj) < tfrz) {
                    double smp = hfus * (tfrz - A(t, c, jj)) / (grav * A(t, c, jj)) * thou;
                    S->supercool[j - 1] = A(a->watsat, c, j) *
                                          clm_pow(smp / A(a->sucsat, c, j), -1.0 / A(a->bsw, c, j));
                    S->supercool[j - 1] *= A(dz, c, jj) * thou;
                }
                if (A(liq, c, jj) > S->supercool[j - 1] && A(t, c, jj) < tfrz) {
                    L(imelt, jj) = 2; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                }
                if (*h2osno_nl > 0.0 && j == 1) {
                    if (A(t, c, jj) > tfrz) {
                        L(imelt, jj) = 1; L(tinc, jj) = tfrz - A(t, c, jj); A(t, c, jj) = tfrz;
                    }
                }
            }
        }
        for (int j = -nlevsno + 1; j <= nlevmaxurbgrnd; ++j) {
            int jj = j + joff;
            if (j <= nlevgrnd && j >= snl + 1) {
                double fct = L(fact, jj);
                if (L(imelt, jj) > 0) {
                    if (j == snl + 1) {
                        if (j > 0) L(hm, jj) = dhsdT * L(tinc, jj) - L(tinc, jj) / fct;
                        else L(hm, jj) = frac_sno * (dhsdT * L(tinc, jj) - L(tinc, jj) / fct);
                        if (j == 1 && frac_h != 0.0) L(hm, jj) -= frac_h * dhsdT * L(tinc, jj);
                    } else if (j == 1) {
                        L(hm, jj) = (1.0 - frac_sno - frac_h) * dhsdT * L(tinc, jj) - L(tinc, jj) / fct;
                    } else {
                        if (j < 1) L(hm, jj) = -frac_sno * L(tinc, jj) / fct;
                        else L(hm, jj) = -L(tinc, jj) / fct;
                    }
                }
                if (L(imelt, jj) == 1 && L(hm, jj) < 0.0) { L(hm, jj) = 0.0; L(imelt, jj) = 0; }
                if (L(imelt, jj) == 2 && L(hm, jj) > 0.0) { L(hm, jj) = 0.0; L(imelt, jj) = 0; }
                if (L(imelt, jj) > 0 && fabs(L(hm, jj)) > 0.0) {
                    L(xm, jj) = L(hm, jj) * dtime / hfus;
                    if (j == 1) {
                        if (*h2osno_nl > 0.0 && L(xm, jj) > 0.0) {
                            double temp1 = *h2osno_nl;
                            *h2osno_nl = JMAX(0.0, temp1 - L(xm, jj));
                            double propor = *h2osno_nl / temp1;
                            *snow_depth = propor * *snow_depth;
                            double heatr = L(hm, jj) - hfus * (temp1 - *h2osno_nl) / dtime;
                            if (heatr > 0.0) {
                                L(xm, jj) = heatr * dtime / hfus;
                                L(hm, jj) = heatr;
                            } else {
                                L(xm, jj) = 0.0;
                                L(hm, jj) = 0.0;
                            }
                            qflx_snomelt = JMAX(0.0, temp1 - *h2osno_nl) / dtime;
                            xmf = hfus * qflx_snomelt;
                            qflx_snow_drain = qflx_snomelt;
                        }
                    }
                    double heatr = 0.0;
                    if (L(xm, jj) > 0.0) {
                        A(ice, c, jj) = JMAX(0.0, L(wice0, jj) - L(xm, jj));
                        heatr = L(hm, jj) - hfus * (L(wice0, jj) - A(ice, c, jj)) / dtime;
                        L(xm2, jj) = L(xm, jj) - A(ice, c, jj);
                        if (A(ice, c, jj) == 0.0) {
                            if (L(wexice0, jj) >= 0.0 && L(xm2, jj) > 0.0 && j >= 2) {
                                A(exc, c, j) = JMAX(0.0, L(wexice0, jj) - L(xm2, jj));
                                heatr = L(hm, jj) - hfus * (L(wexice0, jj) - A(exc, c, j) +
                                                            L(wice0, jj) - A(ice, c, jj)) / dtime;
                            }
                        }
                    } else if (L(xm, jj) < 0.0) {
                        if (j <= 0) {
                            A(ice, c, jj) = JMIN(L(wmass0, jj), L(wice0, jj) - L(xm, jj));
                        } else {
                            if (L(wmass0, jj) - L(wexice0, jj) < S->supercool[j - 1])
                                A(ice, c, jj) = 0.0;
                            else
                                A(ice, c, jj) = JMIN(L(wmass0, jj) - L(wexice0, jj) - S->supercool[j - 1],
                                                     L(wice0, jj) - L(xm, jj));
                        }
                        heatr = L(hm, jj) - hfus * (L(wice0, jj) - A(ice, c, jj)) / dtime;
                    }
                    double ei_val = j >= 1 ? A(exc, c, j) : 0.0;
                    A(liq, c, jj) = JMAX(0.0, L(wmass0, jj) - A(ice, c, jj) - ei_val);
                    if (fabs(heatr) > 0.0) {
                        if (j == snl + 1) {
                            if (j == 1) {
                                A(t, c, jj) += fct * heatr / (1.0 - (1.0 - frac_h) * fct * dhsdT);
                            } else {
                                if (frac_sno > 0.0)
                                    A(t, c, jj) += (fct / frac_sno) * heatr / (1.0 - fct * dhsdT);
                            }
                        } else if (j == 1) {
                            A(t, c, jj) += fct * heatr / (1.0 - (1.0 - frac_sno - frac_h) * fct * dhsdT);
                        } else {
                            if (j > 0) A(t, c, jj) += fct * heatr;
                            else if (frac_sno > 0.0) A(t, c, jj) += (fct / frac_sno) * heatr;
                        }
                        if (j <= 0)
                            if (A(liq, c, jj) * A(ice, c, jj) > 0.0) A(t, c, jj) = tfrz;
                    }
                    if (j >= 1) {
                        xmf += hfus * (L(wice0, jj) - A(ice, c, jj)) / dtime +
                               hfus * (L(wexice0, jj) - A(exc, c, j)) / dtime;
                        if (a->exice_subs)
                            A(a->exice_subs, c, j) = JMAX(0.0, (L(wexice0, jj) - A(exc, c, j)) / DENICE);
                    } else {
                        xmf += hfus * (L(wice0, jj) - A(ice, c, jj)) / dtime;
                    }
                    if (L(imelt, jj) == 1 && j < 1) {
                        double v = JMAX(0.0, L(wice0, jj) - A(ice, c, jj)) / dtime;
                        A(a->qflx_snomelt_lyr, c, jj) = v;
                        qflx_snomelt += v;
                        snomelt_accum += v * dtime * 1.0e-3;
                    }
                    if (L(imelt, jj) == 2 && j < 1) {
                        double v = JMAX(0.0, A(ice, c, jj) - L(wice0, jj)) / dtime;
                        A(a->qflx_snofrz_lyr, c, jj) = v;
                        qflx_snofrz += v;
                    }
                }
            }
        }
        a->xmf[c] = xmf;
        a->qflx_snomelt[c] = qflx_snomelt;
        a->qflx_snofrz[c] = qflx_snofrz;
        a->qflx_snow_drain[c] = qflx_snow_drain;
        if (a->eflx_snomelt) a->eflx_snomelt[c] = qflx_snomelt * hfus;
        if (a->eflx_snomelt_r) a->eflx_snomelt_r[c] = qflx_snomelt * hfus;
        if (a->snomelt_accum) a->snomelt_accum[c] = snomelt_accum;
        if (a->imelt)
            for (int jj = 1; jj <= nlev; ++jj)
                if (jj - nlevsno >= snl + 1) A(a->imelt, c, jj) = L(imelt, jj);
    }
}

/* t_grnd, the soil heat flux diagnostics and the optional copies of the
 * thermal-property scratch (soil_temperature.jl:584-590, 224-269). */
static void heat_finish(const clmsoil_heat_args *a, const clmsoil_levels *lev, int64_t c,
                        double c_h2osfc, heat_scr *S) {
    HEAT_LOCALS;
    /* ---- _tgrnd_kernel! (:224-247) ---- */
    {
        double tg;
        if (snl < 0) {
            if (frac_h2osfc != 0.0)
                tg = frac_sno * A(t, c, snl + 1 + joff) +
                     (1.0 - frac_sno - frac_h2osfc) * A(t, c, 1 + joff) +
                     frac_h2osfc * a->t_h2osfc[c];
            else
                tg = frac_sno * A(t, c, snl + 1 + joff) + (1.0 - frac_sno) * A(t, c, 1 + joff);
        } else {
            if (frac_h2osfc != 0.0)
                tg = (1.0 - frac_h2osfc) * A(t, c, 1 + joff) + frac_h2osfc * a->t_h2osfc[c];
            else
                tg = A(t, c, 1 + joff);
        }
        a->t_grnd[c] = tg;
    }
    /* ---- _eflx_fgr_kernel! (:250-269) ---- */
    a->eflx_fgr12[c] = -cnfac * L(fn, 1 + joff) - (1.0 - cnfac) * L(fn1, 1 + joff);
    if (a->eflx_fgr) {
        for (int j = 1; j <= nlevgrnd; ++j) {
            int jj = j + joff;
            if (j < nlevgrnd) A(a->eflx_fgr, c, j) = -cnfac * L(fn, jj) - (1.0 - cnfac) * L(fn1, jj);
            else A(a->eflx_fgr, c, j) = 0.0;
        }
    }
    /* optional copies of the thermal-property scratch the reference keeps */
    if (a->fact)
        for (int jj = 1; jj <= nlev; ++jj)
            if (jj - nlevsno >= snl + 1) A(a->fact, c, jj) = L(fact, jj);
    if (a->thk)
        for (int jj = 1; jj <= nlev; ++jj) {
            int j = jj - nlevsno;
            if (j >= 1 || j >= snl + 1) A(a->thk, c, jj) = L(thk, jj);
        }
    if (a->bw)
        for (int jj = 1; jj <= nlevsno; ++jj)
            if (jj - nlevsno >= snl + 1) A(a->bw, c, jj) = L(bw, jj);
    if (a->c_h2osfc) a->c_h2osfc[c] = c_h2osfc;
}


/* one column: the stages in the reference's order */
static void heat_column(const clmsoil_heat_args *a, const clmsoil_levels *lev, int64_t c,
                        double dtime, const clmsoil_heat_config *cfg, heat_scr *S) {
    double c_h2osfc, tk_h2osfc;
    heat_assemble_solve(a, lev, c, dtime, cfg, S, &c_h2osfc, &tk_h2osfc);
    heat_pc_h2osfc(a, lev, c, dtime, c_h2osfc, S->fact[lev->nlevsno - 1]);
    heat_pc_beta(a, lev, c, dtime, S);
    heat_finish(a, lev, c, c_h2osfc, S);
}
#undef L
#undef SABG
#undef HEAT_LOCALS

static void heat_scr_alloc(heat_scr *S, int nlev_total) {
    size_t nd = (size_t)nlev_total;
    double *buf = (double *)calloc(nd * (22 + 5) + nd * nd, sizeof(double));
    double **f = (double **)S;
    for (int k = 0; k < 9; ++k) f[k] = buf + (size_t)k * nd;          /* thk..tvector */
    S->bmatrix = buf + 9 * nd;                                          /* 5*nd */
    S->Am = buf + 14 * nd;                                              /* nd*nd */
    double *rest = S->Am + nd * nd;
    S->x = rest; rest += nd;
    S->hm = rest; rest += nd;
    S->xm = rest; rest += nd;
    S->xm2 = rest; rest += nd;
    S->wice0 = rest; rest += nd;
    S->wliq0 = rest; rest += nd;
    S->wexice0 = rest; rest += nd;
    S->wmass0 = rest; rest += nd;
    S->tinc = rest; rest += nd;
    S->supercool = rest; rest += nd;
    S->imelt = (int64_t *)calloc(nd, sizeof(int64_t));
}
static void heat_scr_free(heat_scr *S) {
    free(S->imelt);
    free(S->thk);   /* the buffer starts at thk */
}

/* Test entries: the stages of one heat column on their own, as the reference's
 * tests call soil_therm_prop!, set_matrix!, phase_change_h2osfc! and
 * phase_change_beta! (test/test_soil_temperature.jl). */
int orc_heat_assemble_probe(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                            double dtime, const clmsoil_heat_config *cfg, int64_t c,
                            double *tk, double *cv, double *bmatrix, double *rvector,
                            double *tk_h2osfc) {
    int st = validate_heat(a, lev);
    if (st) return st;
    if (!cfg || c < 0 || c >= a->nc) return CLMSOIL_ERR_ARG;
    const int nlev = lev->nlevsno + lev->nlevgrnd, nlev_total = nlev + 1;
    heat_scr S;
    heat_scr_alloc(&S, nlev_total);
    double c_h2osfc, tkh;
    heat_assemble_solve(a, lev, c, dtime, cfg, &S, &c_h2osfc, &tkh);
    for (int k = 0; k < nlev; ++k) { tk[k] = S.tk[k]; cv[k] = S.cv[k]; }
    for (int k = 0; k < 5 * nlev_total; ++k) bmatrix[k] = S.bmatrix[k];
    for (int k = 0; k < nlev_total; ++k) rvector[k] = S.rvector[k];
    *tk_h2osfc = tkh;
    heat_scr_free(&S);
    return CLMSOIL_OK;
}

int orc_phase_change_h2osfc(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                            double dtime, int64_t c, double c_h2osfc, double fact_top) {
    int st = validate_heat(a, lev);
    if (st) return st;
    if (c < 0 || c >= a->nc) return CLMSOIL_ERR_ARG;
    heat_pc_h2osfc(a, lev, c, dtime, c_h2osfc, fact_top);
    return CLMSOIL_OK;
}

int orc_phase_change_beta(const clmsoil_heat_args *a, const clmsoil_levels *lev, double dtime,
                          int64_t c, const double *fact) {
    int st = validate_heat(a, lev);
    if (st) return st;
    if (c < 0 || c >= a->nc || !fact) return CLMSOIL_ERR_ARG;
    const int nlev = lev->nlevsno + lev->nlevgrnd;
    heat_scr S;
    heat_scr_alloc(&S, nlev + 1);
    for (int k = 0; k < nlev; ++k) S.fact[k] = fact[k];
    heat_pc_beta(a, lev, c, dtime, &S);
    heat_scr_free(&S);
    return CLMSOIL_OK;
}

int orc_soil_temperature(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                         double dtime, const clmsoil_heat_config *cfg) {
    int st = validate_heat(a, lev);
    if (st) return st;
    if (!cfg) return CLMSOIL_ERR_NULL;
    const int nlevsno = lev->nlevsno, nlevgrnd = lev->nlevgrnd;
    const int nlev_total = nlevsno + nlevgrnd + 1;
    for (int64_t c = 0; c < a->nc; ++c) {
        if (a->mask && !a->mask[c]) continue;
        if (a->snl[c] < -nlevsno || a->snl[c] > 0) return CLMSOIL_ERR_SIZE;
    }
#pragma omp parallel
    {
        heat_scr S;
        heat_scr_alloc(&S, nlev_total);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < a->nc; ++c) {
            if (a->mask && !a->mask[c]) continue;
            heat_column(a, lev, c, dtime, cfg, &S);
        }
        heat_scr_free(&S);
    }
    return CLMSOIL_OK;
}

/* ---------------------------------------------------------------------------
 * Zeng-Decker 2009 soil water step, soilwater_zengdecker2009!
 * (soil_water_movement.jl:1616-1764), one column at a time in the order of its
 * kernels; smooth_min/smooth_max/smooth_ifelse on Float64 are Julia's exact
 * min/max/branch (smooth_ad.jl:91-149, 253-262). Scratch index j is the soil
 * layer (1-based, column entry j); zimm[j] is Fortran zimm(c, j), j = 0..nlevsoi.
 * ------------------------------------------------------------------------- */
#define ZMAX 32
/* Test hook of the Zeng-Decker step: with a non-zero seed every pow() of that
 * step is moved one ulp up or down, the direction taken from a hash of its
 * operands and the seed (so the result is deterministic and thread-safe). pow()
 * is the one operation whose last bit differs between the host libm and the HIP
 * device library; the parity tests use this to measure how far such a
 * difference moves the step's results. Zero (the default) leaves pow() exact. */
static uint64_t g_zd_pow_jitter = 0;
void orc_set_zd_pow_jitter(uint64_t seed) { g_zd_pow_jitter = seed; }

static double zd_pow(double x, double y) {
    const double p = pow(x, y);
    if (!g_zd_pow_jitter || !isfinite(p) || p == 0.0) return p;
    uint64_t bx, by;
    memcpy(&bx, &x, 8);
    memcpy(&by, &y, 8);
    uint64_t h = bx ^ (by * 0x9E3779B97F4A7C15ull) ^ g_zd_pow_jitter;
    h ^= h >> 30; h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27; h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return nextafter(p, (h & 1u) ? INFINITY : -INFINITY);
}

static void zd09_column(const clmsoil_zd_args *a, const clmsoil_levels *lev, int64_t c,
                        double dtime, double e_ice) {
    const int64_t ld = a->ld;
    const int n = lev->nlevsoi, ngrnd = lev->nlevgrnd;
    const int joff = lev->nlevsno, joff_zi = lev->nlevsno + 1;
    double zmm[ZMAX], dzmm[ZMAX], zimm[ZMAX], vol_ice[ZMAX], icefrac[ZMAX], vwc_liq[ZMAX];
    double vol_eq[ZMAX], zq[ZMAX], hk[ZMAX], dhkdw[ZMAX], imped[ZMAX], smp[ZMAX], dsmpdw[ZMAX];
    double qin[ZMAX], qout[ZMAX], dqidw0[ZMAX], dqidw1[ZMAX], dqodw1[ZMAX], dqodw2[ZMAX];
    double rmx[ZMAX], amx[ZMAX], bmx[ZMAX], cmx[ZMAX], dwat2[ZMAX], cp[ZMAX], dp[ZMAX];
    for (int j = 0; j < ZMAX; ++j)
        zmm[j] = dzmm[j] = zimm[j] = vol_ice[j] = icefrac[j] = vwc_liq[j] = vol_eq[j] = zq[j] =
            hk[j] = dhkdw[j] = imped[j] = smp[j] = dsmpdw[j] = qin[j] = qout[j] = dqidw0[j] =
                dqidw1[j] = dqodw1[j] = dqodw2[j] = rmx[j] = amx[j] = bmx[j] = cmx[j] = dwat2[j] =
                    0.0;
    /* _soilwm_mm_icefrac_kernel! (:97-121) */
    for (int j = 1; j <= n; ++j) {
        zmm[j] = A(a->z, c, joff + j) * 1.0e3;
        dzmm[j] = A(a->dz, c, joff + j) * 1.0e3;
        zimm[j] = A(a->zi, c, joff_zi + j) * 1.0e3;
        vol_ice[j] = JMIN(A(a->watsat, c, j), A(a->h2osoi_ice, c, joff + j) /
                                                  (A(a->dz, c, joff + j) * DENICE));
        icefrac[j] = JMIN(1.0, vol_ice[j] / A(a->watsat, c, j));
        vwc_liq[j] = JMAX(A(a->h2osoi_liq, c, joff + j), 1.0e-6) /
                     (A(a->dz, c, joff + j) * DENH2O);
        A(a->icefrac, c, j) = icefrac[j];
    }
    /* _soilwm_zwtmm_kernel! (:143-150) */
    zimm[0] = 0.0;
    const double zwtmm = a->zwt[c] * 1.0e3;
    /* _soilwm_jwt_vwczwt_kernel! (:289-324) */
    int jw = n;
    for (int j = 1; j <= n; ++j)
        if (a->zwt[c] <= A(a->zi, c, joff_zi + j)) { jw = j - 1; break; }
    double vwc_zwt = A(a->watsat, c, n);
    if (A(a->t_soisno, c, joff + jw + 1) < TFRZ) {
        double v = vwc_liq[n];
        for (int j = n; j <= ngrnd; ++j) {
            if (a->zwt[c] <= A(a->zi, c, joff_zi + j)) {
                double smp1 = HFUS * (TFRZ - A(a->t_soisno, c, joff + j)) /
                              (GRAV * A(a->t_soisno, c, joff + j)) * 1000.0;
                smp1 = JMAX(A(a->sucsat, c, n), smp1);
                v = A(a->watsat, c, n) * zd_pow(smp1 / A(a->sucsat, c, n), -1.0 / A(a->bsw, c, n));
                v = JMIN(v, 0.5 * (A(a->watsat, c, n) + A(a->h2osoi_vol, c, n)));
                break;
            }
        }
        vwc_zwt = v;
    }
    /* _soilwm_voleq_zq_kernel! (:171-201) */
    const double smpmin = a->smpmin[c];
    for (int j = 1; j <= n; ++j) {
        const double zimm_jm1 = zimm[j - 1], zimm_j = zimm[j];
        const double ws = A(a->watsat, c, j), suc = A(a->sucsat, c, j), b = A(a->bsw, c, j);
        const double zwt = zwtmm;
        const double onem1b = 1.0 - 1.0 / b;
        if (zwt <= zimm_jm1) {
            vol_eq[j] = ws;
        } else if (zwt < zimm_j && zwt > zimm_jm1) {
            const double tempi = 1.0;
            const double temp0 = zd_pow((suc + zwt - zimm_jm1) / suc, onem1b);
            const double voleq1 = -suc * ws / onem1b / (zwt - zimm_jm1) * (tempi - temp0);
            double v = (voleq1 * (zwt - zimm_jm1) + ws * (zimm_j - zwt)) / (zimm_j - zimm_jm1);
            v = JMIN(ws, v);
            vol_eq[j] = JMAX(v, 0.0);
        } else {
            const double tempi = zd_pow((suc + zwt - zimm_j) / suc, onem1b);
            const double temp0 = zd_pow((suc + zwt - zimm_jm1) / suc, onem1b);
            double v = -suc * ws / onem1b / (zimm_j - zimm_jm1) * (tempi - temp0);
            v = JMAX(v, 0.0);
            vol_eq[j] = JMIN(ws, v);
        }
        const double zqv = -suc * zd_pow(JMAX(vol_eq[j] / ws, 0.01), -b);
        zq[j] = JMAX(smpmin, zqv);
    }
    /* _soilwm_voleq_zq_aqu_kernel! (:346-364) */
    if (jw == n) {
        const int j = n;
        const double suc = A(a->sucsat, c, j), ws = A(a->watsat, c, j), b = A(a->bsw, c, j);
        const double zwt = zwtmm, zimm_j = zimm[j];
        const double onem1b = 1.0 - 1.0 / b;
        const double tempi = 1.0;
        const double temp0 = zd_pow((suc + zwt - zimm_j) / suc, onem1b);
        double v = -suc * ws / onem1b / (zwt - zimm_j) * (tempi - temp0);
        v = JMAX(v, 0.0);
        v = JMIN(ws, v);
        vol_eq[j + 1] = v;
        const double zqv = -suc * zd_pow(JMAX(v / ws, 0.01), -b);
        zq[j + 1] = JMAX(smpmin, zqv);
    }
    /* _soilwm_hk_smp_kernel! (:223-262) */
    for (int j = 1; j <= n; ++j) {
        const int jp1 = j + 1 < n ? j + 1 : n;
        const double b = A(a->bsw, c, j);
        double s1 = 0.5 * (vwc_liq[j] + vwc_liq[jp1]) /
                    (0.5 * (A(a->watsat, c, j) + A(a->watsat, c, jp1)));
        s1 = JMIN(1.0, s1);
        const double s2 = A(a->hksat, c, j) * zd_pow(s1, 2.0 * b + 2.0);
        const double imp = zd_pow(10.0, -e_ice * (0.5 * (icefrac[j] + icefrac[jp1])));
        imped[j] = imp;
        hk[j] = imp * s1 * s2;
        dhkdw[j] = imp * (2.0 * b + 3.0) * s2 * (1.0 / (A(a->watsat, c, j) + A(a->watsat, c, jp1)));
        double s_node = JMAX(vwc_liq[j] / A(a->watsat, c, j), 0.01);
        s_node = JMIN(1.0, s_node);
        double smpv = -A(a->sucsat, c, j) * zd_pow(s_node, -b);
        smpv = JMAX(smpmin, smpv);
        smp[j] = smpv;
        dsmpdw[j] = -b * smpv / vwc_liq[j];
        A(a->smp_l, c, j) = smpv;
        A(a->hk_l, c, j) = hk[j];
    }
    /* _soilwm_aqu_zmm_kernel! (:385-396) */
    zmm[n + 1] = 0.5 * (1.0e3 * a->zwt[c] + zmm[n]);
    dzmm[n + 1] = jw < n ? dzmm[n] : 1.0e3 * a->zwt[c] - zmm[n];
    /* _soilwm_tridiag_assemble_kernel! (:430-539), sdamp = 0 (:1700) */
    const double sdamp = 0.0;
    {
        int j = 1;
        qin[j] = a->qflx_infl[c];
        double den = zmm[j + 1] - zmm[j];
        double dzq = zq[j + 1] - zq[j];
        double num = (smp[j + 1] - smp[j]) - dzq;
        qout[j] = -hk[j] * num / den;
        dqodw1[j] = -(-hk[j] * dsmpdw[j] + num * dhkdw[j]) / den;
        dqodw2[j] = -(hk[j] * dsmpdw[j + 1] + num * dhkdw[j]) / den;
        rmx[j] = qin[j] - qout[j] - A(a->qflx_rootsoi, c, j);
        amx[j] = 0.0;
        bmx[j] = dzmm[j] * (sdamp + 1.0 / dtime) + dqodw1[j];
        cmx[j] = dqodw2[j];
        for (j = 2; j <= n - 1; ++j) {
            den = zmm[j] - zmm[j - 1];
            dzq = zq[j] - zq[j - 1];
            num = (smp[j] - smp[j - 1]) - dzq;
            qin[j] = -hk[j - 1] * num / den;
            dqidw0[j] = -(-hk[j - 1] * dsmpdw[j - 1] + num * dhkdw[j - 1]) / den;
            dqidw1[j] = -(hk[j - 1] * dsmpdw[j] + num * dhkdw[j - 1]) / den;
            den = zmm[j + 1] - zmm[j];
            dzq = zq[j + 1] - zq[j];
            num = (smp[j + 1] - smp[j]) - dzq;
            qout[j] = -hk[j] * num / den;
            dqodw1[j] = -(-hk[j] * dsmpdw[j] + num * dhkdw[j]) / den;
            dqodw2[j] = -(hk[j] * dsmpdw[j + 1] + num * dhkdw[j]) / den;
            rmx[j] = qin[j] - qout[j] - A(a->qflx_rootsoi, c, j);
            amx[j] = -dqidw0[j];
            bmx[j] = dzmm[j] / dtime - dqidw1[j] + dqodw1[j];
            cmx[j] = dqodw2[j];
        }
        j = n;
        if (j > jw) {
            den = zmm[j] - zmm[j - 1];
            dzq = zq[j] - zq[j - 1];
            num = (smp[j] - smp[j - 1]) - dzq;
            qin[j] = -hk[j - 1] * num / den;
            dqidw0[j] = -(-hk[j - 1] * dsmpdw[j - 1] + num * dhkdw[j - 1]) / den;
            dqidw1[j] = -(hk[j - 1] * dsmpdw[j] + num * dhkdw[j - 1]) / den;
            qout[j] = 0.0;
            dqodw1[j] = 0.0;
            rmx[j] = qin[j] - qout[j] - A(a->qflx_rootsoi, c, j);
            amx[j] = -dqidw0[j];
            bmx[j] = dzmm[j] / dtime - dqidw1[j] + dqodw1[j];
            cmx[j] = 0.0;
            rmx[j + 1] = 0.0;
            amx[j + 1] = 0.0;
            bmx[j + 1] = dzmm[j + 1] / dtime;
            cmx[j + 1] = 0.0;
        } else {
            double s_node = JMAX(0.5 * ((vwc_zwt + vwc_liq[j]) / A(a->watsat, c, j)), 0.01);
            s_node = JMIN(1.0, s_node);
            double smp1 = -A(a->sucsat, c, j) * zd_pow(s_node, -A(a->bsw, c, j));
            smp1 = JMAX(smpmin, smp1);
            const double dsmpdw1 = -A(a->bsw, c, j) * smp1 / (s_node * A(a->watsat, c, j));
            den = zmm[j] - zmm[j - 1];
            dzq = zq[j] - zq[j - 1];
            num = (smp[j] - smp[j - 1]) - dzq;
            qin[j] = -hk[j - 1] * num / den;
            dqidw0[j] = -(-hk[j - 1] * dsmpdw[j - 1] + num * dhkdw[j - 1]) / den;
            dqidw1[j] = -(hk[j - 1] * dsmpdw[j] + num * dhkdw[j - 1]) / den;
            den = zmm[j + 1] - zmm[j];
            dzq = zq[j + 1] - zq[j];
            num = (smp1 - smp[j]) - dzq;
            qout[j] = -hk[j] * num / den;
            dqodw1[j] = -(-hk[j] * dsmpdw[j] + num * dhkdw[j]) / den;
            dqodw2[j] = -(hk[j] * dsmpdw1 + num * dhkdw[j]) / den;
            rmx[j] = qin[j] - qout[j] - A(a->qflx_rootsoi, c, j);
            amx[j] = -dqidw0[j];
            bmx[j] = dzmm[j] / dtime - dqidw1[j] + dqodw1[j];
            cmx[j] = dqodw2[j];
            qin[j + 1] = qout[j];
            dqidw0[j + 1] = -(-hk[j] * dsmpdw[j] + num * dhkdw[j]) / den;
            dqidw1[j + 1] = -(hk[j] * dsmpdw1 + num * dhkdw[j]) / den;
            qout[j + 1] = 0.0;
            dqodw1[j + 1] = 0.0;
            rmx[j + 1] = qin[j + 1] - qout[j + 1];
            amx[j + 1] = -dqidw0[j + 1];
            bmx[j + 1] = dzmm[j + 1] / dtime - dqidw1[j + 1] + dqodw1[j + 1];
            cmx[j + 1] = 0.0;
        }
    }
    /* tridiagonal_multi!(dwat2, amx, bmx, cmx, rmx, jtop = 1, ..., nlevsoi + 1)
     * (tridiagonal.jl:86-127) */
    {
        const int nl = n + 1;
        const double tiny = 1e-30;
        double bet = bmx[1];
        if (fabs(bet) < tiny) bet = tiny;
        double cp_prev = cmx[1] / bet, dp_prev = rmx[1] / bet;
        cp[1] = cp_prev;
        dp[1] = dp_prev;
        for (int jj = 2; jj <= nl; ++jj) {
            double denom = bmx[jj] - amx[jj] * cp_prev;
            if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0.0 ? 1.0 : denom);
            cp_prev = cmx[jj] / denom;
            dp_prev = (rmx[jj] - amx[jj] * dp_prev) / denom;
            cp[jj] = cp_prev;
            dp[jj] = dp_prev;
        }
        double u_next = dp_prev;
        dwat2[nl] = u_next;
        for (int jj = nl - 1; jj >= 1; --jj) {
            u_next = dp[jj] - cp[jj] * u_next;
            dwat2[jj] = u_next;
        }
    }
    /* _soilwm_renew_qcharge_kernel! (:561-604) */
    for (int j = 1; j <= n; ++j)
        A(a->h2osoi_liq, c, joff + j) = A(a->h2osoi_liq, c, joff + j) + dwat2[j] * dzmm[j];
    if (jw < n) {
        const double wh_zwt = 0.0;
        const double s_node = JMAX(A(a->h2osoi_vol, c, jw + 1) / A(a->watsat, c, jw + 1), 0.01);
        const double s1 = JMIN(1.0, s_node);
        const double ka = imped[jw + 1] * A(a->hksat, c, jw + 1) *
                          zd_pow(s1, 2.0 * A(a->bsw, c, jw + 1) + 3.0);
        const int jm = jw > 1 ? jw : 1;
        const double smp1 = JMAX(smpmin, smp[jm]);
        const double wh = smp1 - zq[jm];
        double q;
        if (jw == 0)
            q = -ka * (wh_zwt - wh) / ((a->zwt[c] + 1.0e-3) * 1000.0);
        else
            q = -ka * (wh_zwt - wh) / ((a->zwt[c] - A(a->z, c, joff + jw)) * 1000.0 * 2.0);
        q = JMAX(-10.0 / dtime, q);
        q = JMIN(10.0 / dtime, q);
        a->qcharge[c] = q;
    } else {
        a->qcharge[c] = dwat2[n + 1] * dzmm[n + 1] / dtime;
    }
    /* _soilwm_deficit_kernel! (:621-633) */
    double acc = 0.0;
    for (int j = 1; j <= n; ++j) {
        const double l = A(a->h2osoi_liq, c, joff + j);
        acc = acc + (l >= 0.0 ? 0.0 : -l);
    }
    a->qflx_deficit[c] = acc;
}

int orc_soilwater_zengdecker2009(const clmsoil_zd_args *a, const clmsoil_levels *lev,
                                 double dtime, const clmsoil_swm_config *cfg) {
    if (!a || !lev || !cfg) return CLMSOIL_ERR_NULL;
    if (cfg->soilwater_movement_method != CLMSOIL_ZENGDECKER_2009) return CLMSOIL_ERR_METHOD;
    if (lev->nlevsoi < 2 || lev->nlevsoi + 1 >= ZMAX || lev->nlevgrnd < lev->nlevsoi + 1)
        return CLMSOIL_ERR_SIZE;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0)) return CLMSOIL_ERR_ARG;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < a->nc; ++c) {
        if (a->mask && !a->mask[c]) continue;
        zd09_column(a, lev, c, dtime, cfg->e_ice);
    }
    return CLMSOIL_OK;
}
