/*
 * lake_oracle.c — CPU restatement of CLM.jl's lake temperature solve: steps 6
 * and 7 of lake_temperature! and the unpack (src/biogeophys/lake_temperature.jl
 * :632-785, 1205-1225, with tridiagonal_multi! of tridiagonal.jl:86-142).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle); see clmsoil_oracle.h.
 * Like the reference, each kernel runs over all columns into [nc, N] scratch
 * arrays (N = nlevsno + nlevlak + nlevgrnd, extended level ji = j_ext + nlevsno)
 * before the next starts. With dbg non-NULL the scratch arrays are copied out
 * (11 blocks of N*nc: zx, cvx, phix, tx before the solve, tkix, factx, fnx, a,
 * b, c1, r) for the tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "clmsoil_oracle.h"

#define LA(X, c, j) (X)[(size_t)(c) + ((size_t)(j) - 1u) * (size_t)a->ld]
#define EX(X, c, ji) (X)[(size_t)(c) + ((size_t)(ji) - 1u) * (size_t)nc]

/* varcon.jl:66 */
static const double L_CNFAC = 0.5;

int orc_lake_temperature_solve(const clmsoil_lake_solve_args *a, const clmsoil_levels *lev,
                               double dtime, double *dbg) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    const int nlevsno = lev->nlevsno, nlevgrnd = lev->nlevgrnd, nlevlak = (int)a->nlevlak;
    const int N = nlevsno + nlevlak + nlevgrnd, extoff = nlevsno;
    const int64_t nc = a->nc;
    for (int64_t c = 0; c < nc; ++c)
        if ((!a->mask_lakec || a->mask_lakec[c]) && (a->snl[c] < -nlevsno || a->snl[c] > 0))
            return CLMSOIL_ERR_SIZE;
    double *buf = (double *)calloc((size_t)13 * (size_t)N * (size_t)(nc > 0 ? nc : 1), sizeof(double));
    if (!buf) return CLMSOIL_ERR_SIZE;
    double *zx = buf, *cvx = zx + (size_t)N * nc, *phix = cvx + (size_t)N * nc,
           *tx = phix + (size_t)N * nc, *tkix = tx + (size_t)N * nc, *factx = tkix + (size_t)N * nc,
           *fnx = factx + (size_t)N * nc, *A = fnx + (size_t)N * nc, *B = A + (size_t)N * nc,
           *C1 = B + (size_t)N * nc, *R = C1 + (size_t)N * nc, *cp = R + (size_t)N * nc,
           *dp = cp + (size_t)N * nc;
#define ON(c) (!a->mask_lakec || a->mask_lakec[c])
#define JTOP(c) ((int)a->snl[c] + 1)
    /* _lake_extcol_setup_kernel! (:632-664) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        for (int j_ext = -nlevsno + 1; j_ext <= nlevlak + nlevgrnd; ++j_ext) {
            const int ji = j_ext + extoff, jprime = j_ext - nlevlak, jj = j_ext + nlevsno;
            if (j_ext < JTOP(c)) continue;
            if (j_ext < 1) {
                EX(zx, c, ji) = LA(a->z, c, jj);
                EX(cvx, c, ji) = LA(a->cv, c, jj);
                EX(phix, c, ji) = j_ext == JTOP(c) ? 0.0 : LA(a->sabg_lyr_col, c, j_ext + nlevsno);
                EX(tx, c, ji) = LA(a->t_soisno, c, jj);
            } else if (j_ext <= nlevlak) {
                EX(zx, c, ji) = LA(a->z_lake, c, j_ext);
                EX(cvx, c, ji) = LA(a->cv_lake, c, j_ext);
                EX(phix, c, ji) = LA(a->phi, c, j_ext);
                EX(tx, c, ji) = LA(a->t_lake, c, j_ext);
            } else {
                const int jpjj = jprime + nlevsno;
                EX(zx, c, ji) = EX(zx, c, nlevlak + extoff) + LA(a->dz_lake, c, nlevlak) / 2.0 +
                                LA(a->z, c, jpjj);
                EX(cvx, c, ji) = LA(a->cv, c, jpjj);
                EX(phix, c, ji) = j_ext == nlevlak + 1 ? a->phi_soil[c] : 0.0;
                EX(tx, c, ji) = LA(a->t_soisno, c, jpjj);
            }
        }
    }
    /* _lake_tkix_kernel! (:673-703) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        for (int ji = 1; ji <= N; ++ji) {
            const int j_ext = ji - nlevsno, jprime = j_ext - nlevlak, jj = j_ext + nlevsno;
            if (j_ext < JTOP(c)) continue;
            if (j_ext < 0) {
                EX(tkix, c, ji) = LA(a->tk, c, jj);
            } else if (j_ext == 0) {
                const double dzp = EX(zx, c, ji + 1) - EX(zx, c, ji);
                EX(tkix, c, ji) = LA(a->tk_lake, c, 1) * LA(a->tk, c, jj) * dzp /
                                  (LA(a->tk, c, jj) * LA(a->z_lake, c, 1) +
                                   LA(a->tk_lake, c, 1) * (-LA(a->z, c, jj)));
            } else if (j_ext < nlevlak) {
                EX(tkix, c, ji) =
                    (LA(a->tk_lake, c, j_ext) * LA(a->tk_lake, c, j_ext + 1) *
                     (LA(a->dz_lake, c, j_ext + 1) + LA(a->dz_lake, c, j_ext))) /
                    (LA(a->tk_lake, c, j_ext) * LA(a->dz_lake, c, j_ext + 1) +
                     LA(a->tk_lake, c, j_ext + 1) * LA(a->dz_lake, c, j_ext));
            } else if (j_ext == nlevlak) {
                const double dzp = EX(zx, c, ji + 1) - EX(zx, c, ji);
                EX(tkix, c, ji) = a->tktopsoillay[c] * LA(a->tk_lake, c, j_ext) * dzp /
                                  (a->tktopsoillay[c] * LA(a->dz_lake, c, j_ext) / 2.0 +
                                   LA(a->tk_lake, c, j_ext) * LA(a->z, c, 1 + nlevsno));
            } else {
                EX(tkix, c, ji) = LA(a->tk, c, jprime + nlevsno);
            }
        }
    }
    /* _lake_tridiag_factors_kernel! (:705-722) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        for (int ji = 1; ji <= N; ++ji) {
            const int j_ext = ji - nlevsno;
            if (j_ext < JTOP(c)) continue;
            EX(factx, c, ji) = dtime / EX(cvx, c, ji);
            if (j_ext < nlevlak + nlevgrnd)
                EX(fnx, c, ji) = EX(tkix, c, ji) * (EX(tx, c, ji + 1) - EX(tx, c, ji)) /
                                 (EX(zx, c, ji + 1) - EX(zx, c, ji));
            else
                EX(fnx, c, ji) = 0.0;
        }
    }
    /* _lake_tridiag_coeffs_kernel! (:724-755) */
    const double cnfac = L_CNFAC, omc = 1.0 - cnfac;
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        for (int ji = 1; ji <= N; ++ji) {
            const int j_ext = ji - nlevsno;
            if (j_ext < JTOP(c)) continue;
            const double f = EX(factx, c, ji);
            if (j_ext == JTOP(c)) {
                const double dzp = EX(zx, c, ji + 1) - EX(zx, c, ji);
                EX(A, c, ji) = 0.0;
                EX(B, c, ji) = 1.0 + omc * f * EX(tkix, c, ji) / dzp;
                EX(C1, c, ji) = -omc * f * EX(tkix, c, ji) / dzp;
                EX(R, c, ji) = EX(tx, c, ji) + f * (a->fin[c] + EX(phix, c, ji) + cnfac * EX(fnx, c, ji));
            } else if (j_ext < nlevlak + nlevgrnd) {
                const double dzm = EX(zx, c, ji) - EX(zx, c, ji - 1);
                const double dzp = EX(zx, c, ji + 1) - EX(zx, c, ji);
                EX(A, c, ji) = -omc * f * EX(tkix, c, ji - 1) / dzm;
                EX(B, c, ji) = 1.0 + omc * f * (EX(tkix, c, ji) / dzp + EX(tkix, c, ji - 1) / dzm);
                EX(C1, c, ji) = -omc * f * EX(tkix, c, ji) / dzp;
                EX(R, c, ji) = EX(tx, c, ji) + cnfac * f * (EX(fnx, c, ji) - EX(fnx, c, ji - 1)) +
                               f * EX(phix, c, ji);
            } else {
                const double dzm = EX(zx, c, ji) - EX(zx, c, ji - 1);
                EX(A, c, ji) = -omc * f * EX(tkix, c, ji - 1) / dzm;
                EX(B, c, ji) = 1.0 + omc * f * EX(tkix, c, ji - 1) / dzm;
                EX(C1, c, ji) = 0.0;
                EX(R, c, ji) = EX(tx, c, ji) - cnfac * f * EX(fnx, c, ji - 1);
            }
        }
    }
    if (dbg) memcpy(dbg, buf, (size_t)11 * (size_t)N * (size_t)nc * sizeof(double));
    /* tridiagonal_multi!(tx, a, b, c1, r, jtop + nlevsno, mask, nc, N) (tridiagonal.jl:86-142) */
    const double tiny = 1e-30;
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        const int j = JTOP(c) + extoff;
        double bet = EX(B, c, j);
        if (fabs(bet) < tiny) bet = tiny;
        double cp_prev = EX(C1, c, j) / bet, dp_prev = EX(R, c, j) / bet;
        EX(cp, c, j) = cp_prev;
        EX(dp, c, j) = dp_prev;
        for (int jj = j + 1; jj <= N; ++jj) {
            double denom = EX(B, c, jj) - EX(A, c, jj) * cp_prev;
            if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0 ? 1.0 : denom);
            cp_prev = EX(C1, c, jj) / denom;
            dp_prev = (EX(R, c, jj) - EX(A, c, jj) * dp_prev) / denom;
            EX(cp, c, jj) = cp_prev;
            EX(dp, c, jj) = dp_prev;
        }
        double u_next = dp_prev;
        EX(tx, c, N) = u_next;
        for (int jj = N - 1; jj >= j; --jj) {
            u_next = EX(dp, c, jj) - EX(cp, c, jj) * u_next;
            EX(tx, c, jj) = u_next;
        }
    }
    /* _lake_unpack_kernel! (:757-785) */
    for (int64_t c = 0; c < nc; ++c) {
        if (!ON(c)) continue;
        for (int ji = 1; ji <= N; ++ji) {
            const int j_ext = ji - nlevsno, jprime = j_ext - nlevlak, jj = j_ext + nlevsno;
            if (j_ext < JTOP(c)) continue;
            if (j_ext < 1)
                LA(a->t_soisno, c, jj) = EX(tx, c, ji);
            else if (j_ext <= nlevlak)
                LA(a->t_lake, c, j_ext) = EX(tx, c, ji);
            else
                LA(a->t_soisno, c, jprime + nlevsno) = EX(tx, c, ji);
        }
    }
#undef ON
#undef JTOP
    free(buf);
    return CLMSOIL_OK;
}
