/*
 * clm_math.h — FP64 log, log10, exp and pow used by the soil-physics kernels.
 *
 * One implementation, compiled both into the HIP kernels (device code) and
 * into the CPU restatement the kernels are checked against, so that the two
 * sides round every transcendental identically. Plain C99 with a qualifier
 * macro: a HIP translation unit defines CLM_MATH_FN / CLM_MATH_TAB before it
 * includes this file (device functions, device-resident tables); a C
 * translation unit gets `static inline` / `static const`.
 *
 * Method: the table-driven logarithm and exponential of P. T. P. Tang
 * ("Table-driven implementation of the exponential function in IEEE
 * floating-point arithmetic", ACM TOMS 15(2), 1989; "Table-driven
 * implementation of the logarithm function in IEEE floating-point
 * arithmetic", ACM TOMS 16(4), 1990), carried in double-double where the
 * result of pow needs it:
 *
 *   log x = k*ln2 + log c_i + log1p(r),  x = 2^k z,  r = z/c_i - 1 (exact)
 *   exp t = 2^m * 2^(j/128) * exp(r),    t = (128m + j) ln2/128 + r
 *   pow(x, y) = exp(y * log x)           (log x and y*log x double-double)
 *
 * log1p(r) and exp(r) - 1 use their Taylor polynomials (|r| < 2^-8.4 and
 * |r| < 2^-8.5); the tables (clm_math_tables.h) come from gen_clm_math.py.
 * The error before the final rounding is below 2^-60 relative for pow, exp
 * and log, so the results are the correctly rounded values except in rare
 * near-halfway cases (the tests measure the agreement rate against results
 * computed at 80 digits).
 *
 * Special values follow C99 Annex F (pow: F.9.4.4). Results in the subnormal
 * range are rounded twice (the scaled result, then the scaling), identically
 * on both sides.
 *
 * Every operation is an IEEE add, multiply, fma, rint or ldexp, evaluated in
 * the written order; compile without contraction (-ffp-contract=off).
 */
#ifndef CLM_MATH_H
#define CLM_MATH_H

#include <math.h>
#include <stdint.h>

#ifndef CLM_MATH_FN
#define CLM_MATH_FN static inline
#endif
#ifndef CLM_MATH_TAB
#define CLM_MATH_TAB static const
#endif

#include "clm_math_tables.h"

/* Table reads. A translation unit may redirect them (e.g. to a copy of the
 * tables in a kernel's on-chip memory) by defining these before including
 * this file; the values, and so the results, are the same. */
#ifndef CLM_LOG_TAB
#define CLM_LOG_TAB(i, k) clm_log_tab[i][k]
#endif
#ifndef CLM_EXP_TAB
#define CLM_EXP_TAB(j, k) clm_exp_tab[j][k]
#endif

typedef struct clm_dd {
    double hi, lo;
} clm_dd;

CLM_MATH_FN uint64_t clm_bits(double x) {
    union {
        double d;
        uint64_t u;
    } v;
    v.d = x;
    return v.u;
}
CLM_MATH_FN double clm_from_bits(uint64_t u) {
    union {
        double d;
        uint64_t u;
    } v;
    v.u = u;
    return v.d;
}

/* Knuth's TwoSum: s + e == a + b exactly. */
CLM_MATH_FN double clm_two_sum(double a, double b, double *e) {
    double s = a + b;
    double bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}
/* Dekker's Fast2Sum (|a| >= |b| or a == 0): s + e == a + b exactly. */
CLM_MATH_FN double clm_fast_two_sum(double a, double b, double *e) {
    double s = a + b;
    *e = b - (s - a);
    return s;
}

/* log x as hi + lo for positive finite x (|error| < 2^-72 |log x| away from
 * x = 1, relative 2^-100 in the subinterval around 1): clm_log_ext below.
 * Subnormal x is scaled by 2^52 first (kadj = -52). Straight-line code: no
 * branches. */
/* The exponent extraction below reads tmp's high word alone. */
typedef char clm_log_off_low_word_is_zero[(CLM_LOG_OFF & 0xffffffffULL) == 0 ? 1 : -1];

/* Core: ix = the bits of a positive normal double, kadj added to its exponent. */
CLM_MATH_FN clm_dd clm_log_ext_bits(uint64_t ix, int kadj) {
    const uint64_t tmp = ix - CLM_LOG_OFF;
    const int i = (int)((tmp >> 44) & 255u);
    /* the exponent from tmp's high word in 32-bit arithmetic (CLM_LOG_OFF's low
     * word is zero): (int64_t)tmp >> 52, which would convert to double as an
     * int64 */
    const int32_t thi = (int32_t)(uint32_t)(tmp >> 32);
    const int k = (thi >> 20) + kadj;
    const uint64_t iz = ix - (tmp & 0xfff0000000000000ULL);
    const double z = clm_from_bits(iz);
    const double invc = CLM_LOG_TAB(i, 0);
    const double logc_hi = CLM_LOG_TAB(i, 1);
    const double logc_lo = CLM_LOG_TAB(i, 2);
    const double kd = (double)k;
    /* r = z * invc - 1 is exact: invc has 9 significant bits and |r| < 2^-8. */
    const double r = fma(z, invc, -1.0);
    const double r2 = r * r;
    const double r2lo = fma(r, r, -r2);
    /* log1p(r) - r + r^2/2 = r^3/3 - r^4/4 + ... - r^8/8 */
    double p = -0.125;
    p = fma(p, r, 0x1.2492492492492p-3);   /* 1/7 */
    p = fma(p, r, -0x1.5555555555555p-3);  /* -1/6 */
    p = fma(p, r, 0x1.999999999999ap-3);   /* 1/5 */
    p = fma(p, r, -0.25);
    p = fma(p, r, 0x1.5555555555555p-2);   /* 1/3 */
    p = p * (r2 * r);
    const double t1 = kd * CLM_LN2_HI; /* exact: CLM_LN2_HI has 42 bits, |k| < 2^11 */
    double e1, e2, e3;
    const double s1 = clm_two_sum(t1, logc_hi, &e1);
    const double s2 = clm_two_sum(s1, r, &e2);
    /* |s2| >= r^2/2: either s1 == 0 and s2 == r, or |log x| > 2^-10 */
    const double s3 = clm_fast_two_sum(s2, -0.5 * r2, &e3);
    double lo = kd * CLM_LN2_LO;
    lo = lo + logc_lo;
    lo = lo + e1;
    lo = lo + e2;
    lo = lo + e3;
    lo = lo + -0.5 * r2lo;
    lo = lo + p;
    clm_dd out;
    out.hi = clm_fast_two_sum(s3, lo, &out.lo);
    return out;
}

CLM_MATH_FN clm_dd clm_log_ext(double x) {
    const int sub = clm_bits(x) < 0x0010000000000000ULL;
    return clm_log_ext_bits(clm_bits(sub ? x * 0x1p52 : x), sub ? -52 : 0);
}

/* clm_log_ext for x >= 2^-1022 (no subnormal scaling): the same value. */
CLM_MATH_FN clm_dd clm_log_ext_normal(double x) { return clm_log_ext_bits(clm_bits(x), 0); }

/* exp(hi + lo) with |lo| <= 2^-52 |hi|, rounded to double. Straight-line
 * code: out-of-range and NaN arguments are clamped for the evaluation and the
 * result is selected at the end. */
CLM_MATH_FN double clm_exp_dd(double hi, double lo) {
    /* above 709.8 the result overflows, below -746 it underflows to zero; a NaN
     * (whose result is selected at the end) is evaluated at -746 */
    const double hc = fmin(fmax(hi, -746.0), 709.8);
    const double lc = (hc == hi) ? lo : 0.0;
    const double kd = rint(hc * CLM_INVLN2N);
    const int ki = (int)kd;   /* |kd| < 2^18 */
    /* hc - kd*ln2/128: exact, kd * CLM_LN2N_HI has at most 53 bits */
    const double rh = fma(-kd, CLM_LN2N_HI, hc);
    /* the low part of kd*ln2/128 reaches 1e-8: r = rh + rl carries it exactly */
    const double rl = lc - kd * CLM_LN2N_LO;
    double rerr;
    const double r = clm_two_sum(rh, rl, &rerr);
    const int j = ki & 127;
    const int m = ki >> 7;    /* arithmetic shift: floor(ki / 128) */
    const double th = CLM_EXP_TAB(j, 0);
    const double tl = CLM_EXP_TAB(j, 1);
    /* exp(r + rerr) - 1 = r + q,  q = rerr + r^2 (1/2 + r/6 + ... + r^4/720) */
    double p = 0x1.6c16c16c16c17p-10;        /* 1/720 */
    p = fma(p, r, 0x1.1111111111111p-7);     /* 1/120 */
    p = fma(p, r, 0x1.5555555555555p-5);     /* 1/24 */
    p = fma(p, r, 0x1.5555555555555p-3);     /* 1/6 */
    p = fma(p, r, 0.5);
    const double q = fma(r * r, p, rerr);
    /* th (1 + r + q) + tl: th + th*r exactly (Fast2Sum), then the tail */
    const double a = th * r;
    const double alo = fma(th, r, -a);
    double e;
    const double u = clm_fast_two_sum(th, a, &e);
    double tail = tl + th * q;
    tail = tail + alo;
    tail = tail + e;
    double res = ldexp(u + tail, m);
    /* log(DBL_MAX) < 709.79 and log of half the smallest subnormal > -745.14 */
    res = (hi > 0x1.62e42fefa39f0p+9) ? INFINITY : res;
    res = (hi < -0x1.74910d52d3052p+9) ? 0.0 : res;
    return (hi == hi) ? res : hi;
}

CLM_MATH_FN double clm_exp(double x) { return clm_exp_dd(x, 0.0); }

/* Special values of log/log10: -inf at 0, x for NaN and +inf, NaN below 0. */
CLM_MATH_FN double clm_log_special(double x, double v) {
    v = (x == 0.0) ? -INFINITY : v;
    v = (x < 0.0) ? NAN : v;
    return (x != x || x == INFINITY) ? x : v;
}

CLM_MATH_FN double clm_log(double x) {
    const int ok = (x > 0.0) & (x <= 0x1.fffffffffffffp+1023);
    return clm_log_special(x, clm_log_ext(ok ? x : 1.0).hi);
}

/* log10 from log x = d.hi + d.lo */
CLM_MATH_FN double clm_log10_dd(clm_dd d) {
    const double p = d.hi * CLM_INVLN10_HI;
    const double plo = fma(d.hi, CLM_INVLN10_HI, -p);
    double t = d.hi * CLM_INVLN10_LO;
    t = t + d.lo * CLM_INVLN10_HI;
    t = t + plo;
    return p + t;
}

CLM_MATH_FN double clm_log10(double x) {
    const int ok = (x > 0.0) & (x <= 0x1.fffffffffffffp+1023);
    return clm_log_special(x, clm_log10_dd(clm_log_ext(ok ? x : 1.0)));
}

/* exp(y * L) for L = log x given as a double-double (x > 0 finite). */
CLM_MATH_FN double clm_pow_log(clm_dd L, double y) {
    const double th = y * L.hi;
    double tl = fma(y, L.hi, -th);
    tl = tl + y * L.lo;
    return clm_exp_dd(th, tl);
}

/* x^y (C99 pow semantics, F.9.4.4). Straight-line code: |x|^y is evaluated
 * for finite non-zero x and finite y (on 1 and 0 otherwise), and the sign and
 * the special values (x < 0, zeros, infinities, NaN, x = 1, y = 0) are selected
 * afterwards. */
CLM_MATH_FN double clm_pow(double x, double y) {
    const double ax = fabs(x);
    const int xfin = (ax > 0.0) & (ax <= 0x1.fffffffffffffp+1023);
    const int yfin = fabs(y) <= 0x1.fffffffffffffp+1023;
    const int ok = xfin & yfin;
    const double r = clm_pow_log(clm_log_ext(ok ? ax : 1.0), ok ? y : 0.0);
    /* y an integer / an odd integer (every double >= 2^53 is an even integer) */
    const double hy = 0.5 * y;
    const int yint = yfin & (floor(y) == y);
    const int yodd = yint & (floor(hy) != hy);
    double v = r;
    v = (x < 0.0 && ok) ? (yint ? (yodd ? -r : r) : NAN) : v;
    v = (x == 0.0 && yfin) ? ((y < 0.0) ? (yodd ? copysign(INFINITY, x) : INFINITY)
                                        : (yodd ? x : 0.0))
                           : v;
    const double vinf = (y < 0.0) ? 0.0 : INFINITY;
    v = (ax == INFINITY && yfin) ? ((x < 0.0 && yodd) ? -vinf : vinf) : v;
    v = (fabs(y) == INFINITY && x == x)
            ? ((ax == 1.0) ? 1.0 : (((ax < 1.0) == (y < 0.0)) ? INFINITY : 0.0))
            : v;
    v = (x != x || y != y) ? x + y : v;
    return (y == 0.0 || x == 1.0) ? 1.0 : v;
}

/* Fast paths for the arguments the soil physics produces. Each returns exactly
 * what the general function returns on its domain, so a caller may take it for
 * a batch of arguments that all satisfy the domain test and the general
 * function otherwise, and get the same bits either way. */

/* Domain of clm_pow_pos: x positive, normal and finite; y finite. */
CLM_MATH_FN int clm_pow_pos_ok(double x, double y) {
    return (x >= 0x1p-1022) & (x <= 0x1.fffffffffffffp+1023) & (fabs(y) <= 0x1.fffffffffffffp+1023);
}
/* clm_pow(x, y) where clm_pow_pos_ok(x, y): none of the special cases but
 * y == 0 and x == 1 can occur there. */
CLM_MATH_FN double clm_pow_pos(double x, double y) {
    const double r = clm_pow_log(clm_log_ext_normal(x), y);
    return (y == 0.0 || x == 1.0) ? 1.0 : r;
}
/* clm_pow(b, y) for a fixed base b (positive, normal, finite, b != 1) given as
 * Lb = clm_log_ext(b), and finite y (the domain test is fabs(y) <= DBL_MAX). */
CLM_MATH_FN double clm_pow_base(clm_dd Lb, double y) {
    const double r = clm_pow_log(Lb, y);
    return (y == 0.0) ? 1.0 : r;
}
/* clm_log10(x) for positive, normal, finite x (clm_pow_pos_ok(x, 0)). */
CLM_MATH_FN double clm_log10_pos(double x) { return clm_log10_dd(clm_log_ext_normal(x)); }

#endif /* CLM_MATH_H */
