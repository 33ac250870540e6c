// heat_adjoint.hip — hand-written reverse mode of the heat step on soil columns.
//
// The reference gets this gradient from Enzyme through soil_temperature!
// (src/biogeophys/soil_temperature.jl:358-593) with the band_solve! rule
// (src/infrastructure/enzyme_rules.jl:72-107) for the solve and Enzyme's own
// reverse pass through the thermal properties and phase_change_beta!
// (:1929-2289); the reverse driver calls it per step (driver_reverse.jl:169-190).
//
// Scope: columns without snow layers, standing surface water, unresolved snow or
// excess ice (snl = 0, frac_h2osfc = 0, h2osno_no_layers = 0, excess_ice = 0),
// i.e. soil columns that may freeze and thaw. Their heat system is the soil
// tridiagonal (the h2osfc row is decoupled: its coupling into soil layer 1 is
// exactly zero). Other columns, and columns whose elimination would interchange
// rows, set DEV_ERR_ADJ and are left out.
//
// One lane owns one column:
//   1. replay the step from its start state with the heat kernel's arithmetic
//      (same expressions, same order, so the same branch decisions): Kersten /
//      Johansen conductivity, heat capacity, interface conductivity, fact, fn,
//      the Crank-Nicolson rows, the elimination and back substitution, and the
//      phase change with its supercooled-water limit;
//   2. walk back: phase change (hm, xm, the new ice and liquid, the temperature
//      correction; max/min pass the adjoint to the branch they selected, the
//      melt/freeze labels are piecewise constant), the transposed solve
//      (lambda = A^-T tbar; rbar = lambda, diagonal bar = -lambda x, off-diagonal
//      bar = -lambda x[neighbour], as in the band_solve! rule), the rows, fn, fact,
//      the interface conductivity, heat capacity and conductivity.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/clmsoil.h"
#include "../../include/clmsoil_adjoint.h"
#include "common.hpp"
#include "state.hpp"

namespace clmsoil {
namespace {

constexpr int NG = MAX_NLEVGRND;

struct HeatAdjParams {
    clmsoil_heat_args a;
    int nlevsno, nlevgrnd;
    double dtime;
    double *t_bar, *liq_bar, *ice_bar;
    double *g_bsw, *g_watsat, *g_scale;
    unsigned *err;
};

__global__ __launch_bounds__(64) void heat_adjoint_kernel(HeatAdjParams P) {
    const clmsoil_heat_args &a = P.a;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    if (a.mask && !a.mask[c]) return;
    const int64_t ld = a.ld;
    const int ns = P.nlevsno, N = P.nlevgrnd;
    const double dtime = P.dtime, cnfac = CNFAC, omc = 1.0 - CNFAC;
    const double tfrz = TFRZ, hfus = HFUS;
    if (a.snl[c] != 0 || a.frac_h2osfc[c] != 0.0 || a.h2osno_no_layers[c] != 0.0) {
        atomicOr(P.err, (unsigned)DEV_ERR_ADJ);
        return;
    }
    const int nbedrock = (int)a.nbedrock[c];
    const double dhsdT = a.dhsdT[c], hs_top = a.hs_top_snow[c];
    const double eflx_bot = a.eflx_bot ? a.eflx_bot[c] : 0.0;
    auto lev = [&](int j) { return c + (int64_t)(ns + j - 1) * ld; };   // soil level j
    auto soil = [&](int j) { return c + (int64_t)(j - 1) * ld; };

    // ---- replay: properties ----
    double t[NG], liq[NG], ice[NG], z[NG], dz[NG], zit[NG], thk[NG], cv[NG];
    for (int j = 1; j <= N; ++j) {
        const int q = j - 1;
        t[q] = a.t_soisno[lev(j)];
        liq[q] = a.h2osoi_liq[lev(j)];
        ice[q] = a.h2osoi_ice[lev(j)];
        z[q] = a.z[lev(j)];
        dz[q] = a.dz[lev(j)];
        zit[q] = a.zi[lev(j)];
        const double exc = a.excess_ice[soil(j)];
        if (exc != 0.0) {
            atomicOr(P.err, (unsigned)DEV_ERR_ADJ);
            return;
        }
        const double wsat = a.watsat[soil(j)];
        const double dzi = dz[q];
        double th;
        double satw = (liq[q] / DENH2O + ice[q] / DENICE + exc / DENICE) / (dzi * wsat);
        satw = jl_min(1.0, satw);
        if (satw > 1.0e-7) {
            double dke;
            if (t[q] >= TFRZ) dke = jl_max(0.0, clm_log10(satw) + 1.0);
            else dke = satw;
            const double fl_denom = liq[q] / (DENH2O * dzi) + ice[q] / (DENICE * dzi) +
                                    exc / (DENICE * dzi);
            const double fl = (liq[q] / (DENH2O * dzi)) / fl_denom;
            const double dksat = a.tkmg[soil(j)] * clm_pow(TKWAT, fl * wsat) *
                                 clm_pow(TKICE, (1.0 - fl) * wsat);
            th = dke * dksat + (1.0 - dke) * a.tkdry[soil(j)];
        } else {
            th = a.tkdry[soil(j)];
        }
        if (j > nbedrock) th = THK_BEDROCK;
        thk[q] = th;
        double v = a.csol[soil(j)] * (1.0 - wsat) * dzi + (ice[q] * CPICE + liq[q] * CPLIQ) +
                   exc * CPICE;
        if (j > nbedrock) v = CSOL_BEDROCK * dzi;
        cv[q] = v;
    }
    // ---- interface conductivity, fact, fn ----
    double tk[NG], fact[NG], fn[NG];
    for (int q = 0; q < N; ++q) {
        const int j = q + 1;
        double tkv = 0.0, fa, f;
        if (j <= N - 1) {
            const double zn = z[q + 1], zin = zit[q + 1];
            tkv = thk[q] * thk[q + 1] * (zn - z[q]) /
                  (thk[q] * (zn - zin) + thk[q + 1] * (zin - z[q]));
        }
        if (j == 1) {
            const double zn = z[q + 1];
            fa = dtime / cv[q] * dz[q] / (0.5 * (z[q] - zit[q] + CAPR * (zn - zit[q])));
            f = tkv * (t[q + 1] - t[q]) / (zn - z[q]);
        } else if (j <= N - 1) {
            fa = dtime / cv[q];
            f = tkv * (t[q + 1] - t[q]) / (z[q + 1] - z[q]);
        } else {
            fa = dtime / cv[q];
            f = eflx_bot;
        }
        tk[q] = tkv;
        fact[q] = fa;
        fn[q] = f;
    }
    // ---- rows (soil layer 1 is the top row) ----
    double rv[NG], dg[NG], lo[NG], up[NG];
    for (int q = 0; q < N; ++q) {
        const int j = q + 1;
        const double zp = j == 1 ? 0.0 : z[q - 1];
        const double zn = j == N ? 0.0 : z[q + 1];
        const double dzp = zn - z[q], dzm = z[q] - zp;
        const double tkp = j == 1 ? 0.0 : tk[q - 1];
        const double fnp = j == 1 ? 0.0 : fn[q - 1];
        const double of = omc * fact[q];
        const double q1 = of * tk[q] / dzp;
        const double q2 = of * tkp / dzm;
        const double q3 = tk[q] / dzp;
        const double q4 = tkp / dzm;
        double r, d, l = 0.0, u = 0.0;
        if (j == 1) {
            r = t[q] + fact[q] * (hs_top - dhsdT * t[q] + cnfac * fn[q]);
            d = 1.0 + q1 - fact[q] * dhsdT;
            u = -q1;
        } else if (j <= N - 1) {
            r = t[q] + cnfac * fact[q] * (fn[q] - fnp);
            u = -q1;
            d = 1.0 + of * (q3 + q4);
            l = -q2;
        } else {
            r = t[q] - cnfac * fact[q] * fnp + fact[q] * fn[q];
            d = 1.0 + q2;
            l = -q2;
        }
        rv[q] = r;
        dg[q] = d;
        lo[q] = l;
        up[q] = u;
    }
    // ---- elimination (no interchange) and back substitution ----
    double u0[NG], y[NG], fm[NG], x[NG];
    u0[0] = dg[0];
    y[0] = rv[0];
    fm[0] = 0.0;
    for (int q = 1; q < N; ++q) {
        if (fabs(lo[q]) > fabs(u0[q - 1])) {
            atomicOr(P.err, (unsigned)DEV_ERR_ADJ);
            return;
        }
        const double f = lo[q] / u0[q - 1];
        fm[q] = f;
        u0[q] = dg[q] - f * up[q - 1];
        y[q] = rv[q] - f * y[q - 1];
    }
    {
        double xn = 0.0;
        for (int q = N - 1; q >= 0; --q) {
            double s = y[q];
            if (q < N - 1) s = s - up[q] * xn;
            x[q] = s / u0[q];
            xn = x[q];
        }
    }

    // ---- reverse: phase change, level by level ----
    double xbar[NG], fbar[NG], lbar[NG], ibar[NG], gw[NG], gb[NG];
    for (int q = 0; q < N; ++q) {
        const int j = q + 1;
        const double tb = P.t_bar[lev(j)], lb = P.liq_bar[lev(j)], ib = P.ice_bar[lev(j)];
        fbar[q] = gw[q] = gb[q] = 0.0;
        const double wsat = a.watsat[soil(j)], bsw = a.bsw[soil(j)], suc = a.sucsat[soil(j)];
        const double wice0 = ice[q], wmass0 = ice[q] + liq[q] + 0.0;
        double tt = x[q], tinc = 0.0, supercool = 0.0;
        int imelt = 0;
        if (ice[q] > 0.0 && tt > tfrz) { imelt = 1; tinc = tfrz - tt; tt = tfrz; }
        double smp = 0.0, sx = 0.0, sp = 0.0;
        const bool sc_on = tt < tfrz;
        if (sc_on) {
            smp = hfus * (tfrz - tt) / (GRAV * tt) * 1000.0;
            sx = smp / suc;
            sp = clm_pow(sx, -1.0 / bsw);
            supercool = wsat * sp;
            supercool *= dz[q] * 1000.0;
        }
        if (liq[q] > supercool && tt < tfrz) { imelt = 2; tinc = tfrz - tt; tt = tfrz; }
        const double fct = fact[q];
        double hm = 0.0, qh = 0.0;
        if (imelt > 0) {
            qh = tinc / fct;
            hm = (j == 1) ? dhsdT * tinc - qh : -qh;
        }
        if (imelt == 1 && hm < 0.0) { hm = 0.0; imelt = 0; }
        if (imelt == 2 && hm > 0.0) { hm = 0.0; imelt = 0; }
        const bool clamped = tt == tfrz && x[q] != tfrz;   // tt set to tfrz above
        if (!(imelt > 0 && fabs(hm) > 0.0)) {
            // no mass change: t_out is x (or tfrz where a label was dropped)
            xbar[q] = clamped ? 0.0 : tb;
            lbar[q] = lb;
            ibar[q] = ib;
            continue;
        }
        const double xm = hm * dtime / hfus;
        double icev = wice0;
        int isel = 0;   // 0: icev = wice0 - xm, 1: 0, 2: wmass0 - supercool
        if (xm > 0.0) {
            const double v = wice0 - xm;
            icev = jl_max(0.0, v);
            isel = (v >= 0.0) ? 0 : 1;
        } else if (xm < 0.0) {
            if (wmass0 - 0.0 < supercool) {
                icev = 0.0;
                isel = 1;
            } else {
                const double v1 = wmass0 - 0.0 - supercool, v2 = wice0 - xm;
                icev = jl_min(v1, v2);
                isel = __builtin_signbit(v1 - v2) ? 2 : 0;
            }
        }
        const double qi = hfus * (wice0 - icev) / dtime;
        const double heatr = (xm > 0.0 || xm < 0.0) ? hm - qi : 0.0;
        const double lv0 = wmass0 - icev - 0.0;
        const bool lsel = !(lv0 < 0.0);   // jl_max(0, lv0) passes lv0 unless negative
        const double den = (j == 1) ? 1.0 - (1.0 - 0.0) * fct * dhsdT : 1.0;
        // t_out = tfrz + fct*heatr/den when |heatr| > 0, else tfrz
        double hmb = 0.0, fctb = 0.0, icevb = ib, wice0b = 0.0, wmass0b = 0.0, scb = 0.0;
        if (fabs(heatr) > 0.0) {
            const double heatrb = tb * fct / den;
            fctb += tb * heatr / den;
            if (j == 1) fctb += tb * fct * heatr * dhsdT / (den * den);
            hmb += heatrb;
            // heatr = hm - hfus*(wice0 - icev)/dtime
            wice0b -= heatrb * hfus / dtime;
            icevb += heatrb * hfus / dtime;
        }
        if (lsel) {   // liq_out = wmass0 - icev
            wmass0b += lb;
            icevb -= lb;
        }
        double xmb = 0.0;
        if (isel == 0) {
            wice0b += icevb;
            xmb -= icevb;
        } else if (isel == 2) {
            wmass0b += icevb;
            scb -= icevb;
        }
        hmb += xmb * dtime / hfus;
        // hm = dhsdT*tinc - tinc/fct (soil layer 1) or -tinc/fct
        double tincb = (j == 1) ? hmb * dhsdT : 0.0;
        tincb -= hmb / fct;
        fctb += hmb * tinc / (fct * fct);
        double xb = -tincb;   // tinc = tfrz - x
        if (sc_on && scb != 0.0) {
            // supercool = wsat * sx^(-1/bsw) * dz * 1000, sx = smp/sucsat,
            // smp = hfus (tfrz - x) / (GRAV x) * 1000
            const double k = dz[q] * 1000.0;
            gw[q] += scb * sp * k;
            const double spb = scb * wsat * k;
            const double e = -1.0 / bsw;
            const double sxb = spb * e * sp / sx;
            gb[q] += spb * sp * clm_log(sx) / (bsw * bsw);
            const double smpb = sxb / suc;
            xb += smpb * (-hfus * 1000.0 * tfrz / (GRAV * x[q] * x[q]));
        }
        xbar[q] = xb;
        fbar[q] += fctb;
        lbar[q] = wmass0b;
        ibar[q] = wmass0b + wice0b;
    }

    // ---- reverse: solve (lambda = A^-T xbar with the forward LU) ----
    double lam[NG];
    lam[0] = xbar[0] / u0[0];
    for (int q = 1; q < N; ++q) lam[q] = (xbar[q] - up[q - 1] * lam[q - 1]) / u0[q];
    for (int q = N - 2; q >= 0; --q) lam[q] = lam[q] - fm[q + 1] * lam[q + 1];

    // ---- reverse: rows -> t, fact, fn, tk ----
    double tbar_in[NG], fnb[NG], tkb[NG];
    for (int q = 0; q < N; ++q) tbar_in[q] = fnb[q] = tkb[q] = 0.0;
    for (int q = 0; q < N; ++q) {
        const int j = q + 1;
        const double rb = lam[q];
        const double db = -lam[q] * x[q];
        const double lob = j == 1 ? 0.0 : -lam[q] * x[q - 1];
        const double upb = j == N ? 0.0 : -lam[q] * x[q + 1];
        const double zp = j == 1 ? 0.0 : z[q - 1];
        const double zn = j == N ? 0.0 : z[q + 1];
        const double dzp = zn - z[q], dzm = z[q] - zp;
        const double tkp = j == 1 ? 0.0 : tk[q - 1];
        const double fnp = j == 1 ? 0.0 : fn[q - 1];
        const double of = omc * fact[q];
        double q1b = 0.0, q2b = 0.0, q3b = 0.0, q4b = 0.0, ofb = 0.0, fab = 0.0;
        if (j == 1) {
            tbar_in[q] += rb * (1.0 - fact[q] * dhsdT);
            fab += rb * (hs_top - dhsdT * t[q] + cnfac * fn[q]);
            fnb[q] += rb * fact[q] * cnfac;
            q1b += db - upb;
            fab -= db * dhsdT;
        } else if (j <= N - 1) {
            tbar_in[q] += rb;
            fab += rb * cnfac * (fn[q] - fnp);
            fnb[q] += rb * cnfac * fact[q];
            fnb[q - 1] -= rb * cnfac * fact[q];
            q1b -= upb;
            ofb += db * (tk[q] / dzp + tkp / dzm);
            q3b += db * of;
            q4b += db * of;
            q2b -= lob;
        } else {
            tbar_in[q] += rb;
            fab += rb * (fn[q] - cnfac * fnp);
            fnb[q - 1] -= rb * cnfac * fact[q];
            q2b += db - lob;
        }
        // q1 = of tk/dzp, q2 = of tkp/dzm, q3 = tk/dzp, q4 = tkp/dzm
        ofb += q1b * tk[q] / dzp + q2b * tkp / dzm;
        tkb[q] += q1b * of / dzp + q3b / dzp;
        if (j > 1) tkb[q - 1] += q2b * of / dzm + q4b / dzm;
        fab += ofb * omc;
        fbar[q] += fab;
    }
    // fn = tk (t[j+1] - t[j]) / (z[j+1] - z[j])
    for (int q = 0; q < N - 1; ++q) {
        const double dzz = z[q + 1] - z[q];
        tkb[q] += fnb[q] * (t[q + 1] - t[q]) / dzz;
        tbar_in[q + 1] += fnb[q] * tk[q] / dzz;
        tbar_in[q] -= fnb[q] * tk[q] / dzz;
    }
    // fact = dtime/cv (times a geometric factor at the top): d fact/d cv = -fact/cv
    double cvb[NG], thkb[NG];
    for (int q = 0; q < N; ++q) {
        cvb[q] = -fbar[q] * fact[q] / cv[q];
        thkb[q] = 0.0;
    }
    // tk = thk thk' D / (thk A + thk' B)
    for (int q = 0; q < N - 1; ++q) {
        const double D = z[q + 1] - z[q], A = z[q + 1] - zit[q + 1], B = zit[q + 1] - z[q];
        const double den = thk[q] * A + thk[q + 1] * B;
        thkb[q] += tkb[q] * (thk[q + 1] * D / den - tk[q] * A / den);
        thkb[q + 1] += tkb[q] * (thk[q] * D / den - tk[q] * B / den);
    }
    // heat capacity and conductivity -> water, ice, porosity
    double sb = 0.0, sw = 0.0;
    for (int q = 0; q < N; ++q) {
        const int j = q + 1;
        const double wsat = a.watsat[soil(j)], dzi = dz[q];
        if (j <= nbedrock) {
            gw[q] -= cvb[q] * a.csol[soil(j)] * dzi;
            ibar[q] += cvb[q] * CPICE;
            lbar[q] += cvb[q] * CPLIQ;
            const double satw0 = (liq[q] / DENH2O + ice[q] / DENICE + 0.0 / DENICE) / (dzi * wsat);
            const double satw = jl_min(1.0, satw0);
            if (satw > 1.0e-7 && thkb[q] != 0.0) {
                const double a1 = liq[q] / (DENH2O * dzi), a2 = ice[q] / (DENICE * dzi);
                const double fl = a1 / (a1 + a2 + 0.0);
                const double lw = clm_log(TKWAT), li = clm_log(TKICE);
                const double dksat = a.tkmg[soil(j)] * clm_pow(TKWAT, fl * wsat) *
                                     clm_pow(TKICE, (1.0 - fl) * wsat);
                double dke, dkes = 0.0;   // d dke / d satw
                if (t[q] >= TFRZ) {
                    const double v = clm_log10(satw) + 1.0;
                    dke = jl_max(0.0, v);
                    if (!(v < 0.0)) dkes = 1.0 / (satw * 2.302585092994046);
                } else {
                    dke = satw;
                    dkes = 1.0;
                }
                const double dkeb = thkb[q] * (dksat - a.tkdry[soil(j)]);
                const double dksatb = thkb[q] * dke;
                // dksat = tkmg exp(wsat (fl lw + (1 - fl) li))
                const double flb = dksatb * dksat * wsat * (lw - li);
                gw[q] += dksatb * dksat * (fl * lw + (1.0 - fl) * li);
                const double s2 = (a1 + a2) * (a1 + a2);
                lbar[q] += flb * a2 / s2 / (DENH2O * dzi);
                ibar[q] -= flb * a1 / s2 / (DENICE * dzi);
                // satw = min(1, (liq/rho_w + ice/rho_i) / (dz wsat))
                if (!(satw0 > 1.0)) {
                    const double satb = dkeb * dkes;
                    lbar[q] += satb / (DENH2O * dzi * wsat);
                    ibar[q] += satb / (DENICE * dzi * wsat);
                    gw[q] -= satb * satw0 / wsat;
                }
            }
        }
        P.t_bar[lev(j)] = tbar_in[q];
        P.liq_bar[lev(j)] = lbar[q];
        P.ice_bar[lev(j)] = ibar[q];
        if (P.g_watsat) P.g_watsat[soil(j)] += gw[q];
        if (P.g_bsw) P.g_bsw[soil(j)] += gb[q];
        sb += gb[q] * a.bsw[soil(j)];
        sw += gw[q] * wsat;
    }
    if (P.g_scale) {
        P.g_scale[c] += sb;
        P.g_scale[2 * ld + c] += sw;
    }
}

}  // namespace
}  // namespace clmsoil

using namespace clmsoil;

extern "C" int clmsoil_soil_temperature_adjoint(const clmsoil_heat_args *a,
                                                const clmsoil_levels *lev, double dtime,
                                                double *t_bar, double *liq_bar, double *ice_bar,
                                                double *grad_bsw, double *grad_watsat,
                                                double *grad_scale, void *stream) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (lev->nlevgrnd > MAX_NLEVGRND) return CLMSOIL_ERR_SIZE;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->snl || !a->nbedrock || !a->z || !a->dz || !a->zi || !a->t_soisno ||
        !a->h2osoi_liq || !a->h2osoi_ice || !a->excess_ice || !a->watsat || !a->tkmg ||
        !a->tkdry || !a->csol || !a->bsw || !a->sucsat || !a->frac_h2osfc ||
        !a->hs_top_snow || !a->dhsdT || !a->h2osno_no_layers || !t_bar || !liq_bar || !ice_bar)
        return CLMSOIL_ERR_NULL;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    HeatAdjParams P{*a, (int)lev->nlevsno, (int)lev->nlevgrnd, dtime, t_bar, liq_bar, ice_bar,
                    grad_bsw, grad_watsat, grad_scale, s->err};
    const int64_t grid = (a->nc + 63) / 64;
    hipLaunchKernelGGL(heat_adjoint_kernel, dim3((unsigned)grid), dim3(64), 0,
                       (hipStream_t)stream, P);
    return hipGetLastError() == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}
