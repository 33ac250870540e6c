// rcolumn.hpp — one column of the moisture-form Richards step held by a group of
// G lanes (layer j in lane (j-1) mod G, register slot (j-1) div G), shared by the
// primal kernels and the adjoint's replay of the primal step.
//
// Register budget. Across substeps a lane keeps, per slot, only what the substep
// loop re-reads on every iteration: the water state, 1000*dz, the impedance-scaled
// conductivity khat, the node-distance denominator den, porosity, and the four
// flux terms of the linear system (qin - qout - root, dqidw0, -dqidw1 + dqodw1,
// dqodw2). bsw, sucsat and the root sink are read again from memory when the
// hydraulic properties are recomputed (after an accepted substep), and smp/hk are
// written through to smp_l/hk_l at that moment instead of being held. The row
// coefficients of the tridiagonal system are formed for all rows in parallel
// before the serial sweep, which then carries only (bet, dwat) between rows.
// Every value is computed with the reference's operation order, so the results
// are the same bits as a straightforward restatement.
//
// EXP = true keeps the separate flux terms the "expensive" error estimate needs
// (flux_calculation = 42). REC = true records every accepted substep (its start
// state and length) to a tape [k][j][c] for the adjoint (k < kmax; kacc counts
// them all).
#pragma once
#include <cstddef>

#include "common.hpp"
#include "mlp_wave.hpp"
#include "richards.hpp"

namespace clmsoil {

// Argument pointers of the Richards kernels (RichardsParams, their first kernel
// argument) are fetched from the kernarg segment where they are used, so that
// they occupy no SGPRs across the substep loop.
#define RP(field) \
    karg_ptr<decltype(RichardsParams::field), (unsigned)offsetof(RichardsParams, field)>()

// Value of layer j+1 (j-1) for the layer in slot i of lane k. In the group
// rotation only the last (first) lane reads across the slot boundary, from the
// first (last) lane, so that lane sends its next (previous) slot instead: one
// select and one DPP move per call.
template <int G>
__device__ __forceinline__ double nxt(const double *x, int i, int k, int S) {
    if constexpr (G == 1) return (i + 1 < S) ? x[i + 1] : 0.0;
    const double up = (i + 1 < S) ? x[i + 1] : 0.0;
    return group_next<G>(k == 0 ? up : x[i]);
}
template <int G>
__device__ __forceinline__ double prv(const double *x, int i, int k) {
    if constexpr (G == 1) return (i > 0) ? x[i - 1] : 0.0;
    const double dn = (i > 0) ? x[i - 1] : 0.0;
    return group_prev<G>(k == G - 1 ? dn : x[i]);
}

template <int G, bool EXP = false, bool REC = false, int LCS = 0, int LCF = 7,
          bool GRID = false>
struct RColumn {
    static constexpr int NMAX = MAX_NLEVSOI;
    static constexpr int S = (NMAX + G - 1) / G;
    static constexpr int SE = EXP ? S : 1;
    // tape of accepted substeps (REC only): tape[(k*(NMAX+1) + j-1)*tape_ld + c] is
    // layer j's liquid water at the start of accepted substep k; row NMAX its length
    double *tape = nullptr;
    int64_t tape_ld = 0;
    int kacc = 0, kmax = 0;

    int64_t c;
    int n, jwt, nsubstep, nrej, iter;
    bool props_valid;
    double zwt, qflx_infl, dtsub, dtdone, qc;
    // bottom-layer drainage qct = qn_a + qn_b * dwat[n] (owner lane of layer n)
    double qn_a, qn_b;
    double liq[S], mdz[S], khat[S], den[S], watsat[S];
    double fnet[S], a0[S], sb[S], c2[S];
    double qin[SE], qout[SE], di1[SE], do1[SE];
    // LDS cache of the column when LCS > 0 (LCS = the block size): bsw, sucsat and
    // the root sink are copied in once per column instead of being read from memory
    // at every property update, and smp/hk are kept there and written to smp_l/hk_l
    // once, when the column is stored. Lane-private: field f, slot i at
    // lc[(f*S + i) * LCS], so a wavefront's access is one contiguous, conflict-free
    // line and every offset is an immediate.
    static constexpr bool LC = LCS > 0;
    // LCF: fields in the cache, 7 (all below) or 5 (khat and den stay in registers)
    static_assert(LCF == 5 || LCF == 7, "cached field sets: 5 or 7");
    static constexpr bool LC_KD = LC && LCF >= 7;
    double *lc = nullptr;
    enum { LC_BSW = 0, LC_SUC = 1, LC_ROOT = 2, LC_SMP = 3, LC_HK = 4, LC_KHAT = 5, LC_DEN = 6,
           LC_FIELDS = LCF };
    __device__ __forceinline__ double &cache(int f, int i) const { return lc[(f * S + i) * LCS]; }
    // khat and den are read once per property update: in LDS when cached
    __device__ __forceinline__ double &kh(int i) {
        if constexpr (LC_KD) return cache(LC_KHAT, i);
        else return khat[i];
    }
    __device__ __forceinline__ double &dn(int i) {
        if constexpr (LC_KD) return cache(LC_DEN, i);
        else return den[i];
    }

    // GRID: the CLM5 grid's level counts (12 snow, 20 soil) as compile-time constants
    __device__ __forceinline__ static int nsoi(const RichardsParams &P) {
        if constexpr (GRID) return MAX_NLEVSOI;
        else return P.nlevsoi;
    }
    __device__ __forceinline__ static int nsno(const RichardsParams &P) {
        if constexpr (GRID) return MAX_NLEVSNO;
        else return P.nlevsno;
    }

    __device__ __forceinline__ int64_t soil(int j) const { return c + (int64_t)(j - 1) * ld_; }
    int64_t ld_;

    // Load column `col`; returns false (and flags the error word) for a bad nbedrock.
    __device__ __forceinline__ bool load(const RichardsParams &P, int64_t col, int k) {
        c = col;
        ld_ = P.ld;
        const int64_t ld = P.ld;
        const int nlevsoi = nsoi(P);
        const int joff = nsno(P), joff_zi = nsno(P) + 1;
        n = (int)RP(nbedrock)[c];
        // water-table index (BC_WATERTABLE, :1346-1350): the first interface jj in
        // 1..nlevsoi with zwt <= zi[jj] gives jwt = jj - 1 (nlevsoi when none).
        // Each lane tests its own interfaces and the group takes the minimum, so
        // every load is in flight at once (a refill stalls the whole wavefront).
        jwt = nlevsoi;
        zwt = 0.0;
        if (P.lower_bc == CLMSOIL_BC_WATERTABLE || P.aquifer_pass) {
            zwt = RP(zwt)[c];
            int jf = nlevsoi;
            const double *p_zi = RP(zi);
#pragma unroll
            for (int i = 0; i < S; ++i) {
                const int jj = i * G + k + 1;
                if (jj <= nlevsoi) {
                    const double zij = p_zi[c + (int64_t)(joff_zi + jj - 1) * ld];
                    if (zwt <= zij && jj - 1 < jf) jf = jj - 1;
                }
            }
            jwt = group_min_i<G>(jf);
        }
        if (n < 1 || n > nlevsoi) {
            if (k == 0) P.flag((unsigned)DEV_ERR_NBEDROCK);
            return false;
        }
        const clm_dd L10 = clm_log_ext_normal(10.0);   // impedance 10^(-e_ice * icefrac)
        const double *p_liq = RP(h2osoi_liq), *p_dz = RP(dz), *p_ws = RP(watsat), *p_hks = RP(hksat);
        const double *p_icf = RP(icefrac), *p_z = RP(z);
        const double *p_bsw = LC ? RP(bsw) : nullptr, *p_suc = LC ? RP(sucsat) : nullptr,
                     *p_root = LC ? RP(qflx_rootsoi) : nullptr;
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            double lq = 0.0, dzv = 1.0, ws = 1.0, hks = 0.0, icf = 0.0, icf1 = 0.0, zj = 0.0,
                   zj1 = 0.0, bs = 1.0, su = 0.0, rt = 0.0;
            // only the levels above bedrock are read (n is known here: the bedrock
            // test above has already waited for it), and layer j+1's ice fraction
            // and depth only where layer j has an interior interface below it
            if (j <= n) {
                const int64_t o = soil(j);
                const int64_t os = c + (int64_t)(joff + j - 1) * ld;
                lq = p_liq[os];
                dzv = p_dz[os];
                ws = p_ws[o];
                hks = p_hks[o];
                icf = p_icf[o];
                zj = p_z[os];
                if (j < n) {
                    icf1 = p_icf[o + ld];
                    zj1 = p_z[os + ld];
                }
                if constexpr (LC) {
                    bs = p_bsw[o];
                    su = p_suc[o];
                    rt = p_root[o];
                }
            }
            if (j <= n) {
                liq[i] = lq;
                mdz[i] = M_TO_MM * dzv;
                watsat[i] = ws;
                double imp;
                if (j == n) {
                    imp = pow_base_w(L10, 10.0, -P.e_ice * icf);
                    dn(i) = M_TO_MM * (zwt - zj);                      // :1357 (WATERTABLE)
                } else {
                    imp = pow_base_w(L10, 10.0, -P.e_ice * (0.5 * (icf + icf1)));
                    dn(i) = M_TO_MM * (zj1 - zj);                      // :1338
                }
                kh(i) = imp * hks;   // (imp * hksat) * s1^(2b+3), the reference's grouping
                if constexpr (LC) {
                    cache(LC_BSW, i) = bs;
                    cache(LC_SUC, i) = su;
                    cache(LC_ROOT, i) = rt;
                }
            } else {
                liq[i] = 0.0; mdz[i] = M_TO_MM; watsat[i] = 1.0; kh(i) = 0.0; dn(i) = 1.0;
            }
            fnet[i] = a0[i] = sb[i] = c2[i] = 0.0;
            if constexpr (EXP) qin[i] = qout[i] = di1[i] = do1[i] = 0.0;
        }
        qn_a = qn_b = 0.0;
        qflx_infl = RP(qflx_infl)[c];
        nsubstep = 0;
        nrej = 0;
        iter = 0;
        const double dtime = P.dtime;
        dtsub = dtime;
        if (const double *p_d0 = RP(dtsub0)) {
            const double d0 = p_d0[c], lo = jl_min(P.dtmin, dtime);
            dtsub = (d0 != d0) ? dtime : jl_min(jl_max(d0, lo), dtime);
        }
        dtdone = 0.0;
        qc = 0.0;
        props_valid = false;
        return true;
    }

    // First substep from the network of clmsoil_learned.h for a column just
    // loaded (`on`): its features are the node saturations of the loaded state
    // (the reference's s2, :1273-1286, 0 below bedrock), log ksat and b of layer
    // 1, the infiltration flux and dtime; dtsub0 = dtime * exp(y_mean + y_scale*y)
    // clamped to [min(dtmin, dtime), dtime], dtime when it is NaN. Every lane of
    // the wavefront calls this; lanes with `on` false contribute nothing.
    __device__ __forceinline__ void learned_first_substep(const RichardsParams &P, int k, bool on) {
        static_assert(G == 4, "the network is evaluated for groups of 4 lanes");
        float fe[6];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            float raw = 0.0f;
            if (on && j <= n) {
                double v = jl_max(liq[i], 1.0e-6) / mdz[i];
                v = v / watsat[i];
                v = jl_min(v, 1.0);
                v = jl_max(0.01, v);
                raw = (float)v;
            }
            fe[i] = mlp_standardise(P.mlp_norm, 4 * i + k, raw);
        }
        float r5 = 0.0f;
        if (on) {
            if (k == 0) r5 = (float)log(RP(hksat)[c]);
            else if (k == 1) r5 = (float)RP(bsw)[c];
            else if (k == 2) r5 = (float)qflx_infl;
            else r5 = (float)P.dtime;
        }
        fe[5] = mlp_standardise(P.mlp_norm, 20 + k, r5);
        const float y = mlp_wave16(P.mlp, fe[0], fe[1], fe[2], fe[3], fe[4], fe[5]);
        if (on) {
            const double dtime = P.dtime;
            const double lr = (double)P.mlp_norm[2 * mlp::NIN] + (double)P.mlp_norm[2 * mlp::NIN + 1] * (double)y;
            const double d0 = dtime * exp(lr), lo = jl_min(P.dtmin, dtime);
            dtsub = (d0 != d0) ? dtime : jl_min(jl_max(d0, lo), dtime);
        }
    }

    // Hydraulic properties, interface fluxes and the flux terms of the system
    // for the current water state (:1273-1378).
    __device__ __forceinline__ void properties(const RichardsParams &P, int k) {
        const int lower_bc = P.lower_bc;
        // node saturation s2 (:1273-1286)
        double s2[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const double vwc = jl_max(liq[i], 1.0e-6) / mdz[i];   // dz * denh2o
            double s = vwc / watsat[i];
            s = jl_min(s, 1.0);
            s = jl_max(0.01, s);
            s2[i] = s;
        }
        // node suction and its derivative (:1300-1309)
        double smp[S], dsmpdw[S], bsw[S];
        const double *p_bsw = LC ? nullptr : RP(bsw), *p_suc = LC ? nullptr : RP(sucsat);
        double *p_smp = LC ? nullptr : RP(smp_l), *p_hk = LC ? nullptr : RP(hk_l);
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            double b = 1.0, suc = 0.0;
            if (j <= n) {
                if constexpr (LC) {
                    b = cache(LC_BSW, i);
                    suc = cache(LC_SUC, i);
                } else {
                    b = p_bsw[soil(j)];
                    suc = p_suc[soil(j)];
                }
            }
            bsw[i] = b;
            const double sn = s2[i];
            const double smpv = -suc * pow_w(sn, -b);
            const double dsmpds = -b * smpv / sn;
            smp[i] = smpv;
            dsmpdw[i] = dsmpds / watsat[i];
            if constexpr (LC) cache(LC_SMP, i) = smpv;
            else if (j <= n && p_smp) p_smp[soil(j)] = smpv;
        }
        // interface conductivity (:1287-1299) and fluxes (:1334-1367)
        double qo[S], d1[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            const double s2n = nxt<G>(s2, i, k, S);
            double s1 = (j == n) ? s2[i] : 0.5 * (s2[i] + s2n);
            s1 = jl_min(s1, 1.0);
            s1 = jl_max(0.01, s1);
            const double b = bsw[i];
            const double hkv = kh(i) * pow_w(s1, 2.0 * b + 3.0);
            const double dhkdw = (2.0 * b + 3.0) * hkv / s1;
            if constexpr (LC) cache(LC_HK, i) = hkv;
            else if (j <= n && p_hk) p_hk[soil(j)] = hkv;
            const double smpn = nxt<G>(smp, i, k, S);
            const double dsmpdwn = nxt<G>(dsmpdw, i, k, S);
            const double watsatn = nxt<G>(watsat, i, k, S);
            // The interior interface (j < n, :1334-1342) and the water-table bottom
            // (j == n, :1346-1367) have the same expression shapes: the bottom uses
            // dhkdw/watsat instead of 0.5*dhkdw/watsat and -smp instead of
            // smpn - smp. Selecting those operands first lets both cases share one
            // set of divisions, each case getting exactly its own expressions.
            const bool last = j == n;
            const double dhkds1 = (last ? dhkdw : 0.5 * dhkdw) / watsat[i];
            const double num = last ? -smp[i] : smpn - smp[i];
            const double dni = dn(i);
            const double qv = -hkv * num / dni + hkv;
            const double e1v = (hkv * dsmpdw[i] - dhkds1 * num) / dni + dhkds1;
            double q = 0.0, e1 = 0.0, e2 = 0.0;
            if (j < n) {
                const double dhkds2 = 0.5 * dhkdw / watsatn;
                q = qv;
                e1 = e1v;
                e2 = (-hkv * dsmpdwn - dhkds2 * num) / dni + dhkds2;
            } else if (last) {
                if (lower_bc == CLMSOIL_BC_WATERTABLE) {
                    if (!(j > jwt)) {
                        q = qv;
                        e1 = e1v;
                    }
                    qn_a = q;          // qcharge = qout + dqodw1 * dwat (:1476-1478)
                    qn_b = e1;
                } else if (lower_bc == CLMSOIL_BC_FLUX) {
                    q = hkv;
                    e1 = dhkds1;       // dhkdw / watsat
                    qn_a = hkv;        // qcharge = hk + dhkdw * dwat (:1471-1472)
                    qn_b = dhkdw;
                }
            }
            qo[i] = q;
            d1[i] = e1;
            c2[i] = e2;
        }
        // inflow at the top of each layer (:1371-1378) and the rows' flux terms
        const double *p_root = LC ? nullptr : RP(qflx_rootsoi);
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            const double po = prv<G>(qo, i, k);
            const double p1 = prv<G>(d1, i, k);
            const double p2 = prv<G>(c2, i, k);
            const double qi = (j == 1) ? qflx_infl : po;
            const double dqidw0 = (j == 1) ? 0.0 : p1;
            const double dqidw1 = (j == 1) ? 0.0 : p2;
            double root = 0.0;
            if (j <= n) {
                if constexpr (LC) root = cache(LC_ROOT, i);
                else root = p_root[soil(j)];
            }
            fnet[i] = qi - qo[i] - root;
            a0[i] = dqidw0;
            sb[i] = -dqidw1 + d1[i];
            if constexpr (EXP) {
                qin[i] = qi;
                qout[i] = qo[i];
                di1[i] = dqidw1;
                do1[i] = d1[i];
            }
        }
        props_valid = true;
    }

    // One pass of the reference's substep loop; returns true when the column's
    // step is complete.
    __device__ __forceinline__ bool iterate(const RichardsParams &P, int k) {
        const int lower_bc = P.lower_bc;
        const double dtime = P.dtime;
        if (iter >= P.max_iter) {
            if (k == 0) P.flag((unsigned)DEV_ERR_ITER);
            return true;
        }
        ++iter;
        nsubstep += 1;
        if (!props_valid) properties(P, k);

        // ---- RHS/LHS of every row (:1381-1397), in parallel ----
        double amx[S], bmx[S], cmx[S], rmx[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            const double dt_dz = dtsub / mdz[i];
            rmx[i] = -fnet[i] * dt_dz;
            amx[i] = (j == 1) ? 0.0 : a0[i] * dt_dz;
            bmx[i] = -1.0 - sb[i] * dt_dz;
            cmx[i] = (j == n) ? 0.0 : -c2[i] * dt_dz;
        }
        double cpv[S];   // the previous row's cmx
#pragma unroll
        for (int i = 0; i < S; ++i) cpv[i] = prv<G>(cmx, i, k);

        // ---- Thomas forward sweep (:1421-1431) ----
        double gam[S], dwat[S];
#pragma unroll
        for (int i = 0; i < S; ++i) gam[i] = dwat[i] = 0.0;
        double bet = 0.0, dwp = 0.0;
#pragma unroll
        for (int j = 1; j <= NMAX; ++j) {
            const int i = (j - 1) / G, kk = (j - 1) % G;
            double ob = 0.0;
            if (k == kk && j <= n) {
                double b2, d;
                if (j == 1) {
                    b2 = bmx[i];
                    if (fabs(b2) < 1.0e-30) b2 = 1.0e-30;
                    d = rmx[i] / b2;
                } else {
                    const double g = cpv[i] / bet;
                    gam[i] = g;
                    b2 = bmx[i] - amx[i] * g;
                    if (fabs(b2) < 1.0e-30) b2 = copysign(1.0e-30, b2 == 0.0 ? 1.0 : b2);
                    d = (rmx[i] - amx[i] * dwp) / b2;
                }
                dwat[i] = d;
                ob = b2;
            }
            bet = group_bcast<G>(ob, kk);
            dwp = group_bcast<G>(dwat[i], kk);
        }
        // ---- back substitution (:1434-1438) ----
        double gn = 0.0, dn = 0.0;
#pragma unroll
        // (a broadcast reads only lane kk's slot; rows below the column's last level
        // hold the zeros they were initialised with)
        for (int j = NMAX; j >= 1; --j) {
            const int i = (j - 1) / G, kk = (j - 1) % G;
            if (k == kk && j < n) dwat[i] = dwat[i] - gn * dn;
            gn = group_bcast<G>(gam[i], kk);
            dn = group_bcast<G>(dwat[i], kk);
        }
        // ---- error estimate (:1441-1457) ----
        double errorMax = 0.0;
        const double *p_root = LC ? nullptr : RP(qflx_rootsoi);
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            double fluxNet0;
            const double dt_dz = dtsub / mdz[i];
            double fluxNet1;
            if constexpr (EXP) {
                double root = 0.0;
                if (j <= n) {
                    if constexpr (LC) root = cache(LC_ROOT, i);
                    else root = p_root[soil(j)];
                }
                const double dwm = prv<G>(dwat, i, k);
                const double dwn = nxt<G>(dwat, i, k, S);
                const double qin_test = (j == 1) ? qin[i] + di1[i] * dwat[i]
                                                 : qin[i] + a0[i] * dwm + di1[i] * dwat[i];
                const double qout_test = (j == n) ? qout[i] + do1[i] * dwat[i]
                                                  : qout[i] + do1[i] * dwat[i] + c2[i] * dwn;
                fluxNet0 = qin_test - qout_test - root;
                fluxNet1 = qin[i] - qout[i] - root;
            } else {
                fluxNet0 = dwat[i] / dt_dz;
                fluxNet1 = fnet[i];
            }
            const double ev = fabs(fluxNet1 - fluxNet0) * dtsub * 0.5;
            if (j <= n) errorMax = jl_max(errorMax, ev);
        }
        errorMax = group_max<G>(errorMax);

        if (errorMax > P.xtu && dtsub > P.dtmin) {                 // :1460-1463
            dtsub = jl_max(dtsub / 2.0, P.dtmin);
            nrej += 1;
            return false;
        }
        // accept (:1466-1478)
        if constexpr (REC) {
            if (kacc < kmax) {
                const int64_t row = (int64_t)kacc * (NMAX + 1);
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    const int j = i * G + k + 1;
                    if (j <= n) tape[(row + j - 1) * tape_ld + c] = liq[i];
                }
                if (k == 0) tape[(row + NMAX) * tape_ld + c] = dtsub;
            }
            ++kacc;
        }
#pragma unroll
        for (int i = 0; i < S; ++i) liq[i] = liq[i] + dwat[i] * mdz[i];
        {
            const int in = (n - 1) / G, kn = (n - 1) % G;
            double qct = 0.0;
#pragma unroll
            for (int i = 0; i < S; ++i) {
                if (i != in || k != kn) continue;
                if (lower_bc != CLMSOIL_BC_ZERO_FLUX) qct = qn_a + qn_b * dwat[i];
            }
            qct = group_bcast<G>(qct, kn);
            qc += qct * (dtsub / dtime);
        }
        props_valid = false;
        dtdone += dtsub;                                            // :1480-1483
        if (fabs(dtime - dtdone) < P.verysmall) return true;
        if (errorMax < P.xtl) dtsub = dtsub * 2.0;
        dtsub = jl_min(dtsub, dtime - dtdone);
        return false;
    }

    __device__ __forceinline__ void store(const RichardsParams &P, int k) {
        const int64_t ld = P.ld;
        const int joff = nsno(P);
        const double dtime = P.dtime;
        {
        double *p_liq = RP(h2osoi_liq), *p_smp = RP(smp_l), *p_hk = RP(hk_l);
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int j = i * G + k + 1;
            if (j <= n) p_liq[c + (int64_t)(joff + j - 1) * ld] = liq[i];
            if constexpr (LC) {
                if (j <= n && p_smp) p_smp[soil(j)] = cache(LC_SMP, i);
                if (j <= n && p_hk) p_hk[soil(j)] = cache(LC_HK, i);
            }
        }
        }
        // liquid and thickness of layer jwt+1, read from its owner lane's registers
        double liq_w = 0.0, mdz_w = M_TO_MM;
        if (P.aquifer_pass && jwt < nsoi(P) && jwt + 1 <= n) {
            const int jw1 = jwt + 1;
#pragma unroll
            for (int i = 0; i < S; ++i)
                if (i * G + k + 1 == jw1) { liq_w = liq[i]; mdz_w = mdz[i]; }
            liq_w = group_bcast<G>(liq_w, (jw1 - 1) % G);
            mdz_w = group_bcast<G>(mdz_w, (jw1 - 1) % G);
        }
        if (k == 0) {
            double qcharge = qc;
            if (P.aquifer_pass && jwt < nsoi(P)) {
                // compute_qcharge! with zero smp/imped (:1153-1217 via :1575-1598)
                const int jw1 = jwt + 1;
                const int64_t o = c + (int64_t)(jw1 - 1) * ld;
                double vwc = 0.0;
                if (jw1 <= n) vwc = jl_max(liq_w, 1.0e-6) / mdz_w;   // dz * denh2o
                const double wh_zwt = -RP(sucsat)[o] - zwt * M_TO_MM;
                double s1 = jl_max(vwc / RP(watsat)[o], 0.01);
                s1 = jl_min(1.0, s1);
                const double ka = 0.0 * RP(hksat)[o] * clm_pow(s1, 2.0 * RP(bsw)[o] + 3.0);
                const double *p_smpmin = RP(smpmin);
                const double smpmin = p_smpmin ? p_smpmin[c] : -1.0e8;
                const double smp1 = jl_max(smpmin, 0.0);
                const int jm = jwt > 1 ? jwt : 1;
                const double wh = smp1 - RP(z)[c + (int64_t)(joff + jm - 1) * ld] * M_TO_MM;
                double q;
                if (jwt == 0)
                    q = -ka * (wh_zwt - wh) / ((zwt + 1.0e-3) * M_TO_MM);
                else
                    q = -ka * (wh_zwt - wh) /
                        ((zwt - RP(z)[c + (int64_t)(joff + jwt - 1) * ld]) * M_TO_MM * 2.0);
                q = jl_max(-10.0 / dtime, q);
                q = jl_min(10.0 / dtime, q);
                qcharge = q;
            }
            RP(qcharge)[c] = qcharge;
            RP(nsubsteps)[c] = (double)nsubstep;
            if (double *p_nr = RP(nrejected)) p_nr[c] = (double)nrej;
        }
    }
};

}  // namespace clmsoil
