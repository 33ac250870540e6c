// adjoint.hip — hand-written reverse mode of the moisture-form Richards step.
//
// The reference gets this gradient from Enzyme differentiating
// _soilwm_moisture_form_kernel! (src/biogeophys/soil_water_movement.jl:1250-1489).
// Here one lane owns one column:
//   1. replay the primal step with RColumn<1, EXP, true>: the same arithmetic as the
//      primal kernels, so the same accept/reject decisions, recording the start
//      state and length of every accepted substep to a tape in the workspace;
//   2. walk the tape backwards. For each accepted substep, recompute its
//      hydraulic properties, fluxes and tridiagonal system from the taped state,
//      then pull the adjoint back through the state update, the tridiagonal
//      solve (lambda = A^-T xbar; abar = -lambda x[j-1], bbar = -lambda x,
//      cbar = -lambda x[j+1], rbar = lambda, as in the reference's
//      tridiagonal_solve! rule, src/infrastructure/enzyme_rules.jl:152-202),
//      the RHS/LHS assembly (:1381-1397), the interface fluxes (:1334-1378) and
//      the Clapp-Hornberger closed forms (:1281-1310).
// Rejected substeps leave the state unchanged and contribute nothing. Substep
// lengths are piecewise constant in the inputs, so their derivative is zero.
// max/min clamps pass the adjoint to the branch they selected (Julia's rule).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/clmsoil.h"
#include "../../include/clmsoil_adjoint.h"
#include "common.hpp"
#include "rcolumn.hpp"
#include "state.hpp"

namespace clmsoil {
namespace {

constexpr int NMAX = MAX_NLEVSOI;
constexpr int KMAX = CLMSOIL_ADJOINT_MAX_SUBSTEPS;

struct AdjointParams {
    double *liq_bar;
    const double *qcharge_bar;
    double *g_bsw, *g_hksat, *g_watsat, *g_scale;
    double *tape;
    int64_t tape_ld;
};

// One column's per-layer inputs (layer j at index j-1).
struct ColIn {
    int n;
    double qflx_infl;
    double mdz[NMAX], watsat[NMAX], khat[NMAX], den[NMAX], bsw[NMAX], sucsat[NMAX], root[NMAX];
};

// Reverse of one accepted substep of column `col` from state w (length dt).
// wbar: in dL/dw after the substep, out dL/dw before it. gb, gk, gw accumulate
// dL/dbsw, dL/dkhat (khat = impedance * hksat), dL/dwatsat per layer.
__device__ void substep_vjp(const ColIn &col, const RichardsParams &P,
                            const double *w, double dt, double qbar, double *wbar,
                            double *gb, double *gk, double *gw) {
    const int n = col.n;
    const bool flux_bc = P.lower_bc == CLMSOIL_BC_FLUX;
    double s2[NMAX], sr[NMAX], s1[NMAX], pw[NMAX], hk[NMAX], dhk[NMAX], q[NMAX], smp[NMAX],
        dsmp[NMAX], qo[NMAX], d1[NMAX], d2[NMAX], tz[NMAX], am[NMAX], bm[NMAX], cm[NMAX],
        x[NMAX], gam[NMAX];
    bool act2[NMAX], act1[NMAX], wpos[NMAX];
    // ---- forward recompute ----
    for (int j = 0; j < n; ++j) {
        const double vwc = jl_max(w[j], 1.0e-6) / col.mdz[j];   // dz * denh2o
        wpos[j] = w[j] > 1.0e-6;
        sr[j] = vwc / col.watsat[j];
        const double m = jl_min(sr[j], 1.0);
        s2[j] = jl_max(0.01, m);
        act2[j] = sr[j] < 1.0 && m > 0.01;
    }
    for (int j = 0; j < n; ++j) {
        const double r1 = (j < n - 1) ? 0.5 * (s2[j] + s2[j + 1]) : s2[j];
        const double m = jl_min(r1, 1.0);
        s1[j] = jl_max(0.01, m);
        act1[j] = r1 < 1.0 && m > 0.01;
        const double b = col.bsw[j];
        const double e = 2.0 * b + 3.0;
        pw[j] = clm_pow(s1[j], e);
        hk[j] = col.khat[j] * pw[j];
        dhk[j] = e * hk[j] / s1[j];
        q[j] = clm_pow(s2[j], -b);
        smp[j] = -col.sucsat[j] * q[j];
        dsmp[j] = (-b * smp[j] / s2[j]) / col.watsat[j];
    }
    for (int j = 0; j < n; ++j) {
        if (j < n - 1) {
            const double num = smp[j + 1] - smp[j], dn = col.den[j];
            const double h1 = 0.5 * dhk[j] / col.watsat[j];
            const double h2 = 0.5 * dhk[j] / col.watsat[j + 1];
            qo[j] = -hk[j] * num / dn + hk[j];
            d1[j] = (hk[j] * dsmp[j] - h1 * num) / dn + h1;
            d2[j] = (-hk[j] * dsmp[j + 1] - h2 * num) / dn + h2;
        } else if (flux_bc) {
            qo[j] = hk[j];
            d1[j] = dhk[j] / col.watsat[j];
            d2[j] = 0.0;
        } else {
            qo[j] = d1[j] = d2[j] = 0.0;
        }
    }
    for (int j = 0; j < n; ++j) {
        tz[j] = dt / col.mdz[j];
        const double qin = j == 0 ? col.qflx_infl : qo[j - 1];
        const double di0 = j == 0 ? 0.0 : d1[j - 1];
        const double di1 = j == 0 ? 0.0 : d2[j - 1];
        am[j] = j == 0 ? 0.0 : di0 * tz[j];
        bm[j] = -1.0 - (-di1 + d1[j]) * tz[j];
        cm[j] = j == n - 1 ? 0.0 : -d2[j] * tz[j];
        x[j] = -(qin - qo[j] - col.root[j]) * tz[j];   // rhs, solved in place below
    }
    {   // Thomas, as the primal (:1421-1438)
        double bet = bm[0];
        if (fabs(bet) < 1.0e-30) bet = 1.0e-30;
        x[0] = x[0] / bet;
        for (int j = 1; j < n; ++j) {
            gam[j] = cm[j - 1] / bet;
            bet = bm[j] - am[j] * gam[j];
            if (fabs(bet) < 1.0e-30) bet = copysign(1.0e-30, bet == 0.0 ? 1.0 : bet);
            x[j] = (x[j] - am[j] * x[j - 1]) / bet;
        }
        for (int j = n - 2; j >= 0; --j) x[j] = x[j] - gam[j + 1] * x[j + 1];
    }
    // ---- reverse: state update liq' = liq + x * (m_to_mm * dz), qc += qct dt/dtime ----
    double lam[NMAX], hkb[NMAX], dhkb[NMAX], smpb[NMAX], dsmpb[NMAX], qob[NMAX], d1b[NMAX],
        d2b[NMAX], s2b[NMAX];
    for (int j = 0; j < n; ++j) {
        lam[j] = wbar[j] * col.mdz[j];   // xbar
        hkb[j] = dhkb[j] = smpb[j] = dsmpb[j] = qob[j] = d1b[j] = d2b[j] = s2b[j] = 0.0;
    }
    if (flux_bc) {   // qct = hk[n] + dhkdw[n] * dwat[n] (:1471-1478)
        const double f = qbar * (dt / P.dtime);
        lam[n - 1] += f * dhk[n - 1];
        hkb[n - 1] += f;
        dhkb[n - 1] += f * x[n - 1];
    }
    {   // transposed solve: A^T has sub c[j-1], diagonal b[j], super a[j+1]
        double bet = bm[0];
        lam[0] = lam[0] / bet;
        for (int j = 1; j < n; ++j) {
            gam[j] = am[j] / bet;
            bet = bm[j] - cm[j - 1] * gam[j];
            lam[j] = (lam[j] - cm[j - 1] * lam[j - 1]) / bet;
        }
        for (int j = n - 2; j >= 0; --j) lam[j] = lam[j] - gam[j + 1] * lam[j + 1];
    }
    // coefficient adjoints -> flux adjoints
    for (int j = 0; j < n; ++j) {
        const double ab = j == 0 ? 0.0 : -lam[j] * x[j - 1];
        const double bb = -lam[j] * x[j];
        const double cb = j == n - 1 ? 0.0 : -lam[j] * x[j + 1];
        const double fnb = -lam[j] * tz[j];   // r = -fluxNet * tz
        qob[j] -= fnb;                        // fluxNet = qin - qout - root
        d1b[j] -= bb * tz[j];                 // b = -1 - (-dqidw1 + dqodw1) tz
        if (j < n - 1) d2b[j] -= cb * tz[j];  // c = -dqodw2 tz
        if (j > 0) {                          // qin, dqidw0, dqidw1 of row j are row j-1's
            qob[j - 1] += fnb;
            d1b[j - 1] += ab * tz[j];
            d2b[j - 1] += bb * tz[j];
        }
    }
    for (int j = 0; j < n; ++j) {
        if (j < n - 1) {
            const double num = smp[j + 1] - smp[j], dn = col.den[j];
            const double h1 = 0.5 * dhk[j] / col.watsat[j];
            const double h2 = 0.5 * dhk[j] / col.watsat[j + 1];
            const double t = 1.0 - num / dn;
            hkb[j] += qob[j] * t + d1b[j] * dsmp[j] / dn - d2b[j] * dsmp[j + 1] / dn;
            const double numb = -(qob[j] * hk[j] + d1b[j] * h1 + d2b[j] * h2) / dn;
            dsmpb[j] += d1b[j] * hk[j] / dn;
            dsmpb[j + 1] -= d2b[j] * hk[j] / dn;
            const double h1b = d1b[j] * t, h2b = d2b[j] * t;
            dhkb[j] += h1b * 0.5 / col.watsat[j] + h2b * 0.5 / col.watsat[j + 1];
            gw[j] -= h1b * h1 / col.watsat[j];
            gw[j + 1] -= h2b * h2 / col.watsat[j + 1];
            smpb[j + 1] += numb;
            smpb[j] -= numb;
        } else if (flux_bc) {
            hkb[j] += qob[j];
            dhkb[j] += d1b[j] / col.watsat[j];
            gw[j] -= d1b[j] * d1[j] / col.watsat[j];
        }
    }
    // Clapp-Hornberger closed forms
    for (int j = 0; j < n; ++j) {
        const double b = col.bsw[j], e = 2.0 * b + 3.0;
        // dsmpdw = (-b smp / s2) / watsat
        const double dsb = dsmpb[j] / col.watsat[j];
        gw[j] -= dsmpb[j] * dsmp[j] / col.watsat[j];
        gb[j] += dsb * (-smp[j] / s2[j]);
        smpb[j] += dsb * (-b / s2[j]);
        s2b[j] += dsb * (b * smp[j] / (s2[j] * s2[j]));
        // smp = -sucsat s2^-b
        s2b[j] += smpb[j] * col.sucsat[j] * b * q[j] / s2[j];
        gb[j] += smpb[j] * col.sucsat[j] * q[j] * clm_log(s2[j]);
        // dhkdw = e hk / s1 ; hk = khat s1^e ; e = 2b + 3
        double eb = dhkb[j] * hk[j] / s1[j];
        const double hkt = hkb[j] + dhkb[j] * e / s1[j];
        double s1b = -dhkb[j] * e * hk[j] / (s1[j] * s1[j]);
        gk[j] += hkt * pw[j];
        s1b += hkt * e * hk[j] / s1[j];
        eb += hkt * hk[j] * clm_log(s1[j]);
        gb[j] += 2.0 * eb;
        const double r1b = act1[j] ? s1b : 0.0;
        if (j < n - 1) {
            s2b[j] += 0.5 * r1b;
            s2b[j + 1] += 0.5 * r1b;
        } else {
            s2b[j] += r1b;
        }
    }
    for (int j = 0; j < n; ++j) {
        const double srb = act2[j] ? s2b[j] : 0.0;
        gw[j] -= srb * sr[j] / col.watsat[j];
        if (wpos[j]) wbar[j] += (srb / col.watsat[j]) / col.mdz[j];
    }
}

template <bool EXP>
__global__ __launch_bounds__(64) void richards_adjoint_kernel(RichardsParams P, AdjointParams A) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.nc) return;
    if (P.mask && !P.mask[c]) return;
    RColumn<1, EXP, true> col;
    col.tape = A.tape;
    col.tape_ld = A.tape_ld;
    col.kmax = KMAX;
    col.kacc = 0;
    if (!col.load(P, c, 0)) return;
    while (!col.iterate(P, 0)) {
    }
    if (col.iter >= P.max_iter) return;   // flagged by iterate
    const int K = col.kacc;
    if (K > KMAX) {
        atomicOr(P.err, (unsigned)DEV_ERR_TAPE);
        return;
    }
    const int n = col.n;
    const int64_t ld = P.ld;
    const int joff = P.nlevsno;
    ColIn in;
    in.n = n;
    in.qflx_infl = col.qflx_infl;
    for (int j = 0; j < n; ++j) {
        const int64_t o = c + (int64_t)j * ld;
        in.mdz[j] = col.mdz[j];
        in.watsat[j] = col.watsat[j];
        in.khat[j] = col.khat[j];
        in.den[j] = col.den[j];
        in.bsw[j] = P.bsw[o];
        in.sucsat[j] = P.sucsat[o];
        in.root[j] = P.qflx_rootsoi[o];
    }
    double wbar[NMAX], gb[NMAX], gk[NMAX], gw[NMAX], w[NMAX];
    for (int j = 0; j < n; ++j) {
        wbar[j] = A.liq_bar[c + (int64_t)(joff + j) * ld];
        gb[j] = gk[j] = gw[j] = 0.0;
    }
    const double qbar = A.qcharge_bar[c];
    for (int k = K - 1; k >= 0; --k) {
        const int64_t row = (int64_t)k * (NMAX + 1);
        for (int j = 0; j < n; ++j) w[j] = A.tape[(row + j) * A.tape_ld + c];
        const double dt = A.tape[(row + NMAX) * A.tape_ld + c];
        substep_vjp(in, P, w, dt, qbar, wbar, gb, gk, gw);
    }
    double sb = 0.0, sk = 0.0, sw = 0.0;
    for (int j = 0; j < n; ++j) {
        A.liq_bar[c + (int64_t)(joff + j) * ld] = wbar[j];
        const int64_t o = c + (int64_t)j * ld;
        const double hks = P.hksat[o];
        const double gh = gk[j] * (in.khat[j] / hks);   // khat = impedance * hksat
        if (A.g_bsw) A.g_bsw[o] += gb[j];
        if (A.g_hksat) A.g_hksat[o] += gh;
        if (A.g_watsat) A.g_watsat[o] += gw[j];
        sb += gb[j] * in.bsw[j];
        sk += gk[j] * in.khat[j];
        sw += gw[j] * in.watsat[j];
    }
    if (A.g_scale) {
        A.g_scale[c] += sb;
        A.g_scale[ld + c] += sk;
        A.g_scale[2 * ld + c] += sw;
    }
}

}  // namespace

RichardsParams richards_params(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                               double dtime, const clmsoil_swm_config *cfg, bool aquifer,
                               unsigned *err);

}  // namespace clmsoil

using namespace clmsoil;

extern "C" {

int64_t clmsoil_adjoint_workspace_bytes(int64_t ld) {
    if (ld < 0) return -1;
    return (int64_t)KMAX * (NMAX + 1) * ld * (int64_t)sizeof(double);
}

int clmsoil_soilwm_moisture_form_adjoint(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                         double dtime, const clmsoil_swm_config *cfg,
                                         double *liq_bar, const double *qcharge_bar,
                                         double *grad_bsw, double *grad_hksat,
                                         double *grad_watsat, double *grad_scale,
                                         void *workspace, int64_t workspace_bytes,
                                         void *stream) {
    if (!a || !cfg) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (cfg->soilwater_movement_method != CLMSOIL_MOISTURE_FORM) return CLMSOIL_ERR_METHOD;
    if (cfg->upper_boundary_condition != CLMSOIL_BC_FLUX) return CLMSOIL_ERR_BC;
    const int bc = cfg->lower_boundary_condition;
    if (bc != CLMSOIL_BC_FLUX && bc != CLMSOIL_BC_ZERO_FLUX) return CLMSOIL_ERR_BC;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (!(cfg->dtmin > 0.0) || !(cfg->verySmall > 0.0)) return CLMSOIL_ERR_ARG;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!a->nbedrock || !a->h2osoi_liq || !a->z || !a->dz || !a->qflx_rootsoi || !a->watsat ||
        !a->hksat || !a->bsw || !a->sucsat || !a->icefrac || !a->qflx_infl || !liq_bar ||
        !qcharge_bar)
        return CLMSOIL_ERR_NULL;
    if (!workspace || workspace_bytes < clmsoil_adjoint_workspace_bytes(a->ld))
        return CLMSOIL_ERR_WORKSPACE;
    DeviceState *s = nullptr;
    if ((st = current_device_state(&s))) return st;
    RichardsParams P = richards_params(a, lev, dtime, cfg, false, s->err);
    P.h2osoi_liq = const_cast<double *>(a->h2osoi_liq);   // read only in the replay
    AdjointParams A{liq_bar, qcharge_bar, grad_bsw, grad_hksat, grad_watsat, grad_scale,
                    static_cast<double *>(workspace), a->ld};
    const int64_t grid = (a->nc + 63) / 64;
    if (P.expensive)
        hipLaunchKernelGGL(richards_adjoint_kernel<true>, dim3((unsigned)grid), dim3(64), 0,
                           (hipStream_t)stream, P, A);
    else
        hipLaunchKernelGGL(richards_adjoint_kernel<false>, dim3((unsigned)grid), dim3(64), 0,
                           (hipStream_t)stream, P, A);
    return hipGetLastError() == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}

}  // extern "C"
