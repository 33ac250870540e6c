// zd09.hip — the Zeng-Decker 2009 soil water step (gfx950), one lane per column.
//
// Follows soilwater_zengdecker2009! (src/biogeophys/soil_water_movement.jl:1616-1764).
// The reference runs it as ten kernels over nc x (nlevsoi+1) scratch matrices
// (zmm, dzmm, vol_eq, zq, hk, dhkdw, smp, dsmpdw, qin/qout and their four
// derivatives, amx/bmx/cmx/rmx, dwat2) followed by tridiagonal_multi!
// (src/infrastructure/tridiagonal.jl:86-127). Here one lane owns one column and
// walks down its layers once:
//
//   node k      layer-k quantities (mm geometry, ice fraction, equilibrium water
//               content and suction, suction and its derivative)       :97-262
//   interface   conductivity between nodes k-1 and k, the flux across it and
//               its two derivatives; the flux leaving row k-1 and the one
//               entering row k are the same expression, so it is formed once
//                                                                    :430-539
//   row k-1     assembled from the two interfaces around it and eliminated at
//               once (Thomas forward sweep, with the reference's 1e-30 guard)
//
// then assembles the bottom row and the aquifer row, back-substitutes, updates
// the liquid water, and computes the recharge (:561-604) and the deficit
// (:621-633). Nothing but the inputs and outputs touches HBM: the sweep's cp/dp
// rows go to LDS, one 8-byte word per lane and row, so each access of a wave is
// a contiguous, conflict-free 512-byte line (2 x nlevsoi x 8 bytes per column;
// the kernel's ~170 VGPRs leave no room for them in registers). Every
// expression keeps the reference's operand order, so results match the
// reference's Float64 arithmetic up to the last-bit differences of pow().
#include "common.hpp"
#include "zd09.hpp"

namespace clmsoil {
namespace {

// Plain (cached) column-major load: several inputs are read twice.
__device__ __forceinline__ double at(const double *p, int64_t c, int j, int64_t ld) {
    return p[c + (int64_t)(j - 1) * ld];
}

struct Node {
    double zmm, dzmm, ws, b, hks, icefrac, vwc, zq, smp, dsmpdw;
};

// Layer k of column c; zimm_above is the top interface depth in mm.
__device__ __forceinline__ Node zd_node(const ZdParams &P, int64_t c, int k, double zimm_above,
                                        double zimm_k, double zwtmm, double smpmin) {
    const int64_t ld = P.ld;
    const int jk = P.nlevsno + k;
    Node n;
    // _soilwm_mm_icefrac_kernel! (:97-121)
    const double dz = at(P.dz, c, jk, ld);
    n.zmm = at(P.z, c, jk, ld) * 1.0e3;
    n.dzmm = dz * 1.0e3;
    n.ws = at(P.watsat, c, k, ld);
    n.b = at(P.bsw, c, k, ld);
    n.hks = at(P.hksat, c, k, ld);
    const double suc = at(P.sucsat, c, k, ld);
    const double vol_ice = jl_min(n.ws, at(P.h2osoi_ice, c, jk, ld) / (dz * DENICE));
    n.icefrac = jl_min(1.0, vol_ice / n.ws);
    n.vwc = jl_max(at(P.h2osoi_liq, c, jk, ld), 1.0e-6) / (dz * DENH2O);
    P.icefrac[c + (int64_t)(k - 1) * ld] = n.icefrac;
    // _soilwm_voleq_zq_kernel! (:171-201)
    const double onem1b = 1.0 - 1.0 / n.b;
    double vol_eq;
    if (zwtmm <= zimm_above) {
        vol_eq = n.ws;
    } else if (zwtmm < zimm_k && zwtmm > zimm_above) {
        const double tempi = 1.0;
        const double temp0 = pow((suc + zwtmm - zimm_above) / suc, onem1b);
        const double voleq1 = -suc * n.ws / onem1b / (zwtmm - zimm_above) * (tempi - temp0);
        double v = (voleq1 * (zwtmm - zimm_above) + n.ws * (zimm_k - zwtmm)) /
                   (zimm_k - zimm_above);
        v = jl_min(n.ws, v);
        vol_eq = jl_max(v, 0.0);
    } else {
        const double tempi = pow((suc + zwtmm - zimm_k) / suc, onem1b);
        const double temp0 = pow((suc + zwtmm - zimm_above) / suc, onem1b);
        double v = -suc * n.ws / onem1b / (zimm_k - zimm_above) * (tempi - temp0);
        v = jl_max(v, 0.0);
        vol_eq = jl_min(n.ws, v);
    }
    n.zq = jl_max(smpmin, -suc * pow(jl_max(vol_eq / n.ws, 0.01), -n.b));
    // node half of _soilwm_hk_smp_kernel! (:223-262)
    double s_node = jl_max(n.vwc / n.ws, 0.01);
    s_node = jl_min(1.0, s_node);
    double smpv = -suc * pow(s_node, -n.b);
    smpv = jl_max(smpmin, smpv);
    n.smp = smpv;
    n.dsmpdw = -n.b * smpv / n.vwc;
    P.smp_l[c + (int64_t)(k - 1) * ld] = smpv;
    return n;
}

// Interface half of _soilwm_hk_smp_kernel!: layer `lo` and jp1 = min(j+1, nlevsoi).
struct Iface {
    double imp, hk, dhkdw;
};
__device__ __forceinline__ Iface zd_iface(const Node &lo, const Node &hi, double e_ice) {
    Iface f;
    double s1 = 0.5 * (lo.vwc + hi.vwc) / (0.5 * (lo.ws + hi.ws));
    s1 = jl_min(1.0, s1);
    const double s2 = lo.hks * pow(s1, 2.0 * lo.b + 2.0);
    f.imp = pow(10.0, -e_ice * (0.5 * (lo.icefrac + hi.icefrac)));
    f.hk = f.imp * s1 * s2;
    f.dhkdw = f.imp * (2.0 * lo.b + 3.0) * s2 * (1.0 / (lo.ws + hi.ws));
    return f;
}

// Flux from node lo to node hi and its derivatives (qout/dqodw1/dqodw2 of the
// upper row = qin/dqidw0/dqidw1 of the lower one, :449-488).
struct Flux {
    double q, d1, d2;
};
__device__ __forceinline__ Flux zd_flux(const Node &lo, const Node &hi, double hk, double dhkdw) {
    Flux f;
    const double den = hi.zmm - lo.zmm;
    const double dzq = hi.zq - lo.zq;
    const double num = (hi.smp - lo.smp) - dzq;
    f.q = -hk * num / den;
    f.d1 = -(-hk * lo.dsmpdw + num * dhkdw) / den;
    f.d2 = -(hk * hi.dsmpdw + num * dhkdw) / den;
    return f;
}

__global__ void __launch_bounds__(256, 2) zd09_kernel(ZdParams P) {
    extern __shared__ double zd_lds[];   // cp rows, then dp rows: [2 * nlevsoi][blockDim.x]
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.nc) return;
    if (P.mask && !P.mask[c]) return;
    const int64_t ld = P.ld;
    const int n = P.nlevsoi, joff = P.nlevsno, joff_zi = P.nlevsno + 1;
    const double dtime = P.dtime, e_ice = P.e_ice;
    const double zwt = P.zwt[c], zwtmm = zwt * 1.0e3, smpmin = P.smpmin[c];

    // _soilwm_jwt_vwczwt_kernel! (:289-324): layer right above the water table
    int jw = n;
    for (int j = 1; j <= n; ++j)
        if (zwt <= at(P.zi, c, joff_zi + j, ld)) { jw = j - 1; break; }
    const int jm = jw > 1 ? jw : 1;   // max(1, jwt) of the recharge (:580-581)

    const int bs = blockDim.x;
    double *cp = zd_lds + threadIdx.x;            // cp of row r at cp[(r - 1) * bs], r = 1..n
    double *dp = zd_lds + threadIdx.x + n * bs;   // dp likewise
    double zimm_above = 0.0;
    double zimm_k = at(P.zi, c, joff_zi + 1, ld) * 1.0e3;
    Node lo = zd_node(P, c, 1, zimm_above, zimm_k, zwtmm, smpmin);
    zimm_above = zimm_k;
    double cap_smp = lo.smp, cap_zq = lo.zq, cap_imp = 1.0;
    double q_up = 0.0, d1_up = 0.0, d2_up = 0.0;   // interface above the current row
    double cp_prev = 0.0, dp_prev = 0.0;
    const double tiny = 1e-30;

    for (int k = 2; k <= n; ++k) {
        zimm_k = at(P.zi, c, joff_zi + k, ld) * 1.0e3;
        const Node hi = zd_node(P, c, k, zimm_above, zimm_k, zwtmm, smpmin);
        zimm_above = zimm_k;
        if (k == jm) { cap_smp = hi.smp; cap_zq = hi.zq; }
        const Iface fi = zd_iface(lo, hi, e_ice);
        P.hk_l[c + (int64_t)(k - 2) * ld] = fi.hk;
        if (k - 1 == jw + 1) cap_imp = fi.imp;
        const Flux fx = zd_flux(lo, hi, fi.hk, fi.dhkdw);
        const int r = k - 1;
        const double root = at(P.qflx_rootsoi, c, r, ld);
        if (r == 1) {
            // node j = 1 (:445-456); sdamp = 0 (:1700)
            const double rmx = P.qflx_infl[c] - fx.q - root;
            const double bmx = lo.dzmm * (0.0 + 1.0 / dtime) + fx.d1;
            double bet = bmx;
            if (fabs(bet) < tiny) bet = tiny;
            cp_prev = fx.d2 / bet;
            dp_prev = rmx / bet;
        } else {
            // interior node (:459-477)
            const double rmx = q_up - fx.q - root;
            const double amx = -d1_up;
            const double bmx = lo.dzmm / dtime - d2_up + fx.d1;
            double denom = bmx - amx * cp_prev;
            if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0.0 ? 1.0 : denom);
            cp_prev = fx.d2 / denom;
            dp_prev = (rmx - amx * dp_prev) / denom;
        }
        cp[(r - 1) * bs] = cp_prev;
        dp[(r - 1) * bs] = dp_prev;
        q_up = fx.q;
        d1_up = fx.d1;
        d2_up = fx.d2;
        lo = hi;
    }
    // interface of the last layer with itself (jp1 = nlevsoi)
    const Iface fn = zd_iface(lo, lo, e_ice);
    P.hk_l[c + (int64_t)(n - 1) * ld] = fn.hk;
    if (n == jw + 1) cap_imp = fn.imp;

    // bottom node and aquifer layer (:480-536, :385-396)
    double rmx_n, amx_n, bmx_n, cmx_n, rmx_a, amx_a, bmx_a, dzmm_a;
    const double root_n = at(P.qflx_rootsoi, c, n, ld);
    if (n > jw) {
        rmx_n = q_up - 0.0 - root_n;
        amx_n = -d1_up;
        bmx_n = lo.dzmm / dtime - d2_up + 0.0;
        cmx_n = 0.0;
        dzmm_a = lo.dzmm;
        rmx_a = 0.0;
        amx_a = 0.0;
        bmx_a = dzmm_a / dtime;
    } else {
        // water content at the water table (:300-321); only this branch reads it
        double vwc_zwt = lo.ws;
        const double suc_n = at(P.sucsat, c, n, ld);
        if (at(P.t_soisno, c, joff + jw + 1, ld) < TFRZ) {
            double v = lo.vwc;
            for (int j = n; j <= P.nlevgrnd; ++j) {
                if (zwt <= at(P.zi, c, joff_zi + j, ld)) {
                    const double t = at(P.t_soisno, c, joff + j, ld);
                    double smp1 = HFUS * (TFRZ - t) / (GRAV * t) * 1000.0;
                    smp1 = jl_max(suc_n, smp1);
                    v = lo.ws * pow(smp1 / suc_n, -1.0 / lo.b);
                    v = jl_min(v, 0.5 * (lo.ws + at(P.h2osoi_vol, c, n, ld)));
                    break;
                }
            }
            vwc_zwt = v;
        }
        // aquifer-layer equilibrium suction (:346-364)
        const double onem1b = 1.0 - 1.0 / lo.b;
        const double tempi = 1.0;
        const double temp0 = pow((suc_n + zwtmm - zimm_above) / suc_n, onem1b);
        double va = -suc_n * lo.ws / onem1b / (zwtmm - zimm_above) * (tempi - temp0);
        va = jl_max(va, 0.0);
        va = jl_min(lo.ws, va);
        const double zq_a = jl_max(smpmin, -suc_n * pow(jl_max(va / lo.ws, 0.01), -lo.b));
        const double zmm_a = 0.5 * (1.0e3 * zwt + lo.zmm);
        dzmm_a = 1.0e3 * zwt - lo.zmm;
        // bottom node (:505-536)
        double s_node = jl_max(0.5 * ((vwc_zwt + lo.vwc) / lo.ws), 0.01);
        s_node = jl_min(1.0, s_node);
        double smp1 = -suc_n * pow(s_node, -lo.b);
        smp1 = jl_max(smpmin, smp1);
        const double dsmpdw1 = -lo.b * smp1 / (s_node * lo.ws);
        const double den = zmm_a - lo.zmm;
        const double dzq = zq_a - lo.zq;
        const double num = (smp1 - lo.smp) - dzq;
        const double qout = -fn.hk * num / den;
        const double dqodw1 = -(-fn.hk * lo.dsmpdw + num * fn.dhkdw) / den;
        const double dqodw2 = -(fn.hk * dsmpdw1 + num * fn.dhkdw) / den;
        rmx_n = q_up - qout - root_n;
        amx_n = -d1_up;
        bmx_n = lo.dzmm / dtime - d2_up + dqodw1;
        cmx_n = dqodw2;
        rmx_a = qout - 0.0;
        amx_a = -dqodw1;
        bmx_a = dzmm_a / dtime - dqodw2 + 0.0;
    }
    {
        double denom = bmx_n - amx_n * cp_prev;
        if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0.0 ? 1.0 : denom);
        cp_prev = cmx_n / denom;
        dp_prev = (rmx_n - amx_n * dp_prev) / denom;
        cp[(n - 1) * bs] = cp_prev;
        dp[(n - 1) * bs] = dp_prev;
        denom = bmx_a - amx_a * cp_prev;   // cmx of the aquifer row is 0
        if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0.0 ? 1.0 : denom);
        dp_prev = (rmx_a - amx_a * dp_prev) / denom;
    }
    // back substitution; cp row j then holds dwat2(j)
    const double dw_a = dp_prev;
    double u = dw_a;
    for (int j = n; j >= 1; --j) {
        u = dp[(j - 1) * bs] - cp[(j - 1) * bs] * u;
        cp[(j - 1) * bs] = u;
    }
    // _soilwm_renew_qcharge_kernel! (:561-604) and _soilwm_deficit_kernel! (:621-633)
    double acc = 0.0;
    for (int j = 1; j <= n; ++j) {
        double *pl = &P.h2osoi_liq[c + (int64_t)(joff + j - 1) * ld];
        const double l = *pl + cp[(j - 1) * bs] * (at(P.dz, c, joff + j, ld) * 1.0e3);
        *pl = l;
        acc = acc + (l >= 0.0 ? 0.0 : -l);
    }
    double qc;
    if (jw < n) {
        const int jp = jw + 1;
        const double ws = at(P.watsat, c, jp, ld);
        const double s_node = jl_max(at(P.h2osoi_vol, c, jp, ld) / ws, 0.01);
        const double s1 = jl_min(1.0, s_node);
        const double ka = cap_imp * at(P.hksat, c, jp, ld) *
                          pow(s1, 2.0 * at(P.bsw, c, jp, ld) + 3.0);
        const double smp1 = jl_max(smpmin, cap_smp);
        const double wh = smp1 - cap_zq;
        const double wh_zwt = 0.0;
        if (jw == 0)
            qc = -ka * (wh_zwt - wh) / ((zwt + 1.0e-3) * 1000.0);
        else
            qc = -ka * (wh_zwt - wh) / ((zwt - at(P.z, c, joff + jw, ld)) * 1000.0 * 2.0);
        qc = jl_max(-10.0 / dtime, qc);
        qc = jl_min(10.0 / dtime, qc);
    } else {
        qc = dw_a * dzmm_a / dtime;
    }
    P.qcharge[c] = qc;
    P.qflx_deficit[c] = acc;
}

}  // namespace

hipError_t launch_zengdecker2009(const ZdParams &P, int block, hipStream_t stream) {
    if (block != 64 && block != 128 && block != 256) block = 128;
    const size_t lds = (size_t)2 * P.nlevsoi * block * sizeof(double);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)zd09_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int64_t grid = (P.nc + block - 1) / block;
    hipLaunchKernelGGL(zd09_kernel, dim3((unsigned)grid), dim3(block), lds, stream, P);
    return hipGetLastError();
}

}  // namespace clmsoil
