// ch4.hip — methane and oxygen transport in the soil column (gfx950), one lane
// per column.
//
// ch4_tran! (src/biogeochem/methane.jl:2005-2092) fills about twenty
// [nc, nlevsoi] scratch arrays over meth_khcc! and one column kernel. Here a lane
// walks its column three times:
//   1. the O2/CH4 competition, the scaled process rates, the ebullition and
//      aerenchyma totals, level by level (each sum in the reference's level order);
//   (then, O2 before CH4, as the two are independent after step 1:)
//   2. per species, the levels top to bottom with a one-level lookahead: the
//      level's Henry partition, epsilon, source, diffusivity and relative
//      concentration are formed once, the interface conductances and the
//      Crank-Nicolson row from the three levels around it, and the Thomas forward
//      elimination consumes the row at once. Multipliers go to a lane-private LDS
//      column ([level][lane]); O2's right-hand sides are parked in its concentration
//      slots and its epsilon in a second LDS column, CH4's right-hand sides in that
//      column, so that CH4's old concentrations stay in memory for its balance;
//   3. per species, the back substitution bottom up, with the CH4 clip or the O2
//      bounds and the absolute concentration of each level as it is solved; then
//      the CH4 balance, whose concentration terms were kept in LDS.
// The arithmetic is the reference's operation by operation; exp, log and pow are
// the device library's.
#include "ch4.hpp"
#include "common.hpp"

namespace clmsoil {
namespace {

constexpr int CH_BLOCK = 64;
constexpr double CH_TFRZ = 273.15, CH_DENH2O = 1000.0, CH_DENICE = 917.0;
constexpr double CH_SN = 1.0e-12;                                  // smallnumber (:2083)
__constant__ double C_H_INV[2] = {1600.0, 1500.0};                 // varcon.jl:219
__constant__ double KH_THETA[2] = {714.29, 769.23};                // varcon.jl:220
__constant__ double D_CON_W[2][3] = {{0.9798, 0.02986, 0.0004381}, {1.172, 0.03443, 0.0005048}};
__constant__ double D_CON_G[2][2] = {{0.1875, 0.0013}, {0.1759, 0.00117}};

// Julia's pow_body(x, n), n >= 0 (Base.Math): compensated squaring, muladd fused
__device__ double jl_pow_int(double x, int64_t n) {
    if (n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            const double err = __builtin_fma(y, xnlo, x * ynlo);
            const double yy = x * y;
            ynlo = __builtin_fma(x, y, -yy);
            y = yy;
            ynlo += err;
        }
        const double err = x * 2.0 * xnlo;
        const double xx = x * x;
        xnlo = __builtin_fma(x, x, -xx);
        x = xx;
        xnlo += err;
        n >>= 1;
    }
    const double err = __builtin_fma(y, xnlo, x * ynlo);
    return (__builtin_isfinite(x) && __builtin_isfinite(err)) ? __builtin_fma(x, y, err) : x * y;
}

// Julia's Float64 ^ Float64 for the non-negative exponents of this step: an
// integer-valued exponent takes the integer power, others the library pow
__device__ __forceinline__ double jl_pow(double x, double y) {
    if (x == 1.0) return 1.0;
    if (y == trunc(y) && y >= 0.0 && y < 0x1.8p62) return jl_pow_int(x, (int64_t)y);
    return pow(x, y);
}

struct Ch4Col {
    const clmsoil_ch4_tran_args &a;
    const clmsoil_ch4_tran_params &p;
    int64_t c;
    int jwt;
    __device__ double at(const double *x, int j) const { return x[c + (int64_t)(j - 1) * a.ld]; }
    __device__ double *ptr(double *x, int j) const { return x + c + (int64_t)(j - 1) * a.ld; }

    // k_h_cc (_meth_khcc_kernel!, :101-117) at temperature t; logth = log(kh_theta)
    __device__ double khcc(double t, int s, double logth) const {
        const double k_h_inv = exp(-C_H_INV[s] * (1.0 / t - 1.0 / 298.0) + logth);
        return t / k_h_inv * 0.0821;
    }
};

// One level of a species' column in the sweep: what its own row and its
// neighbours' rows read.
struct Ch4Lev {
    double dz, D, k, e, src, cw;
};

// The inputs of one level, loaded a step ahead of their use.
struct Ch4In {                 // competition
    double conc_o2, conc_ch4, o2_aere, prod, o2_decomp, o2_oxid, ch4_oxid, ch4_aere, ch4_ebul, dz;
};
struct Ch4Raw {                // a species' sweep; r1..r4 its source terms
    double dz, t, ws, vol, liq, ice, bsw, corg, r1, r2, r3, r4, conc;
};
struct Ch4Fin {                // the clip / bounds pass
    double t, ws, vol, liq, ice, dz, conc;
};

__global__ void __launch_bounds__(CH_BLOCK) ch4_tran_kernel(clmsoil_ch4_tran_args a,
                                                            clmsoil_ch4_tran_params p, double dtime,
                                                            unsigned *err) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * CH_BLOCK + lane;
    if (c >= a.nc || (a.mask && !a.mask[c])) return;     // no barrier in this kernel
    const int nl = p.nlevsoi;
    const int64_t jw = a.jwt[c];
    if (jw < 0 || jw > nl) {
        atomicOr(err, (unsigned)DEV_ERR_JWT);
        return;
    }
    const Ch4Col L{a, p, c, (int)jw};
    const int jwt = (int)jw;
    double *s_cp = lds + lane;                            // [jj - 1] at stride CH_BLOCK
    // [j - 1]: O2's epsilon during its pass, then CH4's parked right-hand side and,
    // once that is used, its balance term (new - old) * dz
    double *s_aux = lds + (int64_t)(nl + 1) * CH_BLOCK + lane;
    const int64_t g = a.col_gridcell[c] - 1;
    const double catm0 = a.c_atm[g], catm1 = a.c_atm[g + a.ldg];
    const double logth0 = log(KH_THETA[0]), logth1 = log(KH_THETA[1]);

    // --- 1. competition for O2 and CH4 (:1671-1724), totals (:1735-1774) ---
    double ebul_total = 0.0, surf_aere = 0.0;
    auto in_at = [&](int j) {
        return Ch4In{L.at(a.conc_o2, j),         L.at(a.conc_ch4, j),       L.at(a.o2_aere_depth, j),
                     L.at(a.ch4_prod_depth, j),  L.at(a.o2_decomp_depth, j), L.at(a.o2_oxid_depth, j),
                     L.at(a.ch4_oxid_depth, j),  L.at(a.ch4_aere_depth, j), L.at(a.ch4_ebul_depth, j),
                     L.at(a.dz, j)};
    };
    Ch4In qa = in_at(1);
    for (int j = 1; j <= nl; ++j) {
        const Ch4In r = qa;
        if (j < nl) qa = in_at(j + 1);
        const double conc_o2 = r.conc_o2, conc_ch4 = r.conc_ch4, o2_aere = r.o2_aere, prod = r.prod;
        double o2_decomp = r.o2_decomp, o2_oxid = r.o2_oxid, ch4_oxid = r.ch4_oxid, ch4_aere = r.ch4_aere;
        double ch4_ebul = r.ch4_ebul;
        const double dz = r.dz;
        const double o2demand = o2_decomp + o2_oxid;
        double o2s, ch4s;
        if (o2demand > 0.0)
            o2s = conc_o2 / dtime + o2_aere > o2demand ? 1.0 : (conc_o2 / dtime + o2_aere) / o2demand;
        else
            o2s = 1.0;
        const double ch4demand = ch4_oxid + ch4_aere + ch4_ebul;
        ch4s = ch4demand > 0.0 ? jl_min((conc_ch4 / dtime + prod) / ch4demand, 1.0) : 1.0;
        if (o2s < 1.0 || ch4s < 1.0) {
            if (ch4s <= o2s) {
                if (o2s < 1.0) {
                    const double o2demand2 = o2_decomp;
                    o2s = o2demand2 > 0.0
                              ? jl_min((conc_o2 / dtime + o2_aere - ch4s * o2_oxid) / o2demand2, 1.0)
                              : 1.0;
                }
                ch4_oxid *= ch4s;
                o2_oxid *= ch4s;
            } else {
                if (ch4s < 1.0) {
                    const double ch4demand2 = ch4_aere + ch4_ebul;
                    ch4s = ch4demand2 > 0.0
                               ? jl_min((conc_ch4 / dtime + prod - o2s * ch4_oxid) / ch4demand2, 1.0)
                               : 1.0;
                }
                ch4_oxid *= o2s;
                o2_oxid *= o2s;
            }
        }
        ch4_aere *= ch4s;
        ch4_ebul *= ch4s;
        o2_decomp *= o2s;
        ebul_total += ch4_ebul * dz;
        if (!p.use_aereoxid_prog) {                       // (:1758-1761)
            ch4_oxid += p.aereoxid * ch4_aere;
            ch4_aere -= p.aereoxid * ch4_aere;
        }
        surf_aere += ch4_aere * dz;
        __builtin_nontemporal_store(o2s, L.ptr(a.o2stress, j));      // final: not read again
        __builtin_nontemporal_store(ch4s, L.ptr(a.ch4stress, j));
        *L.ptr(a.o2_decomp_depth, j) = o2_decomp;
        *L.ptr(a.o2_oxid_depth, j) = o2_oxid;
        *L.ptr(a.ch4_oxid_depth, j) = ch4_oxid;
        *L.ptr(a.ch4_aere_depth, j) = ch4_aere;
        *L.ptr(a.ch4_ebul_depth, j) = ch4_ebul;
    }
    a.ch4_ebul_total[c] = ebul_total;
    a.ch4_surf_aere[c] = surf_aere;

    // the ground conductance floor (:1800-1802), which the reference applies before
    // CH4 and O2 then sees too
    double gcond = a.grnd_ch4_cond[c];
    if (gcond < CH_SN) gcond = CH_SN;
    double spec0 = 0.0, dm1_first = 0.0, cw1_old = 0.0, surf_diff = 0.0, surf_ebul = 0.0;
    // O2 first (the species are independent from here on): CH4's old concentrations
    // stay in memory for its balance while its right-hand sides are parked in LDS
    for (int si = 0; si < 2; ++si) {
        const int s = 1 - si;
        double *conc = s == 0 ? a.conc_ch4 : a.conc_o2;
        const double catm = s == 0 ? catm0 : catm1, logth = s == 0 ? logth0 : logth1;
        // --- 2. snow/pond resistance and ground conductance (:1800-1815) ---
        double pondres = 0.0;
        const double frac = a.frac_h2osfc[c];
        if (!p.lake && p.sat == 1 && frac > 0.0) {
            const double th = a.t_h2osfc[c];
            if (th >= CH_TFRZ) {
                const double tsc = th - CH_TFRZ;
                const double ponddiff = (D_CON_W[s][0] + D_CON_W[s][1] * tsc + D_CON_W[s][2] * (tsc * tsc)) *
                                        1.0e-9 * p.scale_factor_liqdiff;
                const double pondz = a.h2osfc[c] / 1000.0 / frac;
                pondres = pondz / ponddiff;
            } else if (a.h2osfc[c] / frac > p.capthick) {
                pondres = 1.0 / CH_SN;
            }
        }
        const double spec = 1.0 / (1.0 / gcond + 0.0 + pondres);
        if (s == 0) spec0 = spec;

        // the level values of level j (:1741-1793 for epsilon, source and the
        // relative concentration; :1818-1840 for the diffusivity)
        auto raw = [&](int j) {
            Ch4Raw r{};
            r.dz = L.at(a.dz, j);
            r.t = L.at(a.t_soisno, j);
            r.ws = L.at(a.watsat, j);
            r.vol = L.at(a.h2osoi_vol, j);
            if (p.ch4frzout || j > jwt) {
                r.liq = L.at(a.h2osoi_liq, j);
                r.ice = L.at(a.h2osoi_ice, j);
            }
            if (j <= jwt) {
                r.bsw = L.at(a.bsw, j);
                r.corg = L.at(a.cellorg, j);
            }
            if (s == 0) {
                r.r1 = L.at(a.ch4_prod_depth, j);
                r.r2 = L.at(a.ch4_oxid_depth, j);
                r.r3 = L.at(a.ch4_aere_depth, j);
                r.r4 = L.at(a.ch4_ebul_depth, j);
            } else {
                r.r1 = L.at(a.o2_oxid_depth, j);
                r.r2 = L.at(a.o2_decomp_depth, j);
                r.r3 = L.at(a.o2_aere_depth, j);
            }
            r.conc = L.at(conc, j);
            return r;
        };
        auto level = [&](const Ch4Raw &r, int j) {
            Ch4Lev v;
            v.dz = r.dz;
            const double t = r.t;
            v.k = L.khcc(t, s, logth);
            const double vmin = jl_min(r.ws, r.vol);
            const double lf = !p.ch4frzout ? 1.0
                                           : jl_max(0.05, (r.liq / CH_DENH2O + CH_SN) /
                                                              (r.liq / CH_DENH2O + r.ice / CH_DENICE + CH_SN));
            v.e = j <= jwt ? r.ws - (1.0 - v.k) * vmin * lf : r.ws * lf;
            if (s == 0) {
                v.src = r.r1 - r.r2 - r.r3 - r.r4;
                if (j == jwt) v.src += ebul_total / v.dz;
            } else {
                v.src = -r.r1 - r.r2 + r.r3;
            }
            v.cw = r.conc / v.e;
            const double tsc = t - CH_TFRZ;
            double d;
            if (j <= jwt) {
                const double ws = r.ws;
                const double f_a = 1.0 - vmin / ws;
                const double e = ws - vmin;
                const double om_frac = p.organic_max > 0.0 ? jl_min(p.om_frac_sf * r.corg / p.organic_max, 1.0) : 1.0;
                d = (D_CON_G[s][0] + D_CON_G[s][1] * tsc) * 1.0e-4 *
                    (om_frac * jl_pow(f_a, 10.0 / 3.0) / (ws * ws) + (1.0 - om_frac) * (e * e) * jl_pow(f_a, 3.0 / r.bsw)) *
                    p.scale_factor_gasdiff;
            } else {
                const double e = r.ws;
                d = jl_pow(e, p.satpow) * (D_CON_W[s][0] + D_CON_W[s][1] * tsc + D_CON_W[s][2] * (tsc * tsc)) *
                    1.0e-9 * p.scale_factor_liqdiff;
                if (t <= CH_TFRZ)
                    d *= (r.liq / CH_DENH2O + CH_SN) / (r.liq / CH_DENH2O + r.ice / CH_DENICE + CH_SN);
            }
            v.D = jl_max(d, CH_SN);
            return v;
        };

        // ground node (jj = 1): k_h_cc at t_grnd, the atmosphere's concentration
        Ch4Lev prev{0.0, 0.0, L.khcc(a.t_grnd[c], s, logth), 0.0, 0.0, catm}, cur{}, nxt{};
        // Thomas forward sweep from the ground row a = 0, b = 1, c = 0, r = c_atm (:1940-1951)
        double cp_prev = 0.0 / 1.0, dp_prev = catm / 1.0;
        s_cp[0] = cp_prev;
        // iteration j forms level j + 1 and then the row of level j
        Ch4Raw qr = raw(1);                                // inputs of the level after next
        for (int j = 0; j <= nl; ++j) {
            if (j < nl) {
                const Ch4Raw r = qr;
                if (j + 2 <= nl) qr = raw(j + 2);
                nxt = level(r, j + 1);
            }
            if (j == 0) {
                cur = nxt;
                continue;
            }
            // dm1_zm1 / dp1_zp1 (:1843-1871)
            double dm, dp = 0.0;
            if (j == 1 && j != jwt && j != jwt + 1) {
                dm = 1.0 / (1.0 / spec + cur.dz / (cur.D * 2.0));
                dp = j < nl ? 2.0 / (cur.dz / cur.D + nxt.dz / nxt.D) : 0.0;
            } else if (j == 1 && j == jwt) {
                dm = 1.0 / (1.0 / spec + cur.dz / (cur.D * 2.0));
                dp = j < nl ? 2.0 / (cur.dz * cur.k / cur.D + nxt.dz / nxt.D) : 0.0;
            } else if (j == 1) {
                dm = 1.0 / (prev.k / spec + cur.dz / (cur.D * 2.0));
                dp = j < nl ? 2.0 / (cur.dz / cur.D + nxt.dz / nxt.D) : 0.0;
            } else if (j < nl && j != jwt && j != jwt + 1) {
                dm = 2.0 / (cur.dz / cur.D + prev.dz / prev.D);
                dp = 2.0 / (cur.dz / cur.D + nxt.dz / nxt.D);
            } else if (j < nl && j == jwt) {
                dm = 2.0 / (cur.dz / cur.D + prev.dz / prev.D);
                dp = 2.0 / (cur.dz * cur.k / cur.D + nxt.dz / nxt.D);
            } else if (j < nl) {
                dm = 2.0 / (cur.dz / cur.D + prev.dz * prev.k / prev.D);
                dp = 2.0 / (cur.dz / cur.D + nxt.dz / nxt.D);
            } else if (j != jwt + 1) {
                dm = 2.0 / (cur.dz / cur.D + prev.dz / prev.D);
            } else {
                dm = 2.0 / (cur.dz / cur.D + prev.dz * prev.k / prev.D);
            }
            if (j == 1) dm1_first = dm;
            // the row jj = j + 1 (:1874-1924)
            const double dzj = cur.dz, e = cur.e;
            double av, bv, cv, rv;
            if (j < nl && j == jwt) {
                av = -0.5 / dzj * dm;
                bv = e / dtime + 0.5 / dzj * (dp * cur.k + dm);
                cv = -0.5 / dzj * dp;
                rv = e / dtime * cur.cw + 0.5 / dzj * (dp * (nxt.cw - cur.cw * cur.k) - dm * (cur.cw - prev.cw)) +
                     cur.src;
            } else if (j < nl && j == jwt + 1) {
                av = -0.5 / dzj * dm * prev.k;
                bv = e / dtime + 0.5 / dzj * (dp + dm);
                cv = -0.5 / dzj * dp;
                rv = e / dtime * cur.cw + 0.5 / dzj * (dp * (nxt.cw - cur.cw) - dm * (cur.cw - prev.cw * prev.k)) +
                     cur.src;
            } else if (j < nl) {
                av = -0.5 / dzj * dm;
                bv = e / dtime + 0.5 / dzj * (dp + dm);
                cv = -0.5 / dzj * dp;
                rv = e / dtime * cur.cw + 0.5 / dzj * (dp * (nxt.cw - cur.cw) - dm * (cur.cw - prev.cw)) + cur.src;
            } else if (j == nl && j == jwt + 1) {
                av = -0.5 / dzj * dm * prev.k;
                bv = e / dtime + 0.5 / dzj * dm;
                cv = 0.0;
                rv = e / dtime * cur.cw + 0.5 / dzj * (-dm * (cur.cw - prev.cw * prev.k)) + cur.src;
            } else {
                av = -0.5 / dzj * dm;
                bv = e / dtime + 0.5 / dzj * dm;
                cv = 0.0;
                rv = e / dtime * cur.cw + 0.5 / dzj * (-dm * (cur.cw - prev.cw)) + cur.src;
            }
            const double denom = bv - av * cp_prev;
            cp_prev = cv / denom;
            dp_prev = (rv - av * dp_prev) / denom;
            s_cp[(int64_t)j * CH_BLOCK] = cp_prev;
            if (s == 0 && j == 1) cw1_old = cur.cw;
            if (s == 1) {                                 // parked until the back substitution
                *L.ptr(conc, j) = dp_prev;
                s_aux[(int64_t)(j - 1) * CH_BLOCK] = cur.e;
            } else {
                s_aux[(int64_t)(j - 1) * CH_BLOCK] = dp_prev;
            }
            prev = cur;
            cur = nxt;
        }
        // --- 3. back substitution (:1953-1957) bottom up, with the CH4 clip or the
        // O2 bounds and the absolute concentration of each level as it is solved
        // (:1967-1996). The recursion carries the unclipped values, as the
        // reference's solve finishes before its clip. O2 takes its epsilon from LDS;
        // CH4 forms it again. A CH4 level's clipped deficit goes to its multiplier's
        // LDS slot (used by then) and is summed top down below, in the reference's
        // order; its balance term goes to the slot of its right-hand side. The ground
        // node's value is not used.
        auto fin_at = [&](int j) {
            Ch4Fin f{};
            if (s == 0) {
                f.t = L.at(a.t_soisno, j);
                f.ws = L.at(a.watsat, j);
                f.vol = L.at(a.h2osoi_vol, j);
                if (p.ch4frzout) {
                    f.liq = L.at(a.h2osoi_liq, j);
                    f.ice = L.at(a.h2osoi_ice, j);
                }
                f.dz = L.at(a.dz, j);
            }
            f.conc = *L.ptr(conc, j);     // O2: the parked right-hand side; CH4: the old value
            return f;
        };
        double u_next = dp_prev;
        Ch4Fin qf = fin_at(nl);
        for (int j = nl; j >= 1; --j) {
            const Ch4Fin f = qf;
            if (j > 1) qf = fin_at(j - 1);
            double *aux = s_aux + (int64_t)(j - 1) * CH_BLOCK;
            if (j < nl) u_next = (s == 1 ? f.conc : *aux) - s_cp[(int64_t)j * CH_BLOCK] * u_next;
            double cw = u_next;
            if (s == 0) {
                const double k = L.khcc(f.t, s, logth);
                const double vmin = jl_min(f.ws, f.vol);
                const double lf = !p.ch4frzout ? 1.0
                                               : jl_max(0.05, (f.liq / CH_DENH2O + CH_SN) /
                                                                  (f.liq / CH_DENH2O + f.ice / CH_DENICE + CH_SN));
                const double e = j <= jwt ? f.ws - (1.0 - k) * vmin * lf : f.ws * lf;
                double deficit = 0.0;                     // negative-mass clip
                if (cw < 0.0) {
                    deficit = -cw * e * f.dz;
                    cw = 0.0;
                }
                s_cp[(int64_t)j * CH_BLOCK] = deficit;
                const double cnew = cw * e;
                __builtin_nontemporal_store(cnew, L.ptr(conc, j));
                *aux = (cnew - f.conc) * f.dz;
            } else {                                      // O2 bounds
                const double e = *aux;
                cw = jl_max(cw, 1.0e-12);
                cw = jl_min(cw, catm1 / e);
                __builtin_nontemporal_store(cw * e, L.ptr(conc, j));
            }
        }
        if (s == 0) {
            // surface fluxes (:1959-1966); u_next is the new relative CH4 of level 1
            if (jwt != 0) {
                surf_diff = dm1_first * ((u_next + cw1_old) / 2.0 - catm0);
                surf_ebul = 0.0;
            } else {
                surf_diff = dm1_first * ((u_next + cw1_old) / 2.0 - catm0 * L.khcc(a.t_grnd[c], 0, logth0));
                surf_ebul = ebul_total;
            }
            // the clipped deficit, levels in order (unclipped levels add +0.0)
            double deficit_tot = 0.0;
            for (int j = 1; j <= nl; ++j) deficit_tot += s_cp[(int64_t)j * CH_BLOCK];
            surf_diff -= deficit_tot / dtime;
        }
    }

    // CH4 balance and its correction (:1990-2001)
    double errch4 = 0.0;
    double bq[3] = {L.at(a.dz, 1), L.at(a.ch4_prod_depth, 1), L.at(a.ch4_oxid_depth, 1)};
    for (int j = 1; j <= nl; ++j) {
        const double dz = bq[0], pr = bq[1], ox = bq[2];
        if (j < nl) {
            bq[0] = L.at(a.dz, j + 1);
            bq[1] = L.at(a.ch4_prod_depth, j + 1);
            bq[2] = L.at(a.ch4_oxid_depth, j + 1);
        }
        errch4 += s_aux[(int64_t)(j - 1) * CH_BLOCK];     // (new - old) * dz
        errch4 -= pr * dz * dtime;
        errch4 += ox * dz * dtime;
    }
    errch4 += (surf_aere + surf_ebul + surf_diff) * dtime;
    if (fabs(errch4) < 1.0e-8) surf_diff -= errch4 / dtime;
    a.ch4_surf_diff[c] = surf_diff;
    a.ch4_surf_ebul[c] = surf_ebul;
    a.grnd_ch4_cond[c] = spec0;
}

}  // namespace

hipError_t launch_ch4_tran(const clmsoil_ch4_tran_args &a, const clmsoil_ch4_tran_params &p, double dtime,
                           unsigned *err, hipStream_t stream) {
    const dim3 grid((unsigned)((a.nc + CH_BLOCK - 1) / CH_BLOCK));
    const size_t lds = (size_t)(2 * p.nlevsoi + 1) * CH_BLOCK * sizeof(double);
    hipLaunchKernelGGL(ch4_tran_kernel, grid, dim3(CH_BLOCK), lds, stream, a, p, dtime, err);
    return hipGetLastError();
}

}  // namespace clmsoil
