// heat.hip — fused snow+soil heat step for non-urban soil columns (gfx950).
//
// One launch restates the body of soil_temperature!
// (src/biogeophys/soil_temperature.jl:358-593) for columns whose ground-heat-flux
// terms (hs_*, dhsdT, sabg_lyr_col) are given:
//   thermal properties  _soil_snow_tk_kernel!, _soil_tk_interface_kernel!,
//                       _tk_h2osfc_kernel!, _soil_cv_kernel!, _unresolved_snow_cv_kernel!,
//                       _snow_cv_kernel!                        (:599-872)
//   diffusion factor    _heat_diff_kernel!                      (:1271-1310)
//   h2osfc properties   _h2osfc_thermprop_kernel!               (:116-130)
//   RHS / 5-band matrix set_rhs_vec!, set_matrix!               (:1365-1748)
//   banded solve        batched_band_solve!                     (band_diagonal.jl:87-170)
//   unpack, fn1         _tvector_extract_kernel!, _fn1_kernel!  (:149-206)
//   phase change        phase_change_h2osfc!, phase_change_beta! (:1757-2289)
//   diagnostics         _tgrnd_kernel!, _eflx_fgr_kernel!       (:224-269)
//
// Layout: a group of G lanes owns one column; snow+soil level jj (1-based,
// jj = j + nlevsno) lives in lane (jj-1) mod G, register slot (jj-1) div G. The
// standing-surface-water (h2osfc) row of the reference's 5-band system couples
// only to soil layer 1, so it is carried as a column scalar.
//
// Solve: the reference densifies each column and runs Gaussian elimination with
// partial pivoting. Its matrix is tridiagonal in the level order except for the
// snow-0 -> soil-1 (band 1) and h2osfc <-> soil-1 (bands 2/4) couplings, so when
// no row interchange occurs the dense elimination performs exactly the
// operations of the banded sweep below (every other update subtracts f*0). The
// sweep checks the reference's pivot test at every step; a column that would
// interchange rows is re-solved by a banded partially pivoted elimination in its
// wavefront's LDS area. Elimination and back-substitution carries move between
// the lanes of a group by DPP broadcast.
//
// Registers: the kernel is compiled for two wavefronts per SIMD. Inputs are read
// level by level and turned into conductivity and heat capacity at once; the
// inputs needed again after the solve (water, ice, excess ice, dz, porosity, b,
// psi_sat, z) are read from memory again there (they are unchanged until this
// kernel writes its outputs); with G = 4 the original band rows wait for the
// pivoting fallback in LDS; diagnostics are written as soon as they are known.
#include <atomic>
#include <cstddef>

#include <hip/hip_runtime.h>
// The log and exp tables of clm_math.h are read from a per-block copy in LDS
// (8 KB per block; same values, same results): the table gathers of log10 and
// pow sit on the critical path of every level's conductivity. With the pivot
// tiles this fits two 256-thread blocks per CU.
static __shared__ double clm_lds_log_tab[256][3];
static __shared__ double clm_lds_exp_tab[128][2];
#define CLM_LOG_TAB(i, k) clm_lds_log_tab[i][k]
#define CLM_EXP_TAB(j, k) clm_lds_exp_tab[j][k]

#include "common.hpp"
#include "heat.hpp"

namespace clmsoil {

namespace {

template <int G>
__device__ __forceinline__ double hnxt(const double *x, int i, int k, int S) {
    double a = group_next<G>(x[i]);
    double b = (i + 1 < S) ? group_next<G>(x[i + 1]) : 0.0;
    return (k == G - 1) ? b : a;
}
template <int G>
__device__ __forceinline__ double hprv(const double *x, int i, int k) {
    double a = group_prev<G>(x[i]);
    double b = (i > 0) ? group_prev<G>(x[i - 1]) : 0.0;
    return (k == 0) ? b : a;
}
// value of level jj (runtime) broadcast to the whole group
template <int G, int S>
__device__ __forceinline__ double level_value(const double *x, int jj, int k) {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i)
        if (i * G + k + 1 == jj) v = x[i];
    return group_bcast<G>(v, (jj - 1) % G);
}

// LDS tile of one wavefront for the pivoting elimination: 38 rows x 7 band
// entries (columns r-2 .. r+4) plus the right-hand side.
constexpr int BAND_ROWS = MAX_NLEVSNO + MAX_NLEVGRND + 1;
constexpr int BAND_W = 7;
constexpr int TILE_DOUBLES = BAND_ROWS * BAND_W + BAND_ROWS;

// G = 4 keeps the rows for the pivoting fallback in LDS (3 values x S slots x 64
// lanes per wavefront); G = 1 and 2 would need too much LDS for that.
template <int G>
constexpr bool heat_saves_rows() { return G == 4; }
template <int G>
constexpr int heat_lds_doubles_per_wave() {
    constexpr int S = (MAX_NLEVSNO + MAX_NLEVGRND + G - 1) / G;
    return TILE_DOUBLES + (heat_saves_rows<G>() ? 3 * S * 64 : 0);
}


}  // namespace

// Argument pointers are fetched from the kernarg segment in the phase that uses
// them (karg_ptr), so that only that phase's pointers occupy SGPRs.
#define HP(field)                                                                        \
    karg_ptr<decltype(clmsoil_heat_args::field),                                        \
             (unsigned)(offsetof(HeatParams, a) + offsetof(clmsoil_heat_args, field))>()

// NSNO/NGRND > 0: the level counts are compile-time constants (the CLM5 grid),
// 0: read from P.
template <int G, int NSNO = 0, int NGRND = 0>
__device__ __forceinline__ void heat_column(const HeatParams &P, const int64_t c, const int k_lane,
                                            double *band_tile, double *saved_rows) {
    constexpr int NL = MAX_NLEVSNO + MAX_NLEVGRND;   // 37 snow+soil levels
    constexpr int S = (NL + G - 1) / G;
    constexpr bool SAVE_ROWS = heat_saves_rows<G>();
    const clmsoil_heat_args &a = P.a;
    const int64_t ld = a.ld;
    const int nlevsno = NSNO > 0 ? NSNO : P.nlevsno, nlevgrnd = NGRND > 0 ? NGRND : P.nlevgrnd;
    const int nlev = nlevsno + nlevgrnd;
    const int js1 = nlevsno + 1;                       // level index of soil layer 1
    const int64_t snl = HP(snl)[c];
    if (snl < -nlevsno || snl > 0) {
        if (k_lane == 0) atomicOr(P.err, (unsigned)DEV_ERR_SNL);
        return;
    }
    int jtop = (int)snl + nlevsno + 1;                 // top active level
    int k = k_lane;                                    // lane of the group
    // The per-level lane masks (jj >= jtop, the owner lane of a row, ...) are
    // recomputed in each phase rather than kept from the first one: hiding k and
    // jtop from the optimizer at the phase boundaries keeps dozens of 64-bit masks
    // from living through the kernel (where they would be spilled to VGPR lanes).
// The range of k is restated after each boundary, so that with the level counts
// compiled in, each slot's snow/soil tests still resolve at compile time.
#define HEAT_PHASE()                                   \
    do {                                               \
        asm volatile("" : "+v"(k), "+v"(jtop));        \
        __builtin_assume(k >= 0 && k < G);             \
    } while (0)
    const int nbedrock = (int)HP(nbedrock)[c];
    const double dtime = P.dtime;
    const double cnfac = CNFAC, omc = 1.0 - CNFAC;

    // ---- column scalars ----
    const double frac_sno = HP(frac_sno_eff)[c];
    const double frac_h2osfc = HP(frac_h2osfc)[c];
    const double dhsdT = HP(dhsdT)[c];
    const double hs_top_snow = HP(hs_top_snow)[c], hs_soil = HP(hs_soil)[c],
                 hs_h2osfc = HP(hs_h2osfc)[c];
    const double *eflx_bot_p = HP(eflx_bot);
    const double eflx_bot = eflx_bot_p ? eflx_bot_p[c] : 0.0;
    double t_h2osfc = HP(t_h2osfc)[c];
    double h2osfc = HP(h2osfc)[c];
    double h2osno_nl = HP(h2osno_no_layers)[c];
    double snow_depth = HP(snow_depth)[c];
    const double *int_snow_p = HP(int_snow), *snomelt_accum_p = HP(snomelt_accum);
    double int_snow = int_snow_p ? int_snow_p[c] : 0.0;
    double snomelt_accum = snomelt_accum_p ? snomelt_accum_p[c] : 0.0;

    // saturated conductivity tkmg * TKWAT^(fl*watsat) * TKICE^((1-fl)*watsat): the
    // logarithms of the two fixed bases once per column
    const clm_dd L_tkwat = clm_log_ext_normal(TKWAT), L_tkice = clm_log_ext_normal(TKICE);

    // ---- per-level loads, thermal conductivity thk (:647-716) and heat capacity
    // cv (:781-872), one level at a time so that only t, z, zi, dz, sabg, thk and
    // cv stay live afterwards ----
    double t[S], z[S], dz[S], zit[S], sabg[S], thk[S], cv[S];
    {
    const double *p_t = HP(t_soisno), *p_liq = HP(h2osoi_liq), *p_ice = HP(h2osoi_ice);
    const double *p_z = HP(z), *p_dz = HP(dz), *p_zi = HP(zi);
    const double *p_ws = HP(watsat), *p_tkmg = HP(tkmg), *p_tkdry = HP(tkdry), *p_csol = HP(csol);
    const double *p_exc = HP(excess_ice), *p_sabg = HP(sabg_lyr_col);
    double *p_thk = HP(thk), *p_bw = HP(bw);
    // The loads of slot i+1 are issued before the arithmetic of slot i (one slot
    // ahead), so their latency overlaps that arithmetic; the slot's geometry goes
    // straight into its arrays, the rest waits in these registers.
    struct LevelIn {
        double liq, ice, wsat, tkmg, tkdry, csol, exc;
    };
    auto load_slot = [&](int i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        const bool act = jj >= jtop && jj <= nlev;
        LevelIn v{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        t[i] = 0.0;
        z[i] = zit[i] = 0.0;
        dz[i] = 1.0;
        sabg[i] = 0.0;
        const int64_t o = c + (int64_t)(jj - 1) * ld;
        if (act) {
            t[i] = p_t[o];
            v.liq = p_liq[o];
            v.ice = p_ice[o];
            z[i] = __builtin_nontemporal_load(&p_z[o]);
            dz[i] = __builtin_nontemporal_load(&p_dz[o]);
            zit[i] = __builtin_nontemporal_load(&p_zi[o]);          // zi[c, jj] (top interface)
            if (j >= 1) {
                const int64_t os = c + (int64_t)(j - 1) * ld;
                v.wsat = __builtin_nontemporal_load(&p_ws[os]);
                v.tkmg = __builtin_nontemporal_load(&p_tkmg[os]);
                v.tkdry = __builtin_nontemporal_load(&p_tkdry[os]);
                v.csol = __builtin_nontemporal_load(&p_csol[os]);
                v.exc = p_exc[os];
            }
            if (p_sabg && jj <= nlevsno + 1)
                sabg[i] = p_sabg[c + (int64_t)(jj - 1) * ld];
        }
        return v;
    };
    LevelIn nxt_in = load_slot(0);
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        const bool act = jj >= jtop && jj <= nlev;
        const int64_t o = c + (int64_t)(jj - 1) * ld;
        const LevelIn cur = nxt_in;
        if (i + 1 < S) nxt_in = load_slot(i + 1);
        // the next slot's loads stay above this point, this slot's arithmetic below
        __builtin_amdgcn_sched_barrier(0);
        const double liq = cur.liq, ice = cur.ice, wsat = cur.wsat, tkmg = cur.tkmg,
                     tkdry = cur.tkdry, csol = cur.csol, exc = cur.exc;
        const double dzi = dz[i];
        double th = 0.0, bwv = 0.0;
        if (act && j >= 1) {
            double satw = (liq / DENH2O + ice / DENICE + exc / DENICE) / (dzi * wsat);
            satw = jl_min(1.0, satw);
            if (satw > 1.0e-7) {
                double dke;
                if (t[i] >= TFRZ) dke = jl_max(0.0, log10_w(satw) + 1.0);
                else dke = satw;
                const double fl_denom = liq / (DENH2O * dzi) + ice / (DENICE * dzi) +
                                        exc / (DENICE * dzi);
                const double fl = (liq / (DENH2O * dzi)) / fl_denom;
                const double dksat = tkmg * pow_base_w(L_tkwat, TKWAT, fl * wsat) *
                                     pow_base_w(L_tkice, TKICE, (1.0 - fl) * wsat);
                th = dke * dksat + (1.0 - dke) * tkdry;
            } else {
                th = tkdry;
            }
            if (j > nbedrock) th = THK_BEDROCK;
        }
        if (act && snl + 1 < 1 && j <= 0) {
            const double denom = jl_max(frac_sno, 1.0e-6) * jl_max(dzi, 1.0e-6);
            bwv = (ice + liq) / denom;
            if (P.snow_method == 2) {
                if (bwv <= 156.0) th = 0.023 + 0.234 * (bwv / 1000.0);
                else th = 0.138 - 1.01 * (bwv / 1000.0) + 3.233 * ((bwv / 1000.0) * (bwv / 1000.0));
            } else {
                th = TKAIR + (7.75e-5 * bwv + 1.105e-6 * (bwv * bwv)) * (TKICE - TKAIR);
            }
        }
        thk[i] = th;
        if (act) {
            if (p_thk) p_thk[o] = th;
            if (p_bw && j <= 0) p_bw[o] = bwv;
        }
        double v = 1.0;
        if (j >= 1) {
            v = csol * (1.0 - wsat) * dzi + (ice * CPICE + liq * CPLIQ) + exc * CPICE;
            if (j > nbedrock) v = CSOL_BEDROCK * dzi;
            if (j == 1 && h2osno_nl > 0.0) v += CPICE * h2osno_nl;
        } else if (jj >= jtop) {
            if (frac_sno > 0.0) v = jl_max(THIN_SFCLAYER, (CPLIQ * liq + CPICE * ice) / frac_sno);
            else v = THIN_SFCLAYER;
        }
        cv[i] = v;
        // the loads of the slot after next stay below this point
        __builtin_amdgcn_sched_barrier(0);
    }
    }
    HEAT_PHASE();
    // ---- interface conductivity tk, fact, fn (:605-629, :1271-1310) ----
    double tk[S], fact[S], fn[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        const bool act = jj >= jtop && jj <= nlev;
        const double thkn = hnxt<G>(thk, i, k, S);
        const double zn = hnxt<G>(z, i, k, S);
        const double zin = hnxt<G>(zit, i, k, S);
        const double tn = hnxt<G>(t, i, k, S);
        double tkv = 0.0, fa = 0.0, f = 0.0;
        if (act) {
            if (j <= nlevgrnd - 1)
                tkv = thk[i] * thkn * (zn - z[i]) / (thk[i] * (zn - zin) + thkn * (zin - z[i]));
            if (jj == jtop) {
                fa = dtime / cv[i] * dz[i] / (0.5 * (z[i] - zit[i] + CAPR * (zn - zit[i])));
                f = tkv * (tn - t[i]) / (zn - z[i]);
            } else if (j <= nlevgrnd - 1) {
                fa = dtime / cv[i];
                f = tkv * (tn - t[i]) / (zn - z[i]);
            } else {
                fa = dtime / cv[i];
                f = eflx_bot;
            }
        }
        tk[i] = tkv;
        fact[i] = fa;
        fn[i] = f;
    }
    // ---- h2osfc row: conductance, capacity, RHS and band entries ----
    const double thk1 = level_value<G, S>(thk, js1, k);
    const double z1 = level_value<G, S>(z, js1, k);
    const double t1 = level_value<G, S>(t, js1, k);
    double tk_h2osfc;
    {
        const double zh2osfc = 1.0e-3 * (0.5 * h2osfc);
        const double denom = TKWAT * z1 + thk1 * zh2osfc;
        tk_h2osfc = denom > 0.0 ? TKWAT * thk1 * (z1 + zh2osfc) / denom : thk1;
    }
    double c_h2osfc, dz_h2osfc;
    if (h2osfc > THIN_SFCLAYER && frac_h2osfc > THIN_SFCLAYER) {
        c_h2osfc = jl_max(THIN_SFCLAYER, CPLIQ * h2osfc / frac_h2osfc);
        dz_h2osfc = jl_max(THIN_SFCLAYER, 1.0e-3 * h2osfc / frac_h2osfc);
    } else {
        c_h2osfc = THIN_SFCLAYER;
        dz_h2osfc = THIN_SFCLAYER;
    }
    const double dzm_h = 0.5 * dz_h2osfc + z1;
    const double fn_h2osfc = tk_h2osfc * (t1 - t_h2osfc) / dzm_h;
    const double rv_h = t_h2osfc + (dtime / c_h2osfc) * (hs_h2osfc - dhsdT * t_h2osfc + cnfac * fn_h2osfc);
    const double b3_h = 1.0 + omc * (dtime / c_h2osfc) * tk_h2osfc / dzm_h - (dtime / c_h2osfc) * dhsdT;
    const double b2_h = -omc * (dtime / c_h2osfc) * tk_h2osfc / dzm_h;

    HEAT_PHASE();
    // ---- level rows: RHS and bands (:1365-1748) ----
    // sub = coupling to the level above (band 4, or band 5 snow-0 for soil 1),
    // sup = coupling to the next level below (band 2, or band 1 for snow 0),
    // b4h = soil-1 coupling to the h2osfc row (band 4).
    double rv[S], b3[S], sub[S], sup[S];
    double b4h = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        const bool act = jj >= jtop && jj <= nlev;
        const double fnp = hprv<G>(fn, i, k);
        const double tkp = hprv<G>(tk, i, k);
        const double zp = hprv<G>(z, i, k);
        const double zn = hnxt<G>(z, i, k, S);
        double r = 0.0, d = 1.0, lo = 0.0, up = 0.0;
        if (act) {
            const double dzp = zn - z[i];
            const double dzm = z[i] - zp;
            const bool snow = j <= 0, top = jj == jtop;
            const bool s1u = !snow && !top && j == 1;         // soil 1 under snow
            // The rows below need at most four quotients per level. Each is the
            // reference's own subexpression with its own rounding:
            //   q1 = omc*fact*tk/dzp (band 2 is -q1; -omc*fact*tk/dzp == -q1 exactly)
            //   q2 = omc*fact*tkp/dzm, or frac_sno*omc*fact*tkp/dzm for soil 1 under
            //        snow (the lower band is -q2)
            //   q3 = tk/dzp,  q4 = tkp/dzm, or frac_sno*tkp/dzm for soil 1 under snow
            // so that every lane divides four times whatever its case.
            const double of = omc * fact[i];
            const double q1 = of * tk[i] / dzp;
            const double q2 = (s1u ? frac_sno * omc * fact[i] : of) * tkp / dzm;
            const double q3 = tk[i] / dzp;
            const double q4 = (s1u ? frac_sno * tkp : tkp) / dzm;
            if (snow) {                                       // snow (_rhs_snow, _mat_snow)
                if (top) {
                    r = t[i] + fact[i] * (hs_top_snow - dhsdT * t[i] + cnfac * fn[i]);
                    d = 1.0 + q1 - fact[i] * dhsdT;
                } else {
                    r = t[i] + cnfac * fact[i] * (fn[i] - fnp);
                    r += fact[i] * sabg[i];
                    lo = -q2;
                    d = 1.0 + of * (q3 + q4);
                }
                up = -q1;                                     // band 2, or band 1 at j == 0
            } else if (top) {                                 // soil 1, no snow
                r = t[i] + fact[i] * (hs_top_snow - dhsdT * t[i] + cnfac * fn[i]);
                d = 1.0 + q1 - fact[i] * dhsdT;
                up = -q1;
            } else if (s1u) {                                 // soil 1 under snow
                r = t[i] + fact[i] * ((1.0 - frac_sno) * (hs_soil - dhsdT * t[i]) +
                                      cnfac * (fn[i] - frac_sno * fnp));
                r += frac_sno * fact[i] * sabg[i];
                up = -q1;
                d = 1.0 + of * (q3 + q4) - (1.0 - frac_sno) * fact[i] * dhsdT;
                lo = -q2;                                     // band 5 (snow 0)
            } else if (j <= nlevgrnd - 1) {
                r = t[i] + cnfac * fact[i] * (fn[i] - fnp);
                up = -q1;
                d = 1.0 + of * (q3 + q4);
                lo = -q2;
            } else {
                r = t[i] - cnfac * fact[i] * fnp + fact[i] * fn[i];
                d = 1.0 + q2;
                lo = -q2;
            }
            if (j == 1 && frac_h2osfc != 0.0) {               // h2osfc corrections
                r -= frac_h2osfc * fact[i] * ((hs_soil - dhsdT * t[i]) + cnfac * fn_h2osfc);
                d += frac_h2osfc * (omc * fact[i] * tk_h2osfc / dzm_h + fact[i] * dhsdT);
                b4h = -frac_h2osfc * omc * fact[i] * tk_h2osfc / dzm_h;
            }
        }
        rv[i] = r;
        b3[i] = d;
        sub[i] = lo;
        sup[i] = up;
    }
    // b4h stays in the soil-1 owner lane, which is the lane that eliminates that row

    // fn is needed after the solve only for the soil heat flux diagnostics: keep
    // soil layer 1's value and store -cnfac*fn, to which the post-solve part adds
    // -(1-cnfac)*fn1 (the same two roundings as computing the sum at once).
    const double fn_s1 = level_value<G, S>(fn, js1, k);
    if (double *p_fgr = HP(eflx_fgr)) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int jj = i * G + k + 1;
            const int j = jj - nlevsno;
            if (jj >= jtop && jj <= nlev && j >= 1)
                p_fgr[c + (int64_t)(j - 1) * ld] = (j < nlevgrnd) ? -cnfac * fn[i] : 0.0;
        }
    }
    // Compiler-only barrier: the level inputs used again after the solve (water,
    // ice, dz, soil properties, z) are reloaded from memory below instead of being
    // held in registers through the elimination.
    asm volatile("" ::: "memory");
    // The original diagonal, right-hand side and lower coupling of every row are
    // needed again only by the (rare) pivoting fallback below: with G = 4 they go
    // to this wavefront's LDS area instead of staying in registers.
    if constexpr (SAVE_ROWS) {
        const int lane = (int)__lane_id();
#pragma unroll
        for (int i = 0; i < S; ++i) {
            saved_rows[(0 * S + i) * 64 + lane] = b3[i];
            saved_rows[(1 * S + i) * 64 + lane] = rv[i];
            saved_rows[(2 * S + i) * 64 + lane] = sub[i];
        }
    }

    HEAT_PHASE();
    // ---- solve ----
    double x[S];
    double x_h = 0.0;
    bool pivot = false;
    {
        double u0[S], y[S];
#pragma unroll
        for (int i = 0; i < S; ++i) { u0[i] = 1.0; y[i] = 0.0; x[i] = 0.0; }
        double u0p = 0.0, yp = 0.0, supp = 0.0;             // previous level row
        double u0_s0 = 0.0, y_s0 = 0.0, b1_s0 = 0.0;         // snow-0 row
#pragma unroll
        for (int jj = 1; jj <= NL; ++jj) {
            const int i = (jj - 1) / G, kk = (jj - 1) % G;
            bool piv = false;
            if (k == kk && jj >= jtop && jj <= nlev) {
                double diag = b3[i], rhs = rv[i];
                if (jj == js1) {
                    if (jtop <= nlevsno) {                    // pivot column: snow 0
                        piv = piv || (fabs(sub[i]) > fabs(u0_s0));
                        const double f = sub[i] / u0_s0;
                        diag = diag - f * b1_s0;
                        rhs = rhs - f * y_s0;
                    }
                    piv = piv || (fabs(b4h) > fabs(b3_h));    // pivot column: h2osfc
                    const double f2 = b4h / b3_h;
                    diag = diag - f2 * b2_h;
                    rhs = rhs - f2 * rv_h;
                } else if (jj > jtop) {
                    piv = piv || (fabs(sub[i]) > fabs(u0p));
                    const double f = sub[i] / u0p;
                    diag = diag - f * supp;
                    rhs = rhs - f * yp;
                }
                u0[i] = diag;
                y[i] = rhs;
            }
            // a broadcast reads only lane kk's slot; after a row outside the column
            // (above its top level) the carried values are not used
            u0p = group_bcast<G>(u0[i], kk);
            yp = group_bcast<G>(y[i], kk);
            supp = group_bcast<G>(sup[i], kk);
            pivot = pivot || (group_bcast<G>(piv ? 1.0 : 0.0, kk) != 0.0);
            if (jj == nlevsno) { u0_s0 = u0p; y_s0 = yp; b1_s0 = supp; }
        }
        double xn = 0.0;
#pragma unroll
        for (int jj = NL; jj >= 1; --jj) {
            const int i = (jj - 1) / G, kk = (jj - 1) % G;
            if (k == kk && jj >= jtop && jj <= nlev) {
                double s = y[i];
                if (jj < nlev) s = s - sup[i] * xn;
                x[i] = s / u0[i];
            }
            xn = group_bcast<G>(x[i], kk);   // 0 outside the column, as initialised
            if (jj == js1) x_h = (rv_h - b2_h * xn) / b3_h;
        }
        // Columns whose elimination interchanges rows (the reference's partial
        // pivoting, band_diagonal.jl:125-140) are re-solved one at a time per
        // wavefront with the banded form of that elimination in an LDS tile.
        unsigned long long pv = __ballot(pivot && k == 0);
        while (pv) {
            const int leader = __ffsll((long long)pv) - 1;
            const bool mine = ((int)__lane_id() - k) == leader;
            double *tile = band_tile;                   // this wavefront's tile
            double *xt = tile + BAND_ROWS * BAND_W;
            const int jt = jtop;
            const int n = nlev - jtop + 2;
            const int rh = nlevsno - jt + 2;
            if (mine) {
                for (int q = k; q < n * BAND_W; q += G) tile[q] = 0.0;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
#define BT(r, cc) tile[((r) - 1) * BAND_W + ((cc) - (r) + 2)]
            if (mine) {
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    const int jj = i * G + k + 1;
                    if (jj < jtop || jj > nlev) continue;
                    const int r = jj <= nlevsno ? jj - jt + 1 : jj - jt + 2;
                    double b3i = b3[i], rvi = rv[i], subi = sub[i];
                    if constexpr (SAVE_ROWS) {
                        const int lane = (int)__lane_id();
                        b3i = saved_rows[(0 * S + i) * 64 + lane];
                        rvi = saved_rows[(1 * S + i) * 64 + lane];
                        subi = saved_rows[(2 * S + i) * 64 + lane];
                    }
                    BT(r, r) = b3i;
                    xt[r - 1] = rvi;
                    if (jj <= nlevsno) {
                        if (jj > jtop) BT(r, r - 1) = subi;
                        if (jj < nlevsno) BT(r, r + 1) = sup[i];
                        else BT(r, r + 2) = sup[i];
                    } else if (jj == js1) {
                        if (jtop <= nlevsno) BT(r, r - 2) = subi;
                        BT(r, r - 1) = b4h;
                        if (jj < nlev) BT(r, r + 1) = sup[i];
                    } else {
                        BT(r, r - 1) = subi;
                        if (jj < nlev) BT(r, r + 1) = sup[i];
                    }
                }
                if (k == 0) {
                    BT(rh, rh) = b3_h;
                    BT(rh, rh + 1) = b2_h;
                    xt[rh - 1] = rv_h;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            if ((int)__lane_id() == leader) {
                for (int kq = 1; kq <= n - 1; ++kq) {
                    const int rlast = kq + 2 < n ? kq + 2 : n;
                    const int clast = kq + 4 < n ? kq + 4 : n;
                    int pr = kq;
                    double maxv = fabs(BT(kq, kq));
                    for (int r = kq + 1; r <= rlast; ++r) {
                        const double v = fabs(BT(r, kq));
                        if (v > maxv) { maxv = v; pr = r; }
                    }
                    const double akk = BT(pr, kq);
                    if (pr != kq) {
                        for (int cc = kq; cc <= clast; ++cc) {
                            const double akj = BT(kq, cc), apj = BT(pr, cc);
                            BT(kq, cc) = apj;
                            BT(pr, cc) = akj;
                        }
                        const double xk = xt[kq - 1], xp = xt[pr - 1];
                        xt[kq - 1] = xp;
                        xt[pr - 1] = xk;
                    }
                    for (int r = kq + 1; r <= rlast; ++r) {
                        const double f = BT(r, kq) / akk;
                        for (int cc = kq; cc <= clast; ++cc) BT(r, cc) = BT(r, cc) - f * BT(kq, cc);
                        xt[r - 1] = xt[r - 1] - f * xt[kq - 1];
                    }
                }
                for (int r = n; r >= 1; --r) {
                    double sacc = xt[r - 1];
                    const int clast = r + 4 < n ? r + 4 : n;
                    for (int cc = r + 1; cc <= clast; ++cc) sacc = sacc - BT(r, cc) * xt[cc - 1];
                    xt[r - 1] = sacc / BT(r, r);
                }
            }
#undef BT
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            if (mine) {
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    const int jj = i * G + k + 1;
                    if (jj < jtop || jj > nlev) continue;
                    const int r = jj <= nlevsno ? jj - jt + 1 : jj - jt + 2;
                    x[i] = xt[r - 1];
                }
                x_h = xt[rh - 1];
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            pv &= pv - 1;
        }
    }

    HEAT_PHASE();
    // ---- unpack (:149-167) and fn1 (:170-206) ----
    double tnew[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        tnew[i] = (jj >= jtop && jj <= nlev) ? x[i] : 0.0;   // t of inactive levels is 0
    }
    const double x_s1 = level_value<G, S>(tnew, js1, k);
    t_h2osfc = (frac_h2osfc == 0.0) ? x_s1 : x_h;
    double fn1[S];
    {
        double zr[S];
        const double *p_z = HP(z);
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int jj = i * G + k + 1;
            zr[i] = (jj >= jtop && jj <= nlev) ? p_z[c + (int64_t)(jj - 1) * ld] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int jj = i * G + k + 1;
            const int j = jj - nlevsno;
            const double tn = hnxt<G>(tnew, i, k, S);
            const double zn = hnxt<G>(zr, i, k, S);
            double v = 0.0;
            if (jj >= jtop && j <= nlevgrnd - 1) v = tk[i] * (tn - tnew[i]) / (zn - zr[i]);
            fn1[i] = v;
        }
    }
    // water and ice of every level, read again (unchanged in memory until the end)
    double liq[S], ice[S];
    {
    const double *p_liq = HP(h2osoi_liq), *p_ice = HP(h2osoi_ice);
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        liq[i] = ice[i] = 0.0;
        if (jj >= jtop && jj <= nlev) {
            liq[i] = p_liq[c + (int64_t)(jj - 1) * ld];
            ice[i] = p_ice[c + (int64_t)(jj - 1) * ld];
        }
    }
    }

    HEAT_PHASE();
    // ---- phase_change_h2osfc! (:1757-1826) ----
    double xmf_h2osfc = 0.0, qflx_h2osfc_to_ice = 0.0, eflx_h2osfc_to_snow = 0.0;
    bool wrote_t0 = false;
    if (frac_h2osfc > 0.0 && t_h2osfc <= TFRZ) {
        wrote_t0 = true;
        const double fact0 = level_value<G, S>(fact, nlevsno, k);   // snow layer 0
        double t0 = level_value<G, S>(tnew, nlevsno, k);
        double ice0 = level_value<G, S>(ice, nlevsno, k);
        const double tinc = TFRZ - t_h2osfc;
        t_h2osfc = TFRZ;
        const double hm = frac_h2osfc * (dhsdT * tinc - tinc * c_h2osfc / dtime);
        const double xm = hm * dtime / HFUS;
        const double temp1 = h2osfc + xm;
        double h2osno_total = h2osno_nl;
        if (snl < 0) {
#pragma unroll
            for (int jj = 1; jj <= MAX_NLEVSNO; ++jj) {
                if (jj < jtop || jj > nlevsno) continue;
                const int i = (jj - 1) / G, kk = (jj - 1) % G;
                const double w = group_bcast<G>(k == kk ? ice[i] + liq[i] : 0.0, kk);
                h2osno_total += w;
            }
        }
        const double z_avg = frac_sno * snow_depth;
        double rho_avg = z_avg > 0.0 ? jl_min(800.0, h2osno_total / z_avg) : 200.0;
        if (temp1 >= 0.0) {
            int_snow -= xm;
            if (snl == 0) h2osno_nl -= xm;
            else ice0 -= xm;
            h2osno_total -= xm;
            h2osfc += xm;
            xmf_h2osfc = hm;
            qflx_h2osfc_to_ice = -xm / dtime;
            if (frac_sno > 0.0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
            if (snl == 0) {
                t0 = t_h2osfc;
                eflx_h2osfc_to_snow = 0.0;
            } else {
                double c1;
                if (snl == -1) c1 = frac_sno * (dtime / fact0 - dhsdT * dtime);
                else c1 = frac_sno / fact0 * dtime;
                const double c2 = frac_h2osfc != 0.0 ? (-CPLIQ * xm - frac_h2osfc * dhsdT * dtime) : 0.0;
                t0 = (c1 * t0 + c2 * t_h2osfc) / (c1 + c2);
                eflx_h2osfc_to_snow = (t_h2osfc - t0) * c2 / dtime;
            }
        } else {
            rho_avg = (h2osno_total * rho_avg + h2osfc * DENICE) / (h2osno_total + h2osfc);
            int_snow += h2osfc;
            if (snl == 0) h2osno_nl += h2osfc;
            else ice0 += h2osfc;
            h2osno_total += h2osfc;
            qflx_h2osfc_to_ice = h2osfc / dtime;
            t_h2osfc -= temp1 * HFUS / (dtime * dhsdT - c_h2osfc);
            xmf_h2osfc = hm - frac_h2osfc * temp1 * HFUS / dtime;
            if (snl == 0) {
                t0 = t_h2osfc;
            } else if (snl == -1) {
                const double c1 = frac_sno * (dtime / fact0 - dhsdT * dtime);
                const double c2 = frac_h2osfc != 0.0 ? frac_h2osfc * (c_h2osfc - dtime * dhsdT) : 0.0;
                t0 = (c1 * t0 + c2 * t_h2osfc) / (c1 + c2);
                t_h2osfc = t0;
            } else {
                const double c1 = frac_sno / fact0 * dtime;
                const double c2 = frac_h2osfc != 0.0 ? frac_h2osfc * (c_h2osfc - dtime * dhsdT) : 0.0;
                t0 = (c1 * t0 + c2 * t_h2osfc) / (c1 + c2);
                t_h2osfc = t0;
            }
            h2osfc = 0.0;
            if (frac_sno > 0.0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
        }
        // write back the snow-layer-0 slot (level nlevsno); for snl == 0 this is the
        // inactive slot the reference also writes (t_soisno[c, joff])
#pragma unroll
        for (int i = 0; i < S; ++i)
            if (i * G + k + 1 == nlevsno) { tnew[i] = t0; ice[i] = ice0; }
    }

    HEAT_PHASE();
    // ---- phase_change_beta! (:1929-2289), Float64 (hard) branch ----
    const double tfrz = TFRZ, hfus = HFUS;
    int imelt[S];
    double dxmf[S], qmelt_l[S], qfrz_l[S];
    double h2osno_nl_1 = h2osno_nl, snow_depth_1 = snow_depth, qsm_set = 0.0;
    bool unresolved = false;
    bool tgt_written[S];
    {
    double *p_t = HP(t_soisno), *p_liq = HP(h2osoi_liq), *p_ice = HP(h2osoi_ice);
    double *p_exc = HP(excess_ice), *p_qm = HP(qflx_snomelt_lyr), *p_qf = HP(qflx_snofrz_lyr);
    const double *p_ws = HP(watsat), *p_bsw = HP(bsw), *p_suc = HP(sucsat), *p_dz = HP(dz);
    double *p_fact = HP(fact), *p_fgr = HP(eflx_fgr), *p_exs = HP(exice_subs);
    int64_t *p_imelt = HP(imelt);
    // Soil properties of a level, read again (unchanged in memory), and its
    // eflx_fgr (written before the solve): issued one slot ahead of the slot's
    // arithmetic, as in the property phase.
    struct PcIn {
        double exc, wsat, bsw, suc, dz, fgr;
    };
    auto pc_load = [&](int i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        PcIn v{0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
        if (jj >= jtop && jj <= nlev && j >= 1) {
            const int64_t os = c + (int64_t)(j - 1) * ld;
            v.exc = p_exc[os];
            v.wsat = p_ws[os];
            v.bsw = p_bsw[os];
            v.suc = p_suc[os];
            v.dz = p_dz[c + (int64_t)(jj - 1) * ld];
            if (p_fgr && j < nlevgrnd) v.fgr = p_fgr[os];
        }
        return v;
    };
    PcIn pc_next = pc_load(0);
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int jj = i * G + k + 1;
        const int j = jj - nlevsno;
        const bool act = jj >= jtop && jj <= nlev;
        const int64_t o = c + (int64_t)(jj - 1) * ld;
        const PcIn cur = pc_next;
        if (i + 1 < S) pc_next = pc_load(i + 1);
        // the next slot's loads stay above this point, this slot's arithmetic below
        __builtin_amdgcn_sched_barrier(0);
        imelt[i] = 0;
        dxmf[i] = qmelt_l[i] = qfrz_l[i] = 0.0;
        tgt_written[i] = false;
        if (!act) {
            // phase_change_h2osfc! writes t_soisno[c, joff] also without snow layers
            if (jj == nlevsno && wrote_t0) p_t[o] = tnew[i];
            if (j <= 0 && jj >= 1) {
                p_qm[o] = 0.0;
                p_qf[o] = 0.0;
            }
            continue;
        }
        const int64_t os = c + (int64_t)(j - 1) * ld;
        const double exc_i = cur.exc, wsat_i = cur.wsat, bsw_i = cur.bsw, sucsat_i = cur.suc,
                     dz_i = cur.dz;
        const double wice0 = ice[i];
        const double wexice0 = j >= 1 ? exc_i : 0.0;
        const double wmass0 = ice[i] + liq[i] + wexice0;
        double tinc = 0.0;
        double tt = tnew[i];
        double supercool = 0.0;
        if (j <= 0) {
            if (ice[i] > 0.0 && tt > tfrz) { imelt[i] = 1; tinc = tfrz - tt; tt = tfrz; }
            if (liq[i] > 0.0 && tt < tfrz) { imelt[i] = 2; tinc = tfrz - tt; tt = tfrz; }
        } else {
            if (ice[i] > 0.0 && tt > tfrz) { imelt[i] = 1; tinc = tfrz - tt; tt = tfrz; }
            if (exc_i > 0.0 && tt > tfrz) { imelt[i] = 1; tinc = tfrz - tt; tt = tfrz; }
            if (tt < tfrz) {
                const double smp = hfus * (tfrz - tt) / (GRAV * tt) * 1000.0;
                supercool = wsat_i * pow_w(smp / sucsat_i, -1.0 / bsw_i);
                supercool *= dz_i * 1000.0;
            }
            if (liq[i] > supercool && tt < tfrz) { imelt[i] = 2; tinc = tfrz - tt; tt = tfrz; }
            if (h2osno_nl > 0.0 && j == 1) {
                if (tt > tfrz) { imelt[i] = 1; tinc = tfrz - tt; tt = tfrz; }
            }
        }
        // energy surplus/loss and rate of phase change
        double hm = 0.0;
        const double fct = fact[i];
        if (imelt[i] > 0) {
            // one quotient per lane: tinc/fct, or frac_sno*tinc/fct below the top
            // snow layer (-frac_sno*tinc/fct == -(frac_sno*tinc/fct) exactly)
            const bool sni = jj != jtop && j < 1;
            const double qh = (sni ? frac_sno * tinc : tinc) / fct;
            if (jj == jtop) {
                if (j > 0) hm = dhsdT * tinc - qh;
                else hm = frac_sno * (dhsdT * tinc - qh);
                if (j == 1 && frac_h2osfc != 0.0) hm -= frac_h2osfc * dhsdT * tinc;
            } else if (j == 1) {
                hm = (1.0 - frac_sno - frac_h2osfc) * dhsdT * tinc - qh;
            } else {
                hm = -qh;
            }
        }
        if (imelt[i] == 1 && hm < 0.0) { hm = 0.0; imelt[i] = 0; }
        if (imelt[i] == 2 && hm > 0.0) { hm = 0.0; imelt[i] = 0; }
        double icev = ice[i], liqv = liq[i], excv = exc_i;
        if (imelt[i] > 0 && fabs(hm) > 0.0) {
            double xm = hm * dtime / hfus;
            if (j == 1 && h2osno_nl > 0.0 && xm > 0.0) {
                const double temp1 = h2osno_nl;
                h2osno_nl_1 = jl_max(0.0, temp1 - xm);
                const double propor = h2osno_nl_1 / temp1;
                snow_depth_1 = propor * snow_depth;
                const double heatr = hm - hfus * (temp1 - h2osno_nl_1) / dtime;
                if (heatr > 0.0) { xm = heatr * dtime / hfus; hm = heatr; }
                else { xm = 0.0; hm = 0.0; }
                qsm_set = jl_max(0.0, temp1 - h2osno_nl_1) / dtime;
                unresolved = true;
            }
            // new ice (and excess ice) first; the latent heat they release is then
            // hfus*(wice0 - icev)/dtime, or hfus*(wexice0 - excv + wice0 - icev)/dtime
            // when excess ice melts too
            bool exc_melt = false;
            if (xm > 0.0) {
                icev = jl_max(0.0, wice0 - xm);
                const double xm2 = xm - icev;
                if (icev == 0.0 && wexice0 >= 0.0 && xm2 > 0.0 && j >= 2) {
                    excv = jl_max(0.0, wexice0 - xm2);
                    exc_melt = true;
                }
            } else if (xm < 0.0) {
                if (j <= 0) {
                    icev = jl_min(wmass0, wice0 - xm);
                } else {
                    if (wmass0 - wexice0 < supercool) icev = 0.0;
                    else icev = jl_min(wmass0 - wexice0 - supercool, wice0 - xm);
                }
            }
            const double qi = hfus * (wice0 - icev) / dtime;
            const double qx = exc_melt ? hfus * (wexice0 - excv + wice0 - icev) / dtime : qi;
            const double heatr = (xm > 0.0 || xm < 0.0) ? hm - qx : 0.0;
            const double ei_val = j >= 1 ? excv : 0.0;
            liqv = jl_max(0.0, wmass0 - icev - ei_val);
            if (fabs(heatr) > 0.0) {
                // tt += (fct, or fct/frac_sno for snow) * heatr, divided by the
                // implicit factor of the top layer and of soil layer 1 (by 1.0,
                // which is exact, elsewhere); snow layers only when frac_sno > 0
                const bool snowl = j <= 0;
                const double fa = snowl ? fct / frac_sno : fct;
                double den = 1.0;
                if (jj == jtop && j == 1) den = 1.0 - (1.0 - frac_h2osfc) * fct * dhsdT;
                else if (jj == jtop) den = 1.0 - fct * dhsdT;
                else if (j == 1) den = 1.0 - (1.0 - frac_sno - frac_h2osfc) * fct * dhsdT;
                if (!snowl || frac_sno > 0.0) tt += fa * heatr / den;
                if (j <= 0)
                    if (liqv * icev > 0.0) tt = tfrz;
            }
            dxmf[i] = (j >= 1) ? qi + hfus * (wexice0 - excv) / dtime : qi;
            if (j < 1 && (imelt[i] == 1 || imelt[i] == 2)) {
                const double qm = jl_max(0.0, imelt[i] == 1 ? wice0 - icev : icev - wice0) / dtime;
                if (imelt[i] == 1) qmelt_l[i] = qm;
                else qfrz_l[i] = qm;
            }
            tgt_written[i] = true;
        }
        tnew[i] = tt;
        // write back this level (:1929-2289 outputs)
        p_t[o] = tt;
        p_liq[o] = liqv;
        p_ice[o] = icev;
        if (j >= 1 && excv != wexice0) p_exc[os] = excv;
        if (j <= 0) {
            p_qm[o] = qmelt_l[i];
            p_qf[o] = qfrz_l[i];
        }
        if (p_fact) p_fact[o] = fact[i];
        if (p_imelt) p_imelt[o] = imelt[i];
        if (j >= 1) {
            if (p_fgr && j < nlevgrnd) p_fgr[os] = cur.fgr - (1.0 - cnfac) * fn1[i];
            if (p_exs)
                p_exs[os] = tgt_written[i] ? jl_max(0.0, (wexice0 - excv) / DENICE) : 0.0;
        }
        // the loads of the slot after next stay below this point
        __builtin_amdgcn_sched_barrier(0);
    }
    }
    // column accumulators, summed over levels in the reference's order
    double xmf = 0.0, qflx_snomelt = 0.0, qflx_snofrz = 0.0, qflx_snow_drain = 0.0;
    {
        const bool ur = group_bcast<G>(unresolved ? 1.0 : 0.0, (js1 - 1) % G) != 0.0;
        const double qs = group_bcast<G>(qsm_set, (js1 - 1) % G);
        if (ur) {
            h2osno_nl = group_bcast<G>(h2osno_nl_1, (js1 - 1) % G);
            snow_depth = group_bcast<G>(snow_depth_1, (js1 - 1) % G);
        }
        // The reference adds a level's terms only where its phase change was
        // applied (tgt_written). Elsewhere dxmf, qmelt_l and qfrz_l are +0.0, and
        // adding +0.0 changes no sum that cannot be -0.0: xmf, qflx_snomelt and
        // qflx_snofrz start at +0.0 (or a max(0, .) quotient) and only gain terms
        // that are >= +0.0, NaN or (dxmf) of either sign, and a round-to-nearest
        // sum is -0.0 only when both terms are. So those sums take every level's
        // term. snomelt_accum is an input (it may hold -0.0): its term is sent as
        // -1.0 (no max(0, .) quotient is negative) where the reference adds none.
#pragma unroll
        for (int jj = 1; jj <= NL; ++jj) {
            const int i = (jj - 1) / G, kk = (jj - 1) % G;
            if (jj == js1 && ur) {
                qflx_snomelt = qs;
                xmf = hfus * qflx_snomelt;
                qflx_snow_drain = qflx_snomelt;
            }
            xmf += group_bcast<G>(dxmf[i], kk);
            if (jj <= nlevsno) {
                const double qa =
                    group_bcast<G>(tgt_written[i] && imelt[i] == 1 ? qmelt_l[i] : -1.0, kk);
                if (!(qa < 0.0)) {
                    qflx_snomelt += qa;
                    snomelt_accum += qa * dtime * 1.0e-3;
                }
                qflx_snofrz += group_bcast<G>(qfrz_l[i], kk);
            }
        }
    }

    HEAT_PHASE();
    // ---- ground temperature (:224-247) and soil heat flux (:250-269) ----
    const double t_top = level_value<G, S>(tnew, jtop, k);
    const double t_s1 = level_value<G, S>(tnew, js1, k);
    double tg;
    if (snl < 0) {
        if (frac_h2osfc != 0.0)
            tg = frac_sno * t_top + (1.0 - frac_sno - frac_h2osfc) * t_s1 + frac_h2osfc * t_h2osfc;
        else
            tg = frac_sno * t_top + (1.0 - frac_sno) * t_s1;
    } else {
        if (frac_h2osfc != 0.0) tg = (1.0 - frac_h2osfc) * t_s1 + frac_h2osfc * t_h2osfc;
        else tg = t_s1;
    }
    const double fn1_s1 = level_value<G, S>(fn1, js1, k);

    // ---- write back the column scalars (levels were written above) ----
    if (k == 0) {
        HP(t_h2osfc)[c] = t_h2osfc;
        HP(h2osfc)[c] = h2osfc;
        HP(h2osno_no_layers)[c] = h2osno_nl;
        HP(snow_depth)[c] = snow_depth;
        if (double *p = HP(int_snow)) p[c] = int_snow;
        HP(t_grnd)[c] = tg;
        HP(xmf)[c] = xmf;
        HP(xmf_h2osfc)[c] = xmf_h2osfc;
        HP(qflx_snomelt)[c] = qflx_snomelt;
        HP(qflx_snofrz)[c] = qflx_snofrz;
        HP(qflx_snow_drain)[c] = qflx_snow_drain;
        HP(qflx_h2osfc_to_ice)[c] = qflx_h2osfc_to_ice;
        HP(eflx_fgr12)[c] = -cnfac * fn_s1 - (1.0 - cnfac) * fn1_s1;
        if (double *p = HP(c_h2osfc)) p[c] = c_h2osfc;
        if (double *p = HP(eflx_snomelt)) p[c] = qflx_snomelt * hfus;
        if (double *p = HP(eflx_snomelt_r)) p[c] = qflx_snomelt * hfus;
        if (double *p = HP(eflx_h2osfc_to_snow)) p[c] = eflx_h2osfc_to_snow;
        if (double *p = HP(snomelt_accum)) p[c] = snomelt_accum;
    }
}

// Dynamic LDS: one area of heat_lds_doubles_per_wave<G>() doubles per wavefront.
template <int G, int NSNO, int NGRND>
__global__ __launch_bounds__(256, CLMSOIL_HEAT_WAVES) void heat_kernel(HeatParams P) {
    extern __shared__ double heat_lds[];
    for (int t = threadIdx.x; t < 256 * 3; t += blockDim.x)
        (&clm_lds_log_tab[0][0])[t] = (&clm_log_tab[0][0])[t];
    for (int t = threadIdx.x; t < 128 * 2; t += blockDim.x)
        (&clm_lds_exp_tab[0][0])[t] = (&clm_exp_tab[0][0])[t];
    __syncthreads();
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = gtid / G;
    const int k = (int)(threadIdx.x % G);
    if (c >= P.a.nc) return;
    if (const uint8_t *mask = HP(mask)) {
        if (!mask[c]) return;
    }
    double *area = heat_lds + (int64_t)(threadIdx.x / 64) * heat_lds_doubles_per_wave<G>();
    heat_column<G, NSNO, NGRND>(P, c, k, area, area + TILE_DOUBLES);
}

template <int G, int NSNO, int NGRND>
static hipError_t launch_g(const HeatParams &P, hipStream_t stream, int block) {
    const int64_t threads = P.a.nc * G;
    const int64_t grid = (threads + block - 1) / block;
    const size_t lds = (size_t)(block / 64) * heat_lds_doubles_per_wave<G>() * sizeof(double);
    if (lds > 65536) {
        static std::atomic<int> raised{0};
        if (!raised.load(std::memory_order_relaxed)) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&heat_kernel<G, NSNO, NGRND>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds);
            if (e != hipSuccess) return e;
            raised.store(1, std::memory_order_relaxed);
        }
    }
    hipLaunchKernelGGL((heat_kernel<G, NSNO, NGRND>), dim3((unsigned)grid), dim3(block), lds, stream, P);
    return hipGetLastError();
}

hipError_t launch_heat(const HeatParams &P, int G, int block, hipStream_t stream) {
    if (P.a.nc == 0) return hipSuccess;
    // the CLM5 grid (12 snow + 25 ground levels) with its level counts compiled in
    const bool clm5 = P.nlevsno == MAX_NLEVSNO && P.nlevgrnd == MAX_NLEVGRND;
    switch (G) {
        case 1: return launch_g<1, 0, 0>(P, stream, block);
        case 2: return launch_g<2, 0, 0>(P, stream, block);
        default:
            return clm5 ? launch_g<4, MAX_NLEVSNO, MAX_NLEVGRND>(P, stream, block)
                        : launch_g<4, 0, 0>(P, stream, block);
    }
}

}  // namespace clmsoil
