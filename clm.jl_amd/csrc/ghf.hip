// ghf.hip — ground heat flux and its temperature derivative (gfx950).
//
// Follows compute_ground_heat_flux_and_deriv! (src/biogeophys/soil_temperature.jl:
// 1200-1263). The reference runs five kernels (longwave emission per column; rural,
// rural-contribution and urban net flux per patch; SNICAR top layer per patch) with
// atomic patch-to-column adds and per-patch scratch arrays in between. Here one
// launch does it all, with two roles per thread:
//
//   patch role  (thread p < np): the per-patch outputs of the rural and urban
//               kernels (eflx_gnet, dgnetdT, sabg_chk, urban heat terms);
//   column role (thread c < nc): the column's longwave emission, then a walk over
//               its patches patchi..patchf in ascending order that recomputes each
//               patch's contributions and adds them in registers to hs_soil,
//               hs_h2osfc, dhsdT, hs_top, hs_top_snow and the per-layer absorbed
//               solar; the column's sums are written once at the end.
//
// The column owns its sums, so there are no atomics and no zero-fill pass, and the
// additions happen in the order of the reference's sequential CPU path: results are
// bit-identical to it (the reference's own GPU path, with atomics, is not even
// deterministic). Every expression keeps the reference's operand order; x^4 is
// Julia's compensated power by squaring.
#include <climits>

#include "common.hpp"
#include "ghf.hpp"

namespace clmsoil {
namespace {

constexpr double SB = 5.67e-8;                          // varcon.jl:16
constexpr double HVAP = 2.501e6;                        // varcon.jl:24
constexpr int64_t ICOL_ROAD_IMPERV = 74, ICOL_ROAD_PERV = 75;   // column_varcon.jl:10-11

// Julia's x^4 for Float64 (Base.Math.pow_body): square twice, carrying each
// square's rounding error (exact through fma) and adding it back at the end.
__device__ __forceinline__ double jl_pow4(double x) {
    const double x1 = x * x, e1 = __builtin_fma(x, x, -x1);
    const double err = x1 * 2.0 * e1;
    const double x2 = x1 * x1, e2 = __builtin_fma(x1, x1, -x2);
    const double lo = e2 + err;
    if (!(__builtin_isfinite(x2) && __builtin_isfinite(lo))) return x2;
    return x2 + lo;
}

struct Emit {
    double lw, dlw, lw_snow, lw_soil, lw_h2osfc;
};

// _ground_lwrad_emit_kernel! (:946-960); zeros for columns outside the mask, as
// the reference's zero-initialised scratch.
__device__ __forceinline__ Emit column_emission(const clmsoil_ghf_args &a, int64_t c, int joff,
                                                bool snl_ok) {
    Emit e{0.0, 0.0, 0.0, 0.0, 0.0};
    if (!a.mask_nolakec[c] || !snl_ok) return e;
    const double eg = a.emg[c];
    const double tg = a.t_grnd[c];
    const int64_t ld = a.ldc;
    e.lw = eg * SB * jl_pow4(tg);
    e.dlw = 4.0 * eg * SB * (tg * tg * tg);
    e.lw_snow = eg * SB * jl_pow4(a.t_soisno[c + (int64_t)(a.snl[c] + joff) * ld]);
    e.lw_soil = eg * SB * jl_pow4(a.t_soisno[c + (int64_t)joff * ld]);
    e.lw_h2osfc = eg * SB * jl_pow4(a.t_h2osfc[c]);
    return e;
}

// Urban heat terms of _gnet_patch_urban_kernel! (:1058-1073). With neither
// building method set, a road patch keeps its prior eflx_ventilation (the
// reference leaves it unwritten); nothing writes it in that case, so reading it
// here is race-free.
struct Urban {
    double waste, vent, ac, traffic;
};
__device__ __forceinline__ Urban urban_terms(const clmsoil_ghf_args &a, int64_t p, int64_t l,
                                             int64_t it) {
    Urban u;
    const bool simple = a.building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE;
    const bool prog = a.building_temp_method == CLMSOIL_BUILDING_TEMP_PROG;
    const double denom = 1.0 - a.wtlunit_roof[l];
    if (it == ICOL_ROAD_PERV || it == ICOL_ROAD_IMPERV) {
        u.waste = a.eflx_wasteheat_lun[l] / denom;
        u.vent = simple ? 0.0 : (prog ? a.eflx_ventilation_lun[l] / denom : a.eflx_ventilation[p]);
        u.ac = a.eflx_heat_from_ac_lun[l] / denom;
        u.traffic = a.eflx_traffic_lun[l] / denom;
    } else {
        u.waste = u.vent = u.ac = u.traffic = 0.0;
    }
    return u;
}

__device__ __forceinline__ bool snl_in_range(const clmsoil_ghf_args &a, int64_t c, int nlevsno) {
    const int64_t s = a.snl[c];
    return s >= -nlevsno && s <= 0;
}

// One block owns GHF_B consecutive columns and, through them, the contiguous
// patch range patchi[c0]..patchf[c0 + GHF_B - 1]. It walks that range in chunks
// of GHF_B patches: each thread loads one patch (all patch arrays are read once,
// coalesced), writes that patch's own outputs and stages its contributions to the
// column sums in LDS; after a barrier, each thread adds the staged contributions
// of its own column's patches in ascending patch order. The column data every
// patch needs (emission terms, emissivity, forcing, snow state) is computed once
// per column into LDS at the start.
#ifndef GHF_BLOCK
#define GHF_BLOCK 128
#endif
constexpr int GHF_B = GHF_BLOCK;
constexpr int GHF_NL = MAX_NLEVSNO + 1;   // levels of sabg_lyr_* (snow layers + soil layer 1)

struct GhfShared {
    // per column of the block
    double lw[GHF_B], dlw[GHF_B], lw_snow[GHF_B], lw_soil[GHF_B], lw_h2osfc[GHF_B];
    double emg[GHF_B], forc[GHF_B], htvp[GHF_B], fsno[GHF_B];
    int ktop[GHF_B];       // 0-based level of the top layer in sabg_lyr_*; -1: snl invalid
    int itype[GHF_B];
    long long lun[GHF_B];  // landunit of the column's first patch (-1: none)
    int urb[GHF_B];        // urbpoi of that landunit
    // per patch of the current chunk
    double dg[GHF_B], soil[GHF_B], h2o[GHF_B], top[GHF_B], top_snow[GHF_B];
    double sab[GHF_NL][GHF_B];
    int slot[GHF_B];       // owning column's slot in the block, -1: contributes nothing
    int kind[GHF_B];       // 1 rural, 2 urban
    long long lo[GHF_B / 64], hi[GHF_B / 64];
};

__device__ __forceinline__ long long wave_min_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ long long wave_max_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// The patch inputs that need no other load first; loaded one chunk ahead.
struct PatchIn {
    int64_t pc, l;
    bool mask;
    double fvn, wt, sabg, sabg_soil, sabg_snow, dlrad, cgrnd;
    double sh, sh_snow, sh_soil, sh_h2osfc, ev, ev_snow, ev_soil, ev_h2osfc;
};
__device__ __forceinline__ PatchIn load_patch(const clmsoil_ghf_args &a, long long p) {
    PatchIn r;
    r.pc = a.patch_column[p] - 1;
    r.l = a.patch_landunit[p] - 1;
    r.mask = a.mask_nolakep[p];
    r.fvn = (double)a.frac_veg_nosno[p];
    r.wt = a.wtcol[p];
    r.sabg = a.sabg[p];
    r.sabg_soil = a.sabg_soil[p];
    r.sabg_snow = a.sabg_snow[p];
    r.dlrad = a.dlrad[p];
    r.cgrnd = a.cgrnd[p];
    r.sh = a.eflx_sh_grnd[p];
    r.sh_snow = a.eflx_sh_snow[p];
    r.sh_soil = a.eflx_sh_soil[p];
    r.sh_h2osfc = a.eflx_sh_h2osfc[p];
    r.ev = a.qflx_evap_soi[p];
    r.ev_snow = a.qflx_ev_snow[p];
    r.ev_soil = a.qflx_ev_soil[p];
    r.ev_h2osfc = a.qflx_ev_h2osfc[p];
    return r;
}

__global__ void __launch_bounds__(GHF_B) ghf_kernel(GhfParams P) {
    __shared__ GhfShared S;
    const clmsoil_ghf_args &a = P.a;
    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * GHF_B;
    const int64_t c = c0 + t;
    const int nlevsno = P.nlevsno, joff = P.nlevsno;
    const int64_t ldp = a.ldp;
    unsigned bad = 0u;

    // ---- columns: emission (:946-960) and the data their patches read
    long long my_lo = LLONG_MAX, my_hi = -1;
    if (c < a.nc) {
        const bool ok = snl_in_range(a, c, nlevsno);
        if (a.mask_nolakec[c] && !ok) bad |= DEV_ERR_SNL;
        const Emit e = column_emission(a, c, joff, ok);
        S.lw[t] = e.lw;
        S.dlw[t] = e.dlw;
        S.lw_snow[t] = e.lw_snow;
        S.lw_soil[t] = e.lw_soil;
        S.lw_h2osfc[t] = e.lw_h2osfc;
        S.emg[t] = a.emg[c];
        S.forc[t] = a.forc_lwrad[c];
        S.htvp[t] = a.htvp[c];
        S.fsno[t] = a.frac_sno_eff[c];
        S.ktop[t] = ok ? (int)(a.snl[c] + nlevsno) : -1;
        S.itype[t] = (int)a.col_itype[c];
        // patchi = ISPVAL (-9999): no patches; patchf = patchi - 1: an empty range
        const int64_t p0 = a.patchi[c], p1 = a.patchf[c];
        long long lun = -1;
        int urb = 0;
        if (p0 >= 1 && p1 >= p0 && p1 <= a.np) {
            my_lo = p0 - 1;
            my_hi = p1 - 1;
            lun = a.patch_landunit[my_lo] - 1;
            if (lun >= 0 && lun < a.nl) urb = a.urbpoi[lun];
            else lun = -1;
        } else if (p0 != -9999 && !(p0 >= 1 && p1 == p0 - 1)) {
            bad |= DEV_ERR_SUBGRID;
        }
        S.lun[t] = lun;
        S.urb[t] = urb;
    }
    {
        const long long wl = wave_min_i64(my_lo), wh = wave_max_i64(my_hi);
        if ((t & 63) == 0) {
            S.lo[t >> 6] = wl;
            S.hi[t >> 6] = wh;
        }
    }
    __syncthreads();
    long long blo = S.lo[0], bhi = S.hi[0];
    for (int w = 1; w < GHF_B / 64; ++w) {
        blo = S.lo[w] < blo ? S.lo[w] : blo;
        bhi = S.hi[w] > bhi ? S.hi[w] : bhi;
    }

    double s_dhsdT = 0.0, s_soil = 0.0, s_h2osfc = 0.0, s_top = 0.0, s_top_snow = 0.0;
    double s_lyr[GHF_NL];
#pragma unroll
    for (int k = 0; k < GHF_NL; ++k) s_lyr[k] = 0.0;
    const int my_ktop = c < a.nc ? S.ktop[t] : -1;

    // absorbed solar rows k = ktop..nlevsno of a patch of this block (raw values)
    auto load_rows = [&](const PatchIn &in, long long pp, double *rows) {
        const int64_t pc = in.pc;
        if (pp > bhi || pc < c0 || pc >= c0 + GHF_B || pc >= a.nc) return;
        const int ktop = S.ktop[pc - c0];
        if (ktop < 0) return;
#pragma unroll
        for (int k = 0; k < GHF_NL; ++k)
            if (k >= ktop && k <= nlevsno) rows[k] = a.sabg_lyr_patch[pp + (int64_t)k * ldp];
    };
    PatchIn cur{};
    double rows[GHF_NL];
#pragma unroll
    for (int k = 0; k < GHF_NL; ++k) rows[k] = 0.0;
    if (blo + t <= bhi) cur = load_patch(a, blo + t);
    load_rows(cur, blo + t, rows);
    for (long long q = blo; q <= bhi; q += GHF_B) {
        // ---- prefetch the next chunk's independent patch inputs
        PatchIn nxt{};
        if (q + GHF_B + t <= bhi) nxt = load_patch(a, q + GHF_B + t);
        double rows_nxt[GHF_NL];
#pragma unroll
        for (int k = 0; k < GHF_NL; ++k) rows_nxt[k] = 0.0;
        // ---- one patch per thread: its outputs and its staged contributions
        const long long p = q + t;
        int slot = -1, kind = 0;
        if (p <= bhi) {
            const int64_t pc = cur.pc, l = cur.l;
            const bool pmask = cur.mask;
            const double fvn = cur.fvn, wt = cur.wt, sabg = cur.sabg, sabg_soil = cur.sabg_soil;
            const double sabg_snow = cur.sabg_snow, dlrad = cur.dlrad, cgrnd = cur.cgrnd;
            const double sh = cur.sh, sh_snow = cur.sh_snow, sh_soil = cur.sh_soil;
            const double sh_h2osfc = cur.sh_h2osfc, ev = cur.ev, ev_snow = cur.ev_snow;
            const double ev_soil = cur.ev_soil, ev_h2osfc = cur.ev_h2osfc;
            const bool mine = pc >= c0 && pc < c0 + GHF_B && pc < a.nc;
            if (!mine) {
                // belongs to another block's columns: only legal outside every
                // range of this block, which the column loop below checks
            } else if (pmask) {
                const int cs = (int)(pc - c0);
                // the column's landunit is that of its first patch; another
                // landunit on a later patch is looked up on its own
                const bool lok = l >= 0 && l < a.nl;
                const bool urban = lok && (l == S.lun[cs] ? S.urb[cs] != 0 : a.urbpoi[l] != 0);
                if (!lok) {
                    bad |= DEV_ERR_SUBGRID;
                } else if (!urban) {
                    // _gnet_patch_rural_kernel! (:1008-1023) and its contribution
                    // kernel (:1025-1049); _gnet_snicar_kernel! rural branch (:1142-1154)
                    const bool wnz = wt != 0.0;
                    const double htvp = S.htvp[cs];
                    const double base = dlrad + (1.0 - fvn) * S.emg[cs] * S.forc[cs];
                    const double gnet = sabg + base - S.lw[cs] - (sh + ev * htvp);
                    const double dg = -cgrnd - S.dlw[cs];
                    a.eflx_gnet[p] = gnet;
                    a.sabg_chk[p] = S.fsno[cs] * sabg_snow + (1.0 - S.fsno[cs]) * sabg_soil;
                    a.dgnetdT[p] = dg;
                    const double gsoil = sabg_soil + base - S.lw_soil[cs] - (sh_soil + ev_soil * htvp);
                    const double gh2o = sabg_soil + base - S.lw_h2osfc[cs] -
                                        (sh_h2osfc + ev_h2osfc * htvp);
                    S.dg[t] = wnz ? dg * wt : 0.0;
                    S.soil[t] = wnz ? gsoil * wt : 0.0;
                    S.h2o[t] = wnz ? gh2o * wt : 0.0;
                    const int ktop = S.ktop[cs];
                    if (ktop >= 0) {
                        double sab = 0.0;
#pragma unroll
                        for (int k = 0; k < GHF_NL; ++k)
                            if (k == ktop) sab = rows[k];
                        const double gtop = sab + base - S.lw[cs] - (sh + ev * htvp);
                        S.top[t] = gtop * wt;
                        const double gsnow = sab + base - S.lw_snow[cs] - (sh_snow + ev_snow * htvp);
                        S.top_snow[t] = gsnow * wt;
#pragma unroll
                        for (int k = 0; k < GHF_NL; ++k)
                            if (k >= ktop && k <= nlevsno) S.sab[k][t] = rows[k] * wt;
                    }
                    slot = cs;
                    kind = 1;
                } else {
                    // _gnet_patch_urban_kernel! (:1051-1085); urban SNICAR branch (:1155-1160)
                    const bool wnz = wt != 0.0;
                    const int64_t it = S.itype[cs];
                    const Urban u = urban_terms(a, p, l, it);
                    a.eflx_wasteheat[p] = u.waste;
                    if (a.building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE ||
                        a.building_temp_method == CLMSOIL_BUILDING_TEMP_PROG ||
                        !(it == ICOL_ROAD_PERV || it == ICOL_ROAD_IMPERV))
                        a.eflx_ventilation[p] = u.vent;
                    a.eflx_heat_from_ac[p] = u.ac;
                    a.eflx_traffic[p] = u.traffic;
                    const double g = sabg + dlrad - a.eflx_lwrad_net[p] -
                                     (sh + ev * S.htvp[cs] + a.qflx_tran_veg[p] * HVAP) +
                                     u.waste + u.ac + u.traffic + u.vent;
                    a.eflx_gnet[p] = g;
                    if (a.building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE)
                        a.eflx_anthro[p] = u.waste + u.traffic;
                    const double dg = -cgrnd - S.dlw[cs];
                    a.dgnetdT[p] = dg;
                    S.dg[t] = wnz ? dg * wt : 0.0;
                    S.soil[t] = wnz ? g * wt : 0.0;
                    S.h2o[t] = wnz ? g * wt : 0.0;
                    S.top[t] = g * wt;
                    S.top_snow[t] = g * wt;
                    const int ktop = S.ktop[cs];
                    if (ktop >= 0) S.sab[ktop][t] = sabg * wt;
                    slot = cs;
                    kind = 2;
                }
            } else {
                slot = -2;   // the column's own patch, outside the mask: adds nothing
            }
        }
        S.slot[t] = slot;
        S.kind[t] = kind;
        __syncthreads();
        // the next chunk's absorbed-solar rows, in flight during the column loop
        load_rows(nxt, q + GHF_B + t, rows_nxt);
        // ---- the column adds its patches of this chunk, in ascending order
        if (my_hi >= 0 && my_lo <= q + GHF_B - 1 && my_hi >= q) {
            const long long pa = my_lo > q ? my_lo : q;
            const long long pb = my_hi < q + GHF_B - 1 ? my_hi : q + GHF_B - 1;
            // branch-free: every staged value is read, and a patch that adds
            // nothing adds +0.0 (a sum that starts at +0.0 never becomes -0.0,
            // so adding +0.0 leaves it bit-identical)
#pragma unroll 2
            for (long long pp = pa; pp <= pb; ++pp) {
                const int k = (int)(pp - q);
                const int sl = S.slot[k], kind = S.kind[k];
                const double v0 = S.dg[k], v1 = S.soil[k], v2 = S.h2o[k];
                const double v3 = S.top[k], v4 = S.top_snow[k];
                double sv[GHF_NL];
#pragma unroll
                for (int j = 0; j < GHF_NL; ++j) sv[j] = S.sab[j][k];
                if (sl != -2 && sl != t) bad |= DEV_ERR_SUBGRID;   // names another column
                if (sl == t && my_ktop < 0) bad |= DEV_ERR_SNL;
                const bool take = sl == t && my_ktop >= 0;
                s_dhsdT += take ? v0 : 0.0;
                s_soil += take ? v1 : 0.0;
                s_h2osfc += take ? v2 : 0.0;
                s_top += take ? v3 : 0.0;
                s_top_snow += take ? v4 : 0.0;
#pragma unroll
                for (int j = 0; j < GHF_NL; ++j) {
                    const bool in = take && (kind == 1 ? (j >= my_ktop && j <= nlevsno)
                                                       : j == my_ktop);
                    s_lyr[j] += in ? sv[j] : 0.0;
                }
            }
        }
        __syncthreads();
        cur = nxt;
#pragma unroll
        for (int k = 0; k < GHF_NL; ++k) rows[k] = rows_nxt[k];
    }
    if (c < a.nc) {
        const int64_t ldc = a.ldc;
        a.dhsdT[c] = s_dhsdT;
        a.hs_soil[c] = s_soil;
        a.hs_h2osfc[c] = s_h2osfc;
        a.hs_top[c] = s_top;
        a.hs_top_snow[c] = s_top_snow;
#pragma unroll
        for (int k = 0; k < GHF_NL; ++k)
            if (k <= nlevsno) a.sabg_lyr_col[c + (int64_t)k * ldc] = s_lyr[k];
    }
    if (bad) atomicOr(P.err, bad);
}

// ---- soil energy balance error (soil_fluxes.jl:380-447, 499-525)

__device__ __forceinline__ bool is_wall_or_roof(int64_t it) {
    return it == 72 || it == 73 || it == 71;   // ICOL_SUNWALL, ICOL_SHADEWALL, ICOL_ROOF
}

// The column walks its patches in groups of SEB_G: for each group the base
// terms are formed per patch, then the column's layers are read once and each
// layer's stored heat is subtracted from every patch of the group that takes
// that pass, in ascending layer order as the reference does per patch. Layer
// data is thus read once per group instead of once per patch.
constexpr int SEB_G = 8;

__global__ void __launch_bounds__(256) seb_kernel(clmsoil_seb_args a, int nlevsno, int nlevgrnd,
                                                  double dtime, unsigned *err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    unsigned bad = 0u;
    const int64_t p0 = a.patchi[c], p1 = a.patchf[c];
    const bool has = p0 >= 1 && p1 >= p0 && p1 <= a.np;
    if (!has && p0 != -9999 && !(p0 >= 1 && p1 == p0 - 1)) bad |= DEV_ERR_SUBGRID;
    const int64_t snl = a.snl[c];
    const bool snl_ok = snl >= -nlevsno && snl <= 0;
    const bool mc = a.mask_nolakec[c];
    const bool wall = is_wall_or_roof(a.col_itype[c]);
    const int jmax = wall ? a.nlevurb : nlevgrnd;
    const int jend = a.nlevurb > nlevgrnd ? a.nlevurb : nlevgrnd;
    const int64_t ld = a.ldc;
    double acc = 0.0;
    if (has) {
        // base-kernel terms of the column (:387-404)
        const double xmf = a.xmf[c], xmfh = a.xmf_h2osfc[c];
        const double h2o = a.frac_h2osfc[c] * (a.t_h2osfc[c] - a.t_h2osfc_bef[c]) *
                           (a.c_h2osfc[c] / dtime);
        const double to_snow = a.eflx_h2osfc_to_snow[c];
        const double bld = wall ? a.eflx_building_heat_errsoi[c] : 0.0;
        const double fs = a.frac_sno_eff[c];
        for (int64_t g0 = p0 - 1; g0 < p1; g0 += SEB_G) {
            double e[SEB_G];
            bool act[SEB_G], wr[SEB_G], ok[SEB_G];
            bool any = false;
#pragma unroll
            for (int i = 0; i < SEB_G; ++i) {
                const int64_t p = g0 + i;
                ok[i] = p < p1 && a.patch_column[p] == c + 1;
                if (p < p1 && !ok[i]) bad |= DEV_ERR_SUBGRID;
                e[i] = 0.0;
                act[i] = wr[i] = false;
                if (ok[i]) {
                    const bool mp = a.mask_nolakep[p], mu = a.mask_urbanp[p];
                    e[i] = a.errsoi_patch[p];
                    if (mp) {
                        e[i] = a.eflx_soil_grnd[p] - xmf - xmfh - h2o;
                        e[i] += to_snow;
                        if (wall) e[i] += bld;
                    }
                    act[i] = (mp && !wall) || (mu && wall);
                    wr[i] = mp || act[i];
                    if (act[i] && !snl_ok) {
                        bad |= DEV_ERR_SNL;
                        act[i] = false;
                    }
                    any = any || act[i];
                }
            }
            // _soilflux_errsoi_layers_kernel! (:425-447), layers snl+1..jmax. The
            // loop runs over the same levels in every lane (snl and jmax differ per
            // column) so that each load of the wave reads one row: coalesced.
            if (any) {
                for (int jf = -nlevsno + 1; jf <= jend; ++jf) {
                    if (jf < snl + 1 || jf > jmax) continue;
                    const int64_t o = c + (int64_t)(jf + nlevsno - 1) * ld;
                    const double dt = a.t_soisno[o] - a.t_ssbef[o];
                    const double term = jf < 1 ? fs * dt / a.fact[o] : dt / a.fact[o];
#pragma unroll
                    for (int i = 0; i < SEB_G; ++i)
                        if (act[i]) e[i] -= term;
                }
            }
#pragma unroll
            for (int i = 0; i < SEB_G; ++i) {
                if (!ok[i]) continue;
                const int64_t p = g0 + i;
                if (wr[i]) a.errsoi_patch[p] = e[i];
                if (mc) acc += e[i] * a.wtcol[p];
            }
        }
    }
    if (mc) a.errsoi_col[c] = acc;
    if (bad) atomicOr(err, bad);
}

}  // namespace

hipError_t launch_soil_energy_balance(const clmsoil_seb_args &a, int nlevsno, int nlevgrnd,
                                      double dtime, unsigned *err, hipStream_t stream) {
    const int block = 256;
    const int64_t grid = (a.nc + block - 1) / block;
    hipLaunchKernelGGL(seb_kernel, dim3((unsigned)grid), dim3(block), 0, stream, a, nlevsno,
                       nlevgrnd, dtime, err);
    return hipGetLastError();
}

hipError_t launch_ground_heat_flux(const GhfParams &P, int, hipStream_t stream) {
    if (P.a.nc == 0) return hipSuccess;
    const int64_t grid = (P.a.nc + GHF_B - 1) / GHF_B;
    hipLaunchKernelGGL(ghf_kernel, dim3((unsigned)grid), dim3(GHF_B), 0, stream, P);
    return hipGetLastError();
}

}  // namespace clmsoil
