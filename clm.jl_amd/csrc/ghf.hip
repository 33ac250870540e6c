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

__device__ __forceinline__ double urban_gnet(const clmsoil_ghf_args &a, int64_t p, int64_t c,
                                             const Urban &u) {
    return a.sabg[p] + a.dlrad[p] - a.eflx_lwrad_net[p] -
           (a.eflx_sh_grnd[p] + a.qflx_evap_soi[p] * a.htvp[c] + a.qflx_tran_veg[p] * HVAP) +
           u.waste + u.ac + u.traffic + u.vent;
}

__device__ __forceinline__ double rural_base(const clmsoil_ghf_args &a, int64_t p, int64_t c) {
    return a.dlrad[p] + (1.0 - (double)a.frac_veg_nosno[p]) * a.emg[c] * a.forc_lwrad[c];
}

__device__ __forceinline__ bool snl_in_range(const clmsoil_ghf_args &a, int64_t c, int nlevsno) {
    const int64_t s = a.snl[c];
    return s >= -nlevsno && s <= 0;
}

__global__ void __launch_bounds__(256) ghf_kernel(GhfParams P) {
    const clmsoil_ghf_args &a = P.a;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nlevsno = P.nlevsno, joff = P.nlevsno;
    unsigned bad = 0u;

    // ---- patch role: per-patch outputs (:1008-1023, :1051-1080)
    if (t < a.np && a.mask_nolakep[t]) {
        const int64_t p = t;
        const int64_t c = a.patch_column[p] - 1, l = a.patch_landunit[p] - 1;
        if (c < 0 || c >= a.nc || l < 0 || l >= a.nl) {
            bad |= DEV_ERR_SUBGRID;
        } else if (!a.urbpoi[l]) {
            const bool ok = snl_in_range(a, c, nlevsno);
            const Emit e = column_emission(a, c, joff, ok);
            const double base = rural_base(a, p, c);
            a.eflx_gnet[p] = a.sabg[p] + base - e.lw -
                             (a.eflx_sh_grnd[p] + a.qflx_evap_soi[p] * a.htvp[c]);
            a.sabg_chk[p] = a.frac_sno_eff[c] * a.sabg_snow[p] +
                            (1.0 - a.frac_sno_eff[c]) * a.sabg_soil[p];
            a.dgnetdT[p] = -a.cgrnd[p] - e.dlw;
        } else {
            const bool ok = snl_in_range(a, c, nlevsno);
            const Emit e = column_emission(a, c, joff, ok);
            const Urban u = urban_terms(a, p, l, a.col_itype[c]);
            const int64_t it = a.col_itype[c];
            a.eflx_wasteheat[p] = u.waste;
            if (a.building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE ||
                a.building_temp_method == CLMSOIL_BUILDING_TEMP_PROG ||
                !(it == ICOL_ROAD_PERV || it == ICOL_ROAD_IMPERV))
                a.eflx_ventilation[p] = u.vent;
            a.eflx_heat_from_ac[p] = u.ac;
            a.eflx_traffic[p] = u.traffic;
            a.eflx_gnet[p] = urban_gnet(a, p, c, u);
            if (a.building_temp_method == CLMSOIL_BUILDING_TEMP_SIMPLE)
                a.eflx_anthro[p] = u.waste + u.traffic;
            a.dgnetdT[p] = -a.cgrnd[p] - e.dlw;
        }
    }

    // ---- column role: emission and the column's sums over its patches
    if (t < a.nc) {
        const int64_t c = t;
        const int64_t ldc = a.ldc, ldp = a.ldp;
        const bool ok = snl_in_range(a, c, nlevsno);
        if (a.mask_nolakec[c] && !ok) bad |= DEV_ERR_SNL;
        const Emit e = column_emission(a, c, joff, ok);
        double s_dhsdT = 0.0, s_soil = 0.0, s_h2osfc = 0.0, s_top = 0.0, s_top_snow = 0.0;
        double s_lyr[MAX_NLEVSNO + 1];
#pragma unroll
        for (int k = 0; k <= MAX_NLEVSNO; ++k) s_lyr[k] = 0.0;
        // patchi = ISPVAL (-9999): no patches; patchf = patchi - 1: an empty range
        const int64_t p0 = a.patchi[c], p1 = a.patchf[c];
        const bool has = p0 >= 1 && p1 >= p0 && p1 <= a.np;
        if (!has && p0 != -9999 && !(p0 >= 1 && p1 == p0 - 1)) bad |= DEV_ERR_SUBGRID;
        const int64_t lyr_top = ok ? a.snl[c] + 1 : 1;
        for (int64_t p = has ? p0 - 1 : 0; has && p < p1; ++p) {
            if (a.patch_column[p] != c + 1) {
                bad |= DEV_ERR_SUBGRID;
                continue;
            }
            if (!a.mask_nolakep[p]) continue;
            const int64_t l = a.patch_landunit[p] - 1;
            if (l < 0 || l >= a.nl || !ok) {
                bad |= (l < 0 || l >= a.nl) ? DEV_ERR_SUBGRID : DEV_ERR_SNL;
                continue;
            }
            const double wt = a.wtcol[p];
            const bool wnz = wt != 0.0;
            if (!a.urbpoi[l]) {
                // _gnet_patch_rural_contrib_kernel! (:1025-1049)
                const double base = rural_base(a, p, c);
                const double dg = -a.cgrnd[p] - e.dlw;
                const double gsoil = a.sabg_soil[p] + base - e.lw_soil -
                                     (a.eflx_sh_soil[p] + a.qflx_ev_soil[p] * a.htvp[c]);
                const double gh2o = a.sabg_soil[p] + base - e.lw_h2osfc -
                                    (a.eflx_sh_h2osfc[p] + a.qflx_ev_h2osfc[p] * a.htvp[c]);
                s_dhsdT += wnz ? dg * wt : 0.0;
                s_soil += wnz ? gsoil * wt : 0.0;
                s_h2osfc += wnz ? gh2o * wt : 0.0;
                // _gnet_snicar_kernel!, rural branch (:1142-1154)
                const double sab = a.sabg_lyr_patch[p + (lyr_top + joff - 1) * ldp];
                const double gtop = sab + base - e.lw -
                                    (a.eflx_sh_grnd[p] + a.qflx_evap_soi[p] * a.htvp[c]);
                s_top += gtop * wt;
                const double gsnow = sab + base - e.lw_snow -
                                     (a.eflx_sh_snow[p] + a.qflx_ev_snow[p] * a.htvp[c]);
                s_top_snow += gsnow * wt;
#pragma unroll
                for (int k = 0; k <= MAX_NLEVSNO; ++k) {
                    // level k + 1 of sabg_lyr_* is snow layer j = k + 1 - nlevsno
                    if (k <= nlevsno && k + 1 - nlevsno >= lyr_top)
                        s_lyr[k] += a.sabg_lyr_patch[p + (int64_t)k * ldp] * wt;
                }
            } else {
                // _gnet_patch_urban_kernel! contribution (:1080-1084) and the urban
                // branch of _gnet_snicar_kernel! (:1155-1160)
                const Urban u = urban_terms(a, p, l, a.col_itype[c]);
                const double g = urban_gnet(a, p, c, u);
                const double dg = -a.cgrnd[p] - e.dlw;
                s_dhsdT += wnz ? dg * wt : 0.0;
                s_soil += wnz ? g * wt : 0.0;
                s_h2osfc += wnz ? g * wt : 0.0;
                s_top += g * wt;
                s_top_snow += g * wt;
                const double sw = a.sabg[p] * wt;
#pragma unroll
                for (int k = 0; k <= MAX_NLEVSNO; ++k)
                    if (k == lyr_top + nlevsno - 1) s_lyr[k] += sw;
            }
        }
        a.dhsdT[c] = s_dhsdT;
        a.hs_soil[c] = s_soil;
        a.hs_h2osfc[c] = s_h2osfc;
        a.hs_top[c] = s_top;
        a.hs_top_snow[c] = s_top_snow;
#pragma unroll
        for (int k = 0; k <= MAX_NLEVSNO; ++k)
            if (k <= nlevsno) a.sabg_lyr_col[c + (int64_t)k * ldc] = s_lyr[k];
    }
    if (bad) atomicOr(P.err, bad);
}

// ---- soil energy balance error (soil_fluxes.jl:380-447, 499-525)

__device__ __forceinline__ bool is_wall_or_roof(int64_t it) {
    return it == 72 || it == 73 || it == 71;   // ICOL_SUNWALL, ICOL_SHADEWALL, ICOL_ROOF
}

// errsoi of patch p (column c) after the base kernel and both layer passes, in
// the reference's statement order; `e` enters as the patch's prior value.
__device__ __forceinline__ double patch_errsoi(const clmsoil_seb_args &a, int64_t p, int64_t c,
                                               int nlevsno, int nlevgrnd, double dtime,
                                               double e, bool &written) {
    const bool mp = a.mask_nolakep[p], wall = is_wall_or_roof(a.col_itype[c]);
    const bool mu = a.mask_urbanp[p];
    written = mp || (mu && wall);
    if (mp) {
        e = a.eflx_soil_grnd[p] - a.xmf[c] - a.xmf_h2osfc[c] -
            a.frac_h2osfc[c] * (a.t_h2osfc[c] - a.t_h2osfc_bef[c]) * (a.c_h2osfc[c] / dtime);
        e += a.eflx_h2osfc_to_snow[c];
        if (wall) e += a.eflx_building_heat_errsoi[c];
    }
    const bool pass = (mp && !wall) || (mu && wall);
    if (pass) {
        const int jmax = wall ? a.nlevurb : nlevgrnd;
        const int64_t ld = a.ldc;
        const int64_t snl = a.snl[c];
        for (int jf = -nlevsno + 1; jf <= jmax; ++jf) {
            const int64_t o = c + (int64_t)(jf + nlevsno - 1) * ld;
            const double dt = a.t_soisno[o] - a.t_ssbef[o];
            if (jf >= snl + 1 && jf < 1) e -= a.frac_sno_eff[c] * dt / a.fact[o];
            if (jf >= 1) e -= dt / a.fact[o];
        }
    }
    return e;
}

__global__ void __launch_bounds__(256) seb_kernel(clmsoil_seb_args a, int nlevsno, int nlevgrnd,
                                                  double dtime, unsigned *err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    unsigned bad = 0u;
    const int64_t p0 = a.patchi[c], p1 = a.patchf[c];
    const bool has = p0 >= 1 && p1 >= p0 && p1 <= a.np;
    if (!has && p0 != -9999 && !(p0 >= 1 && p1 == p0 - 1)) bad |= DEV_ERR_SUBGRID;
    const int64_t s = a.snl[c];
    const bool snl_ok = s >= -nlevsno && s <= 0;
    const bool mc = a.mask_nolakec[c];
    double acc = 0.0;
    for (int64_t p = has ? p0 - 1 : 0; has && p < p1; ++p) {
        if (a.patch_column[p] != c + 1) {
            bad |= DEV_ERR_SUBGRID;
            continue;
        }
        bool written = false;
        double e = a.errsoi_patch[p];
        if (snl_ok) {
            e = patch_errsoi(a, p, c, nlevsno, nlevgrnd, dtime, e, written);
        } else if (a.mask_nolakep[p] || a.mask_urbanp[p]) {
            bad |= DEV_ERR_SNL;
        }
        if (written) a.errsoi_patch[p] = e;
        if (mc) acc += e * a.wtcol[p];
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

hipError_t launch_ground_heat_flux(const GhfParams &P, int block, hipStream_t stream) {
    if (block != 64 && block != 128 && block != 256) block = 256;
    const int64_t n = P.a.np > P.a.nc ? P.a.np : P.a.nc;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(ghf_kernel, dim3((unsigned)grid), dim3(block), 0, stream, P);
    return hipGetLastError();
}

}  // namespace clmsoil
