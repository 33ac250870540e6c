// hydro.hip — soil water fractions and the water table (gfx950), one lane per column.
//
// set_soil_water_fractions! (src/biogeophys/soil_hydrology.jl:110-136),
// water_table! (:676-893), and the moisture-profile water tables of the
// non-aquifer configuration: perched_water_table! (:1522-1653),
// theta_based_water_table! (:1670-1765) and clamp_zwt_to_bedrock!
// (src/infrastructure/kernels.jl:209-227). The layer loops run over the same layers in every lane,
// so each load of a wavefront reads one row of the column-major arrays; the water
// table's layer-by-layer moves are short data-dependent loops per column.
#include "common.hpp"
#include "hydro.hpp"

namespace clmsoil {
namespace {

__device__ __forceinline__ double at(const double *p, int64_t c, int j, int64_t ld) {
    return p[c + (int64_t)(j - 1) * ld];
}

__global__ void __launch_bounds__(256) fractions_kernel(clmsoil_fractions_args a, int nlevsno,
                                                        int nlevsoi) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    if (a.mask && !a.mask[c]) return;
    const int64_t ld = a.ld;
    for (int j = 1; j <= nlevsoi; ++j) {
        const double exc = a.excess_ice ? at(a.excess_ice, c, j, ld) : 0.0;
        const double ws = at(a.watsat, c, j, ld);
        const double dz_ext = at(a.dz, c, j + nlevsno, ld) + exc / DENICE;
        const double vol_ice = jl_min(ws, (at(a.h2osoi_ice, c, j + nlevsno, ld) + exc) /
                                              (dz_ext * DENICE));
        a.eff_porosity[c + (int64_t)(j - 1) * ld] = jl_max(0.01, ws - vol_ice);
        a.icefrac[c + (int64_t)(j - 1) * ld] = jl_min(1.0, vol_ice / ws);
    }
}

// analytical specific yield of layer j for a water table at zwt (:691-693)
__device__ __forceinline__ double spec_yield(const clmsoil_water_table_args &a, int64_t c, int j,
                                             double zwt) {
    const int64_t ld = a.ld;
    return at(a.watsat, c, j, ld) *
           (1.0 - pow(1.0 + 1.0e3 * zwt / at(a.sucsat, c, j, ld), -1.0 / at(a.bsw, c, j, ld)));
}

__global__ void __launch_bounds__(256) water_table_kernel(clmsoil_water_table_args a, int nlevsno,
                                                          int nlevsoi, double dtime, double aqy,
                                                          int frost) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    if (a.mask && !a.mask[c]) return;
    const int64_t ld = a.ld;
    const int n = nlevsoi, joff = nlevsno, joff_zi = nlevsno + 1;
    a.qflx_drain[c] = 0.0;
    a.qflx_rsub_sat[c] = 0.0;
    a.qflx_drain_perched[c] = 0.0;
    double zwt = a.zwt[c];
    int jwt = n;
    for (int j = 1; j <= n; ++j)
        if (zwt <= at(a.zi, c, joff_zi + j, ld)) { jwt = j - 1; break; }
    // aquifer recharge (:689-746)
    double rous = spec_yield(a, c, n, zwt);
    rous = jl_max(rous, aqy);
    const double qc = a.qcharge[c];
    if (jwt == n) {
        a.wa[c] = a.wa[c] + qc * dtime;
        zwt = zwt - (qc * dtime) / 1000.0 / rous;
    } else {
        double qtot = qc * dtime;
        if (qtot > 0.0) {                              // rising water table
            for (int j = jwt + 1; j >= 1; --j) {
                double s_y = jl_max(spec_yield(a, c, j, zwt), aqy);
                double ql = jl_min(qtot, s_y * (zwt - at(a.zi, c, joff_zi + j - 1, ld)) * 1.0e3);
                ql = jl_max(ql, 0.0);
                if (s_y > 0.0) zwt = zwt - ql / s_y / 1000.0;
                qtot = qtot - ql;
                if (qtot <= 0.0) break;
            }
        } else {                                       // deepening water table
            for (int j = jwt + 1; j <= n; ++j) {
                double s_y = jl_max(spec_yield(a, c, j, zwt), aqy);
                double ql = jl_max(qtot, -(s_y * (at(a.zi, c, joff_zi + j, ld) - zwt) * 1.0e3));
                ql = jl_min(ql, 0.0);
                qtot = qtot - ql;
                if (qtot >= 0.0) {
                    zwt = zwt - ql / s_y / 1000.0;
                    break;
                }
                zwt = at(a.zi, c, joff_zi + j, ld);
            }
            if (qtot > 0.0) zwt = zwt - qtot / 1000.0 / rous;
        }
    }
    a.zwt[c] = zwt;
    if (!frost) return;
    // frost table and perched water table (:757-823)
    int k_frz = at(a.t_soisno, c, joff + 1, ld) > TFRZ ? n : 1;
    for (int k = 2; k <= n; ++k)
        if (at(a.t_soisno, c, joff + k - 1, ld) > TFRZ && at(a.t_soisno, c, joff + k, ld) <= TFRZ) {
            k_frz = k;
            break;
        }
    const double frost_table = at(a.z, c, joff + k_frz, ld);
    a.frost_table[c] = frost_table;
    double zwt_perched = frost_table;
    const bool frozen = at(a.t_soisno, c, joff + k_frz, ld) <= TFRZ;
    if (!(zwt < frost_table && frozen)) {
        const double sat_lev = 0.9;
        int k_perch = 1;
        for (int k = k_frz; k >= 1; --k) {
            const double dzk = at(a.dz, c, joff + k, ld);
            const double vol = at(a.h2osoi_liq, c, joff + k, ld) / (dzk * DENH2O) +
                               at(a.h2osoi_ice, c, joff + k, ld) / (dzk * DENICE);
            a.h2osoi_vol[c + (int64_t)(k - 1) * ld] = vol;
            if (vol / at(a.watsat, c, k, ld) <= sat_lev) { k_perch = k; break; }
        }
        if (!frozen) k_perch = k_frz;
        if (k_frz > k_perch) {
            const int kp = k_perch;
            const double d1 = at(a.dz, c, joff + kp, ld), d2 = at(a.dz, c, joff + kp + 1, ld);
            const double s1 = (at(a.h2osoi_liq, c, joff + kp, ld) / (d1 * DENH2O) +
                               at(a.h2osoi_ice, c, joff + kp, ld) / (d1 * DENICE)) /
                              at(a.watsat, c, kp, ld);
            const double s2 = (at(a.h2osoi_liq, c, joff + kp + 1, ld) / (d2 * DENH2O) +
                               at(a.h2osoi_ice, c, joff + kp + 1, ld) / (d2 * DENICE)) /
                              at(a.watsat, c, kp + 1, ld);
            const double zp1 = at(a.z, c, joff + kp + 1, ld);
            const double m = (zp1 - at(a.z, c, joff + kp, ld)) / (s2 - s1);
            const double b = zp1 - m * s2;
            zwt_perched = jl_max(0.0, m * sat_lev + b);
        }
    }
    a.zwt_perched[c] = zwt_perched;
}

// Volumetric water of soil layer k (the h2osoi_vol the perched and theta-based
// searches store, soil_hydrology.jl:1552-1553 and 1688-1689).
__device__ __forceinline__ double layer_vol(const clmsoil_theta_wt_args &a, int64_t c, int k,
                                            int joff) {
    const int64_t ld = a.ld;
    const double dzk = at(a.dz, c, joff + k, ld);
    return at(a.h2osoi_liq, c, joff + k, ld) / (dzk * DENH2O) +
           at(a.h2osoi_ice, c, joff + k, ld) / (dzk * DENICE);
}

// Depth where the line through (z1, s1) and (z2, s2) reaches 90 % saturation,
// floored at 0 (soil_hydrology.jl:1575-1577 and 1706-1708).
__device__ __forceinline__ double sat_crossing(double s1, double s2, double z1, double z2) {
    const double m = (z2 - z1) / (s2 - s1);
    const double b = z2 - m * s2;
    return jl_max(0.0, m * 0.9 + b);
}

// OPS: 1 perched water table (:1522-1587), 2 theta-based water table
// (:1670-1715), 4 clamp to bedrock (kernels.jl:209-218), applied in that order.
//
// The reference's searches stop at a different layer in every column. Here the
// layer loops are the same in every lane of a wavefront (a lane takes part
// while its own search is running, and the loop ends when no lane's is), so
// each load of a wavefront reads one row of the column-major arrays. One
// top-down sweep of t_soisno finds the frost table; one bottom-up sweep of the
// water contents serves both the perched search (from the frost table up) and
// the theta-based one (from bedrock up), which store the same h2osoi_vol.
template <int OPS>
__global__ void __launch_bounds__(256) theta_wt_kernel(clmsoil_theta_wt_args a, int nlevsno,
                                                       int nlevsoi, unsigned *err) {
    constexpr bool PERCH = (OPS & 1) != 0, THETA = (OPS & 2) != 0, CLAMP = (OPS & 4) != 0;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    if (a.mask && !a.mask[c]) return;
    const int64_t ld = a.ld;
    const int n = nlevsoi, joff = nlevsno, joff_zi = nlevsno + 1;
    const double sat_lev = 0.9;
    int nbed = 0;                                      // 0: theta-based step and clamp skipped
    if constexpr (THETA || CLAMP) {
        const int64_t nb = a.nbedrock[c];
        if (nb < 1 || nb > n) {
            atomicOr(err, (unsigned)DEV_ERR_NBEDROCK);
            if constexpr (!PERCH) return;
        } else {
            nbed = (int)nb;
        }
    }
    // frost table: the first frozen layer under an unfrozen one (:1533-1545)
    int k_frz = 0;
    double t_frz = 0.0, ft = 0.0;
    bool perch = false;
    if constexpr (PERCH) {
        double t_prev = at(a.t_soisno, c, joff + 1, ld);
        const bool top_warm = t_prev > TFRZ;
        k_frz = top_warm ? n : 1;
        t_frz = t_prev;
        bool found = false;
        for (int k = 2; k <= n; ++k) {
            if (!__ballot(!found)) break;
            if (!found) {
                const double tk = at(a.t_soisno, c, joff + k, ld);
                if (t_prev > TFRZ && tk <= TFRZ) {
                    k_frz = k;
                    t_frz = tk;
                    found = true;
                } else if (k == n && top_warm) {
                    t_frz = tk;                        // k_frz = nlevsoi
                }
                t_prev = tk;
            }
        }
        ft = at(a.zi, c, joff_zi + k_frz - 1, ld);    // top of the frozen layer
        perch = !(a.zwt[c] < ft && t_frz <= TFRZ) && k_frz > 1;
    }
    // bottom-up searches for the deepest layer at or below 90 % saturation
    bool pdone = !perch, tdone = !(THETA && nbed >= 1);
    int k_perch = 1, k_zwt = 1;
    double ps1 = 0.0, ps2 = 0.0, ts1 = 0.0, ts2 = 0.0, s_prev = 0.0;
    for (int k = n; k >= 1; --k) {
        if (!__ballot(!pdone || !tdone)) break;
        const bool pa = !pdone && k <= k_frz;
        const bool ta = !tdone && k <= nbed;
        if (pa || ta) {
            const double vol = layer_vol(a, c, k, joff);
            a.h2osoi_vol[c + (int64_t)(k - 1) * ld] = vol;
            const double sk = vol / at(a.watsat, c, k, ld);
            const bool dry = sk <= sat_lev;
            if (pa && (dry || k == 1)) {
                k_perch = k;
                pdone = true;
                ps1 = sk;
                ps2 = s_prev;
            }
            if (ta && dry) {
                k_zwt = k;
                tdone = true;
                ts1 = sk;
                ts2 = s_prev;
            }
            s_prev = sk;
        }
    }
    if constexpr (PERCH) {
        a.frost_table[c] = ft;
        double zp = ft;
        if (perch) {
            if (!(t_frz <= TFRZ)) k_perch = k_frz;
            if (k_frz > k_perch)
                zp = ps1 > ps2 ? at(a.zi, c, joff_zi + k_perch - 1, ld)
                               : sat_crossing(ps1, ps2, at(a.z, c, joff + k_perch, ld),
                                              at(a.z, c, joff + k_perch + 1, ld));
        }
        a.zwt_perched[c] = zp;
    }
    if (nbed < 1) return;
    if constexpr (THETA) {
        double zwt;
        if (k_zwt == 1)
            zwt = at(a.zi, c, joff_zi + 1, ld);
        else if (k_zwt < nbed)
            zwt = sat_crossing(ts1, ts2, at(a.z, c, joff + k_zwt, ld),
                               at(a.z, c, joff + k_zwt + 1, ld));
        else
            zwt = at(a.zi, c, joff_zi + nbed, ld);
        a.zwt[c] = zwt;
    }
    if constexpr (CLAMP) {
        const double zb = at(a.zi, c, joff_zi + nbed, ld);
        if (a.zwt[c] > zb) a.zwt[c] = zb;
    }
}

}  // namespace

hipError_t launch_theta_water_table(const clmsoil_theta_wt_args &a, int nlevsno, int nlevsoi,
                                    int ops, unsigned *err, hipStream_t stream) {
    const int block = 256;
    const dim3 grid((unsigned)((a.nc + block - 1) / block));
    switch (ops) {
    case 1: hipLaunchKernelGGL((theta_wt_kernel<1>), grid, dim3(block), 0, stream, a, nlevsno, nlevsoi, err); break;
    case 2: hipLaunchKernelGGL((theta_wt_kernel<2>), grid, dim3(block), 0, stream, a, nlevsno, nlevsoi, err); break;
    case 4: hipLaunchKernelGGL((theta_wt_kernel<4>), grid, dim3(block), 0, stream, a, nlevsno, nlevsoi, err); break;
    case 7: hipLaunchKernelGGL((theta_wt_kernel<7>), grid, dim3(block), 0, stream, a, nlevsno, nlevsoi, err); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_soil_water_fractions(const clmsoil_fractions_args &a, int nlevsno, int nlevsoi,
                                       hipStream_t stream) {
    const int block = 256;
    const int64_t grid = (a.nc + block - 1) / block;
    hipLaunchKernelGGL(fractions_kernel, dim3((unsigned)grid), dim3(block), 0, stream, a, nlevsno,
                       nlevsoi);
    return hipGetLastError();
}

hipError_t launch_water_table(const clmsoil_water_table_args &a, int nlevsno, int nlevsoi,
                              double dtime, double aqy, int frost, hipStream_t stream) {
    const int block = 256;
    const int64_t grid = (a.nc + block - 1) / block;
    hipLaunchKernelGGL(water_table_kernel, dim3((unsigned)grid), dim3(block), 0, stream, a,
                       nlevsno, nlevsoi, dtime, aqy, frost);
    return hipGetLastError();
}

}  // namespace clmsoil
