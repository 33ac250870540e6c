// surfhyd.hip — the surface hydrology before the Richards step (gfx950): flood
// water, saturated and infiltration excess runoff, the surface water store,
// infiltration, total surface runoff and urban ponding (clm_driver.jl:1939-2012).
//
// The reference launches one kernel per step over all columns; every step reads
// and writes only its own column, so here one lane runs a column through all
// requested steps in the reference's order and the intermediate fluxes stay in
// registers (each is still stored, as the reference's outputs are). Every
// per-column array is read with one coalesced load per wavefront.
#include "common.hpp"
#include "surfhyd.hpp"

namespace clmsoil {
namespace {

// landunit_varcon.jl:7-8, column_varcon.jl:7-11, varcon.jl:85, infilt_excess_runoff.jl:38
constexpr int64_t ISTSOIL = 1, ISTCROP = 2;
constexpr int64_t ICOL_ROOF = 71, ICOL_SUNWALL = 72, ICOL_SHADEWALL = 73, ICOL_ROAD_IMPERV = 74;
constexpr int64_t ISPVAL = -9999;
constexpr double QINMAX_UNLIMITED = 1.0e200;
constexpr double DEG2RAD = 3.141592653589793 / 180.0;    // Julia's pi / 180.0 in Float64

__device__ __forceinline__ bool on(const uint8_t *m, int64_t c) { return !m || m[c]; }

__global__ void __launch_bounds__(256) surfhyd_kernel(clmsoil_surfhyd_args a,
                                                      clmsoil_surfhyd_params p, int nlevsno,
                                                      double dtime) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    const int s = p.steps;
    const bool hyd = on(a.mask_hydrologyc, c);
    // set_floodc! (soil_hydrology.jl:139-152)
    if ((s & CLMSOIL_SH_FLOODC) && on(a.mask_nolakec, c)) {
        const int64_t it = a.col_itype[c];
        a.qflx_floodc[c] = (it != ICOL_SUNWALL && it != ICOL_SHADEWALL)
                               ? a.qflx_floodg[a.col_gridcell[c] - 1]
                               : 0.0;
    }
    if (hyd) {
        // saturated_excess_runoff! (sat_excess_runoff.jl:246-300)
        if (s & CLMSOIL_SH_SAT_EXCESS) {
            double fs;
            if (p.fsat_method == CLMSOIL_FSAT_TOPMODEL) {
                const double ft = a.frost_table[c], zw = a.zwt[c], zwp = a.zwt_perched[c];
                fs = a.wtfact[c] * exp(-0.5 * p.fff * ((ft > zwp && ft <= zw) ? zwp : zw));
            } else {
                const double bi = a.b_infil[c];
                const double ex = bi / (1.0 + bi);
                fs = 1.0 - pow(1.0 - a.top_moist_limited[c] / a.top_max_moist[c], ex);
            }
            if (p.crop_fsat_equals_zero && a.lun_itype[a.col_landunit[c] - 1] == ISTCROP)
                fs = 0.0;
            if (p.hillslope_fsat_equals_zero && a.is_hillslope_column[c] && a.col_active[c] &&
                a.col_cold[c] != ISPVAL)
                fs = 0.0;
            a.fsat[c] = fs;
            a.fcov[c] = fs;
            double q = fs * a.qflx_rain_plus_snomelt[c];
            if (a.col_urbpoi[c]) q = q + a.qflx_floodc[c];
            a.qflx_sat_excess_surf[c] = q;
        }
        // set_qflx_inputs! (soil_hydrology.jl:160-187)
        if (s & CLMSOIL_SH_QFLX_INPUTS) {
            const double top = a.qflx_rain_plus_snomelt[c] + a.qflx_snow_h2osfc[c] +
                               a.qflx_floodc[c];
            a.qflx_top_soil[c] = top;
            const bool nosnow = a.snl[c] >= 0;
            const double fsno = nosnow ? 0.0 : a.frac_sno_eff[c];
            const double evap = nosnow ? a.qflx_liqevap_from_top_layer[c] : a.qflx_ev_soil[c];
            const double fh = a.frac_h2osfc[c];
            const double avail = top - a.qflx_sat_excess_surf[c];
            a.qflx_in_soil[c] = (1.0 - fh) * avail - (1.0 - fsno - fh) * evap;
            a.qflx_top_soil_to_h2osfc[c] = fh * avail - fh * a.qflx_ev_h2osfc[c];
        }
        // infiltration_excess_runoff! (infilt_excess_runoff.jl:142-211, 241-256)
        if (s & CLMSOIL_SH_INFIL_EXCESS) {
            double qu = QINMAX_UNLIMITED;
            if (p.qinmax_method == CLMSOIL_QINMAX_HKSAT) {
                qu = __builtin_huge_val();
                for (int j = 0; j < 3; ++j) {
                    const int64_t o = c + (int64_t)j * a.ld;
                    const double val = pow(10.0, -p.e_ice * a.icefrac[o]) * a.hksat[o];
                    if (val < qu) qu = val;
                }
            }
            const double qm = (1.0 - a.fsat[c]) * qu;
            a.qinmax[c] = qm;
            a.qflx_infl_excess[c] = jl_max(0.0, a.qflx_in_soil[c] - (1.0 - a.frac_h2osfc[c]) * qm);
        }
        // route_infiltration_excess! (soil_hydrology.jl:199-225)
        if (s & CLMSOIL_SH_ROUTE) {
            const int64_t lt = a.lun_itype[a.col_landunit[c] - 1];
            if (lt == ISTSOIL || lt == ISTCROP) {
                const double ie = a.qflx_infl_excess[c];
                a.qflx_in_soil_limited[c] = a.qflx_in_soil[c] - ie;
                if (p.h2osfcflag != 0) {
                    a.qflx_in_h2osfc[c] = a.qflx_top_soil_to_h2osfc[c] + ie;
                    a.qflx_infl_excess_surf[c] = 0.0;
                } else {
                    a.qflx_in_h2osfc[c] = a.qflx_top_soil_to_h2osfc[c];
                    a.qflx_infl_excess_surf[c] = ie;
                }
            } else {
                a.qflx_in_soil_limited[c] = a.qflx_in_soil[c];
                a.qflx_in_h2osfc[c] = 0.0;
                a.qflx_infl_excess_surf[c] = 0.0;
            }
        }
        // update_h2osfc! (surface_water.jl:279-316, 357-374, 400-443, 462-520)
        if (s & CLMSOIL_SH_H2OSFC) {
            const int flag = p.h2osfcflag_update;
            const double h0 = a.h2osfc[c];
            double surf = 0.0;
            if (flag == 1) {
                const double fn = a.frac_h2osfc_nosnow[c];
                const double frac_infclust = fn <= p.pc ? 0.0 : pow(fn - p.pc, p.mu);
                const double thr = a.h2osfc_thresh[c];
                if (h0 > thr) {
                    const double k_wet = 1.0e-4 * sin(DEG2RAD * a.topo_slope[c]);
                    surf = jl_min(k_wet * frac_infclust * (h0 - thr), (h0 - thr) / dtime);
                }
                if (surf < 1.0e-8) surf = 0.0;
            }
            double h1 = h0 + (a.qflx_in_h2osfc[c] - surf) * dtime;
            if (fabs(h1) < 1.0e-13 * fabs(h0)) h1 = 0.0;
            double drain;
            if (h1 < 0.0) {
                drain = h1 / dtime;
            } else {
                drain = jl_min(a.frac_h2osfc[c] * a.qinmax[c], h1 / dtime);
                if (flag == 0) drain = jl_max(0.0, h1 / dtime);
            }
            double h2 = h1 - drain * dtime;
            if (fabs(h2) < 1.0e-13 * fabs(h1)) h2 = 0.0;
            a.h2osfc[c] = h2;
            a.qflx_h2osfc_surf[c] = surf;
            a.qflx_h2osfc_drain[c] = drain;
        }
        // infiltration! (soil_hydrology.jl:236-243)
        if (s & CLMSOIL_SH_INFILTRATION)
            a.qflx_infl[c] = a.qflx_in_soil_limited[c] + a.qflx_h2osfc_drain[c];
        // total_surface_runoff!, hydrologically active columns (:250-258)
        if (s & CLMSOIL_SH_SURF_RUNOFF)
            a.qflx_surf[c] =
                a.qflx_sat_excess_surf[c] + a.qflx_infl_excess_surf[c] + a.qflx_h2osfc_surf[c];
    }
    if (!on(a.mask_urbanc, c) || !(s & (CLMSOIL_SH_SURF_RUNOFF | CLMSOIL_SH_URBAN_POND))) return;
    const int64_t it = a.col_itype[c];
    const bool pond_col = it == ICOL_ROOF || it == ICOL_ROAD_IMPERV;
    double *pond = a.h2osoi_liq + c + (int64_t)nlevsno * a.ld;   // top soil layer
    // total_surface_runoff!, urban columns (:273-301)
    if (s & CLMSOIL_SH_SURF_RUNOFF) {
        if (pond_col) {
            double q;
            if (a.snl[c] < 0) {
                q = jl_max(0.0, a.qflx_rain_plus_snomelt[c]);
            } else {
                q = jl_max(0.0, *pond / dtime + a.qflx_rain_plus_snomelt[c] -
                                    a.qflx_liqevap_from_top_layer[c] - p.pondmx_urban / dtime);
                a.xs_urban[c] = q;
            }
            a.qflx_surf[c] = q + a.qflx_floodc[c];
        } else if (it == ICOL_SUNWALL || it == ICOL_SHADEWALL) {
            a.qflx_surf[c] = 0.0;
        }
    }
    // update_urban_ponding! (:316-338)
    if ((s & CLMSOIL_SH_URBAN_POND) && pond_col && a.snl[c] >= 0) {
        if (a.xs_urban[c] > 0.0)
            *pond = p.pondmx_urban;
        else
            *pond = jl_max(0.0, *pond + (a.qflx_rain_plus_snomelt[c] -
                                         a.qflx_liqevap_from_top_layer[c]) * dtime);
    }
}

// Root water uptake, default method (soil_water_plant_sink.jl:48-81, 341-451).
//
// A lane owns a column and sums its patches in ascending order with the
// reference's operation order, so the result is bit-identical to the
// reference's loop. The patch transpiration, weight and activity of up to
// RS_KMAX patches sit in registers for all layers. CLM stores a column's
// patches contiguously and neighbouring columns' ranges next to each other, so
// a wavefront's 64 columns cover one short window of the patch arrays: when it
// holds at most RS_WCAP patches, each layer's row of rootr_patch over that
// window is read with coalesced loads into the wavefront's LDS slice (no
// workgroup barrier, the wavefront only waits for itself) and the next
// layer's row is already in flight while the lanes sum the current one. A wider
// window (patch ranges scattered over the arrays) or a column with more than
// RS_KMAX patches reads the same values straight from global memory.
constexpr int RS_BLOCK = 256;
constexpr int RS_KMAX = 16;
constexpr int RS_WCAP = 1024;                 // patches a wavefront stages per layer

__device__ __forceinline__ int64_t wave_min(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__global__ void __launch_bounds__(RS_BLOCK) rootsink_kernel(clmsoil_rootsink_args a, int nlevsoi,
                                                            unsigned *err) {
    __shared__ double s_row[RS_BLOCK / 64][RS_WCAP];
    const int lane = threadIdx.x & 63;
    double *row = s_row[threadIdx.x >> 6];
    const int64_t c = (int64_t)blockIdx.x * RS_BLOCK + threadIdx.x;
    bool run = c < a.nc && on(a.mask_hydrologyc, c);
    int64_t lo = 0, n = 0;
    if (run) {
        lo = a.patchi[c] - 1;
        n = a.npatches[c];
        if (n < 0 || lo < 0 || lo > a.np - n) {
            atomicOr(err, (unsigned)DEV_ERR_SUBGRID);
            run = false;
        }
    }
    const bool has = run && n > 0;
    const int64_t p0 = wave_min(has ? lo : INT64_MAX);
    const int64_t p1 = wave_max(has ? lo + n : INT64_MIN);
    const int64_t w = p0 == INT64_MAX ? 0 : p1 - p0;
    const bool staged = w <= RS_WCAP;            // the same in every lane of the wavefront
    const bool small = n <= RS_KMAX;
    double qt[RS_KMAX], wt[RS_KMAX];
    unsigned actm = 0;
    double temp = 0.0;
    if (run) {
        if (small) {
#pragma unroll
            for (int k = 0; k < RS_KMAX; ++k) {
                qt[k] = 0.0;
                wt[k] = 0.0;
                if (k < n) {
                    qt[k] = a.qflx_tran_veg_patch[lo + k];
                    wt[k] = a.wtcol[lo + k];
                    if (a.active[lo + k]) actm |= 1u << k;
                }
            }
#pragma unroll
            for (int k = 0; k < RS_KMAX; ++k)
                if ((actm >> k) & 1u) temp += qt[k] * wt[k];
        } else {
            for (int64_t k = 0; k < n; ++k)
                if (a.active[lo + k]) temp += a.qflx_tran_veg_patch[lo + k] * a.wtcol[lo + k];
        }
    }
    const double qcol = run ? a.qflx_tran_veg_col[c] : 0.0;
    const int64_t o = lo - p0;
    // staged rows: layer j+1 is loaded into registers while layer j is summed
    constexpr int NPRE = RS_WCAP / 64;
    double pre[NPRE];
    if (staged) {
#pragma unroll
        for (int t = 0; t < NPRE; ++t)
            if (lane + 64 * t < w) pre[t] = a.rootr_patch[p0 + lane + 64 * t];
    }
    for (int j = 0; j < nlevsoi; ++j) {
        const double *root = a.rootr_patch + (int64_t)j * a.ldp;
        if (staged) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < NPRE; ++t)
                if (lane + 64 * t < w) row[lane + 64 * t] = pre[t];
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (j + 1 < nlevsoi) {
                const double *next = root + a.ldp;
#pragma unroll
                for (int t = 0; t < NPRE; ++t)
                    if (lane + 64 * t < w) pre[t] = next[p0 + lane + 64 * t];
            }
        }
        if (!run) continue;
        double acc = 0.0;
        if (small && staged) {
#pragma unroll
            for (int k = 0; k < RS_KMAX; ++k)
                if ((actm >> k) & 1u) acc += row[o + k] * qt[k] * wt[k];
        } else if (small) {
#pragma unroll
            for (int k = 0; k < RS_KMAX; ++k)
                if ((actm >> k) & 1u) acc += root[lo + k] * qt[k] * wt[k];
        } else {
            for (int64_t k = 0; k < n; ++k)
                if (a.active[lo + k])
                    acc += root[lo + k] * a.qflx_tran_veg_patch[lo + k] * a.wtcol[lo + k];
        }
        if (temp != 0.0) acc /= temp;
        a.rootr_col[c + (int64_t)j * a.ldc] = acc;
        a.qflx_rootsoi_col[c + (int64_t)j * a.ldc] = acc * qcol;
    }
}

}  // namespace

hipError_t launch_root_sink(const clmsoil_rootsink_args &a, int nlevsoi, unsigned *err,
                            hipStream_t stream) {
    const dim3 grid((unsigned)((a.nc + RS_BLOCK - 1) / RS_BLOCK));
    hipLaunchKernelGGL(rootsink_kernel, grid, dim3(RS_BLOCK), 0, stream, a, nlevsoi, err);
    return hipGetLastError();
}

hipError_t launch_surface_hydrology(const clmsoil_surfhyd_args &a,
                                    const clmsoil_surfhyd_params &p, int nlevsno, double dtime,
                                    hipStream_t stream) {
    const int block = 256;
    const dim3 grid((unsigned)((a.nc + block - 1) / block));
    hipLaunchKernelGGL(surfhyd_kernel, grid, dim3(block), 0, stream, a, p, nlevsno, dtime);
    return hipGetLastError();
}

}  // namespace clmsoil
