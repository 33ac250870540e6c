// lvt.hip — vertical transport of the soil organic matter pools (gfx950).
//
// litter_vert_transp! (src/biogeochem/litter_vert_transp.jl:479-639) without the
// matrix-CN path. The coefficient kernel runs one lane per column. The transport
// kernel runs one lane per (column, pool, tracer type): blockIdx.y is the pool,
// blockIdx.z the tracer, and consecutive lanes take consecutive columns, so every
// load of a wavefront reads one row of the [c, j, s] arrays. A lane walks its
// column once from the surface ghost level down: the diffusivities and fluxes of
// the next level are read ahead, the Patankar coefficients and the tridiagonal
// row are formed in registers, and the Thomas forward elimination consumes the
// row at once; its multipliers and right-hand sides go to a lane-private LDS
// column ([level][lane]). The back substitution leaves the new concentrations in
// the right-hand-side column, and one last pass writes the tendency and the
// concentrations with the bedrock leak. The arithmetic is the reference's
// operation by operation (Julia's x^5, max/min, no pivot guard).
#include "common.hpp"
#include "lvt.hpp"

namespace clmsoil {
namespace {

constexpr int LV_BLOCK = 64;

// Julia's x^5 for Float64 (Base.Math.pow_body): power by squaring with the
// rounding error of each product carried through fma; muladd fused.
__device__ __forceinline__ double jl_pow5(double x) {
    const double x1 = x * x, e1 = __builtin_fma(x, x, -x1);
    const double err1 = x1 * 2.0 * e1;
    const double x2 = x1 * x1, e2 = __builtin_fma(x1, x1, -x2);
    const double xnlo = e2 + err1;
    const double err = __builtin_fma(x, xnlo, x2 * 0.0);
    if (!(__builtin_isfinite(x2) && __builtin_isfinite(err))) return x2 * x;
    return __builtin_fma(x2, x, err);
}

// patankar_A (:66-70)
__device__ __forceinline__ double patankar_A(double pe) {
    return jl_max(0.0, jl_pow5(1.0 - 0.1 * fabs(pe)));
}

// _lvt_coef_kernel! (:93-136)
__global__ void __launch_bounds__(256) lvt_coef_kernel(clmsoil_lvt_args a, clmsoil_lvt_params p,
                                                       unsigned *err) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc || (a.mask_bgc_soilc && !a.mask_bgc_soilc[c])) return;
    const int nd = p.nlevdecomp;
    const int64_t nbr = a.nbedrock[c];
    if (nbr < 1 || nbr > nd) {
        atomicOr(err, (unsigned)DEV_ERR_NBEDROCK);
        return;
    }
    const double *zisoi = a.zisoi - 1;                   // 1-based: zisoi[1] = surface
    const double alt = jl_max(a.altmax[c], a.altmax_lastyear[c]);
    const bool cryo = alt <= p.max_altdepth_cryoturbation && alt > 0.0;
    const double zdeep = jl_min(p.max_depth_cryoturb, zisoi[nbr + 2]);
    for (int j = 1; j <= nd + 1; ++j) {
        double dif = 0.0, adv = 0.0;
        if (cryo) {
            if (j <= nbr + 1)
                dif = zisoi[j + 1] < alt
                          ? p.cryoturb_diffusion_k
                          : jl_max(p.cryoturb_diffusion_k *
                                       (1.0 - (zisoi[j + 1] - alt) / (zdeep - alt)),
                                   0.0);
        } else if (alt > 0.0 && j <= nbr + 1) {
            adv = p.som_adv_flux;
            dif = p.som_diffus;
        }
        a.som_diffus_coef[c + (int64_t)(j - 1) * a.ld] = dif;
        a.som_adv_coef[c + (int64_t)(j - 1) * a.ld] = adv;
    }
}

// _lvt_column_kernel! (:142-471) for pool blockIdx.y + 1 of tracer blockIdx.z
__global__ void __launch_bounds__(LV_BLOCK) lvt_transport_kernel(clmsoil_lvt_args a,
                                                                 clmsoil_lvt_params p,
                                                                 double dtime) {
    extern __shared__ double lds[];                      // [2][nj][LV_BLOCK]
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * LV_BLOCK + lane;
    const int s = blockIdx.y + 1, t = blockIdx.z;
    if (c >= a.nc || (a.mask_bgc_soilc && !a.mask_bgc_soilc[c])) return;
    const int nd = p.nlevdecomp, nj = nd + 2;
    const int64_t nbr64 = a.nbedrock[c];
    if (nbr64 < 1 || nbr64 > nd) return;                 // flagged by the coefficient kernel
    const int nbr = (int)nbr64;
    const int64_t ld = a.ld;
    const int64_t soff = (int64_t)(s - 1) * ld * p.nlevdecomp_full;
    double *conc = a.conc[t] + soff + c;
    const double *src = a.source[t] + soff + c;
    double *tend = a.tendency[t] + soff + c;
    const double *zisoi = a.zisoi - 1, *zsoi = a.zsoi - 1, *dzd = a.dzsoi_decomp - 1;
    // zsoi_ext / dz_node (:566-575), 1-based
    auto zext = [&](int j) { return j <= nd ? zsoi[j] : zisoi[nd + 1]; };
    auto dzn = [&](int j) { return j == 1 ? zext(1) : zext(j) - zext(j - 1); };
    double *cpl = lds + lane, *dpl = lds + (int64_t)nj * LV_BLOCK + lane;   // [jj-1]*LV_BLOCK
    if (p.is_cwd[s - 1]) {
        // CWD: no transport, the source is added in place (:296-300)
        for (int j = 1; j <= nd; ++j)
            dpl[(int64_t)j * LV_BLOCK] = conc[(int64_t)(j - 1) * ld] + src[(int64_t)(j - 1) * ld];
    } else {
        double spinup_term = 1.0;
        if (p.spinup_state >= 1) spinup_term = p.spinup_factor[s - 1];
        if (fabs(spinup_term - 1.0) > 1.0e-6)
            spinup_term = spinup_term *
                          (1.0 + 50.0 / (1.0 + exp(-0.15 * (fabs(a.latdeg[a.col_gridcell[c] - 1]) - 60.0))));
        const double epsilon = 1.0e-30;
        auto scaled = [&](const double *coef, int j) {
            const double v = coef[c + (int64_t)(j - 1) * ld];
            return fabs(v) * spinup_term < epsilon ? epsilon : v * spinup_term;
        };
        // sliding window over levels: diffusivity at j-1, j, j+1; advection at j, j+1
        double dif_m = 0.0, dif_0 = scaled(a.som_diffus_coef, 1), dif_p = scaled(a.som_diffus_coef, 2);
        double adv_0 = scaled(a.som_adv_coef, 1), adv_p = scaled(a.som_adv_coef, 2);
        // row 0 (the surface ghost): a = 0, b = 1, c = -1, r = 0
        double cp_prev = -1.0 / 1.0, dp_prev = 0.0 / 1.0;
        cpl[0] = cp_prev;
        dpl[0] = dp_prev;
        double conc_next = conc[0];                      // concentration of level 1
        for (int j = 1; j <= nd + 1; ++j) {
            // conc_trcr[j+1] (:309-317): every level's concentration, 0 for the ghost
            // level nd + 1
            const double ct = j <= nd ? conc_next : 0.0;
            if (j < nd) conc_next = conc[(int64_t)j * ld];
            double d_m1_zm1, d_p1_zp1, f_m1, f_p1, pe_m1, pe_p1;
            if (j == 1) {
                d_m1_zm1 = 0.0;
                const double w_p1 = (zext(j + 1) - zisoi[j + 1]) / dzn(j + 1);
                const double d_p1 = (dif_p > 0.0 && dif_0 > 0.0)
                                        ? 1.0 / ((1.0 - w_p1) / dif_0 + w_p1 / dif_p)
                                        : 0.0;
                d_p1_zp1 = d_p1 / dzn(j + 1);
                f_m1 = adv_0;
                f_p1 = adv_p;
                pe_m1 = 0.0;
                pe_p1 = f_p1 / d_p1_zp1;
            } else if (j >= nbr + 1) {
                const double w_m1 = (zisoi[j] - zext(j - 1)) / dzn(j);
                const double d_m1 = (dif_0 > 0.0 && dif_m > 0.0)
                                        ? 1.0 / ((1.0 - w_m1) / dif_0 + w_m1 / dif_m)
                                        : 0.0;
                d_m1_zm1 = d_m1 / dzn(j);
                d_p1_zp1 = d_m1_zm1;
                f_m1 = adv_0;
                f_p1 = 0.0;
                pe_m1 = f_m1 / d_m1_zm1;
                pe_p1 = f_p1 / d_p1_zp1;
            } else {
                const double w_m1 = (zisoi[j] - zext(j - 1)) / dzn(j);
                const double d_m1 = (dif_m > 0.0 && dif_0 > 0.0)
                                        ? 1.0 / ((1.0 - w_m1) / dif_0 + w_m1 / dif_m)
                                        : 0.0;
                const double w_p1 = (zext(j + 1) - zisoi[j + 1]) / dzn(j + 1);
                const double d_p1 = (dif_p > 0.0 && dif_0 > 0.0)
                                        ? 1.0 / ((1.0 - w_p1) / dif_0 + w_p1 / dif_p)
                                        : (1.0 - w_m1) * dif_0 + w_p1 * dif_p;
                d_m1_zm1 = d_m1 / dzn(j);
                d_p1_zp1 = d_p1 / dzn(j + 1);
                f_m1 = adv_0;
                f_p1 = adv_p;
                pe_m1 = f_m1 / d_m1_zm1;
                pe_p1 = f_p1 / d_p1_zp1;
            }
            // row jj = j + 1 (:396-426); j = nd + 1 is the ghost row a = -1, b = 1, c = r = 0
            double av, bv, cv, rv;
            if (j <= nd) {
                const double a_p_0 = dzd[j] / dtime;
                av = -(d_m1_zm1 * patankar_A(pe_m1) + jl_max(f_m1, 0.0));
                cv = -(d_p1_zp1 * patankar_A(pe_p1) + jl_max(-f_p1, 0.0));
                bv = -av - cv + a_p_0;
                const double sj = src[(int64_t)(j - 1) * ld];
                rv = j == 1 ? sj * dzd[j] / dtime + (a_p_0 - adv_0) * ct
                            : sj * dzd[j] / dtime + a_p_0 * ct;
            } else {
                av = -1.0;
                bv = 1.0;
                cv = 0.0;
                rv = 0.0;
            }
            // Thomas forward sweep, no pivot guard (:437-446)
            const double denom = bv - av * cp_prev;
            cp_prev = cv / denom;
            dp_prev = (rv - av * dp_prev) / denom;
            cpl[(int64_t)j * LV_BLOCK] = cp_prev;
            dpl[(int64_t)j * LV_BLOCK] = dp_prev;
            // shift the window
            dif_m = dif_0;
            dif_0 = dif_p;
            adv_0 = adv_p;
            if (j + 2 <= nd + 1) {
                dif_p = scaled(a.som_diffus_coef, j + 2);
                adv_p = scaled(a.som_adv_coef, j + 2);
            }
        }
        double u_next = dp_prev;                          // row nj, the ghost below the column
        for (int jj = nj - 1; jj >= 1; --jj) {
            u_next = dpl[(int64_t)(jj - 1) * LV_BLOCK] - cpl[(int64_t)(jj - 1) * LV_BLOCK] * u_next;
            dpl[(int64_t)(jj - 1) * LV_BLOCK] = u_next;
        }
        // transport tendency (:433-435, 462-465)
        for (int j = 1; j <= nd; ++j) {
            const double pre = 0.0 - (conc[(int64_t)(j - 1) * ld] + src[(int64_t)(j - 1) * ld]);
            tend[(int64_t)(j - 1) * ld] = (pre + dpl[(int64_t)j * LV_BLOCK]) / dtime;
        }
    }
    // write back, moving what sits below bedrock into the bedrock layer (:466-475)
    for (int j = 1; j <= nd; ++j) {
        const double v = dpl[(int64_t)j * LV_BLOCK];
        if (j > nbr) {
            conc[(int64_t)(nbr - 1) * ld] = conc[(int64_t)(nbr - 1) * ld] + v * (dzd[j] / dzd[nbr]);
            conc[(int64_t)(j - 1) * ld] = 0.0;
        } else {
            conc[(int64_t)(j - 1) * ld] = v;
        }
    }
}

}  // namespace

hipError_t launch_litter_vert_transp(const clmsoil_lvt_args &a, const clmsoil_lvt_params &p,
                                     double dtime, unsigned *err, hipStream_t stream) {
    const int64_t nc = a.nc;
    hipLaunchKernelGGL(lvt_coef_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream,
                       a, p, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)2 * (p.nlevdecomp + 2) * LV_BLOCK * sizeof(double);
    hipLaunchKernelGGL(lvt_transport_kernel,
                       dim3((unsigned)((nc + LV_BLOCK - 1) / LV_BLOCK), p.npools, p.ntypes),
                       dim3(LV_BLOCK), lds, stream, a, p, dtime);
    return hipGetLastError();
}

}  // namespace clmsoil
