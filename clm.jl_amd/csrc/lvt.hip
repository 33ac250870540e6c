// lvt.hip — vertical transport of the soil organic matter pools (gfx950).
//
// litter_vert_transp! (src/biogeochem/litter_vert_transp.jl:479-639) without the
// matrix-CN path. The coefficient kernel runs one lane per column. Without an
// accelerated-spinup factor the tridiagonal matrix of a column is the same for
// every pool and tracer type; only the right-hand side differs. For the usual
// level counts (10 and 20) the shared-matrix kernel therefore takes a block of 64
// columns and all of their (pool, tracer) systems: its wavefronts first form the
// Patankar rows of different levels in LDS, one wavefront runs the Thomas
// recurrence over them (denominators and multipliers, [level][lane],
// conflict-free), and then every wavefront solves its share of the right-hand
// sides with the column's forward values in registers, finishing the tendency,
// the bedrock leak and the stores from there. Each input and output element of
// the pools is read or written once.
//
// Pools with a spinup factor other than 1 (whose matrix depends on the pool and
// the latitude) and other level counts go to the per-system kernel: one lane per
// (column, pool, tracer), blockIdx.y a pool of its list, blockIdx.z the tracer. A
// lane walks its column once from the surface ghost level down, reading the
// diffusivities, concentrations and sources of the next levels ahead; the
// forward elimination parks its right-hand sides in the concentration slots and
// the pre-solve part of the tendency in the tendency slots, and the back
// substitution reads them back. Consecutive lanes take consecutive columns, so
// every load of a wavefront reads one row of the [c, j, s] arrays. The arithmetic
// of both kernels is the reference's operation by operation (Julia's x^5,
// max/min, no pivot guard).
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

// A list of pools (0-based) for one of the transport kernels.
struct LvtPools {
    int32_t n;
    int8_t pool[CLMSOIL_LVT_MAXPOOLS];
};

// The column's vertical grid (:566-575), 1-based: zsoi_ext, dz_node.
struct LvtGrid {
    const double *zisoi, *zsoi, *dzd;
    int nd;
    __device__ double zext(int j) const { return j <= nd ? zsoi[j] : zisoi[nd + 1]; }
    __device__ double dzn(int j) const { return j == 1 ? zext(1) : zext(j) - zext(j - 1); }
};

// Tridiagonal row jj = j + 1 (:318-426) of level j (1..nd) from the scaled
// diffusivities at levels j-1, j, j+1 and advection at j, j+1.
__device__ __forceinline__ void lvt_row(const LvtGrid &g, int j, int nbr, double dtime,
                                        double dif_m, double dif_0, double dif_p, double adv_0,
                                        double adv_p, double &av, double &bv, double &cv) {
    double d_m1_zm1, d_p1_zp1, f_m1, f_p1, pe_m1, pe_p1;
    if (j == 1) {
        d_m1_zm1 = 0.0;
        const double w_p1 = (g.zext(j + 1) - g.zisoi[j + 1]) / g.dzn(j + 1);
        const double d_p1 =
            (dif_p > 0.0 && dif_0 > 0.0) ? 1.0 / ((1.0 - w_p1) / dif_0 + w_p1 / dif_p) : 0.0;
        d_p1_zp1 = d_p1 / g.dzn(j + 1);
        f_m1 = adv_0;
        f_p1 = adv_p;
        pe_m1 = 0.0;
        pe_p1 = f_p1 / d_p1_zp1;
    } else if (j >= nbr + 1) {
        const double w_m1 = (g.zisoi[j] - g.zext(j - 1)) / g.dzn(j);
        const double d_m1 =
            (dif_0 > 0.0 && dif_m > 0.0) ? 1.0 / ((1.0 - w_m1) / dif_0 + w_m1 / dif_m) : 0.0;
        d_m1_zm1 = d_m1 / g.dzn(j);
        d_p1_zp1 = d_m1_zm1;
        f_m1 = adv_0;
        f_p1 = 0.0;
        pe_m1 = f_m1 / d_m1_zm1;
        pe_p1 = f_p1 / d_p1_zp1;
    } else {
        const double w_m1 = (g.zisoi[j] - g.zext(j - 1)) / g.dzn(j);
        const double d_m1 =
            (dif_m > 0.0 && dif_0 > 0.0) ? 1.0 / ((1.0 - w_m1) / dif_0 + w_m1 / dif_m) : 0.0;
        const double w_p1 = (g.zext(j + 1) - g.zisoi[j + 1]) / g.dzn(j + 1);
        const double d_p1 = (dif_p > 0.0 && dif_0 > 0.0)
                                ? 1.0 / ((1.0 - w_p1) / dif_0 + w_p1 / dif_p)
                                : (1.0 - w_m1) * dif_0 + w_p1 * dif_p;
        d_m1_zm1 = d_m1 / g.dzn(j);
        d_p1_zp1 = d_p1 / g.dzn(j + 1);
        f_m1 = adv_0;
        f_p1 = adv_p;
        pe_m1 = f_m1 / d_m1_zm1;
        pe_p1 = f_p1 / d_p1_zp1;
    }
    const double a_p_0 = g.dzd[j] / dtime;
    av = -(d_m1_zm1 * patankar_A(pe_m1) + jl_max(f_m1, 0.0));
    cv = -(d_p1_zp1 * patankar_A(pe_p1) + jl_max(-f_p1, 0.0));
    bv = -av - cv + a_p_0;
}

// _lvt_column_kernel! (:142-471) for pool pools.pool[blockIdx.y] of tracer blockIdx.z
__global__ void __launch_bounds__(LV_BLOCK) lvt_transport_kernel(clmsoil_lvt_args a,
                                                                 clmsoil_lvt_params p,
                                                                 LvtPools pools, double dtime) {
    extern __shared__ double lds[];                      // [nd + 2][LV_BLOCK] multipliers
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * LV_BLOCK + lane;
    const int s = pools.pool[blockIdx.y] + 1, t = blockIdx.z;
    if (c >= a.nc || (a.mask_bgc_soilc && !a.mask_bgc_soilc[c])) return;
    const int nd = p.nlevdecomp;
    const int64_t nbr64 = a.nbedrock[c];
    if (nbr64 < 1 || nbr64 > nd) return;                 // flagged by the coefficient kernel
    const int nbr = (int)nbr64;
    const int64_t ld = a.ld;
    const int64_t soff = (int64_t)(s - 1) * ld * p.nlevdecomp_full;
    double *conc = a.conc[t] + soff + c;
    const double *src = a.source[t] + soff + c;
    double *tend = a.tendency[t] + soff + c;
    const LvtGrid g{a.zisoi - 1, a.zsoi - 1, a.dzsoi_decomp - 1, nd};
    const double *dzd = g.dzd;
    double *cpl = lds + lane;                            // [jj - 1] at stride LV_BLOCK
    if (p.is_cwd[s - 1]) {
        // CWD: no transport, the source is added in place (:296-300)
        for (int j = 1; j <= nd; ++j)
            conc[(int64_t)(j - 1) * ld] = conc[(int64_t)(j - 1) * ld] + src[(int64_t)(j - 1) * ld];
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
        // sliding window over levels: diffusivity at j-1, j, j+1; advection at j, j+1.
        // The inputs of the next levels are read three levels ahead (coefficients of
        // j+3 .. j+5, concentrations and sources of j+1 .. j+3) so that a level's
        // loads are in flight while the levels above it are solved.
        auto raw = [&](const double *coef, int j) {
            return j <= nd + 1 ? coef[c + (int64_t)(j - 1) * ld] : 0.0;
        };
        auto lvl = [&](const double *x, int j) { return j <= nd ? x[(int64_t)(j - 1) * ld] : 0.0; };
        double dif_m = 0.0, dif_0 = scaled(a.som_diffus_coef, 1), dif_p = scaled(a.som_diffus_coef, 2);
        double adv_0 = scaled(a.som_adv_coef, 1), adv_p = scaled(a.som_adv_coef, 2);
        double qd1 = raw(a.som_diffus_coef, 3), qd2 = raw(a.som_diffus_coef, 4),
               qd3 = raw(a.som_diffus_coef, 5);
        double qa1 = raw(a.som_adv_coef, 3), qa2 = raw(a.som_adv_coef, 4), qa3 = raw(a.som_adv_coef, 5);
        double cq0 = lvl(conc, 1), cq1 = lvl(conc, 2), cq2 = lvl(conc, 3);
        double sq0 = lvl(src, 1), sq1 = lvl(src, 2), sq2 = lvl(src, 3);
        auto sc = [&](double v) { return fabs(v) * spinup_term < epsilon ? epsilon : v * spinup_term; };
        // row 0 (the surface ghost): a = 0, b = 1, c = -1, r = 0
        double cp_prev = -1.0 / 1.0, dp_prev = 0.0 / 1.0;
        cpl[0] = cp_prev;
        for (int j = 1; j <= nd + 1; ++j) {
            // conc_trcr[j+1] (:309-317): every level's concentration, 0 for the ghost
            // level nd + 1
            const double ct = j <= nd ? cq0 : 0.0;
            const double sj = sq0;
            // row jj = j + 1; j = nd + 1 is the ghost row a = -1, b = 1, c = r = 0
            double av, bv, cv, rv;
            if (j <= nd) {
                lvt_row(g, j, nbr, dtime, dif_m, dif_0, dif_p, adv_0, adv_p, av, bv, cv);
                const double a_p_0 = dzd[j] / dtime;
                rv = j == 1 ? sj * dzd[j] / dtime + (a_p_0 - adv_0) * ct
                            : sj * dzd[j] / dtime + a_p_0 * ct;
                // pre-solve part of the tendency (:433-435), parked in its slot
                tend[(int64_t)(j - 1) * ld] = 0.0 - (ct + sj);
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
            if (j <= nd) conc[(int64_t)(j - 1) * ld] = dp_prev;   // parked until the back substitution
            // shift the window and the queues of prefetched inputs
            dif_m = dif_0;
            dif_0 = dif_p;
            adv_0 = adv_p;
            dif_p = sc(qd1);
            adv_p = sc(qa1);
            qd1 = qd2;
            qd2 = qd3;
            qd3 = raw(a.som_diffus_coef, j + 5);
            qa1 = qa2;
            qa2 = qa3;
            qa3 = raw(a.som_adv_coef, j + 5);
            cq0 = cq1;
            cq1 = cq2;
            cq2 = lvl(conc, j + 3);
            sq0 = sq1;
            sq1 = sq2;
            sq2 = lvl(src, j + 3);
        }
        // back substitution from the ghost row below (dp_prev), levels nd .. 1; the
        // surface ghost's value is not needed. The tendency (:462-465) completes here.
        // (the parked values of the next levels up are read two levels ahead)
        double u_next = dp_prev;
        double pd0 = conc[(int64_t)(nd - 1) * ld], pt0 = tend[(int64_t)(nd - 1) * ld];
        double pd1 = nd >= 2 ? conc[(int64_t)(nd - 2) * ld] : 0.0;
        double pt1 = nd >= 2 ? tend[(int64_t)(nd - 2) * ld] : 0.0;
        for (int j = nd; j >= 1; --j) {
            const double pd2 = j >= 3 ? conc[(int64_t)(j - 3) * ld] : 0.0;
            const double pt2 = j >= 3 ? tend[(int64_t)(j - 3) * ld] : 0.0;
            u_next = pd0 - cpl[(int64_t)j * LV_BLOCK] * u_next;
            conc[(int64_t)(j - 1) * ld] = u_next;
            tend[(int64_t)(j - 1) * ld] = (pt0 + u_next) / dtime;
            pd0 = pd1;
            pd1 = pd2;
            pt0 = pt1;
            pt1 = pt2;
        }
    }
    // move what sits below bedrock into the bedrock layer (:466-475)
    for (int j = nbr + 1; j <= nd; ++j) {
        conc[(int64_t)(nbr - 1) * ld] =
            conc[(int64_t)(nbr - 1) * ld] + conc[(int64_t)(j - 1) * ld] * (dzd[j] / dzd[nbr]);
        conc[(int64_t)(j - 1) * ld] = 0.0;
    }
}

constexpr int LV_MAXWAVES = 8;

// Buffer access to the rows of a pool: a wavefront-uniform resource at the
// block's first column of level 1, the level's row offset in a scalar register
// and the lane's byte offset in one vector register for every row.
typedef unsigned int lv_u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t lv_rows(const double *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, -1, 0x00020000);
}
__device__ __forceinline__ double lv_load(__amdgpu_buffer_rsrc_t r, uint32_t boff, uint32_t roff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, boff, roff, 0));
}
// Stores of final values (nothing reads them again in the launch) are
// non-temporal (cache policy bit 1).
__device__ __forceinline__ void lv_store(__amdgpu_buffer_rsrc_t r, uint32_t boff, uint32_t roff,
                                         double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(lv_u32x2, v), r, boff, roff, 2);
}

// _lvt_column_kernel! (:142-471) for the nfast pools packed four bits each in
// `fast` (spinup factor 1 or coarse woody debris) and every tracer type, over a
// block of 64 columns, with nlevdecomp == ND. blockDim.x is 64 * waves; the host
// keeps (ND - 1) * ld * 8 + 512 below 2^32 so that every row offset fits the
// buffer instructions' 32-bit offsets.
template <int ND>
__global__ void __launch_bounds__(LV_MAXWAVES * LV_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
    lvt_shared_kernel(clmsoil_lvt_args a, clmsoil_lvt_params p, int nfast, uint64_t fast,
                      double dtime) {
    // the column's tridiagonal matrix: a, then b and c, which the Thomas recurrence
    // turns into its denominators and multipliers; rows 1..ND+1 ([level][lane])
    __shared__ double s_av[ND + 2][LV_BLOCK], s_dn[ND + 2][LV_BLOCK], s_cp[ND + 2][LV_BLOCK];
    __shared__ double s_ap[ND + 2];                      // dzsoi_decomp / dtime per level
    const int lane = threadIdx.x % LV_BLOCK;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / LV_BLOCK);   // wavefront-uniform
    const int nw = blockDim.x / LV_BLOCK;
    const int64_t c = (int64_t)blockIdx.x * LV_BLOCK + lane;
    const int64_t ld = a.ld;
    bool run = c < a.nc && (!a.mask_bgc_soilc || a.mask_bgc_soilc[c]);
    int nbr = ND;
    if (run) {
        const int64_t nb = a.nbedrock[c];
        if (nb < 1 || nb > ND) run = false;              // flagged by the coefficient kernel
        else nbr = (int)nb;
    }
    const LvtGrid g{a.zisoi - 1, a.zsoi - 1, a.dzsoi_decomp - 1, ND};
    const double *dzd = g.dzd;
    // spinup_term == 1: a coefficient below 1e-30 in magnitude becomes 1e-30 (:322-330)
    auto sc = [](double v) { return fabs(v) * 1.0 < 1.0e-30 ? 1.0e-30 : v * 1.0; };
    // rows of the levels j = 1 + w, 1 + w + nw, ... (every lane writes its slots)
    for (int j = 1 + w; j <= ND; j += nw) {
        double av = 0.0, bv = 1.0, cv = 0.0;
        if (run) {
            const double *dc = a.som_diffus_coef + c, *ac = a.som_adv_coef + c;
            const double dif_m = j > 1 ? sc(dc[(int64_t)(j - 2) * ld]) : 0.0;
            lvt_row(g, j, nbr, dtime, dif_m, sc(dc[(int64_t)(j - 1) * ld]), sc(dc[(int64_t)j * ld]),
                    sc(ac[(int64_t)(j - 1) * ld]), sc(ac[(int64_t)j * ld]), av, bv, cv);
        }
        s_av[j][lane] = av;
        s_dn[j][lane] = bv;
        s_cp[j][lane] = cv;
        if (lane == 0) s_ap[j] = dzd[j] / dtime;
    }
    if (w == 0) {                                        // the ghost row below the column
        s_av[ND + 1][lane] = -1.0;
        s_dn[ND + 1][lane] = 1.0;
        s_cp[ND + 1][lane] = 0.0;
    }
    __syncthreads();
    // Thomas forward recurrence of the matrix (:437-446) from the surface ghost row
    // (a = 0, b = 1, c = -1)
    if (w == 0) {
        double cp_prev = -1.0 / 1.0;
        for (int j = 1; j <= ND + 1; ++j) {
            const double denom = s_dn[j][lane] - s_av[j][lane] * cp_prev;
            cp_prev = s_cp[j][lane] / denom;
            s_dn[j][lane] = denom;
            s_cp[j][lane] = cp_prev;
        }
    }
    __syncthreads();
    if (!run) return;                                    // no barrier follows
    const double ap1m = s_ap[1] - sc(a.som_adv_coef[c]);  // a_p_0 - adv of level 1
    const int nsys = nfast * p.ntypes;
    const uint32_t boff = (uint32_t)lane * (uint32_t)sizeof(double);
    const uint32_t ld8 = (uint32_t)(ld * (int64_t)sizeof(double));
    // system r of this wavefront is tracer t, pool number k of the list (r = t * nfast + k)
    int t = w / nfast, k = w % nfast;
    for (int r = w; r < nsys; r += nw) {
        // read the matrix from LDS level by level in every system rather than
        // holding all of it in registers across the loop
        asm volatile("" ::: "memory");
        const int s = (int)((fast >> (4 * k)) & 15u);
        const int64_t soff = (int64_t)s * ld * p.nlevdecomp_full + (int64_t)blockIdx.x * LV_BLOCK;
        const __amdgpu_buffer_rsrc_t rc = lv_rows(a.conc[t] + soff), rs = lv_rows(a.source[t] + soff),
                                     rt = lv_rows(a.tendency[t] + soff);
        k += nw;
        while (k >= nfast) {
            k -= nfast;
            ++t;
        }
        double x[ND], y[ND];                             // concentrations, sources
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            x[j] = lv_load(rc, boff, (uint32_t)j * ld8);
            y[j] = lv_load(rs, boff, (uint32_t)j * ld8);
        }
        if (p.is_cwd[s]) {
            // CWD: no transport, the source is added in place (:296-300)
#pragma unroll
            for (int j = 0; j < ND; ++j) x[j] = x[j] + y[j];
        } else {
            // right-hand sides (:396-426) and forward elimination; x takes the
            // eliminated right-hand sides, y the pre-solve part of the tendency
            double dp = 0.0 / 1.0;
#pragma unroll
            for (int j = 1; j <= ND; ++j) {
                const double ct = x[j - 1], sj = y[j - 1];
                const double rv = j == 1 ? sj * dzd[j] / dtime + ap1m * ct
                                         : sj * dzd[j] / dtime + s_ap[j] * ct;
                y[j - 1] = 0.0 - (ct + sj);
                dp = (rv - s_av[j][lane] * dp) / s_dn[j][lane];
                x[j - 1] = dp;
                // keep each level's divisions with its level: hoisting the
                // independent ones ahead would hold their results in registers
                __builtin_amdgcn_sched_barrier(0);
            }
            // the ghost row (r = 0), then the back substitution and the tendency (:462-465)
            double u = (0.0 - s_av[ND + 1][lane] * dp) / s_dn[ND + 1][lane];
#pragma unroll
            for (int j = ND; j >= 1; --j) {
                u = x[j - 1] - s_cp[j][lane] * u;
                x[j - 1] = u;
                lv_store(rt, boff, (uint32_t)(j - 1) * ld8, (y[j - 1] + u) / dtime);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // move what sits below bedrock into the bedrock layer (:466-475), in the
        // reference's order of levels
        if (nbr < ND) {
            double acc = 0.0;
#pragma unroll
            for (int j = 1; j <= ND; ++j)
                if (j == nbr) acc = x[j - 1];
            const double dzb = dzd[nbr];
#pragma unroll
            for (int j = 2; j <= ND; ++j)
                if (j > nbr) {
                    acc = acc + x[j - 1] * (dzd[j] / dzb);
                    x[j - 1] = 0.0;
                }
#pragma unroll
            for (int j = 1; j <= ND; ++j)
                if (j == nbr) x[j - 1] = acc;
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) lv_store(rc, boff, (uint32_t)j * ld8, x[j]);
    }
}

template <int ND>
hipError_t launch_shared(const clmsoil_lvt_args &a, const clmsoil_lvt_params &p,
                         const LvtPools &pools, double dtime, hipStream_t stream) {
    const int nsys = pools.n * p.ntypes;
    const int waves = std::min(LV_MAXWAVES, std::max(1, (nsys + 1) / 2));
    uint64_t fast = 0;
    for (int i = 0; i < pools.n; ++i) fast |= (uint64_t)pools.pool[i] << (4 * i);
    hipLaunchKernelGGL(lvt_shared_kernel<ND>, dim3((unsigned)((a.nc + LV_BLOCK - 1) / LV_BLOCK)),
                       dim3(waves * LV_BLOCK), 0, stream, a, p, (int)pools.n, fast, dtime);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_litter_vert_transp(const clmsoil_lvt_args &a, const clmsoil_lvt_params &p,
                                     double dtime, unsigned *err, hipStream_t stream) {
    const int64_t nc = a.nc;
    hipLaunchKernelGGL(lvt_coef_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream,
                       a, p, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // pools whose matrix is the column's shared one go to the shared-matrix kernel
    // when it is built for this level count, the others to the per-system kernel
    const bool shared = (p.nlevdecomp == 20 || p.nlevdecomp == 10) &&
                        (uint64_t)(p.nlevdecomp - 1) * (uint64_t)a.ld * 8u + 512u < (1ull << 32);
    LvtPools fast{}, slow{};
    for (int s = 0; s < p.npools; ++s) {
        const bool unit = p.is_cwd[s] || p.spinup_state < 1 || p.spinup_factor[s] == 1.0;
        LvtPools &l = shared && unit ? fast : slow;
        l.pool[l.n++] = (int8_t)s;
    }
    if (fast.n > 0) {
        e = p.nlevdecomp == 20 ? launch_shared<20>(a, p, fast, dtime, stream)
                               : launch_shared<10>(a, p, fast, dtime, stream);
        if (e != hipSuccess) return e;
    }
    if (slow.n > 0) {
        const size_t lds = (size_t)(p.nlevdecomp + 2) * LV_BLOCK * sizeof(double);
        hipLaunchKernelGGL(lvt_transport_kernel,
                           dim3((unsigned)((nc + LV_BLOCK - 1) / LV_BLOCK), slow.n, p.ntypes),
                           dim3(LV_BLOCK), lds, stream, a, p, slow, dtime);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace clmsoil
