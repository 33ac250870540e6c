// lake.hip — the lake temperature solve (gfx950), one lane per lake column.
//
// The reference builds the extended snow + lake + soil column in ten scratch
// arrays over five kernels, solves it with tridiagonal_multi! and unpacks it
// (src/biogeophys/lake_temperature.jl:632-785, 1205-1225). Here one lane walks
// its column once from the top active level down: the values of the levels
// below are read two steps ahead, the interface conductivity, explicit flux and
// Crank-Nicolson row of the current level are formed in registers, and the
// Thomas forward elimination consumes the row at once. The elimination's
// multipliers go to a lane-private LDS column ([level][lane], conflict-free,
// sized to the level count), its right-hand sides to the temperature arrays
// themselves, which the back substitution then overwrites with the solution.
// The arithmetic is the reference's operation by operation.
#include "common.hpp"
#include "lake.hpp"

namespace clmsoil {
namespace {

constexpr int LK_BLOCK = 64;

struct LakeNode {
    double zx, cvx, phix, tx;
    double tkv;         // tk of a snow or soil level, tk_lake of a lake level
    double dzl;         // dz_lake of a lake level
    double zraw;        // z of a snow or soil level
    double *t;          // where this level's temperature lives
};

struct LakeCol {
    const clmsoil_lake_solve_args &a;
    int64_t c;
    int nlevsno, nlevlak, jtop;
    double zbot;        // z_lake + dz_lake / 2 of the last lake level: the lake bottom
    __device__ double at(const double *p, int j) const { return p[c + (int64_t)(j - 1) * a.ld]; }
    __device__ double *ptr(double *p, int j) const { return p + c + (int64_t)(j - 1) * a.ld; }

    // _lake_extcol_setup_kernel! (lake_temperature.jl:632-664) for level je, with
    // the values _lake_tkix_kernel! (:673-703) reads at that level
    __device__ LakeNode node(int je) const {
        LakeNode n;
        n.dzl = 0.0;
        if (je < 1) {
            const int jj = je + nlevsno;
            n.zraw = at(a.z, jj);
            n.zx = n.zraw;
            n.cvx = at(a.cv, jj);
            n.phix = je == jtop ? 0.0 : at(a.sabg_lyr_col, je + nlevsno);
            n.tkv = at(a.tk, jj);
            n.t = ptr(a.t_soisno, jj);
        } else if (je <= nlevlak) {
            n.zraw = 0.0;
            n.zx = at(a.z_lake, je);
            n.cvx = at(a.cv_lake, je);
            n.phix = at(a.phi, je);
            n.tkv = at(a.tk_lake, je);
            n.dzl = at(a.dz_lake, je);
            n.t = ptr(a.t_lake, je);
        } else {
            const int jj = je - nlevlak + nlevsno;
            n.zraw = at(a.z, jj);
            n.zx = zbot + n.zraw;
            n.cvx = at(a.cv, jj);
            n.phix = je == nlevlak + 1 ? a.phi_soil[c] : 0.0;
            n.tkv = at(a.tk, jj);
            n.t = ptr(a.t_soisno, jj);
        }
        n.tx = *n.t;
        return n;
    }

    // _lake_tkix_kernel! (:673-703) for level je from the nodes je and je+1
    __device__ double tkix(int je, const LakeNode &n0, const LakeNode &n1, double tts) const {
        if (je < 0) return n0.tkv;
        if (je == 0) {
            const double dzp = n1.zx - n0.zx;
            return n1.tkv * n0.tkv * dzp / (n0.tkv * n1.zx + n1.tkv * (-n0.zraw));
        }
        if (je < nlevlak)
            return (n0.tkv * n1.tkv * (n1.dzl + n0.dzl)) / (n0.tkv * n1.dzl + n1.tkv * n0.dzl);
        if (je == nlevlak) {
            const double dzp = n1.zx - n0.zx;
            return tts * n0.tkv * dzp / (tts * n0.dzl / 2.0 + n0.tkv * n1.zraw);
        }
        return n0.tkv;
    }
};

__global__ void __launch_bounds__(LK_BLOCK) lake_solve_kernel(clmsoil_lake_solve_args a, int nlevsno,
                                                              int nlevgrnd, double dtime,
                                                              unsigned *err) {
    extern __shared__ double s_cp[];                 // [level][lane], levels of this grid
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * LK_BLOCK + lane;
    const int nlevlak = (int)a.nlevlak;
    const int jend = nlevlak + nlevgrnd;
    bool run = c < a.nc && (!a.mask_lakec || a.mask_lakec[c]);
    const int64_t snl = run ? a.snl[c] : 0;
    if (run && (snl < -nlevsno || snl > 0)) {
        atomicOr(err, (unsigned)DEV_ERR_SNL);
        run = false;
    }
    // lanes without a column get a top level below the column: they never work
    const int jtop = run ? (int)snl + 1 : jend + 1;
    LakeCol L{a, c, nlevsno, nlevlak, jtop, 0.0};
    const double cnfac = CNFAC, omc = 1.0 - cnfac;
    const double tiny = 1e-30;
    LakeNode cur{}, nxt{}, nn{};
    double fin = 0.0, tts = 0.0;
    if (run) {
        L.zbot = L.at(a.z_lake, nlevlak) + L.at(a.dz_lake, nlevlak) / 2.0;
        fin = a.fin[c];
        tts = a.tktopsoillay[c];
        cur = L.node(jtop);
        nxt = L.node(jtop + 1);
        nn = L.node(jtop + 2);          // jtop + 2 <= 3 <= nlevlak + nlevgrnd
    }
    double zx_prev = 0.0, tkix_prev = 0.0, fnx_prev = 0.0, cp_prev = 0.0, dp_prev = 0.0;
    // The level loop is the same in every lane of the wavefront, from its highest
    // top level down; a lane works from its own top level on, so every load and
    // store of a wavefront touches one row of the column-major arrays.
    int jtop_w = jtop;
    for (int o = 32; o > 0; o >>= 1) jtop_w = min(jtop_w, __shfl_xor(jtop_w, o, 64));
    for (int je = jtop_w; je <= jend; ++je) {
        if (je < jtop) continue;
        // _lake_tridiag_factors_kernel! (:705-722) and _lake_tridiag_coeffs_kernel! (:724-755)
        const double factx = dtime / cur.cvx;
        double tk = 0.0, fnx = 0.0, av, bv, cv, rv;
        if (je < jend) {
            tk = L.tkix(je, cur, nxt, tts);
            fnx = tk * (nxt.tx - cur.tx) / (nxt.zx - cur.zx);
        }
        if (je == jtop) {
            const double dzp = nxt.zx - cur.zx;
            av = 0.0;
            bv = 1.0 + omc * factx * tk / dzp;
            cv = -omc * factx * tk / dzp;
            rv = cur.tx + factx * (fin + cur.phix + cnfac * fnx);
        } else if (je < jend) {
            const double dzm = cur.zx - zx_prev;
            const double dzp = nxt.zx - cur.zx;
            av = -omc * factx * tkix_prev / dzm;
            bv = 1.0 + omc * factx * (tk / dzp + tkix_prev / dzm);
            cv = -omc * factx * tk / dzp;
            rv = cur.tx + cnfac * factx * (fnx - fnx_prev) + factx * cur.phix;
        } else {
            const double dzm = cur.zx - zx_prev;
            av = -omc * factx * tkix_prev / dzm;
            bv = 1.0 + omc * factx * tkix_prev / dzm;
            cv = 0.0;
            rv = cur.tx - cnfac * factx * fnx_prev;
        }
        // Thomas forward sweep (tridiagonal.jl:86-127)
        if (je == jtop) {
            double bet = bv;
            if (fabs(bet) < tiny) bet = tiny;
            cp_prev = cv / bet;
            dp_prev = rv / bet;
        } else {
            double denom = bv - av * cp_prev;
            if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0 ? 1.0 : denom);
            cp_prev = cv / denom;
            dp_prev = (rv - av * dp_prev) / denom;
        }
        s_cp[(je - jtop) * LK_BLOCK + lane] = cp_prev;
        *cur.t = dp_prev;
        zx_prev = cur.zx;
        tkix_prev = tk;
        fnx_prev = fnx;
        if (je < jend) {
            cur = nxt;
            nxt = nn;
            if (je + 3 <= jend) nn = L.node(je + 3);
        }
    }
    // back substitution into t_soisno / t_lake (_lake_unpack_kernel!, :757-785)
    double u_next = dp_prev;
    for (int je = jend - 1; je >= jtop_w; --je) {
        if (je < jtop) continue;
        double *t;
        if (je < 1)
            t = L.ptr(a.t_soisno, je + nlevsno);
        else if (je <= nlevlak)
            t = L.ptr(a.t_lake, je);
        else
            t = L.ptr(a.t_soisno, je - nlevlak + nlevsno);
        u_next = *t - s_cp[(je - jtop) * LK_BLOCK + lane] * u_next;
        __builtin_nontemporal_store(u_next, t);          // final: not read again
    }
}

}  // namespace

hipError_t launch_lake_temperature_solve(const clmsoil_lake_solve_args &a, int nlevsno,
                                         int nlevgrnd, double dtime, unsigned *err,
                                         hipStream_t stream) {
    const dim3 grid((unsigned)((a.nc + LK_BLOCK - 1) / LK_BLOCK));
    const size_t lds = (size_t)(nlevsno + a.nlevlak + nlevgrnd) * LK_BLOCK * sizeof(double);
    hipLaunchKernelGGL(lake_solve_kernel, grid, dim3(LK_BLOCK), lds, stream, a, nlevsno, nlevgrnd,
                       dtime, err);
    return hipGetLastError();
}

}  // namespace clmsoil
